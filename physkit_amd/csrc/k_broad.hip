// Broadphase: per-step LBVH over every leaf's broadphase bounds and the pair-set
// update of broad_phase::calculate_pairs (include/physkit/collision/
// collision_phases.h:377-436), computed in its closed form
//   active'(a,b) = a != b && !(static_a && static_b) && overlap(fat_a, fat_b)
//                  && (active(a,b) || moved_a || moved_b)
// The reference keeps two incremental SAH trees (src/bvh.cpp:235-514); the pair SET
// does not depend on tree shape, so the device rebuilds a linear BVH every step:
//   48-bit Morton keys of leaf centroids -> radix sort -> Karras hierarchy ->
//   escape indices for stackless traversal -> chunked bottom-up refit.
// Internal nodes hold FP32 boxes rounded outward (conservative, 32 bytes per node
// with the child and escape links); every leaf test uses the exact FP64 leaf box, so
// each overlap decision is the reference's aabb::intersects.
#include "pk_access.hpp"

namespace pk
{
namespace
{
constexpr int kBlock = 256;
constexpr int kChunk = static_cast<int>(kRefitChunk); // leaves per refit workgroup
constexpr int kRefitThreads = 1024;
unsigned grid(size_t n, int b = kBlock) { return static_cast<unsigned>((n + b - 1) / b); }

__device__ __forceinline__ unsigned long long ord(double x)
{
    unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(x));
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double unord(unsigned long long o)
{
    unsigned long long u = (o >> 63) ? (o & 0x7fffffffffffffffull) : ~o;
    return __longlong_as_double(static_cast<long long>(u));
}
// float bounds of a double, rounded toward -inf / +inf
__device__ __forceinline__ float f_down(double x)
{
    float f = static_cast<float>(x);
    if (static_cast<double>(f) > x) f = nextafterf(f, -__builtin_inff());
    return f;
}
__device__ __forceinline__ float f_up(double x)
{
    float f = static_cast<float>(x);
    if (static_cast<double>(f) < x) f = nextafterf(f, __builtin_inff());
    return f;
}

__global__ void k_scene_init(unsigned long long *scene)
{
    int t = threadIdx.x;
    if (t < 3) scene[t] = ~0ull;
    else if (t < 6) scene[t] = 0ull;
}

__device__ __forceinline__ double wave_min(double x)
{
    for (int off = 32; off > 0; off >>= 1) x = fmin(x, __shfl_xor(x, off));
    return x;
}
__device__ __forceinline__ double wave_max(double x)
{
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off));
    return x;
}

// centroid bounds: grid-stride per thread, then wavefront and workgroup reductions,
// then one ordered-integer atomic per workgroup and component
__global__ void __launch_bounds__(kBlock) k_scene_bounds(bodies_dev b, unsigned long long *scene)
{
    __shared__ double red[2][3][kBlock / 64];
    double c[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double d[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < b.n; i += gridDim.x * blockDim.x)
        for (int k = 0; k < 3; ++k)
        {
            double v = (b.fat[6 * static_cast<size_t>(i) + k] + b.fat[6 * static_cast<size_t>(i) + k + 3]) * 0.5;
            c[k] = fmin(c[k], v);
            d[k] = fmax(d[k], v);
        }
    int w = threadIdx.x >> 6;
    for (int k = 0; k < 3; ++k)
    {
        c[k] = wave_min(c[k]);
        d[k] = wave_max(d[k]);
        if (__lane_id() == 0)
        {
            red[0][k][w] = c[k];
            red[1][k][w] = d[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < 3)
    {
        int k = threadIdx.x;
        double lo = red[0][k][0], hi = red[1][k][0];
        for (int j = 1; j < kBlock / 64; ++j)
        {
            lo = fmin(lo, red[0][k][j]);
            hi = fmax(hi, red[1][k][j]);
        }
        if (lo <= hi)
        {
            atomicMin(&scene[k], ord(lo));
            atomicMax(&scene[3 + k], ord(hi));
        }
    }
}

__device__ __forceinline__ uint64_t split3(uint64_t x)
{
    x &= 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ void k_morton(bodies_dev b, lbvh_dev t, const unsigned long long *scene, uint64_t *keys, uint32_t *ids)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.n) return;
    uint64_t q[3];
    for (int k = 0; k < 3; ++k)
    {
        double lo = unord(scene[k]), hi = unord(scene[3 + k]);
        double v = (b.fat[6 * static_cast<size_t>(i) + k] + b.fat[6 * static_cast<size_t>(i) + k + 3]) * 0.5;
        double u = hi > lo ? (v - lo) / (hi - lo) : 0.0;
        u = fmin(fmax(u, 0.0), 1.0);
        uint64_t x = static_cast<uint64_t>(u * 65536.0);
        q[k] = x > 65535ull ? 65535ull : x;
    }
    uint64_t m = (split3(q[0]) << 2) | (split3(q[1]) << 1) | split3(q[2]);
    if (t.n_worlds > 1) // world id in the top bits: each world's leaves sort together
        m = (static_cast<uint64_t>(t.world_of[i]) << (kMortonBits - t.wbits)) | (m >> t.wbits);
    keys[i] = m;
    ids[i] = i;
}

// Karras, "Maximizing parallelism in the construction of BVHs, octrees and k-d
// trees" (HPG 2012): internal node i covers a key range found by a binary search on
// the common-prefix length; equal keys fall back to the index bits.
__device__ __forceinline__ int delta(const uint64_t *k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], c = k[j];
    if (a == c) return 64 + __clz(static_cast<unsigned>(i ^ j));
    return __clzll(static_cast<long long>(a ^ c));
}

__global__ void k_build_internal(const uint64_t *k, int n, lbvh_dev t)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(k, n, i, i - d);
    int lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int s = lmax / 2; s >= 1; s /= 2)
        if (delta(k, n, i, i + (l + s) * d) > dmin) l += s;
    int j = i + l * d;
    int dnode = delta(k, n, i, j);
    int s = 0;
    int div = 2;
    while (true)
    {
        int st = (l + div - 1) / div;
        if (delta(k, n, i, i + (s + st) * d) > dnode) s += st;
        if (st == 1) break;
        div *= 2;
    }
    int gamma = i + s * d + (d < 0 ? d : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    int lc = (lo == gamma) ? (n - 1 + gamma) : gamma;
    int rc = (hi == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
    t.left[i] = lc;
    t.right[i] = rc;
    t.first[i] = static_cast<uint32_t>(lo);
    t.last[i] = static_cast<uint32_t>(hi);
    t.nrec[i] = make_uint4(static_cast<uint32_t>(lc), static_cast<uint32_t>(rc), static_cast<uint32_t>(lo),
                           static_cast<uint32_t>(hi));
    t.parent[lc] = static_cast<uint32_t>(i);
    t.parent[rc] = static_cast<uint32_t>(i);
    if (i == 0) t.parent[0] = kNone;
}

__device__ __forceinline__ void gather_leaf(const bodies_dev &b, uint32_t a, uint32_t i, bool mv, const lbvh_dev &t)
{
    t.leaf_body[i] = a;
    t.leaf_moved[i] = mv ? 1 : 0;
    // the body's 48-byte fat record, copied to the leaf's as three 16-byte moves
    const double2 *src = reinterpret_cast<const double2 *>(b.fat + 6 * static_cast<size_t>(a));
    double2 *dst = reinterpret_cast<double2 *>(t.leaf_box + 6 * static_cast<size_t>(i));
    const double2 r0 = src[0], r1 = src[1], r2 = src[2];
    dst[0] = r0;
    dst[1] = r1;
    dst[2] = r2;
    const double f[6] = {r0.x, r0.y, r1.x, r1.y, r2.x, r2.y};
    if (t.leaf_f32)
    {
        t.leaf_f32[2 * static_cast<size_t>(i)] = make_float4(f_down(f[0]), f_down(f[1]), f_down(f[2]), 0.0f);
        t.leaf_f32[2 * static_cast<size_t>(i) + 1] = make_float4(f_up(f[3]), f_up(f[4]), f_up(f[5]), 0.0f);
    }
    if (t.n == 1) t.parent[0] = kNone;
}

// Leaf i's body, moved flag and boxes, and the moved leaves per 256-leaf block. The
// moved leaves are then listed for the pair query in ascending leaf order (offsets of
// the block counts, then a scatter), so the query's concurrent wavefronts walk one
// contiguous range of leaves and share the tree nodes above it in L2.
static_assert(kBlock == 256, "the moved-leaf blocks are the gather's workgroups");
__global__ void __launch_bounds__(kBlock) k_leaf_gather(bodies_dev b, const uint32_t *ids, int all_moved, lbvh_dev t)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = t.n;
    bool mv = false;
    if (i < n)
    {
        const uint32_t a = ids[i];
        mv = !b.is_static[a] && (all_moved || b.moved[a]);
        gather_leaf(b, a, i, mv, t);
    }
    const int cnt = __syncthreads_count(mv);
    if (threadIdx.x == 0) t.moved_wg[blockIdx.x] = static_cast<uint32_t>(cnt);
}

// one workgroup: exclusive offsets of the n_blk block counts (moved_wg[n_blk + b]) and
// the total at counters[C_MLEAF]
__global__ void __launch_bounds__(1024) k_moved_offsets(lbvh_dev t, uint32_t n_blk, unsigned long long *counters)
{
    __shared__ uint32_t part[1024];
    const uint32_t per = (n_blk + 1023) / 1024, b0 = min(threadIdx.x * per, n_blk), b1 = min(b0 + per, n_blk);
    uint32_t s = 0;
    for (uint32_t k = b0; k < b1; ++k) s += t.moved_wg[k];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) // inclusive scan of the thread sums
    {
        const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (uint32_t k = b0; k < b1; ++k)
    {
        const uint32_t c = t.moved_wg[k];
        t.moved_wg[n_blk + k] = run;
        run += c;
    }
    if (threadIdx.x == 1023) counters[C_MLEAF] = part[1023];
}

// moved leaf i at its block's offset plus its rank in the block
__global__ void __launch_bounds__(kBlock) k_moved_scatter(lbvh_dev t, uint32_t n_blk)
{
    __shared__ uint32_t wcnt[kBlock / 64];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool mv = i < t.n && t.leaf_moved[i];
    const unsigned long long mask = __ballot(mv);
    const uint32_t wave = threadIdx.x >> 6;
    if (__lane_id() == 0) wcnt[wave] = static_cast<uint32_t>(__popcll(mask));
    __syncthreads();
    if (!mv) return;
    uint32_t at = t.moved_wg[n_blk + blockIdx.x];
    for (uint32_t w = 0; w < wave; ++w) at += wcnt[w];
    at += static_cast<uint32_t>(__popcll(mask & ((1ull << __lane_id()) - 1ull)));
    t.moved_list[at] = i;
}

// next node in left-first depth-first order after the subtree of u
__global__ void k_escape(lbvh_dev t)
{
    uint32_t n = t.n;
    uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= 2 * n - 1) return;
    uint32_t x = u, esc = kNone;
    for (uint32_t guard = 0; guard < 2 * n; ++guard)
    {
        uint32_t p = t.parent[x];
        if (p == kNone) break;
        if (static_cast<uint32_t>(t.left[p]) == x)
        {
            esc = static_cast<uint32_t>(t.right[p]);
            break;
        }
        x = p;
    }
    t.escape[u] = esc;
}

struct fbox
{
    float lo[3], hi[3];
    uint32_t moved; // 1 when every leaf below is a moved leaf
};

__device__ __forceinline__ fbox child_box(const lbvh_dev &t, uint32_t c)
{
    fbox f;
    uint32_t n = t.n;
    if (c >= n - 1)
    {
        uint32_t j = c - (n - 1);
        const double *p = t.leaf_box + static_cast<size_t>(j) * 6;
        for (int k = 0; k < 3; ++k)
        {
            f.lo[k] = f_down(p[k]);
            f.hi[k] = f_up(p[3 + k]);
        }
        f.moved = t.leaf_moved[j];
    }
    else
    {
        float4 a = t.inode_lo[c], b = t.inode_hi[c];
        f.lo[0] = a.x;
        f.lo[1] = a.y;
        f.lo[2] = a.z;
        f.hi[0] = b.x;
        f.hi[1] = b.y;
        f.hi[2] = b.z;
        f.moved = t.lastf[c] >> 31;
    }
    return f;
}

__device__ __forceinline__ uint32_t wg_arrive(uint32_t *p)
{ return __hip_atomic_fetch_add(p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP); }

// refit state of the nodes inside one workgroup's leaf chunk, in LDS: a node whose leaf
// range lies in [c_lo, c_hi] has its own index in that range (Karras nodes sit at one
// end of their range), so slot = node - c_lo
struct refit_lds
{
    uint32_t visit[kChunk];
    float box[kChunk][6];
    uint8_t moved[kChunk];
    fbox wave_tot[2][kRefitThreads / 64]; // per-wavefront totals of the chunk scans
    // the chunk's internal node records, escapes and parents, and its leaves' parents,
    // read once (coalesced) so the climbs touch only LDS for the tree links
    uint4 rec[kChunk];
    uint32_t esc[kChunk], ipar[kChunk], lpar[kChunk];
};

__device__ __forceinline__ fbox child_box_local(const lbvh_dev &t, uint32_t c, uint32_t c_lo, const refit_lds &L)
{
    uint32_t n = t.n;
    if (c >= n - 1) return child_box(t, c);
    fbox f;
    const float *b = L.box[c - c_lo];
    for (int k = 0; k < 3; ++k)
    {
        f.lo[k] = b[k];
        f.hi[k] = b[3 + k];
    }
    f.moved = L.moved[c - c_lo];
    return f;
}

__device__ __forceinline__ fbox empty_fbox()
{
    fbox f;
    for (int k = 0; k < 3; ++k)
    {
        f.lo[k] = __builtin_inff();
        f.hi[k] = -__builtin_inff();
    }
    f.moved = 1;
    return f;
}
__device__ __forceinline__ void merge_into(fbox &a, const fbox &b)
{
    for (int k = 0; k < 3; ++k)
    {
        a.lo[k] = fminf(a.lo[k], b.lo[k]);
        a.hi[k] = fmaxf(a.hi[k], b.hi[k]);
    }
    a.moved &= b.moved;
}
__device__ __forceinline__ void put_fbox(float4 *lo, float4 *hi, size_t i, const fbox &f)
{
    lo[i] = make_float4(f.lo[0], f.lo[1], f.lo[2], __uint_as_float(f.moved));
    hi[i] = make_float4(f.hi[0], f.hi[1], f.hi[2], 0.0f);
}
__device__ __forceinline__ fbox get_fbox(const float4 *lo, const float4 *hi, size_t i)
{
    const float4 a = lo[i], b = hi[i];
    fbox f;
    f.lo[0] = a.x;
    f.lo[1] = a.y;
    f.lo[2] = a.z;
    f.hi[0] = b.x;
    f.hi[1] = b.y;
    f.hi[2] = b.z;
    f.moved = __float_as_uint(a.w);
    return f;
}
// inclusive scan of boxes across the wavefront: toward higher lanes (UP) or lower
template <bool UP> __device__ __forceinline__ fbox scan_wave(fbox f)
{
    const int lane = static_cast<int>(__lane_id());
    for (int off = 1; off < 64; off <<= 1)
    {
        fbox o;
        for (int k = 0; k < 3; ++k)
        {
            o.lo[k] = UP ? __shfl_up(f.lo[k], off) : __shfl_down(f.lo[k], off);
            o.hi[k] = UP ? __shfl_up(f.hi[k], off) : __shfl_down(f.hi[k], off);
        }
        o.moved = UP ? __shfl_up(f.moved, off) : __shfl_down(f.moved, off);
        if (UP ? lane >= off : lane + off < 64) merge_into(f, o);
    }
    return f;
}

// Bottom-up refit. The second child to arrive at a node computes it.
// k_refit: a workgroup owns leaves [b*kChunk, (b+1)*kChunk) and refits the nodes whose
// leaf range lies inside the chunk, synchronising only within the workgroup
// (workgroup-scope acquire/release on LDS counters). A climb stops at the first node
// whose range crosses the chunk. The workgroup also writes, for every leaf of the chunk,
// the union of the chunk's leaves up to it (prefix) and from it on (suffix), and the
// chunk's total at level 0 of a sparse table over chunks.
// k_span_table: one workgroup fills the table's levels (level k, chunk c: the union of
// chunks [c, c + 2^k)).
// k_span: a node whose range [lo, hi] crosses chunks is the suffix at lo, the prefix at
// hi and at most two table entries covering the whole chunks between, so every such
// node is computed on its own, with no arrivals and no device-scope fences. Min and
// max are exact, so each box equals the union of its leaves' boxes in any order.
__device__ __forceinline__ void store_node(const lbvh_dev &t, uint32_t node, const uint4 &rec, uint32_t esc,
                                           const float *lo, const float *hi, uint32_t mv)
{
    t.inode_lo[node] = make_float4(lo[0], lo[1], lo[2], __uint_as_float(rec.x));
    t.inode_hi[node] = make_float4(hi[0], hi[1], hi[2], __uint_as_float(esc));
    t.lastf[node] = rec.w | (mv << 31);
}

__global__ void __launch_bounds__(kRefitThreads) k_refit(lbvh_dev t)
{
    static_assert(kRefitThreads == kChunk, "one thread per leaf of the chunk");
    __shared__ refit_lds L;
    uint32_t n = t.n;
    uint32_t c_lo = blockIdx.x * kChunk;
    uint32_t c_hi = min(c_lo + kChunk, n) - 1;
    const uint32_t tid = threadIdx.x;
    L.visit[tid] = 0;
    // chunk prefix and suffix unions of the leaf boxes
    const uint32_t i = c_lo + tid;
    const bool valid = i <= c_hi;
    if (i < n - 1 && i <= c_hi)
    {
        L.rec[tid] = t.nrec[i];
        L.esc[tid] = t.escape[i];
        L.ipar[tid] = t.parent[i];
    }
    if (valid) L.lpar[tid] = t.parent[n - 1 + i];
    const fbox leaf = valid ? child_box(t, n - 1 + i) : empty_fbox();
    fbox pre = scan_wave<true>(leaf);
    fbox suf = scan_wave<false>(leaf);
    const int wave = static_cast<int>(tid >> 6), lane = static_cast<int>(tid & 63);
    if (lane == 63) L.wave_tot[0][wave] = pre;
    if (lane == 0) L.wave_tot[1][wave] = suf;
    __syncthreads();
    for (int w = 0; w < wave; ++w) merge_into(pre, L.wave_tot[0][w]);
    for (int w = wave + 1; w < kRefitThreads / 64; ++w) merge_into(suf, L.wave_tot[1][w]);
    if (valid)
    {
        put_fbox(t.span_box, t.span_box + n, i, pre);
        put_fbox(t.span_box + 2 * static_cast<size_t>(n), t.span_box + 3 * static_cast<size_t>(n), i, suf);
        if (i == c_hi) put_fbox(t.span_tab, t.span_tab + t.span_levels * t.n_chunks, blockIdx.x, pre);
    }
    if (!valid) return;
    uint32_t cur = n - 1 + i;
    fbox mine = leaf;
    uint32_t node = L.lpar[tid];
    // a node whose range lies in the chunk has its index there (Karras nodes sit at one
    // end of their range): an index outside is a spanning node
    for (uint32_t guard = 0; node != kNone && guard < n; ++guard)
    {
        if (node < c_lo || node > c_hi) break; // spanning node: k_span computes it
        const uint4 rec = L.rec[node - c_lo];
        if (rec.z < c_lo || rec.w > c_hi) break;
        const uint32_t esc = L.esc[node - c_lo];
        if (wg_arrive(&L.visit[node - c_lo]) == 0) break;
        const uint32_t sib = rec.x == cur ? rec.y : rec.x;
        const fbox o = child_box_local(t, sib, c_lo, L);
        float lo[3] = {fminf(mine.lo[0], o.lo[0]), fminf(mine.lo[1], o.lo[1]), fminf(mine.lo[2], o.lo[2])};
        float hi[3] = {fmaxf(mine.hi[0], o.hi[0]), fmaxf(mine.hi[1], o.hi[1]), fmaxf(mine.hi[2], o.hi[2])};
        const uint32_t mv = mine.moved & o.moved;
        float *lb = L.box[node - c_lo];
        for (int k = 0; k < 3; ++k)
        {
            lb[k] = lo[k];
            lb[3 + k] = hi[k];
        }
        L.moved[node - c_lo] = static_cast<uint8_t>(mv);
        store_node(t, node, rec, esc, lo, hi, mv);
        for (int k = 0; k < 3; ++k)
        {
            mine.lo[k] = lo[k];
            mine.hi[k] = hi[k];
        }
        mine.moved = mv;
        cur = node;
        node = L.ipar[node - c_lo];
    }
}

__global__ void __launch_bounds__(kRefitThreads) k_span_table(lbvh_dev t)
{
    const uint32_t nch = t.n_chunks;
    float4 *lo = t.span_tab, *hi = t.span_tab + static_cast<size_t>(t.span_levels) * nch;
    for (uint32_t k = 1; k < t.span_levels; ++k)
    {
        __syncthreads(); // level k - 1 is complete
        const uint32_t half = 1u << (k - 1);
        const size_t src = static_cast<size_t>(k - 1) * nch, dst = static_cast<size_t>(k) * nch;
        for (uint32_t c = threadIdx.x; c + 2 * half <= nch; c += blockDim.x)
        {
            fbox a = get_fbox(lo, hi, src + c);
            merge_into(a, get_fbox(lo, hi, src + c + half));
            put_fbox(lo, hi, dst + c, a);
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_span(lbvh_dev t)
{
    const uint32_t n = t.n, nch = t.n_chunks;
    const float4 *tlo = t.span_tab, *thi = t.span_tab + static_cast<size_t>(t.span_levels) * nch;
    const uint32_t node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node < n - 1)
    {
        const uint4 rec = t.nrec[node];
        if (rec.z / kChunk == rec.w / kChunk) return; // inside one chunk: k_refit's
        const uint32_t esc = t.escape[node];
        fbox f = get_fbox(t.span_box + 2 * static_cast<size_t>(n), t.span_box + 3 * static_cast<size_t>(n), rec.z);
        merge_into(f, get_fbox(t.span_box, t.span_box + n, rec.w));
        const uint32_t c0 = rec.z / kChunk + 1, c1 = rec.w / kChunk; // whole chunks [c0, c1)
        if (c0 < c1)
        {
            const uint32_t len = c1 - c0;
            const uint32_t lv = 31u - static_cast<uint32_t>(__clz(len));
            const size_t row = static_cast<size_t>(lv) * nch;
            merge_into(f, get_fbox(tlo, thi, row + c0));
            merge_into(f, get_fbox(tlo, thi, row + c1 - (1u << lv)));
        }
        store_node(t, node, rec, esc, f.lo, f.hi, f.moved);
    }
}

// one thread per sorted leaf i; a moved dynamic leaf emits its pair with leaf j when
// j > i, or when j is not a moved leaf (j then never queries): each pair once. A
// subtree whose leaves all lie at or below i and are all moved holds no such j.
// Pairs collect in a per-thread buffer; the output space is reserved after the walk
// (a full buffer is flushed on its own). Same-address atomics serialise across the
// device (about 12 ns each): with one output region the workgroup reserves once; on
// large scenes workgroups write to several regions with a cursor each and reserve per
// wavefront, so no wavefront waits for its workgroup's longest walk.
constexpr int kEmitBuf = 32;

// The node whose leaf range is exactly [lo, hi]: a world's leaves share the world bits
// at the top of their keys, so they form one subtree, whose root (a Karras node sits at
// one end of its range) is node lo or node hi, or the leaf itself for a one-body world.
// kNone if neither matches.
__device__ __forceinline__ uint32_t range_root(const lbvh_dev &t, uint32_t lo, uint32_t hi)
{
    if (lo == hi) return t.n - 1 + lo;
    if (t.first[lo] == lo && t.last[lo] == hi) return lo;
    if (hi < t.n - 1 && t.first[hi] == lo && t.last[hi] == hi) return hi;
    return kNone;
}

__device__ __forceinline__ void flush_own(uint64_t *buf, uint32_t cnt, uint64_t *out_keys, uint32_t cap,
                                          unsigned long long *cursor)
{
    unsigned long long base = atomicAdd(cursor, static_cast<unsigned long long>(cnt));
    for (uint32_t k = 0; k < cnt; ++k)
        if (base + k < cap) out_keys[base + k] = buf[k];
}

__global__ void __launch_bounds__(kBlock)
    k_pair_query(lbvh_dev t, bodies_dev b, uint64_t *regions, uint32_t cap, uint32_t n_regions,
                 unsigned long long *counters)
{
    const uint32_t region = blockIdx.x % n_regions;
    uint64_t *out_keys = regions + static_cast<size_t>(region) * cap;
    unsigned long long *cursor = counters + C_QCUR + region * kCursorStride;
    uint64_t buf[kEmitBuf];
    uint32_t nbuf = 0;
    // thread k walks for the k-th moved leaf (sorted index i), list blocks in XCD order
    const uint32_t n = t.n;
    const uint32_t n_moved = static_cast<uint32_t>(counters[C_MLEAF]);
    const uint32_t k = xcd_block(n_moved, blockDim.x) * blockDim.x + threadIdx.x;
    const uint32_t i = k < n_moved ? t.moved_list[k] : n;
    bool active = false;
    uint32_t a = 0;
    box3 q{};
    float qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
    if (i < n && t.leaf_moved[i])
    {
        active = true;
        a = t.leaf_body[i];
        const double *p = t.leaf_box + static_cast<size_t>(i) * 6;
        q.min = mk(p[0], p[1], p[2]);
        q.max = mk(p[3], p[4], p[5]);
        for (int k = 0; k < 3; ++k)
        {
            qlo[k] = f_down(p[k]);
            qhi[k] = f_up(p[3 + k]);
        }
    }
    // leaves of other worlds are never visited: the walk starts at the root of the
    // querying leaf's world and ends at that subtree's escape, or (no such root) the
    // node's leaf range must meet the sorted range of the world
    const bool batched = t.n_worlds > 1;
    uint32_t wlo = 0, whi = n - 1;
    uint32_t node = active ? 0u : kNone, stop = kNone;
    bool ranged = false;
    if (active && batched)
    {
        uint32_t w = t.world_of[a];
        wlo = t.wstart[w];
        whi = t.wstart[w + 1] - 1;
        const uint32_t root = range_root(t, wlo, whi);
        if (root != kNone)
        {
            node = root;
            stop = t.escape[root];
        }
        else
            ranged = true;
    }
    for (uint32_t guard = 0; node != stop; ++guard)
    {
        if (guard >= 2 * n)
        {
            atomicAdd(&counters[C_TREE_FAIL], 1ull); // malformed tree: reported, never loops
            break;
        }
        uint32_t next;
        if (node < n - 1)
        {
            float4 lo = t.inode_lo[node], hi = t.inode_hi[node];
            bool ov = lo.x <= qhi[0] && hi.x >= qlo[0] && lo.y <= qhi[1] && hi.y >= qlo[1] && lo.z <= qhi[2] &&
                      hi.z >= qlo[2];
            next = __float_as_uint(hi.w);
            if (ov && ranged) ov = t.first[node] <= whi && (t.lastf[node] & 0x7fffffffu) >= wlo;
            if (ov)
            {
                uint32_t lf = t.lastf[node];
                bool skip = (lf & 0x7fffffffu) <= i && (lf >> 31);
                if (!skip) next = __float_as_uint(lo.w);
            }
        }
        else
        {
            uint32_t j = node - (n - 1);
            next = t.escape[node];
            bool near = j != i && j >= wlo && j <= whi;
            bool sure = false;
            if (near && t.leaf_f32)
            {
                // FP32 bounds rounded outward: lo32 <= lo < succ(lo32), pred(hi32) < hi <= hi32.
                // Apart in FP32: apart exactly. The successors of the lower bounds below the
                // query's upper bounds (and the predecessors of the upper bounds above its
                // lower bounds), compared in FP64: overlapping exactly. Only a leaf in
                // between reads its FP64 box.
                const float4 lo = t.leaf_f32[2 * static_cast<size_t>(j)], hi = t.leaf_f32[2 * static_cast<size_t>(j) + 1];
                near = lo.x <= qhi[0] && hi.x >= qlo[0] && lo.y <= qhi[1] && hi.y >= qlo[1] && lo.z <= qhi[2] &&
                       hi.z >= qlo[2];
                const float inf = __builtin_inff();
                sure = near && static_cast<double>(nextafterf(lo.x, inf)) <= q.max.x &&
                       static_cast<double>(nextafterf(lo.y, inf)) <= q.max.y &&
                       static_cast<double>(nextafterf(lo.z, inf)) <= q.max.z &&
                       static_cast<double>(nextafterf(hi.x, -inf)) >= q.min.x &&
                       static_cast<double>(nextafterf(hi.y, -inf)) >= q.min.y &&
                       static_cast<double>(nextafterf(hi.z, -inf)) >= q.min.z;
            }
            if (near)
            {
                bool ov = sure;
                if (!sure)
                {
                    const double *p = t.leaf_box + static_cast<size_t>(j) * 6;
                    box3 lb;
                    lb.min = mk(p[0], p[1], p[2]);
                    lb.max = mk(p[3], p[4], p[5]);
                    ov = box_intersects(lb, q);
                }
                if (ov && (j > i || !t.leaf_moved[j]))
                {
                    if (nbuf == kEmitBuf)
                    {
                        flush_own(buf, nbuf, out_keys, cap, cursor);
                        nbuf = 0;
                    }
                    buf[nbuf++] = pair_key(a, t.leaf_body[j]);
                }
            }
        }
        node = next;
    }
    // converged again: one reservation for the buffered pairs, per workgroup with one
    // region (fewer atomics), per wavefront with several (no wait for the workgroup's
    // longest walk)
    unsigned long long base = 0;
    if (n_regions == 1)
        base = block_reserve(cursor, nbuf);
    else
    {
        int lane = __lane_id();
        uint32_t incl = nbuf;
        for (int off = 1; off < 64; off <<= 1)
        {
            uint32_t v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        uint32_t total = __shfl(incl, 63);
        if (lane == 63 && total) base = atomicAdd(cursor, static_cast<unsigned long long>(total));
        base = __shfl(base, 63) + (incl - nbuf);
    }
    for (uint32_t k = 0; k < nbuf; ++k)
        if (base + k < cap) out_keys[base + k] = buf[k];
}

__device__ __forceinline__ uint32_t lower_bound(const uint64_t *a, uint32_t n, uint64_t key)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi)
    {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// candidate e found in the previous (sorted) pair list: its manifold carries over
__global__ void k_match(const uint64_t *cand, uint32_t m, const uint64_t *prev, uint32_t n_prev, uint32_t *src)
{
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    uint64_t k = cand[e];
    uint32_t j = lower_bound(prev, n_prev, k);
    src[e] = (j < n_prev && prev[j] == k) ? j : kNone;
}

// A previous pair no candidate covers stays while its fat boxes overlap
// (collision_phases.h:383-403). The candidates are every overlapping pair with a moved
// leaf, so an uncovered pair with a moved body no longer overlaps; with neither body
// moved, both fat boxes are the ones under which the pair was active last step (a fat
// box changes only when its body moves, and an upload clears the pair list), so it
// still overlaps. And a pair with neither body moved is never a candidate (each
// candidate has a moved leaf), so the test reduces to the two moved flags.
__global__ void k_prev_keep(const uint64_t *prev, uint32_t n_prev, bodies_dev b, uint32_t *flag)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_prev) return;
    uint64_t key = prev[p];
    uint32_t a = static_cast<uint32_t>(key >> 32), c = static_cast<uint32_t>(key & 0xFFFFFFFFu);
    flag[p] = (!b.moved[a] && !b.moved[c]) ? 1u : 0u;
}

__global__ void k_prev_compact(const uint64_t *prev, uint32_t n_prev, const uint32_t *flag, const uint32_t *pos,
                               uint64_t *akeys, uint32_t *asrc)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_prev || !flag[p]) return;
    akeys[pos[p]] = prev[p];
    asrc[pos[p]] = p;
}

// merge of two sorted, disjoint key lists: an element's output slot is its own index
// plus the number of keys of the other list below it
__global__ void k_merge(const uint64_t *xk, const uint32_t *xs, uint32_t nx, const uint64_t *yk, uint32_t ny,
                        uint64_t *out_keys, uint32_t *out_src)
{
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nx) return;
    uint64_t k = xk[e];
    uint32_t pos = e + lower_bound(yk, ny, k);
    out_keys[pos] = k;
    out_src[pos] = xs[e];
}
} // namespace

void launch_lbvh_build(lbvh_dev &t, const bodies_dev &b, bool all_moved, sort_scratch &sc,
                       unsigned long long *counters, hipStream_t st)
{
    uint32_t n = b.n;
    t.n = n;
    if (n == 0) return;
    k_scene_init<<<1, 64, 0, st>>>(t.scene);
    unsigned sb = grid(n) < 512 ? grid(n) : 512;
    k_scene_bounds<<<sb, kBlock, 0, st>>>(b, t.scene);
    k_morton<<<grid(n), kBlock, 0, st>>>(b, t, t.scene, t.morton, t.ids);
    static const int shifts[kMortonBits / 8] = {0, 8, 16, 24, 32, 40};
    if (radix_sort_u64(t.morton, t.ids, t.morton_alt, t.ids_alt, n, shifts, kMortonBits / 8, sc, st))
    {
        uint64_t *tk = t.morton;
        t.morton = t.morton_alt;
        t.morton_alt = tk;
        uint32_t *ti = t.ids;
        t.ids = t.ids_alt;
        t.ids_alt = ti;
    }
    if (n >= 2) k_build_internal<<<grid(n - 1), kBlock, 0, st>>>(t.morton, static_cast<int>(n), t);
    k_leaf_gather<<<grid(n), kBlock, 0, st>>>(b, t.ids, all_moved ? 1 : 0, t);
    k_moved_offsets<<<1, 1024, 0, st>>>(t, grid(n), counters);
    k_moved_scatter<<<grid(n), kBlock, 0, st>>>(t, grid(n));
    k_escape<<<grid(2 * n - 1), kBlock, 0, st>>>(t);
    if (n >= 2)
    {
        k_refit<<<t.n_chunks, kRefitThreads, 0, st>>>(t);
        if (t.n_chunks > 1)
        {
            if (t.span_levels > 1) k_span_table<<<1, kRefitThreads, 0, st>>>(t);
            k_span<<<grid(n - 1), kBlock, 0, st>>>(t);
        }
    }
}

// The last build's leaf order and hierarchy with this step's leaf boxes and moved flags:
// the pair set and query answers do not depend on the tree's shape, only the walks' cost.
void launch_lbvh_refit(lbvh_dev &t, const bodies_dev &b, bool all_moved, unsigned long long *counters,
                       hipStream_t st)
{
    const uint32_t n = b.n;
    t.n = n;
    if (n == 0) return;
    k_leaf_gather<<<grid(n), kBlock, 0, st>>>(b, t.ids, all_moved ? 1 : 0, t);
    k_moved_offsets<<<1, 1024, 0, st>>>(t, grid(n), counters);
    k_moved_scatter<<<grid(n), kBlock, 0, st>>>(t, grid(n));
    if (n >= 2)
    {
        k_refit<<<t.n_chunks, kRefitThreads, 0, st>>>(t);
        if (t.n_chunks > 1)
        {
            if (t.span_levels > 1) k_span_table<<<1, kRefitThreads, 0, st>>>(t);
            k_span<<<grid(n - 1), kBlock, 0, st>>>(t);
        }
    }
}

__global__ void k_pair_concat(const uint64_t *regions, uint32_t rcap, region_counts rc, uint64_t *dst)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rc.off[kPairRegions]) return;
    uint32_t r = 0;
    while (e >= rc.off[r + 1]) ++r;
    dst[e] = regions[static_cast<size_t>(r) * rcap + (e - rc.off[r])];
}

void launch_pair_concat(const uint64_t *regions, uint32_t rcap, const region_counts &rc, uint64_t *dst,
                        hipStream_t st)
{
    const uint32_t m = rc.off[kPairRegions];
    if (m) k_pair_concat<<<grid(m), kBlock, 0, st>>>(regions, rcap, rc, dst);
}

void launch_pair_query(const lbvh_dev &t, const bodies_dev &b, uint64_t *out_keys, uint32_t cap,
                       uint32_t n_regions, unsigned long long *counters, hipStream_t st)
{
    if (t.n < 2) return;
    // a multiple of 8 workgroups, so every XCD's eighth of the list has its workgroups
    k_pair_query<<<(grid(t.n) + 7) / 8 * 8, kBlock, 0, st>>>(t, b, out_keys, cap, n_regions, counters);
}

void launch_pair_match(const uint64_t *cand, uint32_t m, const uint64_t *prev, uint32_t n_prev, uint32_t *src,
                       hipStream_t st)
{
    if (m) k_match<<<grid(m), kBlock, 0, st>>>(cand, m, prev, n_prev, src);
}

void launch_prev_keep(const uint64_t *prev, uint32_t n_prev, const bodies_dev &b, uint32_t *flag, hipStream_t st)
{
    if (n_prev) k_prev_keep<<<grid(n_prev), kBlock, 0, st>>>(prev, n_prev, b, flag);
}

void launch_prev_compact(const uint64_t *prev, uint32_t n_prev, const uint32_t *flag, const uint32_t *pos,
                         uint64_t *akeys, uint32_t *asrc, hipStream_t st)
{
    if (n_prev) k_prev_compact<<<grid(n_prev), kBlock, 0, st>>>(prev, n_prev, flag, pos, akeys, asrc);
}

void launch_merge(const uint64_t *xk, const uint32_t *xs, uint32_t nx, const uint64_t *yk, uint32_t ny,
                  uint64_t *out_keys, uint32_t *out_src, hipStream_t st)
{
    if (nx) k_merge<<<grid(nx), kBlock, 0, st>>>(xk, xs, nx, yk, ny, out_keys, out_src);
}

} // namespace pk
