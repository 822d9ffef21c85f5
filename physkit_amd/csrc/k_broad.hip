// Broadphase: per-step LBVH over every leaf's broadphase bounds and the pair-set
// update of broad_phase::calculate_pairs (include/physkit/collision/
// collision_phases.h:377-436), computed in its closed form
//   active'(a,b) = a != b && !(static_a && static_b) && overlap(fat_a, fat_b)
//                  && (active(a,b) || moved_a || moved_b)
// The reference keeps two incremental SAH trees (src/bvh.cpp:235-514); the pair SET
// does not depend on tree shape, so the device rebuilds a linear BVH every step:
//   63-bit Morton keys of leaf centroids -> radix sort -> Karras hierarchy ->
//   bottom-up refit with arrival counters -> escape indices for stackless traversal.
// Node boxes are exact FP64 unions, so every overlap test is the reference's test.
#include "pk_access.hpp"

namespace pk
{
namespace
{
constexpr int kBlock = 256;
unsigned grid(size_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

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

__global__ void k_scene_bounds(bodies_dev b, unsigned long long *scene)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double c[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double d[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    if (i < b.n)
    {
        for (int k = 0; k < 3; ++k)
        {
            double v = (b.fat[static_cast<size_t>(k) * b.n + i] + b.fat[static_cast<size_t>(k + 3) * b.n + i]) * 0.5;
            c[k] = v;
            d[k] = v;
        }
    }
    for (int k = 0; k < 3; ++k)
    {
        c[k] = wave_min(c[k]);
        d[k] = wave_max(d[k]);
    }
    if (__lane_id() == 0)
        for (int k = 0; k < 3; ++k)
        {
            if (c[k] <= d[k])
            {
                atomicMin(&scene[k], ord(c[k]));
                atomicMax(&scene[3 + k], ord(d[k]));
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

__global__ void k_morton(bodies_dev b, const unsigned long long *scene, uint64_t *keys, uint32_t *ids)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.n) return;
    uint64_t q[3];
    for (int k = 0; k < 3; ++k)
    {
        double lo = unord(scene[k]), hi = unord(scene[3 + k]);
        double v = (b.fat[static_cast<size_t>(k) * b.n + i] + b.fat[static_cast<size_t>(k + 3) * b.n + i]) * 0.5;
        double u = hi > lo ? (v - lo) / (hi - lo) : 0.0;
        u = fmin(fmax(u, 0.0), 1.0);
        uint64_t x = static_cast<uint64_t>(u * 2097152.0);
        q[k] = x > 2097151ull ? 2097151ull : x;
    }
    keys[i] = (split3(q[0]) << 2) | (split3(q[1]) << 1) | split3(q[2]);
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

__global__ void k_build_internal(const uint64_t *k, int n, int32_t *left, int32_t *right, uint32_t *parent)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(k, n, i, i - d);
    int lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(k, n, i, j);
    int s = 0;
    int div = 2;
    while (true)
    {
        int t = (l + div - 1) / div;
        if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
        div *= 2;
    }
    int gamma = i + s * d + (d < 0 ? d : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    int lc = (lo == gamma) ? (n - 1 + gamma) : gamma;
    int rc = (hi == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
    left[i] = lc;
    right[i] = rc;
    parent[lc] = static_cast<uint32_t>(i);
    parent[rc] = static_cast<uint32_t>(i);
    if (i == 0) parent[0] = kNone;
}

__global__ void k_leaf_gather(bodies_dev b, const uint32_t *ids, uint32_t n, double *node_box,
                              uint32_t *leaf_body, uint32_t *parent)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t a = ids[i];
    leaf_body[i] = a;
    size_t node = static_cast<size_t>(n - 1 + i);
    for (int k = 0; k < 6; ++k) node_box[node * 6 + k] = b.fat[static_cast<size_t>(k) * b.n + a];
    if (n == 1) parent[0] = kNone;
}

__global__ void k_refit(uint32_t n, const int32_t *left, const int32_t *right, const uint32_t *parent,
                        double *node_box, uint32_t *visit)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t node = parent[n - 1 + i];
    for (uint32_t guard = 0; node != kNone && guard < n; ++guard)
    {
        __threadfence();
        uint32_t old = atomicAdd(&visit[node], 1u);
        __threadfence();
        if (old == 0) return; // the sibling subtree is not finished yet
        const double *lb = node_box + static_cast<size_t>(left[node]) * 6;
        const double *rb = node_box + static_cast<size_t>(right[node]) * 6;
        double *o = node_box + static_cast<size_t>(node) * 6;
        for (int k = 0; k < 3; ++k) o[k] = smin(lb[k], rb[k]);
        for (int k = 3; k < 6; ++k) o[k] = smax(lb[k], rb[k]);
        node = parent[node];
    }
}

// next node in left-first depth-first order after the subtree of u
__global__ void k_escape(uint32_t n, const int32_t *left, const int32_t *right, const uint32_t *parent,
                         uint32_t *escape)
{
    uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= 2 * n - 1) return;
    uint32_t x = u, esc = kNone;
    for (uint32_t guard = 0; guard < 2 * n; ++guard)
    {
        uint32_t p = parent[x];
        if (p == kNone) break;
        if (static_cast<uint32_t>(left[p]) == x)
        {
            esc = static_cast<uint32_t>(right[p]);
            break;
        }
        x = p;
    }
    escape[u] = esc;
}

__device__ __forceinline__ box3 node_box_at(const double *nb, uint32_t node)
{
    const double *p = nb + static_cast<size_t>(node) * 6;
    box3 b;
    b.min = mk(p[0], p[1], p[2]);
    b.max = mk(p[3], p[4], p[5]);
    return b;
}

// one thread per sorted leaf i; a moved dynamic leaf emits its pair with leaf j
// when j > i, or when j is not a moved leaf (j then never queries): each pair once.
__global__ void k_pair_query(lbvh_dev t, bodies_dev b, int all_moved, uint64_t *out_keys,
                             uint32_t *out_vals, uint32_t cap, unsigned long long *counters)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = t.n;
    bool active = false;
    uint32_t a = 0;
    box3 q{};
    if (i < n)
    {
        a = t.leaf_body[i];
        active = !b.is_static[a] && (all_moved || b.moved[a]);
        if (active) q = node_box_at(t.node_box, n - 1 + i);
    }
    uint32_t node = active ? 0u : kNone;
    // a depth-first walk visits each of the 2n-1 nodes at most once
    for (uint32_t guard = 0; node != kNone; ++guard)
    {
        if (guard >= 2 * n)
        {
            atomicAdd(&counters[C_TREE_FAIL], 1ull); // malformed tree: reported, never loops
            break;
        }
        box3 nb = node_box_at(t.node_box, node);
        bool emit = false;
        uint64_t key = 0;
        uint32_t next;
        if (box_intersects(nb, q))
        {
            if (node >= n - 1)
            {
                uint32_t j = node - (n - 1);
                if (j != i)
                {
                    uint32_t bb = t.leaf_body[j];
                    bool moved_b = !b.is_static[bb] && (all_moved || b.moved[bb]);
                    if (j > i || !moved_b)
                    {
                        emit = true;
                        key = pair_key(a, bb);
                    }
                }
                next = t.escape[node];
            }
            else
                next = static_cast<uint32_t>(t.left[node]);
        }
        else
            next = t.escape[node];
        unsigned long long slot = wave_append(&counters[C_COMBINED], emit);
        if (emit && slot < cap)
        {
            out_keys[slot] = key;
            out_vals[slot] = kNone;
        }
        node = next;
    }
}

__global__ void k_prev_filter(const uint64_t *prev, uint32_t n_prev, bodies_dev b, uint64_t *out_keys,
                              uint32_t *out_vals, uint32_t cap, unsigned long long *counter)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    bool keep = false;
    uint64_t key = 0;
    if (p < n_prev)
    {
        key = prev[p];
        uint32_t a = static_cast<uint32_t>(key >> 32), c = static_cast<uint32_t>(key & 0xFFFFFFFFu);
        keep = box_intersects(ldbox(b.fat, b.n, a), ldbox(b.fat, b.n, c)); // collision_phases.h:396
    }
    unsigned long long slot = wave_append(counter, keep);
    if (keep && slot < cap)
    {
        out_keys[slot] = key;
        out_vals[slot] = p;
    }
}

__global__ void k_dedupe(const uint64_t *keys, const uint32_t *vals, uint32_t n, uint32_t *flag, uint32_t *src)
{
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    uint64_t k = keys[e];
    bool first = e == 0 || keys[e - 1] != k;
    flag[e] = first ? 1u : 0u;
    if (first)
    {
        uint32_t s = vals[e];
        if (s == kNone && e + 1 < n && keys[e + 1] == k) s = vals[e + 1];
        src[e] = s;
    }
}

__global__ void k_compact(const uint64_t *keys, const uint32_t *src, const uint32_t *flag, const uint32_t *pos,
                          uint32_t n, uint64_t *out_keys, uint32_t *out_src)
{
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n || !flag[e]) return;
    out_keys[pos[e]] = keys[e];
    out_src[pos[e]] = src[e];
}
} // namespace

void launch_lbvh_build(lbvh_dev &t, const bodies_dev &b, sort_scratch &sc, hipStream_t st)
{
    uint32_t n = b.n;
    t.n = n;
    if (n == 0) return;
    k_scene_init<<<1, 64, 0, st>>>(t.scene);
    k_scene_bounds<<<grid(n), kBlock, 0, st>>>(b, t.scene);
    k_morton<<<grid(n), kBlock, 0, st>>>(b, t.scene, t.morton, t.ids);
    static const int shifts[8] = {0, 8, 16, 24, 32, 40, 48, 56};
    if (radix_sort_u64(t.morton, t.ids, t.morton_alt, t.ids_alt, n, shifts, 8, sc, st))
    {
        uint64_t *tk = t.morton;
        t.morton = t.morton_alt;
        t.morton_alt = tk;
        uint32_t *ti = t.ids;
        t.ids = t.ids_alt;
        t.ids_alt = ti;
    }
    if (n >= 2) k_build_internal<<<grid(n - 1), kBlock, 0, st>>>(t.morton, static_cast<int>(n), t.left, t.right, t.parent);
    k_leaf_gather<<<grid(n), kBlock, 0, st>>>(b, t.ids, n, t.node_box, t.leaf_body, t.parent);
    if (n >= 2)
    {
        launch_fill_u32(t.visit, 0u, n - 1, st);
        k_refit<<<grid(n), kBlock, 0, st>>>(n, t.left, t.right, t.parent, t.node_box, t.visit);
    }
    k_escape<<<grid(2 * n - 1), kBlock, 0, st>>>(n, t.left, t.right, t.parent, t.escape);
}

void launch_pair_query(const lbvh_dev &t, const bodies_dev &b, bool all_moved, uint64_t *out_keys,
                       uint32_t *out_vals, uint32_t cap, unsigned long long *counters, hipStream_t st)
{
    if (t.n < 2) return;
    k_pair_query<<<grid(t.n), kBlock, 0, st>>>(t, b, all_moved ? 1 : 0, out_keys, out_vals, cap, counters);
}

void launch_prev_filter(const uint64_t *prev_keys, uint32_t n_prev, const bodies_dev &b,
                        uint64_t *out_keys, uint32_t *out_vals, uint32_t cap,
                        unsigned long long *counter, hipStream_t st)
{
    if (n_prev) k_prev_filter<<<grid(n_prev), kBlock, 0, st>>>(prev_keys, n_prev, b, out_keys, out_vals, cap, counter);
}

void launch_dedupe_flags(const uint64_t *keys, const uint32_t *vals, uint32_t n, uint32_t *flag,
                         uint32_t *src, hipStream_t st)
{
    if (n) k_dedupe<<<grid(n), kBlock, 0, st>>>(keys, vals, n, flag, src);
}

void launch_compact_pairs(const uint64_t *keys, const uint32_t *src, const uint32_t *flag,
                          const uint32_t *pos, uint32_t n, uint64_t *out_keys, uint32_t *out_src,
                          hipStream_t st)
{
    if (n) k_compact<<<grid(n), kBlock, 0, st>>>(keys, src, flag, pos, n, out_keys, out_src);
}

} // namespace pk
