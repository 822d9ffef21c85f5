// Narrowphase: narrow_phase::calculate (include/physkit/collision/collision_phases.h:
// 244-320) split into three passes so each runs dense:
//   1. GJK, one thread per active pair; hits are compacted (wavefront-aggregated
//      append) with their final simplex;
//   2. EPA, one thread per hit, polytope/faces/heap in a private 192-face store;
//      a hit that needs more faces is rerun by a second kernel with a 4096-face store
//      in global memory, so capacity never changes a result;
//   3. manifold update (warm-start transfer, persistence, add/reduce, events), one
//      thread per pair, reading the previous step's manifold by source index.
#include "pk_access.hpp"
#include "pk_gjk.hpp"

#include <type_traits>

namespace pk
{
namespace
{
constexpr int kBlock = 256;
constexpr int kEpaBlock = 64;
unsigned grid(size_t n, int b = kBlock) { return static_cast<unsigned>((n + b - 1) / b); }

// UV: the shapes may include tagged UV spheres (a kernel built without it carries no
// code for their fast support)
template <bool UV> __device__ __forceinline__ gshape body_shape(const bodies_dev &b, const shapes_dev &s, uint32_t id)
{
    gshape g;
    g.kind = 2;
    g.a = ld3(b.pos, b.n, id);
    g.b = mk(0.0, 0.0, 0.0);
    g.q = ldq(b.quat, b.n, id);
    uint32_t sh = b.shape[id];
    g.v = s.verts + 3 * static_cast<size_t>(s.voff[sh]);
    g.nv = s.off[sh + 1] - s.off[sh];
    g.uv = UV ? s.uv[sh] : 0u;
    g.paired = 1;
    return g;
}

// GJK cull. Two bodies are apart when their bounding spheres (radius: the largest
// vertex distance from the shape origin) or their oriented local boxes (the shape's
// local AABB under the body's rotation; separating-axis test over 15 axes) are apart by
// more than the margin 1e-6 m + 1e-6 of the radii. Every point of the Minkowski
// difference then has norm above 1e-6, so GJK's first-support test (|p|^2 < 1e-12,
// collision.cpp:170) cannot fire, and with a gap this far above the rounding of its
// normals the simplex cases report no enclosure (a tetrahedron is only formed with the
// new point past the origin from the previous triangle's plane), so gjk_collision
// returns false: the verdict is the reference's without running it. The tests assume
// the support map rotate(q, .) is a rotation: a quaternion whose squared norm is off 1
// by more than 1e-9 (rotate then scales lengths by up to 1 + 2|q.q - 1|) is never
// culled, nor is anything non-finite (every comparison with NaN is false).
struct cull_body
{
    v3 pos, c, e; // position, oriented box centre and half extents
    v3 ax[3];     // rotate(q, unit axes)
    double r, qn;
};
__device__ __forceinline__ cull_body cull_of(const bodies_dev &b, const shapes_dev &s, uint32_t id)
{
    cull_body o;
    const uint32_t sh = b.shape[id];
    const qt q = ldq(b.quat, b.n, id);
    o.qn = ((q.w * q.w + q.x * q.x) + q.y * q.y) + q.z * q.z;
    o.pos = ld3(b.pos, b.n, id);
    o.r = s.rad[sh];
    const box3 lb = ldbox(s.local_box, s.n, sh);
    o.ax[0] = rotate(q, mk(1.0, 0.0, 0.0));
    o.ax[1] = rotate(q, mk(0.0, 1.0, 0.0));
    o.ax[2] = rotate(q, mk(0.0, 0.0, 1.0));
    const v3 lc = (lb.min + lb.max) * 0.5;
    o.e = (lb.max - lb.min) * 0.5;
    o.c = o.pos + ((o.ax[0] * lc.x + o.ax[1] * lc.y) + o.ax[2] * lc.z);
    return o;
}
// OBB-OBB separating axes (Ericson, Real-Time Collision Detection, 4.4.1); an edge
// axis is not normalised and is at most unit length, so its test is conservative
__device__ __forceinline__ bool obbs_apart(const cull_body &A, const cull_body &B, double m)
{
    double R[3][3], AR[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
        {
            R[i][j] = dot(A.ax[i], B.ax[j]);
            AR[i][j] = fabs(R[i][j]) + 1e-12;
        }
    const v3 d = B.c - A.c;
    const double t[3] = {dot(d, A.ax[0]), dot(d, A.ax[1]), dot(d, A.ax[2])};
    const double ea[3] = {A.e.x, A.e.y, A.e.z}, eb[3] = {B.e.x, B.e.y, B.e.z};
    for (int i = 0; i < 3; ++i)
        if (fabs(t[i]) > ea[i] + ((eb[0] * AR[i][0] + eb[1] * AR[i][1]) + eb[2] * AR[i][2]) + m) return true;
    for (int j = 0; j < 3; ++j)
    {
        const double tj = (t[0] * R[0][j] + t[1] * R[1][j]) + t[2] * R[2][j];
        if (fabs(tj) > ((ea[0] * AR[0][j] + ea[1] * AR[1][j]) + ea[2] * AR[2][j]) + eb[j] + m) return true;
    }
    for (int i = 0; i < 3; ++i)
    {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
        for (int j = 0; j < 3; ++j)
        {
            const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            const double ra = ea[i1] * AR[i2][j] + ea[i2] * AR[i1][j];
            const double rb = eb[j1] * AR[i][j2] + eb[j2] * AR[i][j1];
            if (fabs(t[i2] * R[i1][j] - t[i1] * R[i2][j]) > ra + rb + m) return true;
        }
    }
    return false;
}
// Per-body cull record, written once per step: 20 doubles (160 bytes, 16-byte
// aligned) so a pair reads two contiguous records instead of about twenty scattered
// planar words: pos.xyz, r | c.xyz, 0 | e.xyz, ax[0].x | ax[0].yz, ax[1].xy |
// ax[1].z, ax[2].xyz. A body whose quaternion is off unit norm gets r = infinity, which
// makes every test of its pairs fail.
constexpr int kCullRec = 20;
__device__ __forceinline__ void put_cull(double *rec, const cull_body &o)
{
    double2 *q = reinterpret_cast<double2 *>(rec);
    const bool unit = fabs(o.qn - 1.0) <= 1e-9;
    q[0] = make_double2(o.pos.x, o.pos.y);
    q[1] = make_double2(o.pos.z, unit ? o.r : __builtin_inf());
    q[2] = make_double2(o.c.x, o.c.y);
    q[3] = make_double2(o.c.z, 0.0);
    q[4] = make_double2(o.e.x, o.e.y);
    q[5] = make_double2(o.e.z, o.ax[0].x);
    q[6] = make_double2(o.ax[0].y, o.ax[0].z);
    q[7] = make_double2(o.ax[1].x, o.ax[1].y);
    q[8] = make_double2(o.ax[1].z, o.ax[2].x);
    q[9] = make_double2(o.ax[2].y, o.ax[2].z);
}
__device__ __forceinline__ void get_sphere(const double *rec, cull_body &o)
{
    const double2 *q = reinterpret_cast<const double2 *>(rec);
    const double2 a = q[0], b = q[1];
    o.pos = mk(a.x, a.y, b.x);
    o.r = b.y;
}
__device__ __forceinline__ void get_box(const double *rec, cull_body &o)
{
    const double2 *q = reinterpret_cast<const double2 *>(rec);
    const double2 c0 = q[2], c1 = q[3], e0 = q[4], e1 = q[5], a0 = q[6], a1 = q[7], a2 = q[8], a3 = q[9];
    o.c = mk(c0.x, c0.y, c1.x);
    o.e = mk(e0.x, e0.y, e1.x);
    o.ax[0] = mk(e1.y, a0.x, a0.y);
    o.ax[1] = mk(a1.x, a1.y, a2.x);
    o.ax[2] = mk(a2.y, a3.x, a3.y);
}
// the sphere test on the first 32 bytes of each record; the box test reads the rest
// only for pairs the spheres do not separate
__device__ __forceinline__ bool records_apart(const double *ra, const double *rb)
{
    cull_body A, B;
    get_sphere(ra, A);
    get_sphere(rb, B);
    const double r = A.r + B.r;
    const double m = 1e-6 * r + 1e-6; // infinite for a body off unit norm: nothing passes
    const v3 d = A.pos - B.pos;
    if (sqrt(dot(d, d)) > r + m) return true;
    get_box(ra, A);
    get_box(rb, B);
    return obbs_apart(A, B, m);
}

// the same tests from the planar body arrays (few pairs per body: no records)
__device__ __forceinline__ bool bodies_apart(const cull_body &A, const cull_body &B)
{
    if (!(fabs(A.qn - 1.0) <= 1e-9 && fabs(B.qn - 1.0) <= 1e-9)) return false;
    const double r = A.r + B.r;
    const double m = 1e-6 * r + 1e-6;
    const v3 d = A.pos - B.pos;
    if (sqrt(dot(d, d)) > r + m) return true;
    return obbs_apart(A, B, m);
}

// FP32 sphere records for the mixed form: (position rounded to nearest, radius rounded
// up), 16 bytes per body; infinite radius for a body the FP64 test never culls
__global__ void __launch_bounds__(kBlock) k_cull_prep(bodies_dev b, shapes_dev s, double *rec, float4 *rec32)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.n) return;
    const cull_body o = cull_of(b, s, i);
    put_cull(rec + static_cast<size_t>(kCullRec) * i, o);
    if (rec32)
    {
        const bool unit = fabs(o.qn - 1.0) <= 1e-9;
        float r = static_cast<float>(o.r);
        if (static_cast<double>(r) < o.r) r = nextafterf(r, __builtin_inff());
        rec32[i] = make_float4(static_cast<float>(o.pos.x), static_cast<float>(o.pos.y), static_cast<float>(o.pos.z),
                               unit ? r : __builtin_inff());
    }
}
// The FP32 spheres are certainly apart in the FP64 test's sense: with FP32 rounding of
// the positions (relative 2^-24 each) and of the distance (a few 2^-24), the FP64
// distance is at least d (1 - 1e-5) - 4e-7 (|pA|_1 + |pB|_1); with radii rounded up, the
// FP64 threshold r + 1e-6 r + 1e-6 is at most (rA + rB)(1 + 1e-5) + 2e-6. Infinite or
// NaN values never pass.
__device__ __forceinline__ bool spheres32_apart(float4 a, float4 b)
{
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    const float d = sqrtf(dx * dx + dy * dy + dz * dz);
    const float mag = ((fabsf(a.x) + fabsf(a.y)) + (fabsf(a.z) + fabsf(b.x))) + (fabsf(b.y) + fabsf(b.z));
    return d * (1.0f - 1e-5f) - 4e-7f * mag > (a.w + b.w) * (1.0f + 1e-5f) + 2e-6f;
}

// GJK simplex hand-off, planar: slot 0 = point count, then per point p.xyz and the
// two support features packed in one 64-bit word (kSimplexWords doubles per hit).
__device__ __forceinline__ void st_spt(double *d, uint32_t cap, uint32_t slot, int k, const spt &p)
{
    double *b = d + static_cast<size_t>(1 + 4 * k) * cap + slot;
    b[0] = p.p.x;
    b[cap] = p.p.y;
    b[2 * static_cast<size_t>(cap)] = p.p.z;
    b[3 * static_cast<size_t>(cap)] =
        __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(p.ib) << 32) | p.ia));
}
__device__ __forceinline__ spt ld_spt(const double *d, uint32_t cap, uint32_t slot, int k)
{
    const double *b = d + static_cast<size_t>(1 + 4 * k) * cap + slot;
    spt p;
    p.p = mk(b[0], b[cap], b[2 * static_cast<size_t>(cap)]);
    uint64_t f = static_cast<uint64_t>(__double_as_longlong(b[3 * static_cast<size_t>(cap)]));
    p.ia = static_cast<uint32_t>(f);
    p.ib = static_cast<uint32_t>(f >> 32);
    return p;
}
__device__ __forceinline__ void save_simplex(double *d, uint32_t cap, uint32_t slot, const simplex &x)
{
    d[slot] = static_cast<double>(x.n);
    st_spt(d, cap, slot, 0, x.s0);
    if (x.n > 1) st_spt(d, cap, slot, 1, x.s1);
    if (x.n > 2) st_spt(d, cap, slot, 2, x.s2);
    if (x.n > 3) st_spt(d, cap, slot, 3, x.s3);
}
__device__ __forceinline__ void load_simplex(const double *d, uint32_t cap, uint32_t slot, simplex &x)
{
    x.n = static_cast<int>(d[slot]);
    x.s0 = ld_spt(d, cap, slot, 0);
    if (x.n > 1) x.s1 = ld_spt(d, cap, slot, 1);
    if (x.n > 2) x.s2 = ld_spt(d, cap, slot, 2);
    if (x.n > 3) x.s3 = ld_spt(d, cap, slot, 3);
}

// A lane takes Q consecutive pairs; the pairs whose bodies are not apart (above) are
// listed for k_gjk, which writes their status. With per-body records (many pairs per
// body, most culled by the FP32 spheres) Q is 4: one 32-byte key load, vector stores,
// and the sphere loads in flight together; the cull then also writes every pair's
// next manifold empty (count 0, no event; an empty manifold's slot is never read) and
// lists the culled pairs that had contacts for the manifold update (cull_manifold_out),
// which rewrites those and the GJK list's manifolds. Nothing reads a culled pair's
// status there, so none is written. From the planar arrays (few pairs per body, each
// test reading several arrays) Q is 1, culled pairs are written separated (status 0)
// and the manifold update lists its own pairs (k_manifold_live): there a source and
// count gather in the cull would lengthen each lane's chain of dependent loads. A
// capped grid steps through the pairs, the same number of rounds on every lane.
template <bool REC>
__global__ void __launch_bounds__(kBlock)
    k_gjk_cull(const uint64_t *keys, uint32_t n_pairs, bodies_dev b, shapes_dev s, const double *rec,
               const float4 *rec32, uint8_t *status, uint32_t *list, cull_manifold_out mo,
               unsigned long long *counters)
{
    constexpr uint32_t Q = REC ? 4 : 1;
    const uint64_t n_groups = (static_cast<uint64_t>(n_pairs) + Q - 1) / Q;
    const uint64_t threads = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t base = 0; base < n_groups; base += threads)
    {
        const uint64_t q = base + static_cast<uint64_t>(xcd_block(n_groups, blockDim.x)) * blockDim.x + threadIdx.x;
        const uint32_t p0 = static_cast<uint32_t>(Q * q);
        uint32_t n_here = 0, keepmask = 0, prevmask = 0;
        uint64_t key[Q] = {};
        if (q < n_groups)
        {
            n_here = min(Q, n_pairs - p0);
            int pc[Q] = {};
            if constexpr (Q == 4)
            {
                if (n_here == 4)
                {
                    const ulonglong2 k01 = *reinterpret_cast<const ulonglong2 *>(keys + p0);
                    const ulonglong2 k23 = *reinterpret_cast<const ulonglong2 *>(keys + p0 + 2);
                    key[0] = k01.x;
                    key[1] = k01.y;
                    key[2] = k23.x;
                    key[3] = k23.y;
                    if (mo.src)
                    {
                        const uint4 s4 = *reinterpret_cast<const uint4 *>(mo.src + p0);
                        const uint32_t sp[4] = {s4.x, s4.y, s4.z, s4.w};
#pragma unroll
                        for (int k = 0; k < 4; ++k) pc[k] = sp[k] != kNone ? mo.prev_count[sp[k]] : 0;
                    }
                    else
                    {
                        // the previous list kept whole: the previous counts at the same indices
                        const uint32_t c4 = *reinterpret_cast<const uint32_t *>(mo.prev_count + p0);
#pragma unroll
                        for (int k = 0; k < 4; ++k) pc[k] = static_cast<int>((c4 >> (8 * k)) & 0xFFu);
                    }
                    *reinterpret_cast<uint32_t *>(mo.next.count + p0) = 0u;
                    *reinterpret_cast<uint32_t *>(mo.events + p0) = 0u;
                }
            }
            if (Q != 4 || n_here != 4)
                for (uint32_t k = 0; k < n_here; ++k)
                {
                    const uint32_t p = p0 + k;
                    key[k] = keys[p];
                    if constexpr (REC)
                    {
                        const uint32_t sp = source_of(mo.src, p);
                        pc[k] = sp != kNone ? mo.prev_count[sp] : 0;
                        mo.next.count[p] = 0;
                        mo.events[p] = 0;
                    }
                    else
                        status[p] = 0;
                }
            for (uint32_t k = 0; k < Q; ++k)
                if (k < n_here && pc[k] > 0) prevmask |= 1u << k;
            for (uint32_t k = n_here; k < Q; ++k) key[k] = key[0]; // a valid pair, never listed
            if (REC && rec32)
            {
                // the sphere loads are issued before any test
                float4 sa[Q], sb[Q];
#pragma unroll
                for (uint32_t k = 0; k < Q; ++k)
                {
                    sa[k] = rec32[static_cast<uint32_t>(key[k] >> 32)];
                    sb[k] = rec32[static_cast<uint32_t>(key[k] & 0xFFFFFFFFu)];
                }
#pragma unroll
                for (uint32_t k = 0; k < Q; ++k)
                    if (k < n_here && !spheres32_apart(sa[k], sb[k])) keepmask |= 1u << k;
            }
            else
                keepmask = (1u << n_here) - 1u;
#pragma unroll
            for (uint32_t k = 0; k < Q; ++k)
                if (keepmask & (1u << k))
                {
                    const uint32_t ia = static_cast<uint32_t>(key[k] >> 32);
                    const uint32_t ib = static_cast<uint32_t>(key[k] & 0xFFFFFFFFu);
                    bool apart;
                    if (REC)
                        apart = records_apart(rec + static_cast<size_t>(kCullRec) * ia,
                                              rec + static_cast<size_t>(kCullRec) * ib);
                    else
                        apart = bodies_apart(cull_of(b, s, ia), cull_of(b, s, ib));
                    if (apart) keepmask &= ~(1u << k);
                }
        }
        const uint32_t livemask = prevmask & ~keepmask;
        const uint32_t nk = static_cast<uint32_t>(__popc(keepmask)), nl = static_cast<uint32_t>(__popc(livemask));
        // workgroup-wide appends only where something is listed; each barrier also
        // separates the appends' shared LDS from the one before
        if (__syncthreads_or(nk != 0))
        {
            unsigned long long at = block_reserve(&counters[C_GJK_LIST], nk);
            for (uint32_t k = 0; k < Q; ++k)
                if (keepmask & (1u << k)) list[at++] = p0 + k;
        }
        if (REC && __syncthreads_or(nl != 0))
        {
            unsigned long long at = block_reserve(&counters[C_MLIVE], nl);
            for (uint32_t k = 0; k < Q; ++k)
                if (livemask & (1u << k))
                {
                    if (at < mo.live_cap) mo.live_list[at] = p0 + k; // at most the previous non-empty manifolds
                    ++at;
                }
        }
    }
}

// GJK over the listed pairs, G lanes per pair: the lanes of a group read consecutive
// vertex pairs of the same shape and reduce the support search; the simplex logic runs
// identically in every lane of the group, and lane 0 writes the results. Fewer lanes
// suit shapes with few support candidates (boxes, certified UV spheres): less redundant
// simplex work per lane.
template <int G, bool UV>
__global__ void __launch_bounds__(kBlock, 3) // 3 waves per SIMD fit without spilling
    k_gjk(const uint64_t *keys, const uint32_t *list, int list_slot, bodies_dev b, shapes_dev s, gjk_out_dev g,
          unsigned long long *counters)
{
    const uint32_t n_list = static_cast<uint32_t>(counters[list_slot]);
    const uint32_t t = xcd_block(static_cast<uint64_t>(n_list) * G, blockDim.x) * blockDim.x + threadIdx.x;
    const int gl = static_cast<int>(t % G);
    // groups step through the list together: the same number of rounds on every lane,
    // so each round's workgroup-wide append sees all of them
    const uint64_t per_round = gridDim.x * blockDim.x / G;
    for (uint64_t base = 0; base < n_list; base += per_round)
    {
        const uint64_t i = base + t / G;
        bool hit = false;
        simplex x;
        x.n = 0;
        uint32_t p = 0;
        if (i < n_list)
        {
            p = list[i];
            uint64_t key = keys[p];
            gshape A = body_shape<UV>(b, s, static_cast<uint32_t>(key >> 32));
            gshape B = body_shape<UV>(b, s, static_cast<uint32_t>(key & 0xFFFFFFFFu));
            hit = gjk<G, UV>(A, B, x, gl);
            if (gl == 0) g.status[p] = 0;
        }
        bool lead = hit && gl == 0;
        unsigned long long slot = block_append(&counters[C_HITS], lead);
        if (lead)
        {
            g.hit_of[p] = static_cast<uint32_t>(slot);
            // past the hit list's capacity the hit is only counted: the host sees the
            // count, grows the list and runs the GJK stage again
            if (slot < g.cap)
            {
                g.hit_list[slot] = p;
                save_simplex(g.simplex, g.cap, static_cast<uint32_t>(slot), x);
            }
        }
        __syncthreads(); // the next round's append reuses the same LDS
    }
}

__device__ __forceinline__ void write_info(double *info, uint32_t cap, uint32_t p, const double *o)
{
    for (int k = 0; k < 10; ++k) info[static_cast<size_t>(k) * cap + p] = o[k];
}

// one thread per hit, 192-face store in private memory; a solve that needs more is
// listed for the 4096-face kernel below
template <bool UV, int BS>
__global__ void __launch_bounds__(BS)
    k_epa(const uint64_t *keys, uint32_t n_hits, bodies_dev b, shapes_dev s, gjk_out_dev g,
          uint32_t *ovf_list, unsigned long long *counters)
{
    uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    bool ovf = false;
    if (h < n_hits)
    {
        uint32_t p = g.hit_list[h];
        uint64_t key = keys[p];
        gshape A = body_shape<UV>(b, s, static_cast<uint32_t>(key >> 32));
        gshape B = body_shape<UV>(b, s, static_cast<uint32_t>(key & 0xFFFFFFFFu));
        simplex x;
        load_simplex(g.simplex, g.cap, h, x);
        epa_store_fast store;
        double out[10];
        int r = epa_solve<UV>(A, B, x, store, out);
        if (r == 1) write_info(g.info, g.cap, h, out);
        g.status[p] = static_cast<uint8_t>(r == 1 ? 1 : (r == 0 ? 2 : 3));
        ovf = r == 3;
    }
    unsigned long long slot = wave_append(&counters[C_EPA_OVF], ovf);
    if (ovf) ovf_list[slot] = h;
}

// Second tier: one workgroup per hit that overflowed the private store, its polytope in
// LDS (896 faces): LDS round trips instead of global-memory ones on the solve's long
// chain of dependent accesses. A solve that needs more goes on to k_epa_big.
template <class S, bool UV>
__global__ void __launch_bounds__(64)
    k_epa_lds(const uint64_t *keys, const uint32_t *ovf_list, uint32_t n_ovf, bodies_dev b, shapes_dev s,
              gjk_out_dev g, uint32_t *ovf2_list, unsigned long long *counters, int out_slot)
{
    __shared__ S store;
    const uint32_t t = blockIdx.x;
    if (t >= n_ovf || threadIdx.x != 0) return;
    const uint32_t h = ovf_list ? ovf_list[t] : t;
    const uint32_t p = g.hit_list[h];
    const uint64_t key = keys[p];
    gshape A = body_shape<UV>(b, s, static_cast<uint32_t>(key >> 32));
    gshape B = body_shape<UV>(b, s, static_cast<uint32_t>(key & 0xFFFFFFFFu));
    simplex x;
    load_simplex(g.simplex, g.cap, h, x);
    double out[10];
    const int r = epa_solve<UV>(A, B, x, store, out);
    if (r == 3)
    {
        ovf2_list[atomicAdd(&counters[out_slot], 1ull)] = h;
        return;
    }
    if (r == 1) write_info(g.info, g.cap, h, out);
    g.status[p] = static_cast<uint8_t>(r == 1 ? 1 : 2);
}

template <bool UV>
__global__ void k_epa_big(const uint64_t *keys, const uint32_t *ovf_list, uint32_t n_ovf, bodies_dev b,
                          shapes_dev s, gjk_out_dev g, epa_store_big *pool, unsigned long long *counters)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_ovf) return;
    uint32_t h = ovf_list[t];
    uint32_t p = g.hit_list[h];
    uint64_t key = keys[p];
    gshape A = body_shape<UV>(b, s, static_cast<uint32_t>(key >> 32));
    gshape B = body_shape<UV>(b, s, static_cast<uint32_t>(key & 0xFFFFFFFFu));
    simplex x;
    load_simplex(g.simplex, g.cap, h, x);
    double out[10];
    int r = epa_solve<UV>(A, B, x, pool[t], out);
    if (r == 1) write_info(g.info, g.cap, h, out);
    g.status[p] = static_cast<uint8_t>(r == 1 ? 1 : (r == 0 ? 2 : 3));
    if (r == 3) atomicAdd(&counters[C_EPA_BIG_FAIL], 1ull);
}

// ---- manifolds (collision_phases.h:75-198, 244-320) ----
struct contact
{
    v3 n, la, lb;
    double depth, ln, lt1, lt2;
};

// contact c of the manifold in data slot sl
__device__ __forceinline__ contact ld_contact(const manifolds_dev &m, uint32_t sl, int c)
{
    const double *d = m.data + static_cast<size_t>(c * kMF) * m.dcap + sl;
    size_t s = m.dcap;
    contact k;
    k.n = mk(d[0], d[s], d[2 * s]);
    k.la = mk(d[3 * s], d[4 * s], d[5 * s]);
    k.lb = mk(d[6 * s], d[7 * s], d[8 * s]);
    k.depth = d[9 * s];
    k.ln = d[10 * s];
    k.lt1 = d[11 * s];
    k.lt2 = d[12 * s];
    return k;
}
__device__ __forceinline__ void st_contact(const manifolds_dev &m, uint32_t sl, int c, const contact &k)
{
    double *d = m.data + static_cast<size_t>(c * kMF) * m.dcap + sl;
    size_t s = m.dcap;
    d[0] = k.n.x;
    d[s] = k.n.y;
    d[2 * s] = k.n.z;
    d[3 * s] = k.la.x;
    d[4 * s] = k.la.y;
    d[5 * s] = k.la.z;
    d[6 * s] = k.lb.x;
    d[7 * s] = k.lb.y;
    d[8 * s] = k.lb.z;
    d[9 * s] = k.depth;
    d[10 * s] = k.ln;
    d[11 * s] = k.lt1;
    d[12 * s] = k.lt2;
}

struct mani
{
    contact c[4];
    int n;
};

__device__ void add_reduce(mani &m, const contact &k) // collision_phases.h:139-198
{
    contact pool[5] = {m.c[0], m.c[1], m.c[2], m.c[3], k};
    int best[4] = {0, 1, 2, 3};
    for (int i = 1; i < 5; ++i)
        if (pool[i].depth > pool[best[0]].depth) best[0] = i;
    double md = -1.0;
    for (int i = 0; i < 5; ++i)
    {
        if (i == best[0]) continue;
        double d2 = sqn(pool[i].la - pool[best[0]].la);
        if (d2 > md)
        {
            md = d2;
            best[1] = i;
        }
    }
    double ma = -1.0;
    v3 e0 = pool[best[1]].la - pool[best[0]].la;
    for (int i = 0; i < 5; ++i)
    {
        if (i == best[0] || i == best[1]) continue;
        v3 e1 = pool[i].la - pool[best[0]].la;
        double a2 = sqn(cross(e0, e1));
        if (a2 > ma)
        {
            ma = a2;
            best[2] = i;
        }
    }
    md = -1.0;
    for (int i = 0; i < 5; ++i)
    {
        if (i == best[0] || i == best[1] || i == best[2]) continue;
        double d2 = sqn(pool[i].la - pool[best[2]].la);
        if (d2 > md)
        {
            md = d2;
            best[3] = i;
        }
    }
    for (int i = 0; i < 4; ++i) m.c[i] = pool[best[i]];
    m.n = 4;
}

__device__ __forceinline__ void add_contact(mani &m, const contact &k) // :124-130
{
    if (m.n < 4)
        m.c[m.n++] = k;
    else
        add_reduce(m, k);
}

// Where the cull read the planar arrays: every pair, four consecutive ones per lane,
// written empty (no contacts, no event; vector stores), and those with a GJK hit or
// contacts before listed for k_manifold, which writes them again. A pair with neither stays empty and
// no body state is read for it. A capped grid steps through the pairs, the same number
// of rounds on every lane.
__global__ void __launch_bounds__(kBlock)
    k_manifold_live(const uint32_t *src, uint32_t n_pairs, const uint8_t *status, const uint8_t *prev_count,
                    manifolds_dev next, uint8_t *events, uint32_t *live_list, unsigned long long *counters)
{
    const uint64_t n_quads = (static_cast<uint64_t>(n_pairs) + 3) / 4;
    const uint64_t threads = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t base = 0; base < n_quads; base += threads)
    {
        const uint64_t q = base + static_cast<uint64_t>(xcd_block(n_quads, blockDim.x)) * blockDim.x + threadIdx.x;
        const uint32_t p0 = static_cast<uint32_t>(4 * q);
        uint32_t livemask = 0;
        if (q < n_quads)
        {
            const uint32_t n_here = min(4u, n_pairs - p0);
            if (n_here == 4)
            {
                const uint32_t st4 = *reinterpret_cast<const uint32_t *>(status + p0);
                int pc[4];
                if (src)
                {
                    const uint4 s4 = *reinterpret_cast<const uint4 *>(src + p0);
                    const uint32_t sp[4] = {s4.x, s4.y, s4.z, s4.w};
#pragma unroll
                    for (int k = 0; k < 4; ++k) pc[k] = sp[k] != kNone ? prev_count[sp[k]] : 0;
                }
                else
                {
                    // the previous list kept whole: the previous counts at the same indices
                    const uint32_t c4 = *reinterpret_cast<const uint32_t *>(prev_count + p0);
#pragma unroll
                    for (int k = 0; k < 4; ++k) pc[k] = static_cast<int>((c4 >> (8 * k)) & 0xFFu);
                }
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (((st4 >> (8 * k)) & 0xFFu) == 1u || pc[k] > 0) livemask |= 1u << k;
                *reinterpret_cast<uint32_t *>(next.count + p0) = 0u;
                *reinterpret_cast<uint32_t *>(events + p0) = 0u;
            }
            else
                for (uint32_t k = 0; k < n_here; ++k)
                {
                    const uint32_t p = p0 + k, sp = source_of(src, p);
                    if (status[p] == 1 || (sp != kNone && prev_count[sp] > 0)) livemask |= 1u << k;
                    next.count[p] = 0;
                    events[p] = 0;
                }
        }
        const uint32_t nl = static_cast<uint32_t>(__popc(livemask));
        // a workgroup with no live pair (most of them where pairs far outnumber contacts)
        // skips the append; the barrier also separates the rounds' appends
        if (!__syncthreads_or(nl != 0)) continue;
        unsigned long long at = block_reserve(&counters[C_MLIVE], nl);
        for (int k = 0; k < 4; ++k)
            if (livemask & (1u << k)) live_list[at++] = p0 + static_cast<uint32_t>(k);
    }
}

// The listed pairs' manifolds (collision_phases.h:197-290): live_list (the culled pairs
// that had contacts, or every pair with a hit or contacts before), then gjk_list (the
// pairs GJK ran on; null when live_list holds them)
__global__ void __launch_bounds__(kBlock)
    k_manifold(const uint64_t *keys, const uint32_t *src, const uint32_t *live_list, const uint32_t *gjk_list,
               bodies_dev b, gjk_out_dev g, manifolds_dev prev, manifolds_dev next, uint8_t *events, int epa_guard,
               unsigned long long *counters)
{
    const uint64_t n1 = counters[C_MLIVE], n_live = n1 + (gjk_list ? counters[C_GJK_LIST] : 0ull);
    // the same number of rounds on every lane (each round's append and sum are workgroup-wide)
    const uint64_t threads = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t base = 0; base < n_live; base += threads)
    {
        const uint64_t i = base + static_cast<uint64_t>(xcd_block(n_live, blockDim.x)) * blockDim.x + threadIdx.x;
        const bool live = i < n_live;
        const uint32_t p = !live ? 0u : (i < n1 ? live_list[i] : gjk_list[i - n1]);
        uint32_t count = 0;
        mani nm;
        nm.n = 0;
        if (live)
        {
            const double distance2_eps = 0.005 * 0.005; // :212
            const double breaking = 0.05;               // :213
            const double drift2_eps = 0.06;             // :267
            uint64_t key = keys[p];
            uint32_t ia = static_cast<uint32_t>(key >> 32), ib = static_cast<uint32_t>(key & 0xFFFFFFFFu);
            v3 pa = ld3(b.pos, b.n, ia), pb = ld3(b.pos, b.n, ib);
            qt qa = ldq(b.quat, b.n, ia), qb = ldq(b.quat, b.n, ib);
            // the cull's own entries (first n1 with a GJK list) were culled: no new contact
            bool has_new = !(gjk_list && i < n1) && g.status[p] == 1;
            contact nc{};
            if (has_new)
            {
                const double *inf = g.info + g.hit_of[p];
                size_t s = g.cap;
                nc.n = mk(inf[0], inf[s], inf[2 * s]);
                v3 wa = mk(inf[3 * s], inf[4 * s], inf[5 * s]);
                v3 wb = mk(inf[6 * s], inf[7 * s], inf[8 * s]);
                nc.la = rotate(conj(qa), wa - pa); // particle.h:107-108
                nc.lb = rotate(conj(qb), wb - pb);
                nc.depth = inf[9 * s];
                nc.ln = 0.0;
                nc.lt1 = 0.0;
                nc.lt2 = 0.0;
                if (epa_guard)
                {
                    bool ok = sqn(nc.n) > 0.0 && isfinite(wa.x) && isfinite(wa.y) && isfinite(wa.z) &&
                              isfinite(wb.x) && isfinite(wb.y) && isfinite(wb.z) && isfinite(nc.depth);
                    has_new = ok;
                }
            }
            const uint32_t sp = source_of(src, p);
            int oldn = sp != kNone ? prev.count[sp] : 0;
            const uint32_t osl = oldn > 0 ? prev.slot[sp] : kNone;
            for (int i = 0; i < oldn; ++i)
            {
                contact old = ld_contact(prev, osl, i);
                if (has_new)
                {
                    if (sqn(nc.la - old.la) < distance2_eps || sqn(nc.lb - old.lb) < distance2_eps)
                    {
                        nc.ln = old.ln;
                        nc.lt1 = old.lt1;
                        nc.lt2 = old.lt2;
                        continue;
                    }
                }
                v3 wa = rotate(qa, old.la) + pa; // particle.h:110-111
                v3 wb = rotate(qb, old.lb) + pb;
                v3 nrm = has_new ? nc.n : old.n;
                v3 rel = wb - wa;
                double depth = dot(rel, nrm);
                v3 pja = wa + nrm * depth;
                double drift2 = sqn(pja - wb);
                if (depth > -breaking && drift2 < drift2_eps)
                {
                    old.depth = depth;
                    old.n = nrm;
                    add_contact(nm, old);
                }
            }
            if (has_new) add_contact(nm, nc);
            next.count[p] = static_cast<uint8_t>(nm.n);
            bool was = oldn > 0, is = nm.n > 0;
            events[p] = (!was && is) ? 1 : ((was && !is) ? 2 : 0);
            count = static_cast<uint32_t>(nm.n);
        }
        // a data slot for each non-empty manifold (the host sized the store for every pair
        // that had contacts or a GJK hit, so the slots fit)
        const unsigned long long sl = block_append(&counters[C_MSLOTS], count > 0);
        if (live)
        {
            next.slot[p] = count > 0 ? static_cast<uint32_t>(sl) : kNone;
            if (count > 0) next.owner[sl] = p;
            for (int i = 0; i < nm.n; ++i) st_contact(next, static_cast<uint32_t>(sl), i, nm.c[i]);
        }
        block_add(&counters[C_CONTACTS], count);
        __syncthreads(); // the next round's append and sum reuse the same LDS
    }
}

// ---- standalone batch ----
__device__ gshape pair_shape(const int32_t *kind, const double *prm, const uint32_t *vb, const uint32_t *ve,
                             const double *verts, const uint32_t *uv, uint32_t i)
{
    const double *q = prm + 10 * static_cast<size_t>(i);
    gshape g;
    g.kind = kind[i];
    g.v = nullptr;
    g.nv = 0;
    g.uv = 0;
    g.paired = 0;
    g.b = mk(0.0, 0.0, 0.0);
    g.q = mkq(1.0, 0.0, 0.0, 0.0);
    if (g.kind == 0)
    {
        g.a = mk(q[0], q[1], q[2]);
        g.b = mk(q[3], q[4], q[5]);
    }
    else if (g.kind == 1)
    {
        g.a = mk(q[0], q[1], q[2]);
        g.q = mkq(q[3], q[4], q[5], q[6]);
        g.b = mk(q[7], q[8], q[9]);
    }
    else
    {
        g.a = mk(q[0], q[1], q[2]);
        g.q = mkq(q[3], q[4], q[5], q[6]);
        g.v = verts + 3 * static_cast<size_t>(vb[i]);
        g.nv = ve[i] - vb[i];
        g.uv = uv ? uv[i] : 0u; // no tags: every support scans all vertices (same results)
    }
    return g;
}

template <class S>
__device__ void batch_one(const shape_pair_dev &sp, uint32_t i, double *out, int32_t *hit, S &store)
{
    gshape A = pair_shape(sp.kind_a, sp.prm_a, sp.vbeg_a, sp.vend_a, sp.verts, sp.uv_a, i);
    gshape B = pair_shape(sp.kind_b, sp.prm_b, sp.vbeg_b, sp.vend_b, sp.verts, sp.uv_b, i);
    simplex x;
    if (!gjk<1>(A, B, x, 0))
    {
        hit[i] = 0;
        return;
    }
    double o[10];
    int r = epa_solve(A, B, x, store, o);
    if (r == 1)
        for (int k = 0; k < 10; ++k) out[10 * static_cast<size_t>(i) + k] = o[k];
    hit[i] = r == 1 ? 1 : (r == 0 ? 2 : 3);
}

__global__ void __launch_bounds__(kEpaBlock) k_batch(shape_pair_dev sp, uint32_t n, double *out, int32_t *hit)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    epa_store_fast store;
    batch_one(sp, i, out, hit, store);
}

__global__ void k_batch_big(shape_pair_dev sp, uint32_t n, double *out, int32_t *hit,
                            epa_store_big *pool)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 1) return;
    // single thread walks the overflowed pairs one by one through the large store
    for (uint32_t i = 0; i < n; ++i)
        if (hit[i] == 3) batch_one(sp, i, out, hit, pool[0]);
}
// ---- standalone batch over device arrays with caller-free scratch ----
// Scratch words: [0] pairs with invalid input, [1] pairs listed for the large store,
// [2] pairs that overflowed the large store as well.
__global__ void k_batch_dev_clear(uint32_t *words)
{
    if (threadIdx.x < 4) words[threadIdx.x] = 0u;
}
__global__ void __launch_bounds__(kEpaBlock)
    k_batch_dev(shape_pair_dev sp, uint32_t n, uint32_t n_verts, double *out, int32_t *hit, uint32_t *words,
                uint32_t *ovf_list)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // the host variant checks these on the host (pkg_gjk_epa_batch)
    for (int side = 0; side < 2; ++side)
    {
        const int k = side ? sp.kind_b[i] : sp.kind_a[i];
        const uint32_t vb = side ? sp.vbeg_b[i] : sp.vbeg_a[i], ve = side ? sp.vend_b[i] : sp.vend_a[i];
        if (k < 0 || k > 2 || (k == 2 && (ve <= vb || ve > n_verts)))
        {
            hit[i] = -1;
            atomicAdd(&words[0], 1u);
            return;
        }
    }
    epa_store_fast store;
    batch_one(sp, i, out, hit, store);
    if (hit[i] == 3) ovf_list[atomicAdd(&words[1], 1u)] = i;
}
// one thread per large store, each taking every n_stores-th listed pair
__global__ void k_batch_dev_big(shape_pair_dev sp, double *out, int32_t *hit, uint32_t *words,
                                const uint32_t *ovf_list, epa_store_big *stores, uint32_t n_stores)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_stores) return;
    const uint32_t m = words[1];
    for (uint32_t k = t; k < m; k += n_stores)
    {
        const uint32_t i = ovf_list[k];
        batch_one(sp, i, out, hit, stores[t]);
        if (hit[i] == 3) atomicAdd(&words[2], 1u);
    }
}
} // namespace

size_t epa_big_bytes() { return sizeof(epa_store_big); }
bool epa_first_tier_is_lds(uint32_t n_hits, const tuning &tu) { return n_hits <= tu.epa_lds_max; }

void launch_gjk_epa_batch_dev(const shape_pair_dev &sp, uint32_t n, uint32_t n_verts, double *out, int32_t *hit,
                              uint32_t *words, uint32_t *ovf_list, void *stores, uint32_t n_stores, hipStream_t st)
{
    k_batch_dev_clear<<<1, 64, 0, st>>>(words);
    if (!n) return;
    k_batch_dev<<<grid(n, kEpaBlock), kEpaBlock, 0, st>>>(sp, n, n_verts, out, hit, words, ovf_list);
    k_batch_dev_big<<<1, 64, 0, st>>>(sp, out, hit, words, ovf_list, static_cast<epa_store_big *>(stores), n_stores);
}

void launch_gjk(const uint64_t *keys, uint32_t n_pairs, const bodies_dev &b, const shapes_dev &s, bool uv, int group,
                gjk_out_dev &g, uint32_t *list, double *cull_rec, const gjk32_in &f32, const cull_manifold_out &mo,
                unsigned long long *counters, hipStream_t st)
{
    if (!n_pairs) return;
    if (cull_uses_records(n_pairs, b.n))
    {
        // the mixed form adds the FP32 sphere records after the FP64 ones (16-byte aligned)
        float4 *rec32 = f32.s32 ? reinterpret_cast<float4 *>(cull_rec + static_cast<size_t>(kCullRec) * b.n) : nullptr;
        k_cull_prep<<<grid(b.n), kBlock, 0, st>>>(b, s, cull_rec, rec32);
        k_gjk_cull<true><<<list_grid((n_pairs + 3) / 4), kBlock, 0, st>>>(
            keys, n_pairs, b, s, cull_rec, rec32, g.status, list, mo, counters);
    }
    else
        k_gjk_cull<false><<<list_grid(n_pairs), kBlock, 0, st>>>(
            keys, n_pairs, b, s, cull_rec, nullptr, g.status, list, mo, counters);
    int slot = C_GJK_LIST;
    if (f32.s32)
    {
        launch_gjk32(keys, list, n_pairs, b, *f32.s32, uv, f32.src, f32.prev_count, g.status, f32.list2, counters, st);
        list = f32.list2;
        slot = C_GJK_LIST2;
    }
    // sized for every pair up to kListGrid workgroups, which then step through the list
    // (its count is read on the device); groups past the listed count exit at once
    const unsigned gr = list_grid(static_cast<size_t>(n_pairs) * static_cast<size_t>(group));
    auto go = [&](auto lanes)
    {
        constexpr int G = decltype(lanes)::value;
        if (uv)
            k_gjk<G, true><<<gr, kBlock, 0, st>>>(keys, list, slot, b, s, g, counters);
        else
            k_gjk<G, false><<<gr, kBlock, 0, st>>>(keys, list, slot, b, s, g, counters);
    };
    switch (group)
    {
    case 1: go(std::integral_constant<int, 1>{}); break;
    case 2: go(std::integral_constant<int, 2>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    default: go(std::integral_constant<int, 4>{}); break;
    }
}

void launch_epa(const uint64_t *keys, uint32_t n_hits, const bodies_dev &b, const shapes_dev &s, bool uv,
                const tuning &tu, gjk_out_dev &g, uint32_t *ovf_list, unsigned long long *counters, hipStream_t st)
{
    if (!n_hits) return;
    if (epa_first_tier_is_lds(n_hits, tu))
    {
        if (uv)
            k_epa_lds<epa_store_lds, true><<<n_hits, 64, 0, st>>>(keys, nullptr, n_hits, b, s, g, ovf_list, counters,
                                                                 C_EPA_OVF);
        else
            k_epa_lds<epa_store_lds, false><<<n_hits, 64, 0, st>>>(keys, nullptr, n_hits, b, s, g, ovf_list,
                                                                  counters, C_EPA_OVF);
        return;
    }
    auto go = [&](auto block)
    {
        constexpr int BS = decltype(block)::value;
        if (uv)
            k_epa<true, BS><<<grid(n_hits, BS), BS, 0, st>>>(keys, n_hits, b, s, g, ovf_list, counters);
        else
            k_epa<false, BS><<<grid(n_hits, BS), BS, 0, st>>>(keys, n_hits, b, s, g, ovf_list, counters);
    };
    switch (tu.epa_block)
    {
    case 128: go(std::integral_constant<int, 128>{}); break;
    case 256: go(std::integral_constant<int, 256>{}); break;
    default: go(std::integral_constant<int, 64>{}); break;
    }
}

void launch_epa_lds(const uint64_t *keys, const uint32_t *ovf_list, uint32_t n_ovf, const bodies_dev &b,
                    const shapes_dev &s, const gjk_out_dev &g, uint32_t *ovf2_list, unsigned long long *counters,
                    hipStream_t st)
{
    if (n_ovf)
        k_epa_lds<epa_store_lds, true><<<n_ovf, 64, 0, st>>>(keys, ovf_list, n_ovf, b, s, g, ovf2_list, counters,
                                                            C_EPA_OVF2);
}

void launch_epa_big(const uint64_t *keys, const uint32_t *ovf_list, uint32_t n_ovf,
                    const bodies_dev &b, const shapes_dev &s, gjk_out_dev &g, void *pool,
                    unsigned long long *counters, hipStream_t st)
{
    if (n_ovf)
        k_epa_big<true><<<grid(n_ovf, 64), 64, 0, st>>>(keys, ovf_list, n_ovf, b, s, g,
                                                        static_cast<epa_store_big *>(pool), counters);
}

void launch_manifold_update(const uint64_t *keys, const uint32_t *src, uint32_t n_pairs, uint32_t bound,
                            const bodies_dev &b, const gjk_out_dev &g, manifolds_dev prev, manifolds_dev next,
                            uint8_t *events, bool epa_guard, bool culled, uint32_t *live_list,
                            const uint32_t *gjk_list, unsigned long long *counters, hipStream_t st)
{
    if (!n_pairs) return;
    if (!culled)
        k_manifold_live<<<list_grid((n_pairs + 3) / 4), kBlock, 0, st>>>(src, n_pairs, g.status, prev.count, next,
                                                                         events, live_list, counters);
    if (bound)
        k_manifold<<<list_grid(bound), kBlock, 0, st>>>(keys, src, live_list, culled ? gjk_list : nullptr, b, g,
                                                        prev, next, events, epa_guard ? 1 : 0, counters);
}

bool cull_uses_records(uint32_t n_pairs, uint32_t n_bodies)
{
    // per-body records pay off with several pairs per body (hull clusters: 4; the
    // batched worlds, 0.6, read the planar arrays faster)
    return static_cast<size_t>(n_pairs) >= 2 * static_cast<size_t>(n_bodies);
}

void launch_gjk_epa_batch(const shape_pair_dev &sp, uint32_t n, double *out, int32_t *hit,
                          void *big_pool, hipStream_t st)
{
    if (!n) return;
    k_batch<<<grid(n, kEpaBlock), kEpaBlock, 0, st>>>(sp, n, out, hit);
    if (big_pool) k_batch_big<<<1, 64, 0, st>>>(sp, n, out, hit, static_cast<epa_store_big *>(big_pool));
}

} // namespace pk
