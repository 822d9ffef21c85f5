// FP32 certified separation ahead of the FP64 GJK (mixed-precision narrowphase).
//
// gjk_collision (src/collision.cpp:165-189) returns false for two shapes whose
// distance exceeds the margin m = 1e-6 m + 1e-6 (r_a + r_b) (the argument of the
// bounding-volume cull in k_narrow.hip: no Minkowski point can pass the first-support
// test, and no simplex case reports enclosure with a gap this far above the rounding
// of its normals). This pass proves such a distance in FP32 and writes the reference's
// verdict (separated) without running the FP64 GJK; every pair it cannot prove goes
// on to the FP64 GJK, so verdicts are the reference's bit for bit.
//
// The proof: for any vector n, the shapes are at least
//   gap(n) = (dx . n - h_A(-n) - h_B(n)) / |n|
// apart, with dx = x_A - x_B and h_S the support function of shape S about its
// origin. An FP32 GJK distance iteration (Gilbert-Johnson-Keerthi closest point of the
// Minkowski difference to the origin, simplex reduction after Ericson, Real-Time
// Collision Detection 5.1) supplies good directions n; at each one, gap(n) is
// evaluated in FP32 and accepted only above m plus a bound on its rounding error:
//   |gap_fp32 - gap| <= 28 u (R_A + R_B + |dx|) |n|,   u = 2^-24,
// from the FP32 rounding of the vertices (u R), of the quaternion and the rotation of
// n into each body frame (about 20 u R), of the dot products (5 u R), of dx (5 u |dx|)
// and of the final sum. The test uses 256 u. A tagged UV-sphere mesh is bounded by its
// circumscribed sphere, h(n) <= R |n| (R rounded up to FP32). Non-unit quaternions and
// non-finite values are never certified (comparisons with NaN are false).
#include "pk_access.hpp"
#include "pk_gjk32.hpp"

namespace pk
{
namespace
{
using namespace g32;
constexpr int kBlock = 256;
__device__ __forceinline__ body32 body_of(const bodies_dev &b, const shapes32_dev &s, uint32_t id, bool uv,
                                          bool &unit)
{
    body32 o;
    const qt q = ldq(b.quat, b.n, id);
    const double qn = ((q.w * q.w + q.x * q.x) + q.y * q.y) + q.z * q.z;
    unit = fabs(qn - 1.0) <= 1e-9;
    o.q = {static_cast<float>(q.w), static_cast<float>(q.x), static_cast<float>(q.y), static_cast<float>(q.z)};
    const uint32_t sh = b.shape[id];
    o.r = s.rad[sh];
    if (uv && s.uv[sh])
    {
        o.v = nullptr;
        o.nv = 0;
    }
    else
    {
        o.v = s.v + s.off[sh];
        o.nv = s.off[sh + 1] - s.off[sh];
    }
    return o;
}

// 1 when the pair is proven apart by more than the margin, else 0
template <bool UV> __device__ int certify_apart(const bodies_dev &b, const shapes32_dev &s, uint32_t ia, uint32_t ib)
{
    bool ua, ub;
    const body32 A = body_of(b, s, ia, UV, ua), B = body_of(b, s, ib, UV, ub);
    if (!(ua && ub)) return 0;
    const v3 dx64 = ld3(b.pos, b.n, ia) - ld3(b.pos, b.n, ib);
    const f3 dx = fm(static_cast<float>(dx64.x), static_cast<float>(dx64.y), static_cast<float>(dx64.z));
    return certify(A, B, dx);
}

template <bool UV>
__global__ void __launch_bounds__(kBlock)
    k_gjk32(const uint64_t *keys, const uint32_t *list, bodies_dev b, shapes32_dev s, const uint32_t *src,
            const uint8_t *prev_count, uint8_t *status, uint32_t *list2, unsigned long long *counters)
{
    const uint32_t n_list = static_cast<uint32_t>(counters[C_GJK_LIST]);
    const uint64_t threads = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    // the same number of rounds on every lane (each round's append is workgroup-wide)
    for (uint64_t base = 0; base < n_list; base += threads)
    {
        const uint64_t i = base + static_cast<uint64_t>(xcd_block(n_list, blockDim.x)) * blockDim.x + threadIdx.x;
        bool pass = false;
        uint32_t p = 0;
        if (i < n_list)
        {
            p = list[i];
            // a pair in contact after the previous step is most likely still touching: it
            // goes to the FP64 GJK without an attempt
            const uint32_t sp = source_of(src, p);
            const bool touched = sp != kNone && prev_count[sp] > 0;
            const uint64_t key = keys[p];
            if (!touched && certify_apart<UV>(b, s, static_cast<uint32_t>(key >> 32),
                                              static_cast<uint32_t>(key & 0xFFFFFFFFu)))
                status[p] = 0;
            else
                pass = true;
        }
        const unsigned long long slot = block_append(&counters[C_GJK_LIST2], pass);
        if (pass) list2[slot] = p;
        __syncthreads(); // the next round's append reuses the same LDS
    }
}

__global__ void k_verts32(const uint32_t *off, const uint32_t *voff, const double *verts, uint32_t n_shapes,
                          float4 *out)
{
    const uint32_t sh = blockIdx.x;
    if (sh >= n_shapes) return;
    const uint32_t c = off[sh + 1] - off[sh];
    for (uint32_t k = threadIdx.x; k < c; k += blockDim.x)
    {
        const double *v = verts + 3 * (static_cast<size_t>(voff[sh]) + k);
        out[off[sh] + k] = make_float4(static_cast<float>(v[0]), static_cast<float>(v[1]), static_cast<float>(v[2]), 0.0f);
    }
}
} // namespace

void launch_verts32(const shapes_dev &s, float4 *out, hipStream_t st)
{
    if (s.n) k_verts32<<<s.n, 64, 0, st>>>(s.off, s.voff, s.verts, s.n, out);
}

void launch_gjk32(const uint64_t *keys, const uint32_t *list, uint32_t n_pairs, const bodies_dev &b,
                  const shapes32_dev &s, bool uv, const uint32_t *src, const uint8_t *prev_count, uint8_t *status,
                  uint32_t *list2, unsigned long long *counters, hipStream_t st)
{
    if (!n_pairs) return;
    const unsigned gr = list_grid(n_pairs);
    if (uv)
        k_gjk32<true><<<gr, kBlock, 0, st>>>(keys, list, b, s, src, prev_count, status, list2, counters);
    else
        k_gjk32<false><<<gr, kBlock, 0, st>>>(keys, list, b, s, src, prev_count, status, list2, counters);
}

} // namespace pk
