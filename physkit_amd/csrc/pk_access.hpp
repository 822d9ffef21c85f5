// Planar (structure-of-arrays) loads and stores of vectors, quaternions and 3x3
// matrices, plus wavefront-aggregated counter increments.
#pragma once
#include "pk_internal.hpp"
#include "pk_math.hpp"

namespace pk
{

__device__ __forceinline__ v3 ld3(const double *p, uint32_t n, uint32_t i)
{ return mk(p[i], p[n + i], p[2 * n + i]); }
__device__ __forceinline__ void st3(double *p, uint32_t n, uint32_t i, v3 v)
{
    p[i] = v.x;
    p[n + i] = v.y;
    p[2 * n + i] = v.z;
}
__device__ __forceinline__ qt ldq(const double *p, uint32_t n, uint32_t i)
{ return mkq(p[i], p[n + i], p[2 * n + i], p[3 * n + i]); }
__device__ __forceinline__ void stq(double *p, uint32_t n, uint32_t i, qt q)
{
    p[i] = q.w;
    p[n + i] = q.x;
    p[2 * n + i] = q.y;
    p[3 * n + i] = q.z;
}
__device__ __forceinline__ m3 ld9(const double *p, uint32_t n, uint32_t i)
{
    m3 r;
#pragma unroll
    for (int k = 0; k < 9; ++k) r.m[k] = p[static_cast<size_t>(k) * n + i];
    return r;
}
__device__ __forceinline__ void st9(double *p, uint32_t n, uint32_t i, const m3 &a)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) p[static_cast<size_t>(k) * n + i] = a.m[k];
}
__device__ __forceinline__ box3 ldbox(const double *p, uint32_t n, uint32_t i)
{
    box3 b;
    b.min = mk(p[i], p[n + i], p[2 * n + i]);
    b.max = mk(p[3 * n + i], p[4 * n + i], p[5 * n + i]);
    return b;
}
__device__ __forceinline__ void stbox(double *p, uint32_t n, uint32_t i, const box3 &b)
{
    p[i] = b.min.x;
    p[n + i] = b.min.y;
    p[2 * n + i] = b.min.z;
    p[3 * n + i] = b.max.x;
    p[4 * n + i] = b.max.y;
    p[5 * n + i] = b.max.z;
}

// one atomic per wavefront: lanes with `pred` get consecutive slots; returns the
// slot of this lane (undefined when !pred)
__device__ __forceinline__ unsigned long long wave_append(unsigned long long *counter, bool pred)
{
    unsigned long long mask = __ballot(pred);
    if (mask == 0) return 0;
    int lane = __lane_id();
    int leader = __ffsll(static_cast<long long>(mask)) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, static_cast<unsigned long long>(__popcll(mask)));
    base = __shfl(base, leader);
    unsigned long long below = mask & ((1ull << lane) - 1ull);
    return base + static_cast<unsigned long long>(__popcll(below));
}

// adds the sum of v (0 <= v < 8) over the active lanes with one atomic per wavefront
__device__ __forceinline__ void wave_sum3(unsigned long long *counter, uint32_t v)
{
    unsigned long long mask = __ballot(true);
    unsigned long long s = static_cast<unsigned long long>(__popcll(__ballot((v & 1u) != 0))) +
                           2ull * static_cast<unsigned long long>(__popcll(__ballot((v & 2u) != 0))) +
                           4ull * static_cast<unsigned long long>(__popcll(__ballot((v & 4u) != 0)));
    int leader = __ffsll(static_cast<long long>(mask)) - 1;
    if (__lane_id() == leader && s) atomicAdd(counter, s);
}

} // namespace pk
