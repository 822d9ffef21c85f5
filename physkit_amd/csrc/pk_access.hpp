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
// a body's fat box: one 48-byte record per body (min xyz, max xyz), so a read at a
// body id (the tree's leaf gather, in Morton order) touches one or two lines, not six
__device__ __forceinline__ box3 ld_fat(const double *p, uint32_t i)
{
    const double2 *r = reinterpret_cast<const double2 *>(p + 6 * static_cast<size_t>(i));
    const double2 a = r[0], c = r[1], d = r[2];
    box3 b;
    b.min = mk(a.x, a.y, c.x);
    b.max = mk(c.y, d.x, d.y);
    return b;
}
__device__ __forceinline__ void st_fat(double *p, uint32_t i, const box3 &b)
{
    double2 *r = reinterpret_cast<double2 *>(p + 6 * static_cast<size_t>(i));
    r[0] = make_double2(b.min.x, b.min.y);
    r[1] = make_double2(b.min.z, b.max.x);
    r[2] = make_double2(b.max.y, b.max.z);
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

// Workgroup-aggregated counters. Device-scope atomics on one address serialise across
// the device at about 12 ns each on MI355X (65,536 wavefronts adding once each to one
// counter take 0.79 ms; one add per 256-thread workgroup, 0.20 ms), so a kernel that
// adds once per wavefront is bounded by 12 ns x its wavefronts. These add once per
// workgroup. Every thread of the workgroup must call them (none may have returned),
// the workgroup size must be a multiple of 64 and at most 1024. Their LDS is static: a
// kernel that calls one again (a loop over rounds) needs a workgroup barrier between
// the calls, after every thread has returned from the earlier one.

// exclusive offset of this thread's v slots in a range reserved at *counter
__device__ __forceinline__ unsigned long long block_reserve(unsigned long long *counter, uint32_t v)
{
    __shared__ uint32_t wsum[16];
    __shared__ unsigned long long wbase;
    const int lane = static_cast<int>(__lane_id());
    const uint32_t wave = threadIdx.x >> 6;
    uint32_t incl = v;
    for (int off = 1; off < 64; off <<= 1)
    {
        uint32_t o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long t = 0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) t += wsum[w];
        wbase = t ? atomicAdd(counter, t) : 0ull;
    }
    __syncthreads();
    unsigned long long base = wbase;
    for (uint32_t w = 0; w < wave; ++w) base += wsum[w];
    return base + (incl - v);
}

// XCD-aware order of a list's blocks: MI355X hands workgroups to its 8 XCDs in turn
// (workgroup b to XCD b mod 8), so workgroup b takes list block (b mod 8) * per + b / 8,
// per = ceil(active / 8), where active = the workgroups a round of the list fills (at
// most the grid). Each XCD's L2 then serves one contiguous eighth of the list. A
// workgroup past the list gets block `active` (no items). The grid must be a multiple
// of 8; when the list fills every workgroup the order is a permutation of the grid.
__device__ __forceinline__ uint32_t xcd_block(uint64_t n_items, uint32_t items_per_wg)
{
    const uint64_t need = (n_items + items_per_wg - 1) / items_per_wg;
    const uint32_t active = need < gridDim.x ? static_cast<uint32_t>(need) : gridDim.x;
    const uint32_t per = (active + 7) / 8, xcd = blockIdx.x % 8, local = blockIdx.x / 8;
    return local < per ? xcd * per + local : active;
}

// lanes with `pred` get consecutive slots; returns this lane's (undefined when !pred)
__device__ __forceinline__ unsigned long long block_append(unsigned long long *counter, bool pred)
{
    __shared__ uint32_t wcnt[16];
    __shared__ unsigned long long wbase;
    const unsigned long long mask = __ballot(pred);
    const uint32_t wave = threadIdx.x >> 6;
    if (__lane_id() == 0) wcnt[wave] = static_cast<uint32_t>(__popcll(mask));
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long t = 0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) t += wcnt[w];
        wbase = t ? atomicAdd(counter, t) : 0ull;
    }
    __syncthreads();
    unsigned long long base = wbase;
    for (uint32_t w = 0; w < wave; ++w) base += wcnt[w];
    return base + static_cast<unsigned long long>(__popcll(mask & ((1ull << __lane_id()) - 1ull)));
}

// adds the workgroup's sum of v
__device__ __forceinline__ void block_add(unsigned long long *counter, uint32_t v)
{
    __shared__ uint32_t wsum[16];
    uint32_t sum = v;
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if (__lane_id() == 0) wsum[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long t = 0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) t += wsum[w];
        if (t) atomicAdd(counter, t);
    }
}

// raises *counter to the workgroup's maximum of v
__device__ __forceinline__ void block_max(unsigned long long *counter, uint32_t v)
{
    __shared__ uint32_t wmax[16];
    uint32_t m = v;
    for (int off = 32; off > 0; off >>= 1)
    {
        const uint32_t o = __shfl_xor(m, off);
        m = o > m ? o : m;
    }
    if (__lane_id() == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        uint32_t t = 0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) t = wmax[w] > t ? wmax[w] : t;
        if (t) atomicMax(counter, static_cast<unsigned long long>(t));
    }
}

} // namespace pk
