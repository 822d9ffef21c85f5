// LSD radix sort of (uint64 key, uint32 value) on 8-bit digits, and a device-wide
// exclusive scan. Per pass: (1) per-tile digit histograms in LDS, (2) exclusive scan
// of the digit-major histogram table, (3) stable scatter where each key's rank
// among equal digits inside a 64-lane wavefront comes from 8 ballots (one per digit
// bit) and the wavefronts of a tile are chained through LDS counters.
#include "pk_access.hpp"

#include <stdexcept>

namespace pk
{
namespace
{
constexpr int kThreads = 256;
constexpr int kItems = 16;               // keys per thread and tile, at most
constexpr int kTile = kThreads * kItems; // keys per tile, at most
constexpr int kScanItems = 8;
constexpr int kScanTile = kThreads * kScanItems;

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

// a tile is kThreads * items keys; fewer items per thread give more tiles (workgroups)
// when n is small, so a pass fills the device
__global__ void __launch_bounds__(kThreads) k_rs_hist(const uint64_t *keys, uint32_t n, int shift,
                                                      uint32_t *hist, uint32_t nblocks, int items)
{
    __shared__ uint32_t h[256];
    int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    size_t base = static_cast<size_t>(blockIdx.x) * kThreads * items;
    for (int it = 0; it < items; ++it)
    {
        size_t idx = base + static_cast<size_t>(it) * kThreads + tid;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[static_cast<size_t>(tid) * nblocks + blockIdx.x] = h[tid];
}

// Stable scatter of one tile of kThreads * items keys starting at `base`. run[d]: the
// output position of the tile's first key with digit d (advanced in place); wc: LDS
// workspace.
__device__ __forceinline__ void scatter_tile(const uint64_t *kin, const uint32_t *vin, uint64_t *kout,
                                             uint32_t *vout, uint32_t n, int shift, size_t base, int items,
                                             uint32_t *run, uint32_t (*wc)[256])
{
    int tid = threadIdx.x;
    int wave = tid >> 6;
    for (int it = 0; it < items; ++it)
    {
        size_t idx = base + static_cast<size_t>(it) * kThreads + tid;
        bool valid = idx < n;
        uint64_t key = valid ? kin[idx] : 0;
        uint32_t val = valid ? vin[idx] : 0;
        uint32_t d = static_cast<uint32_t>((key >> shift) & 255u);
        wc[0][tid] = 0;
        wc[1][tid] = 0;
        wc[2][tid] = 0;
        wc[3][tid] = 0;
        __syncthreads();
        uint64_t mask = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit)
        {
            bool b = ((d >> bit) & 1u) != 0;
            uint64_t bm = __ballot(b);
            mask &= b ? bm : ~bm;
        }
        uint32_t rank = static_cast<uint32_t>(__popcll(mask & lanemask_lt()));
        if (valid && rank == 0) wc[wave][d] = static_cast<uint32_t>(__popcll(mask));
        __syncthreads();
        {
            uint32_t r = run[tid];
            uint32_t c0 = wc[0][tid], c1 = wc[1][tid], c2 = wc[2][tid], c3 = wc[3][tid];
            wc[0][tid] = r;
            wc[1][tid] = r + c0;
            wc[2][tid] = r + c0 + c1;
            wc[3][tid] = r + c0 + c1 + c2;
            run[tid] = r + c0 + c1 + c2 + c3;
        }
        __syncthreads();
        if (valid)
        {
            uint32_t dst = wc[wave][d] + rank;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kThreads)
    k_rs_scatter(const uint64_t *kin, const uint32_t *vin, uint64_t *kout, uint32_t *vout, uint32_t n,
                 int shift, const uint32_t *offs, uint32_t nblocks, int items)
{
    __shared__ uint32_t run[256];
    __shared__ uint32_t wc[4][256];
    run[threadIdx.x] = offs[static_cast<size_t>(threadIdx.x) * nblocks + blockIdx.x];
    scatter_tile(kin, vin, kout, vout, n, shift, static_cast<size_t>(blockIdx.x) * kThreads * items, items, run,
                 wc);
}

// ---- scan ----
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
    int lane = __lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        uint32_t t = __shfl_up(x, off);
        if (lane >= off) x += t;
    }
    return x;
}

// exclusive block scan of one value per thread; returns the block total in *total
__device__ uint32_t block_excl_scan(uint32_t x, uint32_t *total)
{
    __shared__ uint32_t wsum[kThreads / 64];
    int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
    uint32_t incl = wave_incl_scan(x);
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (int w = 0; w < kThreads / 64; ++w)
    {
        if (w < wave) woff += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return woff + incl - x;
}

__global__ void __launch_bounds__(kThreads) k_scan_reduce(const uint32_t *in, uint32_t n, uint32_t *bsums)
{
    size_t base = static_cast<size_t>(blockIdx.x) * kScanTile + static_cast<size_t>(threadIdx.x) * kScanItems;
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += in[base + k];
    uint32_t tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) bsums[blockIdx.x] = tot;
}

// single block: exclusive scan of nb block sums in place, grand total to *total
__global__ void __launch_bounds__(kThreads) k_scan_top(uint32_t *bsums, uint32_t nb, uint32_t *total)
{
    uint32_t carry = 0;
    for (uint32_t start = 0; start < nb; start += kThreads)
    {
        uint32_t i = start + threadIdx.x;
        uint32_t x = i < nb ? bsums[i] : 0;
        uint32_t tot;
        uint32_t ex = block_excl_scan(x, &tot);
        if (i < nb) bsums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(kThreads)
    k_scan_down(const uint32_t *in, uint32_t *out, uint32_t n, const uint32_t *bsums)
{
    size_t base = static_cast<size_t>(blockIdx.x) * kScanTile + static_cast<size_t>(threadIdx.x) * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; ++k)
    {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan(s, &tot) + bsums[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k)
    {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
}

// n <= kTile: every pass in one workgroup (histogram, scan of the 256 digit counts,
// scatter), the passes ping-ponging through the two global buffers with workgroup
// barriers between them: one launch instead of five per pass
struct shift_list
{
    int s[16];
    int n;
};
__global__ void __launch_bounds__(kThreads)
    k_rs_small(uint64_t *k, uint32_t *v, uint64_t *k_alt, uint32_t *v_alt, uint32_t n, shift_list sl)
{
    __shared__ uint32_t h[256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t wc[4][256];
    const int tid = threadIdx.x;
    bool alt = false;
    for (int p = 0; p < sl.n; ++p)
    {
        const uint64_t *ki = alt ? k_alt : k;
        const uint32_t *vi = alt ? v_alt : v;
        uint64_t *ko = alt ? k : k_alt;
        uint32_t *vo = alt ? v : v_alt;
        const int shift = sl.s[p];
        h[tid] = 0;
        __syncthreads();
        for (int it = 0; it < kItems; ++it)
        {
            uint32_t idx = static_cast<uint32_t>(it * kThreads + tid);
            if (idx < n) atomicAdd(&h[(ki[idx] >> shift) & 255u], 1u);
        }
        __syncthreads();
        uint32_t tot;
        run[tid] = block_excl_scan(h[tid], &tot);
        __syncthreads();
        scatter_tile(ki, vi, ko, vo, n, shift, 0, kItems, run, wc); // ends with a barrier
        alt = !alt;
    }
}

// n <= kScanTile: the three scan launches in one
__global__ void __launch_bounds__(kThreads) k_scan_small(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *total)
{
    const uint32_t base = threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; ++k)
    {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan(s, &tot);
    for (int k = 0; k < kScanItems; ++k)
    {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0) *total = tot;
}

uint32_t nblk(size_t n, size_t tile) { return static_cast<uint32_t>((n + tile - 1) / tile); }

void check(hipError_t e)
{
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}
} // namespace

size_t scan_block_sums(uint32_t n) { return nblk(n, kScanTile) + 1; }

void exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *block_sums,
                        uint32_t *total, hipStream_t st)
{
    if (n <= static_cast<uint32_t>(kScanTile))
    {
        k_scan_small<<<1, kThreads, 0, st>>>(in, out, n, total);
        return;
    }
    uint32_t nb = nblk(n, kScanTile);
    k_scan_reduce<<<nb, kThreads, 0, st>>>(in, n, block_sums);
    k_scan_top<<<1, kThreads, 0, st>>>(block_sums, nb, total);
    k_scan_down<<<nb, kThreads, 0, st>>>(in, out, n, block_sums);
}

// keys per thread for a pass over n keys: the most (up to kItems) that still give at
// least kFillTiles tiles
constexpr uint32_t kFillTiles = 512;
int tile_items(uint32_t n)
{
    int items = kItems;
    while (items > 1 && nblk(n, static_cast<size_t>(kThreads) * items) < kFillTiles) items >>= 1;
    return items;
}

bool radix_sort_u64(uint64_t *k, uint32_t *v, uint64_t *k_alt, uint32_t *v_alt, uint32_t n,
                    const int *shifts, int n_shifts, sort_scratch &sc, hipStream_t st)
{
    if (n <= 1 || n_shifts == 0) return false;
    if (n <= static_cast<uint32_t>(kTile) && n_shifts <= 16)
    {
        shift_list sl{};
        for (int p = 0; p < n_shifts; ++p) sl.s[p] = shifts[p];
        sl.n = n_shifts;
        k_rs_small<<<1, kThreads, 0, st>>>(k, v, k_alt, v_alt, n, sl);
        return (n_shifts & 1) != 0;
    }
    const int items = tile_items(n);
    uint32_t nb = nblk(n, static_cast<size_t>(kThreads) * items);
    size_t entries = static_cast<size_t>(256) * nb;
    size_t need = 2 * entries + scan_block_sums(static_cast<uint32_t>(entries)) + 1;
    if (sc.hist_cap < need)
    {
        if (sc.hist) check(hipFree(sc.hist));
        sc.hist = nullptr;
        check(hipMalloc(&sc.hist, need * sizeof(uint32_t)));
        sc.hist_cap = need;
    }
    uint32_t *hist = sc.hist;
    uint32_t *hscan = hist + entries;
    uint32_t *bsums = hscan + entries;
    uint32_t *total = bsums + scan_block_sums(static_cast<uint32_t>(entries));
    bool alt = false;
    for (int p = 0; p < n_shifts; ++p)
    {
        const uint64_t *ki = alt ? k_alt : k;
        const uint32_t *vi = alt ? v_alt : v;
        uint64_t *ko = alt ? k : k_alt;
        uint32_t *vo = alt ? v : v_alt;
        k_rs_hist<<<nb, kThreads, 0, st>>>(ki, n, shifts[p], hist, nb, items);
        exclusive_scan_u32(hist, hscan, static_cast<uint32_t>(entries), bsums, total, st);
        k_rs_scatter<<<nb, kThreads, 0, st>>>(ki, vi, ko, vo, n, shifts[p], hscan, nb, items);
        alt = !alt;
    }
    return alt;
}

} // namespace pk
