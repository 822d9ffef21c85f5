// Internal interfaces between the host orchestration (api.cpp) and the kernel
// translation units. Device pointers only; every array is structure-of-arrays
// ("planar": component c of element i lives at ptr[c * stride + i]) so that a
// wavefront reading one component of 64 consecutive elements issues one
// coalesced 512-byte access.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>

namespace pk
{

constexpr uint32_t kNone = 0xFFFFFFFFu;
// the previous step's index of pair p (kNone for a new pair); src is null where the
// previous list was kept whole, each pair at its own index
__host__ __device__ __forceinline__ uint32_t source_of(const uint32_t *src, uint32_t p) { return src ? src[p] : p; }

// per-body state, n bodies, body id = index
struct bodies_dev
{
    uint32_t n;
    double *pos;        // 3 planes
    double *vel;        // 3
    double *ang;        // 3
    double *acc;        // 3 (external force accumulator / m, cleared every step)
    double *torque;     // 3
    double *quat;       // 4 planes: w, x, y, z
    double *mass;       // 1
    double *inv_mass;   // 1
    double *I_local;    // 9 planes, row-major element order
    double *Iinv_local; // 9
    double *Iw;         // 9
    double *Iw_inv;     // 9
    double *restitution;
    double *friction;
    double *fat; // 6 doubles per body: min xyz, max xyz
    uint32_t *shape;
    uint8_t *is_static;
    uint8_t *moved;
};

struct shapes_dev
{
    uint32_t n;
    const uint32_t *off; // n+1 vertex offsets (the caller's numbering: counts)
    const uint32_t *voff; // n: first vertex of each shape's block in verts. Blocks start at
                          // an even vertex (16-byte aligned) and hold an even count, the
                          // last vertex repeated when odd, so a scan reads vertex pairs
                          // as three aligned 16-byte loads
    const double *verts;  // xyz interleaved
    const double *rad;    // n: largest vertex distance from the shape origin
    double *local_box;   // 6 planes over shapes
    const uint32_t *uv;  // n: UV-sphere tag ((stacks << 16) | sectors) or 0
};

// manifold cache: a contact count (0-4, one byte) and a data slot per pair; up to 4
// contacts x 13 fields per slot, planar with stride dcap: field f of contact c of the
// manifold in slot s at data[(c * 13 + f) * dcap + s]. Only pairs with contacts hold a
// slot, so the data scales with the contacts, not with the broadphase pairs. A pair's
// slot is read only where its count is above 0 (an empty manifold's slot is not written).
struct manifolds_dev
{
    uint32_t cap;  // pairs
    uint8_t *count;
    uint32_t *slot;
    uint32_t dcap; // slots
    double *data;
    uint32_t *owner; // dcap: the pair whose contacts a slot holds (written by the manifold update)
};
constexpr int kMF = 13; // normal3 local_a3 local_b3 depth ln lt1 lt2
constexpr int kSimplexWords = 17; // GJK simplex hand-off per hit: count + 4 x (p.xyz, packed features)

// device counters (one 64-bit slot each)
enum counter_slot
{
    C_MOVED = 0,
    C_CAND = 1,
    // 2: unused
    C_PAIRS = 3,
    C_HITS = 4,
    C_CONTACTS = 5,
    C_ROWS = 6,
    C_REMAIN = 7,
    C_EPA_OVF = 8,
    C_EPA_BIG_FAIL = 9,
    C_EVENTS = 10,
    C_TREE_FAIL = 11,
    C_EPA_OVF2 = 12, // overflows of the LDS store, rerun in global memory
    C_KEEP = 13,
    C_ISL_MAX = 14,
    C_NCOL = 15,
    C_NLIST = 16,
    C_GJK_LIST = 17, // pairs left for GJK after the bounding-sphere cull
    C_GJK_LIST2 = 18, // of those, pairs the FP32 certified-separation pass left to the FP64 GJK
    C_MSLOTS = 19,    // manifold data slots taken by the manifold update
    C_MLIVE = 20,     // pairs listed for the manifold update (see launch_manifold_update)
    C_MLEAF = 21,     // leaves listed for the pair query (dynamic and moved)
    // pair query cursors: one per output region, each in its own 256-byte line, so
    // wavefronts reserving output space spread their atomics over kPairRegions
    // addresses instead of serialising on one
    C_QCUR = 32,
    C_NCOUNTERS = 32 + 8 * 32
};
constexpr uint32_t kPairRegions = 8;
constexpr uint32_t kCursorStride = 32;

// ---------------- k_body.hip ----------------
void launch_body_init(const bodies_dev &b, const shapes_dev &s, hipStream_t st);
void launch_shape_boxes(const shapes_dev &s, hipStream_t st);
void launch_integrate_vel(const bodies_dev &b, const shapes_dev &s, double gx, double gy, double gz,
                          double dt, unsigned long long *counters, hipStream_t st);
void launch_integrate_pos(const bodies_dev &b, double dt, hipStream_t st);
// integrate_pos with dt_prev, then integrate_vel with dt, in one pass (bit-identical to
// the two launches)
void launch_integrate_pos_vel(const bodies_dev &b, const shapes_dev &s, double dt_prev, double gx, double gy,
                              double gz, double dt, unsigned long long *counters, hipStream_t st);
void launch_fill_u64(unsigned long long *p, unsigned long long v, size_t n, hipStream_t st);
void launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t st);

// ---------------- k_sort.hip ----------------
// LSD radix sort of (key, value) pairs on 8-bit digits at the given bit shifts
// (ascending). Buffers ping-pong; returns true when the result is in (k_alt, v_alt).
struct sort_scratch
{
    uint32_t *hist = nullptr; // 256 * max_blocks
    size_t hist_cap = 0;
};
bool radix_sort_u64(uint64_t *k, uint32_t *v, uint64_t *k_alt, uint32_t *v_alt, uint32_t n,
                    const int *shifts, int n_shifts, sort_scratch &sc, hipStream_t st);
// exclusive scan of n uint32 flags into out, total written to *total (device)
void exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *block_sums,
                        uint32_t *total, hipStream_t st);
size_t scan_block_sums(uint32_t n);

// ---------------- k_broad.hip ----------------
// Node indices are unified: internal nodes [0, n-1), leaf i is node n-1+i; node 0 is
// the root when n >= 2. Leaves are in Morton order.
struct lbvh_dev
{
    uint32_t n; // leaves
    uint64_t *morton;
    uint32_t *ids;
    uint64_t *morton_alt;
    uint32_t *ids_alt;
    int32_t *left;       // n-1: children (unified)
    int32_t *right;      // n-1
    uint32_t *first;     // n-1: leaf range of the node
    uint32_t *last;      // n-1
    uint32_t *lastf;     // n-1: last | (every leaf below is moved) << 31
    uint4 *nrec;         // n-1: (left, right, first, last) in one record for the refit climb
    float4 *inode_lo;    // n-1: FP32 box min rounded down, w = left child bits
    float4 *inode_hi;    // n-1: FP32 box max rounded up, w = escape bits
    uint32_t *parent;    // 2n-1
    uint32_t *escape;    // 2n-1: next node after the subtree in left-first order
    double *leaf_box;    // 6n: exact FP64 leaf bounds, leaf-major
    float4 *leaf_f32;    // 2n or null: leaf bounds rounded outward to FP32 (min, max; w unused),
                         // tested before the exact bounds in the pair query
    uint32_t *leaf_body; // n: body id of leaf i
    uint8_t *leaf_moved; // n: leaf queries this step (dynamic and moved)
    uint32_t *moved_list; // n: the querying leaves' sorted indices, ascending (count at counters[C_MLEAF])
    uint32_t *moved_wg;   // 2 * ceil(n / 256): moved leaves per 256-leaf block, then their exclusive offsets
    // refit chunks of kRefitChunk leaves: per leaf, the union of its chunk's leaf boxes
    // up to it and from it on (4n: prefix lo, prefix hi, suffix lo, suffix hi; w of lo =
    // every-leaf-moved flag), and a sparse table over chunk totals (2 * levels * chunks)
    uint32_t n_chunks, span_levels;
    float4 *span_box;
    float4 *span_tab;
    unsigned long long *scene; // 6 ordered-integer bounds of leaf centroids
    // batched independent worlds (n_worlds > 1): bodies [wstart[w], wstart[w+1]) form
    // world w; Morton keys carry the world in their top wbits bits, so world w's leaves
    // are sorted positions [wstart[w], wstart[w+1]) and the walk prunes other worlds
    uint32_t n_worlds;
    int wbits;
    const uint32_t *world_of; // n: world of each body
    const uint32_t *wstart;   // n_worlds + 1
};
// Morton keys: 16 bits per axis (bit 3k+2 x, 3k+1 y, 3k z), six 8-bit sort passes; a
// batched context puts the world in the top bits and drops as many low Morton bits
constexpr int kMortonBits = 48;
constexpr uint32_t kRefitChunk = 1024;
inline uint32_t refit_chunks(uint32_t n) { return (n + kRefitChunk - 1) / kRefitChunk; }
inline uint32_t refit_levels(uint32_t n)
{
    uint32_t c = refit_chunks(n), l = 1;
    while ((2u << (l - 1)) <= c) ++l;
    return l;
}
void launch_lbvh_build(lbvh_dev &t, const bodies_dev &b, bool all_moved, sort_scratch &sc,
                       unsigned long long *counters, hipStream_t st);
// the last build's hierarchy refitted to the current leaf boxes (no Morton codes, sort
// or topology); the same ids, node count and world layout as that build
void launch_lbvh_refit(lbvh_dev &t, const bodies_dev &b, bool all_moved, unsigned long long *counters,
                       hipStream_t st);
// candidate keys (pairs with at least one moved leaf), appended per region (below);
// a malformed tree is reported in counters[C_TREE_FAIL]
// workgroup w writes its candidates to region r = w % n_regions: out_keys[r * rcap, ...)
// with the count at counters[C_QCUR + r * kCursorStride]
void launch_pair_query(const lbvh_dev &t, const bodies_dev &b, uint64_t *out_keys, uint32_t rcap,
                       uint32_t n_regions, unsigned long long *counters, hipStream_t st);
// concatenates the regions' counts[r] keys into dst (off: exclusive prefix of counts)
struct region_counts
{
    uint32_t off[kPairRegions + 1];
};
void launch_pair_concat(const uint64_t *regions, uint32_t rcap, const region_counts &rc, uint64_t *dst,
                        hipStream_t st);
// src[e] = index of candidate e in the sorted previous list (kNone if new)
void launch_pair_match(const uint64_t *cand, uint32_t m, const uint64_t *prev, uint32_t n_prev,
                       uint32_t *src, hipStream_t st);
// flag the previous pairs that stay without being candidates (neither body moved), then
// compact them
void launch_prev_keep(const uint64_t *prev, uint32_t n_prev, const bodies_dev &b, uint32_t *flag, hipStream_t st);
void launch_prev_compact(const uint64_t *prev, uint32_t n_prev, const uint32_t *flag,
                         const uint32_t *pos, uint64_t *akeys, uint32_t *asrc, hipStream_t st);
// writes list x (sorted) into its slots of the merge of x and y (sorted, disjoint)
void launch_merge(const uint64_t *xk, const uint32_t *xs, uint32_t nx, const uint64_t *yk,
                  uint32_t ny, uint64_t *out_keys, uint32_t *out_src, hipStream_t st);

// ---------------- k_narrow.hip ----------------
struct gjk_out_dev
{
    uint32_t *hit_list;   // compacted pair indices whose GJK found an intersection (cap entries;
                          // hits past cap are counted, not written)
    double *simplex;      // per hit-list entry: n, then 4 x 9 doubles (p, pa, pb), planar stride cap
    uint32_t cap;
    double *info;         // per hit-list entry: normal3 wa3 wb3 depth, planar stride cap
    uint32_t *hit_of;     // per pair: its hit-list entry (read only where status is 1)
    uint8_t *status;      // per pair: 0 separated, 1 contact, 2 degenerate, 3 overflow
};
// FP32 copies of the shapes for the certified-separation pass (k_gjk32.hip): vertex i
// of shape s at v[off[s] + i] (xyz, w = 0), rounded to nearest; rad rounded up
struct shapes32_dev
{
    const float4 *v;
    const uint32_t *off;
    const float *rad;
    const uint32_t *uv;
};
void launch_verts32(const shapes_dev &s, float4 *out, hipStream_t st);
// Pairs the cull listed (counters[C_GJK_LIST] entries of `list`, at most n_pairs) that
// FP32 proves apart by more than GJK's margin are written separated (status 0); the
// rest, and every pair with contacts after the previous step (prev_count of its source,
// source_of(src, p), above 0), are listed in list2, count in counters[C_GJK_LIST2].
void launch_gjk32(const uint64_t *keys, const uint32_t *list, uint32_t n_pairs, const bodies_dev &b,
                  const shapes32_dev &s, bool uv, const uint32_t *src, const uint8_t *prev_count, uint8_t *status,
                  uint32_t *list2, unsigned long long *counters, hipStream_t st);
// GJK over every pair with `group` (2 or 4) lanes per pair; pairs whose bounding
// spheres are clearly apart are written separated first and the rest listed in `list`
// (capacity n_pairs), so GJK wavefronts run only pairs that need it; cull_rec: 22
// doubles per body of scratch for the per-body cull records. With s32 (the mixed-
// precision form), the cull first tests FP32 spheres (certainly apart or decided by the
// FP64 records), and the FP32 pass above runs between the cull and the FP64 GJK, which
// then takes list2 (capacity n_pairs).
// list kernels (GJK passes over the cull's list, whose count only the device knows)
// launch at most this many workgroups, which step through longer lists
constexpr unsigned kListGrid = 16384;
// workgroups of 256 for n items, rounded up to a multiple of 8 (one per XCD in each
// turn of the dispatcher) and capped at kListGrid; the kernels step through the rest
inline unsigned list_grid(size_t n)
{
    const size_t g = (n + 255) / 256;
    return static_cast<unsigned>(std::min<size_t>((g + 7) / 8 * 8, kListGrid));
}
struct gjk32_in
{
    const shapes32_dev *s32; // null: no FP32 pass
    const uint32_t *src;
    const uint8_t *prev_count;
    uint32_t *list2;
};
// What the cull also settles of the manifold update: a pair whose bodies are apart has
// no new contact, so its next manifold is empty unless it had contacts (prev_count of
// its source, source_of(src, p), above 0). The cull writes every pair's next count,
// slot and event empty, and lists the culled pairs that had contacts in live_list
// (count at counters[C_MLIVE], at most live_cap: the previous step's non-empty
// manifolds); the manifold update then runs over live_list and the GJK list.
struct cull_manifold_out
{
    const uint32_t *src;
    const uint8_t *prev_count;
    manifolds_dev next;
    uint8_t *events;
    uint32_t *live_list;
    uint32_t live_cap;
};
void launch_gjk(const uint64_t *keys, uint32_t n_pairs, const bodies_dev &b, const shapes_dev &s, bool uv, int group,
                gjk_out_dev &g, uint32_t *list, double *cull_rec, const gjk32_in &f32, const cull_manifold_out &mo,
                unsigned long long *counters, hipStream_t st);
// Launch shapes the autotuner picks (pkg_autotune); none changes a result.
struct tuning
{
    int gjk_group = 4;           // lanes per pair in k_gjk (1, 2, 4, 8): 64 / gjk_group pairs per wavefront
    int epa_block = 64;          // threads per workgroup of k_epa (64, 128, 256)
    uint32_t epa_lds_max = 512;  // batches of at most this many GJK hits run one EPA solve per workgroup
                                 // with the LDS store (k_epa_lds); larger ones one per thread
    uint32_t tree_refits = 7;    // steps that refit the last LBVH hierarchy between full builds
};
void launch_epa(const uint64_t *keys, uint32_t n_hits, const bodies_dev &b, const shapes_dev &s, bool uv,
                const tuning &tu, gjk_out_dev &g, uint32_t *ovf_list, unsigned long long *counters, hipStream_t st);
// overflowed hits rerun with an LDS store; its own overflows are listed in ovf2_list
// (count at counters[C_EPA_OVF2]) for launch_epa_big
void launch_epa_lds(const uint64_t *keys, const uint32_t *ovf_list, uint32_t n_ovf, const bodies_dev &b,
                    const shapes_dev &s, const gjk_out_dev &g, uint32_t *ovf2_list, unsigned long long *counters,
                    hipStream_t st);
size_t epa_big_bytes();
// launch_epa runs batches this small with the LDS store already (its overflows skip
// launch_epa_lds)
bool epa_first_tier_is_lds(uint32_t n_hits, const tuning &tu);
void launch_epa_big(const uint64_t *keys, const uint32_t *ovf_list, uint32_t n_ovf,
                    const bodies_dev &b, const shapes_dev &s, gjk_out_dev &g, void *pool,
                    unsigned long long *counters, hipStream_t st);
// whether the cull of n_pairs pairs reads per-body records, and with them settles the
// culled pairs' manifolds (cull_manifold_out)
bool cull_uses_records(uint32_t n_pairs, uint32_t n_bodies);
// culled (the cull settled the manifolds): the manifolds of the pairs the cull listed
// (live_list, counters[C_MLIVE] entries) and of the pairs GJK ran on (gjk_list,
// counters[C_GJK_LIST] entries); every other pair's manifold the cull wrote empty.
// Otherwise a first pass writes every pair empty and lists those with a GJK hit or
// contacts before in live_list (n_pairs entries), and the update runs over that list.
// bound (at least the listed count) sizes the grid; the kernels read the counts.
void launch_manifold_update(const uint64_t *keys, const uint32_t *src, uint32_t n_pairs, uint32_t bound,
                            const bodies_dev &b, const gjk_out_dev &g, manifolds_dev prev, manifolds_dev next,
                            uint8_t *events, bool epa_guard, bool culled, uint32_t *live_list,
                            const uint32_t *gjk_list, unsigned long long *counters, hipStream_t st);

// standalone narrowphase over shape pairs (pkg_gjk_epa_batch)
struct shape_pair_dev
{
    const int32_t *kind_a, *kind_b;
    const double *prm_a, *prm_b; // 10 doubles per shape
    const uint32_t *vbeg_a, *vend_a, *vbeg_b, *vend_b;
    const double *verts;
    const uint32_t *uv_a, *uv_b; // UV-sphere tags of mesh shapes (0: none)
};
void launch_gjk_epa_batch(const shape_pair_dev &sp, uint32_t n, double *out, int32_t *hit,
                          void *big_pool, hipStream_t st);
// The same over device arrays with scratch the caller does not allocate: words (4 x u32:
// invalid pairs, pairs listed for the large store, pairs that overflowed it too; cleared
// by the first launch), ovf_list (n entries) and n_stores large EPA stores. Invalid pairs
// get hit = -1. uv tags may be null.
void launch_gjk_epa_batch_dev(const shape_pair_dev &sp, uint32_t n, uint32_t n_verts, double *out, int32_t *hit,
                              uint32_t *words, uint32_t *ovf_list, void *stores, uint32_t n_stores, hipStream_t st);

// ---------------- k_solver.hip ----------------
struct rows_dev
{
    uint32_t cap4;   // 4 * pair capacity
    double *d;       // kRowF planes of cap4
    uint8_t *valid;  // cap4
};
constexpr int kRowF = 44; // 38 row words + the contact's lever arms (6)
constexpr uint32_t kUncoloured = 0xFFFFFFFEu; // colour of an item still to be coloured
// colour[p] = kUncoloured for manifolds that make solver rows (else kNone), their row
// count at counters[C_ROWS], and those manifolds listed in list (count in C_NLIST).
// Only pairs holding contacts can make rows: the pass runs over the data slots the
// manifold update filled (counters[C_MSLOTS] of at most m.dcap) after every colour has
// been set to kNone.
void launch_row_flags(const uint64_t *keys, uint32_t n_pairs, const bodies_dev &b, manifolds_dev m, uint32_t *colour,
                      uint32_t *list, unsigned long long *counters, hipStream_t st);
// rows of manifold ord[k], k < n, at slots c * cap + k of r (solve order); with ord
// null, of manifold k at slots 4k + c (pair order). f32: r.d holds FP32 planes (the
// rows computed in FP64 and rounded once)
void launch_contact_setup(const uint64_t *keys, uint32_t n, const uint32_t *ord, uint32_t cap, const bodies_dev &b,
                          manifolds_dev m, rows_dev r, double dt, double threshold, bool f32, hipStream_t st);
// With batched worlds, a manifold's colouring priority uses its index within its own
// world's pair list (p - pair_base[world]), so each world colours as if stepped alone;
// world_of and pair_base are null for a single world.
void launch_pair_base(const uint64_t *keys, uint32_t n_pairs, const uint32_t *world_of, uint32_t *pair_base,
                      hipStream_t st);
// Colouring rounds over a list of items (null: all n). n_dev, when not null, holds the
// list's count on the device and n only bounds it (the grid). The claim array is
// cleared once before the first round; `used` (colours 0-63 taken at each body) too;
// each round the claim holders take the smallest colour free at their dynamic bodies,
// and C_NCOL ends as the colour count.
void launch_colour_claim(const uint64_t *keys, const uint32_t *list, uint32_t n, const unsigned long long *n_dev,
                         const bodies_dev &b, const uint32_t *colour, const uint32_t *world_of,
                         const uint32_t *pair_base, unsigned long long *claim, uint32_t round, hipStream_t st);
void launch_colour_assign(const uint64_t *keys, const uint32_t *list, uint32_t n, const unsigned long long *n_dev,
                          const bodies_dev &b, uint32_t *colour, const uint32_t *world_of, const uint32_t *pair_base,
                          const unsigned long long *claim, unsigned long long *used, uint32_t round,
                          unsigned long long *counters, hipStream_t st);
// the listed solver manifolds (list[0, nl), or 0..nl-1 when list is null; every one
// coloured below n_colours) by colour, each colour in index order: returns the order
// array (one of the two given, nl entries each) and colour_start[0..n_colours];
// n_pairs bounds the indices
uint32_t *launch_colour_bucket(const uint32_t *colour, const uint32_t *list, uint32_t nl, uint32_t n_pairs,
                               uint32_t n_colours, uint64_t *keys, uint64_t *keys_alt, uint32_t *order,
                               uint32_t *order_alt, uint32_t *colour_start, sort_scratch &sc, hipStream_t st);
void launch_solve_colour(const uint64_t *keys, const uint32_t *order, uint32_t begin, uint32_t end,
                         const bodies_dev &b, rows_dev r, uint32_t cap, bool warm, hipStream_t st);
void launch_impulse_writeback(const uint32_t *order, uint32_t n_act, manifolds_dev m, rows_dev r, uint32_t cap,
                              hipStream_t st);
// Contact islands (connected components of the row manifolds over dynamic bodies,
// each rooted at its smallest body id), manifolds sorted by (island, colour): a
// workgroup solves a chunk of whole islands, every colour and pass, in one launch.
struct island_sort
{
    const uint64_t *sk;  // sorted keys (island << cbits | colour; `last` for no rows)
    const uint32_t *ord; // manifold index at each sorted position
    int cbits;
    uint64_t last;
};
constexpr uint32_t kIslandMaxManifolds = 4096; // larger islands use the per-colour launches
constexpr uint32_t kIslandMaxColours = 64;  // above: the per-colour launches
// `fail` becomes all-ones if the islands could not be built (the caller then falls back)
// over the nl row manifolds list[0..nl) (the colouring's list); positions of the sorted
// order index that list's manifolds, so n_pairs below is nl
island_sort launch_island_sort(const uint64_t *keys, const uint32_t *colour, const uint32_t *list, uint32_t nl,
                               uint32_t n_colours, const bodies_dev &b, uint32_t *par, uint64_t *ik, uint32_t *iv,
                               uint64_t *ik_alt, uint32_t *iv_alt, unsigned long long *fail, sort_scratch &sc,
                               hipStream_t st);
// island starts (positions where the island changes) and the row-manifold count
void launch_island_segments(const island_sort &is, uint32_t n_pairs, uint32_t *flags, uint32_t *pos, uint32_t *bsums,
                            uint32_t *total, uint32_t *starts, uint32_t *n_act, hipStream_t st);
void launch_island_maxlen(const uint32_t *starts, const uint32_t *totals, uint32_t max_islands,
                          unsigned long long *mx, hipStream_t st);
// longest: manifolds of the longest island (chunks are sized so each fits one workgroup)
void launch_solve_islands(const uint64_t *keys, const island_sort &is, const uint32_t *starts, uint32_t n_isl,
                          uint32_t n_act, uint32_t n_colours, const bodies_dev &b, const rows_dev &r, uint32_t cap,
                          uint32_t iters, uint32_t longest, bool f32, hipStream_t st);
// FP32 solve (rows_dev whose d holds FP32 planes, written by launch_contact_setup with
// f32; body velocities accumulate in FP64)
void launch_solve_colour32(const uint64_t *keys, const uint32_t *order, uint32_t begin, uint32_t end,
                           const bodies_dev &b, rows_dev r32, uint32_t cap, bool warm, hipStream_t st);
void launch_impulse_writeback32(const uint32_t *order, uint32_t n_act, manifolds_dev m, rows_dev r32, uint32_t cap,
                                hipStream_t st);

// ---------------- k_query.hip ----------------
// hits of a batch of rays / boxes: body, distance (rays), query index; sort buffers
struct query_scratch
{
    uint32_t *hid, *hq, *perm, *perm_alt;
    double *hd;
    uint64_t *keys, *keys_alt;
};
// per internal node (left, right, parent, split axis) for the ordered closest-hit walk
void launch_ray_nodes(const lbvh_dev &t, uint4 *rn, hipStream_t st);
void launch_ray_closest(const lbvh_dev &t, const uint4 *rn, const bodies_dev &b, uint32_t nr, const double *org,
                        const double *dir, const double *maxd, const uint32_t *rw, uint32_t *out_id, double *out_d,
                        unsigned long long *fail, hipStream_t st);
void launch_ray_count(const lbvh_dev &t, const bodies_dev &b, uint32_t nr, const double *org, const double *dir,
                      const double *maxd, const uint32_t *rw, uint32_t *cnt, unsigned long long *fail, hipStream_t st);
void launch_ray_fill(const lbvh_dev &t, const bodies_dev &b, uint32_t nr, const double *org, const double *dir,
                     const double *maxd, const uint32_t *rw, const uint32_t *off, uint32_t *hid, double *hd,
                     uint32_t *hq, unsigned long long *fail, hipStream_t st);
void launch_box_count(const lbvh_dev &t, const bodies_dev &b, uint32_t nq, const double *boxes, const uint32_t *qw,
                      uint32_t *cnt, unsigned long long *fail, hipStream_t st);
void launch_box_fill(const lbvh_dev &t, const bodies_dev &b, uint32_t nq, const double *boxes, const uint32_t *qw,
                     const uint32_t *off, uint32_t *hid, uint32_t *hq, unsigned long long *fail, hipStream_t st);
// sorts the m hits by (query, distance, body) (by_distance) or (query, body) and writes
// bodies (and distances) in that order
void launch_hits_sort(query_scratch &qs, uint32_t m, uint32_t n_bodies, uint32_t n_queries, bool by_distance,
                      uint32_t *out_id, double *out_d, sort_scratch &sc, hipStream_t st);

// ---------------- k_joint.hip ----------------
// Joint records: kJW planes of `cap` doubles (word w of joint j at rec[w * cap + j]),
// joints dense in key order. Words: la 0-2, lb 3-5, ax_a 6-8, ax_b 9-11, f1 12-14,
// f2 15-17, qrel 18-21 (w x y z), distance 22, stiffness 23, damping 24, acc 25-30.
constexpr int kJW = 31;
constexpr int kJR = 13; // row words: Jv 0-2, Jwa 3-5, Jwb 6-8, Meff 9, bias 10, gamma 11, acc 12
struct joints_dev
{
    uint32_t n, cap;
    double *rec;
    uint8_t *type;
    uint32_t *a, *b;
    uint64_t *keys;  // (a << 32) | b, the colouring input
    double *rows;    // 6 * kJR planes of cap: word w of row r of joint j at rows[(r * kJR + w) * cap + j]
    uint8_t *nrows;  // rows built this step
};
// Records after an add or remove: position i of the new list copies joint src[i] of
// the old one, or, when src[i] has bit 31 set, constructs joint (src[i] & 0x7fffffff)
// of `desc` (21 doubles each: local_a, local_b, local_axis_a, local_axis_b, anchor,
// axis, distance, stiffness, damping) with `dint` (type, a, b) from the current poses.
void launch_joint_rebuild(const joints_dev &old_j, const joints_dev &new_j, const uint32_t *src,
                          const double *desc, const uint32_t *dint, const bodies_dev &b, hipStream_t st);
void launch_joint_rows(const joints_dev &j, const bodies_dev &b, double dt, hipStream_t st);
void launch_joint_solve(const joints_dev &j, const uint32_t *order, uint32_t begin, uint32_t end,
                        const bodies_dev &b, bool warm, hipStream_t st);
void launch_joint_writeback(const joints_dev &j, hipStream_t st);

} // namespace pk
