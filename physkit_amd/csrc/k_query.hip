// World queries over the broadphase LBVH: raycast (world_base::raycast,
// include/physkit/core/world.h:259-319, over dynamic_bvh::raycast, bvh.h:406-450) and
// AABB query (dynamic_bvh::query_aabb, bvh.h:313-344). The reference walks its static
// and dynamic trees; here one thread per ray or box walks the LBVH of all bodies
// stacklessly (left child on a hit, escape pointer otherwise). Internal nodes hold
// FP32 boxes rounded outward, tested in FP64 with the same slab arithmetic: the slab
// interval only widens as the box grows, so no leaf is pruned that the exact test would
// accept. Leaves hold the exact FP64 tree bounds (fat boxes), tested with
// ray::intersect_distance (bvh.h:58-91) operation for operation.
#include "pk_access.hpp"

#include <utility>

namespace pk
{
namespace
{
constexpr int kBlock = 256;
unsigned grid(size_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

// ray::intersect_distance: slab test with NaN products (0 * inf) replaced by the
// unconstrained limits; the entry distance clamped at 0
__device__ __forceinline__ bool ray_slab(v3 o, v3 inv, v3 mn, v3 mx, double maxd, double &t)
{
    double lo, hi, tminx, tmaxx, tminy, tmaxy, tminz, tmaxz;
    lo = (mn.x - o.x) * inv.x;
    hi = (mx.x - o.x) * inv.x;
    if (isnan(lo)) lo = -__builtin_inf();
    if (isnan(hi)) hi = __builtin_inf();
    tminx = fmin(lo, hi);
    tmaxx = fmax(lo, hi);
    lo = (mn.y - o.y) * inv.y;
    hi = (mx.y - o.y) * inv.y;
    if (isnan(lo)) lo = -__builtin_inf();
    if (isnan(hi)) hi = __builtin_inf();
    tminy = fmin(lo, hi);
    tmaxy = fmax(lo, hi);
    lo = (mn.z - o.z) * inv.z;
    hi = (mx.z - o.z) * inv.z;
    if (isnan(lo)) lo = -__builtin_inf();
    if (isnan(hi)) hi = __builtin_inf();
    tminz = fmin(lo, hi);
    tmaxz = fmax(lo, hi);
    double tmin = tminx;
    if (tmin < tminy) tmin = tminy;
    if (tmin < tminz) tmin = tminz;
    double tmax = tmaxx;
    if (tmaxy < tmax) tmax = tmaxy;
    if (tmaxz < tmax) tmax = tmaxz;
    if (tmax >= 0.0 && tmin <= tmax && tmin <= maxd)
    {
        t = (0.0 < tmin) ? tmin : 0.0;
        return true;
    }
    return false;
}

struct qray
{
    v3 o, inv;
    double maxd;
    uint32_t wlo, whi; // sorted leaf range searched (one world's leaves, or all)
};

__device__ __forceinline__ qray load_ray(const lbvh_dev &t, uint32_t r, const double *org, const double *dir,
                                         const double *maxd, const uint32_t *rw)
{
    qray q;
    q.o = mk(org[3 * r], org[3 * r + 1], org[3 * r + 2]);
    v3 d = normalized(mk(dir[3 * r], dir[3 * r + 1], dir[3 * r + 2])); // ray::ray, bvh.h:46-49
    q.inv = mk(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
    q.maxd = maxd[r];
    q.wlo = 0;
    q.whi = t.n - 1;
    if (rw && t.n_worlds > 1)
    {
        uint32_t w = rw[r];
        if (w < t.n_worlds)
        {
            q.wlo = t.wstart[w];
            q.whi = t.wstart[w + 1] - 1; // an empty world gives whi < wlo: nothing is visited
        }
        else
        {
            q.wlo = 1; // no such world: nothing is visited
            q.whi = 0;
        }
    }
    return q;
}

__device__ __forceinline__ bool node_in_range(const lbvh_dev &t, uint32_t node, uint32_t wlo, uint32_t whi)
{ return t.first[node] <= whi && (t.lastf[node] & 0x7fffffffu) >= wlo; }

// Calls leaf(body, distance) for every live leaf of the range the ray meets within
// max(); max() may shrink as leaves are accepted (closest-hit pruning).
template <class Leaf, class Max>
__device__ __forceinline__ void walk_ray(const lbvh_dev &t, const bodies_dev &b, const qray &q, Leaf &&leaf,
                                         Max &&maxd, unsigned long long *fail)
{
    const uint32_t n = t.n;
    if (n == 0 || q.whi < q.wlo || q.whi == 0xFFFFFFFFu) return;
    const bool ranged = q.wlo != 0 || q.whi != n - 1;
    uint32_t node = 0;
    for (uint32_t guard = 0; node != kNone; ++guard)
    {
        if (guard >= 2 * n)
        {
            atomicAdd(fail, 1ull); // malformed tree: reported, never loops
            return;
        }
        uint32_t next;
        if (node + 1 < n) // internal
        {
            float4 lo = t.inode_lo[node], hi = t.inode_hi[node];
            next = __float_as_uint(hi.w);
            double td;
            bool hit = ray_slab(q.o, q.inv, mk(lo.x, lo.y, lo.z), mk(hi.x, hi.y, hi.z), maxd(), td);
            if (hit && ranged) hit = node_in_range(t, node, q.wlo, q.whi);
            if (hit) next = __float_as_uint(lo.w);
        }
        else
        {
            const uint32_t j = node - (n - 1);
            next = n == 1 ? kNone : t.escape[node];
            if (j >= q.wlo && j <= q.whi)
            {
                const uint32_t body = t.leaf_body[j];
                const double *p = t.leaf_box + static_cast<size_t>(j) * 6;
                double td;
                if (b.is_static[body] != 2 &&
                    ray_slab(q.o, q.inv, mk(p[0], p[1], p[2]), mk(p[3], p[4], p[5]), maxd(), td))
                    leaf(body, td);
            }
        }
        node = next;
    }
}

// Walk topology for ordered ray walks: per internal node (left, right, parent, split
// axis). The split axis is the axis of the highest Morton bit that differs between the
// last leaf of the left child and the first of the right child (bit 3L+2 is x, 3L+1 y,
// 3L z); a ray enters the child on its own side of that split first.
__global__ void k_ray_nodes(lbvh_dev t, uint4 *rn)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = t.n;
    if (i + 1 >= n) return;
    uint32_t lc = static_cast<uint32_t>(t.left[i]), rc = static_cast<uint32_t>(t.right[i]);
    uint32_t g = lc >= n - 1 ? lc - (n - 1) : t.last[lc];
    uint64_t x = t.morton[g] ^ t.morton[g + 1];
    uint32_t axis = 0;
    if (x)
    {
        int bit = 63 - __clzll(static_cast<long long>(x));
        int wb = t.n_worlds > 1 ? t.wbits : 0;
        if (bit < kMortonBits - wb) axis = static_cast<uint32_t>(2 - (bit + wb) % 3);
    }
    rn[i] = make_uint4(lc, rc, t.parent[i], axis);
}

// Closest hit with an ordered stackless walk (Hapala, Davidovic, Wald, Havran,
// Slusallek, "Efficient Stack-less BVH Traversal for Ray Tracing", SCCG 2011): the
// near child first, moving through parent links and a three-state machine, so the
// first hits found are close and prune the rest of the walk with their distance.
__global__ void __launch_bounds__(kBlock)
    k_ray_closest(lbvh_dev t, const uint4 *rn, bodies_dev b, uint32_t nr, const double *org, const double *dir,
                  const double *maxd, const uint32_t *rw, uint32_t *out_id, double *out_d,
                  unsigned long long *fail)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    const qray q = load_ray(t, r, org, dir, maxd, rw);
    const uint32_t n = t.n;
    double best = q.maxd;
    uint32_t best_id = kNone;
    const bool neg[3] = {q.inv.x < 0.0, q.inv.y < 0.0, q.inv.z < 0.0};
    const bool ranged = q.wlo != 0 || q.whi != n - 1;
    // box test of a node (leaves: exact FP64, accepting the hit)
    auto visit = [&](uint32_t node) -> bool
    {
        double td;
        if (node + 1 < n)
        {
            float4 lo = t.inode_lo[node], hi = t.inode_hi[node];
            bool hit = ray_slab(q.o, q.inv, mk(lo.x, lo.y, lo.z), mk(hi.x, hi.y, hi.z), best, td);
            if (hit && ranged) hit = node_in_range(t, node, q.wlo, q.whi);
            return hit;
        }
        const uint32_t j = node - (n - 1);
        if (j >= q.wlo && j <= q.whi)
        {
            const uint32_t body = t.leaf_body[j];
            const double *p = t.leaf_box + static_cast<size_t>(j) * 6;
            if (b.is_static[body] != 2 && ray_slab(q.o, q.inv, mk(p[0], p[1], p[2]), mk(p[3], p[4], p[5]), best, td) &&
                (td < best || (td == best && body < best_id)))
            {
                best = td;
                best_id = body;
            }
        }
        return false; // a leaf has nothing below it
    };
    if (n == 1)
        visit(0);
    else if (n > 1 && q.whi >= q.wlo && q.whi != 0xFFFFFFFFu && visit(0))
    {
        enum
        {
            FROM_PARENT,
            FROM_SIBLING,
            FROM_CHILD
        } state = FROM_PARENT;
        uint4 pr = rn[0]; // the node whose children are being walked
        uint32_t cur = neg[pr.w] ? pr.y : pr.x;
        uint32_t par = 0;
        for (uint32_t guard = 0;; ++guard)
        {
            if (guard >= 6 * n)
            {
                atomicAdd(fail, 1ull); // malformed tree: reported, never loops
                break;
            }
            if (state == FROM_CHILD)
            {
                if (cur == 0) break; // back at the root: done
                par = cur + 1 < n ? rn[cur].z : t.parent[cur];
                pr = rn[par];
                uint32_t near = neg[pr.w] ? pr.y : pr.x, far = neg[pr.w] ? pr.x : pr.y;
                if (cur == near)
                {
                    cur = far;
                    state = FROM_SIBLING;
                }
                else
                    cur = par;
                continue;
            }
            const bool inside = visit(cur);
            if (!inside)
            {
                if (state == FROM_PARENT)
                {
                    cur = neg[pr.w] ? pr.x : pr.y; // the far sibling
                    state = FROM_SIBLING;
                }
                else
                {
                    cur = par;
                    state = FROM_CHILD;
                }
                continue;
            }
            par = cur; // an internal node the ray enters: walk its children, near first
            pr = rn[cur];
            cur = neg[pr.w] ? pr.y : pr.x;
            state = FROM_PARENT;
        }
    }
    out_id[r] = best_id;
    out_d[r] = best_id == kNone ? __builtin_inf() : best;
}

// every hit: FILL = false counts per ray, FILL = true writes (body, distance, ray) at
// the ray's offset in walk order
template <bool FILL>
__global__ void __launch_bounds__(kBlock)
    k_ray_all(lbvh_dev t, bodies_dev b, uint32_t nr, const double *org, const double *dir, const double *maxd,
              const uint32_t *rw, uint32_t *cnt, const uint32_t *off, uint32_t *hid, double *hd, uint32_t *hq,
              unsigned long long *fail)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    const qray q = load_ray(t, r, org, dir, maxd, rw);
    uint32_t k = FILL ? off[r] : 0u;
    const uint32_t end = FILL ? off[r + 1] : 0u;
    walk_ray(
        t, b, q,
        [&](uint32_t body, double d)
        {
            if (FILL)
            {
                if (k < end) // the fill walk repeats the counting walk; never past the ray's slots
                {
                    hid[k] = body;
                    hd[k] = d;
                    hq[k] = r;
                }
            }
            ++k;
        },
        [&]() { return q.maxd; }, fail);
    if (!FILL) cnt[r] = k;
}

// AABB query: leaves whose tree bounds intersect the box (bounds.h:87-92)
template <bool FILL>
__global__ void __launch_bounds__(kBlock)
    k_box_all(lbvh_dev t, bodies_dev b, uint32_t nq, const double *boxes, const uint32_t *qw, uint32_t *cnt,
              const uint32_t *off, uint32_t *hid, uint32_t *hq, unsigned long long *fail)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nq) return;
    const uint32_t n = t.n;
    box3 q;
    q.min = mk(boxes[6 * r], boxes[6 * r + 1], boxes[6 * r + 2]);
    q.max = mk(boxes[6 * r + 3], boxes[6 * r + 4], boxes[6 * r + 5]);
    float qlo[3], qhi[3];
    const double *qd = boxes + 6 * static_cast<size_t>(r);
    for (int c = 0; c < 3; ++c)
    {
        float f = static_cast<float>(qd[c]);
        if (static_cast<double>(f) > qd[c]) f = nextafterf(f, -__builtin_inff());
        qlo[c] = f;
        float g = static_cast<float>(qd[3 + c]);
        if (static_cast<double>(g) < qd[3 + c]) g = nextafterf(g, __builtin_inff());
        qhi[c] = g;
    }
    uint32_t wlo = 0, whi = n - 1;
    if (qw && t.n_worlds > 1)
    {
        if (qw[r] < t.n_worlds)
        {
            wlo = t.wstart[qw[r]];
            whi = t.wstart[qw[r] + 1] - 1;
        }
        else
        {
            wlo = 1;
            whi = 0;
        }
    }
    uint32_t k = FILL ? off[r] : 0u;
    const uint32_t end = FILL ? off[r + 1] : 0u;
    const bool ranged = wlo != 0 || whi != n - 1;
    uint32_t node = (n == 0 || whi < wlo || whi == 0xFFFFFFFFu) ? kNone : 0u;
    for (uint32_t guard = 0; node != kNone; ++guard)
    {
        if (guard >= 2 * n)
        {
            atomicAdd(fail, 1ull);
            break;
        }
        uint32_t next;
        if (node + 1 < n)
        {
            float4 lo = t.inode_lo[node], hi = t.inode_hi[node];
            bool ov = lo.x <= qhi[0] && hi.x >= qlo[0] && lo.y <= qhi[1] && hi.y >= qlo[1] && lo.z <= qhi[2] &&
                      hi.z >= qlo[2];
            if (ov && ranged) ov = node_in_range(t, node, wlo, whi);
            next = __float_as_uint(ov ? lo.w : hi.w);
        }
        else
        {
            const uint32_t j = node - (n - 1);
            next = n == 1 ? kNone : t.escape[node];
            if (j >= wlo && j <= whi)
            {
                const uint32_t body = t.leaf_body[j];
                const double *p = t.leaf_box + static_cast<size_t>(j) * 6;
                box3 lb;
                lb.min = mk(p[0], p[1], p[2]);
                lb.max = mk(p[3], p[4], p[5]);
                if (b.is_static[body] != 2 && box_intersects(lb, q))
                {
                    if (FILL && k < end)
                    {
                        hid[k] = body;
                        hq[k] = r;
                    }
                    ++k;
                }
            }
        }
        node = next;
    }
    if (!FILL) cnt[r] = k;
}

// sort keys: u32 or the bits of a non-negative double (monotone as unsigned integers),
// read through the current permutation
__global__ void k_key_u32(const uint32_t *src, const uint32_t *perm, uint32_t n, uint64_t *keys)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = src[perm ? perm[i] : i];
}
__global__ void k_key_f64(const double *src, const uint32_t *perm, uint32_t n, uint64_t *keys)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = static_cast<uint64_t>(__double_as_longlong(src[perm[i]]));
}
__global__ void k_iota(uint32_t *v, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}
__global__ void k_gather_hits(const uint32_t *perm, uint32_t n, const uint32_t *hid, const double *hd,
                              uint32_t *out_id, double *out_d)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t s = perm[i];
    out_id[i] = hid[s];
    if (hd) out_d[i] = hd[s];
}

int bits_for(uint64_t maxv)
{
    int b = 1;
    while (b < 64 && (1ull << b) <= maxv) ++b;
    return b;
}

// stable sort of the permutation by key (keys rebuilt by the caller); returns the
// buffer holding the sorted permutation
uint32_t *sort_perm(query_scratch &qs, uint32_t m, int bits, sort_scratch &sc, hipStream_t st)
{
    int shifts[8];
    int ns = 0;
    for (int s = 0; s < bits && ns < 8; s += 8) shifts[ns++] = s;
    bool alt = radix_sort_u64(qs.keys, qs.perm, qs.keys_alt, qs.perm_alt, m, shifts, ns, sc, st);
    if (alt)
    {
        std::swap(qs.keys, qs.keys_alt);
        std::swap(qs.perm, qs.perm_alt);
    }
    return qs.perm;
}
} // namespace

void launch_ray_nodes(const lbvh_dev &t, uint4 *rn, hipStream_t st)
{
    if (t.n >= 2) k_ray_nodes<<<grid(t.n - 1), kBlock, 0, st>>>(t, rn);
}
void launch_ray_closest(const lbvh_dev &t, const uint4 *rn, const bodies_dev &b, uint32_t nr, const double *org,
                        const double *dir, const double *maxd, const uint32_t *rw, uint32_t *out_id, double *out_d,
                        unsigned long long *fail, hipStream_t st)
{
    if (nr) k_ray_closest<<<grid(nr), kBlock, 0, st>>>(t, rn, b, nr, org, dir, maxd, rw, out_id, out_d, fail);
}
void launch_ray_count(const lbvh_dev &t, const bodies_dev &b, uint32_t nr, const double *org, const double *dir,
                      const double *maxd, const uint32_t *rw, uint32_t *cnt, unsigned long long *fail, hipStream_t st)
{
    if (nr)
        k_ray_all<false><<<grid(nr), kBlock, 0, st>>>(t, b, nr, org, dir, maxd, rw, cnt, nullptr, nullptr, nullptr,
                                                       nullptr, fail);
}
void launch_ray_fill(const lbvh_dev &t, const bodies_dev &b, uint32_t nr, const double *org, const double *dir,
                     const double *maxd, const uint32_t *rw, const uint32_t *off, uint32_t *hid, double *hd,
                     uint32_t *hq, unsigned long long *fail, hipStream_t st)
{
    if (nr)
        k_ray_all<true><<<grid(nr), kBlock, 0, st>>>(t, b, nr, org, dir, maxd, rw, nullptr, off, hid, hd, hq, fail);
}
void launch_box_count(const lbvh_dev &t, const bodies_dev &b, uint32_t nq, const double *boxes, const uint32_t *qw,
                      uint32_t *cnt, unsigned long long *fail, hipStream_t st)
{
    if (nq) k_box_all<false><<<grid(nq), kBlock, 0, st>>>(t, b, nq, boxes, qw, cnt, nullptr, nullptr, nullptr, fail);
}
void launch_box_fill(const lbvh_dev &t, const bodies_dev &b, uint32_t nq, const double *boxes, const uint32_t *qw,
                     const uint32_t *off, uint32_t *hid, uint32_t *hq, unsigned long long *fail, hipStream_t st)
{
    if (nq) k_box_all<true><<<grid(nq), kBlock, 0, st>>>(t, b, nq, boxes, qw, nullptr, off, hid, hq, fail);
}

void launch_hits_sort(query_scratch &qs, uint32_t m, uint32_t n_bodies, uint32_t n_queries, bool by_distance,
                      uint32_t *out_id, double *out_d, sort_scratch &sc, hipStream_t st)
{
    if (!m) return;
    // stable passes from the least significant criterion: body, then distance, then query
    k_key_u32<<<grid(m), kBlock, 0, st>>>(qs.hid, nullptr, m, qs.keys);
    k_iota<<<grid(m), kBlock, 0, st>>>(qs.perm, m);
    uint32_t *perm = sort_perm(qs, m, bits_for(n_bodies), sc, st);
    if (by_distance)
    {
        k_key_f64<<<grid(m), kBlock, 0, st>>>(qs.hd, perm, m, qs.keys);
        perm = sort_perm(qs, m, 64, sc, st);
    }
    k_key_u32<<<grid(m), kBlock, 0, st>>>(qs.hq, perm, m, qs.keys);
    perm = sort_perm(qs, m, bits_for(n_queries), sc, st);
    k_gather_hits<<<grid(m), kBlock, 0, st>>>(perm, m, qs.hid, by_distance ? qs.hd : nullptr, out_id, out_d);
}

} // namespace pk
