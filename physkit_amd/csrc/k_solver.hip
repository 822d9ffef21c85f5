// Contact solver: constraint_solver::setup_contacts / solve_constraints
// (include/physkit/collision/constraint.h:1052-1201, build_contact_jacobian :874-953,
// jacobian_row :29-87). The reference runs projected Gauss-Seidel sequentially over
// all contact points. Here the manifolds are coloured so that no two manifolds of a
// colour share a dynamic body; within a colour every manifold is independent, so one
// thread per manifold reproduces, bit for bit, a sequential sweep in colour order.
// Colouring: each round, every uncoloured manifold whose priority (mix32(p) << 32 | p)
// is the minimum claim at each of its dynamic bodies takes the smallest colour free at
// those bodies.
#include "pk_access.hpp"

#include <cfloat>
#include <type_traits>

namespace pk
{
namespace
{
constexpr int kBlock = 256;
unsigned grid(size_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

struct rowv
{
    v3 Jv, Jwa, Jwb;
    double Meff, bias, acc;
};
struct body_state
{
    v3 vel, ang;
    double im;
    m3 Iwi;
    bool stat;
};

__device__ __forceinline__ double row_jv(const rowv &r, const body_state &A, const body_state &B)
{ return ((dot(r.Jv, A.vel) + dot(r.Jwa, A.ang)) - dot(r.Jv, B.vel)) + dot(r.Jwb, B.ang); }

__device__ __forceinline__ void row_apply(rowv &r, double lo, double hi, body_state &A, body_state &B) // :49-72
{
    double jv = row_jv(r, A, B);
    double cfp = 0.0 * r.acc;
    double lambda = -((jv + r.bias) + cfp) * r.Meff;
    double old = r.acc;
    double v = old + lambda;
    r.acc = (v < lo) ? lo : ((hi < v) ? hi : v);
    double ap = r.acc - old;
    if (!A.stat)
    {
        A.vel = A.vel + (r.Jv * ap) * A.im;
        A.ang = A.ang + mul(A.Iwi, r.Jwa * ap);
    }
    if (!B.stat)
    {
        B.vel = B.vel + ((-r.Jv) * ap) * B.im;
        B.ang = B.ang + mul(B.Iwi, r.Jwb * ap);
    }
}

__device__ __forceinline__ void row_warm(const rowv &r, body_state &A, body_state &B) // :74-87
{
    if (r.acc == 0.0) return;
    if (!A.stat)
    {
        A.vel = A.vel + (r.Jv * r.acc) * A.im;
        A.ang = A.ang + mul(A.Iwi, r.Jwa * r.acc);
    }
    if (!B.stat)
    {
        B.vel = B.vel + ((-r.Jv) * r.acc) * B.im;
        B.ang = B.ang + mul(B.Iwi, r.Jwb * r.acc);
    }
}

__device__ __forceinline__ void friction_block(rowv &t1, rowv &t2, double mu, double nacc, double i11,
                                               double i12, double i22, body_state &A, body_state &B) // :1131-1187
{
    double maxf = mu * nacc;
    double jv1 = row_jv(t1, A, B);
    double jv2 = row_jv(t2, A, B);
    double b1 = jv1 + 0.0;
    double b2 = jv2 + 0.0;
    double l1 = -(i11 * b1 + i12 * b2);
    double l2 = -(i12 * b1 + i22 * b2);
    double o1 = t1.acc, o2 = t2.acc;
    double n1 = o1 + l1, n2 = o2 + l2;
    double mag = hypot_pk(n1, n2);
    if (mag > maxf)
    {
        n1 = n1 * (maxf / mag);
        n2 = n2 * (maxf / mag);
    }
    t1.acc = n1;
    t2.acc = n2;
    double a1 = n1 - o1, a2 = n2 - o2;
    if (!A.stat)
    {
        A.vel = A.vel + (t1.Jv * a1 + t2.Jv * a2) * A.im;
        A.ang = A.ang + mul(A.Iwi, t1.Jwa * a1 + t2.Jwa * a2);
    }
    if (!B.stat)
    {
        B.vel = B.vel + ((-t1.Jv) * a1 - t2.Jv * a2) * B.im;
        B.ang = B.ang + mul(B.Iwi, t1.Jwb * a1 + t2.Jwb * a2);
    }
}

__device__ __forceinline__ void basis(v3 n, v3 &t1, v3 &t2) // :104-113
{
    double sign = copysign(1.0, n.z);
    double a = -1.0 / (sign + n.z);
    double b = n.x * n.y * a;
    t1 = mk(1.0 + sign * n.x * n.x * a, sign * b, -sign * n.x);
    t2 = mk(b, sign + n.y * n.y * a, -n.y);
}

__device__ __forceinline__ rowv make_row(const body_state &A, const body_state &B, v3 dir, v3 ra, v3 rb)
{
    rowv r;
    r.Jv = dir;
    r.Jwa = cross(ra, dir);
    r.Jwb = -cross(rb, dir);
    v3 ka = mul(A.Iwi, r.Jwa);
    v3 kb = mul(B.Iwi, r.Jwb);
    r.Meff = 1.0 / (((A.im + B.im) + dot(r.Jwa, ka)) + dot(r.Jwb, kb));
    r.bias = 0.0;
    r.acc = 0.0;
    return r;
}

// Body velocities seen by a manifold solve: in device memory by body id (vel_global),
// or in a workgroup's LDS table by slot for its dynamic bodies (vel_lds; static bodies,
// never written, are read from device memory).
constexpr uint16_t kNoSlot = 0xFFFF;
struct vel_global
{
    __device__ __forceinline__ void load(const bodies_dev &b, uint32_t id, uint16_t, v3 &v, v3 &w) const
    {
        v = ld3(b.vel, b.n, id);
        w = ld3(b.ang, b.n, id);
    }
    __device__ __forceinline__ void store(const bodies_dev &b, uint32_t id, uint16_t, v3 v, v3 w) const
    {
        st3(b.vel, b.n, id, v);
        st3(b.ang, b.n, id, w);
    }
};
struct vel_lds
{
    double (*L)[6];
    __device__ __forceinline__ void load(const bodies_dev &b, uint32_t id, uint16_t slot, v3 &v, v3 &w) const
    {
        if (slot == kNoSlot)
        {
            v = ld3(b.vel, b.n, id);
            w = ld3(b.ang, b.n, id);
            return;
        }
        v = mk(L[slot][0], L[slot][1], L[slot][2]);
        w = mk(L[slot][3], L[slot][4], L[slot][5]);
    }
    __device__ __forceinline__ void store(const bodies_dev &, uint32_t, uint16_t slot, v3 v, v3 w) const
    {
        L[slot][0] = v.x;
        L[slot][1] = v.y;
        L[slot][2] = v.z;
        L[slot][3] = w.x;
        L[slot][4] = w.y;
        L[slot][5] = w.z;
    }
};

// inverse mass, inverse world inertia and the static flag (velocities come separately)
__device__ __forceinline__ body_state ld_body(const bodies_dev &b, uint32_t i)
{
    body_state s;
    s.vel = mk(0.0, 0.0, 0.0);
    s.ang = mk(0.0, 0.0, 0.0);
    s.im = b.inv_mass[i];
    s.Iwi = ld9(b.Iw_inv, b.n, i);
    s.stat = b.is_static[i] != 0;
    return s;
}

// row field planes
enum
{
    F_NJV = 0, F_NJWA = 3, F_NJWB = 6, F_NMEFF = 9, F_NBIAS = 10, F_NACC = 11,
    F_T1JV = 12, F_T1JWA = 15, F_T1JWB = 18, F_T1MEFF = 21, F_T1ACC = 22,
    F_T2JV = 23, F_T2JWA = 26, F_T2JWB = 29, F_T2MEFF = 32, F_T2ACC = 33,
    F_MU = 34, F_I11 = 35, F_I12 = 36, F_I22 = 37,
    F_RA = 38, F_RB = 41 // lever arms: the island solve rebuilds the Jacobians from them
};
__device__ __forceinline__ double &rf(const rows_dev &r, int f, uint32_t slot)
{ return r.d[static_cast<size_t>(f) * r.cap4 + slot]; }
__device__ __forceinline__ v3 rf3(const rows_dev &r, int f, uint32_t slot)
{ return mk(rf(r, f, slot), rf(r, f + 1, slot), rf(r, f + 2, slot)); }

// Contact c of manifold p: world normal n and lever arms as build_contact_jacobian
// forms them (:883-890) and whether the contact makes rows (a penetration that is not
// <= 0, so a NaN one does too, as in the reference's `if (penetration <= 0) continue`).
__device__ __forceinline__ bool contact_frame(const manifolds_dev &m, uint32_t sl, int c, v3 pa, v3 pb, qt qa,
                                              qt qb, v3 &n, v3 &ra, v3 &rb, double &pen)
{
    const double *d = m.data + static_cast<size_t>(c * kMF) * m.dcap + sl;
    const size_t s = m.dcap;
    const v3 nw = normalized(mk(d[0], d[s], d[2 * s]));
    const v3 la = mk(d[3 * s], d[4 * s], d[5 * s]);
    const v3 lb = mk(d[6 * s], d[7 * s], d[8 * s]);
    const v3 lnb = rotate(conj(qb), nw);
    ra = rotate(qa, la);
    rb = rotate(qb, lb);
    n = rotate(qb, lnb);
    pen = dot((pb + rb) - (pa + ra), n);
    return !(pen <= 0.0);
}

// which manifolds make solver rows: each is listed and its colour[p] set to kUncoloured
// (no other pair's colour is read), and the row count; the rows themselves are written
// by k_setup once the solve order is known
__global__ void __launch_bounds__(kBlock)
    k_row_flags(const uint64_t *keys, bodies_dev b, manifolds_dev m, uint32_t *colour, uint32_t *list,
                unsigned long long *counters)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n_slots = static_cast<uint32_t>(counters[C_MSLOTS]);
    uint32_t nrows = 0, p = 0;
    if (s < n_slots && s < m.dcap)
    {
        p = m.owner[s];
        const int cnt = m.count[p];
        const uint64_t key = keys[p];
        const uint32_t ia = static_cast<uint32_t>(key >> 32), ib = static_cast<uint32_t>(key & 0xFFFFFFFFu);
        const v3 pa = ld3(b.pos, b.n, ia), pb = ld3(b.pos, b.n, ib);
        const qt qa = ldq(b.quat, b.n, ia), qb = ldq(b.quat, b.n, ib);
        for (int c = 0; c < cnt; ++c)
        {
            v3 n, ra, rb;
            double pen;
            if (contact_frame(m, s, c, pa, pb, qa, qb, n, ra, rb, pen)) ++nrows;
        }
        if (nrows) colour[p] = kUncoloured;
    }
    block_add(&counters[C_ROWS], nrows);
    const unsigned long long at = block_append(&counters[C_NLIST], nrows > 0);
    if (nrows) list[at] = p;
}

// Rows of manifold ord[k] in solve order: contact c at slot c * cap + k of every row
// plane, so the lanes solving one colour read each field at consecutive addresses.
// Without ord, manifold k's contact c at slot 4k + c (pair order).
__device__ __forceinline__ uint32_t row_slot(const uint32_t *ord, uint32_t cap, uint32_t k, int c)
{ return ord ? static_cast<uint32_t>(c) * cap + k : 4 * k + static_cast<uint32_t>(c); }

// row planes of element type T (double: the FP64 solve; float: the FP32 solve, whose
// rows are computed in FP64 and rounded once when stored)
template <class T> struct rows_t
{
    uint32_t cap4;
    T *d;
    uint8_t *valid;
};
template <class T> __host__ __device__ __forceinline__ rows_t<T> rows_as(const rows_dev &r)
{
    rows_t<T> o;
    o.cap4 = r.cap4;
    o.d = reinterpret_cast<T *>(r.d);
    o.valid = r.valid;
    return o;
}
template <class T> __device__ __forceinline__ void put(const rows_t<T> &r, int f, uint32_t slot, double v)
{ r.d[static_cast<size_t>(f) * r.cap4 + slot] = static_cast<T>(v); }
template <class T> __device__ __forceinline__ void put3(const rows_t<T> &r, int f, uint32_t slot, v3 v)
{
    put(r, f, slot, v.x);
    put(r, f + 1, slot, v.y);
    put(r, f + 2, slot, v.z);
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_setup(const uint64_t *keys, uint32_t n, const uint32_t *ord, uint32_t cap, bodies_dev b, manifolds_dev m,
            rows_t<T> r, double dt, double thr)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
    {
        const uint32_t p = ord ? ord[k] : k;
        for (int c = 0; c < 4; ++c) r.valid[row_slot(ord, cap, k, c)] = 0;
        int cnt = m.count[p];
        if (cnt > 0)
        {
            uint64_t key = keys[p];
            uint32_t ia = static_cast<uint32_t>(key >> 32), ib = static_cast<uint32_t>(key & 0xFFFFFFFFu);
            body_state A = ld_body(b, ia), B = ld_body(b, ib);
            vel_global().load(b, ia, kNoSlot, A.vel, A.ang);
            vel_global().load(b, ib, kNoSlot, B.vel, B.ang);
            v3 pa = ld3(b.pos, b.n, ia), pb = ld3(b.pos, b.n, ib);
            qt qa = ldq(b.quat, b.n, ia), qb = ldq(b.quat, b.n, ib);
            double e = smax(b.restitution[ia], b.restitution[ib]);
            double mu = sqrt(b.friction[ia] * b.friction[ib]);
            const uint32_t sl = m.slot[p];
            for (int c = 0; c < cnt; ++c)
            {
                const double *d = m.data + static_cast<size_t>(c * kMF) * m.dcap + sl;
                const size_t s = m.dcap;
                // build_contact_jacobian, :874-953
                v3 n, ra, rb;
                double pen;
                if (!contact_frame(m, sl, c, pa, pb, qa, qb, n, ra, rb, pen)) continue;
                rowv nr = make_row(A, B, n, ra, rb);
                v3 vca = A.vel + cross(A.ang, ra);
                v3 vcb = B.vel + cross(B.ang, rb);
                double vrel = dot(vca - vcb, n);
                double rbias = 0.0;
                if (vrel < -thr) rbias = e * vrel;
                double baum = (-0.1 / dt) * smax(0.0, pen - 0.005);
                nr.bias = smin(baum, rbias);
                v3 t1, t2;
                basis(n, t1, t2);
                rowv r1 = make_row(A, B, t1, ra, rb);
                rowv r2 = make_row(A, B, t2, ra, rb);
                // setup_contacts, :1075-1102
                double m11 = 1.0 / r1.Meff, m22 = 1.0 / r2.Meff;
                double m12 = dot(r1.Jwa, mul(A.Iwi, r2.Jwa)) + dot(r1.Jwb, mul(B.Iwi, r2.Jwb));
                double det = m11 * m22 - m12 * m12;
                double i11, i12, i22;
                if (det > 0.0)
                {
                    i11 = m22 / det;
                    i22 = m11 / det;
                    i12 = -m12 / det;
                }
                else
                {
                    i11 = r1.Meff;
                    i22 = r2.Meff;
                    i12 = 0.0;
                }
                const uint32_t slot = row_slot(ord, cap, k, c);
                put3(r, F_NJV, slot, nr.Jv);
                put3(r, F_NJWA, slot, nr.Jwa);
                put3(r, F_NJWB, slot, nr.Jwb);
                put(r, F_NMEFF, slot, nr.Meff);
                put(r, F_NBIAS, slot, nr.bias);
                put(r, F_NACC, slot, d[10 * s]);
                put3(r, F_T1JV, slot, r1.Jv);
                put3(r, F_T1JWA, slot, r1.Jwa);
                put3(r, F_T1JWB, slot, r1.Jwb);
                put(r, F_T1MEFF, slot, r1.Meff);
                put(r, F_T1ACC, slot, d[11 * s]);
                put3(r, F_T2JV, slot, r2.Jv);
                put3(r, F_T2JWA, slot, r2.Jwa);
                put3(r, F_T2JWB, slot, r2.Jwb);
                put(r, F_T2MEFF, slot, r2.Meff);
                put(r, F_T2ACC, slot, d[12 * s]);
                put(r, F_MU, slot, mu);
                put(r, F_I11, slot, i11);
                put(r, F_I12, slot, i12);
                put(r, F_I22, slot, i22);
                put3(r, F_RA, slot, ra);
                put3(r, F_RB, slot, rb);
                r.valid[slot] = 1;
            }
        }
    }
}

// Claim of a manifold in a colouring round: the round (from 1) above the complement of
// mix32(p). mix32 is a bijection, so ordering by it equals ordering by the priority
// (mix32(p) << 32) | p; an atomic max keeps this round's lowest priority at each body,
// and claims left from earlier rounds lose to any claim of the current one, so the
// claim array is cleared once per colouring, not every round.
__device__ __forceinline__ uint64_t claim_of(uint32_t p, uint32_t round)
{ return (static_cast<uint64_t>(round + 1u) << 32) | (0xFFFFFFFFu - mix32(p)); }

__device__ __forceinline__ uint32_t local_index(uint32_t p, uint32_t ia, const uint32_t *world_of,
                                                const uint32_t *pair_base)
{ return pair_base ? p - pair_base[world_of[ia]] : p; }

// first pair of each world in the sorted pair list (worlds without pairs are not written)
__global__ void k_pair_base(const uint64_t *keys, uint32_t n_pairs, const uint32_t *world_of, uint32_t *pair_base)
{
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pairs) return;
    uint32_t w = world_of[static_cast<uint32_t>(keys[p] >> 32)];
    if (p == 0 || world_of[static_cast<uint32_t>(keys[p - 1] >> 32)] != w) pair_base[w] = p;
}

// the items coloured: list[0, n) (or 0..n-1 when list is null); n_dev, when given,
// holds the list's count on the device (n then bounds it)
__device__ __forceinline__ uint32_t colour_count(uint32_t n, const unsigned long long *n_dev)
{
    return n_dev ? static_cast<uint32_t>(min(static_cast<unsigned long long>(n), *n_dev)) : n;
}

__global__ void k_claim(const uint64_t *keys, const uint32_t *list, uint32_t n, const unsigned long long *n_dev,
                        bodies_dev b, const uint32_t *colour, const uint32_t *world_of, const uint32_t *pair_base,
                        unsigned long long *claim, uint32_t round)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= colour_count(n, n_dev)) return;
    uint32_t p = list ? list[i] : i;
    if (colour[p] != kUncoloured) return;
    uint64_t key = keys[p];
    uint32_t ia = static_cast<uint32_t>(key >> 32), ib = static_cast<uint32_t>(key & 0xFFFFFFFFu);
    unsigned long long pr = claim_of(local_index(p, ia, world_of, pair_base), round);
    if (!b.is_static[ia]) atomicMax(&claim[ia], pr);
    if (!b.is_static[ib]) atomicMax(&claim[ib], pr);
}

// the row manifolds (colour kUncoloured after contact setup), listed for the rounds
// A manifold holding the claim at each of its dynamic bodies takes the smallest colour
// that no coloured manifold at those bodies holds (`used`: colours 0-63 taken at each
// body), or 64 + the round number when all 64 are taken. Manifolds coloured in one
// round share no dynamic body, so their choices are independent.
__global__ void k_assign(const uint64_t *keys, const uint32_t *list, uint32_t n_bound, const unsigned long long *n_dev,
                         bodies_dev b, uint32_t *colour, const uint32_t *world_of, const uint32_t *pair_base,
                         const unsigned long long *claim, unsigned long long *used, uint32_t round,
                         unsigned long long *counters)
{
    const uint32_t n = colour_count(n_bound, n_dev);
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool remain = false;
    uint32_t ncol = 0;
    uint32_t p = i < n ? (list ? list[i] : i) : 0u;
    if (i < n && colour[p] == kUncoloured)
    {
        uint64_t key = keys[p];
        uint32_t ia = static_cast<uint32_t>(key >> 32), ib = static_cast<uint32_t>(key & 0xFFFFFFFFu);
        unsigned long long pr = claim_of(local_index(p, ia, world_of, pair_base), round);
        const bool sa = b.is_static[ia] != 0, sb = b.is_static[ib] != 0;
        bool ok = (sa || claim[ia] == pr) && (sb || claim[ib] == pr);
        if (ok)
        {
            unsigned long long u = (sa ? 0ull : used[ia]) | (sb ? 0ull : used[ib]);
            uint32_t k = ~u ? static_cast<uint32_t>(__ffsll(static_cast<long long>(~u)) - 1) : 64u + round;
            colour[p] = k;
            if (k < 64)
            {
                if (!sa) atomicOr(&used[ia], 1ull << k);
                if (!sb) atomicOr(&used[ib], 1ull << k);
            }
            ncol = k + 1;
        }
        else
            remain = true;
    }
    block_add(&counters[C_REMAIN], remain ? 1u : 0u);
    block_max(&counters[C_NCOL], ncol);
}

// colour buckets by a radix sort of the listed items' (colour, index) keys: each
// colour's items in index order (list null: items 0..nl-1)
__global__ void k_colour_keys(const uint32_t *colour, const uint32_t *list, uint32_t nl, uint64_t *keys,
                              uint32_t *vals)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nl) return;
    const uint32_t p = list ? list[i] : i;
    keys[i] = (static_cast<uint64_t>(colour[p]) << 32) | p;
    vals[i] = p;
}

// start[c]: the first sorted key of colour c or above (start[n_colours] = n)
__global__ void k_colour_bounds(const uint64_t *sorted, uint32_t n, uint32_t n_colours, uint32_t *start)
{
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > n_colours) return;
    const uint64_t first = static_cast<uint64_t>(c) << 32;
    uint32_t lo = 0, hi = n;
    while (lo < hi)
    {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (sorted[mid] < first)
            lo = mid + 1;
        else
            hi = mid;
    }
    start[c] = lo;
}

__device__ __forceinline__ rowv ld_row(const rows_dev &r, int f_jv, int f_meff, int f_bias, int f_acc, uint32_t slot)
{
    rowv v;
    v.Jv = rf3(r, f_jv, slot);
    v.Jwa = rf3(r, f_jv + 3, slot);
    v.Jwb = rf3(r, f_jv + 6, slot);
    v.Meff = rf(r, f_meff, slot);
    v.bias = f_bias >= 0 ? rf(r, f_bias, slot) : 0.0;
    v.acc = rf(r, f_acc, slot);
    return v;
}

// the manifold at position t of the solve order (rows written in that order by k_setup)
template <class V = vel_global>
__device__ __forceinline__ void solve_one(const uint64_t *keys, const uint32_t *order, uint32_t t, const bodies_dev &b,
                                          const rows_dev &r, uint32_t cap, int warm, const V &vs = V(),
                                          uint16_t sa = kNoSlot, uint16_t sb = kNoSlot)
{
    uint32_t p = order[t];
    uint64_t key = keys[p];
    uint32_t ia = static_cast<uint32_t>(key >> 32), ib = static_cast<uint32_t>(key & 0xFFFFFFFFu);
    body_state A = ld_body(b, ia), B = ld_body(b, ib);
    vs.load(b, ia, sa, A.vel, A.ang);
    vs.load(b, ib, sb, B.vel, B.ang);
    for (int c = 0; c < 4; ++c)
    {
        uint32_t slot = static_cast<uint32_t>(c) * cap + t;
        if (!r.valid[slot]) continue;
        rowv nr = ld_row(r, F_NJV, F_NMEFF, F_NBIAS, F_NACC, slot);
        rowv t1 = ld_row(r, F_T1JV, F_T1MEFF, -1, F_T1ACC, slot);
        rowv t2 = ld_row(r, F_T2JV, F_T2MEFF, -1, F_T2ACC, slot);
        if (warm)
        {
            row_warm(nr, A, B);
            row_warm(t1, A, B);
            row_warm(t2, A, B);
        }
        else
        {
            row_apply(nr, 0.0, DBL_MAX, A, B);
            friction_block(t1, t2, rf(r, F_MU, slot), nr.acc, rf(r, F_I11, slot), rf(r, F_I12, slot),
                           rf(r, F_I22, slot), A, B);
            rf(r, F_NACC, slot) = nr.acc;
            rf(r, F_T1ACC, slot) = t1.acc;
            rf(r, F_T2ACC, slot) = t2.acc;
        }
    }
    if (!A.stat) vs.store(b, ia, sa, A.vel, A.ang);
    if (!B.stat) vs.store(b, ib, sb, B.vel, B.ang);
}

// one colour: positions [begin, end) of the solve order
__global__ void __launch_bounds__(kBlock)
    k_solve(const uint64_t *keys, const uint32_t *order, uint32_t begin, uint32_t end, bodies_dev b, rows_dev r,
            uint32_t cap, int warm)
{
    uint32_t t = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (t < end) solve_one(keys, order, t, b, r, cap, warm);
}

// ---- FP32 solve with FP64 per-body velocity accumulation ----
// Rows are computed in FP64 by k_setup<float> and stored once as FP32 in solve order;
// the row arithmetic of the solve (J.v, lambda, clamping, the 2x2 friction block) runs
// in FP32 and each body's velocity increment is added to its FP64 velocity.
struct f3
{
    float x, y, z;
};
__device__ __forceinline__ f3 f3m(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 to_f3(v3 a) { return {static_cast<float>(a.x), static_cast<float>(a.y), static_cast<float>(a.z)}; }
__device__ __forceinline__ float dotf(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ f3 scl(f3 a, float k) { return {a.x * k, a.y * k, a.z * k}; }
__device__ __forceinline__ f3 addf(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3 widen(f3 a) { return mk(static_cast<double>(a.x), static_cast<double>(a.y), static_cast<double>(a.z)); }

struct rows32_dev
{
    uint32_t cap4;
    float *d;
    uint8_t *valid;
};
// FP32 rows live in a rows_dev buffer whose planes hold floats
__host__ __device__ __forceinline__ rows32_dev as32(const rows_dev &r)
{
    rows32_dev o;
    o.cap4 = r.cap4;
    o.d = reinterpret_cast<float *>(r.d);
    o.valid = r.valid;
    return o;
}
__device__ __forceinline__ float &rg(const rows32_dev &r, int f, uint32_t slot)
{ return r.d[static_cast<size_t>(f) * r.cap4 + slot]; }
__device__ __forceinline__ f3 rg3(const rows32_dev &r, int f, uint32_t slot)
{ return f3m(rg(r, f, slot), rg(r, f + 1, slot), rg(r, f + 2, slot)); }

struct body32
{
    v3 vel, ang; // FP64 accumulators
    float im;
    float I[9];
    bool stat;
};
__device__ __forceinline__ f3 mulI(const float *I, f3 v)
{ return f3m((I[0] * v.x + I[1] * v.y) + I[2] * v.z, (I[3] * v.x + I[4] * v.y) + I[5] * v.z, (I[6] * v.x + I[7] * v.y) + I[8] * v.z); }
__device__ __forceinline__ body32 ld_body32(const bodies_dev &b, uint32_t i)
{
    body32 s;
    s.im = static_cast<float>(b.inv_mass[i]);
    m3 I = ld9(b.Iw_inv, b.n, i);
    for (int k = 0; k < 9; ++k) s.I[k] = static_cast<float>(I.m[k]);
    s.stat = b.is_static[i] != 0;
    return s;
}
__device__ __forceinline__ float jv32(f3 Jv, f3 Jwa, f3 Jwb, const body32 &A, const body32 &B)
{ return ((dotf(Jv, to_f3(A.vel)) + dotf(Jwa, to_f3(A.ang))) - dotf(Jv, to_f3(B.vel))) + dotf(Jwb, to_f3(B.ang)); }
// impulse a along a row (or two rows): FP32 increments, FP64 accumulation
__device__ __forceinline__ void push32(body32 &A, body32 &B, f3 lin, f3 angA, f3 angB)
{
    if (!A.stat)
    {
        A.vel = A.vel + widen(scl(lin, A.im));
        A.ang = A.ang + widen(mulI(A.I, angA));
    }
    if (!B.stat)
    {
        B.vel = B.vel - widen(scl(lin, B.im));
        B.ang = B.ang + widen(mulI(B.I, angB));
    }
}

template <class V = vel_global>
__device__ __forceinline__ void solve_one32(const uint64_t *keys, const uint32_t *order, uint32_t t, const bodies_dev &b,
                                            const rows32_dev &r, uint32_t cap, int warm, const V &vs = V(),
                                            uint16_t sa = kNoSlot, uint16_t sb = kNoSlot)
{
    uint32_t p = order[t];
    uint64_t key = keys[p];
    uint32_t ia = static_cast<uint32_t>(key >> 32), ib = static_cast<uint32_t>(key & 0xFFFFFFFFu);
    body32 A = ld_body32(b, ia), B = ld_body32(b, ib);
    vs.load(b, ia, sa, A.vel, A.ang);
    vs.load(b, ib, sb, B.vel, B.ang);
    for (int c = 0; c < 4; ++c)
    {
        uint32_t slot = static_cast<uint32_t>(c) * cap + t;
        if (!r.valid[slot]) continue;
        f3 nJv = rg3(r, F_NJV, slot), nJa = rg3(r, F_NJWA, slot), nJb = rg3(r, F_NJWB, slot);
        f3 aJv = rg3(r, F_T1JV, slot), aJa = rg3(r, F_T1JWA, slot), aJb = rg3(r, F_T1JWB, slot);
        f3 bJv = rg3(r, F_T2JV, slot), bJa = rg3(r, F_T2JWA, slot), bJb = rg3(r, F_T2JWB, slot);
        float nacc = rg(r, F_NACC, slot), acc1 = rg(r, F_T1ACC, slot), acc2 = rg(r, F_T2ACC, slot);
        if (warm)
        {
            if (nacc != 0.0f) push32(A, B, scl(nJv, nacc), scl(nJa, nacc), scl(nJb, nacc));
            if (acc1 != 0.0f) push32(A, B, scl(aJv, acc1), scl(aJa, acc1), scl(aJb, acc1));
            if (acc2 != 0.0f) push32(A, B, scl(bJv, acc2), scl(bJa, acc2), scl(bJb, acc2));
            continue;
        }
        // normal row, lambda >= 0 (constraint.h:49-72)
        float jv = jv32(nJv, nJa, nJb, A, B);
        float lambda = -(jv + rg(r, F_NBIAS, slot)) * rg(r, F_NMEFF, slot);
        float nv = nacc + lambda;
        nv = nv < 0.0f ? 0.0f : nv;
        float ap = nv - nacc;
        nacc = nv;
        push32(A, B, scl(nJv, ap), scl(nJa, ap), scl(nJb, ap));
        // friction block on the cone (constraint.h:1131-1187)
        float maxf = rg(r, F_MU, slot) * nacc;
        float b1 = jv32(aJv, aJa, aJb, A, B), b2 = jv32(bJv, bJa, bJb, A, B);
        float i11 = rg(r, F_I11, slot), i12 = rg(r, F_I12, slot), i22 = rg(r, F_I22, slot);
        float n1 = acc1 - (i11 * b1 + i12 * b2), n2 = acc2 - (i12 * b1 + i22 * b2);
        float mag = sqrtf(n1 * n1 + n2 * n2);
        if (mag > maxf)
        {
            n1 = n1 * (maxf / mag);
            n2 = n2 * (maxf / mag);
        }
        float a1 = n1 - acc1, a2 = n2 - acc2;
        acc1 = n1;
        acc2 = n2;
        push32(A, B, addf(scl(aJv, a1), scl(bJv, a2)), addf(scl(aJa, a1), scl(bJa, a2)), addf(scl(aJb, a1), scl(bJb, a2)));
        rg(r, F_NACC, slot) = nacc;
        rg(r, F_T1ACC, slot) = acc1;
        rg(r, F_T2ACC, slot) = acc2;
    }
    if (!A.stat) vs.store(b, ia, sa, A.vel, A.ang);
    if (!B.stat) vs.store(b, ib, sb, B.vel, B.ang);
}

__global__ void __launch_bounds__(kBlock)
    k_solve32(const uint64_t *keys, const uint32_t *order, uint32_t begin, uint32_t end, bodies_dev b, rows32_dev r,
              uint32_t cap, int warm)
{
    uint32_t t = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (t < end) solve_one32(keys, order, t, b, r, cap, warm);
}

__global__ void k_writeback32(const uint32_t *order, uint32_t n_act, manifolds_dev m, rows32_dev r, uint32_t cap)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_act) return;
    uint32_t p = order[t];
    for (int c = 0; c < 4; ++c)
    {
        uint32_t slot = static_cast<uint32_t>(c) * cap + t;
        if (!r.valid[slot]) continue;
        double *d = m.data + static_cast<size_t>(c * kMF) * m.dcap + m.slot[p];
        size_t s = m.dcap;
        d[10 * s] = static_cast<double>(rg(r, F_NACC, slot));
        d[11 * s] = static_cast<double>(rg(r, F_T1ACC, slot));
        d[12 * s] = static_cast<double>(rg(r, F_T2ACC, slot));
    }
}

__global__ void k_writeback(const uint32_t *order, uint32_t n_act, manifolds_dev m, rows_dev r, uint32_t cap) // :1191-1196
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_act) return;
    uint32_t p = order[t];
    for (int c = 0; c < 4; ++c)
    {
        uint32_t slot = static_cast<uint32_t>(c) * cap + t;
        if (!r.valid[slot]) continue;
        double *d = m.data + static_cast<size_t>(c * kMF) * m.dcap + m.slot[p];
        size_t s = m.dcap;
        d[10 * s] = rf(r, F_NACC, slot);
        d[11 * s] = rf(r, F_T1ACC, slot);
        d[12 * s] = rf(r, F_T2ACC, slot);
    }
}

// ---- contact islands: connected components of the row manifolds over dynamic bodies ----
// Union-find with hooking of the larger root under the smaller: every component's root
// is its smallest body id, whatever order the unions run in.
__device__ __forceinline__ uint32_t par_ld(uint32_t *par, uint32_t x)
{ return __hip_atomic_load(&par[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ uint32_t uf_find(uint32_t *par, uint32_t x)
{
    uint32_t p = par_ld(par, x);
    while (p != x)
    {
        uint32_t gp = par_ld(par, p);
        if (gp != p) __hip_atomic_store(&par[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); // path halving
        x = p;
        p = gp;
    }
    return x;
}

__global__ void k_uf_init(uint32_t *par, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) par[i] = i;
}

// the row manifolds are list[0..nl)
__global__ void k_uf_union(const uint64_t *keys, const uint32_t *list, uint32_t nl, bodies_dev b, uint32_t *par)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nl) return;
    const uint32_t p = list[i];
    uint64_t key = keys[p];
    uint32_t a = static_cast<uint32_t>(key >> 32), c = static_cast<uint32_t>(key & 0xFFFFFFFFu);
    if (b.is_static[a] || b.is_static[c]) return;
    for (uint32_t guard = 0; guard < 4096u; ++guard)
    {
        uint32_t ra = uf_find(par, a), rc = uf_find(par, c);
        if (ra == rc) return;
        uint32_t hi = ra > rc ? ra : rc, lo = ra > rc ? rc : ra;
        uint32_t expect = hi;
        if (__hip_atomic_compare_exchange_strong(&par[hi], &expect, lo, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

__global__ void k_uf_flatten(uint32_t *par, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) par[i] = uf_find(par, i);
}

// a union that gave up (never seen; bounded so no thread can spin forever) leaves two
// bodies of one manifold in different islands: flagged, and the caller falls back
__global__ void k_uf_verify(const uint64_t *keys, const uint32_t *list, uint32_t nl, bodies_dev b, const uint32_t *par,
                            unsigned long long *fail)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nl) return;
    const uint32_t p = list[i];
    uint64_t key = keys[p];
    uint32_t a = static_cast<uint32_t>(key >> 32), c = static_cast<uint32_t>(key & 0xFFFFFFFFu);
    if (!b.is_static[a] && !b.is_static[c] && par[a] != par[c]) atomicMax(fail, ~0ull);
}

// sort key of row manifold list[i]: (island root << cbits) | colour
__global__ void k_island_keys(const uint64_t *keys, const uint32_t *colour, const uint32_t *list, uint32_t nl,
                              bodies_dev b, const uint32_t *par, int cbits, uint64_t *out, uint32_t *vals)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nl) return;
    const uint32_t p = list[i];
    const uint64_t key = keys[p];
    uint32_t a = static_cast<uint32_t>(key >> 32);
    if (b.is_static[a]) a = static_cast<uint32_t>(key & 0xFFFFFFFFu);
    out[i] = (static_cast<uint64_t>(par[a]) << cbits) | colour[p];
    vals[i] = p;
}

__global__ void k_island_flags(const uint64_t *sk, uint32_t n, uint64_t last, int cbits, uint32_t *flags)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    flags[t] = (sk[t] != last && (t == 0 || (sk[t] >> cbits) != (sk[t - 1] >> cbits))) ? 1u : 0u;
}

__global__ void k_island_starts(const uint32_t *flags, const uint32_t *pos, uint32_t n, uint32_t *starts)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && flags[t]) starts[pos[t]] = t;
}

// totals = (island count, row manifold count), read on the device; the grid covers an
// upper bound of the island count
__global__ void k_island_maxlen(const uint32_t *starts, const uint32_t *totals, unsigned long long *mx)
{
    const uint32_t n_isl = totals[0], n_act = totals[1];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t len = 0;
    if (i < n_isl) len = (i + 1 < n_isl ? starts[i + 1] : n_act) - starts[i];
    for (int off = 32; off > 0; off >>= 1)
    {
        uint32_t o = __shfl_xor(len, off);
        len = o > len ? o : len;
    }
    if (__lane_id() == 0 && len) atomicMax(mx, static_cast<unsigned long long>(len));
}

// number of row manifolds: position of the first `last` key
__global__ void k_count_below(const uint64_t *sk, uint32_t n, uint64_t last, uint32_t *out)
{
    if (blockIdx.x || threadIdx.x) return;
    uint32_t lo = 0, hi = n;
    while (lo < hi)
    {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (sk[mid] < last)
            lo = mid + 1;
        else
            hi = mid;
    }
    *out = lo;
}

// One workgroup per chunk of whole islands: chunk k holds the islands whose first
// position lies in [k W, (k+1) W) of the island-sorted order. Warm start then `iters`
// passes; each pass steps through the colours with a workgroup barrier between them,
// every thread solving its positions of the current colour. Manifolds of one colour
// share no dynamic body, and the islands of a chunk share none at all, so this is the
// colour-ordered solve restricted to the chunk. A chunk holds at most W + L - 1
// manifolds for islands of at most L, so W = CAP + 1 - L keeps every chunk within the
// workgroup's CAP threads (one manifold each, register-resident) and its LDS.
__device__ __forceinline__ uint32_t lower_u32(const uint32_t *a, uint32_t n, uint32_t key)
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
// Chunks of at most CAP manifolds keep the velocities of their dynamic bodies in LDS
// for all passes: body ids are sorted in LDS (bitonic), a body's slot is its first
// position in the sorted list, and the velocities are written back at the end.
__device__ __forceinline__ uint16_t slot_of(const uint32_t *ids, uint32_t n_ids, uint32_t id)
{
    uint32_t lo = 0, hi = n_ids;
    while (lo < hi)
    {
        uint32_t mid = (lo + hi) >> 1;
        if (ids[mid] < id)
            lo = mid + 1;
        else
            hi = mid;
    }
    return static_cast<uint16_t>(lo);
}

// ---- register-resident island solve ----
// A chunk of at most kIslandThreads manifolds gives each thread one manifold for the
// whole solve: its contacts' normals, lever arms, effective mass, bias, friction block
// and impulses, and both bodies' inverse mass and world inverse inertia, stay in
// registers across every pass and colour; only the bodies' velocities move, through
// LDS. The Jacobians are rebuilt from the normal and lever arms with the calls
// k_setup made (make_row, basis), so the FP64 solve is the one solve_one runs, bit for
// bit. The FP32 solve rebuilds them in FP32 from the rounded normal and lever arms.
template <class T> struct ctc_t
{
    T n[3], ra[3], rb[3];
    T meff, bias, mu, i11, i12, i22, an, a1, a2;
};
template <class T> __device__ __forceinline__ void ctc_load(ctc_t<T> &k, const rows_t<T> &r, uint32_t slot)
{
    for (int c = 0; c < 3; ++c)
    {
        k.n[c] = r.d[static_cast<size_t>(F_NJV + c) * r.cap4 + slot];
        k.ra[c] = r.d[static_cast<size_t>(F_RA + c) * r.cap4 + slot];
        k.rb[c] = r.d[static_cast<size_t>(F_RB + c) * r.cap4 + slot];
    }
    auto g = [&](int f) { return r.d[static_cast<size_t>(f) * r.cap4 + slot]; };
    k.meff = g(F_NMEFF);
    k.bias = g(F_NBIAS);
    k.mu = g(F_MU);
    k.i11 = g(F_I11);
    k.i12 = g(F_I12);
    k.i22 = g(F_I22);
    k.an = g(F_NACC);
    k.a1 = g(F_T1ACC);
    k.a2 = g(F_T2ACC);
}
template <class T> __device__ __forceinline__ void ctc_store_acc(const ctc_t<T> &k, const rows_t<T> &r, uint32_t slot)
{
    r.d[static_cast<size_t>(F_NACC) * r.cap4 + slot] = k.an;
    r.d[static_cast<size_t>(F_T1ACC) * r.cap4 + slot] = k.a1;
    r.d[static_cast<size_t>(F_T2ACC) * r.cap4 + slot] = k.a2;
}
// FP64: the rows of solve_one rebuilt (Jacobians from make_row's and basis' operations)
__device__ __forceinline__ void ctc_rows(const ctc_t<double> &k, rowv &nr, rowv &r1, rowv &r2)
{
    const v3 n = mk(k.n[0], k.n[1], k.n[2]), ra = mk(k.ra[0], k.ra[1], k.ra[2]), rb = mk(k.rb[0], k.rb[1], k.rb[2]);
    v3 t1, t2;
    basis(n, t1, t2);
    nr.Jv = n;
    nr.Jwa = cross(ra, n);
    nr.Jwb = -cross(rb, n);
    nr.Meff = k.meff;
    nr.bias = k.bias;
    nr.acc = k.an;
    r1.Jv = t1;
    r1.Jwa = cross(ra, t1);
    r1.Jwb = -cross(rb, t1);
    r1.Meff = 0.0;
    r1.bias = 0.0;
    r1.acc = k.a1;
    r2.Jv = t2;
    r2.Jwa = cross(ra, t2);
    r2.Jwb = -cross(rb, t2);
    r2.Meff = 0.0;
    r2.bias = 0.0;
    r2.acc = k.a2;
}
__device__ __forceinline__ void ctc_solve(ctc_t<double> &k, bool warm, body_state &A, body_state &B)
{
    rowv nr, t1, t2;
    ctc_rows(k, nr, t1, t2);
    if (warm)
    {
        row_warm(nr, A, B);
        row_warm(t1, A, B);
        row_warm(t2, A, B);
        return;
    }
    row_apply(nr, 0.0, DBL_MAX, A, B);
    friction_block(t1, t2, k.mu, nr.acc, k.i11, k.i12, k.i22, A, B);
    k.an = nr.acc;
    k.a1 = t1.acc;
    k.a2 = t2.acc;
}
// FP32: solve_one32's arithmetic with the Jacobians rebuilt in FP32
__device__ __forceinline__ f3 fcr(f3 a, f3 b) { return f3m(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
__device__ __forceinline__ void fbasis(f3 n, f3 &t1, f3 &t2)
{
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    t1 = f3m(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    t2 = f3m(b, sign + n.y * n.y * a, -n.y);
}
__device__ __forceinline__ void ctc_solve(ctc_t<float> &k, bool warm, body32 &A, body32 &B)
{
    const f3 n = f3m(k.n[0], k.n[1], k.n[2]), ra = f3m(k.ra[0], k.ra[1], k.ra[2]), rb = f3m(k.rb[0], k.rb[1], k.rb[2]);
    f3 t1, t2;
    fbasis(n, t1, t2);
    const f3 nJa = fcr(ra, n), nJb = scl(fcr(rb, n), -1.0f);
    const f3 aJa = fcr(ra, t1), aJb = scl(fcr(rb, t1), -1.0f);
    const f3 bJa = fcr(ra, t2), bJb = scl(fcr(rb, t2), -1.0f);
    if (warm)
    {
        if (k.an != 0.0f) push32(A, B, scl(n, k.an), scl(nJa, k.an), scl(nJb, k.an));
        if (k.a1 != 0.0f) push32(A, B, scl(t1, k.a1), scl(aJa, k.a1), scl(aJb, k.a1));
        if (k.a2 != 0.0f) push32(A, B, scl(t2, k.a2), scl(bJa, k.a2), scl(bJb, k.a2));
        return;
    }
    const float jv = jv32(n, nJa, nJb, A, B);
    const float lambda = -(jv + k.bias) * k.meff;
    float nv = k.an + lambda;
    nv = nv < 0.0f ? 0.0f : nv;
    const float ap = nv - k.an;
    k.an = nv;
    push32(A, B, scl(n, ap), scl(nJa, ap), scl(nJb, ap));
    const float maxf = k.mu * k.an;
    const float b1 = jv32(t1, aJa, aJb, A, B), b2 = jv32(t2, bJa, bJb, A, B);
    float n1 = k.a1 - (k.i11 * b1 + k.i12 * b2), n2 = k.a2 - (k.i12 * b1 + k.i22 * b2);
    const float mag = sqrtf(n1 * n1 + n2 * n2);
    if (mag > maxf)
    {
        n1 = n1 * (maxf / mag);
        n2 = n2 * (maxf / mag);
    }
    const float a1 = n1 - k.a1, a2 = n2 - k.a2;
    k.a1 = n1;
    k.a2 = n2;
    push32(A, B, addf(scl(t1, a1), scl(t2, a2)), addf(scl(aJa, a1), scl(bJa, a2)), addf(scl(aJb, a1), scl(bJb, a2)));
}
__device__ __forceinline__ void body_consts(const bodies_dev &b, uint32_t i, body_state &s) { s = ld_body(b, i); }
__device__ __forceinline__ void body_consts(const bodies_dev &b, uint32_t i, body32 &s) { s = ld_body32(b, i); }

// the whole solve of one chunk, one manifold per thread (m <= CAP)
template <bool F32, int CAP>
__device__ __forceinline__ void solve_chunk_reg(const uint64_t *keys, const uint32_t *order, const uint64_t *sk,
                                                uint32_t s0, uint32_t m, uint64_t cmask, uint32_t n_colours,
                                                const bodies_dev &b, const rows_dev &rd, uint32_t cap, uint32_t iters,
                                                const vel_lds &L, const uint16_t (*sl)[CAP])
{
    using T = typename std::conditional<F32, float, double>::type;
    using B = typename std::conditional<F32, body32, body_state>::type;
    const rows_t<T> r = rows_as<T>(rd);
    const uint32_t i = threadIdx.x;
    const bool own = i < m;
    const uint32_t t = s0 + i;
    uint32_t col = kNone, ia = 0, ib = 0;
    uint16_t sa = kNoSlot, sb = kNoSlot;
    B A, Bb;
    ctc_t<T> k[4];
    bool valid[4] = {false, false, false, false};
    v3 sva = mk(0.0, 0.0, 0.0), swa = sva, svb = sva, swb = sva; // velocities of static bodies
    if (own)
    {
        col = static_cast<uint32_t>(sk[t] & cmask);
        const uint64_t key = keys[order[t]];
        ia = static_cast<uint32_t>(key >> 32);
        ib = static_cast<uint32_t>(key & 0xFFFFFFFFu);
        sa = sl[0][i];
        sb = sl[1][i];
        body_consts(b, ia, A);
        body_consts(b, ib, Bb);
        if (sa == kNoSlot) vel_global().load(b, ia, kNoSlot, sva, swa);
        if (sb == kNoSlot) vel_global().load(b, ib, kNoSlot, svb, swb);
#pragma unroll
        for (int c = 0; c < 4; ++c)
        {
            const uint32_t slot = static_cast<uint32_t>(c) * cap + t;
            valid[c] = r.valid[slot] != 0;
            if (valid[c]) ctc_load(k[c], r, slot);
        }
    }
    for (uint32_t pass = 0; pass <= iters; ++pass)
    {
        const bool warm = pass == 0;
        for (uint32_t c = 0; c < n_colours; ++c)
        {
            if (own && col == c)
            {
                if (sa == kNoSlot)
                {
                    A.vel = sva;
                    A.ang = swa;
                }
                else
                    L.load(b, ia, sa, A.vel, A.ang);
                if (sb == kNoSlot)
                {
                    Bb.vel = svb;
                    Bb.ang = swb;
                }
                else
                    L.load(b, ib, sb, Bb.vel, Bb.ang);
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (valid[q]) ctc_solve(k[q], warm, A, Bb);
                if (!A.stat) L.store(b, ia, sa, A.vel, A.ang);
                if (!Bb.stat) L.store(b, ib, sb, Bb.vel, Bb.ang);
            }
            __syncthreads();
        }
    }
    if (own)
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (valid[q]) ctc_store_acc(k[q], r, static_cast<uint32_t>(q) * cap + t);
}

// kIslandThreads threads. A chunk of at most kIslandThreads manifolds is solved one
// manifold per thread in registers, one of at most CAP manifolds with the velocities in
// LDS and each manifold's rows read per pass, a larger one (an island of more than CAP
// manifolds) from global memory. CAP 256 (every island at most 256 manifolds, windows
// sized to match) halves the LDS, and its FP32 form is held to two waves per SIMD, so
// two workgroups share a CU; CAP 512 keeps every register for its long islands.
constexpr int kIslandThreads = 256;
template <bool F32, int CAP>
__global__ void __launch_bounds__(kIslandThreads, (F32 && CAP == 256) ? 2 : 1)
    k_solve_islands(const uint64_t *keys, const uint32_t *order, const uint64_t *sk, const uint32_t *starts,
                    uint32_t n_isl, uint32_t n_act, uint32_t n_colours, int cbits, bodies_dev b, rows_dev r,
                    uint32_t cap, uint32_t iters, uint32_t window)
{
    constexpr uint32_t kLdsIds = 2 * CAP;
    __shared__ uint32_t ids[kLdsIds];
    __shared__ double lv[kLdsIds][6];
    __shared__ uint16_t sl[2][CAP];
    const uint32_t k = blockIdx.x;
    const uint32_t lo = lower_u32(starts, n_isl, k * window);
    const uint32_t hi = lower_u32(starts, n_isl, (k + 1) * window);
    if (lo >= hi) return; // no island starts in this chunk
    const uint32_t s0 = starts[lo], e0 = hi < n_isl ? starts[hi] : n_act;
    const uint64_t cmask = (1ull << cbits) - 1;
    const uint32_t m = e0 - s0;
    const bool use_lds = m <= static_cast<uint32_t>(CAP);
    // the sorted id list spans the smallest power of two holding both bodies of every
    // manifold (sentinels sort last)
    uint32_t n_ids = 2;
    while (n_ids < 2 * m) n_ids <<= 1;
    if (use_lds)
    {
        for (uint32_t i = threadIdx.x; i < n_ids; i += kIslandThreads)
        {
            uint32_t id = 0xFFFFFFFFu;
            if (i < 2 * m)
            {
                uint64_t key = keys[order[s0 + i / 2]];
                uint32_t x = (i & 1) ? static_cast<uint32_t>(key & 0xFFFFFFFFu) : static_cast<uint32_t>(key >> 32);
                if (!b.is_static[x]) id = x;
            }
            ids[i] = id;
        }
        __syncthreads();
        for (uint32_t size = 2; size <= n_ids; size <<= 1) // bitonic sort
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1)
            {
                for (uint32_t i = threadIdx.x; i < n_ids; i += kIslandThreads)
                {
                    uint32_t j = i ^ stride;
                    if (j > i)
                    {
                        bool up = (i & size) == 0;
                        uint32_t x = ids[i], y = ids[j];
                        if ((x > y) == up)
                        {
                            ids[i] = y;
                            ids[j] = x;
                        }
                    }
                }
                __syncthreads();
            }
        for (uint32_t i = threadIdx.x; i < n_ids; i += kIslandThreads)
        {
            uint32_t id = ids[i];
            if (id != 0xFFFFFFFFu && (i == 0 || ids[i - 1] != id))
            {
                v3 v = ld3(b.vel, b.n, id), w = ld3(b.ang, b.n, id);
                lv[i][0] = v.x;
                lv[i][1] = v.y;
                lv[i][2] = v.z;
                lv[i][3] = w.x;
                lv[i][4] = w.y;
                lv[i][5] = w.z;
            }
        }
        for (uint32_t i = threadIdx.x; i < m; i += kIslandThreads)
        {
            uint64_t key = keys[order[s0 + i]];
            uint32_t x = static_cast<uint32_t>(key >> 32), y = static_cast<uint32_t>(key & 0xFFFFFFFFu);
            sl[0][i] = b.is_static[x] ? kNoSlot : slot_of(ids, n_ids, x);
            sl[1][i] = b.is_static[y] ? kNoSlot : slot_of(ids, n_ids, y);
        }
        __syncthreads();
    }
    vel_lds L{lv};
    if (use_lds && m <= static_cast<uint32_t>(kIslandThreads))
        solve_chunk_reg<F32, CAP>(keys, order, sk, s0, m, cmask, n_colours, b, r, cap, iters, L, sl);
    else
    for (uint32_t pass = 0; pass <= iters; ++pass)
    {
        const int warm = pass == 0 ? 1 : 0;
        for (uint32_t c = 0; c < n_colours; ++c)
        {
            for (uint32_t t = s0 + threadIdx.x; t < e0; t += kIslandThreads)
            {
                if ((sk[t] & cmask) != c) continue;
                if (use_lds)
                {
                    uint16_t sa = sl[0][t - s0], sb = sl[1][t - s0];
                    if (F32)
                        solve_one32(keys, order, t, b, as32(r), cap, warm, L, sa, sb);
                    else
                        solve_one(keys, order, t, b, r, cap, warm, L, sa, sb);
                }
                else if (F32)
                    solve_one32(keys, order, t, b, as32(r), cap, warm);
                else
                    solve_one(keys, order, t, b, r, cap, warm);
            }
            __syncthreads();
        }
    }
    if (use_lds)
        for (uint32_t i = threadIdx.x; i < n_ids; i += kIslandThreads)
        {
            uint32_t id = ids[i];
            if (id != 0xFFFFFFFFu && (i == 0 || ids[i - 1] != id))
            {
                st3(b.vel, b.n, id, mk(lv[i][0], lv[i][1], lv[i][2]));
                st3(b.ang, b.n, id, mk(lv[i][3], lv[i][4], lv[i][5]));
            }
        }
}
} // namespace

void launch_row_flags(const uint64_t *keys, uint32_t n_pairs, const bodies_dev &b, manifolds_dev m, uint32_t *colour,
                      uint32_t *list, unsigned long long *counters, hipStream_t st)
{
    if (!n_pairs) return;
    if (m.dcap) k_row_flags<<<grid(m.dcap), kBlock, 0, st>>>(keys, b, m, colour, list, counters);
}
void launch_contact_setup(const uint64_t *keys, uint32_t n, const uint32_t *ord, uint32_t cap, const bodies_dev &b,
                          manifolds_dev m, rows_dev r, double dt, double threshold, bool f32, hipStream_t st)
{
    if (!n) return;
    if (f32)
        k_setup<float><<<grid(n), kBlock, 0, st>>>(keys, n, ord, cap, b, m, rows_as<float>(r), dt, threshold);
    else
        k_setup<double><<<grid(n), kBlock, 0, st>>>(keys, n, ord, cap, b, m, rows_as<double>(r), dt, threshold);
}
void launch_pair_base(const uint64_t *keys, uint32_t n_pairs, const uint32_t *world_of, uint32_t *pair_base,
                      hipStream_t st)
{
    if (n_pairs) k_pair_base<<<grid(n_pairs), kBlock, 0, st>>>(keys, n_pairs, world_of, pair_base);
}
void launch_colour_claim(const uint64_t *keys, const uint32_t *list, uint32_t n, const unsigned long long *n_dev,
                         const bodies_dev &b, const uint32_t *colour, const uint32_t *world_of,
                         const uint32_t *pair_base, unsigned long long *claim, uint32_t round, hipStream_t st)
{
    if (n)
        k_claim<<<grid(n), kBlock, 0, st>>>(keys, list, n, n_dev, b, colour, world_of, pair_base, claim, round);
}
void launch_colour_assign(const uint64_t *keys, const uint32_t *list, uint32_t n, const unsigned long long *n_dev,
                          const bodies_dev &b, uint32_t *colour, const uint32_t *world_of, const uint32_t *pair_base,
                          const unsigned long long *claim, unsigned long long *used, uint32_t round,
                          unsigned long long *counters, hipStream_t st)
{
    if (n)
        k_assign<<<grid(n), kBlock, 0, st>>>(keys, list, n, n_dev, b, colour, world_of, pair_base, claim, used,
                                             round, counters);
}
uint32_t *launch_colour_bucket(const uint32_t *colour, const uint32_t *list, uint32_t nl, uint32_t n_pairs,
                               uint32_t n_colours, uint64_t *keys, uint64_t *keys_alt, uint32_t *order,
                               uint32_t *order_alt, uint32_t *colour_start, sort_scratch &sc, hipStream_t st)
{
    if (!nl || !n_colours) return order;
    k_colour_keys<<<grid(nl), kBlock, 0, st>>>(colour, list, nl, keys, order);
    int pbits = 1; // pair indices below n_pairs
    while (pbits < 32 && (1ull << pbits) < n_pairs) ++pbits;
    int cbits = 1; // colours below n_colours
    while (cbits < 32 && (1ull << cbits) < n_colours) ++cbits;
    int shifts[8];
    int ns = 0;
    for (int sft = 0; sft < pbits; sft += 8) shifts[ns++] = sft;
    for (int sft = 32; sft < 32 + cbits; sft += 8) shifts[ns++] = sft;
    bool alt = radix_sort_u64(keys, order, keys_alt, order_alt, nl, shifts, ns, sc, st);
    const uint64_t *sk = alt ? keys_alt : keys;
    k_colour_bounds<<<grid(n_colours + 1), kBlock, 0, st>>>(sk, nl, n_colours, colour_start);
    return alt ? order_alt : order;
}
void launch_solve_colour(const uint64_t *keys, const uint32_t *order, uint32_t begin, uint32_t end,
                         const bodies_dev &b, rows_dev r, uint32_t cap, bool warm, hipStream_t st)
{
    if (end > begin) k_solve<<<grid(end - begin), kBlock, 0, st>>>(keys, order, begin, end, b, r, cap, warm ? 1 : 0);
}
void launch_impulse_writeback(const uint32_t *order, uint32_t n_act, manifolds_dev m, rows_dev r, uint32_t cap,
                              hipStream_t st)
{
    if (n_act) k_writeback<<<grid(n_act), kBlock, 0, st>>>(order, n_act, m, r, cap);
}

island_sort launch_island_sort(const uint64_t *keys, const uint32_t *colour, const uint32_t *list, uint32_t nl,
                               uint32_t n_colours, const bodies_dev &b, uint32_t *par, uint64_t *ik, uint32_t *iv,
                               uint64_t *ik_alt, uint32_t *iv_alt, unsigned long long *fail, sort_scratch &sc,
                               hipStream_t st)
{
    island_sort is{};
    const uint32_t n = b.n;
    k_uf_init<<<grid(n), kBlock, 0, st>>>(par, n);
    if (nl) k_uf_union<<<grid(nl), kBlock, 0, st>>>(keys, list, nl, b, par);
    k_uf_flatten<<<grid(n), kBlock, 0, st>>>(par, n);
    if (nl) k_uf_verify<<<grid(nl), kBlock, 0, st>>>(keys, list, nl, b, par, fail);
    int cbits = 1;
    while (cbits < 31 && (1u << cbits) <= n_colours) ++cbits;
    int ibits = 1;
    while (ibits < 32 && (1ull << ibits) < n) ++ibits;
    const int bits = cbits + ibits;
    is.cbits = cbits;
    is.last = (1ull << bits) - 1;
    if (nl) k_island_keys<<<grid(nl), kBlock, 0, st>>>(keys, colour, list, nl, b, par, cbits, ik, iv);
    int shifts[8];
    int ns = 0;
    for (int sft = 0; sft < bits; sft += 8) shifts[ns++] = sft;
    bool alt = nl > 1 && radix_sort_u64(ik, iv, ik_alt, iv_alt, nl, shifts, ns, sc, st);
    is.sk = alt ? ik_alt : ik;
    is.ord = alt ? iv_alt : iv;
    return is;
}
void launch_island_segments(const island_sort &is, uint32_t n_pairs, uint32_t *flags, uint32_t *pos, uint32_t *bsums,
                            uint32_t *total, uint32_t *starts, uint32_t *n_act, hipStream_t st)
{
    k_count_below<<<1, 64, 0, st>>>(is.sk, n_pairs, is.last, n_act);
    k_island_flags<<<grid(n_pairs), kBlock, 0, st>>>(is.sk, n_pairs, is.last, is.cbits, flags);
    exclusive_scan_u32(flags, pos, n_pairs, bsums, total, st);
    k_island_starts<<<grid(n_pairs), kBlock, 0, st>>>(flags, pos, n_pairs, starts);
}
void launch_island_maxlen(const uint32_t *starts, const uint32_t *totals, uint32_t max_islands,
                          unsigned long long *mx, hipStream_t st)
{
    if (max_islands) k_island_maxlen<<<grid(max_islands), kBlock, 0, st>>>(starts, totals, mx);
}
void launch_solve_islands(const uint64_t *keys, const island_sort &is, const uint32_t *starts, uint32_t n_isl,
                          uint32_t n_act, uint32_t n_colours, const bodies_dev &b, const rows_dev &r, uint32_t cap,
                          uint32_t iters, uint32_t longest, bool f32, hipStream_t st)
{
    if (!n_isl) return;
    // the smaller workgroup (and LDS) when every island fits it, so more chunks are
    // resident at once; windows shrink with the longest island so every chunk fits
    auto go = [&](auto capacity, uint32_t window)
    {
        constexpr int CAP = decltype(capacity)::value;
        const unsigned nb = static_cast<unsigned>((n_act + window - 1) / window);
        if (f32)
            k_solve_islands<true, CAP><<<nb, kIslandThreads, 0, st>>>(keys, is.ord, is.sk, starts, n_isl, n_act,
                                                                       n_colours, is.cbits, b, r, cap, iters, window);
        else
            k_solve_islands<false, CAP><<<nb, kIslandThreads, 0, st>>>(keys, is.ord, is.sk, starts, n_isl, n_act,
                                                                        n_colours, is.cbits, b, r, cap, iters, window);
    };
    if (longest <= static_cast<uint32_t>(kIslandThreads))
        go(std::integral_constant<int, kIslandThreads>{}, kIslandThreads + 1 - std::max<uint32_t>(longest, 1));
    else
        go(std::integral_constant<int, 2 * kIslandThreads>{}, kIslandThreads);
}
void launch_solve_colour32(const uint64_t *keys, const uint32_t *order, uint32_t begin, uint32_t end,
                           const bodies_dev &b, rows_dev r32, uint32_t cap, bool warm, hipStream_t st)
{
    if (end > begin)
        k_solve32<<<grid(end - begin), kBlock, 0, st>>>(keys, order, begin, end, b, as32(r32), cap, warm ? 1 : 0);
}
void launch_impulse_writeback32(const uint32_t *order, uint32_t n_act, manifolds_dev m, rows_dev r32, uint32_t cap,
                                hipStream_t st)
{
    if (n_act) k_writeback32<<<grid(n_act), kBlock, 0, st>>>(order, n_act, m, as32(r32), cap);
}

} // namespace pk
