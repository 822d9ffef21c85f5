// GJK + EPA on the device (src/collision.cpp:12-518, include/physkit/collision/
// collision.h:27-59). Same branch structure and operation order as the reference, so
// verdicts, depths, normals and witness points are reproducible bit for bit.
// The GJK simplex lives in registers (only constant indices); a group of lanes can
// share one GJK, splitting its support searches. The EPA polytope, faces and min-heap
// live in a caller-provided store of compile-time capacity; running out of capacity
// is reported (status 3) and never truncates silently.
#pragma once
#include "pk_math.hpp"

namespace pk
{

// A support point keeps the Minkowski point and the support feature of each shape
// (a mesh vertex index, or the corner code of a box); the witness points pa, pb
// are recomputed from the features when needed, bit for bit what support() returned.
struct spt
{
    v3 p;
    uint32_t ia, ib;
};

// shape kinds: 0 aabb (a = min, b = max), 1 obb (a = center, b = half extents, q),
// 2 mesh instance (a = position, q, vertices v[0..nv) xyz interleaved)
struct gshape
{
    int kind;
    v3 a, b;
    qt q;
    const double *v;
    uint32_t nv;
    uint32_t uv;     // mesh built as a UV sphere: (stacks << 16) | sectors, else 0
    uint32_t paired; // v is 16-byte aligned and readable up to an even vertex count
};

__device__ __forceinline__ v3 ldv(const double *v, uint32_t i) { return mk(v[3 * i], v[3 * i + 1], v[3 * i + 2]); }

constexpr uint32_t kNoIdx = 0xFFFFFFFFu;

// Support feature of s in direction d. G lanes of a group (G = 1: one thread) share
// a shape: lane gl scans vertices gl, gl+G, ... and the group reduces. The result is
// the reference's (mesh::support, mesh.cpp:341-358): the first vertex with a strictly
// greater dot product, i.e. the lowest index among the maxima, except that a NaN dot
// at vertex 0 keeps vertex 0 (no later value compares greater) and NaN dots elsewhere
// are never taken.
// Support of a UV-sphere mesh (tag s.uv, stacks and sectors at most 1024) from the
// vertices around the analytic optimum, with a certificate. Direction d in the mesh
// frame has polar angle phi* from +y and azimuth alpha; the candidates are both poles
// and the sectors nearest alpha (3) on the rings nearest phi* (3), evaluated with the
// same dot product as the full scan and ranked by (value, lowest index). phi* and
// alpha are located in single precision from the normalised direction (error ~1e-7
// rad, far below a tenth of a cell), so every other vertex lies at least 1.3 ring
// steps (pi/stacks) from phi* or, on a candidate ring, 1.3 sector steps
// (2 pi/sectors) from alpha, which bounds its dot product by
//   r |d| cos(1.3 pi/stacks)                                   (other rings)
//   r sin(phi_i) rho cos(1.3 2pi/sectors) + r cos(phi_i) d.y   (other sectors of ring i)
// with rho = |(d.x, d.z)| (the true gap is 1.5 steps; the difference absorbs the
// single-precision cosines). If the best candidate exceeds both bounds by a slack far
// above every rounding error of the dot products and the vertex coordinates, no other
// vertex can equal or beat it, so the result is the full scan's; otherwise
// (degenerate directions) the caller scans every vertex.
__device__ __forceinline__ bool uv_support(const gshape &s, v3 ld, uint32_t &out)
{
    const int S = static_cast<int>(s.uv >> 16), T = static_cast<int>(s.uv & 0xFFFFu);
    if (!(isfinite(ld.x) && isfinite(ld.y) && isfinite(ld.z))) return false;
    const double rho = sqrt(ld.x * ld.x + ld.z * ld.z);
    const double dn = sqrt(rho * rho + ld.y * ld.y);
    if (!(dn > 0.0)) return false;
    const float pif = 3.14159265f;
    const float phi = atan2f(static_cast<float>(rho / dn), static_cast<float>(ld.y / dn));
    int i0 = static_cast<int>(rintf(phi * (static_cast<float>(S) / pif)));
    i0 = i0 < 1 ? 1 : (i0 > S - 1 ? S - 1 : i0);
    int j0 = 0;
    if (rho > 0.0)
    {
        const float al = atan2f(static_cast<float>(ld.z / rho), static_cast<float>(ld.x / rho));
        j0 = static_cast<int>(rintf(al * (static_cast<float>(T) / (2.0f * pif))));
        j0 = ((j0 % T) + T) % T;
    }
    const int ilo = i0 > 1 ? i0 - 1 : 1, ihi = i0 < S - 1 ? i0 + 1 : S - 1;
    double bv = dot(ldv(s.v, 0), ld);
    uint32_t bi = 0;
    for (int i = ilo; i <= ihi; ++i)
        for (int dj = -1; dj <= 1; ++dj)
        {
            uint32_t k = 1u + static_cast<uint32_t>((i - 1) * T + (j0 + dj + T) % T);
            double t = dot(ldv(s.v, k), ld);
            if (t > bv || (t == bv && k < bi))
            {
                bv = t;
                bi = k;
            }
        }
    {
        uint32_t k = s.nv - 1;
        double t = dot(ldv(s.v, k), ld);
        if (t > bv) // the largest index: it wins no tie
        {
            bv = t;
            bi = k;
        }
    }
    const double r = s.v[1];
    const double slack = 1e-12 * r * (fabs(ld.x) + fabs(ld.y) + fabs(ld.z));
    double ub = -__builtin_inf();
    if (ilo > 1 || ihi < S - 1) ub = r * dn * static_cast<double>(cosf(1.3f * pif / static_cast<float>(S)));
    if (T > 3)
    {
        const double cs = static_cast<double>(cosf(1.3f * 2.0f * pif / static_cast<float>(T)));
        for (int i = ilo; i <= ihi; ++i)
        {
            const double *p = s.v + 3 * (1 + static_cast<size_t>(i - 1) * T); // sector 0: (r sin phi, r cos phi, 0)
            double u = p[0] * rho * cs + p[1] * ld.y;
            ub = u > ub ? u : ub;
        }
    }
    if (!(bv > ub + slack)) return false;
    out = bi;
    return true;
}

template <int G> __device__ __forceinline__ uint32_t support_idx(const gshape &s, v3 d, int gl)
{
    if (s.kind == 2) // mesh::instance::support, mesh.cpp:442-448
    {
        v3 ld = rotate(conj(s.q), d);
        uint32_t k;
        if (s.uv && uv_support(s, ld, k)) return k;
        double bv = 0.0;
        uint32_t bi = kNoIdx;
        bool nan0 = false;
        if (s.paired)
        {
            // lane gl scans vertex pairs gl, gl+G, ...: both vertices of a pair in three
            // aligned 16-byte loads, taken in index order
            const uint32_t np = (s.nv + 1) >> 1;
            for (uint32_t j = static_cast<uint32_t>(gl); j < np; j += G)
            {
                const double2 *q = reinterpret_cast<const double2 *>(s.v + 6 * static_cast<size_t>(j));
                const double2 a = q[0], b = q[1], c = q[2];
                const uint32_t i = 2 * j;
                double t = dot(mk(a.x, a.y, b.x), ld);
                if (i == 0) nan0 = t != t;
                if (t == t && (bi == kNoIdx || t > bv))
                {
                    bv = t;
                    bi = i;
                }
                if (i + 1 < s.nv)
                {
                    t = dot(mk(b.y, c.x, c.y), ld);
                    if (t == t && (bi == kNoIdx || t > bv))
                    {
                        bv = t;
                        bi = i + 1;
                    }
                }
            }
        }
        else
            for (uint32_t i = static_cast<uint32_t>(gl); i < s.nv; i += G)
            {
                double t = dot(ldv(s.v, i), ld);
                if (i == 0) nan0 = t != t;
                if (t == t && (bi == kNoIdx || t > bv))
                {
                    bv = t;
                    bi = i;
                }
            }
        for (int m = 1; m < G; m <<= 1)
        {
            double ov = __shfl_xor(bv, m, G);
            uint32_t oi = __shfl_xor(bi, m, G);
            int on = __shfl_xor(static_cast<int>(nan0), m, G);
            bool take = oi != kNoIdx && (bi == kNoIdx || ov > bv || (!(bv > ov) && oi < bi));
            if (take)
            {
                bv = ov;
                bi = oi;
            }
            nan0 = nan0 || on != 0;
        }
        return nan0 ? 0u : bi;
    }
    if (s.kind == 1) // obb::support, bounds.h:539-548: sign of the local direction
    {
        v3 ld = rotate(conj(s.q), d);
        return (ld.x >= 0 ? 1u : 0u) | (ld.y >= 0 ? 2u : 0u) | (ld.z >= 0 ? 4u : 0u);
    }
    // aabb::support, bounds.h:188-199
    return (d.x >= 0 ? 1u : 0u) | (d.y >= 0 ? 2u : 0u) | (d.z >= 0 ? 4u : 0u);
}

__device__ __forceinline__ v3 support_at(const gshape &s, uint32_t f)
{
    if (s.kind == 2) return rotate(s.q, ldv(s.v, f)) + s.a;
    if (s.kind == 1)
    {
        v3 sg = mk(((f & 1u) ? 1.0 : -1.0) * s.b.x, ((f & 2u) ? 1.0 : -1.0) * s.b.y,
                   ((f & 4u) ? 1.0 : -1.0) * s.b.z);
        return s.a + rotate(s.q, sg);
    }
    return mk((f & 1u) ? s.b.x : s.a.x, (f & 2u) ? s.b.y : s.a.y, (f & 4u) ? s.b.z : s.a.z);
}

__device__ __forceinline__ v3 sel3(bool c, v3 a, v3 b) { return mk(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z); }
__device__ __forceinline__ gshape sel_shape(bool c, const gshape &a, const gshape &b)
{
    gshape r;
    r.kind = c ? a.kind : b.kind;
    r.a = sel3(c, a.a, b.a);
    r.b = sel3(c, a.b, b.b);
    r.q = mkq(c ? a.q.w : b.q.w, c ? a.q.x : b.q.x, c ? a.q.y : b.q.y, c ? a.q.z : b.q.z);
    r.v = c ? a.v : b.v;
    r.nv = c ? a.nv : b.nv;
    r.uv = c ? a.uv : b.uv;
    r.paired = c ? a.paired : b.paired;
    return r;
}

// s_A(d) - s_B(-d), collision.h:41-49. With UV-sphere shapes possible (UV) the two
// support searches share one loop body (kept rolled) so the larger search code is
// emitted once per call site; without, both are inlined directly.
template <int G, bool UV = true>
__device__ __forceinline__ spt msupport(const gshape &A, const gshape &B, v3 d, int gl)
{
    spt r;
    if (!UV)
    {
        r.ia = support_idx<G>(A, d, gl);
        r.ib = support_idx<G>(B, -d, gl);
        r.p = support_at(A, r.ia) - support_at(B, r.ib);
        return r;
    }
    r.ia = 0;
    r.ib = 0;
#pragma nounroll
    for (int side = 0; side < 2; ++side)
    {
        gshape sh = sel_shape(side == 0, A, B);
        uint32_t f = support_idx<G>(sh, side == 0 ? d : -d, gl);
        if (side == 0)
            r.ia = f;
        else
            r.ib = f;
    }
    r.p = support_at(A, r.ia) - support_at(B, r.ib);
    return r;
}

struct simplex
{
    spt s0, s1, s2, s3;
    int n;
};
// Written as value selects: a branch per slot lets the optimiser merge the four
// stores into one store through a selected pointer, which pins the simplex in
// private memory (scratch) instead of registers.
__device__ __forceinline__ void sx_put(spt &dst, bool c, const spt &p)
{
    dst.p = sel3(c, p.p, dst.p);
    dst.ia = c ? p.ia : dst.ia;
    dst.ib = c ? p.ib : dst.ib;
}
__device__ __forceinline__ void sx_push(simplex &x, const spt &p)
{
    int n = x.n;
    sx_put(x.s0, n == 0, p);
    sx_put(x.s1, n == 1, p);
    sx_put(x.s2, n == 2, p);
    sx_put(x.s3, n >= 3, p);
    x.n = n + 1;
}

__device__ __forceinline__ void handle_line(simplex &x, v3 &dir) // collision.cpp:12-41
{
    const double eps = 1e-12;
    spt a = x.s1, b = x.s0;
    v3 ab = b.p - a.p, ao = -a.p;
    if (dot(ab, ao) > 0.0)
    {
        v3 triple = cross(cross(ab, ao), ab);
        if (sqn(triple) < eps)
        {
            v3 abh = normalized(ab);
            v3 perp = cross(abh, mk(0.0, 1.0, 0.0));
            if (sqn(perp) < eps) perp = cross(abh, mk(0.0, 0.0, 1.0));
            dir = normalized(perp);
        }
        else
            dir = normalized(triple);
    }
    else
    {
        x.s0 = x.s1;
        x.n = 1;
        dir = normalized(ao);
    }
}

__device__ __forceinline__ void handle_triangle(simplex &x, v3 &dir) // collision.cpp:43-88
{
    spt a = x.s2, b = x.s1, c = x.s0;
    v3 ab = b.p - a.p, ac = c.p - a.p, ao = -a.p;
    v3 abc = cross(ab, ac);
    v3 ab_perp = cross(ab, abc);
    if (dot(ab_perp, ao) > 0.0)
    {
        x.s0 = b;
        x.s1 = a;
        x.n = 2;
        v3 triple = cross(cross(ab, ao), ab);
        dir = (sqn(triple) < 1e-12) ? normalized(ao) : normalized(triple);
        return;
    }
    v3 ac_perp = cross(abc, ac);
    if (dot(ac_perp, ao) > 0.0)
    {
        x.s0 = c;
        x.s1 = a;
        x.n = 2;
        v3 triple = cross(cross(ac, ao), ac);
        dir = (sqn(triple) < 1e-12) ? normalized(ao) : normalized(triple);
        return;
    }
    if (dot(abc, ao) <= 0.0)
    {
        x.s0 = b;
        x.s1 = c;
        dir = normalized(-abc);
    }
    else
        dir = normalized(abc);
}

__device__ __forceinline__ bool handle_tetrahedron(simplex &x, v3 &dir) // collision.cpp:90-147
{
    spt a = x.s3, b = x.s2, c = x.s1, d = x.s0;
    v3 ao = -a.p;
    v3 abc = cross(b.p - a.p, c.p - a.p);
    v3 acd = cross(c.p - a.p, d.p - a.p);
    v3 adb = cross(d.p - a.p, b.p - a.p);
    if (dot(abc, d.p - a.p) > 0.0) abc = -abc;
    if (dot(acd, b.p - a.p) > 0.0) acd = -acd;
    if (dot(adb, c.p - a.p) > 0.0) adb = -adb;
    if (dot(abc, ao) > 0.0)
    {
        x.s0 = c;
        x.s1 = b;
        x.s2 = a;
        x.n = 3;
        dir = normalized(abc);
        handle_triangle(x, dir);
        return false;
    }
    if (dot(acd, ao) > 0.0)
    {
        x.s0 = d;
        x.s1 = c;
        x.s2 = a;
        x.n = 3;
        dir = normalized(acd);
        handle_triangle(x, dir);
        return false;
    }
    if (dot(adb, ao) > 0.0)
    {
        x.s0 = b;
        x.s1 = d;
        x.s2 = a;
        x.n = 3;
        dir = normalized(adb);
        handle_triangle(x, dir);
        return false;
    }
    return true;
}

// boolean GJK, collision.cpp:165-189; returns true with the final simplex on a hit
// (inlined so the simplex stays in registers). With G > 1 the G lanes of a group run
// it together: support searches are split over the lanes, everything else is
// computed identically in every lane of the group.
template <int G, bool UV = true> __device__ __forceinline__ bool gjk(const gshape &A, const gshape &B, simplex &x, int gl)
{
    const double eps = 1e-12;
    x.n = 0;
    v3 dir = mk(1.0, 0.0, 0.0);
    // pass -1 is the initial support (with UV spheres possible, one call site keeps the
    // larger search code single; otherwise the first search is peeled off the loop)
    int first = -1;
    if (!UV)
    {
        spt p = msupport<G, UV>(A, B, dir, gl);
        sx_push(x, p);
        if (sqn(p.p) < eps) return true;
        dir = -normalized(p.p);
        first = 0;
    }
    for (int it = first; it < 100; ++it)
    {
        spt np = msupport<G, UV>(A, B, dir, gl);
        if (it < 0)
        {
            sx_push(x, np);
            if (sqn(np.p) < eps) return true;
            dir = -normalized(np.p);
            continue;
        }
        if (dot(np.p, dir) <= 0.0) return false;
        sx_push(x, np);
        if (x.n == 2)
            handle_line(x, dir);
        else if (x.n == 3)
            handle_triangle(x, dir);
        else if (handle_tetrahedron(x, dir))
            return true;
    }
    return false;
}

// collision.cpp:191-248
template <int G, bool UV = true>
__device__ __forceinline__ bool pad_simplex(const gshape &A, const gshape &B, simplex &x, int gl)
{
    // stage k fills slot k (1: s1 from direction +x, 2: s2 off the line s0-s1, 3: s3
    // off the plane), each retried once in the opposite direction; written as one
    // loop so the support search is emitted once
    for (int k = x.n < 1 ? 1 : x.n; k <= 3; ++k)
    {
        v3 line = x.s1.p - x.s0.p;
        v3 dir;
        if (k == 1)
            dir = mk(1.0, 0.0, 0.0);
        else if (k == 2)
        {
            dir = cross(normalized(line), mk(0.0, 1.0, 0.0));
            if (sqn(dir) < 1e-6) dir = cross(normalized(line), mk(0.0, 0.0, 1.0));
            dir = normalized(dir);
        }
        else
            dir = normalized(cross(line, x.s2.p - x.s0.p));
        spt p;
#pragma nounroll
        for (int att = 0; att < 2; ++att)
        {
            p = msupport<G, UV>(A, B, att == 0 ? dir : -dir, gl);
            if (att == 1) break;
            bool again;
            if (k == 1)
                again = sqn(p.p - x.s0.p) < 1e-6;
            else if (k == 2)
                again = sqn(cross(line, p.p - x.s0.p)) < 1e-6;
            else
                again = fabs(dot(p.p - x.s0.p, dir)) < 1e-6;
            if (!again) break;
        }
        sx_put(x.s1, k == 1, p);
        sx_put(x.s2, k == 2, p);
        sx_put(x.s3, k == 3, p);
        x.n = k + 1;
    }
    v3 ad = x.s0.p - x.s3.p;
    v3 bd = x.s1.p - x.s3.p;
    v3 cd = x.s2.p - x.s3.p;
    double triple = dot(ad, cross(bd, cd));
    return fabs(triple) > 1e-12;
}


// ---------------- EPA (collision.cpp:251-509) ----------------
// One thread per solve. The polytope lives in a store of compile-time capacity
// (FC faces, PC points, index type I) laid out field by field: the face normals and
// plane distances, vertex and adjacency indices, the obsolete flags, and a min-heap
// that keeps each slot's plane distance beside its face index so a sift step reads
// one word instead of chasing the face. Running out of capacity returns 3 and the
// caller reruns the solve with a larger store; capacity never changes a result.
template <int FC, int PC, class I> struct epa_store
{
    static_assert(FC < (1 << (8 * sizeof(I))) && PC < (1 << (8 * sizeof(I))), "index type too small");
    static constexpr int fcap = FC, pcap = PC;
    using idx_t = I;
    static constexpr I null = static_cast<I>(~I(0));
    struct geo
    {
        double nx, ny, nz, d; // face normal and plane distance: one 32-byte record
    } g[FC];
    double hd[FC]; // heap: plane distance of the face in each slot
    double px[PC], py[PC], pz[PC];        // polytope points
    uint32_t pia[PC], pib[PC];            // their support features
    // vertices, face across edge (v[i], v[i+1]) and the obsolete flag of a face, packed
    // in one aligned word, so a face's bookkeeping is one memory access, not seven
    struct alignas(8 * sizeof(I)) meta
    {
        I v[3], a[3], obs, pad;
    } fm[FC];
    struct alignas(4 * sizeof(I)) edge
    {
        I s, e, adj, pad; // horizon edge (s, e) and the face across it
    } eg[FC];
    I hf[FC];  // heap: face in each slot
    I stk[FC]; // silhouette stack

    __device__ __forceinline__ v3 nrm(int f) const { return mk(g[f].nx, g[f].ny, g[f].nz); }
    __device__ __forceinline__ double dist(int f) const { return g[f].d; }
    __device__ __forceinline__ void set_face(int f, v3 n, double d)
    {
        g[f].nx = n.x;
        g[f].ny = n.y;
        g[f].nz = n.z;
        g[f].d = d;
    }
    __device__ __forceinline__ double hkey(int i) const { return hd[i]; }
    __device__ __forceinline__ void set_hkey(int i, double v) { hd[i] = v; }
    __device__ __forceinline__ v3 pt(int i) const { return mk(px[i], py[i], pz[i]); }
    __device__ __forceinline__ void set_pt(int i, v3 p)
    {
        px[i] = p.x;
        py[i] = p.y;
        pz[i] = p.z;
    }
};
using epa_store_fast = epa_store<192, 68, uint8_t>;    // every hull solve seen so far fits
using epa_store_lds = epa_store<896, 68, uint16_t>;    // overflow path, in LDS
using epa_store_big = epa_store<4096, 68, uint16_t>;   // last overflow path, global memory
static_assert(sizeof(epa_store_lds) <= 65536, "the LDS store must fit one workgroup's 64 KiB");

template <class S> struct epa_solver
{
    S &m;
    int npoly = 0, nfaces = 0, nheap = 0;
    __device__ explicit epa_solver(S &store) : m(store) {}

    __device__ __forceinline__ v3 P(int i) const { return m.pt(i); }
    __device__ __forceinline__ v3 N(int f) const { return m.nrm(f); }

    // init_face(fi, s, e, k, -1) (:273-297), the link to adj across (s, e) (:305-313)
    // and push(fi) for the new face (s, e, k) on a horizon edge. k is the new point,
    // whose coordinates pk are still in registers. The face's vertices stay in the order
    // (s, e, k), so s is at index 0 and the link to adj is a[0].
    __device__ __forceinline__ void new_face(int fi, int s, int e, int k, v3 pk, int adj)
    {
        v3 ps = P(s);
        v3 ab = P(e) - ps;
        v3 ac = pk - ps;
        v3 nrm = cross(ab, ac);
        if (sqn(nrm) < 1e-12)
            nrm = mk(0.0, 0.0, 0.0);
        else
            nrm = normalized(nrm);
        const double d = dot(nrm, ps);
        typename S::meta mt;
        mt.v[0] = static_cast<typename S::idx_t>(s);
        mt.v[1] = static_cast<typename S::idx_t>(e);
        mt.v[2] = static_cast<typename S::idx_t>(k);
        mt.a[0] = static_cast<typename S::idx_t>(adj);
        mt.a[1] = S::null;
        mt.a[2] = S::null;
        mt.obs = 0;
        mt.pad = 0;
        m.fm[fi] = mt;
        m.set_face(fi, nrm, d);
        m.fm[adj].a[vidx(adj, e)] = static_cast<typename S::idx_t>(fi);
        ++nheap;
        sift_up(nheap - 1, fi, d);
    }
    __device__ __forceinline__ int vidx(int f, int va) const
    {
        typename S::meta mt = m.fm[f];
        return (mt.v[0] == va) ? 0 : (mt.v[1] == va ? 1 : 2);
    }
    // min-heap on plane distance, in the sift order of libstdc++'s push_heap/pop_heap
    __device__ __forceinline__ void sift_up(int hole, int vf, double vd)
    {
        int parent = (hole - 1) / 2;
        while (hole > 0)
        {
            double pd = m.hkey(parent);
            if (!(pd > vd)) break;
            m.set_hkey(hole, pd);
            m.hf[hole] = m.hf[parent];
            hole = parent;
            parent = (hole - 1) / 2;
        }
        m.set_hkey(hole, vd);
        m.hf[hole] = static_cast<typename S::idx_t>(vf);
    }
    __device__ __forceinline__ void push(int f)
    {
        ++nheap;
        sift_up(nheap - 1, f, m.dist(f));
    }
    __device__ int pop()
    {
        while (nheap > 0)
        {
            int len = nheap;
            int top = m.hf[0];
            if (len > 1)
            {
                int vf = m.hf[len - 1];
                double vd = m.hkey(len - 1);
                int n = len - 1;
                int hole = 0, child = 0;
                while (child < (n - 1) / 2)
                {
                    child = 2 * (child + 1);
                    double rd = m.hkey(child), ld = m.hkey(child - 1);
                    if (rd > ld)
                    {
                        child--;
                        rd = ld;
                    }
                    m.set_hkey(hole, rd);
                    m.hf[hole] = m.hf[child];
                    hole = child;
                }
                if ((n & 1) == 0 && child == (n - 2) / 2)
                {
                    child = 2 * (child + 1);
                    m.set_hkey(hole, m.hkey(child - 1));
                    m.hf[hole] = m.hf[child - 1];
                    hole = child - 1;
                }
                sift_up(hole, vf, vd);
            }
            --nheap;
            if (!m.fm[top].obs) return top;
        }
        return -1;
    }
    // :315-353; horizon edges into es/ee/eadj; returns their count, -1 on overflow
    __device__ int silhouette(int f0, v3 p)
    {
        const double tol = 1e-6;
        int sp = 0, nh = 0;
        m.fm[f0].obs = 1;
        // f0 first, then the faces on the stack (f0 needs no trip through the stack)
        int cur = f0;
        for (;;)
        {
            typename S::meta cm = m.fm[cur];
            for (int i = 0; i < 3; ++i)
            {
                int nb = cm.a[i];
                if (nb == S::null) continue;
                if (m.fm[nb].obs) continue;
                if (dot(N(nb), p) > m.dist(nb) + tol)
                {
                    m.fm[nb].obs = 1;
                    m.stk[sp++] = static_cast<typename S::idx_t>(nb); // each face is pushed at most once
                }
                else
                {
                    if (nh >= S::fcap) return -1;
                    typename S::edge ed;
                    ed.s = cm.v[i];
                    ed.e = cm.v[(i + 1) % 3];
                    ed.adj = static_cast<typename S::idx_t>(nb);
                    ed.pad = 0;
                    m.eg[nh] = ed;
                    ++nh;
                }
            }
            if (sp == 0) break;
            cur = m.stk[--sp];
        }
        return nh;
    }
    // :355-388, from the four points q (already stored as points 0-3), held in registers:
    // the faces' vertices and links are worked out in registers and each face record is
    // written once
    __device__ void initial_tetrahedron(const v3 (&q)[4])
    {
        const int def[4][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2}, {1, 3, 2, 0}};
        int fv[4][3];
#pragma unroll
        for (int f = 0; f < 4; ++f)
        {
            // init_face(f, def[f][0], def[f][1], def[f][2], opposite point def[f][3]) (:273-297)
            const int i = def[f][0];
            int v1 = def[f][1], v2 = def[f][2];
            const v3 pi = q[i];
            v3 nrm = cross(q[v1] - pi, q[v2] - pi);
            if (sqn(nrm) < 1e-12)
                nrm = mk(0.0, 0.0, 0.0);
            else
                nrm = normalized(nrm);
            if (dot(nrm, q[def[f][3]] - pi) > 0.0)
            {
                const int t = v1;
                v1 = v2;
                v2 = t;
                nrm = -nrm;
            }
            fv[f][0] = i;
            fv[f][1] = v1;
            fv[f][2] = v2;
            const double d = dot(nrm, pi);
            m.set_face(f, nrm, d);
            nfaces = f + 1;
            ++nheap;
            sift_up(nheap - 1, f, d);
        }
        int fa[4][3];
#pragma unroll
        for (int f = 0; f < 4; ++f)
            fa[f][0] = fa[f][1] = fa[f][2] = S::null;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = i + 1; j < 4; ++j)
#pragma unroll
                for (int e1 = 0; e1 < 3; ++e1)
                {
                    const int u1 = fv[i][e1], w1 = fv[i][(e1 + 1) % 3];
#pragma unroll
                    for (int e2 = 0; e2 < 3; ++e2)
                    {
                        const int u2 = fv[j][e2], w2 = fv[j][(e2 + 1) % 3];
                        if (u1 == w2 && w1 == u2)
                        {
                            fa[i][e1] = j;
                            fa[j][e2] = i;
                        }
                    }
                }
#pragma unroll
        for (int f = 0; f < 4; ++f)
        {
            typename S::meta mt;
            for (int c = 0; c < 3; ++c)
            {
                mt.v[c] = static_cast<typename S::idx_t>(fv[f][c]);
                mt.a[c] = static_cast<typename S::idx_t>(fa[f][c]);
            }
            mt.obs = 0;
            mt.pad = 0;
            m.fm[f] = mt;
        }
    }
    // :424-454; out = normal3, world_a3, world_b3, depth
    __device__ void barycentric(int fi, const gshape &A, const gshape &B, double *out) const
    {
        typename S::meta fmt = m.fm[fi];
        int i0 = fmt.v[0], i1 = fmt.v[1], i2 = fmt.v[2];
        const v3 q0 = P(i0), q1 = P(i1), q2 = P(i2);
        const v3 p0a = support_at(A, m.pia[i0]), p1a = support_at(A, m.pia[i1]), p2a = support_at(A, m.pia[i2]);
        const v3 p0b = support_at(B, m.pib[i0]), p1b = support_at(B, m.pib[i1]), p2b = support_at(B, m.pib[i2]);
        v3 nrm = N(fi);
        double fd = m.dist(fi);
        v3 pm = nrm * fd;
        v3 v0 = q1 - q0, v1 = q2 - q0, v2 = pm - q0;
        double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
        double d20 = dot(v2, v0), d21 = dot(v2, v1);
        double denom = d00 * d11 - d01 * d01;
        double v = (d11 * d20 - d01 * d21) / denom;
        double w = (d00 * d21 - d01 * d20) / denom;
        double u = 1.0 - v - w;
        v3 wa = (p0a * u + p1a * v) + p2a * w;
        v3 wb = (p0b * u + p1b * v) + p2b * w;
        out[0] = -nrm.x;
        out[1] = -nrm.y;
        out[2] = -nrm.z;
        out[3] = wa.x;
        out[4] = wa.y;
        out[5] = wa.z;
        out[6] = wb.x;
        out[7] = wb.y;
        out[8] = wb.z;
        out[9] = fd;
    }
    __device__ __forceinline__ void put_point(int i, const spt &p)
    {
        m.set_pt(i, p.p);
        m.pia[i] = p.ia;
        m.pib[i] = p.ib;
    }
};

// epa_solver::solve (collision.cpp:411-504) from a GJK simplex.
// returns 0 no contact (degenerate simplex or empty heap), 1 contact written to out,
// 3 the store's capacity was exceeded (rerun with a larger store)
template <bool UV = true, class S>
__device__ int epa_solve(const gshape &A, const gshape &B, simplex &x, S &store, double *out)
{
    static_assert(S::pcap <= 256, "point indices are packed in bytes");
    if (x.n < 4 && !pad_simplex<1, UV>(A, B, x, 0)) return 0;
    epa_solver<S> e(store);
    e.put_point(0, x.s0);
    e.put_point(1, x.s1);
    e.put_point(2, x.s2);
    e.put_point(3, x.s3);
    e.npoly = 4;
    const v3 q[4] = {x.s0.p, x.s1.p, x.s2.p, x.s3.p};
    e.initial_tetrahedron(q);
    const double tol = 1e-6;
    for (int it = 0; it < 64; ++it)
    {
        int mf = e.pop();
        if (mf < 0) break;
        v3 nrm = e.N(mf);
        double fd = store.dist(mf);
        spt p = msupport<1, UV>(A, B, nrm, 0);
        double pd = dot(nrm, p.p);
        if (pd - fd < tol)
        {
            e.barycentric(mf, A, B, out);
            return 1;
        }
        int nh = e.silhouette(mf, p.p);
        if (nh < 0) return 3;
        if (nh == 0) break;
        if (e.npoly >= S::pcap || e.nfaces + nh > S::fcap) return 3;
        int pi = e.npoly++;
        e.put_point(pi, p);
        // new faces on the horizon, linked and pushed in edge order (:481-498)
        int base = e.nfaces;
        // the first 8 horizon edges' end points (point indices, below 256) are also kept
        // in registers, one byte each, for the linking loop below
        uint64_t hs8 = 0, he8 = 0;
        for (int h = 0; h < nh; ++h)
        {
            typename S::edge ed = store.eg[h];
            e.new_face(base + h, ed.s, ed.e, pi, p.p, ed.adj);
            if (h < 8)
            {
                hs8 |= static_cast<uint64_t>(ed.s) << (8 * h);
                he8 |= static_cast<uint64_t>(ed.e) << (8 * h);
            }
        }
        e.nfaces += nh;
        // link(base + i, base + j, he_i, pi) and its mirror: a new face is (s, e, pi), so
        // its edge (e, pi) is a[1] and its edge (pi, s) is a[2]
        for (int i = 0; i < nh; ++i)
        {
            int hs_i, he_i;
            if (nh <= 8)
            {
                hs_i = static_cast<int>((hs8 >> (8 * i)) & 0xFFu);
                he_i = static_cast<int>((he8 >> (8 * i)) & 0xFFu);
            }
            else
            {
                typename S::edge ei = store.eg[i];
                hs_i = ei.s;
                he_i = ei.e;
            }
            for (int j = i + 1; j < nh; ++j)
            {
                int hs_j, he_j;
                if (nh <= 8)
                {
                    hs_j = static_cast<int>((hs8 >> (8 * j)) & 0xFFu);
                    he_j = static_cast<int>((he8 >> (8 * j)) & 0xFFu);
                }
                else
                {
                    typename S::edge ej = store.eg[j];
                    hs_j = ej.s;
                    he_j = ej.e;
                }
                if (he_i == hs_j)
                {
                    store.fm[base + i].a[1] = static_cast<typename S::idx_t>(base + j);
                    store.fm[base + j].a[2] = static_cast<typename S::idx_t>(base + i);
                }
                else if (hs_i == he_j)
                {
                    store.fm[base + j].a[1] = static_cast<typename S::idx_t>(base + i);
                    store.fm[base + i].a[2] = static_cast<typename S::idx_t>(base + j);
                }
            }
        }
    }
    int mf = e.pop();
    if (mf < 0) return 0;
    e.barycentric(mf, A, B, out);
    return 1;
}

} // namespace pk
