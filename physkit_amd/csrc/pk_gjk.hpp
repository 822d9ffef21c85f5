// GJK + EPA on the device (src/collision.cpp:12-518, include/physkit/collision/
// collision.h:27-59). Same branch structure and operation order as the reference, so
// verdicts, depths, normals and witness points are reproducible bit for bit.
// The GJK simplex lives in registers (only constant indices). The EPA polytope,
// faces and min-heap live in a caller-provided store of compile-time capacity;
// running out of capacity is reported (status 3) and never truncates silently.
#pragma once
#include "pk_math.hpp"

namespace pk
{

struct spt
{
    v3 p, pa, pb;
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
};

__device__ __forceinline__ v3 ldv(const double *v, uint32_t i) { return mk(v[3 * i], v[3 * i + 1], v[3 * i + 2]); }

__device__ __forceinline__ v3 support(const gshape &s, v3 d)
{
    if (s.kind == 2) // mesh::instance::support, mesh.cpp:442-448 -> mesh::support :341-358
    {
        v3 ld = rotate(conj(s.q), d);
        uint32_t best = 0;
        double bd = dot(ldv(s.v, 0), ld);
        for (uint32_t i = 1; i < s.nv; ++i)
        {
            double t = dot(ldv(s.v, i), ld);
            if (t > bd)
            {
                bd = t;
                best = i;
            }
        }
        return rotate(s.q, ldv(s.v, best)) + s.a;
    }
    if (s.kind == 1) // obb::support, bounds.h:539-548
    {
        v3 ld = rotate(conj(s.q), d);
        v3 sg = mk((ld.x >= 0 ? 1.0 : -1.0) * s.b.x, (ld.y >= 0 ? 1.0 : -1.0) * s.b.y,
                   (ld.z >= 0 ? 1.0 : -1.0) * s.b.z);
        return s.a + rotate(s.q, sg);
    }
    // aabb::support, bounds.h:188-199
    return mk(d.x >= 0 ? s.b.x : s.a.x, d.y >= 0 ? s.b.y : s.a.y, d.z >= 0 ? s.b.z : s.a.z);
}

__device__ __forceinline__ spt msupport(const gshape &A, const gshape &B, v3 d) // collision.h:41-49
{
    spt r;
    r.pa = support(A, d);
    r.pb = support(B, -d);
    r.p = r.pa - r.pb;
    return r;
}

struct simplex
{
    spt s0, s1, s2, s3;
    int n;
};
__device__ __forceinline__ void sx_push(simplex &x, const spt &p)
{
    if (x.n == 0)
        x.s0 = p;
    else if (x.n == 1)
        x.s1 = p;
    else if (x.n == 2)
        x.s2 = p;
    else
        x.s3 = p;
    ++x.n;
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
// (inlined so the simplex stays in registers)
__device__ __forceinline__ bool gjk(const gshape &A, const gshape &B, simplex &x)
{
    const double eps = 1e-12;
    x.n = 0;
    v3 dir = mk(1.0, 0.0, 0.0);
    spt p = msupport(A, B, dir);
    sx_push(x, p);
    if (sqn(p.p) < eps) return true;
    dir = -normalized(p.p);
    for (int it = 0; it < 100; ++it)
    {
        spt np = msupport(A, B, dir);
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
__device__ __forceinline__ bool pad_simplex(const gshape &A, const gshape &B, simplex &x)
{
    int start = x.n;
    if (start <= 1)
    {
        v3 dir = mk(1.0, 0.0, 0.0);
        spt p2 = msupport(A, B, dir);
        if (sqn(p2.p - x.s0.p) < 1e-6) p2 = msupport(A, B, -dir);
        x.s1 = p2;
        x.n = 2;
    }
    if (start <= 2)
    {
        v3 line = x.s1.p - x.s0.p;
        v3 dir = cross(normalized(line), mk(0.0, 1.0, 0.0));
        if (sqn(dir) < 1e-6) dir = cross(normalized(line), mk(0.0, 0.0, 1.0));
        dir = normalized(dir);
        spt p3 = msupport(A, B, dir);
        if (sqn(cross(line, p3.p - x.s0.p)) < 1e-6) p3 = msupport(A, B, -dir);
        x.s2 = p3;
        x.n = 3;
    }
    if (start <= 3)
    {
        v3 ab = x.s1.p - x.s0.p;
        v3 ac = x.s2.p - x.s0.p;
        v3 dir = normalized(cross(ab, ac));
        spt p4 = msupport(A, B, dir);
        if (fabs(dot(p4.p - x.s0.p, dir)) < 1e-6) p4 = msupport(A, B, -dir);
        x.s3 = p4;
        x.n = 4;
    }
    v3 ad = x.s0.p - x.s3.p;
    v3 bd = x.s1.p - x.s3.p;
    v3 cd = x.s2.p - x.s3.p;
    double triple = dot(ad, cross(bd, cd));
    return fabs(triple) > 1e-12;
}


// ---------------- EPA (collision.cpp:251-509) ----------------
constexpr uint16_t kFNull = 0xFFFF;
constexpr int kPolyCap = 68; // 4 simplex points + at most one per iteration (64)

// one record per face: a face test reads one 48-byte record
struct epa_face
{
    v3 n;        // unit normal (zero for a degenerate triangle)
    double d;    // distance of the face plane from the origin
    uint16_t v[3];
    uint16_t a[3]; // adjacent face across edge (v[i], v[i+1])
    uint8_t obsolete;
};
struct epa_edge
{
    uint16_t s, e, adj;
};

// A store is private to one EPA solve and contiguous, so the working set of a solve
// is compact in memory.
template <int FCAP> struct epa_store
{
    static constexpr int cap = FCAP;
    epa_face f[FCAP];
    spt poly[kPolyCap];
    uint16_t heap[FCAP];
    uint16_t stack[FCAP];
    epa_edge h[FCAP];
    uint16_t nf[FCAP];
};

template <class S> struct epa_ctx
{
    S &m;
    int npoly = 0, nfaces = 0, nheap = 0;

    __device__ explicit epa_ctx(S &store) : m(store) {}

    __device__ void init_face(int fi, int i, int j, int k, int opp) // :273-297
    {
        epa_face &f = m.f[fi];
        uint16_t v1 = static_cast<uint16_t>(j), v2 = static_cast<uint16_t>(k);
        v3 pi = m.poly[i].p;
        v3 ab = m.poly[j].p - pi;
        v3 ac = m.poly[k].p - pi;
        v3 nrm = cross(ab, ac);
        if (sqn(nrm) < 1e-12)
            nrm = mk(0.0, 0.0, 0.0);
        else
            nrm = normalized(nrm);
        if (opp >= 0 && dot(nrm, m.poly[opp].p - pi) > 0.0)
        {
            uint16_t t = v1;
            v1 = v2;
            v2 = t;
            nrm = -nrm;
        }
        f.v[0] = static_cast<uint16_t>(i);
        f.v[1] = v1;
        f.v[2] = v2;
        f.a[0] = kFNull;
        f.a[1] = kFNull;
        f.a[2] = kFNull;
        f.obsolete = 0;
        f.n = nrm;
        f.d = dot(nrm, pi);
    }
    __device__ void link(int f1, int f2, int va, int vb) // :305-313
    {
        epa_face &a = m.f[f1];
        epa_face &b = m.f[f2];
        int e1 = (a.v[0] == va) ? 0 : (a.v[1] == va ? 1 : 2);
        int e2 = (b.v[0] == vb) ? 0 : (b.v[1] == vb ? 1 : 2);
        a.a[e1] = static_cast<uint16_t>(f2);
        b.a[e2] = static_cast<uint16_t>(f1);
    }
    // binary min-heap on face distance; sift order of libstdc++'s push_heap/pop_heap
    __device__ bool before(int x, int y) const { return m.f[x].d > m.f[y].d; }
    __device__ void sift_up(int hole, int top, uint16_t val)
    {
        double dv = m.f[val].d;
        int parent = (hole - 1) / 2;
        while (hole > top)
        {
            uint16_t pv = m.heap[parent];
            if (!(m.f[pv].d > dv)) break;
            m.heap[hole] = pv;
            hole = parent;
            parent = (hole - 1) / 2;
        }
        m.heap[hole] = val;
    }
    __device__ void push(int f)
    {
        m.heap[nheap] = static_cast<uint16_t>(f);
        ++nheap;
        sift_up(nheap - 1, 0, static_cast<uint16_t>(f));
    }
    __device__ int pop()
    {
        while (nheap > 0)
        {
            int len = nheap;
            uint16_t top = m.heap[0];
            if (len > 1)
            {
                uint16_t val = m.heap[len - 1];
                m.heap[len - 1] = top;
                int n = len - 1;
                int hole = 0, child = 0;
                while (child < (n - 1) / 2)
                {
                    child = 2 * (child + 1);
                    uint16_t r = m.heap[child], l = m.heap[child - 1];
                    if (before(r, l))
                    {
                        child--;
                        r = l;
                    }
                    m.heap[hole] = r;
                    hole = child;
                }
                if ((n & 1) == 0 && child == (n - 2) / 2)
                {
                    child = 2 * (child + 1);
                    m.heap[hole] = m.heap[child - 1];
                    hole = child - 1;
                }
                sift_up(hole, 0, val);
            }
            --nheap;
            if (!m.f[top].obsolete) return top;
        }
        return -1;
    }
    __device__ int alloc_face()
    {
        if (nfaces >= S::cap) return -1;
        return nfaces++;
    }
    // :315-353; returns the number of horizon edges, -1 on overflow
    __device__ int silhouette(int f0, v3 p)
    {
        const double tol = 1e-6;
        int sp = 0, nh = 0;
        m.stack[sp++] = static_cast<uint16_t>(f0);
        m.f[f0].obsolete = 1;
        while (sp > 0)
        {
            int cur = m.stack[--sp];
            uint16_t cv[3] = {m.f[cur].v[0], m.f[cur].v[1], m.f[cur].v[2]};
            uint16_t ca[3] = {m.f[cur].a[0], m.f[cur].a[1], m.f[cur].a[2]};
            for (int i = 0; i < 3; ++i)
            {
                int nb = ca[i];
                if (nb == kFNull) continue;
                epa_face &g = m.f[nb];
                if (g.obsolete) continue;
                if (dot(g.n, p) > g.d + tol)
                {
                    g.obsolete = 1;
                    m.stack[sp++] = static_cast<uint16_t>(nb); // each face is pushed at most once
                }
                else
                {
                    if (nh >= S::cap) return -1;
                    epa_edge ed;
                    ed.s = cv[i];
                    ed.e = cv[(i + 1) % 3];
                    ed.adj = static_cast<uint16_t>(nb);
                    m.h[nh++] = ed;
                }
            }
        }
        return nh;
    }
    __device__ void initial_tetrahedron() // :355-388
    {
        const int def[4][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2}, {1, 3, 2, 0}};
        for (int f = 0; f < 4; ++f)
        {
            init_face(f, def[f][0], def[f][1], def[f][2], def[f][3]);
            nfaces = f + 1;
            push(f);
        }
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j)
                for (int e1 = 0; e1 < 3; ++e1)
                {
                    int u1 = m.f[i].v[e1], v1 = m.f[i].v[(e1 + 1) % 3];
                    for (int e2 = 0; e2 < 3; ++e2)
                    {
                        int u2 = m.f[j].v[e2], v2 = m.f[j].v[(e2 + 1) % 3];
                        if (u1 == v2 && v1 == u2)
                        {
                            m.f[i].a[e1] = static_cast<uint16_t>(j);
                            m.f[j].a[e2] = static_cast<uint16_t>(i);
                        }
                    }
                }
    }
    // :424-454; out = normal3, world_a3, world_b3, depth
    __device__ void barycentric(int fi, double *out) const
    {
        const epa_face &f = m.f[fi];
        const spt &p0 = m.poly[f.v[0]];
        const spt &p1 = m.poly[f.v[1]];
        const spt &p2 = m.poly[f.v[2]];
        v3 nrm = f.n;
        double fd = f.d;
        v3 pm = nrm * fd;
        v3 v0 = p1.p - p0.p, v1 = p2.p - p0.p, v2 = pm - p0.p;
        double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
        double d20 = dot(v2, v0), d21 = dot(v2, v1);
        double denom = d00 * d11 - d01 * d01;
        double v = (d11 * d20 - d01 * d21) / denom;
        double w = (d00 * d21 - d01 * d20) / denom;
        double u = 1.0 - v - w;
        v3 wa = (p0.pa * u + p1.pa * v) + p2.pa * w;
        v3 wb = (p0.pb * u + p1.pb * v) + p2.pb * w;
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
};

// epa_solver::solve (collision.cpp:411-504) from a GJK simplex.
// returns 0 no contact (degenerate simplex or empty heap), 1 contact written to out,
// 3 the store's capacity was exceeded (rerun with a larger store)
template <class S>
__device__ int epa_solve(const gshape &A, const gshape &B, simplex &x, S &store, double *out)
{
    if (x.n < 4 && !pad_simplex(A, B, x)) return 0;
    epa_ctx<S> e(store);
    store.poly[0] = x.s0;
    store.poly[1] = x.s1;
    store.poly[2] = x.s2;
    store.poly[3] = x.s3;
    e.npoly = 4;
    e.initial_tetrahedron();
    const double tol = 1e-6;
    for (int it = 0; it < 64; ++it)
    {
        int mf = e.pop();
        if (mf < 0) break;
        v3 nrm = store.f[mf].n;
        double fd = store.f[mf].d;
        spt p = msupport(A, B, nrm);
        double pd = dot(nrm, p.p);
        if (pd - fd < tol)
        {
            e.barycentric(mf, out);
            return 1;
        }
        int nh = e.silhouette(mf, p.p);
        if (nh < 0) return 3;
        if (nh == 0) break;
        store.poly[e.npoly] = p;
        int pi = e.npoly++;
        for (int h = 0; h < nh; ++h)
        {
            int f = e.alloc_face();
            if (f < 0) return 3;
            epa_edge ed = store.h[h];
            e.init_face(f, ed.s, ed.e, pi, -1);
            e.link(f, ed.adj, ed.s, ed.e);
            e.push(f);
            store.nf[h] = static_cast<uint16_t>(f);
        }
        for (int i = 0; i < nh; ++i)
            for (int j = i + 1; j < nh; ++j)
            {
                epa_edge hi = store.h[i], hj = store.h[j];
                if (hi.e == hj.s)
                    e.link(store.nf[i], store.nf[j], hi.e, pi);
                else if (hi.s == hj.e)
                    e.link(store.nf[j], store.nf[i], hj.e, pi);
            }
    }
    int mf = e.pop();
    if (mf < 0) return 0;
    e.barycentric(mf, out);
    return 1;
}

} // namespace pk
