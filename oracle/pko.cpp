// Oracle (test infrastructure only): FP64 CPU restatement of PhysKit's world step.
// Each function cites the reference lines it follows. See pko.hpp for the
// determinism choices.
#include "pko.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace pko
{

// ---------------- bounds ----------------
aabb aabb_from_points(const v3 *p, std::size_t n) // bounds.h:33-49
{
    aabb b{p[0], p[0]};
    for (std::size_t i = 1; i < n; ++i)
    {
        b.min.x = smin(b.min.x, p[i].x);
        b.min.y = smin(b.min.y, p[i].y);
        b.min.z = smin(b.min.z, p[i].z);
        b.max.x = smax(b.max.x, p[i].x);
        b.max.y = smax(b.max.y, p[i].y);
        b.max.z = smax(b.max.z, p[i].z);
    }
    return b;
}

aabb aabb_rotate(const aabb &b, qt q) // bounds.h:142-158 (corner 0 first, then 1..7)
{
    v3 p0 = rotate(q, b.min);
    aabb r{p0, p0};
    for (unsigned i = 1; i < 8; ++i)
    {
        v3 pt = rotate(q, b.point(i));
        r.min.x = smin(r.min.x, pt.x);
        r.min.y = smin(r.min.y, pt.y);
        r.min.z = smin(r.min.z, pt.z);
        r.max.x = smax(r.max.x, pt.x);
        r.max.y = smax(r.max.y, pt.y);
        r.max.z = smax(r.max.z, pt.z);
    }
    return r;
}

// ---------------- meshes ----------------
mesh make_mesh(const v3 *verts, std::size_t n)
{
    mesh m;
    m.verts.assign(verts, verts + n);
    m.local = aabb_from_points(verts, n);
    return m;
}

mesh mesh_box(v3 h) // mesh.cpp:25-58, vertex order 0..7
{
    v3 v[8] = {{-h.x, -h.y, -h.z}, {+h.x, -h.y, -h.z}, {+h.x, +h.y, -h.z}, {-h.x, +h.y, -h.z},
               {-h.x, -h.y, +h.z}, {+h.x, -h.y, +h.z}, {+h.x, +h.y, +h.z}, {-h.x, +h.y, +h.z}};
    return make_mesh(v, 8);
}

mesh mesh_sphere(double r, unsigned stacks, unsigned sectors) // mesh.cpp:60-115 (vertices)
{
    std::vector<v3> v;
    v.push_back({0.0, r, 0.0});
    const double pi = 3.141592653589793;
    for (unsigned i = 1; i < stacks; ++i)
    {
        double phi = pi * static_cast<double>(i) / static_cast<double>(stacks);
        for (unsigned j = 0; j < sectors; ++j)
        {
            double theta = 2.0 * pi * static_cast<double>(j) / static_cast<double>(sectors);
            v.push_back({r * std::sin(phi) * std::cos(theta), r * std::cos(phi),
                         r * std::sin(phi) * std::sin(theta)});
        }
    }
    v.push_back({0.0, -r, 0.0});
    return make_mesh(v.data(), v.size());
}

mesh mesh_pyramid(double bh, double height) // mesh.cpp:117-137
{
    v3 v[5] = {{-bh, 0.0, -bh}, {+bh, 0.0, -bh}, {+bh, 0.0, +bh}, {-bh, 0.0, +bh}, {0.0, height, 0.0}};
    return make_mesh(v, 5);
}

// ---------------- GJK (collision.cpp:12-189) ----------------
namespace
{
struct simplex
{
    support_pt s[4];
    int n = 0;
    void push(const support_pt &p) { s[n++] = p; }
    void erase(int i)
    {
        for (int k = i; k + 1 < n; ++k) s[k] = s[k + 1];
        --n;
    }
};

bool handle_line(simplex &sx, v3 &dir) // collision.cpp:12-41
{
    const double eps = 1e-12;
    const support_pt a = sx.s[1], b = sx.s[0];
    v3 ab = b.p - a.p, ao = -a.p;
    if (dot(ab, ao) > 0.0)
    {
        v3 triple = cross(cross(ab, ao), ab);
        if (sqn(triple) < eps)
        {
            v3 abh = normalized(ab);
            v3 perp = cross(abh, v3{0.0, 1.0, 0.0});
            if (sqn(perp) < eps) perp = cross(abh, v3{0.0, 0.0, 1.0});
            dir = normalized(perp);
        }
        else
            dir = normalized(triple);
    }
    else
    {
        sx.erase(0);
        dir = normalized(ao);
    }
    return false;
}

bool handle_triangle(simplex &sx, v3 &dir) // collision.cpp:43-88
{
    const support_pt a = sx.s[2], b = sx.s[1], c = sx.s[0];
    v3 ab = b.p - a.p, ac = c.p - a.p, ao = -a.p;
    v3 abc = cross(ab, ac);
    v3 ab_perp = cross(ab, abc);
    if (dot(ab_perp, ao) > 0.0)
    {
        sx.erase(0);
        v3 triple = cross(cross(ab, ao), ab);
        dir = (sqn(triple) < 1e-12) ? normalized(ao) : normalized(triple);
        return false;
    }
    v3 ac_perp = cross(abc, ac);
    if (dot(ac_perp, ao) > 0.0)
    {
        sx.erase(1);
        v3 triple = cross(cross(ac, ao), ac);
        dir = (sqn(triple) < 1e-12) ? normalized(ao) : normalized(triple);
        return false;
    }
    if (dot(abc, ao) <= 0.0)
    {
        std::swap(sx.s[0], sx.s[1]);
        dir = normalized(-abc);
    }
    else
        dir = normalized(abc);
    return false;
}

bool handle_tetrahedron(simplex &sx, v3 &dir) // collision.cpp:90-147
{
    const support_pt a = sx.s[3], b = sx.s[2], c = sx.s[1], d = sx.s[0];
    v3 ao = -a.p;
    v3 abc = cross(b.p - a.p, c.p - a.p);
    v3 acd = cross(c.p - a.p, d.p - a.p);
    v3 adb = cross(d.p - a.p, b.p - a.p);
    auto orient = [](v3 &nrm, const support_pt &fa, const support_pt &opp)
    {
        if (dot(nrm, opp.p - fa.p) > 0.0) nrm = -nrm;
    };
    orient(abc, a, d);
    orient(acd, a, b);
    orient(adb, a, c);
    if (dot(abc, ao) > 0.0)
    {
        sx.n = 0;
        sx.push(c);
        sx.push(b);
        sx.push(a);
        dir = normalized(abc);
        return handle_triangle(sx, dir);
    }
    if (dot(acd, ao) > 0.0)
    {
        sx.n = 0;
        sx.push(d);
        sx.push(c);
        sx.push(a);
        dir = normalized(acd);
        return handle_triangle(sx, dir);
    }
    if (dot(adb, ao) > 0.0)
    {
        sx.n = 0;
        sx.push(b);
        sx.push(d);
        sx.push(a);
        dir = normalized(adb);
        return handle_triangle(sx, dir);
    }
    return true;
}

bool handle_simplex(simplex &sx, v3 &dir) // collision.cpp:149-162
{
    switch (sx.n)
    {
    case 2: return handle_line(sx, dir);
    case 3: return handle_triangle(sx, dir);
    case 4: return handle_tetrahedron(sx, dir);
    default: return false;
    }
}

template <class A, class B> support_pt msupport(const A &a, const B &b, v3 d) // collision.h:41-49
{
    v3 pa = a.support(d);
    v3 pb = b.support(-d);
    return {pa - pb, pa, pb};
}

template <class A, class B>
bool gjk(const A &a, const B &b, simplex &sx, gjk_stats *st) // collision.cpp:165-189
{
    const double eps = 1e-12;
    v3 dir{1.0, 0.0, 0.0};
    support_pt p = msupport(a, b, dir);
    sx.push(p);
    if (sqn(p.p) < eps) return true;
    dir = -normalized(p.p);
    for (int it = 0; it < 100; ++it)
    {
        if (st) st->gjk_iters = it + 1;
        support_pt np = msupport(a, b, dir);
        if (dot(np.p, dir) <= 0.0) return false;
        sx.push(np);
        if (handle_simplex(sx, dir)) return true;
    }
    return false;
}

template <class A, class B> bool pad_simplex(const A &a, const B &b, simplex &sx) // :191-248
{
    int start = sx.n;
    if (start <= 1)
    {
        v3 dir{1.0, 0.0, 0.0};
        support_pt p2 = msupport(a, b, dir);
        if (sqn(p2.p - sx.s[0].p) < 1e-6) p2 = msupport(a, b, -dir);
        sx.push(p2);
    }
    if (start <= 2)
    {
        v3 line = sx.s[1].p - sx.s[0].p;
        v3 dir = cross(normalized(line), v3{0.0, 1.0, 0.0});
        if (sqn(dir) < 1e-6) dir = cross(normalized(line), v3{0.0, 0.0, 1.0});
        dir = normalized(dir);
        support_pt p3 = msupport(a, b, dir);
        if (sqn(cross(line, p3.p - sx.s[0].p)) < 1e-6) p3 = msupport(a, b, -dir);
        sx.push(p3);
    }
    if (start <= 3)
    {
        v3 ab = sx.s[1].p - sx.s[0].p;
        v3 ac = sx.s[2].p - sx.s[0].p;
        v3 dir = normalized(cross(ab, ac));
        support_pt p4 = msupport(a, b, dir);
        if (std::fabs(dot(p4.p - sx.s[0].p, dir)) < 1e-6) p4 = msupport(a, b, -dir);
        sx.push(p4);
    }
    v3 ad = sx.s[0].p - sx.s[3].p;
    v3 bd = sx.s[1].p - sx.s[3].p;
    v3 cd = sx.s[2].p - sx.s[3].p;
    double triple = dot(ad, cross(bd, cd));
    return std::fabs(triple) > 1e-12;
}

// ---------------- EPA (collision.cpp:251-509) ----------------
constexpr uint32_t kNull = 0xFFFF;

struct face
{
    uint32_t v[3];
    uint32_t adj[3] = {kNull, kNull, kNull};
    v3 normal;
    double dist = 0;
    bool obsolete = false;
};

struct epa
{
    std::vector<support_pt> poly;
    std::vector<face> faces;
    std::vector<uint32_t> heap;

    void init_face(uint32_t f, uint32_t i, uint32_t j, uint32_t k, uint32_t opp) // :273-297
    {
        face &fc = faces[f];
        fc.v[0] = i;
        fc.v[1] = j;
        fc.v[2] = k;
        v3 ab = poly[j].p - poly[i].p;
        v3 ac = poly[k].p - poly[i].p;
        fc.normal = cross(ab, ac);
        if (sqn(fc.normal) < 1e-12)
            fc.normal = v3{0.0, 0.0, 0.0};
        else
            fc.normal = normalized(fc.normal);
        if (opp != kNull && dot(fc.normal, poly[opp].p - poly[i].p) > 0.0)
        {
            std::swap(fc.v[1], fc.v[2]);
            fc.normal = -fc.normal;
        }
        fc.dist = dot(fc.normal, poly[i].p);
    }
    uint32_t alloc_face()
    {
        faces.emplace_back();
        return static_cast<uint32_t>(faces.size() - 1);
    }
    void link(uint32_t f1, uint32_t f2, uint32_t va, uint32_t vb) // :305-313
    {
        face &a = faces[f1];
        face &b = faces[f2];
        int e1 = (a.v[0] == va) ? 0 : (a.v[1] == va ? 1 : 2);
        int e2 = (b.v[0] == vb) ? 0 : (b.v[1] == vb ? 1 : 2);
        a.adj[e1] = f2;
        b.adj[e2] = f1;
    }
    // binary min-heap on distance with the sift order of libstdc++'s push_heap/pop_heap
    // (comparator "less" = faces[x].dist > faces[y].dist), :390-408
    bool before(uint32_t x, uint32_t y) const { return faces[x].dist > faces[y].dist; }
    void sift_up(std::size_t hole, std::size_t top, uint32_t val)
    {
        std::size_t parent = (hole - 1) / 2;
        while (hole > top && before(heap[parent], val))
        {
            heap[hole] = heap[parent];
            hole = parent;
            parent = (hole - 1) / 2;
        }
        heap[hole] = val;
    }
    void push(uint32_t f)
    {
        heap.push_back(f);
        sift_up(heap.size() - 1, 0, f);
    }
    uint32_t pop()
    {
        while (!heap.empty())
        {
            std::size_t len = heap.size();
            uint32_t top = heap[0];
            if (len > 1)
            {
                uint32_t val = heap[len - 1];
                heap[len - 1] = top;
                std::size_t n = len - 1; // heap range [0, n)
                std::size_t hole = 0, child = 0;
                while (child < (n - 1) / 2)
                {
                    child = 2 * (child + 1);
                    if (before(heap[child], heap[child - 1])) child--;
                    heap[hole] = heap[child];
                    hole = child;
                }
                if ((n & 1) == 0 && child == (n - 2) / 2)
                {
                    child = 2 * (child + 1);
                    heap[hole] = heap[child - 1];
                    hole = child - 1;
                }
                sift_up(hole, 0, val);
            }
            heap.pop_back();
            if (!faces[top].obsolete) return top;
        }
        return kNull;
    }
    struct edge
    {
        uint32_t s, e, adj;
    };
    void silhouette(uint32_t f0, v3 p, std::vector<edge> &horizon) // :315-353
    {
        const double tol = 1e-6;
        std::vector<uint32_t> stack{f0};
        faces[f0].obsolete = true;
        horizon.clear();
        while (!stack.empty())
        {
            uint32_t cur = stack.back();
            stack.pop_back();
            for (int i = 0; i < 3; ++i)
            {
                uint32_t nb = faces[cur].adj[i];
                if (nb == kNull) continue;
                if (faces[nb].obsolete) continue;
                if (dot(faces[nb].normal, p) > faces[nb].dist + tol)
                {
                    faces[nb].obsolete = true;
                    stack.push_back(nb);
                }
                else
                    horizon.push_back({faces[cur].v[i], faces[cur].v[(i + 1) % 3], nb});
            }
        }
    }
    void initial_tetrahedron() // :355-388
    {
        auto mk = [&](uint32_t i, uint32_t j, uint32_t k, uint32_t opp)
        {
            uint32_t f = alloc_face();
            init_face(f, i, j, k, opp);
            push(f);
        };
        mk(0, 1, 2, 3);
        mk(0, 2, 3, 1);
        mk(0, 3, 1, 2);
        mk(1, 3, 2, 0);
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j)
                for (int e1 = 0; e1 < 3; ++e1)
                {
                    uint32_t u1 = faces[i].v[e1], v1 = faces[i].v[(e1 + 1) % 3];
                    for (int e2 = 0; e2 < 3; ++e2)
                    {
                        uint32_t u2 = faces[j].v[e2], v2 = faces[j].v[(e2 + 1) % 3];
                        if (u1 == v2 && v1 == u2)
                        {
                            faces[i].adj[e1] = static_cast<uint32_t>(j);
                            faces[j].adj[e2] = static_cast<uint32_t>(i);
                        }
                    }
                }
    }
    collision_info barycentric(const face &f) const // :424-454
    {
        const support_pt &p0 = poly[f.v[0]], &p1 = poly[f.v[1]], &p2 = poly[f.v[2]];
        v3 pm = f.normal * f.dist;
        v3 v0 = p1.p - p0.p, v1 = p2.p - p0.p, v2 = pm - p0.p;
        double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
        double d20 = dot(v2, v0), d21 = dot(v2, v1);
        double denom = d00 * d11 - d01 * d01;
        double v = (d11 * d20 - d01 * d21) / denom;
        double w = (d00 * d21 - d01 * d20) / denom;
        double u = 1.0 - v - w;
        collision_info ci;
        ci.normal = -f.normal;
        ci.world_a = (p0.pa * u + p1.pa * v) + p2.pa * w;
        ci.world_b = (p0.pb * u + p1.pb * v) + p2.pb * w;
        ci.depth = f.dist;
        return ci;
    }
};

template <class A, class B>
std::optional<collision_info> epa_solve(const A &a, const B &b, simplex &sx, gjk_stats *st) // :411-504
{
    if (sx.n < 4 && !pad_simplex(a, b, sx)) return std::nullopt;
    epa s;
    s.poly.assign(sx.s, sx.s + 4);
    s.initial_tetrahedron();
    const double tol = 1e-6;
    std::vector<epa::edge> horizon;
    std::vector<uint32_t> nf;
    for (int it = 0; it < 64; ++it)
    {
        if (st) st->epa_iters = it + 1;
        uint32_t mf = s.pop();
        if (mf == kNull) break;
        v3 nrm = s.faces[mf].normal;
        double fd = s.faces[mf].dist;
        support_pt p = msupport(a, b, nrm);
        double pd = dot(nrm, p.p);
        if (pd - fd < tol)
        {
            if (st)
            {
                st->polytope = static_cast<int>(s.poly.size());
                st->faces = static_cast<int>(s.faces.size());
            }
            return s.barycentric(s.faces[mf]);
        }
        s.silhouette(mf, p.p, horizon);
        if (horizon.empty()) break;
        s.poly.push_back(p);
        uint32_t pi = static_cast<uint32_t>(s.poly.size() - 1);
        nf.clear();
        for (const auto &e : horizon)
        {
            uint32_t f = s.alloc_face();
            s.init_face(f, e.s, e.e, pi, kNull);
            s.link(f, e.adj, e.s, e.e);
            s.push(f);
            nf.push_back(f);
        }
        std::size_t n = nf.size();
        for (std::size_t i = 0; i < n; ++i)
            for (std::size_t j = i + 1; j < n; ++j)
            {
                if (horizon[i].e == horizon[j].s)
                    s.link(nf[i], nf[j], horizon[i].e, pi);
                else if (horizon[i].s == horizon[j].e)
                    s.link(nf[j], nf[i], horizon[j].e, pi);
            }
    }
    if (st)
    {
        st->polytope = static_cast<int>(s.poly.size());
        st->faces = static_cast<int>(s.faces.size());
    }
    uint32_t mf = s.pop();
    if (mf == kNull) return std::nullopt;
    return s.barycentric(s.faces[mf]);
}

template <class A, class B>
std::optional<collision_info> gjk_epa_t(const A &a, const B &b, gjk_stats *st) // :512-518
{
    simplex sx;
    if (gjk(a, b, sx, st))
    {
        if (st) st->hit = true;
        return epa_solve(a, b, sx, st);
    }
    return std::nullopt;
}
} // namespace

std::optional<collision_info> gjk_epa(const shape_ref &a, const shape_ref &b, gjk_stats *st)
{ return gjk_epa_t(a, b, st); }
std::optional<collision_info> gjk_epa(const mesh_instance &a, const mesh_instance &b,
                                      gjk_stats *st)
{ return gjk_epa_t(a, b, st); }

// ---------------- bodies ----------------
void body::update_derived() // particle.h:140-146
{
    R = rotation_matrix(q);
    m3 rt = transpose(R);
    Iw = mul(mul(R, I_local), rt);
    Iw_inv = mul(mul(R, Iinv_local), rt);
}

qt exp_map(v3 w, double dt) // integrate.h:20-31
{
    v3 ang = w * dt;
    double mag = norm(ang);
    if (mag < 1e-12)
    {
        v3 half = ang * 0.5;
        return qnormalized(qt{1.0, half.x, half.y, half.z});
    }
    v3 axis = ang / mag;
    double ha = 0.5 * mag;
    double s, c;
    sincos_pk(ha, &s, &c);
    return qt{c, s * axis.x, s * axis.y, s * axis.z};
}

// ---------------- manifold (collision_phases.h:124-198) ----------------
void manifold::add_contact(const contact &k)
{
    if (n < 4)
        c[n++] = k;
    else
        add_reduce(k);
}

void manifold::add_reduce(const contact &k)
{
    contact pool[5] = {c[0], c[1], c[2], c[3], k};
    int best[4] = {0, 1, 2, 3};
    for (int i = 1; i < 5; ++i)
        if (pool[i].depth > pool[best[0]].depth) best[0] = i;
    double md = -1.0;
    for (int i = 0; i < 5; ++i)
    {
        if (i == best[0]) continue;
        double d2 = sqn(pool[i].local_a - pool[best[0]].local_a);
        if (d2 > md)
        {
            md = d2;
            best[1] = i;
        }
    }
    double ma = -1.0;
    v3 e0 = pool[best[1]].local_a - pool[best[0]].local_a;
    for (int i = 0; i < 5; ++i)
    {
        if (i == best[0] || i == best[1]) continue;
        v3 e1 = pool[i].local_a - pool[best[0]].local_a;
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
        double d2 = sqn(pool[i].local_a - pool[best[2]].local_a);
        if (d2 > md)
        {
            md = d2;
            best[3] = i;
        }
    }
    n = 0;
    for (int i : best) c[n++] = pool[i];
}

uint32_t colour_mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// ---------------- world ----------------
void world::finalize_world() { gnorm = norm(gravity); }

uint32_t world::add_body(const body &in) // world.h:202-208, particle.h:16-28
{
    body b = in;
    b.inv_mass = 1.0 / b.mass;
    b.Iinv_local = inverse(b.I_local);
    if (b.inv_mass == 0.0) b.Iinv_local = m3_zero();
    b.acc = {};
    b.torque = {};
    b.update_derived();
    mesh_instance inst{&shapes[b.shape], b.pos, b.q};
    b.fat = inst.bounds(); // dynamic_bvh::add stores exact bounds
    if (b.disabled) // a removed body's slot has no leaf (dynamic_bvh::remove_leaf, collision/bvh.h:307)
    {
        const double nan = std::numeric_limits<double>::quiet_NaN();
        b.fat.min = {nan, nan, nan};
        b.fat.max = b.fat.min;
    }
    b.moved = false;
    bodies.push_back(b);
    return static_cast<uint32_t>(bodies.size() - 1);
}

void world::stage_integrate_vel(double dt) // world.cpp:22-34
{
    counters.moved = 0;
    for (auto &b : bodies)
    {
        b.moved = false;
        if (b.is_static) continue;
        v3 force = gravity * b.mass; // particle.h:78-79
        b.acc = b.acc + force * b.inv_mass;
        b.vel = b.vel + b.acc * dt; // integrate.h:36-40
        v3 aa = mul(b.Iw_inv, b.torque - cross(b.angvel, mul(b.Iw, b.angvel))); // particle.h:72-76
        b.angvel = b.angvel + aa * dt;
        mesh_instance inst{&shapes[b.shape], b.pos, b.q};
        aabb tb = inst.bounds();
        v3 disp = b.vel * dt;
        if (!b.fat.contains(tb)) // bvh.cpp:475-514
        {
            const double margin = 0.1;
            aabb nb{tb.min - v3{margin, margin, margin}, tb.max + v3{margin, margin, margin}};
            if (disp.x < 0.0) nb.min.x = nb.min.x + disp.x; else nb.max.x = nb.max.x + disp.x;
            if (disp.y < 0.0) nb.min.y = nb.min.y + disp.y; else nb.max.y = nb.max.y + disp.y;
            if (disp.z < 0.0) nb.min.z = nb.min.z + disp.z; else nb.max.z = nb.max.z + disp.z;
            b.fat = nb;
            b.moved = true;
            counters.moved++;
        }
        b.acc = {};
        b.torque = {};
    }
}

// collision_phases.h:377-436 in closed form:
// active'(a,b) = a!=b && !(static_a && static_b) && overlap(fat_a, fat_b)
//                && (active(a,b) || moved_a || moved_b)
void world::stage_pairs()
{
    // a box with a NaN coordinate fails every comparison of aabb::intersects, so it
    // overlaps nothing; leaving it out keeps the sort below a strict weak ordering
    auto has_nan = [](const aabb &b)
    {
        return std::isnan(b.min.x) || std::isnan(b.min.y) || std::isnan(b.min.z) || std::isnan(b.max.x) ||
               std::isnan(b.max.y) || std::isnan(b.max.z);
    };
    std::vector<uint32_t> order;
    order.reserve(bodies.size());
    for (std::size_t i = 0; i < bodies.size(); ++i)
        if (!has_nan(bodies[i].fat)) order.push_back(static_cast<uint32_t>(i));
    const std::size_t n = order.size();
    std::sort(order.begin(), order.end(),
              [&](uint32_t x, uint32_t y)
              {
                  if (bodies[x].fat.min.x != bodies[y].fat.min.x)
                      return bodies[x].fat.min.x < bodies[y].fat.min.x;
                  return x < y;
              });
    auto moved = [&](uint32_t i)
    { return !bodies[i].is_static && (all_moved || bodies[i].moved); };
    std::vector<uint64_t> next;
    for (std::size_t oi = 0; oi < n; ++oi)
    {
        uint32_t i = order[oi];
        const aabb &bi = bodies[i].fat;
        for (std::size_t oj = oi + 1; oj < n; ++oj)
        {
            uint32_t j = order[oj];
            const aabb &bj = bodies[j].fat;
            if (bj.min.x > bi.max.x) break;
            if (bodies[i].is_static && bodies[j].is_static) continue;
            if (!bi.intersects(bj)) continue;
            uint64_t k = pair_key(i, j);
            if (moved(i) || moved(j) || std::binary_search(active.begin(), active.end(), k))
                next.push_back(k);
        }
    }
    std::sort(next.begin(), next.end());
    // carry manifolds of surviving pairs; new pairs start empty (on_pair_added :218-223)
    std::vector<manifold> nm(next.size());
    std::size_t p = 0;
    for (std::size_t i = 0; i < next.size(); ++i)
    {
        while (p < active.size() && active[p] < next[i]) ++p;
        if (p < active.size() && active[p] == next[i])
            nm[i] = manifolds[p];
        else
        {
            nm[i] = manifold{};
            nm[i].a = static_cast<uint32_t>(next[i] >> 32);
            nm[i].b = static_cast<uint32_t>(next[i] & 0xFFFFFFFFu);
        }
    }
    active.swap(next);
    manifolds.swap(nm);
    counters.pairs = active.size();
}

void world::stage_narrow() // collision_phases.h:244-320
{
    events.clear();
    counters.gjk_hits = 0;
    counters.contacts = 0;
    const double distance2_eps = 0.005 * 0.005;
    const double breaking = 0.05;
    const double drift2_eps = 0.06;
    for (auto &man : manifolds)
    {
        body &A = bodies[man.a];
        body &B = bodies[man.b];
        mesh_instance ia{&shapes[A.shape], A.pos, A.q};
        mesh_instance ib{&shapes[B.shape], B.pos, B.q};
        gjk_stats st;
        auto col = gjk_epa(ia, ib, &st);
        if (st.hit) counters.gjk_hits++;
        std::optional<contact> nc;
        if (col && epa_guard)
        {
            const collision_info &ci = *col;
            bool ok = sqn(ci.normal) > 0.0 && std::isfinite(ci.world_a.x) && std::isfinite(ci.world_a.y) &&
                      std::isfinite(ci.world_a.z) && std::isfinite(ci.world_b.x) && std::isfinite(ci.world_b.y) &&
                      std::isfinite(ci.world_b.z) && std::isfinite(ci.depth);
            if (!ok) col.reset();
        }
        if (col)
        {
            contact k;
            k.normal = col->normal;
            k.local_a = rotate(conj(A.q), col->world_a - A.pos); // particle.h:107-108
            k.local_b = rotate(conj(B.q), col->world_b - B.pos);
            k.depth = col->depth;
            nc = k;
        }
        manifold nmf;
        nmf.a = man.a;
        nmf.b = man.b;
        for (int i = 0; i < man.n; ++i)
        {
            contact old = man.c[i];
            if (nc)
            {
                if (sqn(nc->local_a - old.local_a) < distance2_eps ||
                    sqn(nc->local_b - old.local_b) < distance2_eps)
                {
                    nc->ln = old.ln;
                    nc->lt1 = old.lt1;
                    nc->lt2 = old.lt2;
                    continue;
                }
            }
            v3 wa = rotate(A.q, old.local_a) + A.pos; // particle.h:110-111
            v3 wb = rotate(B.q, old.local_b) + B.pos;
            v3 nrm = nc ? nc->normal : old.normal;
            v3 rel = wb - wa;
            double depth = dot(rel, nrm);
            v3 pa = wa + nrm * depth;
            double drift2 = sqn(pa - wb);
            if (depth > -breaking && drift2 < drift2_eps)
            {
                old.depth = depth;
                old.normal = nrm;
                nmf.add_contact(old);
            }
        }
        if (nc) nmf.add_contact(*nc);
        bool was = man.n > 0, is = nmf.n > 0;
        man = nmf;
        uint64_t key = pair_key(man.a, man.b);
        if (!was && is)
            events.push_back({key, 1});
        else if (was && !is)
            events.push_back({key, 2});
        counters.contacts += static_cast<uint64_t>(man.n);
    }
}

namespace
{
struct jrow // constraint.h:29-47
{
    v3 Jv, Jwa, Jwb;
    double Meff = 0, bias = 0, gamma = 0, acc = 0, lo = -DBL_MAX, hi = DBL_MAX;
};
struct cpoint // constraint.h:1203-1212
{
    jrow n, t1, t2;
    double mu = 0, i11 = 0, i12 = 0, i22 = 0;
    contact *cache = nullptr;
    body *A = nullptr, *B = nullptr;
};

inline double row_jv(const jrow &r, const body &A, const body &B) // constraint.h:52-53
{ return ((dot(r.Jv, A.vel) + dot(r.Jwa, A.angvel)) - dot(r.Jv, B.vel)) + dot(r.Jwb, B.angvel); }

void row_apply(jrow &r, body &A, body &B) // constraint.h:49-72
{
    double jv = row_jv(r, A, B);
    double cfp = r.gamma * r.acc;
    double lambda = -((jv + r.bias) + cfp) * r.Meff;
    double old = r.acc;
    double v = old + lambda;
    r.acc = (v < r.lo) ? r.lo : ((r.hi < v) ? r.hi : v);
    double ap = r.acc - old;
    if (!A.is_static)
    {
        A.vel = A.vel + (r.Jv * ap) * A.inv_mass;
        A.angvel = A.angvel + mul(A.Iw_inv, r.Jwa * ap);
    }
    if (!B.is_static)
    {
        B.vel = B.vel + ((-r.Jv) * ap) * B.inv_mass;
        B.angvel = B.angvel + mul(B.Iw_inv, r.Jwb * ap);
    }
}

void row_warm(const jrow &r, body &A, body &B) // constraint.h:74-87
{
    if (r.acc == 0.0) return;
    if (!A.is_static)
    {
        A.vel = A.vel + (r.Jv * r.acc) * A.inv_mass;
        A.angvel = A.angvel + mul(A.Iw_inv, r.Jwa * r.acc);
    }
    if (!B.is_static)
    {
        B.vel = B.vel + ((-r.Jv) * r.acc) * B.inv_mass;
        B.angvel = B.angvel + mul(B.Iw_inv, r.Jwb * r.acc);
    }
}

void basis(v3 n, v3 &t1, v3 &t2) // constraint.h:104-113
{
    double sign = std::copysign(1.0, n.z);
    double a = -1.0 / (sign + n.z);
    double b = n.x * n.y * a;
    t1 = {1.0 + sign * n.x * n.x * a, sign * b, -sign * n.x};
    t2 = {b, sign + n.y * n.y * a, -n.y};
}

jrow make_row(const body &A, const body &B, v3 dir, v3 ra, v3 rb)
{
    jrow r;
    r.Jv = dir;
    r.Jwa = cross(ra, dir);
    r.Jwb = -cross(rb, dir);
    v3 ka = mul(A.Iw_inv, r.Jwa);
    v3 kb = mul(B.Iw_inv, r.Jwb);
    r.Meff = 1.0 / (((A.inv_mass + B.inv_mass) + dot(r.Jwa, ka)) + dot(r.Jwb, kb));
    return r;
}

// constraint.h:874-953
bool build_contact(const body &A, const body &B, const contact &k, double dt, double thr, cpoint &out)
{
    const double bias_factor = 0.1, slop = 0.005;
    v3 nw = normalized(k.normal);
    v3 lnb = rotate(conj(B.q), nw);
    v3 ra = rotate(A.q, k.local_a);
    v3 rb = rotate(B.q, k.local_b);
    v3 n = rotate(B.q, lnb);
    double pen = dot((B.pos + rb) - (A.pos + ra), n);
    if (pen <= 0.0) return false;
    out.n = make_row(A, B, n, ra, rb);
    out.n.lo = 0.0;
    v3 vca = A.vel + cross(A.angvel, ra);
    v3 vcb = B.vel + cross(B.angvel, rb);
    double vrel = dot(vca - vcb, n);
    double e = smax(A.restitution, B.restitution);
    double rbias = 0.0;
    if (vrel < -thr) rbias = e * vrel;
    double baum = (-bias_factor / dt) * smax(0.0, pen - slop);
    out.n.bias = smin(baum, rbias);
    v3 t1, t2;
    basis(n, t1, t2);
    out.t1 = make_row(A, B, t1, ra, rb);
    out.t2 = make_row(A, B, t2, ra, rb);
    return true;
}

void friction_block(cpoint &pt) // constraint.h:1131-1187
{
    body &A = *pt.A;
    body &B = *pt.B;
    double maxf = pt.mu * pt.n.acc;
    jrow &t1 = pt.t1;
    jrow &t2 = pt.t2;
    double jv1 = row_jv(t1, A, B);
    double jv2 = row_jv(t2, A, B);
    double b1 = jv1 + t1.bias;
    double b2 = jv2 + t2.bias;
    double l1 = -(pt.i11 * b1 + pt.i12 * b2);
    double l2 = -(pt.i12 * b1 + pt.i22 * b2);
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
    if (!A.is_static)
    {
        A.vel = A.vel + (t1.Jv * a1 + t2.Jv * a2) * A.inv_mass;
        A.angvel = A.angvel + mul(A.Iw_inv, t1.Jwa * a1 + t2.Jwa * a2);
    }
    if (!B.is_static)
    {
        B.vel = B.vel + ((-t1.Jv) * a1 - t2.Jv * a2) * B.inv_mass;
        B.angvel = B.angvel + mul(B.Iw_inv, t1.Jwb * a1 + t2.Jwb * a2);
    }
}
// effective mass denominators of the joint rows, each in the reference's term order
inline v3 ka_of(const body &A, const jrow &r) { return mul(A.Iw_inv, r.Jwa); }
inline v3 kb_of(const body &B, const jrow &r) { return mul(B.Iw_inv, r.Jwb); }

jrow joint_linear_row(v3 n, v3 ra, v3 rb) // J = [n, ra x n, -n, -(rb x n)]
{
    jrow r;
    r.Jv = n;
    r.Jwa = cross(ra, n);
    r.Jwb = -cross(rb, n);
    return r;
}

jrow joint_angular_row(v3 n) // J = [0, n, 0, -n] (angular_scale 1 m)
{
    jrow r;
    v3 jt = n * 1.0;
    r.Jv = v3{0.0, 0.0, 0.0};
    r.Jwa = jt;
    r.Jwb = -jt;
    return r;
}

double angular_meff(const body &A, const body &B, const jrow &r)
{ return 1.0 / (dot(r.Jwa, ka_of(A, r)) + dot(r.Jwb, kb_of(B, r))); }

// orientation_error of slider / weld (constraint.h:689-697, 812-820), rotated to world
v3 orientation_error_world(const body &A, const body &B, qt qrel)
{
    qt cur = qmul(conj(A.q), B.q);
    qt oe = qmul(qrel, conj(cur));
    if (oe.w < 0) oe = qt{-oe.w, -oe.x, -oe.y, -oe.z};
    return rotate(A.q, v3{oe.x, oe.y, oe.z} * 2.0);
}

// build_jacobian_impl of every joint type; returns the number of rows written to out
int build_joint_rows(const joint &j, const body &A, const body &B, double dt, jrow *out)
{
    const v3 axes[3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    v3 ra = rotate(A.q, j.la);
    v3 rb = rotate(B.q, j.lb);
    switch (j.type)
    {
    case j_distance: // constraint.h:186-228
    case j_spring:   // constraint.h:291-335
    {
        v3 pa = A.pos + ra, pb = B.pos + rb;
        v3 delta = pa - pb;
        double cur = norm(delta);
        if (cur < 1e-6) return 0;
        v3 u = delta / cur;
        jrow r;
        r.Jv = u;
        r.Jwa = cross(ra, u);
        r.Jwb = -cross(rb, u);
        v3 ka = ka_of(A, r), kb = kb_of(B, r);
        double mc = ((A.inv_mass + B.inv_mass) + dot(r.Jwa, ka)) + dot(r.Jwb, kb);
        if (j.type == j_distance)
        {
            r.Meff = 1.0 / mc;
            r.bias = (0.1 / dt) * (cur - j.distance);
        }
        else
        {
            double gamma = 1.0 / (dt * (j.damping + dt * j.stiffness));
            double beta = (dt * j.stiffness) / (j.damping + dt * j.stiffness);
            r.Meff = 1.0 / (mc + gamma);
            r.bias = (beta / dt) * (cur - j.distance);
            r.gamma = gamma;
        }
        r.acc = j.acc[0];
        out[0] = r;
        return 1;
    }
    case j_ball_socket: // constraint.h:386-425
    case j_hinge:       // constraint.h:509-573
    case j_weld:        // constraint.h:763-834
    {
        const double bf = j.type == j_hinge ? 0.2 : 0.1;
        v3 delta = (A.pos + ra) - (B.pos + rb);
        int k = 0;
        for (const v3 &n : axes)
        {
            jrow r = joint_linear_row(n, ra, rb);
            v3 ka = ka_of(A, r), kb = kb_of(B, r);
            r.Meff = 1.0 / (((A.inv_mass + B.inv_mass) + dot(r.Jwa, ka)) + dot(r.Jwb, kb));
            r.bias = (bf / dt) * dot(delta, n);
            r.acc = j.acc[static_cast<std::size_t>(k)];
            out[k++] = r;
        }
        if (j.type == j_hinge)
        {
            v3 ua = rotate(A.q, j.ax_a), ub = rotate(B.q, j.ax_b);
            v3 aerr = -cross(ua, ub);
            const v3 dirs[2] = {rotate(A.q, j.f1), rotate(A.q, j.f2)};
            for (const v3 &n : dirs)
            {
                jrow r = joint_angular_row(n);
                r.Meff = angular_meff(A, B, r);
                r.bias = (bf / dt) * (dot(aerr, n) * 1.0);
                r.acc = j.acc[static_cast<std::size_t>(k)];
                out[k++] = r;
            }
        }
        else if (j.type == j_weld)
        {
            v3 aerr = orientation_error_world(A, B, j.qrel);
            for (const v3 &n : axes)
            {
                jrow r = joint_angular_row(n);
                r.Meff = angular_meff(A, B, r);
                r.bias = ((bf / dt) * dot(aerr, n)) * 1.0;
                r.acc = j.acc[static_cast<std::size_t>(k)];
                out[k++] = r;
            }
        }
        return k;
    }
    case j_slider: // constraint.h:630-713
    {
        v3 delta = (A.pos + ra) - (B.pos + rb);
        v3 axis_a = rotate(A.q, j.ax_a);
        v3 t1 = rotate(A.q, j.f1), t2 = rotate(A.q, j.f2);
        v3 lever = ra - delta;
        int k = 0;
        for (const v3 &n : {t1, t2})
        {
            jrow r;
            r.Jv = n;
            r.Jwa = cross(lever, n);
            r.Jwb = -cross(rb, n);
            v3 ka = ka_of(A, r), kb = kb_of(B, r);
            double lin = A.inv_mass + B.inv_mass;
            double ang = dot(r.Jwa, ka) + dot(r.Jwb, kb);
            double reg = 0.01 * ang;
            r.gamma = reg;
            r.Meff = 1.0 / ((lin + ang) + reg);
            double e = dot(delta, n);
            e = (e < -0.2) ? -0.2 : ((0.2 < e) ? 0.2 : e); // std::clamp
            r.bias = (0.1 / dt) * e;
            r.acc = j.acc[static_cast<std::size_t>(k)];
            out[k++] = r;
        }
        v3 aerr = orientation_error_world(A, B, j.qrel);
        for (const v3 &n : {axis_a, t1, t2})
        {
            jrow r = joint_angular_row(n);
            r.Meff = angular_meff(A, B, r);
            r.bias = (0.1 / dt) * (dot(aerr, n) * 1.0);
            r.acc = j.acc[static_cast<std::size_t>(k)];
            out[k++] = r;
        }
        return k;
    }
    default:
        return 0;
    }
}
} // namespace

int joint_dimension(int type)
{
    switch (type)
    {
    case j_distance:
    case j_spring:
        return 1;
    case j_ball_socket:
        return 3;
    case j_hinge:
    case j_slider:
        return 5;
    case j_weld:
        return 6;
    default:
        return 0;
    }
}

bool world::add_joint(const joint_desc &d) // the constructors of constraint.h:180-184, 287-292, 377-382, 502-507, 612-628, 757-762
{
    if (d.type < 0 || d.type > j_spring || d.slot >= (1u << 24) || d.a >= bodies.size() || d.b >= bodies.size())
        return false;
    if (d.a == d.b || bodies[d.a].disabled || bodies[d.b].disabled) return false;
    joint j;
    j.type = d.type;
    j.slot = d.slot;
    j.a = d.a;
    j.b = d.b;
    auto it = std::lower_bound(joints.begin(), joints.end(), j.key(),
                               [](const joint &x, uint32_t k) { return x.key() < k; });
    if (it != joints.end() && it->key() == j.key()) return false;
    const body &A = bodies[d.a];
    const body &B = bodies[d.b];
    j.distance = d.distance;
    j.stiffness = d.stiffness;
    j.damping = d.damping;
    switch (d.type)
    {
    case j_distance:
    case j_spring:
        j.la = d.local_a;
        j.lb = d.local_b;
        break;
    case j_hinge:
        j.la = d.local_a;
        j.lb = d.local_b;
        j.ax_a = normalized(d.local_axis_a);
        j.ax_b = normalized(d.local_axis_b);
        basis(j.ax_a, j.f1, j.f2);
        break;
    case j_ball_socket:
    case j_weld:
    case j_slider:
        // particle::project_to_local (particle.h:107-108)
        j.la = rotate(conj(A.q), d.anchor - A.pos);
        j.lb = rotate(conj(B.q), d.anchor - B.pos);
        if (d.type != j_ball_socket) j.qrel = qmul(conj(A.q), B.q);
        if (d.type == j_slider)
        {
            v3 axn = normalized(d.axis);
            j.ax_a = rotate(conj(A.q), axn);
            v3 w1, w2;
            basis(axn, w1, w2);
            j.f1 = rotate(conj(A.q), w1);
            j.f2 = rotate(conj(A.q), w2);
        }
        break;
    default:
        break;
    }
    joints.insert(it, j);
    return true;
}

std::optional<double> ray_box_distance(v3 o, v3 inv, const aabb &box, double max_dist) // bvh.h:58-91
{
    auto slab = [](double lo, double hi, double &tl, double &th)
    {
        if (std::isnan(lo)) lo = -INFINITY;
        if (std::isnan(hi)) hi = INFINITY;
        tl = std::fmin(lo, hi);
        th = std::fmax(lo, hi);
    };
    double tminx, tmaxx, tminy, tmaxy, tminz, tmaxz;
    slab((box.min.x - o.x) * inv.x, (box.max.x - o.x) * inv.x, tminx, tmaxx);
    slab((box.min.y - o.y) * inv.y, (box.max.y - o.y) * inv.y, tminy, tmaxy);
    slab((box.min.z - o.z) * inv.z, (box.max.z - o.z) * inv.z, tminz, tmaxz);
    // std::max / std::min over {x, y, z}: the leftmost of equal values
    double tmin = tminx;
    if (tmin < tminy) tmin = tminy;
    if (tmin < tminz) tmin = tminz;
    double tmax = tmaxx;
    if (tmaxy < tmax) tmax = tmaxy;
    if (tmaxz < tmax) tmax = tmaxz;
    if (tmax >= 0.0 && tmin <= tmax && tmin <= max_dist) return (0.0 < tmin) ? tmin : 0.0; // std::max(0.0, tmin)
    return std::nullopt;
}

std::vector<std::pair<uint32_t, double>> world::raycast(v3 origin, v3 dir, double max_dist) const
{
    v3 d = normalized(dir); // ray::ray, bvh.h:46-49
    v3 inv{1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
    std::vector<std::pair<uint32_t, double>> out;
    for (std::size_t i = 0; i < bodies.size(); ++i)
    {
        if (bodies[i].disabled) continue;
        if (auto t = ray_box_distance(origin, inv, bodies[i].fat, max_dist))
            out.emplace_back(static_cast<uint32_t>(i), *t);
    }
    std::sort(out.begin(), out.end(), [](const auto &x, const auto &y)
              { return x.second != y.second ? x.second < y.second : x.first < y.first; });
    return out;
}

std::vector<uint32_t> world::query_aabb(const aabb &box) const
{
    std::vector<uint32_t> out;
    for (std::size_t i = 0; i < bodies.size(); ++i)
        if (!bodies[i].disabled && bodies[i].fat.intersects(box)) out.push_back(static_cast<uint32_t>(i));
    return out;
}

bool world::remove_joint(uint32_t key)
{
    auto it = std::lower_bound(joints.begin(), joints.end(), key,
                               [](const joint &x, uint32_t k) { return x.key() < k; });
    if (it == joints.end() || it->key() != key) return false;
    joints.erase(it);
    return true;
}

void world::stage_solve(double dt) // constraint.h:1052-1201
{
    const double thr = 2.0 * gnorm * dt;
    // rows grouped by manifold, manifolds in sorted order
    std::vector<cpoint> pts;
    std::vector<std::size_t> mstart, midx; // per contributing manifold: first row, manifold idx
    for (std::size_t mi = 0; mi < manifolds.size(); ++mi)
    {
        manifold &man = manifolds[mi];
        if (man.n == 0) continue;
        body &A = bodies[man.a];
        body &B = bodies[man.b];
        std::size_t first = pts.size();
        for (int ci = 0; ci < man.n; ++ci)
        {
            contact &k = man.c[ci];
            cpoint p;
            if (!build_contact(A, B, k, dt, thr, p)) continue;
            p.mu = std::sqrt(A.friction * B.friction);
            double m11 = 1.0 / p.t1.Meff, m22 = 1.0 / p.t2.Meff;
            double m12 = dot(p.t1.Jwa, mul(A.Iw_inv, p.t2.Jwa)) + dot(p.t1.Jwb, mul(B.Iw_inv, p.t2.Jwb));
            double det = m11 * m22 - m12 * m12;
            if (det > 0.0)
            {
                p.i11 = m22 / det;
                p.i22 = m11 / det;
                p.i12 = -m12 / det;
            }
            else
            {
                p.i11 = p.t1.Meff;
                p.i22 = p.t2.Meff;
                p.i12 = 0.0;
            }
            p.n.acc = k.ln;
            p.t1.acc = k.lt1;
            p.t2.acc = k.lt2;
            p.cache = &k;
            p.A = &A;
            p.B = &B;
            pts.push_back(p);
        }
        if (pts.size() > first)
        {
            mstart.push_back(first);
            midx.push_back(mi);
        }
    }
    mstart.push_back(pts.size());
    const std::size_t M = midx.size();
    counters.rows = pts.size();

    // processing order over contributing manifolds
    std::vector<std::size_t> order;
    colours.assign(M, 0);
    colour_keys.resize(M);
    for (std::size_t i = 0; i < M; ++i) colour_keys[i] = active[midx[i]];
    if (solver_mode == 0)
    {
        for (std::size_t i = 0; i < M; ++i) order.push_back(i);
        counters.colours = M ? 1 : 0;
    }
    else
    {
        const uint64_t none = ~0ull;
        std::vector<uint64_t> claim(bodies.size(), none);
        std::vector<uint64_t> used(bodies.size(), 0); // colours taken at each dynamic body (bits 0-63)
        std::vector<int> col(M, -1);
        std::size_t left = M;
        uint32_t round = 0, ncol = 0;
        while (left > 0)
        {
            std::fill(claim.begin(), claim.end(), none);
            for (std::size_t i = 0; i < M; ++i)
            {
                if (col[i] >= 0) continue;
                const manifold &man = manifolds[midx[i]];
                uint64_t pr = (static_cast<uint64_t>(colour_mix32(static_cast<uint32_t>(midx[i]))) << 32) | midx[i];
                if (!bodies[man.a].is_static) claim[man.a] = std::min(claim[man.a], pr);
                if (!bodies[man.b].is_static) claim[man.b] = std::min(claim[man.b], pr);
            }
            for (std::size_t i = 0; i < M; ++i)
            {
                if (col[i] >= 0) continue;
                const manifold &man = manifolds[midx[i]];
                uint64_t pr = (static_cast<uint64_t>(colour_mix32(static_cast<uint32_t>(midx[i]))) << 32) | midx[i];
                const bool sa = bodies[man.a].is_static, sb = bodies[man.b].is_static;
                bool ok = (sa || claim[man.a] == pr) && (sb || claim[man.b] == pr);
                if (ok)
                {
                    // the smallest colour no coloured manifold at either dynamic body holds
                    uint64_t u = (sa ? 0 : used[man.a]) | (sb ? 0 : used[man.b]);
                    uint32_t k = ~u ? static_cast<uint32_t>(__builtin_ctzll(~u)) : 64u + round;
                    col[i] = static_cast<int>(k);
                    if (k < 64)
                    {
                        if (!sa) used[man.a] |= 1ull << k;
                        if (!sb) used[man.b] |= 1ull << k;
                    }
                    ncol = std::max(ncol, k + 1);
                    --left;
                }
            }
            ++round;
        }
        const uint32_t c = ncol;
        counters.colours = c;
        for (uint32_t cc = 0; cc < c; ++cc)
            for (std::size_t i = 0; i < M; ++i)
                if (col[i] == static_cast<int>(cc)) order.push_back(i);
        for (std::size_t i = 0; i < M; ++i) colours[i] = static_cast<uint32_t>(col[i]);
    }

    // joint rows (constraint.h:1110-1121), all joints in key order; coloured like the
    // manifolds, over the joints' dynamic bodies, with priority from the joint index
    const std::size_t J = joints.size();
    std::vector<jrow> jr(6 * J);
    for (std::size_t k = 0; k < J; ++k)
        joints[k].rows = build_joint_rows(joints[k], bodies[joints[k].a], bodies[joints[k].b], dt, &jr[6 * k]);
    std::vector<std::size_t> jorder;
    joint_colours.clear();
    if (solver_mode == 0)
    {
        for (std::size_t k = 0; k < J; ++k) jorder.push_back(k);
    }
    else if (J)
    {
        const uint64_t none = ~0ull;
        std::vector<uint64_t> claim(bodies.size(), none);
        std::vector<uint64_t> used(bodies.size(), 0);
        std::vector<int> col(J, -1);
        std::size_t left = J;
        uint32_t round = 0, c = 0;
        while (left > 0)
        {
            std::fill(claim.begin(), claim.end(), none);
            for (std::size_t k = 0; k < J; ++k)
            {
                if (col[k] >= 0) continue;
                const joint &jj = joints[k];
                uint64_t pr = (static_cast<uint64_t>(colour_mix32(static_cast<uint32_t>(k))) << 32) | k;
                if (!bodies[jj.a].is_static) claim[jj.a] = std::min(claim[jj.a], pr);
                if (!bodies[jj.b].is_static) claim[jj.b] = std::min(claim[jj.b], pr);
            }
            for (std::size_t k = 0; k < J; ++k)
            {
                if (col[k] >= 0) continue;
                const joint &jj = joints[k];
                uint64_t pr = (static_cast<uint64_t>(colour_mix32(static_cast<uint32_t>(k))) << 32) | k;
                const bool sa = bodies[jj.a].is_static, sb = bodies[jj.b].is_static;
                if ((sa || claim[jj.a] == pr) && (sb || claim[jj.b] == pr))
                {
                    uint64_t u = (sa ? 0 : used[jj.a]) | (sb ? 0 : used[jj.b]);
                    uint32_t kc = ~u ? static_cast<uint32_t>(__builtin_ctzll(~u)) : 64u + round;
                    col[k] = static_cast<int>(kc);
                    if (kc < 64)
                    {
                        if (!sa) used[jj.a] |= 1ull << kc;
                        if (!sb) used[jj.b] |= 1ull << kc;
                    }
                    c = std::max(c, kc + 1);
                    --left;
                }
            }
            ++round;
        }
        for (uint32_t cc = 0; cc < c; ++cc)
            for (std::size_t k = 0; k < J; ++k)
                if (col[k] == static_cast<int>(cc)) jorder.push_back(k);
        for (std::size_t k = 0; k < J; ++k) joint_colours.push_back(static_cast<uint32_t>(col[k]));
    }

    for (std::size_t k : jorder) // warm start, joints first (constraint.h:1124-1125)
        for (int r = 0; r < joints[k].rows; ++r)
            row_warm(jr[6 * k + static_cast<std::size_t>(r)], bodies[joints[k].a], bodies[joints[k].b]);
    for (std::size_t oi : order) // warm start
        for (std::size_t r = mstart[oi]; r < mstart[oi + 1]; ++r)
        {
            cpoint &p = pts[r];
            row_warm(p.n, *p.A, *p.B);
            row_warm(p.t1, *p.A, *p.B);
            row_warm(p.t2, *p.A, *p.B);
        }
    for (int it = 0; it < iterations; ++it)
    {
        for (std::size_t k : jorder) // joint rows first (constraint.h:1134)
            for (int r = 0; r < joints[k].rows; ++r)
                row_apply(jr[6 * k + static_cast<std::size_t>(r)], bodies[joints[k].a], bodies[joints[k].b]);
        for (std::size_t oi : order)
            for (std::size_t r = mstart[oi]; r < mstart[oi + 1]; ++r)
            {
                cpoint &p = pts[r];
                row_apply(p.n, *p.A, *p.B);
                friction_block(p);
            }
    }
    for (std::size_t k = 0; k < J; ++k) // constraint.h:1196-1198
        for (int r = 0; r < joints[k].rows; ++r)
            joints[k].acc[static_cast<std::size_t>(r)] = jr[6 * k + static_cast<std::size_t>(r)].acc;
    for (auto &p : pts)
    {
        p.cache->ln = p.n.acc;
        p.cache->lt1 = p.t1.acc;
        p.cache->lt2 = p.t2.acc;
    }
}

void world::stage_integrate_pos(double dt) // world.cpp:50-55, integrate.h:42-46
{
    for (auto &b : bodies)
    {
        if (b.is_static) continue;
        b.pos = b.pos + b.vel * dt;
        b.q = qmul(exp_map(b.angvel, dt), b.q);
        b.update_derived();
    }
}

void world::step(double dt) // src/world.cpp:20-56
{
    if (run_integrate)
        stage_integrate_vel(dt);
    else
    {
        counters.moved = 0;
        for (auto &b : bodies) b.moved = false;
    }
    stage_pairs();
    stage_narrow();
    if (run_solver)
        stage_solve(dt);
    else
    {
        counters.rows = 0;
        counters.colours = 0;
    }
    if (run_integrate) stage_integrate_pos(dt);
}

} // namespace pko
