// Oracle (test infrastructure only): C ABI over the FP64 restatement so pytest can
// drive it through ctypes. Never linked into the product.
#include "pko.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

using namespace pko;

namespace
{
v3 ld3(const double *p) { return {p[0], p[1], p[2]}; }
void st3(double *p, v3 v)
{
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
shape_ref make_shape(int kind, const double *prm, const double *verts, int nv, mesh &storage)
{
    shape_ref s{};
    s.kind = kind;
    if (kind == 0)
        s.box = {ld3(prm), ld3(prm + 3)};
    else if (kind == 1)
        s.obb = {ld3(prm), qt{prm[3], prm[4], prm[5], prm[6]}, ld3(prm + 7)};
    else
    {
        std::vector<v3> vv(static_cast<std::size_t>(nv));
        for (int i = 0; i < nv; ++i) vv[static_cast<std::size_t>(i)] = ld3(verts + 3 * i);
        storage = make_mesh(vv.data(), vv.size());
        s.inst = {&storage, ld3(prm), qt{prm[3], prm[4], prm[5], prm[6]}};
    }
    return s;
}
} // namespace

extern "C"
{

// flags: 1 = every dynamic leaf counts as moved, 2 = skip the contact solver,
//        4 = skip velocity/position integration, 8 = discard degenerate EPA results
void *pko_world_new(const double *gravity, int iterations, int solver_mode, int flags)
{
    auto *w = new world();
    w->gravity = ld3(gravity);
    w->iterations = iterations;
    w->solver_mode = solver_mode;
    w->all_moved = (flags & 1) != 0;
    w->run_solver = (flags & 2) == 0;
    w->run_integrate = (flags & 4) == 0;
    w->epa_guard = (flags & 8) != 0;
    w->finalize_world();
    return w;
}

void pko_world_free(void *h) { delete static_cast<world *>(h); }

int pko_add_shape(void *h, const double *verts, int n)
{
    auto *w = static_cast<world *>(h);
    std::vector<v3> vv(static_cast<std::size_t>(n));
    for (int i = 0; i < n; ++i) vv[static_cast<std::size_t>(i)] = ld3(verts + 3 * i);
    w->shapes.push_back(make_mesh(vv.data(), vv.size()));
    return static_cast<int>(w->shapes.size() - 1);
}

int pko_add_bodies(void *h, int n, const uint8_t *is_static, const uint32_t *shape,
                   const double *pos, const double *vel, const double *angvel, const double *quat,
                   const double *mass, const double *inertia, const double *restitution,
                   const double *friction)
{
    auto *w = static_cast<world *>(h);
    for (int i = 0; i < n; ++i)
    {
        body b;
        b.is_static = is_static[i] != 0;
        b.disabled = is_static[i] == 2;
        b.shape = shape[i];
        if (b.shape >= w->shapes.size()) return -1;
        b.pos = ld3(pos + 3 * i);
        b.vel = ld3(vel + 3 * i);
        b.angvel = ld3(angvel + 3 * i);
        b.q = qt{quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3]};
        b.mass = b.is_static ? __builtin_inf() : mass[i]; // object.h:81-85
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) b.I_local.m[r][c] = inertia[9 * i + 3 * r + c];
        b.restitution = restitution[i];
        b.friction = friction[i];
        w->add_body(b);
    }
    return static_cast<int>(w->bodies.size());
}

double pko_step(void *h, double dt, int nsteps)
{
    auto *w = static_cast<world *>(h);
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < nsteps; ++i) w->step(dt);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

int pko_num_bodies(void *h) { return static_cast<int>(static_cast<world *>(h)->bodies.size()); }

void pko_get_bodies(void *h, double *pos, double *vel, double *angvel, double *quat, double *fat,
                    uint8_t *moved)
{
    auto *w = static_cast<world *>(h);
    for (std::size_t i = 0; i < w->bodies.size(); ++i)
    {
        const body &b = w->bodies[i];
        st3(pos + 3 * i, b.pos);
        st3(vel + 3 * i, b.vel);
        st3(angvel + 3 * i, b.angvel);
        quat[4 * i] = b.q.w;
        quat[4 * i + 1] = b.q.x;
        quat[4 * i + 2] = b.q.y;
        quat[4 * i + 3] = b.q.z;
        st3(fat + 6 * i, b.fat.min);
        st3(fat + 6 * i + 3, b.fat.max);
        moved[i] = b.moved ? 1 : 0;
    }
}

// world inertia and its inverse, 9 row-major doubles per body
void pko_get_derived(void *h, double *iw, double *iw_inv, double *inv_mass)
{
    auto *w = static_cast<world *>(h);
    for (std::size_t i = 0; i < w->bodies.size(); ++i)
    {
        const body &b = w->bodies[i];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
            {
                iw[9 * i + 3 * r + c] = b.Iw.m[r][c];
                iw_inv[9 * i + 3 * r + c] = b.Iw_inv.m[r][c];
            }
        inv_mass[i] = b.inv_mass;
    }
}

size_t pko_num_pairs(void *h) { return static_cast<world *>(h)->active.size(); }

void pko_get_pairs(void *h, uint64_t *keys)
{
    auto *w = static_cast<world *>(h);
    std::copy(w->active.begin(), w->active.end(), keys);
}

// per pair: count, then 4 contacts x 13 doubles
// (normal3, local_a3, local_b3, depth, normal impulse, tangent impulses 2)
void pko_get_manifolds(void *h, int32_t *counts, double *data)
{
    auto *w = static_cast<world *>(h);
    for (std::size_t i = 0; i < w->manifolds.size(); ++i)
    {
        const manifold &m = w->manifolds[i];
        counts[i] = m.n;
        for (int c = 0; c < 4; ++c)
        {
            double *d = data + (i * 4 + static_cast<std::size_t>(c)) * 13;
            const contact &k = m.c[static_cast<std::size_t>(c)];
            if (c >= m.n)
            {
                std::memset(d, 0, 13 * sizeof(double));
                continue;
            }
            st3(d, k.normal);
            st3(d + 3, k.local_a);
            st3(d + 6, k.local_b);
            d[9] = k.depth;
            d[10] = k.ln;
            d[11] = k.lt1;
            d[12] = k.lt2;
        }
    }
}

size_t pko_get_events(void *h, uint64_t *keys, int32_t *kinds, size_t cap)
{
    auto *w = static_cast<world *>(h);
    size_t n = std::min(cap, w->events.size());
    for (size_t i = 0; i < n; ++i)
    {
        keys[i] = w->events[i].key;
        kinds[i] = w->events[i].kind;
    }
    return w->events.size();
}

// moved, pairs, gjk_hits, contacts, rows, colours
// a caller moving bodies between steps (pkg_set_body_state): null arrays are left alone;
// a new orientation refreshes the derived state (particle::orientation, particle.h:40-44)
void pko_set_bodies(void *h, int count, const uint32_t *ids, const double *pos, const double *vel,
                    const double *angvel, const double *quat)
{
    auto *w = static_cast<world *>(h);
    for (int t = 0; t < count; ++t)
    {
        body &b = w->bodies[ids[t]];
        if (pos) b.pos = ld3(pos + 3 * t);
        if (vel) b.vel = ld3(vel + 3 * t);
        if (angvel) b.angvel = ld3(angvel + 3 * t);
        if (quat)
        {
            b.q = qt{quat[4 * t], quat[4 * t + 1], quat[4 * t + 2], quat[4 * t + 3]};
            b.update_derived();
        }
    }
}

// world queries: hits of one ray (sorted by distance, then body) / one box (ascending);
// returns the hit count, writes up to cap entries
size_t pko_raycast(void *h, const double *origin, const double *dir, double max_dist, uint32_t *ids,
                   double *dists, size_t cap)
{
    auto hits = static_cast<world *>(h)->raycast(ld3(origin), ld3(dir), max_dist);
    for (size_t k = 0; k < std::min(cap, hits.size()); ++k)
    {
        if (ids) ids[k] = hits[k].first;
        if (dists) dists[k] = hits[k].second;
    }
    return hits.size();
}

size_t pko_query_aabb(void *h, const double *box, uint32_t *ids, size_t cap)
{
    auto hits = static_cast<world *>(h)->query_aabb(aabb{ld3(box), ld3(box + 3)});
    for (size_t k = 0; k < std::min(cap, hits.size()); ++k)
        if (ids) ids[k] = hits[k];
    return hits.size();
}

// Closest hit per ray with a CPU bounding-volume hierarchy over the same leaves
// (median split of the longest centroid axis, 4 leaves per node, built once per call),
// walked front to back with the hit distance pruning the walk, as dynamic_bvh::raycast
// does with a callback that returns the hit distance (bvh.h:355-404). Same result as
// the first element of world::raycast; used to time a tree-walking CPU raycast.
// Returns the seconds spent walking (the build is not timed).
double pko_raycast_closest_bvh(void *h, size_t n_rays, const double *origins, const double *dirs,
                               const double *max_dist, uint32_t *ids, double *dists)
{
    const world &w = *static_cast<world *>(h);
    struct node
    {
        aabb box;
        uint32_t left = 0, right = 0, first = 0, count = 0; // count > 0: leaf range of `items`
    };
    std::vector<uint32_t> items;
    for (std::size_t i = 0; i < w.bodies.size(); ++i)
        if (!w.bodies[i].disabled) items.push_back(static_cast<uint32_t>(i));
    std::vector<node> nodes;
    auto centre = [&](uint32_t i, int ax)
    {
        const aabb &b = w.bodies[i].fat;
        return ax == 0 ? b.min.x + b.max.x : (ax == 1 ? b.min.y + b.max.y : b.min.z + b.max.z);
    };
    auto build = [&](auto &&self, uint32_t first, uint32_t count) -> uint32_t
    {
        uint32_t id = static_cast<uint32_t>(nodes.size());
        nodes.emplace_back();
        aabb box = w.bodies[items[first]].fat;
        double cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t k = first; k < first + count; ++k)
        {
            const aabb &b = w.bodies[items[k]].fat;
            box.min = {std::fmin(box.min.x, b.min.x), std::fmin(box.min.y, b.min.y), std::fmin(box.min.z, b.min.z)};
            box.max = {std::fmax(box.max.x, b.max.x), std::fmax(box.max.y, b.max.y), std::fmax(box.max.z, b.max.z)};
            for (int ax = 0; ax < 3; ++ax)
            {
                cmin[ax] = std::fmin(cmin[ax], centre(items[k], ax));
                cmax[ax] = std::fmax(cmax[ax], centre(items[k], ax));
            }
        }
        nodes[id].box = box;
        if (count <= 4)
        {
            nodes[id].first = first;
            nodes[id].count = count;
            return id;
        }
        int ax = 0;
        for (int a = 1; a < 3; ++a)
            if (cmax[a] - cmin[a] > cmax[ax] - cmin[ax]) ax = a;
        uint32_t mid = first + count / 2;
        std::nth_element(items.begin() + first, items.begin() + mid, items.begin() + first + count,
                         [&](uint32_t x, uint32_t y) { return centre(x, ax) < centre(y, ax); });
        uint32_t l = self(self, first, mid - first);
        uint32_t r = self(self, mid, first + count - mid);
        nodes[id].left = l;
        nodes[id].right = r;
        return id;
    };
    if (!items.empty()) build(build, 0, static_cast<uint32_t>(items.size()));
    auto t0 = std::chrono::steady_clock::now();
    std::vector<uint32_t> stack;
    for (size_t r = 0; r < n_rays; ++r)
    {
        v3 o = ld3(origins + 3 * r);
        v3 d = normalized(ld3(dirs + 3 * r));
        v3 inv{1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
        double best = max_dist[r];
        uint32_t best_id = 0xFFFFFFFFu;
        stack.clear();
        if (!nodes.empty()) stack.push_back(0);
        while (!stack.empty())
        {
            const node &nd = nodes[stack.back()];
            stack.pop_back();
            if (!ray_box_distance(o, inv, nd.box, best)) continue;
            if (nd.count)
            {
                for (uint32_t k = nd.first; k < nd.first + nd.count; ++k)
                {
                    uint32_t i = items[k];
                    auto t = ray_box_distance(o, inv, w.bodies[i].fat, best);
                    if (t && (*t < best || (*t == best && i < best_id)))
                    {
                        best = *t;
                        best_id = i;
                    }
                }
                continue;
            }
            auto dl = ray_box_distance(o, inv, nodes[nd.left].box, best);
            auto dr = ray_box_distance(o, inv, nodes[nd.right].box, best);
            if (dl && dr)
            {
                if (*dl > *dr)
                {
                    stack.push_back(nd.left);
                    stack.push_back(nd.right);
                }
                else
                {
                    stack.push_back(nd.right);
                    stack.push_back(nd.left);
                }
            }
            else if (dl)
                stack.push_back(nd.left);
            else if (dr)
                stack.push_back(nd.right);
        }
        ids[r] = best_id;
        dists[r] = best_id == 0xFFFFFFFFu ? INFINITY : best;
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ray::intersect_distance on one box; returns 0 when missed
int pko_ray_box(const double *origin, const double *dir, const double *box, double max_dist, double *t)
{
    v3 d = normalized(ld3(dir));
    auto r = ray_box_distance(ld3(origin), v3{1.0 / d.x, 1.0 / d.y, 1.0 / d.z}, aabb{ld3(box), ld3(box + 3)},
                              max_dist);
    if (!r) return 0;
    *t = *r;
    return 1;
}

// prm: local_a 3, local_b 3, local_axis_a 3, local_axis_b 3, anchor 3, axis 3,
//      distance, stiffness, damping (21 doubles)
int pko_add_joint(void *h, int type, uint32_t slot, uint32_t a, uint32_t b, const double *prm)
{
    joint_desc d;
    d.type = type;
    d.slot = slot;
    d.a = a;
    d.b = b;
    d.local_a = ld3(prm);
    d.local_b = ld3(prm + 3);
    d.local_axis_a = ld3(prm + 6);
    d.local_axis_b = ld3(prm + 9);
    d.anchor = ld3(prm + 12);
    d.axis = ld3(prm + 15);
    d.distance = prm[18];
    d.stiffness = prm[19];
    d.damping = prm[20];
    return static_cast<world *>(h)->add_joint(d) ? 1 : 0;
}

int pko_remove_joint(void *h, uint32_t key) { return static_cast<world *>(h)->remove_joint(key) ? 1 : 0; }

// live joints in key order: key, the 6 accumulated impulses, rows built by the last solve
size_t pko_get_joints(void *h, uint32_t *keys, double *acc, int32_t *rows, size_t cap)
{
    auto *w = static_cast<world *>(h);
    size_t n = std::min(cap, w->joints.size());
    for (size_t k = 0; k < n; ++k)
    {
        const joint &j = w->joints[k];
        if (keys) keys[k] = j.key();
        if (acc)
            for (int r = 0; r < 6; ++r) acc[6 * k + static_cast<size_t>(r)] = j.acc[static_cast<size_t>(r)];
        if (rows) rows[k] = j.rows;
    }
    return w->joints.size();
}

size_t pko_get_joint_colours(void *h, uint32_t *out, size_t cap)
{
    auto *w = static_cast<world *>(h);
    size_t n = std::min(cap, w->joint_colours.size());
    for (size_t k = 0; k < n; ++k) out[k] = w->joint_colours[k];
    return w->joint_colours.size();
}

void pko_get_counters(void *h, uint64_t *out)
{
    auto *w = static_cast<world *>(h);
    out[0] = w->counters.moved;
    out[1] = w->counters.pairs;
    out[2] = w->counters.gjk_hits;
    out[3] = w->counters.contacts;
    out[4] = w->counters.rows;
    out[5] = w->counters.colours;
}

// colours of the manifolds that produced solver rows in the last step
size_t pko_get_colours(void *h, uint32_t *out, uint64_t *keys, size_t cap)
{
    auto *w = static_cast<world *>(h);
    size_t n = std::min(cap, w->colours.size());
    std::copy(w->colours.begin(), w->colours.begin() + static_cast<std::ptrdiff_t>(n), out);
    if (keys) std::copy(w->colour_keys.begin(), w->colour_keys.begin() + static_cast<std::ptrdiff_t>(n), keys);
    return w->colours.size();
}

// Standalone narrowphase (collision.cpp:512-518). kind: 0 aabb (min3,max3),
// 1 obb (center3, quat wxyz, half3), 2 mesh instance (pos3, quat wxyz) over verts.
// out: normal3, world_a3, world_b3, depth. stats: gjk iters, epa iters, polytope, faces.
int pko_gjk_epa(int kind_a, const double *pa, const double *verts_a, int nva, int kind_b,
                const double *pb, const double *verts_b, int nvb, double *out, int *stats)
{
    mesh ma, mb;
    shape_ref a = make_shape(kind_a, pa, verts_a, nva, ma);
    shape_ref b = make_shape(kind_b, pb, verts_b, nvb, mb);
    gjk_stats st;
    auto r = gjk_epa(a, b, &st);
    if (stats)
    {
        stats[0] = st.gjk_iters;
        stats[1] = st.epa_iters;
        stats[2] = st.polytope;
        stats[3] = st.faces;
    }
    if (!r) return st.hit ? 2 : 0; // 2: GJK hit but EPA found a degenerate simplex
    st3(out, r->normal);
    st3(out + 3, r->world_a);
    st3(out + 6, r->world_b);
    out[9] = r->depth;
    return 1;
}

// mesh generators (mesh.cpp:25-137): writes vertices, returns the vertex count
int pko_mesh_box(const double *half, double *out)
{
    mesh m = mesh_box(ld3(half));
    for (std::size_t i = 0; i < m.verts.size(); ++i) st3(out + 3 * i, m.verts[i]);
    return static_cast<int>(m.verts.size());
}
int pko_mesh_sphere(double r, int stacks, int sectors, double *out)
{
    mesh m = mesh_sphere(r, static_cast<unsigned>(stacks), static_cast<unsigned>(sectors));
    for (std::size_t i = 0; i < m.verts.size(); ++i) st3(out + 3 * i, m.verts[i]);
    return static_cast<int>(m.verts.size());
}
int pko_mesh_pyramid(double bh, double h, double *out)
{
    mesh m = mesh_pyramid(bh, h);
    for (std::size_t i = 0; i < m.verts.size(); ++i) st3(out + 3 * i, m.verts[i]);
    return static_cast<int>(m.verts.size());
}

// world AABB of a mesh instance (mesh.cpp:398)
void pko_instance_bounds(const double *verts, int nv, const double *pos, const double *quat,
                         double *out)
{
    std::vector<v3> vv(static_cast<std::size_t>(nv));
    for (int i = 0; i < nv; ++i) vv[static_cast<std::size_t>(i)] = ld3(verts + 3 * i);
    mesh m = make_mesh(vv.data(), vv.size());
    mesh_instance inst{&m, ld3(pos), qt{quat[0], quat[1], quat[2], quat[3]}};
    aabb b = inst.bounds();
    st3(out, b.min);
    st3(out + 3, b.max);
}

void pko_instance_support(const double *verts, int nv, const double *pos, const double *quat,
                          const double *dir, double *out)
{
    std::vector<v3> vv(static_cast<std::size_t>(nv));
    for (int i = 0; i < nv; ++i) vv[static_cast<std::size_t>(i)] = ld3(verts + 3 * i);
    mesh m = make_mesh(vv.data(), vv.size());
    mesh_instance inst{&m, ld3(pos), qt{quat[0], quat[1], quat[2], quat[3]}};
    st3(out, inst.support(ld3(dir)));
}

void pko_obb_support(const double *prm, const double *dir, double *out)
{
    obb_shape o{ld3(prm), qt{prm[3], prm[4], prm[5], prm[6]}, ld3(prm + 7)};
    st3(out, o.support(ld3(dir)));
}

void pko_sincos(const double *x, double *s, double *c, int n)
{
    for (int i = 0; i < n; ++i) sincos_pk(x[i], s + i, c + i);
}

void pko_exp_map(const double *w, double dt, double *q)
{
    qt r = exp_map(ld3(w), dt);
    q[0] = r.w;
    q[1] = r.x;
    q[2] = r.y;
    q[3] = r.z;
}

void pko_inverse3(const double *m, double *out)
{
    m3 a;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a.m[r][c] = m[3 * r + c];
    m3 b = inverse(a);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[3 * r + c] = b.m[r][c];
}

// ---- stage-level hooks for the contact-path pins ----
// One stage of world::step (src/world.cpp:20-56): 0 integrate_vel and leaf update,
// 1 pair-set update, 2 narrowphase, 3 contact solve, 4 integrate_pos.
int pko_run_stage(void *h, int stage, double dt)
{
    auto *w = static_cast<world *>(h);
    switch (stage)
    {
    case 0: w->stage_integrate_vel(dt); return 1;
    case 1: w->stage_pairs(); return 1;
    case 2: w->stage_narrow(); return 1;
    case 3: w->stage_solve(dt); return 1;
    case 4: w->stage_integrate_pos(dt); return 1;
    default: return 0;
    }
}

namespace
{
contact ld_contact(const double *d) // the 13-double layout of pko_get_manifolds
{
    contact k;
    k.normal = ld3(d);
    k.local_a = ld3(d + 3);
    k.local_b = ld3(d + 6);
    k.depth = d[9];
    k.ln = d[10];
    k.lt1 = d[11];
    k.lt2 = d[12];
    return k;
}
void st_contact(double *d, const contact &k)
{
    st3(d, k.normal);
    st3(d + 3, k.local_a);
    st3(d + 6, k.local_b);
    d[9] = k.depth;
    d[10] = k.ln;
    d[11] = k.lt1;
    d[12] = k.lt2;
}
} // namespace

// Replaces the cached manifold of the active pair `key` by n contacts (13 doubles each,
// as pko_get_manifolds writes them). Returns 0 if the pair is not active or n > 4.
int pko_set_manifold(void *h, uint64_t key, int n, const double *data)
{
    auto *w = static_cast<world *>(h);
    auto it = std::lower_bound(w->active.begin(), w->active.end(), key);
    if (it == w->active.end() || *it != key || n < 0 || n > 4) return 0;
    manifold &m = w->manifolds[static_cast<std::size_t>(it - w->active.begin())];
    m.n = n;
    for (int i = 0; i < n; ++i) m.c[static_cast<std::size_t>(i)] = ld_contact(data + 13 * i);
    return 1;
}

// Leaf bounds of bodies ids[0..count) (6 doubles each: min xyz, max xyz), as a caller
// restoring a saved world state sets them.
void pko_set_fat(void *h, int count, const uint32_t *ids, const double *fat)
{
    auto *w = static_cast<world *>(h);
    for (int t = 0; t < count; ++t)
    {
        body &b = w->bodies[ids[t]];
        b.fat.min = ld3(fat + 6 * t);
        b.fat.max = ld3(fat + 6 * t + 3);
    }
}

// Replaces the active pair list (n keys, ascending) and its manifolds (counts and 4 x 13
// doubles per pair, the pko_get_manifolds layout). Returns 0 if the keys are not
// ascending or a count is out of range.
int pko_set_pairs(void *h, size_t n, const uint64_t *keys, const int32_t *counts, const double *data)
{
    auto *w = static_cast<world *>(h);
    for (size_t i = 0; i < n; ++i)
        if ((i && keys[i] <= keys[i - 1]) || counts[i] < 0 || counts[i] > 4) return 0;
    w->active.assign(keys, keys + n);
    w->manifolds.assign(n, manifold{});
    for (size_t i = 0; i < n; ++i)
    {
        manifold &m = w->manifolds[i];
        m.a = static_cast<uint32_t>(keys[i] >> 32);
        m.b = static_cast<uint32_t>(keys[i] & 0xFFFFFFFFu);
        m.n = counts[i];
        for (int c = 0; c < m.n; ++c)
            m.c[static_cast<std::size_t>(c)] = ld_contact(data + (i * 4 + static_cast<size_t>(c)) * 13);
    }
    return 1;
}

// manifold::add_contact (collision_phases.h:124-130) on a manifold holding n_old
// contacts; writes the result (up to 4 contacts) to out and returns its count.
int pko_manifold_add(int n_old, const double *old_contacts, const double *new_contact, double *out)
{
    if (n_old < 0 || n_old > 4) return -1;
    manifold m;
    m.n = n_old;
    for (int i = 0; i < n_old; ++i) m.c[static_cast<std::size_t>(i)] = ld_contact(old_contacts + 13 * i);
    m.add_contact(ld_contact(new_contact));
    for (int i = 0; i < m.n; ++i) st_contact(out + 13 * i, m.c[static_cast<std::size_t>(i)]);
    return m.n;
}

} // extern "C"
