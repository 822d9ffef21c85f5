// physkit_amd C++ facade: PhysKit's world API (include/physkit/core/world.h:23-49,
// 184-424; include/physkit/core/object.h:15-144; include/physkit/collision/mesh.h)
// over the C ABI of include/physkit_amd.h. Header-only, C++17.
//
//   auto w = physkit_amd::world(physkit_amd::world_desc::make().with_gravity({0, -9.81, 0}));
//   auto box = physkit_amd::mesh::box({0.5, 0.5, 0.5});
//   auto h = w.create_rigid(physkit_amd::object_desc::dynam().with_mesh(box).with_pos({0, 5, 0}));
//   w.step(1.0 / 60.0);                 // world_base::step -> world::step_impl on the GPU
//   auto p = w.get_rigid(h)->pos();
//
// Handles are generational arena handles (include/physkit/detail/arena.h:24-122): the
// slot index is the body id on the device, so pair keys match the reference's.
// Removing a body frees its slot for the next create_rigid (LIFO, as arena::add).
// Errors throw std::runtime_error / std::invalid_argument, as the reference does.
#pragma once
#include "../physkit_amd.h"

#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <limits>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace physkit_amd
{

struct vec3
{
    double x = 0, y = 0, z = 0;
};
struct quat // (w, x, y, z)
{
    double w = 1, x = 0, y = 0, z = 0;
    static quat identity() { return {}; }
    static quat from_angle_axis(double angle, vec3 axis) // lin_alg.h:569-576
    {
        double n = std::sqrt(axis.x * axis.x + axis.y * axis.y + axis.z * axis.z);
        double s = std::sin(0.5 * angle) / n;
        return {std::cos(0.5 * angle), axis.x * s, axis.y * s, axis.z * s};
    }
};
using mat3 = std::array<double, 9>; // row-major

inline mat3 identity3() { return {1, 0, 0, 0, 1, 0, 0, 0, 1}; }

class mesh // src/mesh.cpp:16-137
{
public:
    explicit mesh(std::vector<vec3> v) : M_vertices(std::move(v))
    {
        if (M_vertices.empty()) throw std::invalid_argument("mesh needs at least one vertex");
    }
    static std::shared_ptr<const mesh> make(std::vector<vec3> vertices)
    { return std::make_shared<const mesh>(std::move(vertices)); }
    static std::shared_ptr<const mesh> box(vec3 h)
    {
        return make({{-h.x, -h.y, -h.z}, {+h.x, -h.y, -h.z}, {+h.x, +h.y, -h.z}, {-h.x, +h.y, -h.z},
                     {-h.x, -h.y, +h.z}, {+h.x, -h.y, +h.z}, {+h.x, +h.y, +h.z}, {-h.x, +h.y, +h.z}});
    }
    static std::shared_ptr<const mesh> sphere(double r, unsigned stacks = 16, unsigned sectors = 32)
    {
        const double pi = 3.141592653589793;
        std::vector<vec3> v{{0.0, r, 0.0}};
        for (unsigned i = 1; i < stacks; ++i)
        {
            double phi = pi * static_cast<double>(i) / static_cast<double>(stacks);
            for (unsigned j = 0; j < sectors; ++j)
            {
                double th = 2.0 * pi * static_cast<double>(j) / static_cast<double>(sectors);
                v.push_back({r * std::sin(phi) * std::cos(th), r * std::cos(phi), r * std::sin(phi) * std::sin(th)});
            }
        }
        v.push_back({0.0, -r, 0.0});
        return make(std::move(v));
    }
    static std::shared_ptr<const mesh> pyramid(double b, double height)
    { return make({{-b, 0.0, -b}, {+b, 0.0, -b}, {+b, 0.0, +b}, {-b, 0.0, +b}, {0.0, height, 0.0}}); }
    const std::vector<vec3> &vertices() const { return M_vertices; }

private:
    std::vector<vec3> M_vertices;
};

class world_desc // world.h:23-49
{
public:
    static world_desc make() { return {}; }
    world_desc &with_gravity(vec3 g)
    {
        M_gravity = g;
        return *this;
    }
    world_desc &with_solver_iterations(uint32_t n)
    {
        M_iterations = n;
        return *this;
    }
    vec3 gravity() const { return M_gravity; }
    uint32_t solver_iterations() const { return M_iterations; }
    // not in the reference: see PKG_FLAG_EPA_GUARD in physkit_amd.h
    world_desc &with_epa_guard(bool on)
    {
        M_flags = on ? (M_flags | PKG_FLAG_EPA_GUARD) : (M_flags & ~PKG_FLAG_EPA_GUARD);
        return *this;
    }
    uint32_t flags() const { return M_flags; }

private:
    vec3 M_gravity{0.0, -9.81, 0.0};
    uint32_t M_iterations = 10;
    uint32_t M_flags = 0;
};

enum class body_type : uint8_t
{
    dynam,
    stat
};

class object_desc // object.h:15-109
{
public:
    explicit object_desc(body_type t) : M_type(t) {}
    static object_desc dynam() { return object_desc(body_type::dynam); }
    static object_desc stat() { return object_desc(body_type::stat); }
    object_desc &with_pos(vec3 p) { M_pos = p; return *this; }
    object_desc &with_vel(vec3 v) { M_vel = v; return *this; }
    object_desc &with_mass(double m) { M_mass = m; return *this; }
    object_desc &with_orientation(quat q) { M_q = q; return *this; }
    object_desc &with_ang_vel(vec3 w) { M_w = w; return *this; }
    object_desc &with_inertia_tensor(const mat3 &I) { M_I = I; return *this; }
    object_desc &with_mesh(std::shared_ptr<const mesh> m) { M_mesh = std::move(m); return *this; }
    object_desc &with_restitution(double e) { M_e = e; return *this; }
    object_desc &with_friction(double f) { M_mu = f; return *this; }

    body_type type() const { return M_type; }
    vec3 pos() const { return M_pos; }
    vec3 vel() const { return M_vel; }
    double mass() const { return M_type == body_type::dynam ? M_mass : INFINITY; }
    quat orientation() const { return M_q; }
    vec3 ang_vel() const { return M_w; }
    const mat3 &inertia_tensor() const { return M_I; }
    const std::shared_ptr<const mesh> &mesh_ptr() const { return M_mesh; }
    double restitution() const { return M_e; }
    double friction() const { return M_mu; }

private:
    body_type M_type;
    vec3 M_pos{}, M_vel{}, M_w{};
    double M_mass = 1.0;
    quat M_q{};
    mat3 M_I = identity3();
    std::shared_ptr<const mesh> M_mesh;
    double M_e = 0.5, M_mu = 0.5;
};

class handle // arena<T>::handle, arena.h:34-61
{
public:
    handle() = default;
    handle(uint32_t idx, uint32_t gen) : M_idx(idx), M_gen(gen) {}
    uint32_t index() const { return M_idx; }
    uint32_t generation() const { return M_gen; }
    uint64_t id() const { return (static_cast<uint64_t>(M_gen) << 32) | M_idx; }
    bool operator==(const handle &o) const { return M_idx == o.M_idx && M_gen == o.M_gen; }

private:
    uint32_t M_idx = 0xFFFFFFFFu, M_gen = 0;
};

// joint types, numbered as constraint_type (include/physkit/collision/constraint.h:14-23)
enum class joint_type : int32_t
{
    distance = PKG_JOINT_DISTANCE,
    ball_socket = PKG_JOINT_BALL_SOCKET,
    hinge = PKG_JOINT_HINGE,
    slider = PKG_JOINT_SLIDER,
    weld = PKG_JOINT_WELD,
    spring = PKG_JOINT_SPRING,
};

// The fields of the six desc structs (constraint.h:144-175 distance, 239-285 spring,
// 357-374 ball_socket, 446-497 hinge, 585-609 slider, 737-755 weld); a type reads only
// its own fields. Built as in the reference: <type>_constraint::desc::make(a, b).with_*.
class joint_desc
{
public:
    joint_desc(joint_type t, handle a, handle b) : M_type(t), M_a(a), M_b(b) {}
    joint_desc &with_local_anchor_a(const vec3 &p) { M_local_a = p; return *this; }
    joint_desc &with_local_anchor_b(const vec3 &p) { M_local_b = p; return *this; }
    joint_desc &with_local_axis_a(const vec3 &v) { M_axis_a = v; return *this; }
    joint_desc &with_local_axis_b(const vec3 &v) { M_axis_b = v; return *this; }
    joint_desc &with_anchor(const vec3 &p) { M_anchor = p; return *this; }
    joint_desc &with_axis(const vec3 &v) { M_axis = v; return *this; }
    joint_desc &with_distance(double d) { M_distance = d; return *this; }
    joint_desc &with_stiffness(double k) { M_stiffness = k; return *this; }
    joint_desc &with_damping(double c) { M_damping = c; return *this; }

    joint_type type() const { return M_type; }
    handle a() const { return M_a; }
    handle b() const { return M_b; }
    pkg_joint_desc native(uint32_t slot) const
    {
        pkg_joint_desc d{};
        d.type = static_cast<int32_t>(M_type);
        d.slot = slot;
        d.a = M_a.index();
        d.b = M_b.index();
        auto put = [](double *o, const vec3 &v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; };
        put(d.local_a, M_local_a);
        put(d.local_b, M_local_b);
        put(d.local_axis_a, M_axis_a);
        put(d.local_axis_b, M_axis_b);
        put(d.anchor, M_anchor);
        put(d.axis, M_axis);
        d.distance = M_distance;
        d.stiffness = M_stiffness;
        d.damping = M_damping;
        return d;
    }

private:
    joint_type M_type;
    handle M_a, M_b;
    vec3 M_local_a, M_local_b, M_axis_a, M_axis_b, M_anchor, M_axis;
    double M_distance = 0, M_stiffness = 0, M_damping = 0;
};

template <joint_type T> struct joint_of
{
    struct desc
    {
        static joint_desc make(handle a, handle b) { return joint_desc(T, a, b); }
    };
};
using distance_constraint = joint_of<joint_type::distance>;
using spring_constraint = joint_of<joint_type::spring>;
using ball_socket_constraint = joint_of<joint_type::ball_socket>;
using hinge_constraint = joint_of<joint_type::hinge>;
using slider_constraint = joint_of<joint_type::slider>;
using weld_constraint = joint_of<joint_type::weld>;

// handle into the joint type's arena (arena<constraint>::handle, arena.h:34-61)
struct joint_handle
{
    joint_type type;
    uint32_t index, generation;
    uint32_t key() const { return PKG_JOINT_KEY(static_cast<int32_t>(type), index); }
};

// snapshot of one body after the last step
struct object_state
{
    vec3 p, v, w;
    quat q;
    const vec3 &pos() const { return p; }
    const vec3 &vel() const { return v; }
    const vec3 &ang_vel() const { return w; }
    const quat &orientation() const { return q; }
};

class world
{
public:
    explicit world(const world_desc &desc = world_desc::make()) : M_desc(desc)
    {
        pkg_config cfg{};
        vec3 g = desc.gravity();
        cfg.gravity[0] = g.x;
        cfg.gravity[1] = g.y;
        cfg.gravity[2] = g.z;
        cfg.solver_iterations = desc.solver_iterations();
        cfg.flags = desc.flags();
        check(pkg_create(&M_ctx, &cfg), "pkg_create");
    }
    world(const world &) = delete;
    world &operator=(const world &) = delete;
    ~world()
    {
        if (M_ctx) pkg_destroy(M_ctx);
    }

    // world_base::create_rigid (world.h:202-208); the slot index is the body id
    handle create_rigid(const object_desc &d)
    {
        if (!d.mesh_ptr()) throw std::invalid_argument("Object must have a mesh"); // object.h:121
        uint32_t idx;
        if (!M_free.empty())
        {
            idx = M_free.back();
            M_free.pop_back();
            M_slots[idx].desc = d;
            M_slots[idx].live = true;
            M_slots[idx].fresh = true;
        }
        else
        {
            idx = static_cast<uint32_t>(M_slots.size());
            M_slots.push_back({d, 0, true, true});
        }
        M_dirty = true;
        return {idx, M_slots[idx].gen};
    }

    // world_base::remove_rigid (world.h:209-218): the slot becomes a disabled body until
    // reused, so ids of other bodies do not change; its pairs end with it
    bool remove_rigid(handle h)
    {
        if (h.index() >= M_slots.size()) return false;
        slot &s = M_slots[h.index()];
        if (!s.live || s.gen != h.generation()) return false;
        s.live = false;
        s.gen++;
        M_free.push_back(h.index());
        M_dirty = true;
        return true;
    }

    std::optional<object_state> get_rigid(handle h)
    {
        if (h.index() >= M_slots.size()) return std::nullopt;
        const slot &s = M_slots[h.index()];
        if (!s.live || s.gen != h.generation()) return std::nullopt;
        sync();
        return M_state[h.index()];
    }

    // world::add_constraint (world.h:387-395). As the reference's add command, the joint
    // is built at the next step, before step_impl, from the bodies' poses then.
    joint_handle add_constraint(const joint_desc &d)
    {
        for (handle h : {d.a(), d.b()})
            if (h.index() >= M_slots.size() || !M_slots[h.index()].live ||
                M_slots[h.index()].gen != h.generation())
                throw std::invalid_argument("add_constraint: stale body handle");
        const size_t t = static_cast<size_t>(d.type());
        uint32_t idx;
        if (!M_jfree[t].empty())
        {
            idx = M_jfree[t].back();
            M_jfree[t].pop_back();
        }
        else
        {
            idx = static_cast<uint32_t>(M_jgen[t].size());
            M_jgen[t].push_back(0);
            M_jlive[t].push_back(false);
        }
        M_jlive[t][idx] = true;
        M_jadd.push_back(d.native(idx));
        return {d.type(), idx, M_jgen[t][idx]};
    }

    // world::remove_constraint (world.h:397-398); false for a stale handle
    bool remove_constraint(joint_handle h)
    {
        const size_t t = static_cast<size_t>(h.type);
        if (h.index >= M_jgen[t].size() || !M_jlive[t][h.index] || M_jgen[t][h.index] != h.generation) return false;
        M_jlive[t][h.index] = false;
        M_jgen[t][h.index]++;
        M_jfree[t].push_back(h.index);
        const uint32_t key = h.key();
        for (size_t i = 0; i < M_jadd.size(); ++i)
            if (PKG_JOINT_KEY(M_jadd[i].type, M_jadd[i].slot) == key)
            {
                M_jadd.erase(M_jadd.begin() + static_cast<std::ptrdiff_t>(i));
                return true;
            }
        M_jdel.push_back(key);
        return true;
    }

    // accumulated impulses of a live joint (6 entries; rows beyond its dimension are 0)
    std::optional<std::array<double, 6>> joint_impulses(joint_handle h)
    {
        const size_t t = static_cast<size_t>(h.type);
        if (h.index >= M_jgen[t].size() || !M_jlive[t][h.index] || M_jgen[t][h.index] != h.generation)
            return std::nullopt;
        size_t n = 0;
        check(pkg_get_joints(M_ctx, nullptr, nullptr, 0, &n), "pkg_get_joints");
        std::vector<uint32_t> keys(n);
        std::vector<double> imp(6 * n);
        check(pkg_get_joints(M_ctx, keys.data(), imp.data(), n, &n), "pkg_get_joints");
        for (size_t i = 0; i < n; ++i)
            if (keys[i] == h.key())
            {
                std::array<double, 6> out{};
                for (size_t r = 0; r < 6; ++r) out[r] = imp[6 * i + r];
                return out;
            }
        return std::nullopt; // added, not yet built (no step since)
    }

    // world_base::raycast (world.h:259-319): every body whose tree bounds the ray meets
    // within max_dist, with the entry distance, sorted by (distance, body)
    std::vector<std::pair<handle, double>> raycast(const vec3 &origin, const vec3 &dir,
                                                   double max_dist = std::numeric_limits<double>::max())
    {
        if (M_dirty) upload();
        const double o[3] = {origin.x, origin.y, origin.z}, d[3] = {dir.x, dir.y, dir.z};
        uint32_t off[2] = {0, 0};
        size_t total = 0;
        check(pkg_raycast(M_ctx, 1, o, d, &max_dist, nullptr, off, nullptr, nullptr, 0, &total), "pkg_raycast");
        std::vector<uint32_t> ids(total);
        std::vector<double> dist(total);
        if (total)
            check(pkg_raycast(M_ctx, 1, o, d, &max_dist, nullptr, off, ids.data(), dist.data(), total, &total),
                  "pkg_raycast");
        std::vector<std::pair<handle, double>> out;
        for (size_t i = 0; i < total; ++i) out.emplace_back(handle{ids[i], M_slots[ids[i]].gen}, dist[i]);
        return out;
    }

    // bodies whose tree bounds intersect the box (dynamic_bvh::query_aabb, bvh.h:313-344)
    std::vector<handle> query_aabb(const vec3 &mn, const vec3 &mx)
    {
        if (M_dirty) upload();
        const double box[6] = {mn.x, mn.y, mn.z, mx.x, mx.y, mx.z};
        uint32_t off[2] = {0, 0};
        size_t total = 0;
        check(pkg_query_aabb(M_ctx, 1, box, nullptr, off, nullptr, 0, &total), "pkg_query_aabb");
        std::vector<uint32_t> ids(total);
        if (total) check(pkg_query_aabb(M_ctx, 1, box, nullptr, off, ids.data(), total, &total), "pkg_query_aabb");
        std::vector<handle> out;
        for (uint32_t i : ids) out.emplace_back(i, M_slots[i].gen);
        return out;
    }

    // begin/end callbacks: what world_base::on_collision / on_collision_exit receive
    // (world.h:327-338), in pair-key order after each step
    std::function<void(handle, handle)> on_collision;
    std::function<void(handle, handle)> on_collision_exit;

    // world_base::step(dt) -> world::step_impl (src/world.cpp:20-56)
    void step(double dt)
    {
        if (M_failed)
            throw std::runtime_error("physkit_amd: a previous step failed part-way and the world's state is lost");
        if (M_dirty) upload();
        flush_joints();
        // from here the device state is the step's, whether it completes or not
        M_state_valid = false;
        try
        {
            check(pkg_step(M_ctx, dt, 1), "pkg_step");
        }
        catch (...)
        {
            M_failed = true;
            throw;
        }
        dispatch_events();
    }

    pkg_counters counters() const
    {
        pkg_counters c{};
        pkg_get_counters(M_ctx, &c);
        return c;
    }
    pkg_ctx *native() { return M_ctx; }

private:
    struct slot
    {
        object_desc desc;
        uint32_t gen;
        bool live;
        bool fresh; // created since the last upload: starts from its object_desc
    };
    world_desc M_desc;
    pkg_ctx *M_ctx = nullptr;
    std::vector<slot> M_slots;
    std::vector<uint32_t> M_free;
    std::vector<object_state> M_state;
    bool M_dirty = true, M_state_valid = false, M_uploaded = false;
    bool M_failed = false; // a step threw: the device state is part-way through it
    size_t M_n_dev = 0; // bodies on the device (slot count at the last upload)
    // one arena per joint type: generation and liveness per slot, LIFO free slots
    std::array<std::vector<uint32_t>, 6> M_jgen, M_jfree;
    std::array<std::vector<bool>, 6> M_jlive;
    std::vector<pkg_joint_desc> M_jadd; // built at the next step
    std::vector<uint32_t> M_jdel;       // removed at the next step

    void flush_joints()
    {
        if (!M_jdel.empty())
        {
            check(pkg_remove_joints(M_ctx, M_jdel.data(), M_jdel.size()), "pkg_remove_joints");
            M_jdel.clear();
        }
        if (!M_jadd.empty())
        {
            check(pkg_add_joints(M_ctx, M_jadd.data(), M_jadd.size()), "pkg_add_joints");
            M_jadd.clear();
        }
    }

    void check(pkg_status st, const char *what) const
    {
        if (st == PKG_OK) return;
        std::string msg = std::string(what) + ": " +
                          (M_ctx ? pkg_last_error(M_ctx) : (st == PKG_NO_DEVICE ? "no HIP device" : "error"));
        if (st == PKG_INVALID_ARG) throw std::invalid_argument(msg);
        throw std::runtime_error(msg);
    }

    void upload()
    {
        if (M_uploaded && !M_state_valid) download_state(); // keep the state reached so far
        const size_t n = M_slots.size();
        std::vector<uint32_t> offs{0};
        std::vector<double> verts;
        std::vector<uint8_t> is_static(n);
        std::vector<uint32_t> shape(n);
        std::vector<double> pos(3 * n), vel(3 * n), ang(3 * n), q(4 * n), mass(n), I(9 * n), e(n), mu(n);
        for (size_t i = 0; i < n; ++i)
        {
            const slot &s = M_slots[i];
            const object_desc &d = s.desc;
            for (const vec3 &v : d.mesh_ptr()->vertices())
            {
                verts.push_back(v.x);
                verts.push_back(v.y);
                verts.push_back(v.z);
            }
            offs.push_back(static_cast<uint32_t>(verts.size() / 3));
            shape[i] = static_cast<uint32_t>(i);
            bool live = s.live;
            // a free slot stays as a disabled body (is_static = 2): it never pairs
            is_static[i] = !live ? 2 : (d.type() == body_type::stat ? 1 : 0);
            vec3 p = d.pos(), v = d.vel(), w = d.ang_vel();
            quat o = d.orientation();
            if (M_uploaded && i < M_state.size() && live && !s.fresh)
            {
                p = M_state[i].p;
                v = M_state[i].v;
                w = M_state[i].w;
                o = M_state[i].q;
            }
            double pv[3] = {p.x, p.y, p.z}, vv[3] = {v.x, v.y, v.z}, wv[3] = {w.x, w.y, w.z};
            for (int k = 0; k < 3; ++k)
            {
                pos[3 * i + k] = pv[k];
                vel[3 * i + k] = vv[k];
                ang[3 * i + k] = wv[k];
            }
            q[4 * i] = o.w;
            q[4 * i + 1] = o.x;
            q[4 * i + 2] = o.y;
            q[4 * i + 3] = o.z;
            mass[i] = d.type() == body_type::dynam ? d.mass() : 0.0;
            for (int k = 0; k < 9; ++k) I[9 * i + k] = d.inertia_tensor()[k];
            e[i] = d.restitution();
            mu[i] = d.friction();
        }
        check(pkg_upload_shapes(M_ctx, static_cast<uint32_t>(n), offs.data(), verts.data()), "pkg_upload_shapes");
        pkg_body_desc bd{is_static.data(), shape.data(), pos.data(), vel.data(), ang.data(), q.data(),
                         mass.data(), I.data(), e.data(), mu.data()};
        check(pkg_upload_bodies(M_ctx, static_cast<uint32_t>(n), &bd), "pkg_upload_bodies");
        for (slot &s : M_slots) s.fresh = false;
        M_dirty = false;
        M_uploaded = true;
        M_n_dev = n;
        M_state_valid = false;
    }

    void sync()
    {
        if (M_dirty) upload();
        if (!M_state_valid) download_state();
    }

    // poses of the uploaded bodies (the slot count of the last upload)
    void download_state()
    {
        const size_t n = M_n_dev;
        std::vector<double> pos(3 * n), vel(3 * n), ang(3 * n), q(4 * n);
        pkg_body_state out{pos.data(), vel.data(), ang.data(), q.data(), nullptr, nullptr};
        check(pkg_download_bodies(M_ctx, &out), "pkg_download_bodies");
        M_state.resize(n);
        for (size_t i = 0; i < n; ++i)
            M_state[i] = {{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]},
                          {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]},
                          {ang[3 * i], ang[3 * i + 1], ang[3 * i + 2]},
                          {q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]}};
        M_state_valid = true;
    }

    void dispatch_events()
    {
        if (!on_collision && !on_collision_exit) return;
        size_t n = 0;
        check(pkg_get_events(M_ctx, nullptr, nullptr, 0, &n), "pkg_get_events");
        if (!n) return;
        std::vector<uint64_t> keys(n);
        std::vector<int32_t> kinds(n);
        check(pkg_get_events(M_ctx, keys.data(), kinds.data(), n, &n), "pkg_get_events");
        for (size_t i = 0; i < n; ++i)
        {
            uint32_t a = static_cast<uint32_t>(keys[i] >> 32), b = static_cast<uint32_t>(keys[i] & 0xFFFFFFFFu);
            handle ha{a, M_slots[a].gen}, hb{b, M_slots[b].gen};
            if (kinds[i] == 1 && on_collision) on_collision(ha, hb);
            if (kinds[i] == 2 && on_collision_exit) on_collision_exit(ha, hb);
        }
    }
};

} // namespace physkit_amd
