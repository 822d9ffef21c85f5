// Oracle (test infrastructure only): FP64 CPU restatement of PhysKit's world step
// (src/world.cpp:20-56 and its callees). Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it, and only as the checker.
//
// Deliberate, documented determinism choices (the reference's order depends on
// abseil hash-set iteration, collision_phases.h:383-403, and is not reproducible):
//   * the active pair list and the manifold list are kept sorted by pair key;
//   * the contact solver runs either sequentially in that order or in the colour
//     order produced by colour_manifolds() (the same integer algorithm the device
//     runs), selected by world::solver_mode.
#pragma once
#include "pko_math.hpp"

#include <array>
#include <cstddef>
#include <cstdint>
#include <optional>
#include <vector>

namespace pko
{

// ---------------- bounds (include/physkit/collision/bounds.h:27-205) ----------------
struct aabb
{
    v3 min, max;
    v3 point(unsigned i) const
    {
        return {(i & 1u) ? max.x : min.x, (i & 2u) ? max.y : min.y, (i & 4u) ? max.z : min.z};
    }
    bool contains(const aabb &o) const // bounds.h:80-85
    {
        return (o.min.x >= min.x && o.max.x <= max.x) && (o.min.y >= min.y && o.max.y <= max.y) &&
               (o.min.z >= min.z && o.max.z <= max.z);
    }
    bool intersects(const aabb &o) const // bounds.h:87-92
    {
        return (min.x <= o.max.x && max.x >= o.min.x) && (min.y <= o.max.y && max.y >= o.min.y) &&
               (min.z <= o.max.z && max.z >= o.min.z);
    }
    // analytic support, bounds.h:188-199
    v3 support(v3 d) const
    {
        double sx = d.x >= 0 ? 1.0 : -1.0, sy = d.y >= 0 ? 1.0 : -1.0, sz = d.z >= 0 ? 1.0 : -1.0;
        return {sx > 0 ? max.x : min.x, sy > 0 ? max.y : min.y, sz > 0 ? max.z : min.z};
    }
};
aabb aabb_from_points(const v3 *p, std::size_t n);     // bounds.h:33-49
aabb aabb_rotate(const aabb &b, qt q);                  // bounds.h:142-158
inline aabb aabb_translate(const aabb &b, v3 o) { return {b.min + o, b.max + o}; }

struct obb_shape // bounds.h:539-548
{
    v3 center;
    qt q;
    v3 he;
    v3 support(v3 d) const
    {
        v3 ld = rotate(conj(q), d);
        v3 s{(ld.x >= 0 ? 1.0 : -1.0) * he.x, (ld.y >= 0 ? 1.0 : -1.0) * he.y,
             (ld.z >= 0 ? 1.0 : -1.0) * he.z};
        return center + rotate(q, s);
    }
};

// ---------------- mesh (src/mesh.cpp, include/physkit/collision/mesh.h) ----------------
struct mesh
{
    std::vector<v3> verts;
    aabb local; // aabb::from_points(vertices), mesh.h:191
    // first vertex with strictly greater dot wins (mesh.cpp:341-358)
    v3 support(v3 d) const
    {
        std::size_t best = 0;
        double bd = dot(verts[0], d);
        for (std::size_t i = 1; i < verts.size(); ++i)
        {
            double t = dot(verts[i], d);
            if (t > bd)
            {
                bd = t;
                best = i;
            }
        }
        return verts[best];
    }
};
mesh make_mesh(const v3 *verts, std::size_t n);
mesh mesh_box(v3 half);                                       // mesh.cpp:25-58
mesh mesh_sphere(double r, unsigned stacks, unsigned sectors); // mesh.cpp:60-115
mesh mesh_pyramid(double base_half, double height);            // mesh.cpp:117-137

struct mesh_instance // mesh.cpp:442-448, 398
{
    const mesh *m;
    v3 pos;
    qt q;
    v3 support(v3 d) const
    {
        v3 ld = rotate(conj(q), d);
        return rotate(q, m->support(ld)) + pos;
    }
    aabb bounds() const { return aabb_translate(aabb_rotate(m->local, q), pos); }
};

// ---------------- GJK / EPA (src/collision.cpp) ----------------
struct support_pt
{
    v3 p, pa, pb;
};
struct collision_info // collision.h:52-59
{
    v3 normal, world_a, world_b;
    double depth = 0;
};
struct gjk_stats
{
    int gjk_iters = 0, epa_iters = 0, polytope = 0, faces = 0;
    bool hit = false;
};

struct shape_ref // tagged union used by the standalone gjk_epa entry
{
    int kind; // 0 aabb, 1 obb, 2 mesh instance
    aabb box;
    obb_shape obb;
    mesh_instance inst;
    v3 support(v3 d) const
    {
        if (kind == 0) return box.support(d);
        if (kind == 1) return obb.support(d);
        return inst.support(d);
    }
};
std::optional<collision_info> gjk_epa(const shape_ref &a, const shape_ref &b,
                                      gjk_stats *st = nullptr);
std::optional<collision_info> gjk_epa(const mesh_instance &a, const mesh_instance &b,
                                      gjk_stats *st = nullptr);

// ---------------- bodies and world ----------------
struct body // particle.h:13-147 + object.h:111-144
{
    bool is_static = false;
    bool disabled = false; // a removed body's arena slot: static, never paired
    uint32_t shape = 0;
    v3 pos, vel, acc, angvel, torque;
    qt q;
    double mass = 1, inv_mass = 1;
    m3 I_local, Iinv_local, R, Iw, Iw_inv;
    double restitution = 0.5, friction = 0.5;
    aabb fat;        // dynamic tree leaf bounds (bvh.h:294-302 stores exact bounds at add)
    bool moved = false;
    void update_derived(); // particle.h:140-146
};

struct contact // collision_phases.h:75-99
{
    v3 normal, local_a, local_b;
    double depth = 0;
    double ln = 0, lt1 = 0, lt2 = 0;
};
struct manifold
{
    uint32_t a = 0, b = 0;
    int n = 0;
    std::array<contact, 4> c{};
    void add_contact(const contact &k); // collision_phases.h:124-130
    void add_reduce(const contact &k);  // collision_phases.h:139-198
};

struct event
{
    uint64_t key;
    int kind; // 1 begin, 2 end
};

// ---------------- joints (include/physkit/collision/constraint.h:139-834) ----------------
// Types numbered as constraint_type (constraint.h:14-23). A joint's key is
// (type << 24) | slot, slot being its index in the per-type arena (arena.h:64-80); the
// reference builds joint rows arena by arena in type order, slot order within one
// (constraint.h:1110-1121), which is ascending key order.
enum joint_type : int
{
    j_distance = 0,
    j_ball_socket = 1,
    j_hinge = 2,
    j_slider = 3,
    j_weld = 4,
    j_spring = 5,
};
struct joint_desc // the desc structs of constraint.h:144-175, 239-285, 357-374, 446-497, 585-609, 737-755
{
    int type = 0;
    uint32_t slot = 0, a = 0, b = 0;
    v3 local_a, local_b;          // distance, spring, hinge
    v3 local_axis_a, local_axis_b; // hinge
    v3 anchor;                    // ball_socket, slider, weld (world frame)
    v3 axis;                      // slider (world frame)
    double distance = 0, stiffness = 0, damping = 0;
};
struct joint // state the constructors keep (constraint.h:180-184, 287-292, 377-382, 502-507, 612-628, 757-762)
{
    int type = 0;
    uint32_t slot = 0, a = 0, b = 0;
    v3 la, lb;     // body-frame anchors
    v3 ax_a, ax_b; // hinge: normalized body-frame axes; slider: ax_a = body-frame axis of a
    v3 f1, f2;     // hinge: basis of ax_a; slider: body-frame t1, t2 of a
    qt qrel{1, 0, 0, 0}; // slider, weld: conj(q_a) q_b at construction
    double distance = 0, stiffness = 0, damping = 0;
    std::array<double, 6> acc{}; // accumulated impulses, kept across steps for warm starts
    int rows = 0;                // rows built in the last solve (0: none)
    uint32_t key() const { return (static_cast<uint32_t>(type) << 24) | slot; }
};
int joint_dimension(int type); // jacobian_dimension of each type

struct step_counters
{
    uint64_t moved = 0, pairs = 0, gjk_hits = 0, contacts = 0, rows = 0, colours = 0;
};

inline uint64_t pair_key(uint32_t a, uint32_t b) // collision_phases.h:63-69
{
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return (static_cast<uint64_t>(lo) << 32) | hi;
}

// deterministic manifold colouring shared with the device (Jones-Plassmann order with
// the smallest free colour): each round, every uncoloured manifold whose priority is
// the minimum, over all uncoloured manifolds, at each of its dynamic bodies takes the
// smallest colour that no coloured manifold at those bodies holds (64 + the round
// number when colours 0-63 are all taken there). priority = (mix32(p) << 32) | p, with
// p the manifold's position in the sorted pair list.
uint32_t colour_mix32(uint32_t x);

struct world
{
    v3 gravity{0.0, -9.81, 0.0};
    double gnorm = 9.81;
    int iterations = 10;
    int solver_mode = 1; // 0: sequential in manifold order, 1: colour order
    bool all_moved = false;
    bool run_solver = true;
    bool run_integrate = true;
    bool epa_guard = false; // discard zero-normal / non-finite EPA results (not in the reference)

    std::vector<mesh> shapes;
    std::vector<body> bodies;
    std::vector<uint64_t> active;     // sorted pair keys
    std::vector<manifold> manifolds;  // parallel to `active`
    std::vector<event> events;        // begin/end events of the last step, manifold order
    std::vector<joint> joints;        // ascending key
    std::vector<uint32_t> joint_colours; // colour of each joint in the last solve (colour order)
    std::vector<uint32_t> colours;    // colour of each manifold that produced rows (last step)
    std::vector<uint64_t> colour_keys; // pair key of each of those manifolds
    step_counters counters;

    void finalize_world();            // |g| as world.h:383 computes it
    uint32_t add_body(const body &b); // world_base::create_rigid, world.h:202-208
    void step(double dt);             // world::step_impl, src/world.cpp:20-56
    // world::add_constraint (world.h:387-395): the constructor reads the bodies' current
    // poses. False if the key is taken, an argument is out of range, a body is a
    // disabled slot, or both ends are the same body.
    bool add_joint(const joint_desc &d);
    bool remove_joint(uint32_t key); // world::remove_constraint, world.h:397-398

    // World queries over the broadphase leaves (every live body's tree bounds: the fat
    // box of a dynamic body, the exact box a static body was added with). Brute force:
    // the trees only prune, so their result sets are these (bvh.h:313-344, 406-450).
    // raycast: (body, entry distance) of every leaf the ray meets within max_dist
    // (world.h:259-319), sorted by (distance, body) -- the reference's order follows its
    // trees' shapes and is not reproduced. query_aabb: bodies whose leaf box intersects
    // the box (bounds.h:87-92), ascending.
    std::vector<std::pair<uint32_t, double>> raycast(v3 origin, v3 dir, double max_dist) const;
    std::vector<uint32_t> query_aabb(const aabb &box) const;

    // stages (exposed for stage tests)
    void stage_integrate_vel(double dt);
    void stage_pairs();
    void stage_narrow();
    void stage_solve(double dt);
    void stage_integrate_pos(double dt);
};

// ray::intersect_distance (bvh.h:58-91): slab test of a ray with normalized direction,
// inverse direction 1/d (IEEE infinities for zero components); nullopt when missed
std::optional<double> ray_box_distance(v3 origin, v3 inv_dir, const aabb &box, double max_dist);

// exponential map of integrate.h:20-31
qt exp_map(v3 w, double dt);

} // namespace pko
