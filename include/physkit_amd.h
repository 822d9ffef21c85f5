/* physkit_amd: MI355X (gfx950) implementation of PhysKit's collision-and-contact world
 * step behind a C ABI. Plain pointers and sizes only; no exceptions cross this
 * boundary. Every entry point returns a pkg_status; pkg_last_error() gives the text.
 *
 * The plug point this library replaces is PhysKit's
 *   virtual void world_base::step_impl(quantity<si::second> dt)
 * (include/physkit/core/world.h:340, implemented by world::step_impl in
 * src/world.cpp:20-56). A world hands its bodies to a pkg_ctx once, then calls
 * pkg_step() where it would call step_impl(), and reads poses, pairs, manifolds
 * and begin/end events back.
 *
 * Body ids are the reference's arena slot indices (include/physkit/detail/arena.h:
 * 64-80, used as broadphase object ids at world.h:205-207), so pair keys are the
 * reference's keys: (min_id << 32) | max_id (collision_phases.h:63-69).
 *
 * Threading: one pkg_ctx per host thread; calls on one ctx are not re-entrant. All
 * device work of a ctx is ordered on one HIP stream owned by the ctx.
 */
#ifndef PHYSKIT_AMD_H
#define PHYSKIT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pkg_status
{
    PKG_OK = 0,
    PKG_INVALID_ARG = 1,
    PKG_OOM = 2,
    PKG_OVERFLOW_PAIRS = 3,
    PKG_OVERFLOW_EPA = 4,
    PKG_DEVICE_ERROR = 5,
    PKG_NO_DEVICE = 6
} pkg_status;

typedef struct pkg_ctx pkg_ctx;

/* pkg_config.flags */
#define PKG_FLAG_ALL_MOVED 1u    /* every dynamic leaf counts as moved each step       */
#define PKG_FLAG_NO_SOLVER 2u    /* skip contact setup and solve                        */
#define PKG_FLAG_NO_INTEGRATE 4u /* skip velocity/position integration and leaf update */
/* Discard an EPA result whose normal has zero length or whose witness points or depth
 * are not finite, as if GJK had found no contact. Not in the reference: its EPA can
 * return such a result from a face spanned by three collinear support points
 * (collision.cpp:284-285, 442-445), which then propagates NaN through the solver. */
#define PKG_FLAG_EPA_GUARD 8u
/* Mixed-precision contact solve (not in the reference, whose solver is FP64): the
 * rows are built in FP64 and solved in FP32 (J.v, lambda, clamps, friction block),
 * each body's velocity increment added to its FP64 velocity. Geometry (pairs, GJK,
 * EPA, manifolds) stays FP64 and bit-exact; results drift from the FP64 solve. */
#define PKG_FLAG_FP32_SOLVE 16u
/* FP32 certified separation ahead of the FP64 GJK (not in the reference): a pair whose
 * shapes FP32 proves farther apart than GJK's margin, with a bound on the FP32
 * rounding error, is given the reference's verdict (separated) without the FP64 GJK;
 * every other pair runs the FP64 GJK. Verdicts are bit-exact. */
#define PKG_FLAG_FP32_GJK 32u
/* FP32 leaf boxes in the broadphase walk (not in the reference): leaf boxes rounded
 * outward to FP32 are tested first and only their overlaps take the exact FP64 test,
 * so the pair set is bit-exact. */
#define PKG_FLAG_FP32_BROAD 64u
/* The mixed-precision form: FP32 broadphase leaf tests, FP32 certified GJK separation,
 * FP32 contact solve; EPA, manifolds and body state stay FP64. */
#define PKG_FLAG_MIXED (PKG_FLAG_FP32_SOLVE | PKG_FLAG_FP32_GJK | PKG_FLAG_FP32_BROAD)

/* Mirrors world_desc (include/physkit/core/world.h:23-49): gravity and solver
 * iterations; defaults there are (0,-9.81,0) and 10. */
typedef struct pkg_config
{
    double gravity[3];
    uint32_t solver_iterations;
    uint32_t flags;
} pkg_config;

/* Bodies, structure of arrays, n entries each; quaternions are (w, x, y, z);
 * inertia is the local 3x3 tensor row-major (object_desc, include/physkit/core/
 * object.h:15-109). A static body (is_static=1) has infinite mass, as
 * object_desc::mass() returns for body_type::stat (object.h:81-85); its mass entry
 * is ignored. is_static=2 marks a disabled slot (a removed body's arena slot kept so
 * that the other ids do not change): it never moves and never pairs. */
typedef struct pkg_body_desc
{
    const uint8_t *is_static;
    const uint32_t *shape;
    const double *pos;         /* 3n */
    const double *vel;         /* 3n */
    const double *angvel;      /* 3n */
    const double *quat;        /* 4n */
    const double *mass;        /* n  */
    const double *inertia;     /* 9n */
    const double *restitution; /* n  */
    const double *friction;    /* n  */
} pkg_body_desc;

/* Per-body state written by pkg_download_bodies (any pointer may be NULL). */
typedef struct pkg_body_state
{
    double *pos;    /* 3n */
    double *vel;    /* 3n */
    double *angvel; /* 3n */
    double *quat;   /* 4n */
    double *fat;    /* 6n: broadphase leaf bounds min xyz, max xyz */
    uint8_t *moved; /* n: leaf reinserted in the last step */
} pkg_body_state;

/* Counters of the last step, and running totals since the bodies were uploaded (a
 * caller stepping many steps in one pkg_step call reads the pairs it went through). */
typedef struct pkg_counters
{
    uint64_t moved;    /* leaves reinserted (dynamic_bvh::update_leaf returned true) */
    uint64_t pairs;    /* active broadphase pairs = pairs through GJK               */
    uint64_t gjk_hits; /* pairs whose GJK found an intersection                    */
    uint64_t contacts; /* manifold contact points after the narrowphase            */
    uint64_t rows;     /* contact solver points (penetration > 0)                   */
    uint64_t colours;  /* solver colours                                            */
    uint64_t epa_overflow; /* pairs whose EPA needed the large-polytope path         */
    uint64_t candidates;   /* candidate pairs emitted by the tree query              */
    uint64_t gjk_listed;   /* pairs the bounding-volume cull left to the GJK stage   */
    uint64_t gjk_fp64;     /* of those, pairs the FP64 GJK ran (all of them unless
                              PKG_FLAG_FP32_GJK certified the others apart)          */
    uint64_t steps_total;    /* steps since pkg_upload_bodies                          */
    uint64_t pairs_total;    /* the pairs of those steps, summed                       */
    uint64_t gjk_hits_total; /* their GJK hits, summed                                 */
} pkg_counters;

/* Replaces constructing a physkit::world (world.h:382-385). */
pkg_status pkg_create(pkg_ctx **out, const pkg_config *cfg);
pkg_status pkg_destroy(pkg_ctx *ctx);
const char *pkg_last_error(const pkg_ctx *ctx);

/* Convex meshes (mesh::make / box / sphere / pyramid, src/mesh.cpp:16-137):
 * n_shapes shapes, shape s owning vertices [offsets[s], offsets[s+1]) of verts (xyz).
 * The local AABB is aabb::from_points over the vertices (mesh.h:191). With bodies
 * loaded, an upload that drops a shape a body uses is refused (PKG_INVALID_ARG). */
pkg_status pkg_upload_shapes(pkg_ctx *ctx, uint32_t n_shapes, const uint32_t *offsets,
                             const double *verts);

/* Replaces world_base::create_rigid for n bodies with ids 0..n-1 (world.h:202-208);
 * resets pairs and manifolds. Inverse mass/inertia and the world inertia follow
 * particle::particle (include/physkit/core/particle.h:16-28, 140-146); each leaf
 * starts with its exact world bounds, as dynamic_bvh::add stores them (bvh.h:294-302). */
pkg_status pkg_upload_bodies(pkg_ctx *ctx, uint32_t n, const pkg_body_desc *bodies);

/* Batched independent worlds. The reference steps one world per physkit::world
 * (world.h:379-424); a caller with many worlds concatenates their bodies, uploads them
 * once, and declares bodies [first_body[w], first_body[w+1]) as world w
 * (first_body has n_worlds + 1 entries, from 0 to the body count). Bodies of
 * different worlds never pair, and each world's result is the one it would reach
 * stepped alone: pair keys stay global body ids (subtract first_body[w] for the
 * world's own keys), and the solver colours each world from its own pair order.
 * pkg_upload_bodies resets to a single world; this call resets pairs and manifolds. */
pkg_status pkg_set_worlds(pkg_ctx *ctx, uint32_t n_worlds, const uint32_t *first_body);

/* Overwrites kinematic state of existing bodies (a caller moving objects between
 * steps). ids: count entries; arrays as in pkg_body_desc (pos/vel/angvel/quat). */
pkg_status pkg_set_body_state(pkg_ctx *ctx, uint32_t count, const uint32_t *ids,
                              const double *pos, const double *vel, const double *angvel,
                              const double *quat);

/* world::step_impl(dt) (src/world.cpp:20-56), nsteps times. Synchronous: returns
 * when the device results of the last step are final. Within one call each step's
 * position integration runs in the next step's first kernel (one pass over the
 * bodies instead of two); the results are bit-identical to nsteps calls of one step.
 * A step that fails (any status but PKG_OK: PKG_OOM, PKG_OVERFLOW_*, PKG_DEVICE_ERROR)
 * leaves the bodies part-way through it (velocities may carry its gravity, leaf bounds
 * may be updated and their moved flags used up). The ctx then refuses pkg_step and
 * pkg_step_timed (PKG_DEVICE_ERROR) until pkg_upload_bodies loads bodies again;
 * pkg_download_bodies still returns the part-way state. */
pkg_status pkg_step(pkg_ctx *ctx, double dt, int nsteps);

pkg_status pkg_download_bodies(pkg_ctx *ctx, const pkg_body_state *out);

/* Sorted active pair keys (pair_manager::active_pairs, collision_phases.h:54). */
pkg_status pkg_get_pairs(pkg_ctx *ctx, uint64_t *keys, size_t cap, size_t *n);

/* Manifolds in pair-key order (narrow_phase::manifolds, collision_phases.h:322):
 * counts[p] contacts for pair p and data[(p*4+c)*13 + f] with f = normal xyz,
 * local_a xyz, local_b xyz, depth, normal impulse, tangent impulses (2). */
pkg_status pkg_get_manifolds(pkg_ctx *ctx, int32_t *counts, double *data, size_t cap,
                             size_t *n);

/* Begin (1) / end (2) events of the last step in pair-key order: what
 * narrow_phase::calculate reports through on_coll_beg / on_coll_end
 * (collision_phases.h:314-318, world.h:327-338). */
pkg_status pkg_get_events(pkg_ctx *ctx, uint64_t *keys, int32_t *kinds, size_t cap,
                          size_t *n);

pkg_status pkg_get_counters(pkg_ctx *ctx, pkg_counters *out);

/* Colours of the solver manifolds of the last step (manifolds with at least one
 * contact row, in pair-key order), and optionally their pair keys. No counterpart in
 * the reference (its solver is sequential, constraint.h:1107-1201): a test hook. */
pkg_status pkg_get_colours(pkg_ctx *ctx, uint32_t *colours, uint64_t *keys, size_t cap, size_t *n);

/* ---- joint constraints (include/physkit/collision/constraint.h:139-834) ----
 * Types are numbered as constraint_type (constraint.h:14-23). A joint is identified by
 * its key (type << 24) | slot, slot being the index the caller's per-type arena gave
 * it (arena.h:64-80; the reference keeps one arena per type, world.h:401,
 * constraint.h:1018-1021). Each step the joint rows are built after the contact rows
 * and, as in constraint_solver::solve_constraints (constraint.h:1107-1201), warm
 * started and iterated before the contacts; joints are solved in FP64 whatever the
 * flags. The device colours joints like manifolds (no two joints of a colour share a
 * dynamic body) and sweeps the colours in order. */
#define PKG_JOINT_DISTANCE 0
#define PKG_JOINT_BALL_SOCKET 1
#define PKG_JOINT_HINGE 2
#define PKG_JOINT_SLIDER 3
#define PKG_JOINT_WELD 4
#define PKG_JOINT_SPRING 5
#define PKG_JOINT_KEY(type, slot) ((((uint32_t)(type)) << 24) | ((uint32_t)(slot)))

/* The fields of the six desc structs (constraint.h:144-175 distance, 239-285 spring,
 * 357-374 ball_socket, 446-497 hinge, 585-609 slider, 737-755 weld); a type reads only
 * its own fields. Vectors are xyz. */
typedef struct pkg_joint_desc
{
    int32_t type;            /* PKG_JOINT_* */
    uint32_t slot;           /* < 2^24 */
    uint32_t a, b;           /* body ids, a != b */
    double local_a[3];       /* distance, spring, hinge: anchor in body a's frame */
    double local_b[3];       /* distance, spring, hinge: anchor in body b's frame */
    double local_axis_a[3];  /* hinge: axis in a's frame */
    double local_axis_b[3];  /* hinge: axis in b's frame */
    double anchor[3];        /* ball_socket, slider, weld: world anchor */
    double axis[3];          /* slider: world axis */
    double distance;         /* distance, spring: rest length */
    double stiffness;        /* spring */
    double damping;          /* spring */
} pkg_joint_desc;

/* world::add_constraint (world.h:387-395) for n joints. As the reference's
 * constructors do, ball_socket, slider and weld fix their frames from the bodies'
 * current poses (the state after the last step or upload). PKG_INVALID_ARG if a key
 * is already live or repeated, a type or slot is out of range, a body id is out of
 * range, names a disabled slot, or a == b; or if the ctx runs batched worlds. */
pkg_status pkg_add_joints(pkg_ctx *ctx, const pkg_joint_desc *descs, size_t n);

/* world::remove_constraint (world.h:397-398). PKG_INVALID_ARG if a key is not live. */
pkg_status pkg_remove_joints(pkg_ctx *ctx, const uint32_t *keys, size_t n);

/* Live joints in key order and, optionally, their accumulated impulses (6 per joint,
 * rows beyond the type's jacobian_dimension are 0). *n = number of live joints. */
pkg_status pkg_get_joints(pkg_ctx *ctx, uint32_t *keys, double *impulses, size_t cap, size_t *n);

/* ---- world queries over the broadphase tree ----
 * They see every live body's tree bounds: a dynamic body's fat box, the exact box a
 * static body was added with (bvh.h:294-302; bvh.cpp:475-514), as they stand after the
 * last step (between steps the tree is the last step's; after an upload it is built on
 * first use). ray_world / box_world (optional, batched contexts): query i searches only
 * world ray_world[i]; null searches every body. Arrays are host pointers unless the
 * name ends in _device.
 *
 * Rays: origins and dirs xyz (dirs are normalized, as ray::ray does, bvh.h:46-49),
 * max_dist per ray; a body is hit when ray::intersect_distance(bounds, max_dist)
 * (bvh.h:58-91) returns a distance: the entry distance, 0 for an origin inside. */

/* world_base::raycast (world.h:259-319) for n_rays rays: the hits of ray i are
 * ids/dists[offsets[i] .. offsets[i+1]) sorted by (distance, body). The reference yields
 * them in its trees' walk order, merged from the static and dynamic trees; that order
 * depends on the trees' shapes and is not reproduced. offsets has n_rays + 1 entries;
 * *total = all hits. If cap < *total only offsets and *total are written. */
pkg_status pkg_raycast(pkg_ctx *ctx, uint32_t n_rays, const double *origins, const double *dirs,
                       const double *max_dist, const uint32_t *ray_world, uint32_t *offsets, uint32_t *ids,
                       double *dists, size_t cap, size_t *total);

/* The first hit of each ray in (distance, body) order: the first element of the sorted
 * hit list (dynamic_bvh::raycast with a callback that returns the hit distance,
 * bvh.h:355-404, prunes the same way). ids[i] = 0xFFFFFFFF and dists[i] = +inf when
 * ray i hits nothing. */
pkg_status pkg_raycast_closest(pkg_ctx *ctx, uint32_t n_rays, const double *origins, const double *dirs,
                               const double *max_dist, const uint32_t *ray_world, uint32_t *ids, double *dists);
/* The same with every array in device memory (the ctx's device); runs on the ctx's
 * stream and returns when the results are written. kernel_ms (optional, host): device
 * time of the query kernel, from HIP events around it on the ctx stream. */
pkg_status pkg_raycast_closest_device(pkg_ctx *ctx, uint32_t n_rays, const double *origins, const double *dirs,
                                      const double *max_dist, const uint32_t *ray_world, uint32_t *ids,
                                      double *dists, double *kernel_ms);

/* dynamic_bvh::query_aabb (bvh.h:313-344) over both trees for n_boxes boxes (min xyz,
 * max xyz): bodies whose bounds intersect box i (aabb::intersects, bounds.h:87-92) are
 * ids[offsets[i] .. offsets[i+1]), ascending. Same cap protocol as pkg_raycast. */
pkg_status pkg_query_aabb(pkg_ctx *ctx, uint32_t n_boxes, const double *boxes, const uint32_t *box_world,
                          uint32_t *offsets, uint32_t *ids, size_t cap, size_t *total);

/* pkg_step with per-stage device time, measured with HIP events on the ctx stream.
 * Stages: 0 integrate_vel + leaf update, 1 LBVH build, 2 pair-set update, 3 GJK
 * stage (cull, FP32 pass, GJK kernel), 4 EPA kernel(s), 5 manifold update kernel,
 * 6 contact solver (rows, colouring, islands, iterations, impulse write-back),
 * 7 integrate_pos, 8 the solver's iteration kernels alone (part of stage 6).
 * stage_ms[i] is the sum over the nsteps steps (up to 9 entries). */
pkg_status pkg_step_timed(pkg_ctx *ctx, double dt, int nsteps, double *stage_ms, int n_stages);

/* ---- launch shapes (not in the reference) ----
 * Work-list, tile and pairs-per-wavefront shapes of the narrowphase kernels. None of
 * them changes a result (the support search returns the lowest-index maximum however
 * its lanes are grouped), only the time a step takes. */
typedef struct pkg_tuning
{
    int32_t gjk_lanes;     /* lanes per pair in the GJK kernel: 1, 2, 4 or 8 (64 / lanes pairs per
                              wavefront); default 2 for shapes averaging at most 14 support
                              candidates, else 4 */
    int32_t epa_block;     /* threads per workgroup of the one-solve-per-thread EPA kernel: 64, 128
                              or 256 (default 64) */
    uint32_t epa_lds_max;  /* GJK-hit batches of at most this many run one EPA solve per workgroup
                              with the polytope in LDS; larger ones one solve per thread (512) */
    uint32_t tree_refits;  /* steps between full LBVH builds that refit the last build's hierarchy
                              to the new leaf boxes instead (0-64; default 7; 0 builds every
                              step). Pairs and query answers do not depend on the tree's shape,
                              so this changes only the walks' cost. Not chosen by pkg_autotune */
    float gjk_ms[4];       /* pkg_autotune: GJK stage ms per step with 1, 2, 4, 8 lanes per pair */
    float epa_block_ms[3]; /* EPA stage ms per step with workgroups of 64, 128, 256 threads */
    float epa_lds_ms[3];   /* EPA stage ms per step with epa_lds_max 0, 512, 4096 */
} pkg_tuning;

/* Picks the launch shapes above by timing them on this world. It takes `steps` real
 * steps (the world advances as under pkg_step, with the same results); in each, after
 * the GJK and EPA stages have run, it runs them again with every candidate shape on
 * that step's own pairs and hits (both stages are idempotent) and times each run. It
 * keeps the fastest of each: GJK lanes on the GJK stage, the EPA block size with every
 * batch per thread, then the LDS threshold with the step's fastest block size.
 * pkg_upload_bodies returns to the defaults. out (optional) receives the choice and the
 * mean times measured. */
pkg_status pkg_autotune(pkg_ctx *ctx, double dt, int steps, pkg_tuning *out);
pkg_status pkg_get_tuning(pkg_ctx *ctx, pkg_tuning *out);
/* Sets the shapes directly (the *_ms fields are ignored); PKG_INVALID_ARG for values
 * outside the lists above. */
pkg_status pkg_set_tuning(pkg_ctx *ctx, const pkg_tuning *t);

/* Standalone narrowphase: gjk_epa(a, b) (src/collision.cpp:512-518) over n shape
 * pairs. Shape i of side a: kind_a[i] (0 aabb, 1 obb, 2 mesh instance), params
 * 10 doubles at prm_a[10*i] (aabb: min3 max3; obb: center3 quat4 half3; mesh:
 * pos3 quat4) and, for meshes, the vertex range [vbeg_a[i], vend_a[i]) of verts.
 * out[10*i]: normal3, world_a3, world_b3, depth; hit[i]: 0 separated, 1 contact,
 * 2 GJK hit with a degenerate EPA simplex (no contact). All pointers are host
 * pointers. */
pkg_status pkg_gjk_epa_batch(pkg_ctx *ctx, uint32_t n, const int32_t *kind_a,
                             const double *prm_a, const uint32_t *vbeg_a, const uint32_t *vend_a,
                             const int32_t *kind_b, const double *prm_b, const uint32_t *vbeg_b,
                             const uint32_t *vend_b, const double *verts, uint32_t n_verts,
                             double *out, int32_t *hit);

/* No counterpart in the reference. For callers that cannot allocate device memory
 * (code under stream capture, tight serving loops, harnesses that forbid hipMalloc):
 * pkg_gjk_epa_batch over arrays that are already in device memory (the ctx's device),
 * with the same arguments, results and errors, run on the ctx's stream; returns when the
 * results are written. It calls no hipMalloc or hipFree: its only scratch is a static
 * device array of the library (about 0.8 MB per GPU, held until the process ends, set up
 * by pkg_create), which limits n to PKG_NOALLOC_MAX_PAIRS (PKG_INVALID_ARG above it).
 * Every call on one GPU shares that array, so calls that overlap on one GPU (from
 * several ctx or threads) race. Without UV-sphere tags every support scans all
 * vertices: identical results, slower for UV spheres. A pair with an invalid kind or
 * vertex range gets hit = -1 and the call returns PKG_INVALID_ARG. */
#define PKG_NOALLOC_MAX_PAIRS 65536u
pkg_status pkg_gjk_epa_batch_device(pkg_ctx *ctx, uint32_t n, const int32_t *kind_a,
                                    const double *prm_a, const uint32_t *vbeg_a, const uint32_t *vend_a,
                                    const int32_t *kind_b, const double *prm_b, const uint32_t *vbeg_b,
                                    const uint32_t *vend_b, const double *verts, uint32_t n_verts,
                                    double *out, int32_t *hit);

/* ---- synthetic scenes (input generation for tests and benchmarks) ---- */
typedef struct pkg_scene pkg_scene;
typedef struct pkg_scene_info
{
    uint32_t n_shapes;
    uint32_t n_verts;
    uint32_t n_bodies;
    uint32_t n_static;
    double dt;
    uint32_t iterations;
    uint32_t steps;
    uint32_t flags; /* pkg_config.flags the scene is meant to run with */
    uint32_t n_worlds; /* independent worlds (kind 3), else 1 */
} pkg_scene_info;

/* kind 0: 1,024 random boxes over a ground box; kind 1: convex-hull clusters
 * (narrowphase stress); kind 2: box pyramids (solver stress); kind 3: 4,096
 * independent worlds of 16-4,096 bodies (spheres, boxes, hulls), for pkg_set_worlds;
 * kind 4: granular pile of 2,097,152 small spheres (16x32 UV meshes) and boxes falling
 * into a 64-box hopper; kind 5: the same with 8x12 spheres; kind 6: kind 4 with every
 * length times 10. n_override != 0 changes the body count (kind 0, 1, 4-6), pyramid
 * count (kind 2) or world count (kind 3). */
pkg_status pkg_scene_create(pkg_scene **out, int kind, uint64_t seed, uint32_t n_override);
pkg_status pkg_scene_info_get(const pkg_scene *s, pkg_scene_info *info);
/* Copies the scene into caller buffers sized from pkg_scene_info. */
pkg_status pkg_scene_copy(const pkg_scene *s, uint32_t *shape_offsets, double *verts,
                          uint8_t *is_static, uint32_t *shape, double *pos, double *vel,
                          double *angvel, double *quat, double *mass, double *inertia,
                          double *restitution, double *friction);
/* first_body: n_worlds + 1 entries (the world partition for pkg_set_worlds). */
pkg_status pkg_scene_copy_worlds(const pkg_scene *s, uint32_t *first_body);
pkg_status pkg_scene_destroy(pkg_scene *s);

#ifdef __cplusplus
}
#endif

#endif /* PHYSKIT_AMD_H */
