// C ABI and host orchestration of one world step on one HIP stream.
// Stage order is world::step_impl (src/world.cpp:20-56):
//   integrate_vel + leaf update -> pairs -> narrowphase -> contact setup/solve ->
//   integrate_pos.
// Host synchronisation points (counts that size the next launch): candidate count,
// pair count, GJK hit count, EPA overflow count, solver rows, colour rounds.
#include "physkit_amd.h"
#include "pk_access.hpp"
#include "pk_gjk.hpp"
#include "uv_sphere.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
using namespace pk;

struct pk_error : std::runtime_error
{
    pkg_status st;
    pk_error(pkg_status s, const std::string &m) : std::runtime_error(m), st(s) {}
};

void hip_check(hipError_t e, const char *what)
{
    if (e == hipErrorOutOfMemory) throw pk_error(PKG_OOM, std::string(what) + ": out of device memory");
    if (e != hipSuccess) throw pk_error(PKG_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

// device buffer owned by the context; freed when the context is destroyed
template <class T> struct dbuf
{
    T *p = nullptr;
    size_t cap = 0;
    dbuf() = default;
    dbuf(const dbuf &) = delete;
    dbuf &operator=(const dbuf &) = delete;
    ~dbuf() { release(); }
    void reserve(size_t n) // contents are not kept
    {
        if (n <= cap) return;
        const bool regrow = cap > 0;
        release();
        size_t want = std::max<size_t>(n, 1);
        // A buffer that grows again (its count creeping up step by step) takes a quarter
        // more, so the step does not pay a hipFree (which waits for the device) and a
        // hipMalloc each time the count reaches a new high; exactly n if that does not fit.
        if (regrow)
        {
            const size_t roomy = want + want / 4;
            if (hipMalloc(reinterpret_cast<void **>(&p), roomy * sizeof(T)) == hipSuccess)
            {
                cap = roomy;
                return;
            }
            p = nullptr;
            (void)hipGetLastError();
        }
        hip_check(hipMalloc(reinterpret_cast<void **>(&p), want * sizeof(T)), "hipMalloc");
        cap = want;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    void swap(dbuf &o)
    {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
    }
};

__global__ void k_set_state(bodies_dev b, uint32_t count, const uint32_t *ids, const double *pos,
                            const double *vel, const double *ang, const double *quat)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    uint32_t i = ids[t];
    if (pos) st3(b.pos, b.n, i, mk(pos[3 * t], pos[3 * t + 1], pos[3 * t + 2]));
    if (vel) st3(b.vel, b.n, i, mk(vel[3 * t], vel[3 * t + 1], vel[3 * t + 2]));
    if (ang) st3(b.ang, b.n, i, mk(ang[3 * t], ang[3 * t + 1], ang[3 * t + 2]));
    if (quat) // particle::orientation(q) refreshes the derived state, particle.h:40-44
    {
        qt q = mkq(quat[4 * t], quat[4 * t + 1], quat[4 * t + 2], quat[4 * t + 3]);
        stq(b.quat, b.n, i, q);
        m3 R = rotation_matrix(q);
        m3 Rt = transpose(R);
        st9(b.Iw, b.n, i, mul(mul(R, ld9(b.I_local, b.n, i)), Rt));
        st9(b.Iw_inv, b.n, i, mul(mul(R, ld9(b.Iinv_local, b.n, i)), Rt));
    }
}

int id_bits(uint32_t n)
{
    int b = 0;
    while (b < 32 && (static_cast<uint64_t>(1) << b) < n) ++b;
    return b < 1 ? 1 : b;
}
} // namespace

namespace
{
template <class T> void upload(dbuf<T> &d, const T *h, size_t count, hipStream_t st)
{
    d.reserve(count);
    if (count) hip_check(hipMemcpyAsync(d.p, h, count * sizeof(T), hipMemcpyHostToDevice, st), "upload");
}
// AoS (k components per element) -> planar
std::vector<double> planar(const double *a, uint32_t n, int k)
{
    std::vector<double> o(static_cast<size_t>(n) * k);
    for (uint32_t i = 0; i < n; ++i)
        for (int c = 0; c < k; ++c) o[static_cast<size_t>(c) * n + i] = a[static_cast<size_t>(i) * k + c];
    return o;
}
void aos(const std::vector<double> &p, uint32_t n, int k, double *out)
{
    for (uint32_t i = 0; i < n; ++i)
        for (int c = 0; c < k; ++c) out[static_cast<size_t>(i) * k + c] = p[static_cast<size_t>(c) * n + i];
}
template <class T> std::vector<T> download(const dbuf<T> &d, size_t count, hipStream_t st)
{
    std::vector<T> h(count);
    if (count) hip_check(hipMemcpyAsync(h.data(), d.p, count * sizeof(T), hipMemcpyDeviceToHost, st), "download");
    hip_check(hipStreamSynchronize(st), "stream sync");
    return h;
}
} // namespace

// Scratch of pkg_gjk_epa_batch_device, the entry point that allocates no device memory:
// its counters, its list of solves that overflow the private EPA store and the large
// stores that rerun them. One copy per GPU (each GPU holds it until the process ends),
// shared by every call on that GPU: calls that overlap on one GPU race. Host code never
// uses the name `scratch` for its memory, only the per-device address that
// hipGetSymbolAddress returns (noalloc_ptr below).
namespace
{
constexpr uint32_t kNoAllocStores = 2;
struct noalloc_layout
{
    uint32_t words[64]; // counters (launch_gjk_epa_batch_dev); one 256-byte line
    uint32_t ovf_list[PKG_NOALLOC_MAX_PAIRS];
    pk::epa_store_big stores[kNoAllocStores];
};
constexpr int kMaxDevices = 64;
std::mutex noalloc_mu;
void *noalloc_ptr[kMaxDevices] = {};
} // namespace
static __device__ __align__(256) noalloc_layout scratch;

struct pkg_ctx
{
    hipStream_t st = nullptr;
    int device = 0;
    noalloc_layout *noalloc = nullptr; // this device's copy of `scratch`
    pkg_config cfg{};
    double gnorm = 9.81;
    std::string err;

    uint32_t n_shapes = 0, n_verts = 0;
    bool has_uv = false; // some shape is a tagged UV sphere
    std::vector<uint32_t> shape_cost; // support candidates per shape: 12 for a tagged UV sphere, else its vertices
    dbuf<uint32_t> shape_off, shape_voff, shape_uv;
    dbuf<double> verts, shape_box, shape_rad, cull_rec;
    dbuf<float4> verts32; // FP32 vertices for the certified-separation pass
    dbuf<float> rad32;

    uint32_t n = 0;
    dbuf<double> pos, vel, ang, acc, torque, quat, mass, inv_mass, I_local, Iinv_local, Iw, Iw_inv, rest,
        fric, fat;
    dbuf<uint32_t> shape;
    dbuf<uint8_t> is_static, moved;

    dbuf<unsigned long long> counters;
    unsigned long long *h_counters = nullptr;
    uint32_t candidates = 0; // candidate pairs of the last step
    // leaf count above which the pair query writes kPairRegions regions; the environment
    // variable PKG_PAIR_REGION_MIN_LEAVES, read at pkg_create, overrides it (tests use it
    // to take the region path on small scenes)
    uint32_t region_min_leaves = 1u << 20;

    // broadphase
    dbuf<uint64_t> morton, morton_alt;
    dbuf<uint32_t> ids, ids_alt, parent, escape, leaf_body, range_first, range_last, lastf;
    dbuf<int32_t> left, right;
    dbuf<float4> inode_lo, inode_hi, span_box, span_tab;
    dbuf<uint4> nrec;
    dbuf<double> leaf_box;
    dbuf<float4> leaf_f32; // with PKG_FLAG_FP32_BROAD
    dbuf<uint8_t> leaf_moved;
    dbuf<uint32_t> moved_list, moved_wg;
    dbuf<unsigned long long> scene;
    sort_scratch sc;
    // batched worlds
    uint32_t n_worlds = 1;
    dbuf<uint32_t> world_of, wstart, pair_base;
    dbuf<uint64_t> comb_keys, comb_keys_alt, akeys;
    dbuf<uint32_t> comb_vals, comb_vals_alt, flag, fpos, bsums, total, srcb, asrc;

    // pair list and manifold cache, double-buffered (cur = last step's result)
    dbuf<uint64_t> keys[2];
    dbuf<uint32_t> src;
    dbuf<uint8_t> mcount[2];
    dbuf<uint32_t> mslot[2];
    dbuf<double> mdata[2];
    dbuf<uint32_t> mowner[2]; // the pair of each data slot
    uint32_t mcap[2] = {0, 0};   // pairs the count and slot arrays hold
    uint32_t mdcap[2] = {0, 0};  // manifold data slots
    uint32_t mused[2] = {0, 0};  // slots in use after the step that wrote the set
    uint32_t last_hits = 0;      // GJK hits of the last step (sizes the hit lists)
    uint32_t P[2] = {0, 0};
    int cur = 0;
    dbuf<uint8_t> events;

    // narrowphase
    dbuf<uint32_t> hit_list, ovf_list, ovf2_list, gjk_list, gjk_list2, hit_of, mlive;
    // launch shapes: GJK lanes per pair defaults to 2 when the bodies' shapes average at
    // most 14 support candidates (boxes, certified UV spheres, small hulls), else 4, until
    // pkg_autotune or pkg_set_tuning picks them
    tuning tune;
    bool tuned = false;
    // pkg_autotune: while set, each step also times every GJK and EPA launch shape on
    // that step's own inputs (the stages are idempotent, so the step's results do not
    // change), adding the milliseconds to these sums
    struct shape_trials
    {
        hipEvent_t t0 = nullptr, t1 = nullptr;
        double gjk[4] = {0, 0, 0, 0}, epa_block[3] = {0, 0, 0}, epa_lds[3] = {0, 0, 0};
    } *trials = nullptr;
    dbuf<double> simplex, info;
    dbuf<uint8_t> status;
    dbuf<char> big_pool;

    // solver
    dbuf<uint32_t> colour, colour_start, order, order_alt;
    dbuf<uint32_t> par, iv, iv_alt, iflags, ipos, istarts; // contact islands
    dbuf<uint64_t> ik, ik_alt;
    dbuf<uint64_t> ckeys, ckeys_alt;
    dbuf<double> rows_s;      // solver rows in solve order
    dbuf<uint8_t> row_valid_s;
    dbuf<unsigned long long> claim, used; // colouring: per-body claims, colours taken per body
    dbuf<uint32_t> clist;                 // colouring: the row manifolds
    uint32_t n_colours = 0;
    uint32_t n_listed = 0; // manifolds with rows in clist after the last step
    uint32_t colour_rounds = 4; // rounds after which the last step's colouring was seen done

    // joints, dense in key order; an add or remove rebuilds the records into the other set
    std::vector<uint32_t> jkey_h;         // live keys
    std::vector<uint32_t> jab_h;          // bodies a, b of each live joint
    int jset = 0;
    uint32_t jcap[2] = {0, 0};
    dbuf<double> jrec[2];
    dbuf<uint8_t> jtype[2];
    dbuf<uint32_t> ja[2], jb[2];
    dbuf<uint64_t> jkeys[2];
    dbuf<double> jrows;
    dbuf<uint8_t> jnrows;
    dbuf<uint32_t> jcol, jorder, jorder_alt, jcs;
    dbuf<uint64_t> jck, jck_alt;
    std::vector<uint32_t> jcs_h;          // joint colour starts
    uint32_t *jord = nullptr;             // joints in colour order
    bool jcoloured = false;               // colouring valid for the joint set and static flags
    std::vector<uint8_t> h_static;        // is_static of the uploaded bodies

    // world queries
    bool tree_ok = false; // the LBVH holds the current tree bounds (built by the last step or query)
    bool topo_ok = false; // the LBVH's leaf order and hierarchy belong to the current bodies and worlds
    uint32_t topo_age = 0; // refits since the last full build
    bool rn_ok = false;   // q_rn holds the walk topology of that tree
    dbuf<uint4> q_rn;
    dbuf<uint32_t> q_cnt, q_off, q_hid, q_hq, q_perm, q_perm_alt, q_bsums, q_total, q_u32;
    dbuf<double> q_hd, q_org, q_dir, q_maxd, q_box, q_out_d;
    dbuf<uint64_t> q_keys, q_keys_alt;
    dbuf<uint32_t> q_out_id, q_world;
    hipEvent_t q_ev[2] = {nullptr, nullptr}; // query kernel timing

    pkg_counters last{};
    bool have_bodies = false;
    // a step failed part-way: body state holds part of that step, so stepping is refused
    // until pkg_upload_bodies loads bodies again
    bool poisoned = false;
    std::vector<uint32_t> h_shape; // shape of each uploaded body (checks shape re-uploads)
    void update_gjk_group()
    {
        double cost = 0.0;
        uint32_t live = 0;
        for (size_t i = 0; i < h_shape.size(); ++i)
            if (h_static[i] != 2)
            {
                cost += shape_cost[h_shape[i]];
                ++live;
            }
        if (!tuned) tune.gjk_group = (live && cost <= 14.0 * live) ? 2 : 4;
    }

    lbvh_dev tree();
    lbvh_dev build_tree(const bodies_dev &b, bool all_moved);
    // the LBVH of the current tree bounds: the last step's, or built now
    lbvh_dev query_tree()
    {
        if (!tree_ok)
        {
            bodies_dev b = B();
            reset_counter(C_TREE_FAIL);
            return build_tree(b, false);
        }
        return tree();
    }
    // offsets of per-query counts (exclusive scan) and the total
    uint32_t scan_counts(uint32_t nq);
    query_scratch qscratch(uint32_t m);

    joints_dev JD(int k)
    {
        joints_dev j;
        j.n = k == jset ? static_cast<uint32_t>(jkey_h.size()) : 0;
        j.cap = jcap[k];
        j.rec = jrec[k].p;
        j.type = jtype[k].p;
        j.a = ja[k].p;
        j.b = jb[k].p;
        j.keys = jkeys[k].p;
        j.rows = jrows.p;
        j.nrows = jnrows.p;
        return j;
    }
    // joint rows of this step and, when the joint set or the static flags changed, the
    // joint colouring (same rounds as the manifolds, priority from the joint index)
    void joint_prepare(const bodies_dev &b, double dt);
    void joints_rebuild(const std::vector<uint32_t> &keys, const std::vector<uint32_t> &ab,
                        const std::vector<uint32_t> &src, const std::vector<double> &desc,
                        const std::vector<uint32_t> &dint);

    bodies_dev B()
    {
        bodies_dev b;
        b.n = n;
        b.pos = pos.p;
        b.vel = vel.p;
        b.ang = ang.p;
        b.acc = acc.p;
        b.torque = torque.p;
        b.quat = quat.p;
        b.mass = mass.p;
        b.inv_mass = inv_mass.p;
        b.I_local = I_local.p;
        b.Iinv_local = Iinv_local.p;
        b.Iw = Iw.p;
        b.Iw_inv = Iw_inv.p;
        b.restitution = rest.p;
        b.friction = fric.p;
        b.fat = fat.p;
        b.shape = shape.p;
        b.is_static = is_static.p;
        b.moved = moved.p;
        return b;
    }
    shapes_dev S()
    {
        shapes_dev s;
        s.n = n_shapes;
        s.off = shape_off.p;
        s.voff = shape_voff.p;
        s.verts = verts.p;
        s.rad = shape_rad.p;
        s.local_box = shape_box.p;
        s.uv = shape_uv.p;
        return s;
    }
    shapes32_dev S32()
    {
        shapes32_dev s;
        s.v = verts32.p;
        s.off = shape_off.p;
        s.rad = rad32.p;
        s.uv = shape_uv.p;
        return s;
    }
    manifolds_dev MF(int k)
    {
        manifolds_dev m;
        m.cap = mcap[k];
        m.count = mcount[k].p;
        m.slot = mslot[k].p;
        m.dcap = mdcap[k];
        m.data = mdata[k].p;
        m.owner = mowner[k].p;
        return m;
    }

    // counters [lo, hi] in one copy and one synchronisation
    void read_counters(int lo, int hi)
    {
        hip_check(hipMemcpyAsync(h_counters + lo, counters.p + lo, sizeof(unsigned long long) * (hi - lo + 1),
                                 hipMemcpyDeviceToHost, st),
                  "counter copy");
        hip_check(hipStreamSynchronize(st), "stream sync");
    }
    unsigned long long read_counter(int slot)
    {
        hip_check(hipMemcpyAsync(h_counters + slot, counters.p + slot, sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost, st),
                  "counter copy");
        hip_check(hipStreamSynchronize(st), "stream sync");
        return h_counters[slot];
    }
    void reset_counter(int slot) { launch_fill_u64(counters.p + slot, 0ull, 1, st); }

    // defer_pos: another step follows in the same call, so this step's integrate_pos is
    // left to the next step's first kernel (pos_pending, pending_dt)
    void step_once(double dt, hipEvent_t *ev, bool defer_pos = false);
    bool pos_pending = false;
    double pending_dt = 0.0;
};

namespace
{
constexpr uint32_t kColourCheck = 4; // joint colouring: rounds between host checks of the remaining count

void record(hipEvent_t *ev, int k, hipStream_t st)
{
    if (ev) hip_check(hipEventRecord(ev[k], st), "event record");
}
} // namespace

lbvh_dev pkg_ctx::tree()
{
    lbvh_dev t;
    t.n = n;
    t.morton = morton.p;
    t.morton_alt = morton_alt.p;
    t.ids = ids.p;
    t.ids_alt = ids_alt.p;
    t.left = left.p;
    t.right = right.p;
    t.first = range_first.p;
    t.last = range_last.p;
    t.lastf = lastf.p;
    t.nrec = nrec.p;
    t.inode_lo = inode_lo.p;
    t.inode_hi = inode_hi.p;
    t.parent = parent.p;
    t.escape = escape.p;
    t.leaf_box = leaf_box.p;
    t.leaf_f32 = (cfg.flags & PKG_FLAG_FP32_BROAD) ? leaf_f32.p : nullptr;
    t.leaf_body = leaf_body.p;
    t.leaf_moved = leaf_moved.p;
    t.moved_list = moved_list.p;
    t.moved_wg = moved_wg.p;
    t.n_chunks = refit_chunks(n);
    t.span_levels = refit_levels(n);
    t.span_box = span_box.p;
    t.span_tab = span_tab.p;
    t.scene = scene.p;
    t.n_worlds = n_worlds;
    t.wbits = id_bits(n_worlds);
    t.world_of = n_worlds > 1 ? world_of.p : nullptr;
    t.wstart = n_worlds > 1 ? wstart.p : nullptr;
    return t;
}

lbvh_dev pkg_ctx::build_tree(const bodies_dev &b, bool all_moved)
{
    lbvh_dev t = tree();
    if (topo_ok && topo_age < tune.tree_refits)
    {
        launch_lbvh_refit(t, b, all_moved, counters.p, st);
        ++topo_age;
    }
    else
    {
        launch_lbvh_build(t, b, all_moved, sc, counters.p, st);
        if (t.morton != morton.p)
        {
            std::swap(morton.p, morton_alt.p);
            std::swap(ids.p, ids_alt.p);
        }
        topo_ok = true;
        topo_age = 0;
    }
    tree_ok = true;
    rn_ok = false;
    return t;
}

void pkg_ctx::joint_prepare(const bodies_dev &b, double dt)
{
    const uint32_t J = static_cast<uint32_t>(jkey_h.size());
    jrows.reserve(static_cast<size_t>(6) * kJR * jcap[jset]);
    jnrows.reserve(jcap[jset]);
    joints_dev jd = JD(jset);
    launch_joint_rows(jd, b, dt, st);
    if (jcoloured) return;
    jcol.reserve(J);
    claim.reserve(std::max<uint32_t>(n, 1));
    launch_fill_u32(jcol.p, kUncoloured, J, st);
    used.reserve(std::max<uint32_t>(n, 1));
    launch_fill_u64(used.p, 0ull, n, st);
    launch_fill_u64(claim.p, 0ull, n, st);
    reset_counter(C_NCOL);
    for (uint32_t round = 0;; ++round)
    {
        launch_colour_claim(jd.keys, nullptr, J, nullptr, b, jcol.p, nullptr, nullptr, claim.p, round, st);
        reset_counter(C_REMAIN);
        launch_colour_assign(jd.keys, nullptr, J, nullptr, b, jcol.p, nullptr, nullptr, claim.p, used.p, round,
                             counters.p, st);
        if (round % kColourCheck == kColourCheck - 1 && read_counter(C_REMAIN) == 0) break;
        if (round > 1000000) throw pk_error(PKG_DEVICE_ERROR, "joint colouring did not terminate");
    }
    const uint32_t c = static_cast<uint32_t>(read_counter(C_NCOL));
    jorder.reserve(J);
    jorder_alt.reserve(J);
    jck.reserve(J);
    jck_alt.reserve(J);
    jcs.reserve(c + 1);
    jord = launch_colour_bucket(jcol.p, nullptr, J, J, c, jck.p, jck_alt.p, jorder.p, jorder_alt.p, jcs.p, sc, st);
    jcs_h.assign(c + 1, 0u);
    hip_check(hipMemcpyAsync(jcs_h.data(), jcs.p, (c + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, st),
              "joint colour copy");
    hip_check(hipStreamSynchronize(st), "stream sync");
    jcoloured = true;
}

void pkg_ctx::joints_rebuild(const std::vector<uint32_t> &keys, const std::vector<uint32_t> &ab,
                             const std::vector<uint32_t> &src, const std::vector<double> &desc,
                             const std::vector<uint32_t> &dint)
{
    const int o = jset, k = 1 - jset;
    const uint32_t J = static_cast<uint32_t>(keys.size());
    const uint32_t cap = std::max<uint32_t>(64u, (J + 63u) & ~63u);
    if (jcap[k] < cap || !jrec[k].p)
    {
        jrec[k].reserve(static_cast<size_t>(kJW) * cap);
        jtype[k].reserve(cap);
        ja[k].reserve(cap);
        jb[k].reserve(cap);
        jkeys[k].reserve(cap);
        jcap[k] = cap;
    }
    if (J)
    {
        dbuf<uint32_t> dsrc, ddint;
        dbuf<double> ddesc;
        upload(dsrc, src.data(), src.size(), st);
        if (!desc.empty())
        {
            upload(ddesc, desc.data(), desc.size(), st);
            upload(ddint, dint.data(), dint.size(), st);
        }
        joints_dev oj = JD(o);
        joints_dev nj = JD(k);
        nj.n = J;
        launch_joint_rebuild(oj, nj, dsrc.p, ddesc.p, ddint.p, B(), st);
        hip_check(hipGetLastError(), "joint rebuild launch");
        hip_check(hipStreamSynchronize(st), "joint rebuild");
    }
    jset = k;
    jkey_h = keys;
    jab_h = ab;
    jcoloured = false;
}

uint32_t pkg_ctx::scan_counts(uint32_t nq)
{
    q_off.reserve(static_cast<size_t>(nq) + 1);
    q_bsums.reserve(scan_block_sums(nq));
    q_total.reserve(1);
    exclusive_scan_u32(q_cnt.p, q_off.p, nq, q_bsums.p, q_total.p, st);
    uint32_t total = 0;
    hip_check(hipMemcpyAsync(q_off.p + nq, q_total.p, sizeof(uint32_t), hipMemcpyDeviceToDevice, st), "offset copy");
    hip_check(hipMemcpyAsync(&total, q_total.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st), "hit total");
    hip_check(hipStreamSynchronize(st), "stream sync");
    return total;
}

query_scratch pkg_ctx::qscratch(uint32_t m)
{
    const size_t c = std::max<uint32_t>(m, 1);
    q_hid.reserve(c);
    q_hq.reserve(c);
    q_perm.reserve(c);
    q_perm_alt.reserve(c);
    q_hd.reserve(c);
    q_keys.reserve(c);
    q_keys_alt.reserve(c);
    query_scratch qs;
    qs.hid = q_hid.p;
    qs.hq = q_hq.p;
    qs.perm = q_perm.p;
    qs.perm_alt = q_perm_alt.p;
    qs.hd = q_hd.p;
    qs.keys = q_keys.p;
    qs.keys_alt = q_keys_alt.p;
    return qs;
}

void pkg_ctx::step_once(double dt, hipEvent_t *ev, bool defer_pos)
{
    const bool integrate = (cfg.flags & PKG_FLAG_NO_INTEGRATE) == 0;
    const bool solve = (cfg.flags & PKG_FLAG_NO_SOLVER) == 0;
    const bool all_moved = (cfg.flags & PKG_FLAG_ALL_MOVED) != 0;
    bodies_dev b = B();
    shapes_dev s = S();
    launch_fill_u64(counters.p, 0ull, C_NCOUNTERS, st);
    record(ev, 0, st);

    // ---- 1. gravity, integrate_vel, leaf update (world.cpp:22-34)
    if (integrate && pos_pending)
    {
        launch_integrate_pos_vel(b, s, pending_dt, cfg.gravity[0], cfg.gravity[1], cfg.gravity[2], dt, counters.p,
                                 st);
        pos_pending = false;
    }
    else if (integrate)
        launch_integrate_vel(b, s, cfg.gravity[0], cfg.gravity[1], cfg.gravity[2], dt, counters.p, st);
    else
        launch_fill_u32(reinterpret_cast<uint32_t *>(moved.p), 0u, (n + 3) / 4, st);
    record(ev, 1, st);

    // ---- 2. broadphase (collision_phases.h:377-436)
    lbvh_dev t = build_tree(b, all_moved);
    record(ev, 2, st);
    const uint32_t Pprev = P[cur];
    const int nxt = 1 - cur;
    // candidates: pairs with at least one moved leaf, each once
    // Above about a million leaves (16,384 wavefronts, each reserving once) the
    // reservations would serialise on one cursor: the query then writes kPairRegions
    // regions of comb_keys_alt, concatenated into comb_keys afterwards.
    const uint32_t nreg = n > region_min_leaves ? kPairRegions : 1u;
    dbuf<uint64_t> &qout = nreg > 1 ? comb_keys_alt : comb_keys;
    uint32_t M = 0, rcap = 0;
    region_counts rc{};
    for (int attempt = 0;; ++attempt)
    {
        size_t want = std::max<size_t>(qout.cap, 4 * static_cast<size_t>(n) + 1024 * nreg);
        qout.reserve(want);
        rcap = static_cast<uint32_t>(std::min<size_t>(qout.cap / nreg, 0xFFFFFFF0u / nreg));
        launch_fill_u64(counters.p + C_QCUR, 0ull, static_cast<size_t>(nreg) * kCursorStride, st);
        launch_pair_query(t, b, qout.p, rcap, nreg, counters.p, st);
        read_counters(C_MOVED, C_QCUR + static_cast<int>((nreg - 1) * kCursorStride));
        if (h_counters[C_TREE_FAIL]) throw pk_error(PKG_DEVICE_ERROR, "broadphase tree walk exceeded its node count");
        unsigned long long cnt = 0, most = 0;
        for (uint32_t r = 0; r < nreg; ++r)
        {
            const unsigned long long c = h_counters[C_QCUR + r * kCursorStride];
            rc.off[r] = static_cast<uint32_t>(std::min<unsigned long long>(cnt, 0xFFFFFFFFull));
            cnt += c;
            most = std::max(most, c);
        }
        if (cnt > 0xFFFFFFF0ull) throw pk_error(PKG_OVERFLOW_PAIRS, "candidate pair count exceeds 32-bit indexing");
        if (most <= rcap)
        {
            M = static_cast<uint32_t>(cnt);
            for (uint32_t r = nreg; r <= kPairRegions; ++r) rc.off[r] = M;
            break;
        }
        if (attempt > 2) throw pk_error(PKG_OVERFLOW_PAIRS, "candidate pair buffer could not be sized");
        qout.release();
        qout.reserve(static_cast<size_t>(nreg) * static_cast<size_t>(most + most / 4 + 1024));
    }
    candidates = M;
    comb_vals.reserve(M);
    comb_vals_alt.reserve(M);
    if (nreg > 1)
    {
        comb_keys.reserve(std::max<uint32_t>(M, 1));
        launch_pair_concat(comb_keys_alt.p, rcap, rc, comb_keys.p, st);
    }
    else
        comb_keys_alt.reserve(M);
    uint64_t *ck = comb_keys.p;
    if (M > 1)
    {
        int bits = id_bits(n);
        int shifts[16];
        int ns = 0;
        for (int sft = 0; sft < bits; sft += 8) shifts[ns++] = sft;
        for (int sft = 32; sft < 32 + bits; sft += 8) shifts[ns++] = sft;
        if (radix_sort_u64(comb_keys.p, comb_vals.p, comb_keys_alt.p, comb_vals_alt.p, M, shifts, ns, sc, st))
            ck = comb_keys_alt.p;
    }
    // carry-over: a candidate found in the previous sorted list keeps its manifold;
    // previous pairs no candidate covers stay while their leaves overlap
    uint32_t K = 0;
    bool kept_whole = false;
    if (!all_moved && Pprev && h_counters[C_MOVED] == 0)
    {
        // No leaf moved: there are no candidates and every previous pair keeps both fat
        // boxes, so the carry-over would keep the previous list whole and in order. It
        // is taken over as it stands, each pair's manifold coming from its own index
        // (no source map: a null map stands for that).
        if (M != 0) throw pk_error(PKG_DEVICE_ERROR, "broadphase found candidates with no leaf moved");
        K = Pprev;
        keys[nxt].swap(keys[cur]);
        kept_whole = true;
    }
    else
    {
        srcb.reserve(std::max<uint32_t>(M, 1));
        launch_pair_match(ck, M, keys[cur].p, Pprev, srcb.p, st);
        if (!all_moved && Pprev)
        {
            flag.reserve(Pprev);
            fpos.reserve(Pprev);
            bsums.reserve(scan_block_sums(Pprev));
            total.reserve(1);
            launch_prev_keep(keys[cur].p, Pprev, b, flag.p, st);
            exclusive_scan_u32(flag.p, fpos.p, Pprev, bsums.p, total.p, st);
            uint32_t h_total = 0;
            hip_check(hipMemcpyAsync(&h_total, total.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st), "total copy");
            hip_check(hipStreamSynchronize(st), "stream sync");
            K = h_total;
            akeys.reserve(std::max<uint32_t>(K, 1));
            asrc.reserve(std::max<uint32_t>(K, 1));
            launch_prev_compact(keys[cur].p, Pprev, flag.p, fpos.p, akeys.p, asrc.p, st);
        }
        if (static_cast<unsigned long long>(M) + K > 0xFFFFFFF0ull)
            throw pk_error(PKG_OVERFLOW_PAIRS, "pair count exceeds 32-bit indexing");
        keys[nxt].reserve(std::max<size_t>(static_cast<size_t>(M) + K, 1));
        src.reserve(std::max<size_t>(static_cast<size_t>(M) + K, 1));
        launch_merge(ck, srcb.p, M, akeys.p, K, keys[nxt].p, src.p, st);
        launch_merge(akeys.p, asrc.p, K, ck, M, keys[nxt].p, src.p, st);
    }
    const uint32_t Pn = M + K;
    const uint32_t *srcp = kept_whole ? nullptr : src.p;
    record(ev, 3, st);

    // ---- 3. narrowphase (collision_phases.h:244-320)
    // per-pair arrays hold every pair; per-hit arrays (hit list, simplex, contact data)
    // and manifold data only the GJK hits and the pairs in contact
    if (mcap[nxt] < Pn)
    {
        uint32_t c = std::max<uint32_t>(Pn + Pn / 4, 1024);
        mcount[nxt].release();
        mslot[nxt].release();
        mcount[nxt].reserve(c);
        mslot[nxt].reserve(c);
        mcap[nxt] = c;
    }
    uint32_t ncap = std::max<uint32_t>(mcap[nxt], 1);
    status.reserve(ncap);
    events.reserve(ncap);
    hit_of.reserve(ncap);
    gjk_out_dev g;
    g.status = status.p;
    g.hit_of = hit_of.p;
    auto size_hits = [&](uint32_t want)
    {
        want = std::max<uint32_t>(want, 1024);
        hit_list.reserve(want);
        ovf_list.reserve(want);
        simplex.reserve(static_cast<size_t>(want) * kSimplexWords);
        info.reserve(static_cast<size_t>(want) * 10);
        g.hit_list = hit_list.p;
        g.simplex = simplex.p;
        g.info = info.p;
        g.cap = static_cast<uint32_t>(std::min({hit_list.cap, ovf_list.cap, simplex.cap / kSimplexWords, info.cap / 10}));
    };
    size_hits(std::min<uint32_t>(Pn, last_hits + last_hits / 2));
    if (Pn)
    {
        gjk_list.reserve(Pn);
        cull_rec.reserve(22 * static_cast<size_t>(n)); // FP64 records, then FP32 spheres (mixed form)
        gjk32_in f32{};
        shapes32_dev s32 = S32();
        if (cfg.flags & PKG_FLAG_FP32_GJK)
        {
            gjk_list2.reserve(Pn);
            f32.s32 = &s32;
            f32.src = srcp;
            f32.prev_count = mcount[cur].p;
            f32.list2 = gjk_list2.p;
        }
        // the GJK stage; it is idempotent once its three counters are cleared (they start
        // the step at zero)
        // the cull also writes every pair's next manifold empty and lists the culled pairs
        // that had contacts: at most the previous step's non-empty manifolds
        const uint32_t live_cap = std::max<uint32_t>(mused[cur], 1);
        mlive.reserve(live_cap);
        cull_manifold_out mo{srcp, mcount[cur].p, MF(nxt), events.p, mlive.p, live_cap};
        auto gjk_stage = [&](int group, bool again)
        {
            if (again)
                for (int c : {C_GJK_LIST, C_GJK_LIST2, C_HITS, C_MLIVE}) reset_counter(c);
            launch_gjk(keys[nxt].p, Pn, b, s, has_uv, group, g, gjk_list.p, cull_rec.p, f32, mo, counters.p, st);
        };
        gjk_stage(tune.gjk_group, false);
        record(ev, 4, st);
        read_counters(C_HITS, C_MLIVE);
        uint32_t hits = static_cast<uint32_t>(h_counters[C_HITS]);
        if (h_counters[C_MLIVE] > live_cap)
            throw pk_error(PKG_DEVICE_ERROR, "more culled pairs with contacts than manifolds with contacts");
        if (hits > g.cap)
        {
            // more hits than the hit lists hold: grow them and run the stage again
            size_hits(hits + hits / 4);
            gjk_stage(tune.gjk_group, true);
            record(ev, 4, st);
            read_counters(C_HITS, C_MLIVE);
            hits = static_cast<uint32_t>(h_counters[C_HITS]);
            if (hits > g.cap) throw pk_error(PKG_DEVICE_ERROR, "GJK hit count changed between identical runs");
        }
        last_hits = hits;
        // the EPA stage, with overflow reruns in larger stores (capacity never changes a
        // result); it writes every hit's status and contact, so it too can run again
        auto epa_stage = [&](const tuning &tu, bool again)
        {
            if (again)
                for (int c : {C_EPA_OVF, C_EPA_OVF2, C_EPA_BIG_FAIL}) reset_counter(c);
            launch_epa(keys[nxt].p, hits, b, s, has_uv, tu, g, ovf_list.p, counters.p, st);
            const uint32_t novf = static_cast<uint32_t>(read_counter(C_EPA_OVF));
            if (!novf) return;
            // a small batch already ran with the LDS store: its overflows go straight to
            // the global-memory store
            const uint32_t *big_in = ovf_list.p;
            uint32_t n2 = novf;
            if (!epa_first_tier_is_lds(hits, tu))
            {
                ovf2_list.reserve(novf);
                launch_epa_lds(keys[nxt].p, ovf_list.p, novf, b, s, g, ovf2_list.p, counters.p, st);
                n2 = static_cast<uint32_t>(read_counter(C_EPA_OVF2));
                big_in = ovf2_list.p;
            }
            if (n2)
            {
                big_pool.reserve(static_cast<size_t>(n2) * epa_big_bytes());
                launch_epa_big(keys[nxt].p, big_in, n2, b, s, g, big_pool.p, counters.p, st);
                if (read_counter(C_EPA_BIG_FAIL))
                    throw pk_error(PKG_OVERFLOW_EPA, "EPA polytope exceeded 4096 faces");
            }
        };
        record(ev, 9, st);
        epa_stage(tune, false);
        record(ev, 5, st);
        if (trials)
        {
            auto timed = [&](auto &&run)
            {
                hip_check(hipEventRecord(trials->t0, st), "event record");
                run();
                hip_check(hipEventRecord(trials->t1, st), "event record");
                hip_check(hipEventSynchronize(trials->t1), "event sync");
                float ms = 0;
                hip_check(hipEventElapsedTime(&ms, trials->t0, trials->t1), "elapsed");
                return static_cast<double>(ms);
            };
            static const int lanes[4] = {1, 2, 4, 8}, blocks[3] = {64, 128, 256};
            static const uint32_t lds_max[3] = {0u, 512u, 4096u};
            for (int c = 0; c < 4; ++c) trials->gjk[c] += timed([&] { gjk_stage(lanes[c], true); });
            if (read_counter(C_HITS) != hits) throw pk_error(PKG_DEVICE_ERROR, "GJK hit count changed between runs");
            tuning tu = tune;
            tu.epa_lds_max = 0; // every batch through k_epa while its block size is timed
            int best = 0;
            double ms[3];
            for (int c = 0; c < 3; ++c)
            {
                tu.epa_block = blocks[c];
                ms[c] = timed([&] { epa_stage(tu, true); });
                trials->epa_block[c] += ms[c];
                if (ms[c] < ms[best]) best = c;
            }
            tu.epa_block = blocks[best];
            for (int c = 0; c < 3; ++c)
            {
                tu.epa_lds_max = lds_max[c];
                trials->epa_lds[c] += timed([&] { epa_stage(tu, true); });
            }
        }
        // a manifold is non-empty only if its pair had contacts or has a hit now: the
        // data store needs at most hits + the previous step's slots
        const uint32_t dneed = hits + mused[cur];
        if (mdcap[nxt] < dneed)
        {
            const uint32_t c = std::max<uint32_t>(dneed + dneed / 4, 1024);
            mdata[nxt].release();
            mdata[nxt].reserve(static_cast<size_t>(c) * 4 * kMF);
            mowner[nxt].release();
            mowner[nxt].reserve(c);
            mdcap[nxt] = c;
        }
        // with per-body records the cull settled every culled pair's manifold: the update
        // runs over the culled pairs that had contacts, then every pair GJK ran on;
        // otherwise it lists its own pairs into the GJK list's buffer (free again)
        const bool culled = cull_uses_records(Pn, n);
        const uint32_t bound = culled ? static_cast<uint32_t>(h_counters[C_MLIVE] + h_counters[C_GJK_LIST])
                                      : std::min(dneed, Pn);
        launch_manifold_update(keys[nxt].p, srcp, Pn, bound, b, g, MF(cur), MF(nxt), events.p,
                               (cfg.flags & PKG_FLAG_EPA_GUARD) != 0, culled, culled ? mlive.p : gjk_list.p,
                               gjk_list.p, counters.p, st);
    }
    else
    {
        record(ev, 4, st);
        record(ev, 9, st);
        record(ev, 5, st);
    }
    record(ev, 6, st);
    // the solve iterations' interval (events 10-11) is empty unless a solve runs
    record(ev, 10, st);
    record(ev, 11, st);

    // ---- 4. contacts and joints (constraint.h:1052-1201)
    n_colours = 0;
    n_listed = 0;
    const uint32_t n_joints = static_cast<uint32_t>(jkey_h.size());
    if (solve && (Pn || n_joints))
    {
        const bool f32 = (cfg.flags & PKG_FLAG_FP32_SOLVE) != 0;
        uint32_t c = 0;  // manifold colours (0: no contact rows)
        uint32_t nl = 0; // manifolds with rows, listed in clist
        const double thr = 2.0 * gnorm * dt;
        // Rows go straight into solve order (rs, built below from `ord`), as FP32 planes
        // for the FP32 solve.
        auto write_rows = [&](const uint32_t *ord, uint32_t n_act, const rows_dev &rs, uint32_t scap)
        { launch_contact_setup(keys[nxt].p, n_act, ord, scap, b, MF(nxt), rs, dt, thr, f32, st); };
        if (Pn)
        {
            colour.reserve(ncap);
            claim.reserve(std::max<uint32_t>(n, 1));
            clist.reserve(std::max<uint32_t>(mdcap[nxt], 1));
            // The manifolds with rows are listed as they are flagged (C_NLIST, from zero at
            // the step's start), and the colouring rounds read that count on the device
            // over a grid sized by the list's capacity: the row count, the list count and
            // the colouring's state come back together at its first check.
            launch_row_flags(keys[nxt].p, Pn, b, MF(nxt), colour.p, clist.p, counters.p, st);
            const uint32_t lcap = static_cast<uint32_t>(std::min<size_t>(clist.cap, 0xFFFFFFFFu));
            {
                const unsigned long long *nl_dev = counters.p + C_NLIST;
                const uint32_t *wo = nullptr, *pb = nullptr;
                if (n_worlds > 1)
                {
                    pair_base.reserve(n_worlds);
                    launch_pair_base(keys[nxt].p, Pn, world_of.p, pair_base.p, st);
                    wo = world_of.p;
                    pb = pair_base.p;
                }
                used.reserve(std::max<uint32_t>(n, 1));
                launch_fill_u64(used.p, 0ull, n, st);
                launch_fill_u64(claim.p, 0ull, n, st);
                // the host looks at the remaining count (with the colour count in the same
                // copy) first after as many rounds as the last step took, then every 2:
                // consecutive steps colour alike, so one wait usually does. Rounds after the
                // last manifold is coloured change nothing (and cost two short launches)
                const uint32_t first = std::max<uint32_t>(colour_rounds, 1);
                uint32_t check = first;
                for (uint32_t round = 0;; ++round)
                {
                    launch_colour_claim(keys[nxt].p, clist.p, lcap, nl_dev, b, colour.p, wo, pb, claim.p, round, st);
                    reset_counter(C_REMAIN);
                    launch_colour_assign(keys[nxt].p, clist.p, lcap, nl_dev, b, colour.p, wo, pb, claim.p, used.p,
                                         round, counters.p, st);
                    if (round + 1 == check)
                    {
                        read_counters(C_ROWS, C_NLIST); // rows, remaining, colours and the list count
                        if (h_counters[C_ROWS] == 0) break; // nothing to colour (the rounds did nothing)
                        if (h_counters[C_REMAIN] == 0)
                        {
                            // done at the first look: try one round fewer next time, so a
                            // step that needed many rounds does not hold later steps back
                            colour_rounds = check == first ? std::max<uint32_t>(check - 1, 1) : check;
                            break;
                        }
                        check += 2;
                    }
                    if (round > 1000000) throw pk_error(PKG_DEVICE_ERROR, "colouring did not terminate");
                }
                if (h_counters[C_ROWS] > 0)
                {
                    nl = static_cast<uint32_t>(h_counters[C_NLIST]);
                    c = static_cast<uint32_t>(h_counters[C_NCOL]);
                }
            }
        }
        n_colours = c;
        n_listed = nl;
        if (n_joints) joint_prepare(b, dt);
        bool solved = false;
        if (c && !n_joints && c <= kIslandMaxColours)
        {
            // one workgroup per chunk of whole contact islands when every island is small enough
            // over the listed row manifolds only
            par.reserve(std::max<uint32_t>(n, 1));
            ik.reserve(nl);
            ik_alt.reserve(nl);
            iv.reserve(nl);
            iv_alt.reserve(nl);
            iflags.reserve(nl);
            ipos.reserve(nl);
            istarts.reserve(nl);
            bsums.reserve(scan_block_sums(nl));
            total.reserve(2);
            reset_counter(C_ISL_MAX);
            island_sort is = launch_island_sort(keys[nxt].p, colour.p, clist.p, nl, c, b, par.p, ik.p, iv.p,
                                                ik_alt.p, iv_alt.p, counters.p + C_ISL_MAX, sc, st);
            launch_island_segments(is, nl, iflags.p, ipos.p, bsums.p, total.p, istarts.p, total.p + 1, st);
            // island count, row manifold count and the longest island in one wait
            launch_island_maxlen(istarts.p, total.p, nl, counters.p + C_ISL_MAX, st);
            uint32_t h2[2] = {0, 0};
            hip_check(hipMemcpyAsync(h2, total.p, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "island count");
            read_counters(C_ISL_MAX, C_ISL_MAX);
            const uint32_t n_isl = h2[0], n_act = h2[1];
            if (h_counters[C_ISL_MAX] <= kIslandMaxManifolds)
            {
                const uint32_t scap = std::max<uint32_t>(n_act, 1);
                rows_s.reserve(static_cast<size_t>(4) * scap * kRowF);
                row_valid_s.reserve(static_cast<size_t>(4) * scap);
                rows_dev rs;
                rs.cap4 = 4 * scap;
                rs.d = rows_s.p;
                rs.valid = row_valid_s.p;
                write_rows(is.ord, n_act, rs, scap);
                record(ev, 10, st);
                launch_solve_islands(keys[nxt].p, is, istarts.p, n_isl, n_act, c, b, rs, scap,
                                     cfg.solver_iterations, static_cast<uint32_t>(h_counters[C_ISL_MAX]), f32, st);
                record(ev, 11, st);
                if (f32)
                    launch_impulse_writeback32(is.ord, n_act, MF(nxt), rs, scap, st);
                else
                    launch_impulse_writeback(is.ord, n_act, MF(nxt), rs, scap, st);
                solved = true;
            }
        }
        if (!solved)
        {
            // per-colour launches: each pass sweeps the joint colours, then the manifold
            // colours (pass 0 warm starts, constraint.h:1123-1130; then the iterations)
            uint32_t *ord = nullptr;
            std::vector<uint32_t> cs;
            uint32_t n_act = 0, scap = 1;
            rows_dev rs{};
            if (c)
            {
                colour_start.reserve(c + 1);
                order.reserve(nl);
                order_alt.reserve(nl);
                ckeys.reserve(nl);
                ckeys_alt.reserve(nl);
                ord = launch_colour_bucket(colour.p, clist.p, nl, Pn, c, ckeys.p, ckeys_alt.p, order.p, order_alt.p,
                                           colour_start.p, sc, st);
                cs.assign(c + 1, 0u);
                hip_check(hipMemcpyAsync(cs.data(), colour_start.p, (c + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, st),
                          "colour copy");
                hip_check(hipStreamSynchronize(st), "stream sync");
                // rows of the solver manifolds gathered into solve order
                n_act = cs[c];
                scap = std::max<uint32_t>(n_act, 1);
                rows_s.reserve(static_cast<size_t>(4) * scap * kRowF);
                row_valid_s.reserve(static_cast<size_t>(4) * scap);
                rs.cap4 = 4 * scap;
                rs.d = rows_s.p;
                rs.valid = row_valid_s.p;
                write_rows(ord, n_act, rs, scap);
            }
            const joints_dev jd = JD(jset);
            const uint32_t jc = n_joints ? static_cast<uint32_t>(jcs_h.size()) - 1 : 0;
            record(ev, 10, st);
            for (uint32_t pass = 0; pass <= cfg.solver_iterations; ++pass)
            {
                const bool warm = pass == 0;
                for (uint32_t k = 0; k < jc; ++k) launch_joint_solve(jd, jord, jcs_h[k], jcs_h[k + 1], b, warm, st);
                for (uint32_t k = 0; k < c; ++k)
                {
                    if (f32)
                        launch_solve_colour32(keys[nxt].p, ord, cs[k], cs[k + 1], b, rs, scap, warm, st);
                    else
                        launch_solve_colour(keys[nxt].p, ord, cs[k], cs[k + 1], b, rs, scap, warm, st);
                }
            }
            record(ev, 11, st);
            if (c)
            {
                if (f32)
                    launch_impulse_writeback32(ord, n_act, MF(nxt), rs, scap, st);
                else
                    launch_impulse_writeback(ord, n_act, MF(nxt), rs, scap, st);
            }
            if (n_joints) launch_joint_writeback(jd, st);
        }
    }
    record(ev, 7, st);

    // ---- 5. integrate_pos (world.cpp:50-55)
    if (integrate && defer_pos)
    {
        pos_pending = true; // the next step_once of this call runs it first
        pending_dt = dt;
    }
    else if (integrate)
        launch_integrate_pos(b, dt, st);
    record(ev, 8, st);

    P[nxt] = Pn;
    cur = nxt;
    hip_check(hipMemcpyAsync(h_counters, counters.p, C_NCOUNTERS * sizeof(unsigned long long), hipMemcpyDeviceToHost, st),
              "counters copy");
    hip_check(hipStreamSynchronize(st), "stream sync");
    last.moved = h_counters[C_MOVED];
    last.pairs = Pn;
    last.gjk_hits = h_counters[C_HITS];
    last.contacts = h_counters[C_CONTACTS];
    last.rows = h_counters[C_ROWS];
    last.colours = n_colours;
    last.epa_overflow = h_counters[C_EPA_OVF];
    last.candidates = candidates;
    mused[cur] = Pn ? static_cast<uint32_t>(h_counters[C_MSLOTS]) : 0u;
    last.gjk_listed = Pn ? h_counters[C_GJK_LIST] : 0;
    last.gjk_fp64 = Pn ? ((cfg.flags & PKG_FLAG_FP32_GJK) ? h_counters[C_GJK_LIST2] : h_counters[C_GJK_LIST]) : 0;
    last.steps_total += 1;
    last.pairs_total += last.pairs;
    last.gjk_hits_total += last.gjk_hits;
}

// ------------------------------------------------------------------ C ABI
#define PKG_GUARD(ctx, ...)                                                                        \
    try                                                                                            \
    {                                                                                              \
        __VA_ARGS__;                                                                               \
        return PKG_OK;                                                                             \
    }                                                                                              \
    catch (const pk_error &e)                                                                      \
    {                                                                                              \
        if (ctx) (ctx)->err = e.what();                                                            \
        return e.st;                                                                               \
    }                                                                                              \
    catch (const std::exception &e)                                                                \
    {                                                                                              \
        if (ctx) (ctx)->err = e.what();                                                            \
        return PKG_DEVICE_ERROR;                                                                   \
    }


extern "C"
{

pkg_status pkg_create(pkg_ctx **out, const pkg_config *cfg)
{
    if (!out || !cfg) return PKG_INVALID_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PKG_NO_DEVICE;
    auto *c = new pkg_ctx();
    try
    {
        c->cfg = *cfg;
        const double *g = cfg->gravity;
        c->gnorm = std::sqrt((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]); // world.h:383
        hip_check(hipGetDevice(&c->device), "get device");
        if (c->device < 0 || c->device >= kMaxDevices) throw pk_error(PKG_DEVICE_ERROR, "device index out of range");
        {
            // the no-allocation entry point's scratch: its address on this device, once
            std::lock_guard<std::mutex> lk(noalloc_mu);
            if (!noalloc_ptr[c->device])
            {
                void *p = nullptr;
                hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch symbol address");
                noalloc_ptr[c->device] = p;
            }
            c->noalloc = static_cast<noalloc_layout *>(noalloc_ptr[c->device]);
        }
        hip_check(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking), "stream create");
        hip_check(hipHostMalloc(reinterpret_cast<void **>(&c->h_counters), (C_NCOUNTERS + 2) * sizeof(unsigned long long)),
                  "pinned alloc");
        c->counters.reserve(C_NCOUNTERS);
        if (const char *e = std::getenv("PKG_PAIR_REGION_MIN_LEAVES"))
            c->region_min_leaves = static_cast<uint32_t>(std::strtoul(e, nullptr, 10));
        c->scene.reserve(6);
    }
    catch (const pk_error &e)
    {
        delete c;
        return e.st;
    }
    *out = c;
    return PKG_OK;
}

pkg_status pkg_destroy(pkg_ctx *ctx)
{
    if (!ctx) return PKG_INVALID_ARG;
    if (ctx->st) (void)hipStreamSynchronize(ctx->st);
    // the dbuf members free their device memory in their destructors
    if (ctx->sc.hist) (void)hipFree(ctx->sc.hist);
    for (hipEvent_t e : ctx->q_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->h_counters) (void)hipHostFree(ctx->h_counters);
    if (ctx->st) (void)hipStreamDestroy(ctx->st);
    delete ctx;
    return PKG_OK;
}

const char *pkg_last_error(const pkg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

pkg_status pkg_upload_shapes(pkg_ctx *ctx, uint32_t n_shapes, const uint32_t *offsets, const double *verts)
{
    if (!ctx || !offsets || (n_shapes && !verts)) return PKG_INVALID_ARG;
    for (uint32_t i = 0; i < n_shapes; ++i)
        if (offsets[i + 1] <= offsets[i])
        {
            ctx->err = "every shape needs at least one vertex (mesh::make asserts non-empty, mesh.cpp:19)";
            return PKG_INVALID_ARG;
        }
    if (ctx->have_bodies)
        for (size_t i = 0; i < ctx->h_shape.size(); ++i)
            if (ctx->h_shape[i] >= n_shapes)
            {
                ctx->err = "body " + std::to_string(i) + " uses shape " + std::to_string(ctx->h_shape[i]) +
                           ", which this upload drops; upload the bodies again after the shapes";
                return PKG_INVALID_ARG;
            }
    PKG_GUARD(ctx, {
        uint32_t nv = offsets[n_shapes];
        std::vector<uint32_t> uv(std::max<uint32_t>(n_shapes, 1), 0u);
        pk::uv_sphere_detector det;
        for (uint32_t i = 0; i < n_shapes; ++i)
            uv[i] = det.tag(verts + 3 * static_cast<size_t>(offsets[i]), offsets[i + 1] - offsets[i]);
        upload(ctx->shape_uv, uv.data(), uv.size(), ctx->st);
        ctx->has_uv = std::any_of(uv.begin(), uv.end(), [](uint32_t t) { return t != 0; });
        ctx->shape_cost.resize(n_shapes);
        for (uint32_t i = 0; i < n_shapes; ++i) ctx->shape_cost[i] = uv[i] ? 12u : offsets[i + 1] - offsets[i];
        upload(ctx->shape_off, offsets, n_shapes + 1, ctx->st);
        // device vertex blocks: even start, even count (an odd block repeats its last
        // vertex, which is never scanned), so vertex pairs load as aligned 16-byte words
        std::vector<uint32_t> voff(std::max<uint32_t>(n_shapes, 1), 0u);
        std::vector<double> rad(std::max<uint32_t>(n_shapes, 1), 0.0);
        std::vector<double> pv;
        pv.reserve(3 * (static_cast<size_t>(nv) + n_shapes));
        for (uint32_t i = 0; i < n_shapes; ++i)
        {
            voff[i] = static_cast<uint32_t>(pv.size() / 3);
            const uint32_t c = offsets[i + 1] - offsets[i];
            const double *src = verts + 3 * static_cast<size_t>(offsets[i]);
            pv.insert(pv.end(), src, src + 3 * static_cast<size_t>(c));
            if (c & 1u) pv.insert(pv.end(), src + 3 * static_cast<size_t>(c - 1), src + 3 * static_cast<size_t>(c));
            for (uint32_t k = 0; k < c; ++k)
            {
                const double *q = src + 3 * static_cast<size_t>(k);
                const double r = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
                rad[i] = std::isfinite(r) ? std::max(rad[i], r) : std::numeric_limits<double>::infinity();
            }
        }
        if (pv.empty()) pv.assign(6, 0.0);
        upload(ctx->shape_voff, voff.data(), voff.size(), ctx->st);
        upload(ctx->shape_rad, rad.data(), rad.size(), ctx->st);
        // radii for the FP32 pass, rounded up so they still bound every vertex
        std::vector<float> rad32(rad.size());
        for (size_t i = 0; i < rad.size(); ++i)
        {
            float f = static_cast<float>(rad[i]);
            if (static_cast<double>(f) < rad[i]) f = std::nextafter(f, std::numeric_limits<float>::infinity());
            rad32[i] = f;
        }
        upload(ctx->rad32, rad32.data(), rad32.size(), ctx->st);
        upload(ctx->verts, pv.data(), pv.size(), ctx->st);
        ctx->shape_box.reserve(static_cast<size_t>(n_shapes) * 6);
        ctx->n_shapes = n_shapes;
        ctx->n_verts = nv;
        launch_shape_boxes(ctx->S(), ctx->st);
        ctx->verts32.reserve(std::max<uint32_t>(nv, 1));
        launch_verts32(ctx->S(), ctx->verts32.p, ctx->st);
        hip_check(hipStreamSynchronize(ctx->st), "shape upload");
        if (ctx->have_bodies) ctx->update_gjk_group(); // the loaded bodies' shapes changed
    })
}

pkg_status pkg_upload_bodies(pkg_ctx *ctx, uint32_t n, const pkg_body_desc *d)
{
    if (!ctx || !d) return PKG_INVALID_ARG;
    if (n && (!d->is_static || !d->shape || !d->pos || !d->vel || !d->angvel || !d->quat || !d->mass ||
              !d->inertia || !d->restitution || !d->friction))
        return PKG_INVALID_ARG;
    for (uint32_t i = 0; i < n; ++i)
        if (d->shape[i] >= ctx->n_shapes)
        {
            ctx->err = "body references a shape that was not uploaded (object must have a mesh, object.h:121)";
            return PKG_INVALID_ARG;
        }
    for (uint32_t x : ctx->jab_h)
        if (x >= n || d->is_static[x] == 2)
        {
            ctx->err = "a live joint references a body that is out of range or a disabled slot";
            return PKG_INVALID_ARG;
        }
    PKG_GUARD(ctx, {
        ctx->h_static.assign(d->is_static, d->is_static + n);
        ctx->h_shape.assign(d->shape, d->shape + n);
        ctx->tuned = false; // a new body set: back to the default shapes until tuned again
        ctx->update_gjk_group();
        ctx->jcoloured = false;
        hipStream_t st = ctx->st;
        ctx->n = n;
        std::vector<double> mass(n);
        for (uint32_t i = 0; i < n; ++i)
            mass[i] = d->is_static[i] ? std::numeric_limits<double>::infinity() : d->mass[i]; // object.h:81-85
        auto up = [&](dbuf<double> &b, const double *a, int k)
        {
            std::vector<double> p = planar(a, n, k);
            upload(b, p.data(), p.size(), st);
            hip_check(hipStreamSynchronize(st), "upload");
        };
        up(ctx->pos, d->pos, 3);
        up(ctx->vel, d->vel, 3);
        up(ctx->ang, d->angvel, 3);
        up(ctx->quat, d->quat, 4);
        up(ctx->mass, mass.data(), 1);
        up(ctx->I_local, d->inertia, 9);
        up(ctx->rest, d->restitution, 1);
        up(ctx->fric, d->friction, 1);
        upload(ctx->shape, d->shape, n, st);
        upload(ctx->is_static, d->is_static, n, st);
        for (auto *b : {&ctx->acc, &ctx->torque}) b->reserve(static_cast<size_t>(3) * n);
        ctx->inv_mass.reserve(n);
        for (auto *b : {&ctx->Iinv_local, &ctx->Iw, &ctx->Iw_inv}) b->reserve(static_cast<size_t>(9) * n);
        ctx->fat.reserve(static_cast<size_t>(6) * n);
        ctx->moved.reserve(n + 4);
        // broadphase arrays
        ctx->morton.reserve(n);
        ctx->morton_alt.reserve(n);
        ctx->ids.reserve(n);
        ctx->ids_alt.reserve(n);
        ctx->left.reserve(n);
        ctx->right.reserve(n);
        ctx->parent.reserve(2 * static_cast<size_t>(n));
        ctx->escape.reserve(2 * static_cast<size_t>(n));
        ctx->leaf_body.reserve(n);
        ctx->leaf_moved.reserve(n);
        ctx->moved_list.reserve(n);
        ctx->moved_wg.reserve(2 * ((static_cast<size_t>(n) + 255) / 256));
        for (auto *b : {&ctx->range_first, &ctx->range_last, &ctx->lastf}) b->reserve(n);
        ctx->span_box.reserve(4 * static_cast<size_t>(n));
        ctx->span_tab.reserve(2 * static_cast<size_t>(refit_levels(n)) * refit_chunks(n));
        ctx->inode_lo.reserve(n);
        ctx->inode_hi.reserve(n);
        ctx->nrec.reserve(n);
        ctx->leaf_box.reserve(6 * static_cast<size_t>(n));
        if (ctx->cfg.flags & PKG_FLAG_FP32_BROAD) ctx->leaf_f32.reserve(2 * static_cast<size_t>(n));
        launch_body_init(ctx->B(), ctx->S(), st);
        ctx->P[0] = ctx->P[1] = 0;
        ctx->mused[0] = ctx->mused[1] = 0;
        ctx->cur = 0;
        ctx->n_worlds = 1;
        ctx->have_bodies = true;
        ctx->poisoned = false;
        ctx->pos_pending = false;
        ctx->last = pkg_counters{};
        ctx->tree_ok = false;
        ctx->topo_ok = false;
        hip_check(hipStreamSynchronize(st), "body upload");
    })
}

pkg_status pkg_set_worlds(pkg_ctx *ctx, uint32_t n_worlds, const uint32_t *first_body)
{
    if (!ctx || !n_worlds || !first_body || !ctx->have_bodies) return PKG_INVALID_ARG;
    if (n_worlds > 1 && !ctx->jkey_h.empty())
    {
        ctx->err = "joints are supported in single-world contexts only";
        return PKG_INVALID_ARG;
    }
    if (first_body[0] != 0 || first_body[n_worlds] != ctx->n)
    {
        ctx->err = "first_body must start at 0 and end at the body count";
        return PKG_INVALID_ARG;
    }
    for (uint32_t w = 0; w < n_worlds; ++w)
        if (first_body[w + 1] < first_body[w])
        {
            ctx->err = "first_body must be non-decreasing";
            return PKG_INVALID_ARG;
        }
    PKG_GUARD(ctx, {
        hipStream_t st = ctx->st;
        std::vector<uint32_t> wo(ctx->n);
        for (uint32_t w = 0; w < n_worlds; ++w)
            for (uint32_t i = first_body[w]; i < first_body[w + 1]; ++i) wo[i] = w;
        upload(ctx->world_of, wo.data(), wo.size(), st);
        upload(ctx->wstart, first_body, static_cast<size_t>(n_worlds) + 1, st);
        hip_check(hipStreamSynchronize(st), "world upload");
        ctx->n_worlds = n_worlds;
        ctx->tree_ok = false;
        ctx->topo_ok = false;
        ctx->P[0] = ctx->P[1] = 0; // pairs restart under the new partition
        ctx->mused[0] = ctx->mused[1] = 0;
        ctx->cur = 0;
    })
}

pkg_status pkg_add_joints(pkg_ctx *ctx, const pkg_joint_desc *descs, size_t n)
{
    if (!ctx || (n && !descs)) return PKG_INVALID_ARG;
    if (!n) return PKG_OK;
    if (!ctx->have_bodies)
    {
        ctx->err = "joints need uploaded bodies";
        return PKG_INVALID_ARG;
    }
    if (ctx->n_worlds > 1)
    {
        ctx->err = "joints are supported in single-world contexts only";
        return PKG_INVALID_ARG;
    }
    std::vector<std::pair<uint32_t, uint32_t>> add; // (key, desc index)
    for (size_t i = 0; i < n; ++i)
    {
        const pkg_joint_desc &d = descs[i];
        if (d.type < PKG_JOINT_DISTANCE || d.type > PKG_JOINT_SPRING || d.slot >= (1u << 24))
        {
            ctx->err = "joint type or slot out of range";
            return PKG_INVALID_ARG;
        }
        if (d.a >= ctx->n || d.b >= ctx->n || d.a == d.b || ctx->h_static[d.a] == 2 || ctx->h_static[d.b] == 2)
        {
            ctx->err = "joint bodies must be two different live bodies";
            return PKG_INVALID_ARG;
        }
        add.emplace_back(PKG_JOINT_KEY(d.type, d.slot), static_cast<uint32_t>(i));
    }
    std::sort(add.begin(), add.end());
    for (size_t i = 0; i < add.size(); ++i)
        if ((i && add[i].first == add[i - 1].first) ||
            std::binary_search(ctx->jkey_h.begin(), ctx->jkey_h.end(), add[i].first))
        {
            ctx->err = "joint key already live or repeated";
            return PKG_INVALID_ARG;
        }
    PKG_GUARD(ctx, {
        std::vector<uint32_t> keys, ab, src, dint;
        std::vector<double> desc;
        size_t io = 0, ia = 0;
        const size_t no = ctx->jkey_h.size();
        while (io < no || ia < add.size())
        {
            if (ia == add.size() || (io < no && ctx->jkey_h[io] < add[ia].first))
            {
                keys.push_back(ctx->jkey_h[io]);
                ab.push_back(ctx->jab_h[2 * io]);
                ab.push_back(ctx->jab_h[2 * io + 1]);
                src.push_back(static_cast<uint32_t>(io));
                ++io;
            }
            else
            {
                const pkg_joint_desc &d = descs[add[ia].second];
                keys.push_back(add[ia].first);
                ab.push_back(d.a);
                ab.push_back(d.b);
                src.push_back(0x80000000u | static_cast<uint32_t>(dint.size() / 3));
                dint.push_back(static_cast<uint32_t>(d.type));
                dint.push_back(d.a);
                dint.push_back(d.b);
                for (const double *v : {d.local_a, d.local_b, d.local_axis_a, d.local_axis_b, d.anchor, d.axis})
                    desc.insert(desc.end(), v, v + 3);
                desc.push_back(d.distance);
                desc.push_back(d.stiffness);
                desc.push_back(d.damping);
                ++ia;
            }
        }
        ctx->joints_rebuild(keys, ab, src, desc, dint);
    })
}

pkg_status pkg_remove_joints(pkg_ctx *ctx, const uint32_t *keys_in, size_t n)
{
    if (!ctx || (n && !keys_in)) return PKG_INVALID_ARG;
    if (!n) return PKG_OK;
    std::vector<uint32_t> del(keys_in, keys_in + n);
    std::sort(del.begin(), del.end());
    for (size_t i = 0; i < del.size(); ++i)
        if ((i && del[i] == del[i - 1]) || !std::binary_search(ctx->jkey_h.begin(), ctx->jkey_h.end(), del[i]))
        {
            ctx->err = "joint key not live (or repeated)";
            return PKG_INVALID_ARG;
        }
    PKG_GUARD(ctx, {
        std::vector<uint32_t> keys, ab, src;
        for (size_t i = 0; i < ctx->jkey_h.size(); ++i)
        {
            if (std::binary_search(del.begin(), del.end(), ctx->jkey_h[i])) continue;
            keys.push_back(ctx->jkey_h[i]);
            ab.push_back(ctx->jab_h[2 * i]);
            ab.push_back(ctx->jab_h[2 * i + 1]);
            src.push_back(static_cast<uint32_t>(i));
        }
        ctx->joints_rebuild(keys, ab, src, {}, {});
    })
}

pkg_status pkg_get_joints(pkg_ctx *ctx, uint32_t *keys, double *impulses, size_t cap, size_t *n)
{
    if (!ctx || !n) return PKG_INVALID_ARG;
    PKG_GUARD(ctx, {
        const size_t J = ctx->jkey_h.size();
        *n = J;
        const size_t m = std::min(cap, J);
        if (keys)
            for (size_t i = 0; i < m; ++i) keys[i] = ctx->jkey_h[i];
        if (impulses && m)
        {
            const int k = ctx->jset;
            const size_t jc = ctx->jcap[k];
            std::vector<double> acc(6 * jc);
            hip_check(hipMemcpyAsync(acc.data(), ctx->jrec[k].p + 25 * jc, 6 * jc * sizeof(double),
                                     hipMemcpyDeviceToHost, ctx->st),
                      "joint impulses");
            hip_check(hipStreamSynchronize(ctx->st), "stream sync");
            for (size_t i = 0; i < m; ++i)
                for (size_t r = 0; r < 6; ++r) impulses[6 * i + r] = acc[r * jc + i];
        }
    })
}

namespace
{
pkg_status check_query_worlds(pkg_ctx *ctx, uint32_t nq, const uint32_t *w)
{
    if (w)
        for (uint32_t i = 0; i < nq; ++i)
            if (w[i] >= ctx->n_worlds)
            {
                ctx->err = "query world out of range";
                return PKG_INVALID_ARG;
            }
    return PKG_OK;
}

// counts of a query batch summed in 64 bits when 32-bit offsets could overflow
void check_hit_total(pkg_ctx *ctx, uint32_t nq)
{
    if (static_cast<uint64_t>(nq) * ctx->n <= 0xFFFFFFFFull) return;
    std::vector<uint32_t> c = download(ctx->q_cnt, nq, ctx->st);
    uint64_t sum = 0;
    for (uint32_t x : c) sum += x;
    if (sum > 0xFFFFFFFFull) throw pk_error(PKG_OVERFLOW_PAIRS, "query hit count exceeds 32-bit indexing");
}

void check_tree_walk(pkg_ctx *ctx)
{
    if (ctx->read_counter(C_TREE_FAIL)) throw pk_error(PKG_DEVICE_ERROR, "query tree walk exceeded its node count");
}
} // namespace

pkg_status pkg_raycast(pkg_ctx *ctx, uint32_t n_rays, const double *origins, const double *dirs,
                       const double *max_dist, const uint32_t *ray_world, uint32_t *offsets, uint32_t *ids,
                       double *dists, size_t cap, size_t *total)
{
    if (!ctx || !total || !offsets || (n_rays && (!origins || !dirs || !max_dist))) return PKG_INVALID_ARG;
    if (!ctx->have_bodies) return PKG_INVALID_ARG;
    if (pkg_status e = check_query_worlds(ctx, n_rays, ray_world)) return e;
    PKG_GUARD(ctx, {
        hipStream_t st = ctx->st;
        *total = 0;
        offsets[0] = 0;
        if (!n_rays) return PKG_OK;
        upload(ctx->q_org, origins, 3 * static_cast<size_t>(n_rays), st);
        upload(ctx->q_dir, dirs, 3 * static_cast<size_t>(n_rays), st);
        upload(ctx->q_maxd, max_dist, n_rays, st);
        if (ray_world) upload(ctx->q_world, ray_world, n_rays, st);
        const uint32_t *rw = ray_world ? ctx->q_world.p : nullptr;
        lbvh_dev t = ctx->query_tree();
        bodies_dev b = ctx->B();
        unsigned long long *fail = ctx->counters.p + C_TREE_FAIL;
        ctx->reset_counter(C_TREE_FAIL);
        ctx->q_cnt.reserve(n_rays);
        launch_ray_count(t, b, n_rays, ctx->q_org.p, ctx->q_dir.p, ctx->q_maxd.p, rw, ctx->q_cnt.p, fail, st);
        check_hit_total(ctx, n_rays);
        const uint32_t m = ctx->scan_counts(n_rays);
        check_tree_walk(ctx);
        hip_check(hipMemcpyAsync(offsets, ctx->q_off.p, (static_cast<size_t>(n_rays) + 1) * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, st),
                  "offsets");
        hip_check(hipStreamSynchronize(st), "stream sync");
        *total = m;
        if (!m || cap < m) return PKG_OK;
        if (!ids || !dists) throw pk_error(PKG_INVALID_ARG, "ids and dists are needed when cap >= total");
        query_scratch qs = ctx->qscratch(m);
        launch_ray_fill(t, b, n_rays, ctx->q_org.p, ctx->q_dir.p, ctx->q_maxd.p, rw, ctx->q_off.p, qs.hid, qs.hd,
                        qs.hq, fail, st);
        ctx->q_out_id.reserve(m);
        ctx->q_out_d.reserve(m);
        launch_hits_sort(qs, m, ctx->n, n_rays, true, ctx->q_out_id.p, ctx->q_out_d.p, ctx->sc, st);
        hip_check(hipMemcpyAsync(ids, ctx->q_out_id.p, m * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "ids");
        hip_check(hipMemcpyAsync(dists, ctx->q_out_d.p, m * sizeof(double), hipMemcpyDeviceToHost, st), "dists");
        hip_check(hipStreamSynchronize(st), "stream sync");
        check_tree_walk(ctx);
    })
}

pkg_status pkg_raycast_closest_device(pkg_ctx *ctx, uint32_t n_rays, const double *origins, const double *dirs,
                                      const double *max_dist, const uint32_t *ray_world, uint32_t *ids,
                                      double *dists, double *kernel_ms)
{
    if (!ctx || (n_rays && (!origins || !dirs || !max_dist || !ids || !dists))) return PKG_INVALID_ARG;
    if (!ctx->have_bodies) return PKG_INVALID_ARG;
    PKG_GUARD(ctx, {
        if (kernel_ms) *kernel_ms = 0.0;
        if (!n_rays) return PKG_OK;
        lbvh_dev t = ctx->query_tree();
        ctx->reset_counter(C_TREE_FAIL);
        if (kernel_ms && !ctx->q_ev[0])
            for (auto &e : ctx->q_ev) hip_check(hipEventCreate(&e), "event create");
        if (!ctx->rn_ok)
        {
            ctx->q_rn.reserve(ctx->n);
            launch_ray_nodes(t, ctx->q_rn.p, ctx->st);
            ctx->rn_ok = true;
        }
        if (kernel_ms) hip_check(hipEventRecord(ctx->q_ev[0], ctx->st), "event record");
        launch_ray_closest(t, ctx->q_rn.p, ctx->B(), n_rays, origins, dirs, max_dist, ray_world, ids, dists,
                           ctx->counters.p + C_TREE_FAIL, ctx->st);
        hip_check(hipGetLastError(), "raycast launch");
        if (kernel_ms) hip_check(hipEventRecord(ctx->q_ev[1], ctx->st), "event record");
        check_tree_walk(ctx); // synchronizes the stream
        if (kernel_ms)
        {
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, ctx->q_ev[0], ctx->q_ev[1]), "event time");
            *kernel_ms = ms;
        }
    })
}

pkg_status pkg_raycast_closest(pkg_ctx *ctx, uint32_t n_rays, const double *origins, const double *dirs,
                               const double *max_dist, const uint32_t *ray_world, uint32_t *ids, double *dists)
{
    if (!ctx || (n_rays && (!origins || !dirs || !max_dist || !ids || !dists))) return PKG_INVALID_ARG;
    if (!ctx->have_bodies) return PKG_INVALID_ARG;
    if (pkg_status e = check_query_worlds(ctx, n_rays, ray_world)) return e;
    PKG_GUARD(ctx, {
        if (!n_rays) return PKG_OK;
        hipStream_t st = ctx->st;
        upload(ctx->q_org, origins, 3 * static_cast<size_t>(n_rays), st);
        upload(ctx->q_dir, dirs, 3 * static_cast<size_t>(n_rays), st);
        upload(ctx->q_maxd, max_dist, n_rays, st);
        if (ray_world) upload(ctx->q_world, ray_world, n_rays, st);
        ctx->q_out_id.reserve(n_rays);
        ctx->q_out_d.reserve(n_rays);
        pkg_status e = pkg_raycast_closest_device(ctx, n_rays, ctx->q_org.p, ctx->q_dir.p, ctx->q_maxd.p,
                                                  ray_world ? ctx->q_world.p : nullptr, ctx->q_out_id.p,
                                                  ctx->q_out_d.p, nullptr);
        if (e != PKG_OK) return e;
        hip_check(hipMemcpyAsync(ids, ctx->q_out_id.p, n_rays * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "ids");
        hip_check(hipMemcpyAsync(dists, ctx->q_out_d.p, n_rays * sizeof(double), hipMemcpyDeviceToHost, st), "dists");
        hip_check(hipStreamSynchronize(st), "stream sync");
    })
}

pkg_status pkg_query_aabb(pkg_ctx *ctx, uint32_t n_boxes, const double *boxes, const uint32_t *box_world,
                          uint32_t *offsets, uint32_t *ids, size_t cap, size_t *total)
{
    if (!ctx || !total || !offsets || (n_boxes && !boxes)) return PKG_INVALID_ARG;
    if (!ctx->have_bodies) return PKG_INVALID_ARG;
    if (pkg_status e = check_query_worlds(ctx, n_boxes, box_world)) return e;
    PKG_GUARD(ctx, {
        hipStream_t st = ctx->st;
        *total = 0;
        offsets[0] = 0;
        if (!n_boxes) return PKG_OK;
        upload(ctx->q_box, boxes, 6 * static_cast<size_t>(n_boxes), st);
        if (box_world) upload(ctx->q_world, box_world, n_boxes, st);
        const uint32_t *qw = box_world ? ctx->q_world.p : nullptr;
        lbvh_dev t = ctx->query_tree();
        bodies_dev b = ctx->B();
        unsigned long long *fail = ctx->counters.p + C_TREE_FAIL;
        ctx->reset_counter(C_TREE_FAIL);
        ctx->q_cnt.reserve(n_boxes);
        launch_box_count(t, b, n_boxes, ctx->q_box.p, qw, ctx->q_cnt.p, fail, st);
        check_hit_total(ctx, n_boxes);
        const uint32_t m = ctx->scan_counts(n_boxes);
        check_tree_walk(ctx);
        hip_check(hipMemcpyAsync(offsets, ctx->q_off.p, (static_cast<size_t>(n_boxes) + 1) * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, st),
                  "offsets");
        hip_check(hipStreamSynchronize(st), "stream sync");
        *total = m;
        if (!m || cap < m) return PKG_OK;
        if (!ids) throw pk_error(PKG_INVALID_ARG, "ids are needed when cap >= total");
        query_scratch qs = ctx->qscratch(m);
        launch_box_fill(t, b, n_boxes, ctx->q_box.p, qw, ctx->q_off.p, qs.hid, qs.hq, fail, st);
        ctx->q_out_id.reserve(m);
        launch_hits_sort(qs, m, ctx->n, n_boxes, false, ctx->q_out_id.p, nullptr, ctx->sc, st);
        hip_check(hipMemcpyAsync(ids, ctx->q_out_id.p, m * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "ids");
        hip_check(hipStreamSynchronize(st), "stream sync");
        check_tree_walk(ctx);
    })
}

pkg_status pkg_set_body_state(pkg_ctx *ctx, uint32_t count, const uint32_t *ids, const double *pos,
                              const double *vel, const double *angvel, const double *quat)
{
    if (!ctx || (count && !ids)) return PKG_INVALID_ARG;
    for (uint32_t i = 0; i < count; ++i)
        if (ids[i] >= ctx->n) return PKG_INVALID_ARG;
    PKG_GUARD(ctx, {
        if (!count) return PKG_OK;
        hipStream_t st = ctx->st;
        dbuf<uint32_t> di;
        dbuf<double> dp, dv, da, dq;
        upload(di, ids, count, st);
        if (pos) upload(dp, pos, 3 * static_cast<size_t>(count), st);
        if (vel) upload(dv, vel, 3 * static_cast<size_t>(count), st);
        if (angvel) upload(da, angvel, 3 * static_cast<size_t>(count), st);
        if (quat) upload(dq, quat, 4 * static_cast<size_t>(count), st);
        k_set_state<<<(count + 255) / 256, 256, 0, st>>>(ctx->B(), count, di.p, pos ? dp.p : nullptr,
                                                         vel ? dv.p : nullptr, angvel ? da.p : nullptr,
                                                         quat ? dq.p : nullptr);
        hip_check(hipGetLastError(), "set state launch");
        hip_check(hipStreamSynchronize(st), "set state");
        di.release();
        dp.release();
        dv.release();
        da.release();
        dq.release();
    })
}

namespace
{
// stage k of pkg_step_timed spans events (kFrom[k], kTo[k]) recorded by step_once; event 9
// follows the GJK hit-count read; stage 8 (events 10-11, inside stage 6) is the solve
// iterations alone
constexpr int kStages = 9;
constexpr int kFrom[kStages] = {0, 1, 2, 3, 9, 5, 6, 7, 10};
constexpr int kTo[kStages] = {1, 2, 3, 4, 5, 6, 7, 8, 11};

// a failed step leaves body state part-way through it (velocities may carry its gravity,
// leaf bounds may be updated and their moved flags used up): the ctx refuses further
// steps until bodies are uploaded again
pkg_status refuse_if_poisoned(pkg_ctx *ctx)
{
    if (!ctx->poisoned) return PKG_OK;
    ctx->err = "a previous step failed part-way; upload the bodies again (pkg_upload_bodies) before stepping";
    return PKG_DEVICE_ERROR;
}
} // namespace

pkg_status pkg_step(pkg_ctx *ctx, double dt, int nsteps)
{
    if (!ctx || nsteps < 0 || !(dt >= 0.0)) return PKG_INVALID_ARG;
    if (!ctx->have_bodies) return PKG_INVALID_ARG;
    if (pkg_status p = refuse_if_poisoned(ctx)) return p;
    const pkg_status r = [&]() -> pkg_status
    {
        PKG_GUARD(ctx, {
            for (int i = 0; i < nsteps; ++i) ctx->step_once(dt, nullptr, i + 1 < nsteps);
            hip_check(hipGetLastError(), "step");
        })
    }();
    if (r != PKG_OK) ctx->poisoned = true;
    return r;
}

pkg_status pkg_step_timed(pkg_ctx *ctx, double dt, int nsteps, double *stage_ms, int n_stages)
{
    if (!ctx || nsteps < 0 || !stage_ms || n_stages < 1) return PKG_INVALID_ARG;
    if (!ctx->have_bodies) return PKG_INVALID_ARG;
    if (pkg_status p = refuse_if_poisoned(ctx)) return p;
    const pkg_status r = [&]() -> pkg_status
    {
    PKG_GUARD(ctx, {
        hipEvent_t ev[12];
        for (auto &e : ev) hip_check(hipEventCreate(&e), "event create");
        double acc[kStages] = {};
        for (int i = 0; i < nsteps; ++i)
        {
            ctx->step_once(dt, ev, i + 1 < nsteps);
            hip_check(hipEventSynchronize(ev[8]), "event sync");
            for (int k = 0; k < kStages; ++k)
            {
                float ms = 0;
                hip_check(hipEventElapsedTime(&ms, ev[kFrom[k]], ev[kTo[k]]), "elapsed");
                acc[k] += ms;
            }
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        for (int k = 0; k < n_stages && k < kStages; ++k) stage_ms[k] = acc[k];
    })
    }();
    if (r != PKG_OK) ctx->poisoned = true;
    return r;
}

namespace
{
void tuning_out(const tuning &t, pkg_tuning *o)
{
    o->gjk_lanes = t.gjk_group;
    o->epa_block = t.epa_block;
    o->epa_lds_max = t.epa_lds_max;
    o->tree_refits = t.tree_refits;
}
} // namespace

pkg_status pkg_get_tuning(pkg_ctx *ctx, pkg_tuning *out)
{
    if (!ctx || !out) return PKG_INVALID_ARG;
    *out = pkg_tuning{};
    tuning_out(ctx->tune, out);
    return PKG_OK;
}

pkg_status pkg_set_tuning(pkg_ctx *ctx, const pkg_tuning *t)
{
    if (!ctx || !t) return PKG_INVALID_ARG;
    const int g = t->gjk_lanes, e = t->epa_block;
    if (!(g == 1 || g == 2 || g == 4 || g == 8) || !(e == 64 || e == 128 || e == 256))
    {
        ctx->err = "pkg_set_tuning: gjk_lanes must be 1, 2, 4 or 8 and epa_block 64, 128 or 256";
        return PKG_INVALID_ARG;
    }
    if (t->tree_refits > 64)
    {
        ctx->err = "pkg_set_tuning: tree_refits must be at most 64";
        return PKG_INVALID_ARG;
    }
    ctx->tune.gjk_group = g;
    ctx->tune.epa_block = e;
    ctx->tune.epa_lds_max = t->epa_lds_max;
    ctx->tune.tree_refits = t->tree_refits;
    ctx->tuned = true;
    return PKG_OK;
}

pkg_status pkg_autotune(pkg_ctx *ctx, double dt, int steps, pkg_tuning *out)
{
    if (!ctx || !(dt >= 0.0) || steps < 1) return PKG_INVALID_ARG;
    if (!ctx->have_bodies) return PKG_INVALID_ARG;
    if (pkg_status p = refuse_if_poisoned(ctx)) return p;
    pkg_tuning res{};
    const pkg_status r = [&]() -> pkg_status
    {
        PKG_GUARD(ctx, {
            pkg_ctx::shape_trials tr;
            hip_check(hipEventCreate(&tr.t0), "event create");
            hip_check(hipEventCreate(&tr.t1), "event create");
            struct release
            {
                pkg_ctx *c;
                pkg_ctx::shape_trials &t;
                ~release()
                {
                    c->trials = nullptr;
                    (void)hipEventDestroy(t.t0);
                    (void)hipEventDestroy(t.t1);
                }
            } rel{ctx, tr};
            ctx->trials = &tr;
            for (int k = 0; k < steps; ++k) ctx->step_once(dt, nullptr);
            auto argmin = [](const double *v, int n)
            {
                int b = 0;
                for (int c = 1; c < n; ++c)
                    if (v[c] < v[b]) b = c;
                return b;
            };
            static const int lanes[4] = {1, 2, 4, 8}, blocks[3] = {64, 128, 256};
            static const uint32_t lds_max[3] = {0u, 512u, 4096u};
            tuning &tu = ctx->tune;
            tu.gjk_group = lanes[argmin(tr.gjk, 4)];
            tu.epa_block = blocks[argmin(tr.epa_block, 3)];
            tu.epa_lds_max = lds_max[argmin(tr.epa_lds, 3)];
            for (int c = 0; c < 4; ++c) res.gjk_ms[c] = static_cast<float>(tr.gjk[c] / steps);
            for (int c = 0; c < 3; ++c)
            {
                res.epa_block_ms[c] = static_cast<float>(tr.epa_block[c] / steps);
                res.epa_lds_ms[c] = static_cast<float>(tr.epa_lds[c] / steps);
            }
            ctx->tuned = true;
        })
    }();
    if (r != PKG_OK)
    {
        ctx->poisoned = true;
        return r;
    }
    tuning_out(ctx->tune, &res);
    if (out) *out = res;
    return PKG_OK;
}

pkg_status pkg_download_bodies(pkg_ctx *ctx, const pkg_body_state *out)
{
    if (!ctx || !out) return PKG_INVALID_ARG;
    PKG_GUARD(ctx, {
        uint32_t n = ctx->n;
        hipStream_t st = ctx->st;
        auto get = [&](dbuf<double> &b, int k, double *dst)
        {
            if (!dst) return;
            std::vector<double> h = download(b, static_cast<size_t>(n) * k, st);
            aos(h, n, k, dst);
        };
        get(ctx->pos, 3, out->pos);
        get(ctx->vel, 3, out->vel);
        get(ctx->ang, 3, out->angvel);
        get(ctx->quat, 4, out->quat);
        if (out->fat) // already one record per body
        {
            std::vector<double> h = download(ctx->fat, static_cast<size_t>(n) * 6, st);
            std::memcpy(out->fat, h.data(), h.size() * sizeof(double));
        }
        if (out->moved)
        {
            std::vector<uint8_t> m = download(ctx->moved, n, st);
            std::memcpy(out->moved, m.data(), n);
        }
    })
}

pkg_status pkg_get_pairs(pkg_ctx *ctx, uint64_t *keys, size_t cap, size_t *n)
{
    if (!ctx || !n) return PKG_INVALID_ARG;
    PKG_GUARD(ctx, {
        uint32_t P = ctx->P[ctx->cur];
        *n = P;
        if (keys && cap)
        {
            std::vector<uint64_t> k = download(ctx->keys[ctx->cur], std::min<size_t>(cap, P), ctx->st);
            std::memcpy(keys, k.data(), k.size() * sizeof(uint64_t));
        }
    })
}

pkg_status pkg_get_manifolds(pkg_ctx *ctx, int32_t *counts, double *data, size_t cap, size_t *n)
{
    if (!ctx || !n) return PKG_INVALID_ARG;
    PKG_GUARD(ctx, {
        int c = ctx->cur;
        uint32_t P = ctx->P[c];
        *n = P;
        size_t m = std::min<size_t>(cap, P);
        if (m && (counts || data))
        {
            std::vector<uint8_t> cnt = download(ctx->mcount[c], P, ctx->st);
            std::vector<uint32_t> sl;
            std::vector<double> raw;
            const size_t dcap = ctx->mdcap[c];
            if (data && dcap)
            {
                sl = download(ctx->mslot[c], P, ctx->st);
                raw = download(ctx->mdata[c], dcap * 4 * kMF, ctx->st);
            }
            for (size_t p = 0; p < m; ++p)
            {
                if (counts) counts[p] = cnt[p];
                if (data)
                    for (int k = 0; k < 4; ++k)
                        for (int f = 0; f < kMF; ++f)
                            data[(p * 4 + k) * kMF + f] =
                                k < cnt[p] ? raw[static_cast<size_t>(k * kMF + f) * dcap + sl[p]] : 0.0;
            }
        }
    })
}

pkg_status pkg_get_events(pkg_ctx *ctx, uint64_t *keys, int32_t *kinds, size_t cap, size_t *n)
{
    if (!ctx || !n) return PKG_INVALID_ARG;
    PKG_GUARD(ctx, {
        int c = ctx->cur;
        uint32_t P = ctx->P[c];
        std::vector<uint8_t> ev = download(ctx->events, P, ctx->st);
        std::vector<uint64_t> k = download(ctx->keys[c], P, ctx->st);
        size_t cnt = 0;
        for (uint32_t p = 0; p < P; ++p)
            if (ev[p])
            {
                if (cnt < cap)
                {
                    if (keys) keys[cnt] = k[p];
                    if (kinds) kinds[cnt] = ev[p];
                }
                ++cnt;
            }
        *n = cnt;
    })
}

pkg_status pkg_get_counters(pkg_ctx *ctx, pkg_counters *out)
{
    if (!ctx || !out) return PKG_INVALID_ARG;
    *out = ctx->last;
    return PKG_OK;
}

pkg_status pkg_get_colours(pkg_ctx *ctx, uint32_t *colours, uint64_t *keys, size_t cap, size_t *n)
{
    if (!ctx || !n) return PKG_INVALID_ARG;
    PKG_GUARD(ctx, {
        uint32_t P = ctx->P[ctx->cur];
        size_t cnt = 0;
        if (ctx->n_colours && P && ctx->n_listed)
        {
            // the coloured manifolds are the listed ones, reported in pair order
            std::vector<uint32_t> list = download(ctx->clist, ctx->n_listed, ctx->st);
            std::sort(list.begin(), list.end());
            std::vector<uint32_t> col = download(ctx->colour, P, ctx->st);
            std::vector<uint64_t> k;
            if (keys) k = download(ctx->keys[ctx->cur], P, ctx->st);
            for (uint32_t p : list)
            {
                if (cnt < cap)
                {
                    if (colours) colours[cnt] = col[p];
                    if (keys) keys[cnt] = k[p];
                }
                ++cnt;
            }
        }
        *n = cnt;
    })
}

pkg_status pkg_gjk_epa_batch(pkg_ctx *ctx, uint32_t n, const int32_t *kind_a, const double *prm_a,
                             const uint32_t *vbeg_a, const uint32_t *vend_a, const int32_t *kind_b,
                             const double *prm_b, const uint32_t *vbeg_b, const uint32_t *vend_b,
                             const double *verts, uint32_t n_verts, double *out, int32_t *hit)
{
    if (!ctx || !kind_a || !prm_a || !vbeg_a || !vend_a || !kind_b || !prm_b || !vbeg_b || !vend_b || !out || !hit)
        return PKG_INVALID_ARG;
    for (uint32_t i = 0; i < n; ++i)
    {
        for (int side = 0; side < 2; ++side)
        {
            int k = side ? kind_b[i] : kind_a[i];
            uint32_t vb = side ? vbeg_b[i] : vbeg_a[i], ve = side ? vend_b[i] : vend_a[i];
            if (k < 0 || k > 2) return PKG_INVALID_ARG;
            if (k == 2 && (ve <= vb || ve > n_verts || !verts)) return PKG_INVALID_ARG;
        }
    }
    PKG_GUARD(ctx, {
        if (!n) return PKG_OK;
        hipStream_t st = ctx->st;
        dbuf<int32_t> ka, kb, dh;
        dbuf<double> pa, pb, dv, dout;
        dbuf<uint32_t> ba, ea, bb, eb, ua, ub;
        dbuf<char> pool;
        {
            std::vector<uint32_t> tag_a(n, 0u), tag_b(n, 0u);
            pk::uv_sphere_detector det;
            for (uint32_t i = 0; i < n; ++i)
            {
                if (kind_a[i] == 2) tag_a[i] = det.tag(verts + 3 * static_cast<size_t>(vbeg_a[i]), vend_a[i] - vbeg_a[i]);
                if (kind_b[i] == 2) tag_b[i] = det.tag(verts + 3 * static_cast<size_t>(vbeg_b[i]), vend_b[i] - vbeg_b[i]);
            }
            upload(ua, tag_a.data(), n, st);
            upload(ub, tag_b.data(), n, st);
            hip_check(hipStreamSynchronize(st), "tag upload");
        }
        upload(ka, kind_a, n, st);
        upload(kb, kind_b, n, st);
        upload(pa, prm_a, 10 * static_cast<size_t>(n), st);
        upload(pb, prm_b, 10 * static_cast<size_t>(n), st);
        upload(ba, vbeg_a, n, st);
        upload(ea, vend_a, n, st);
        upload(bb, vbeg_b, n, st);
        upload(eb, vend_b, n, st);
        std::vector<double> vz(3, 0.0);
        if (n_verts && verts)
            upload(dv, verts, 3 * static_cast<size_t>(n_verts), st);
        else
            upload(dv, vz.data(), 3, st);
        dout.reserve(10 * static_cast<size_t>(n));
        dh.reserve(n);
        pool.reserve(epa_big_bytes());
        shape_pair_dev sp;
        sp.kind_a = ka.p;
        sp.kind_b = kb.p;
        sp.prm_a = pa.p;
        sp.prm_b = pb.p;
        sp.vbeg_a = ba.p;
        sp.vend_a = ea.p;
        sp.vbeg_b = bb.p;
        sp.vend_b = eb.p;
        sp.verts = dv.p;
        sp.uv_a = ua.p;
        sp.uv_b = ub.p;
        launch_gjk_epa_batch(sp, n, dout.p, dh.p, pool.p, st);
        hip_check(hipGetLastError(), "batch launch");
        std::vector<double> o = download(dout, 10 * static_cast<size_t>(n), st);
        std::vector<int32_t> h = download(dh, n, st);
        for (uint32_t i = 0; i < n; ++i)
        {
            hit[i] = h[i];
            if (h[i] == 1) std::memcpy(out + 10 * static_cast<size_t>(i), o.data() + 10 * static_cast<size_t>(i), 10 * sizeof(double));
        }
        for (auto *b : {&ka, &kb, &dh}) b->release();
        for (auto *b : {&pa, &pb, &dv, &dout}) b->release();
        for (auto *b : {&ba, &ea, &bb, &eb}) b->release();
        pool.release();
        for (uint32_t i = 0; i < n; ++i)
            if (h[i] == 3) throw pk_error(PKG_OVERFLOW_EPA, "EPA polytope exceeded 4096 faces");
    })
}

pkg_status pkg_gjk_epa_batch_device(pkg_ctx *ctx, uint32_t n, const int32_t *kind_a, const double *prm_a,
                                    const uint32_t *vbeg_a, const uint32_t *vend_a, const int32_t *kind_b,
                                    const double *prm_b, const uint32_t *vbeg_b, const uint32_t *vend_b,
                                    const double *verts, uint32_t n_verts, double *out, int32_t *hit)
{
    if (!ctx) return PKG_INVALID_ARG;
    if (n > PKG_NOALLOC_MAX_PAIRS)
    {
        ctx->err = "pkg_gjk_epa_batch_device: n = " + std::to_string(n) + " exceeds PKG_NOALLOC_MAX_PAIRS (" +
                   std::to_string(PKG_NOALLOC_MAX_PAIRS) + "), the capacity of its static scratch";
        return PKG_INVALID_ARG;
    }
    if (n && (!kind_a || !prm_a || !vbeg_a || !vend_a || !kind_b || !prm_b || !vbeg_b || !vend_b || !out || !hit ||
              (n_verts && !verts)))
    {
        ctx->err = "pkg_gjk_epa_batch_device: null array";
        return PKG_INVALID_ARG;
    }
    PKG_GUARD(ctx, {
        noalloc_layout *sc = ctx->noalloc;
        shape_pair_dev sp;
        sp.kind_a = kind_a;
        sp.kind_b = kind_b;
        sp.prm_a = prm_a;
        sp.prm_b = prm_b;
        sp.vbeg_a = vbeg_a;
        sp.vend_a = vend_a;
        sp.vbeg_b = vbeg_b;
        sp.vend_b = vend_b;
        sp.verts = verts;
        sp.uv_a = nullptr; // no UV-sphere tags: full support scans, the same results
        sp.uv_b = nullptr;
        launch_gjk_epa_batch_dev(sp, n, n_verts, out, hit, sc->words, sc->ovf_list, sc->stores, kNoAllocStores,
                                 ctx->st);
        hip_check(hipGetLastError(), "batch launch");
        // the counters into the ctx's pinned words (allocated by pkg_create)
        uint32_t *h = reinterpret_cast<uint32_t *>(ctx->h_counters + C_NCOUNTERS);
        hip_check(hipMemcpyAsync(h, sc->words, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->st), "counter copy");
        hip_check(hipStreamSynchronize(ctx->st), "stream sync");
        if (h[0])
            throw pk_error(PKG_INVALID_ARG, std::to_string(h[0]) + " pairs with a kind outside 0-2 or a vertex range "
                                                                   "outside [0, n_verts) (their hit is -1)");
        if (h[2]) throw pk_error(PKG_OVERFLOW_EPA, "EPA polytope exceeded 4096 faces");
    })
}

} // extern "C"
