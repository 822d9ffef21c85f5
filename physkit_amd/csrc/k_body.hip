// Per-body kernels: body set-up (particle ctor), velocity integration with the
// broadphase leaf update, and position integration. One thread per body over
// planar FP64 arrays. HBM-bound: each body reads and writes a fixed number of
// doubles with unit-stride, coalesced accesses.
#include "pk_access.hpp"

namespace pk
{
namespace
{
constexpr int kBlock = 256;

__global__ void k_shape_boxes(shapes_dev s) // aabb::from_points, bounds.h:33-49
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n) return;
    uint32_t b = s.voff[i], e = b + (s.off[i + 1] - s.off[i]);
    v3 p0 = mk(s.verts[3 * b], s.verts[3 * b + 1], s.verts[3 * b + 2]);
    box3 r;
    r.min = p0;
    r.max = p0;
    for (uint32_t k = b + 1; k < e; ++k)
    {
        v3 p = mk(s.verts[3 * k], s.verts[3 * k + 1], s.verts[3 * k + 2]);
        r.min.x = smin(r.min.x, p.x);
        r.min.y = smin(r.min.y, p.y);
        r.min.z = smin(r.min.z, p.z);
        r.max.x = smax(r.max.x, p.x);
        r.max.y = smax(r.max.y, p.y);
        r.max.z = smax(r.max.z, p.z);
    }
    stbox(s.local_box, s.n, i, r);
}

__device__ void derived(const bodies_dev &b, uint32_t i, qt q) // particle.h:140-146
{
    m3 R = rotation_matrix(q);
    m3 Rt = transpose(R);
    st9(b.Iw, b.n, i, mul(mul(R, ld9(b.I_local, b.n, i)), Rt));
    st9(b.Iw_inv, b.n, i, mul(mul(R, ld9(b.Iinv_local, b.n, i)), Rt));
}

__global__ void k_body_init(bodies_dev b, shapes_dev s) // particle.h:16-28, bvh.h:294-302
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.n) return;
    double im = 1.0 / b.mass[i];
    b.inv_mass[i] = im;
    m3 Ii = inverse(ld9(b.I_local, b.n, i));
    if (im == 0.0)
        for (double &e : Ii.m) e = 0.0;
    st9(b.Iinv_local, b.n, i, Ii);
    qt q = ldq(b.quat, b.n, i);
    derived(b, i, q);
    uint32_t sh = b.shape[i];
    box3 fat = instance_bounds(ldbox(s.local_box, s.n, sh), q, ld3(b.pos, b.n, i));
    if (b.is_static[i] == 2) // a disabled (freed) slot: NaN bounds overlap nothing
    {
        double nan = __builtin_nan("");
        fat.min = mk(nan, nan, nan);
        fat.max = fat.min;
    }
    st_fat(b.fat, i, fat);
    st3(b.acc, b.n, i, mk(0.0, 0.0, 0.0));
    st3(b.torque, b.n, i, mk(0.0, 0.0, 0.0));
    b.moved[i] = 0;
}

// a force or torque accumulator cleared (clear_forces, particle.h:101-105); a store
// only where a bit is set, so untouched accumulators cost no write
__device__ __forceinline__ void clear3(double *p, uint32_t n, uint32_t i, v3 v)
{
    if (__double_as_longlong(v.x) | __double_as_longlong(v.y) | __double_as_longlong(v.z))
        st3(p, n, i, mk(0.0, 0.0, 0.0));
}

// src/world.cpp:22-34 for dynamic body i: gravity force, integrate_vel (integrate.h:36-40,
// particle.h:72-76), world bounds, dynamic_bvh::update_leaf (bvh.cpp:475-514),
// clear_forces; from its velocities, pose and world inertia (read or just computed).
// Returns whether the leaf moved.
__device__ __forceinline__ bool integrate_vel_body(const bodies_dev &b, const shapes_dev &s, uint32_t i, v3 g,
                                                   double dt, v3 vel, v3 w, qt q, v3 pos, const m3 &Iw, const m3 &Iwi)
{
    bool moved = false;
    {
        double m = b.mass[i], im = b.inv_mass[i];
        v3 force = g * m;
        const v3 acc_in = ld3(b.acc, b.n, i), torque_in = ld3(b.torque, b.n, i);
        v3 acc = acc_in + force * im;
        vel = vel + acc * dt;
        v3 aa = mul(Iwi, torque_in - cross(w, mul(Iw, w)));
        w = w + aa * dt;
        st3(b.vel, b.n, i, vel);
        st3(b.ang, b.n, i, w);
        box3 tb = instance_bounds(ldbox(s.local_box, s.n, b.shape[i]), q, pos);
        box3 fat = ld_fat(b.fat, i);
        if (!box_contains(fat, tb))
        {
            const double margin = 0.1;
            v3 disp = vel * dt;
            box3 nb;
            nb.min = tb.min - mk(margin, margin, margin);
            nb.max = tb.max + mk(margin, margin, margin);
            if (disp.x < 0.0) nb.min.x = nb.min.x + disp.x; else nb.max.x = nb.max.x + disp.x;
            if (disp.y < 0.0) nb.min.y = nb.min.y + disp.y; else nb.max.y = nb.max.y + disp.y;
            if (disp.z < 0.0) nb.min.z = nb.min.z + disp.z; else nb.max.z = nb.max.z + disp.z;
            st_fat(b.fat, i, nb);
            moved = true;
        }
        clear3(b.acc, b.n, i, acc_in);
        clear3(b.torque, b.n, i, torque_in);
    }
    return moved;
}

__global__ void k_integrate_vel(bodies_dev b, shapes_dev s, double gx, double gy, double gz,
                                double dt, unsigned long long *counters)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool moved = false;
    if (i < b.n && !b.is_static[i])
        moved = integrate_vel_body(b, s, i, mk(gx, gy, gz), dt, ld3(b.vel, b.n, i), ld3(b.ang, b.n, i),
                                   ldq(b.quat, b.n, i), ld3(b.pos, b.n, i), ld9(b.Iw, b.n, i), ld9(b.Iw_inv, b.n, i));
    if (i < b.n) b.moved[i] = moved ? 1 : 0;
    block_add(&counters[C_MOVED], moved ? 1u : 0u);
}

// src/world.cpp:50-55, integrate.h:42-46, particle.h:40-44
__global__ void k_integrate_pos(bodies_dev b, double dt)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.n || b.is_static[i]) return;
    st3(b.pos, b.n, i, ld3(b.pos, b.n, i) + ld3(b.vel, b.n, i) * dt);
    qt q = qmul(exp_map(ld3(b.ang, b.n, i), dt), ldq(b.quat, b.n, i));
    stq(b.quat, b.n, i, q);
    derived(b, i, q);
}

// One step's integrate_pos (dt_prev) and the next step's integrate_vel (dt) in one pass
// over the bodies, for consecutive steps with nothing between them: the same arithmetic
// as k_integrate_pos then k_integrate_vel, with the new pose, velocities and world
// inertia kept in registers instead of written and read back.
__global__ void k_integrate_pos_vel(bodies_dev b, shapes_dev s, double dt_prev, double gx, double gy, double gz,
                                    double dt, unsigned long long *counters)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool moved = false;
    if (i < b.n && !b.is_static[i])
    {
        const v3 vel = ld3(b.vel, b.n, i), w = ld3(b.ang, b.n, i);
        const v3 pos = ld3(b.pos, b.n, i) + vel * dt_prev;
        st3(b.pos, b.n, i, pos);
        const qt q = qmul(exp_map(w, dt_prev), ldq(b.quat, b.n, i));
        stq(b.quat, b.n, i, q);
        // derived(b, i, q), kept
        const m3 R = rotation_matrix(q);
        const m3 Rt = transpose(R);
        const m3 Iw = mul(mul(R, ld9(b.I_local, b.n, i)), Rt);
        const m3 Iwi = mul(mul(R, ld9(b.Iinv_local, b.n, i)), Rt);
        st9(b.Iw, b.n, i, Iw);
        st9(b.Iw_inv, b.n, i, Iwi);
        moved = integrate_vel_body(b, s, i, mk(gx, gy, gz), dt, vel, w, q, pos, Iw, Iwi);
    }
    if (i < b.n) b.moved[i] = moved ? 1 : 0;
    block_add(&counters[C_MOVED], moved ? 1u : 0u);
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, size_t n)
{
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n)
{
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

unsigned grid(size_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }
} // namespace

void launch_shape_boxes(const shapes_dev &s, hipStream_t st)
{
    if (s.n) k_shape_boxes<<<grid(s.n), kBlock, 0, st>>>(s);
}
void launch_body_init(const bodies_dev &b, const shapes_dev &s, hipStream_t st)
{
    if (b.n) k_body_init<<<grid(b.n), kBlock, 0, st>>>(b, s);
}
void launch_integrate_vel(const bodies_dev &b, const shapes_dev &s, double gx, double gy, double gz,
                          double dt, unsigned long long *counters, hipStream_t st)
{
    if (b.n) k_integrate_vel<<<grid(b.n), kBlock, 0, st>>>(b, s, gx, gy, gz, dt, counters);
}
void launch_integrate_pos_vel(const bodies_dev &b, const shapes_dev &s, double dt_prev, double gx, double gy,
                              double gz, double dt, unsigned long long *counters, hipStream_t st)
{
    if (b.n) k_integrate_pos_vel<<<grid(b.n), kBlock, 0, st>>>(b, s, dt_prev, gx, gy, gz, dt, counters);
}
void launch_integrate_pos(const bodies_dev &b, double dt, hipStream_t st)
{
    if (b.n) k_integrate_pos<<<grid(b.n), kBlock, 0, st>>>(b, dt);
}
void launch_fill_u64(unsigned long long *p, unsigned long long v, size_t n, hipStream_t st)
{
    if (n) k_fill_u64<<<grid(n), kBlock, 0, st>>>(p, v, n);
}
void launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t st)
{
    if (n) k_fill_u32<<<grid(n), kBlock, 0, st>>>(p, v, n);
}

} // namespace pk
