// Joint constraints: the constructors and build_jacobian_impl of distance, spring,
// ball_socket, hinge, slider and weld (include/physkit/collision/constraint.h:139-834),
// and their rows in constraint_solver::solve_constraints (constraint.h:1107-1201):
// warm started and iterated before the contact rows, accumulated impulses written
// back for the next step's warm start. Joints are coloured like manifolds (no two
// joints of a colour share a dynamic body), so one thread per joint of a colour
// reproduces a sequential sweep in colour order. A joint's rows stay in one thread,
// in the reference's row order, because they share the joint's two bodies.
#include "pk_access.hpp"

#include <cfloat>

namespace pk
{
namespace
{
constexpr int kBlock = 256;
unsigned grid(size_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

enum : int
{
    J_DISTANCE = 0,
    J_BALL_SOCKET = 1,
    J_HINGE = 2,
    J_SLIDER = 3,
    J_WELD = 4,
    J_SPRING = 5,
};

// "Building an Orthonormal Basis, Revisited" (Duff et al. 2017), constraint.h:104-113
__device__ __forceinline__ void onb(v3 n, v3 &t1, v3 &t2)
{
    double sign = copysign(1.0, n.z);
    double a = -1.0 / (sign + n.z);
    double b = n.x * n.y * a;
    t1 = mk(1.0 + sign * n.x * n.x * a, sign * b, -sign * n.x);
    t2 = mk(b, sign + n.y * n.y * a, -n.y);
}

__device__ __forceinline__ v3 rd3(const double *rec, uint32_t cap, int w, uint32_t j)
{ return mk(rec[static_cast<size_t>(w) * cap + j], rec[static_cast<size_t>(w + 1) * cap + j],
            rec[static_cast<size_t>(w + 2) * cap + j]); }
__device__ __forceinline__ void wr3(double *rec, uint32_t cap, int w, uint32_t j, v3 v)
{
    rec[static_cast<size_t>(w) * cap + j] = v.x;
    rec[static_cast<size_t>(w + 1) * cap + j] = v.y;
    rec[static_cast<size_t>(w + 2) * cap + j] = v.z;
}

__global__ void __launch_bounds__(kBlock) k_joint_rebuild(joints_dev o, joints_dev nj, const uint32_t *src, const double *desc,
                                const uint32_t *dint, bodies_dev b)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nj.n) return;
    uint32_t s = src[j];
    if ((s & 0x80000000u) == 0)
    {
        for (int w = 0; w < kJW; ++w) nj.rec[static_cast<size_t>(w) * nj.cap + j] = o.rec[static_cast<size_t>(w) * o.cap + s];
        nj.type[j] = o.type[s];
        nj.a[j] = o.a[s];
        nj.b[j] = o.b[s];
        nj.keys[j] = o.keys[s];
        return;
    }
    s &= 0x7FFFFFFFu;
    const double *d = desc + 21 * static_cast<size_t>(s);
    const int type = static_cast<int>(dint[3 * s]);
    const uint32_t ia = dint[3 * s + 1], ib = dint[3 * s + 2];
    for (int w = 0; w < kJW; ++w) nj.rec[static_cast<size_t>(w) * nj.cap + j] = 0.0;
    const v3 local_a = mk(d[0], d[1], d[2]), local_b = mk(d[3], d[4], d[5]);
    const v3 axis_a = mk(d[6], d[7], d[8]), axis_b = mk(d[9], d[10], d[11]);
    const v3 anchor = mk(d[12], d[13], d[14]), axis = mk(d[15], d[16], d[17]);
    v3 la = mk(0.0, 0.0, 0.0), lb = la, ax_a = la, ax_b = la, f1 = la, f2 = la;
    qt qrel = mkq(1.0, 0.0, 0.0, 0.0);
    if (type == J_DISTANCE || type == J_SPRING) // constraint.h:180-184, 287-292
    {
        la = local_a;
        lb = local_b;
    }
    else if (type == J_HINGE) // constraint.h:502-507
    {
        la = local_a;
        lb = local_b;
        ax_a = normalized(axis_a);
        ax_b = normalized(axis_b);
        onb(ax_a, f1, f2);
    }
    else // ball_socket :377-382, slider :612-628, weld :757-762: particle::project_to_local
    {
        const v3 pa = ld3(b.pos, b.n, ia), pb = ld3(b.pos, b.n, ib);
        const qt qa = ldq(b.quat, b.n, ia), qb = ldq(b.quat, b.n, ib);
        la = rotate(conj(qa), anchor - pa);
        lb = rotate(conj(qb), anchor - pb);
        if (type != J_BALL_SOCKET) qrel = qmul(conj(qa), qb);
        if (type == J_SLIDER)
        {
            v3 axn = normalized(axis);
            ax_a = rotate(conj(qa), axn);
            v3 w1, w2;
            onb(axn, w1, w2);
            f1 = rotate(conj(qa), w1);
            f2 = rotate(conj(qa), w2);
        }
    }
    wr3(nj.rec, nj.cap, 0, j, la);
    wr3(nj.rec, nj.cap, 3, j, lb);
    wr3(nj.rec, nj.cap, 6, j, ax_a);
    wr3(nj.rec, nj.cap, 9, j, ax_b);
    wr3(nj.rec, nj.cap, 12, j, f1);
    wr3(nj.rec, nj.cap, 15, j, f2);
    nj.rec[18 * static_cast<size_t>(nj.cap) + j] = qrel.w;
    nj.rec[19 * static_cast<size_t>(nj.cap) + j] = qrel.x;
    nj.rec[20 * static_cast<size_t>(nj.cap) + j] = qrel.y;
    nj.rec[21 * static_cast<size_t>(nj.cap) + j] = qrel.z;
    nj.rec[22 * static_cast<size_t>(nj.cap) + j] = d[18];
    nj.rec[23 * static_cast<size_t>(nj.cap) + j] = d[19];
    nj.rec[24 * static_cast<size_t>(nj.cap) + j] = d[20];
    nj.type[j] = static_cast<uint8_t>(type);
    nj.a[j] = ia;
    nj.b[j] = ib;
    nj.keys[j] = (static_cast<uint64_t>(ia) << 32) | ib;
}

struct jbody
{
    v3 pos;
    qt q;
    double im;
    m3 Iwi;
};
__device__ __forceinline__ jbody ld_jbody(const bodies_dev &b, uint32_t i)
{
    jbody s;
    s.pos = ld3(b.pos, b.n, i);
    s.q = ldq(b.quat, b.n, i);
    s.im = b.inv_mass[i];
    s.Iwi = ld9(b.Iw_inv, b.n, i);
    return s;
}

struct jrow
{
    v3 Jv, Jwa, Jwb;
    double Meff, bias, gamma, acc;
};

__device__ __forceinline__ void put_row(const joints_dev &J, uint32_t j, int r, const jrow &w)
{
    double *p = J.rows + static_cast<size_t>(r) * kJR * J.cap;
    wr3(p, J.cap, 0, j, w.Jv);
    wr3(p, J.cap, 3, j, w.Jwa);
    wr3(p, J.cap, 6, j, w.Jwb);
    p[9 * static_cast<size_t>(J.cap) + j] = w.Meff;
    p[10 * static_cast<size_t>(J.cap) + j] = w.bias;
    p[11 * static_cast<size_t>(J.cap) + j] = w.gamma;
    p[12 * static_cast<size_t>(J.cap) + j] = w.acc;
}

__device__ __forceinline__ double acc_of(const joints_dev &J, uint32_t j, int r)
{ return J.rec[static_cast<size_t>(25 + r) * J.cap + j]; }

// J = [n, ra x n, -n, -(rb x n)]
__device__ __forceinline__ jrow linear_row(v3 n, v3 ra, v3 rb)
{
    jrow r;
    r.Jv = n;
    r.Jwa = cross(ra, n);
    r.Jwb = -cross(rb, n);
    r.Meff = r.bias = r.gamma = r.acc = 0.0;
    return r;
}
// J = [0, n, 0, -n] (angular_scale 1 m)
__device__ __forceinline__ jrow angular_row(v3 n, const jbody &A, const jbody &B)
{
    jrow r;
    v3 jt = n * 1.0;
    r.Jv = mk(0.0, 0.0, 0.0);
    r.Jwa = jt;
    r.Jwb = -jt;
    r.Meff = 1.0 / (dot(r.Jwa, mul(A.Iwi, r.Jwa)) + dot(r.Jwb, mul(B.Iwi, r.Jwb)));
    r.bias = r.gamma = r.acc = 0.0;
    return r;
}
// orientation error of slider / weld (constraint.h:689-697, 812-820), in world frame
__device__ __forceinline__ v3 orientation_error(const jbody &A, const jbody &B, qt qrel)
{
    qt cur = qmul(conj(A.q), B.q);
    qt oe = qmul(qrel, conj(cur));
    if (oe.w < 0) oe = mkq(-oe.w, -oe.x, -oe.y, -oe.z);
    return rotate(A.q, mk(oe.x, oe.y, oe.z) * 2.0);
}

// build_jacobian_impl of each type (constraint.h:186-228, 291-335, 386-425, 509-573,
// 630-713, 763-834), from the poses and world inverse inertia at the solve
__global__ void __launch_bounds__(kBlock) k_joint_rows(joints_dev J, bodies_dev b, double dt)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= J.n) return;
    const int type = J.type[j];
    const jbody A = ld_jbody(b, J.a[j]), B = ld_jbody(b, J.b[j]);
    const v3 ra = rotate(A.q, rd3(J.rec, J.cap, 0, j));
    const v3 rb = rotate(B.q, rd3(J.rec, J.cap, 3, j));
    const v3 ex = mk(1.0, 0.0, 0.0), ey = mk(0.0, 1.0, 0.0), ez = mk(0.0, 0.0, 1.0);
    int k = 0;
    if (type == J_DISTANCE || type == J_SPRING)
    {
        v3 pa = A.pos + ra, pb = B.pos + rb;
        v3 delta = pa - pb;
        double cur = norm(delta);
        if (!(cur < 1e-6))
        {
            const double dist = J.rec[22 * static_cast<size_t>(J.cap) + j];
            v3 u = delta / cur;
            jrow r = linear_row(u, ra, rb);
            double mc = ((A.im + B.im) + dot(r.Jwa, mul(A.Iwi, r.Jwa))) + dot(r.Jwb, mul(B.Iwi, r.Jwb));
            if (type == J_DISTANCE)
            {
                r.Meff = 1.0 / mc;
                r.bias = (0.1 / dt) * (cur - dist);
            }
            else
            {
                const double ks = J.rec[23 * static_cast<size_t>(J.cap) + j];
                const double kd = J.rec[24 * static_cast<size_t>(J.cap) + j];
                double gamma = 1.0 / (dt * (kd + dt * ks));
                double beta = (dt * ks) / (kd + dt * ks);
                r.Meff = 1.0 / (mc + gamma);
                r.bias = (beta / dt) * (cur - dist);
                r.gamma = gamma;
            }
            r.acc = acc_of(J, j, 0);
            put_row(J, j, k++, r);
        }
    }
    else if (type == J_SLIDER)
    {
        v3 delta = (A.pos + ra) - (B.pos + rb);
        v3 axis_a = rotate(A.q, rd3(J.rec, J.cap, 6, j));
        v3 t1 = rotate(A.q, rd3(J.rec, J.cap, 12, j)), t2 = rotate(A.q, rd3(J.rec, J.cap, 15, j));
        v3 lever = ra - delta;
        for (int s = 0; s < 2; ++s)
        {
            v3 n = s == 0 ? t1 : t2;
            jrow r;
            r.Jv = n;
            r.Jwa = cross(lever, n);
            r.Jwb = -cross(rb, n);
            double lin = A.im + B.im;
            double ang = dot(r.Jwa, mul(A.Iwi, r.Jwa)) + dot(r.Jwb, mul(B.Iwi, r.Jwb));
            double reg = 0.01 * ang;
            r.gamma = reg;
            r.Meff = 1.0 / ((lin + ang) + reg);
            double e = dot(delta, n);
            e = (e < -0.2) ? -0.2 : ((0.2 < e) ? 0.2 : e); // std::clamp
            r.bias = (0.1 / dt) * e;
            r.acc = acc_of(J, j, k);
            put_row(J, j, k++, r);
        }
        const qt qrel = mkq(J.rec[18 * static_cast<size_t>(J.cap) + j], J.rec[19 * static_cast<size_t>(J.cap) + j],
                            J.rec[20 * static_cast<size_t>(J.cap) + j], J.rec[21 * static_cast<size_t>(J.cap) + j]);
        v3 aerr = orientation_error(A, B, qrel);
        for (int s = 0; s < 3; ++s)
        {
            v3 n = s == 0 ? axis_a : (s == 1 ? t1 : t2);
            jrow r = angular_row(n, A, B);
            r.bias = (0.1 / dt) * (dot(aerr, n) * 1.0);
            r.acc = acc_of(J, j, k);
            put_row(J, j, k++, r);
        }
    }
    else // ball_socket, hinge, weld: three positional rows first
    {
        const double bf = type == J_HINGE ? 0.2 : 0.1;
        v3 delta = (A.pos + ra) - (B.pos + rb);
        for (int s = 0; s < 3; ++s)
        {
            v3 n = s == 0 ? ex : (s == 1 ? ey : ez);
            jrow r = linear_row(n, ra, rb);
            r.Meff = 1.0 / (((A.im + B.im) + dot(r.Jwa, mul(A.Iwi, r.Jwa))) + dot(r.Jwb, mul(B.Iwi, r.Jwb)));
            r.bias = (bf / dt) * dot(delta, n);
            r.acc = acc_of(J, j, k);
            put_row(J, j, k++, r);
        }
        if (type == J_HINGE)
        {
            v3 ua = rotate(A.q, rd3(J.rec, J.cap, 6, j)), ub = rotate(B.q, rd3(J.rec, J.cap, 9, j));
            v3 aerr = -cross(ua, ub);
            v3 va = rotate(A.q, rd3(J.rec, J.cap, 12, j)), wa = rotate(A.q, rd3(J.rec, J.cap, 15, j));
            for (int s = 0; s < 2; ++s)
            {
                v3 n = s == 0 ? va : wa;
                jrow r = angular_row(n, A, B);
                r.bias = (bf / dt) * (dot(aerr, n) * 1.0);
                r.acc = acc_of(J, j, k);
                put_row(J, j, k++, r);
            }
        }
        else if (type == J_WELD)
        {
            const qt qrel = mkq(J.rec[18 * static_cast<size_t>(J.cap) + j], J.rec[19 * static_cast<size_t>(J.cap) + j],
                                J.rec[20 * static_cast<size_t>(J.cap) + j], J.rec[21 * static_cast<size_t>(J.cap) + j]);
            v3 aerr = orientation_error(A, B, qrel);
            for (int s = 0; s < 3; ++s)
            {
                v3 n = s == 0 ? ex : (s == 1 ? ey : ez);
                jrow r = angular_row(n, A, B);
                r.bias = ((bf / dt) * dot(aerr, n)) * 1.0;
                r.acc = acc_of(J, j, k);
                put_row(J, j, k++, r);
            }
        }
    }
    J.nrows[j] = static_cast<uint8_t>(k);
}

struct jvel
{
    v3 vel, ang;
    double im;
    m3 Iwi;
    bool stat;
};

// jacobian_row::apply (constraint.h:49-72) with CFM gamma and the default bounds
// -max .. max of a joint row
__device__ __forceinline__ void joint_apply(v3 Jv, v3 Jwa, v3 Jwb, double Meff, double bias, double gamma,
                                            double &acc, jvel &A, jvel &B)
{
    double jv = ((dot(Jv, A.vel) + dot(Jwa, A.ang)) - dot(Jv, B.vel)) + dot(Jwb, B.ang);
    double cfp = gamma * acc;
    double lambda = -((jv + bias) + cfp) * Meff;
    double old = acc;
    double v = old + lambda;
    acc = (v < -DBL_MAX) ? -DBL_MAX : ((DBL_MAX < v) ? DBL_MAX : v);
    double ap = acc - old;
    if (!A.stat)
    {
        A.vel = A.vel + (Jv * ap) * A.im;
        A.ang = A.ang + mul(A.Iwi, Jwa * ap);
    }
    if (!B.stat)
    {
        B.vel = B.vel + ((-Jv) * ap) * B.im;
        B.ang = B.ang + mul(B.Iwi, Jwb * ap);
    }
}

// jacobian_row::warm_start (constraint.h:74-87)
__device__ __forceinline__ void joint_warm(v3 Jv, v3 Jwa, v3 Jwb, double acc, jvel &A, jvel &B)
{
    if (acc == 0.0) return;
    if (!A.stat)
    {
        A.vel = A.vel + (Jv * acc) * A.im;
        A.ang = A.ang + mul(A.Iwi, Jwa * acc);
    }
    if (!B.stat)
    {
        B.vel = B.vel + ((-Jv) * acc) * B.im;
        B.ang = B.ang + mul(B.Iwi, Jwb * acc);
    }
}

__device__ __forceinline__ jvel ld_jvel(const bodies_dev &b, uint32_t i)
{
    jvel s;
    s.vel = ld3(b.vel, b.n, i);
    s.ang = ld3(b.ang, b.n, i);
    s.im = b.inv_mass[i];
    s.Iwi = ld9(b.Iw_inv, b.n, i);
    s.stat = b.is_static[i] != 0;
    return s;
}

__global__ void __launch_bounds__(kBlock) k_joint_solve(joints_dev J, const uint32_t *order, uint32_t begin, uint32_t end, bodies_dev b,
                              int warm)
{
    uint32_t t = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= end) return;
    const uint32_t j = order[t];
    const int nr = J.nrows[j];
    if (nr == 0) return;
    const uint32_t ia = J.a[j], ib = J.b[j];
    jvel A = ld_jvel(b, ia), B = ld_jvel(b, ib);
    for (int r = 0; r < nr; ++r)
    {
        const double *p = J.rows + static_cast<size_t>(r) * kJR * J.cap;
        v3 Jv = rd3(p, J.cap, 0, j), Jwa = rd3(p, J.cap, 3, j), Jwb = rd3(p, J.cap, 6, j);
        double acc = p[12 * static_cast<size_t>(J.cap) + j];
        if (warm)
            joint_warm(Jv, Jwa, Jwb, acc, A, B);
        else
        {
            joint_apply(Jv, Jwa, Jwb, p[9 * static_cast<size_t>(J.cap) + j], p[10 * static_cast<size_t>(J.cap) + j],
                        p[11 * static_cast<size_t>(J.cap) + j], acc, A, B);
            J.rows[(static_cast<size_t>(r) * kJR + 12) * J.cap + j] = acc;
        }
    }
    if (!A.stat)
    {
        st3(b.vel, b.n, ia, A.vel);
        st3(b.ang, b.n, ia, A.ang);
    }
    if (!B.stat)
    {
        st3(b.vel, b.n, ib, B.vel);
        st3(b.ang, b.n, ib, B.ang);
    }
}

// constraint.h:1196-1198: accumulated impulses back into the joint for the next step
__global__ void k_joint_writeback(joints_dev J)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= J.n) return;
    const int nr = J.nrows[j];
    for (int r = 0; r < nr; ++r)
        J.rec[static_cast<size_t>(25 + r) * J.cap + j] = J.rows[(static_cast<size_t>(r) * kJR + 12) * J.cap + j];
}
} // namespace

void launch_joint_rebuild(const joints_dev &old_j, const joints_dev &new_j, const uint32_t *src,
                          const double *desc, const uint32_t *dint, const bodies_dev &b, hipStream_t st)
{
    if (new_j.n) k_joint_rebuild<<<grid(new_j.n), kBlock, 0, st>>>(old_j, new_j, src, desc, dint, b);
}
void launch_joint_rows(const joints_dev &j, const bodies_dev &b, double dt, hipStream_t st)
{
    if (j.n) k_joint_rows<<<grid(j.n), kBlock, 0, st>>>(j, b, dt);
}
void launch_joint_solve(const joints_dev &j, const uint32_t *order, uint32_t begin, uint32_t end,
                        const bodies_dev &b, bool warm, hipStream_t st)
{
    if (end > begin) k_joint_solve<<<grid(end - begin), kBlock, 0, st>>>(j, order, begin, end, b, warm ? 1 : 0);
}
void launch_joint_writeback(const joints_dev &j, hipStream_t st)
{
    if (j.n) k_joint_writeback<<<grid(j.n), kBlock, 0, st>>>(j);
}

} // namespace pk
