// FP64 vector/quaternion/matrix arithmetic for the device kernels.
// Every function fixes one operation order (FMA contraction is disabled for all
// device code), so the results are reproducible bit for bit:
//   dot(a,b) = (ax*bx + ay*by) + az*bz; cross as (ay*bz - az*by, ...);
//   normalized(a) = a / sqrt(|a|^2) when |a|^2 > 0 else a;
//   q*v = v + w*uv + qv x uv with uv = 2 (qv x v);
//   q1*q2 = Hamilton product with terms summed left to right;
//   R(q): tx = 2x, twx = tx*w, ..., r00 = 1 - (tyy + tzz), ...;
//   3x3 inverse: cofactors of column 0, det = sum c_i0 m_i0, scaled by 1/det.
// These are the operations PhysKit performs through Eigen
// (include/physkit/algebra/lin_alg.h:196-235, 478-576).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#define PK_HD __host__ __device__ __forceinline__

namespace pk
{

struct v3
{
    double x, y, z;
};
PK_HD v3 mk(double x, double y, double z)
{
    v3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
PK_HD v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
PK_HD v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
PK_HD v3 operator-(v3 a) { return mk(-a.x, -a.y, -a.z); }
PK_HD v3 operator*(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
PK_HD v3 operator/(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
PK_HD double dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
PK_HD v3 cross(v3 a, v3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
PK_HD double sqn(v3 a) { return dot(a, a); }
PK_HD double norm(v3 a) { return sqrt(sqn(a)); }
PK_HD v3 normalized(v3 a)
{
    double z = sqn(a);
    return z > 0.0 ? a / sqrt(z) : a;
}
// std::min / std::max semantics
PK_HD double smin(double a, double b) { return (b < a) ? b : a; }
PK_HD double smax(double a, double b) { return (a < b) ? b : a; }

struct qt
{
    double w, x, y, z;
};
PK_HD qt mkq(double w, double x, double y, double z)
{
    qt q;
    q.w = w;
    q.x = x;
    q.y = y;
    q.z = z;
    return q;
}
PK_HD qt conj(qt q) { return mkq(q.w, -q.x, -q.y, -q.z); }
PK_HD v3 rotate(qt q, v3 v)
{
    v3 qv = mk(q.x, q.y, q.z);
    v3 uv = cross(qv, v);
    uv = uv + uv;
    return (v + uv * q.w) + cross(qv, uv);
}
PK_HD qt qmul(qt a, qt b)
{
    return mkq(a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
               a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
               a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z,
               a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x);
}
PK_HD qt qnormalized(qt q) // coefficients summed in storage order x, y, z, w
{
    double z = ((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w;
    if (!(z > 0.0)) return q;
    double s = sqrt(z);
    return mkq(q.w / s, q.x / s, q.y / s, q.z / s);
}

struct m3
{
    double m[9]; // row-major
};
PK_HD v3 mul(const m3 &a, v3 v)
{
    return mk((a.m[0] * v.x + a.m[1] * v.y) + a.m[2] * v.z,
              (a.m[3] * v.x + a.m[4] * v.y) + a.m[5] * v.z,
              (a.m[6] * v.x + a.m[7] * v.y) + a.m[8] * v.z);
}
PK_HD m3 mul(const m3 &a, const m3 &b)
{
    m3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] =
                (a.m[3 * i] * b.m[j] + a.m[3 * i + 1] * b.m[3 + j]) + a.m[3 * i + 2] * b.m[6 + j];
    return r;
}
PK_HD m3 transpose(const m3 &a)
{
    m3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[3 * i + j] = a.m[3 * j + i];
    return r;
}
PK_HD m3 rotation_matrix(qt q)
{
    double tx = 2.0 * q.x, ty = 2.0 * q.y, tz = 2.0 * q.z;
    double twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
    double txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    double tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    m3 r;
    r.m[0] = 1.0 - (tyy + tzz);
    r.m[1] = txy - twz;
    r.m[2] = txz + twy;
    r.m[3] = txy + twz;
    r.m[4] = 1.0 - (txx + tzz);
    r.m[5] = tyz - twx;
    r.m[6] = txz - twy;
    r.m[7] = tyz + twx;
    r.m[8] = 1.0 - (txx + tyy);
    return r;
}
PK_HD double cofactor(const m3 &a, int i, int j)
{
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return a.m[3 * i1 + j1] * a.m[3 * i2 + j2] - a.m[3 * i1 + j2] * a.m[3 * i2 + j1];
}
PK_HD m3 inverse(const m3 &a)
{
    double c0 = cofactor(a, 0, 0), c1 = cofactor(a, 1, 0), c2 = cofactor(a, 2, 0);
    double det = (c0 * a.m[0] + c1 * a.m[3]) + c2 * a.m[6];
    double invdet = 1.0 / det;
    m3 r;
    r.m[0] = c0 * invdet;
    r.m[1] = c1 * invdet;
    r.m[2] = c2 * invdet;
    r.m[3] = cofactor(a, 0, 1) * invdet;
    r.m[4] = cofactor(a, 1, 1) * invdet;
    r.m[5] = cofactor(a, 2, 1) * invdet;
    r.m[6] = cofactor(a, 0, 2) * invdet;
    r.m[7] = cofactor(a, 1, 2) * invdet;
    r.m[8] = cofactor(a, 2, 2) * invdet;
    return r;
}

// sin and cos of x: Cody-Waite reduction by pi/2 with a three-part constant
// (p1, p2 carry 33 significant bits so k*p1 and k*p2 are exact for |k| < 2^20),
// then Taylor polynomials (1/n! coefficients) on [-pi/4, pi/4] in Horner form.
// Within 1 ulp of glibc on |x| < 100.
PK_HD void sincos_pk(double x, double *s_out, double *c_out)
{
    const double two_over_pi = 0x1.45f306dc9c883p-1;
    const double p1 = 0x1.921fb544p+0;
    const double p2 = 0x1.0b4611a6p-34;
    const double p3 = 0x1.3198a2e037073p-69;
    double k = rint(x * two_over_pi);
    double r = ((x - k * p1) - k * p2) - k * p3;
    double r2 = r * r;
    double ps = -0x1.2f49b46814157p-57;
    ps = 0x1.952c77030ad4ap-49 + r2 * ps;
    ps = -0x1.ae7f3e733b81fp-41 + r2 * ps;
    ps = 0x1.6124613a86d09p-33 + r2 * ps;
    ps = -0x1.ae64567f544e4p-26 + r2 * ps;
    ps = 0x1.71de3a556c734p-19 + r2 * ps;
    ps = -0x1.a01a01a01a01ap-13 + r2 * ps;
    ps = 0x1.1111111111111p-7 + r2 * ps;
    ps = -0x1.5555555555555p-3 + r2 * ps;
    double s = r + (r * r2) * ps;
    double pc = 0x1.e542ba4020225p-62;
    pc = -0x1.6827863b97d97p-53 + r2 * pc;
    pc = 0x1.ae7f3e733b81fp-45 + r2 * pc;
    pc = -0x1.93974a8c07c9dp-37 + r2 * pc;
    pc = 0x1.1eed8eff8d898p-29 + r2 * pc;
    pc = -0x1.27e4fb7789f5cp-22 + r2 * pc;
    pc = 0x1.a01a01a01a01ap-16 + r2 * pc;
    pc = -0x1.6c16c16c16c17p-10 + r2 * pc;
    pc = 0x1.5555555555555p-5 + r2 * pc;
    pc = -0.5 + r2 * pc;
    double c = 1.0 + r2 * pc;
    long q = static_cast<long>(k) & 3;
    if (q == 0)
    {
        *s_out = s;
        *c_out = c;
    }
    else if (q == 1)
    {
        *s_out = c;
        *c_out = -s;
    }
    else if (q == 2)
    {
        *s_out = -s;
        *c_out = -c;
    }
    else
    {
        *s_out = -c;
        *c_out = s;
    }
}

PK_HD double hypot_pk(double a, double b) { return sqrt(a * a + b * b); }

// exponential map of include/physkit/detail/integrate.h:20-31
PK_HD qt exp_map(v3 w, double dt)
{
    v3 ang = w * dt;
    double mag = norm(ang);
    if (mag < 1e-12)
    {
        v3 half = ang * 0.5;
        return qnormalized(mkq(1.0, half.x, half.y, half.z));
    }
    v3 axis = ang / mag;
    double ha = 0.5 * mag;
    double s, c;
    sincos_pk(ha, &s, &c);
    return mkq(c, s * axis.x, s * axis.y, s * axis.z);
}

struct box3
{
    v3 min, max;
};
PK_HD bool box_contains(const box3 &outer, const box3 &in) // bounds.h:80-85
{
    return (in.min.x >= outer.min.x && in.max.x <= outer.max.x) &&
           (in.min.y >= outer.min.y && in.max.y <= outer.max.y) &&
           (in.min.z >= outer.min.z && in.max.z <= outer.max.z);
}
PK_HD bool box_intersects(const box3 &a, const box3 &b) // bounds.h:87-92
{
    return (a.min.x <= b.max.x && a.max.x >= b.min.x) && (a.min.y <= b.max.y && a.max.y >= b.min.y) &&
           (a.min.z <= b.max.z && a.max.z >= b.min.z);
}
// world bounds of a mesh instance: local AABB corners 0..7 rotated, min/max, translated
// (mesh.cpp:398, bounds.h:142-158, 112-113)
PK_HD box3 instance_bounds(const box3 &local, qt q, v3 pos)
{
    v3 p0 = rotate(q, local.min);
    box3 r;
    r.min = p0;
    r.max = p0;
    for (unsigned i = 1; i < 8; ++i)
    {
        v3 c = mk((i & 1u) ? local.max.x : local.min.x, (i & 2u) ? local.max.y : local.min.y,
                  (i & 4u) ? local.max.z : local.min.z);
        v3 pt = rotate(q, c);
        r.min.x = smin(r.min.x, pt.x);
        r.min.y = smin(r.min.y, pt.y);
        r.min.z = smin(r.min.z, pt.z);
        r.max.x = smax(r.max.x, pt.x);
        r.max.y = smax(r.max.y, pt.y);
        r.max.z = smax(r.max.z, pt.z);
    }
    r.min = r.min + pos;
    r.max = r.max + pos;
    return r;
}

PK_HD uint64_t pair_key(uint32_t a, uint32_t b)
{
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return (static_cast<uint64_t>(lo) << 32) | hi;
}

PK_HD uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

} // namespace pk
