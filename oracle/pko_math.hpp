// Oracle (test infrastructure only): FP64 CPU restatement of PhysKit's world step.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
// and only as the checker. The product (physkit_amd/) never includes or links it.
//
// This header restates the linear algebra the reference delegates to Eigen 5.0.1
// (include/physkit/algebra/lin_alg.h:53-611) with one fixed operation order per
// function, written from the mathematics:
//   * dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z          (squaredNorm, dot)
//   * cross(a,b)    = (ay*bz - az*by, az*bx - ax*bz, ax*by - ay*bx)
//   * normalized(a) = a / sqrt(|a|^2) when |a|^2 > 0, else a  (Eigen normalized())
//   * q * v         = v + w*uv + qv x uv, uv = 2 (qv x v)    (lin_alg.h:493-499)
//   * q1 * q2       = Hamilton product, terms left to right     (lin_alg.h:478-483)
//   * R(q)          = 1 - 2(..) form with tx=2x, twx=tx*w, ...  (lin_alg.h:530-535)
//   * M^-1 (3x3)    = cofactors of column 0, det = sum c_i0*m_i0, 1/det scaling
//                                                              (lin_alg.h:204-209)
//   * AngleAxis     = (cos(a/2), sin(a/2) axis)                (lin_alg.h:569-576)
// Eigen's own source is not available here, so agreement with the reference at the
// last bit is "parity unpinned"; the tests pin these against closed forms.
// sin/cos: Cody-Waite reduction by pi/2 (three-part constant) and Taylor polynomials
// on [-pi/4, pi/4] evaluated in Horner form; the device code carries the same
// algorithm so both sides agree bit for bit. Error vs libm is checked in tests
// (<= 2 ulp).
#pragma once
#include <cmath>
#include <cstdint>

namespace pko
{

struct v3
{
    double x = 0, y = 0, z = 0;
};
inline v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
inline v3 operator*(v3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline v3 operator/(v3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline v3 cross(v3 a, v3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double sqn(v3 a) { return dot(a, a); }
inline double norm(v3 a) { return std::sqrt(sqn(a)); }
inline v3 normalized(v3 a)
{
    double z = sqn(a);
    return z > 0.0 ? a / std::sqrt(z) : a;
}
// std::min / std::max semantics: min(a,b) = (b < a) ? b : a ; max(a,b) = (a < b) ? b : a
inline double smin(double a, double b) { return (b < a) ? b : a; }
inline double smax(double a, double b) { return (a < b) ? b : a; }

struct qt
{
    double w = 1, x = 0, y = 0, z = 0;
};
inline qt conj(qt q) { return {q.w, -q.x, -q.y, -q.z}; }
inline v3 rotate(qt q, v3 v)
{
    v3 qv{q.x, q.y, q.z};
    v3 uv = cross(qv, v);
    uv = uv + uv;
    return (v + uv * q.w) + cross(qv, uv);
}
inline qt qmul(qt a, qt b)
{
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
            a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z,
            a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x};
}
// coefficient storage order of the reference's quaternion is (x, y, z, w)
inline qt qnormalized(qt q)
{
    double z = ((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w;
    if (!(z > 0.0)) return q;
    double s = std::sqrt(z);
    return {q.w / s, q.x / s, q.y / s, q.z / s};
}

struct m3
{
    double m[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
};
inline m3 m3_zero()
{
    m3 r;
    for (auto &row : r.m)
        for (double &e : row) e = 0.0;
    return r;
}
inline v3 mul(const m3 &a, v3 v)
{
    return {(a.m[0][0] * v.x + a.m[0][1] * v.y) + a.m[0][2] * v.z,
            (a.m[1][0] * v.x + a.m[1][1] * v.y) + a.m[1][2] * v.z,
            (a.m[2][0] * v.x + a.m[2][1] * v.y) + a.m[2][2] * v.z};
}
inline m3 mul(const m3 &a, const m3 &b)
{
    m3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[i][j] = (a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j];
    return r;
}
inline m3 transpose(const m3 &a)
{
    m3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}
inline m3 rotation_matrix(qt q)
{
    double tx = 2.0 * q.x, ty = 2.0 * q.y, tz = 2.0 * q.z;
    double twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
    double txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    double tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    m3 r;
    r.m[0][0] = 1.0 - (tyy + tzz);
    r.m[0][1] = txy - twz;
    r.m[0][2] = txz + twy;
    r.m[1][0] = txy + twz;
    r.m[1][1] = 1.0 - (txx + tzz);
    r.m[1][2] = tyz - twx;
    r.m[2][0] = txz - twy;
    r.m[2][1] = tyz + twx;
    r.m[2][2] = 1.0 - (txx + tyy);
    return r;
}
// cofactor C(i,j) = m(i1,j1) m(i2,j2) - m(i1,j2) m(i2,j1), i1=(i+1)%3, i2=(i+2)%3
inline double cof(const m3 &a, int i, int j)
{
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return a.m[i1][j1] * a.m[i2][j2] - a.m[i1][j2] * a.m[i2][j1];
}
inline m3 inverse(const m3 &a)
{
    double c0 = cof(a, 0, 0), c1 = cof(a, 1, 0), c2 = cof(a, 2, 0);
    double det = (c0 * a.m[0][0] + c1 * a.m[1][0]) + c2 * a.m[2][0];
    double invdet = 1.0 / det;
    m3 r;
    r.m[0][0] = c0 * invdet;
    r.m[0][1] = c1 * invdet;
    r.m[0][2] = c2 * invdet;
    r.m[1][0] = cof(a, 0, 1) * invdet;
    r.m[1][1] = cof(a, 1, 1) * invdet;
    r.m[1][2] = cof(a, 2, 1) * invdet;
    r.m[2][0] = cof(a, 0, 2) * invdet;
    r.m[2][1] = cof(a, 1, 2) * invdet;
    r.m[2][2] = cof(a, 2, 2) * invdet;
    return r;
}

// ---- sin / cos (own implementation; see header comment) ----
inline void sincos_pk(double x, double *s_out, double *c_out)
{
    const double two_over_pi = 0x1.45f306dc9c883p-1;
    const double p1 = 0x1.921fb544p+0;       // 33 significant bits
    const double p2 = 0x1.0b4611a6p-34;      // 33 significant bits
    const double p3 = 0x1.3198a2e037073p-69; // remainder
    double k = std::nearbyint(x * two_over_pi);
    double r = ((x - k * p1) - k * p2) - k * p3;
    double r2 = r * r;
    // sin: r + r^3 * (s1 + r2 (s2 + ...)), Taylor coefficients (-1)^n/(2n+1)!
    double ps = -0x1.2f49b46814157p-57;    // -1/19!
    ps = 0x1.952c77030ad4ap-49 + r2 * ps;  //  1/17!
    ps = -0x1.ae7f3e733b81fp-41 + r2 * ps; // -1/15!
    ps = 0x1.6124613a86d09p-33 + r2 * ps;  //  1/13!
    ps = -0x1.ae64567f544e4p-26 + r2 * ps; // -1/11!
    ps = 0x1.71de3a556c734p-19 + r2 * ps;  //  1/9!
    ps = -0x1.a01a01a01a01ap-13 + r2 * ps; // -1/7!
    ps = 0x1.1111111111111p-7 + r2 * ps;   //  1/5!
    ps = -0x1.5555555555555p-3 + r2 * ps;  // -1/3!
    double s = r + (r * r2) * ps;
    // cos: 1 + r2 * (c1 + r2 (c2 + ...)), coefficients (-1)^n/(2n)!
    double pc = 0x1.e542ba4020225p-62;     //  1/20!
    pc = -0x1.6827863b97d97p-53 + r2 * pc; // -1/18!
    pc = 0x1.ae7f3e733b81fp-45 + r2 * pc;  //  1/16!
    pc = -0x1.93974a8c07c9dp-37 + r2 * pc; // -1/14!
    pc = 0x1.1eed8eff8d898p-29 + r2 * pc;  //  1/12!
    pc = -0x1.27e4fb7789f5cp-22 + r2 * pc; // -1/10!
    pc = 0x1.a01a01a01a01ap-16 + r2 * pc;  //  1/8!
    pc = -0x1.6c16c16c16c17p-10 + r2 * pc; // -1/6!
    pc = 0x1.5555555555555p-5 + r2 * pc;   //  1/4!
    pc = -0.5 + r2 * pc;                   // -1/2!
    double c = 1.0 + r2 * pc;
    long q = static_cast<long>(k) & 3;
    switch (q)
    {
    case 0: *s_out = s; *c_out = c; break;
    case 1: *s_out = c; *c_out = -s; break;
    case 2: *s_out = -s; *c_out = -c; break;
    default: *s_out = -c; *c_out = s; break;
    }
}

inline double hypot_pk(double a, double b) { return std::sqrt(a * a + b * b); }

} // namespace pko
