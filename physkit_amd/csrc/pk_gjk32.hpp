// FP32 certified separation: the arithmetic (host and device), shared by the kernel in
// k_gjk32.hip and the host-side checks of the tests. What it proves and the error bound
// it uses are stated in k_gjk32.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace pk
{
namespace g32
{
constexpr int kMaxIter = 24;
constexpr float kU = 5.9604645e-8f; // 2^-24

struct f3
{
    float x, y, z;
};
__host__ __device__ __forceinline__ f3 fm(float x, float y, float z) { return {x, y, z}; }
__host__ __device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__host__ __device__ __forceinline__ f3 operator*(f3 a, float k) { return {a.x * k, a.y * k, a.z * k}; }
__host__ __device__ __forceinline__ float fdot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__host__ __device__ __forceinline__ f3 fcross(f3 a, f3 b)
{ return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct q4
{
    float w, x, y, z;
};
// rotate(q, v) = v + w t + u x t with t = 2 (u x v): the rotation of a unit quaternion
__host__ __device__ __forceinline__ f3 frot(q4 q, f3 v)
{
    const f3 u = fm(q.x, q.y, q.z);
    const f3 t = fcross(u, v) * 2.0f;
    return (v + t * q.w) + fcross(u, t);
}
__host__ __device__ __forceinline__ f3 frot_inv(q4 q, f3 v) { return frot({q.w, -q.x, -q.y, -q.z}, v); }

struct body32
{
    q4 q;
    const float4 *v; // vertices, or null for a shape bounded by its sphere
    uint32_t nv;
    float r; // largest vertex distance from the origin, rounded up
};

// support of one body about its origin in world direction d: the extreme point (world
// frame, relative to the body's position) and its value h = max v . d
__host__ __device__ __forceinline__ f3 support(const body32 &s, f3 d, float &h)
{
    if (!s.v)
    {
        const float len = sqrtf(fdot(d, d));
        h = s.r * len * (1.0f + 4.0f * kU);
        return len > 0.0f ? d * (s.r / len) : fm(s.r, 0.0f, 0.0f);
    }
    const f3 dl = frot_inv(s.q, d);
    float best = -__builtin_inff();
    f3 bv = fm(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < s.nv; ++i)
    {
        const float4 v = s.v[i];
        const float t = (v.x * dl.x + v.y * dl.y) + v.z * dl.z;
        if (t > best)
        {
            best = t;
            bv = fm(v.x, v.y, v.z);
        }
    }
    h = best;
    return frot(s.q, bv);
}

// closest point of the triangle (a, b, c) to the origin with its barycentric weights
// (Ericson 5.1.5)
__host__ __device__ __forceinline__ f3 closest_tri(f3 a, f3 b, f3 c, float &wa, float &wb, float &wc)
{
    const f3 ab = b - a, ac = c - a;
    const float d1 = -fdot(ab, a), d2 = -fdot(ac, a);
    if (d1 <= 0.0f && d2 <= 0.0f)
    {
        wa = 1.0f, wb = 0.0f, wc = 0.0f;
        return a;
    }
    const float d3 = -fdot(ab, b), d4 = -fdot(ac, b);
    if (d3 >= 0.0f && d4 <= d3)
    {
        wa = 0.0f, wb = 1.0f, wc = 0.0f;
        return b;
    }
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f)
    {
        const float v = d1 / (d1 - d3);
        wa = 1.0f - v, wb = v, wc = 0.0f;
        return a + ab * v;
    }
    const float d5 = -fdot(ab, c), d6 = -fdot(ac, c);
    if (d6 >= 0.0f && d5 <= d6)
    {
        wa = 0.0f, wb = 0.0f, wc = 1.0f;
        return c;
    }
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f)
    {
        const float w = d2 / (d2 - d6);
        wa = 1.0f - w, wb = 0.0f, wc = w;
        return a + ac * w;
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f)
    {
        const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        wa = 0.0f, wb = 1.0f - w, wc = w;
        return b + (c - b) * w;
    }
    const float den = 1.0f / (va + vb + vc);
    const float v = vb * den, w = vc * den;
    wa = 1.0f - v - w, wb = v, wc = w;
    // interior: the origin's projection on the plane along the face normal. Its
    // direction keeps FP32 accuracy when the origin is close to a large face, where
    // a + ab v + ac w would lose it to cancellation.
    const f3 nrm = fcross(ab, ac);
    const float nn = fdot(nrm, nrm);
    if (nn > 0.0f) return nrm * (fdot(nrm, a) / nn);
    return (a + ab * v) + ac * w;
}

// the simplex's closest point to the origin; the simplex is reduced to the points with
// non-zero weight. Returns false when the origin is enclosed (a tetrahedron with the
// origin on the inner side of every face).
__host__ __device__ inline bool closest_simplex(f3 (&y)[4], int &n, f3 &v)
{
    if (n == 1)
    {
        v = y[0];
        return true;
    }
    if (n == 2)
    {
        const f3 ab = y[1] - y[0];
        const float den = fdot(ab, ab);
        float t = den > 0.0f ? -fdot(y[0], ab) / den : 0.0f;
        if (!(t > 0.0f))
        {
            n = 1;
            v = y[0];
            return true;
        }
        if (t >= 1.0f)
        {
            y[0] = y[1];
            n = 1;
            v = y[0];
            return true;
        }
        v = y[0] + ab * t;
        return true;
    }
    if (n == 3)
    {
        float w[3];
        v = closest_tri(y[0], y[1], y[2], w[0], w[1], w[2]);
        int k = 0;
        for (int i = 0; i < 3; ++i)
            if (w[i] > 0.0f) y[k++] = y[i];
        n = k > 0 ? k : 1;
        return true;
    }
    // tetrahedron: each face that has the origin on the far side from the fourth point
    const int fc[4][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2}, {1, 3, 2, 0}};
    float best = __builtin_inff();
    int bf = -1;
    float bw[3] = {0.0f, 0.0f, 0.0f};
    f3 bp = fm(0.0f, 0.0f, 0.0f);
    for (int f = 0; f < 4; ++f)
    {
        const f3 a = y[fc[f][0]], b = y[fc[f][1]], c = y[fc[f][2]], d = y[fc[f][3]];
        const f3 nrm = fcross(b - a, c - a);
        const float so = -fdot(a, nrm), sd = fdot(d - a, nrm);
        if (!(so * sd < 0.0f) && sd != 0.0f) continue; // origin on d's side of this face
        float w0, w1, w2;
        const f3 p = closest_tri(a, b, c, w0, w1, w2);
        const float dd = fdot(p, p);
        if (dd < best)
        {
            best = dd;
            bf = f;
            bp = p;
            bw[0] = w0, bw[1] = w1, bw[2] = w2;
        }
    }
    if (bf < 0) return false;
    f3 z[3];
    int k = 0;
    for (int i = 0; i < 3; ++i)
        if (bw[i] > 0.0f) z[k++] = y[fc[bf][i]];
    for (int i = 0; i < k; ++i) y[i] = z[i];
    n = k > 0 ? k : 1;
    v = bp;
    return true;
}


// 1 when shapes A (its origin at dx from B's) and B are proven apart by more than the
// margin 1e-6 m + 1e-6 (R_A + R_B) plus the FP32 rounding bound, else 0
__host__ __device__ inline int certify(const body32 &A, const body32 &B, f3 dx)
{
    const float dxn = sqrtf(fdot(dx, dx));
    const float rsum = A.r + B.r;
    const float margin = 1e-6f * rsum + 1e-6f;
    const float err = 256.0f * kU * (rsum + dxn);
    f3 y[4];
    int n = 0;
    f3 v = dxn > 0.0f ? dx : fm(1.0f, 0.0f, 0.0f);
    for (int it = 0; it < kMaxIter; ++it)
    {
        const float vv = fdot(v, v);
        if (!(vv > 0.0f)) return 0; // the origin is on the simplex: touching or overlapping
        const float len = sqrtf(vv);
        const f3 nrm = v * (1.0f / len);
        // w: the point of the Minkowski difference A - B extreme toward -n; gap along n
        float ha, hb;
        const f3 sa = support(A, -nrm, ha);
        const f3 sb = support(B, nrm, hb);
        const float gap = (fdot(dx, nrm) - ha) - hb;
        const float nlen = sqrtf(fdot(nrm, nrm));
        if (gap > (margin + err) * (nlen * (1.0f + 4.0f * kU))) return 1;
        const f3 w = (sa + dx) - sb;
        // no further progress toward the origin: the distance is about |v|, below the
        // certifiable gap
        if (vv - fdot(v, w) <= 1e-6f * vv) return 0;
        for (int i = 0; i < n; ++i)
            if (y[i].x == w.x && y[i].y == w.y && y[i].z == w.z) return 0;
        y[n++] = w;
        if (!closest_simplex(y, n, v)) return 0; // origin enclosed
    }
    return 0;
}

} // namespace g32
} // namespace pk
