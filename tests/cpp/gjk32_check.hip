// Host-side check of the FP32 certified-separation pass (physkit_amd/csrc/pk_gjk32.hpp):
// runs the same certification arithmetic on the CPU for shape pairs read from stdin, so
// a test can compare it with the oracle's FP64 GJK verdicts. No GPU is used.
//
// stdin: pair count, then per pair and per side (A then B):
//   nv, nv x (x y z), quaternion (w x y z), position (x y z), uv flag (1: bound the
//   shape by its circumscribed sphere, as the device does for a tagged UV-sphere mesh)
// stdout: one line per pair, 1 (proven apart) or 0
#include "pk_gjk32.hpp"

#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

using namespace pk::g32;

struct side
{
    std::vector<float4> v;
    body32 b;
    double pos[3];
};

static bool read_side(side &s)
{
    int nv = 0;
    if (std::scanf("%d", &nv) != 1 || nv <= 0) return false;
    s.v.resize(static_cast<size_t>(nv));
    double rad = 0.0;
    for (int i = 0; i < nv; ++i)
    {
        double x, y, z;
        if (std::scanf("%lf %lf %lf", &x, &y, &z) != 3) return false;
        // the device's FP32 pool: vertices rounded to nearest, radius rounded up
        s.v[static_cast<size_t>(i)] = make_float4(static_cast<float>(x), static_cast<float>(y), static_cast<float>(z), 0.0f);
        rad = std::fmax(rad, std::sqrt(x * x + y * y + z * z));
    }
    double q[4];
    int uv = 0;
    if (std::scanf("%lf %lf %lf %lf %lf %lf %lf %d", &q[0], &q[1], &q[2], &q[3], &s.pos[0], &s.pos[1], &s.pos[2],
                   &uv) != 8)
        return false;
    float r = static_cast<float>(rad);
    if (static_cast<double>(r) < rad) r = std::nextafter(r, std::numeric_limits<float>::infinity());
    s.b.q = {static_cast<float>(q[0]), static_cast<float>(q[1]), static_cast<float>(q[2]), static_cast<float>(q[3])};
    s.b.r = r;
    s.b.v = uv ? nullptr : s.v.data();
    s.b.nv = uv ? 0u : static_cast<uint32_t>(nv);
    const double qn = ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3];
    if (!(std::fabs(qn - 1.0) <= 1e-9)) s.b.r = std::numeric_limits<float>::quiet_NaN(); // never certified
    return true;
}

int main()
{
    int n = 0;
    if (std::scanf("%d", &n) != 1) return 1;
    for (int i = 0; i < n; ++i)
    {
        side a, b;
        if (!read_side(a) || !read_side(b)) return 1;
        const f3 dx = fm(static_cast<float>(a.pos[0] - b.pos[0]), static_cast<float>(a.pos[1] - b.pos[1]),
                         static_cast<float>(a.pos[2] - b.pos[2]));
        const int ok = (a.b.r == a.b.r && b.b.r == b.b.r) ? certify(a.b, b.b, dx) : 0;
        std::printf("%d\n", ok);
    }
    return 0;
}
