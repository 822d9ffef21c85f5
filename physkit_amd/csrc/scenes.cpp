// Synthetic scenes for tests and benchmarks (input generation only).
// RNG: SplitMix64-seeded xoshiro256**; U[a,b) = a + (b-a) * (x >> 11) * 2^-53;
// normals by Box-Muller; uniform rotations by Shoemake's method.
// PhysKit cannot compute mesh inertia (mesh::inertia_tensor throws, src/mesh.cpp:
// 154-166), so the generators supply closed-form inertia tensors.
#include "physkit_amd.h"

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

struct pkg_scene
{
    std::vector<uint32_t> shape_offsets{0};
    std::vector<double> verts;
    std::vector<uint8_t> is_static;
    std::vector<uint32_t> shape;
    std::vector<double> pos, vel, angvel, quat, mass, inertia, restitution, friction;
    std::vector<uint32_t> first_body{0}; // world w = bodies [first_body[w], first_body[w+1])
    pkg_scene_info info{};
};

namespace
{
constexpr double kPi = 3.141592653589793;

struct rng
{
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x)
    {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    explicit rng(uint64_t seed)
    {
        for (auto &w : s) w = splitmix(seed);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double u01() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
    double uni(double a, double b) { return a + (b - a) * u01(); }
    double normal()
    {
        double u1 = 1.0 - u01(); // (0, 1]
        double u2 = u01();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
    void quat(double *q) // (w, x, y, z)
    {
        double u1 = u01(), u2 = u01(), u3 = u01();
        double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
        q[1] = a * std::sin(2.0 * kPi * u2);
        q[2] = a * std::cos(2.0 * kPi * u2);
        q[3] = b * std::sin(2.0 * kPi * u3);
        q[0] = b * std::cos(2.0 * kPi * u3);
    }
};

// vertex order of mesh::box (src/mesh.cpp:31-40)
uint32_t add_box_shape(pkg_scene &s, double hx, double hy, double hz)
{
    const double v[8][3] = {{-hx, -hy, -hz}, {+hx, -hy, -hz}, {+hx, +hy, -hz}, {-hx, +hy, -hz},
                            {-hx, -hy, +hz}, {+hx, -hy, +hz}, {+hx, +hy, +hz}, {-hx, +hy, +hz}};
    for (const auto &p : v) s.verts.insert(s.verts.end(), p, p + 3);
    s.shape_offsets.push_back(static_cast<uint32_t>(s.verts.size() / 3));
    return static_cast<uint32_t>(s.shape_offsets.size() - 2);
}

void add_body(pkg_scene &s, bool is_static, uint32_t shape, const double *p, const double *v,
              const double *q, double m, const double *I, double e, double mu)
{
    s.is_static.push_back(is_static ? 1 : 0);
    s.shape.push_back(shape);
    s.pos.insert(s.pos.end(), p, p + 3);
    s.vel.insert(s.vel.end(), v, v + 3);
    const double zero[3] = {0, 0, 0};
    s.angvel.insert(s.angvel.end(), zero, zero + 3);
    s.quat.insert(s.quat.end(), q, q + 4);
    s.mass.push_back(m);
    s.inertia.insert(s.inertia.end(), I, I + 9);
    s.restitution.push_back(e);
    s.friction.push_back(mu);
}

void diag(double *I, double a, double b, double c)
{
    for (int i = 0; i < 9; ++i) I[i] = 0.0;
    I[0] = a;
    I[4] = b;
    I[8] = c;
}

const double kIdentityQ[4] = {1, 0, 0, 0};
const double kZero3[3] = {0, 0, 0};

// 1,024 boxes, half extents U[0.1,0.5] per axis, centres x,z in U[-10,10], y in
// U[1,21], uniform rotations, v ~ N(0,1), density 1000, over a static 100x1x100 m box
void scene_boxes(pkg_scene &s, uint64_t seed, uint32_t n)
{
    rng r(seed);
    double I[9];
    uint32_t g = add_box_shape(s, 50.0, 0.5, 50.0);
    const double gp[3] = {0.0, -0.5, 0.0};
    diag(I, 1, 1, 1);
    add_body(s, true, g, gp, kZero3, kIdentityQ, 0.0, I, 0.5, 0.5);
    for (uint32_t i = 0; i < n; ++i)
    {
        double hx = r.uni(0.1, 0.5), hy = r.uni(0.1, 0.5), hz = r.uni(0.1, 0.5);
        double p[3] = {r.uni(-10.0, 10.0), r.uni(1.0, 21.0), r.uni(-10.0, 10.0)};
        double q[4];
        r.quat(q);
        double v[3] = {r.normal(), r.normal(), r.normal()};
        double m = 1000.0 * 8.0 * hx * hy * hz;
        diag(I, m / 3.0 * (hy * hy + hz * hz), m / 3.0 * (hx * hx + hz * hz),
             m / 3.0 * (hx * hx + hy * hy));
        uint32_t sh = add_box_shape(s, hx, hy, hz);
        add_body(s, false, sh, p, v, q, m, I, 0.5, 0.5);
    }
    s.info.dt = 1.0 / 60.0;
    s.info.iterations = 10;
    s.info.steps = 120;
    s.info.flags = 0;
}

// convex hulls of 8-32 points on ellipsoids with semi-axes U[0.05,0.3]; every sample
// point lies on the strictly convex ellipsoid, so every point is a hull vertex.
// Centres from Gaussian clusters of 512 bodies each, cluster means in a 200 m cube.
// kClusterSigma = 0.4 m scaled by kSpread, tuned so the mean pairs per body is ~4.
constexpr double kClusterSigma = 0.4;
constexpr double kSpread = 2.65;
void scene_hulls(pkg_scene &s, uint64_t seed, uint32_t n)
{
    rng r(seed);
    uint32_t nc = n / 512 ? n / 512 : 1;
    std::vector<double> means(3 * nc);
    for (auto &m : means) m = r.uni(-100.0, 100.0);
    double I[9];
    for (uint32_t i = 0; i < n; ++i)
    {
        uint32_t k = 8 + static_cast<uint32_t>(r.next() % 25);
        double a = r.uni(0.05, 0.3), b = r.uni(0.05, 0.3), c = r.uni(0.05, 0.3);
        for (uint32_t j = 0; j < k; ++j)
        {
            double z = r.uni(-1.0, 1.0), phi = r.uni(0.0, 2.0 * kPi);
            double rho = std::sqrt(1.0 - z * z);
            s.verts.push_back(a * rho * std::cos(phi));
            s.verts.push_back(b * rho * std::sin(phi));
            s.verts.push_back(c * z);
        }
        s.shape_offsets.push_back(static_cast<uint32_t>(s.verts.size() / 3));
        uint32_t cl = static_cast<uint32_t>(r.next() % nc);
        double sig = kClusterSigma * kSpread;
        double p[3] = {means[3 * cl] + sig * r.normal(), means[3 * cl + 1] + sig * r.normal(),
                       means[3 * cl + 2] + sig * r.normal()};
        double q[4];
        r.quat(q);
        double m = 1000.0 * (4.0 / 3.0) * kPi * a * b * c;
        diag(I, m / 5.0 * (b * b + c * c), m / 5.0 * (a * a + c * c), m / 5.0 * (a * a + b * b));
        add_body(s, false, i, p, kZero3, q, m, I, 0.5, 0.5);
    }
    s.info.dt = 1.0 / 60.0;
    s.info.iterations = 10;
    s.info.steps = 60;
    s.info.flags = PKG_FLAG_ALL_MOVED | PKG_FLAG_NO_SOLVER | PKG_FLAG_NO_INTEGRATE;
}

// pyramids of 15 layers of unit boxes (layer k holds 15-k boxes), on an 8x8 grid
// 20 m apart, x/z jitter N(0, 0.002), mass 1, I = 1/6, mu 0.6, e 0, over a ground box
void scene_pyramids(pkg_scene &s, uint64_t seed, uint32_t n_pyr)
{
    rng r(seed);
    double I[9];
    uint32_t g = add_box_shape(s, 100.0, 0.5, 100.0);
    uint32_t ub = add_box_shape(s, 0.5, 0.5, 0.5);
    const double gp[3] = {0.0, -0.5, 0.0};
    diag(I, 1, 1, 1);
    add_body(s, true, g, gp, kZero3, kIdentityQ, 0.0, I, 0.0, 0.6);
    diag(I, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0);
    for (uint32_t pi = 0; pi < n_pyr; ++pi)
    {
        double cx = (static_cast<double>(pi % 8) - 3.5) * 20.0;
        double cz = (static_cast<double>(pi / 8) - 3.5) * 20.0;
        for (int layer = 0; layer < 15; ++layer)
        {
            int cnt = 15 - layer;
            for (int j = 0; j < cnt; ++j)
            {
                double p[3] = {cx + (j - (cnt - 1) * 0.5) + 0.002 * r.normal(),
                               0.5 + static_cast<double>(layer), cz + 0.002 * r.normal()};
                add_body(s, false, ub, p, kZero3, kIdentityQ, 1.0, I, 0.0, 0.6);
            }
        }
    }
    s.info.dt = 1.0 / 120.0;
    s.info.iterations = 20;
    s.info.steps = 600;
    // exactly stacked axis-aligned boxes make the reference EPA emit zero-normal,
    // NaN-witness contacts; this scene runs with the documented guard
    s.info.flags = PKG_FLAG_EPA_GUARD;
}
// vertices of mesh::sphere(r, stacks, sectors) (src/mesh.cpp:60-115): north pole, the
// rings from the top down, south pole
uint32_t add_sphere_shape(pkg_scene &s, double rad, unsigned stacks, unsigned sectors)
{
    s.verts.push_back(0.0);
    s.verts.push_back(rad);
    s.verts.push_back(0.0);
    for (unsigned i = 1; i < stacks; ++i)
    {
        double phi = kPi * static_cast<double>(i) / static_cast<double>(stacks);
        for (unsigned j = 0; j < sectors; ++j)
        {
            double th = 2.0 * kPi * static_cast<double>(j) / static_cast<double>(sectors);
            s.verts.push_back(rad * std::sin(phi) * std::cos(th));
            s.verts.push_back(rad * std::cos(phi));
            s.verts.push_back(rad * std::sin(phi) * std::sin(th));
        }
    }
    s.verts.push_back(0.0);
    s.verts.push_back(-rad);
    s.verts.push_back(0.0);
    s.shape_offsets.push_back(static_cast<uint32_t>(s.verts.size() / 3));
    return static_cast<uint32_t>(s.shape_offsets.size() - 2);
}

// hull of k points on an ellipsoid (all of them are vertices of the hull)
uint32_t add_ellipsoid_hull(pkg_scene &s, rng &r, uint32_t k, double a, double b, double c)
{
    for (uint32_t j = 0; j < k; ++j)
    {
        double z = r.uni(-1.0, 1.0), phi = r.uni(0.0, 2.0 * kPi);
        double rho = std::sqrt(1.0 - z * z);
        s.verts.push_back(a * rho * std::cos(phi));
        s.verts.push_back(b * rho * std::sin(phi));
        s.verts.push_back(c * z);
    }
    s.shape_offsets.push_back(static_cast<uint32_t>(s.verts.size() / 3));
    return static_cast<uint32_t>(s.shape_offsets.size() - 2);
}

// Batched uneven worlds: n_worlds independent worlds with body counts log-uniform in
// [16, 4096]; each is the boxes scene scaled to its size (cube edge 20 m x
// (N/1024)^(1/3), ground box half extent 2.5 x edge) with a body mix of 40% UV
// spheres (r U[0.05,0.5], 16 stacks x 32 sectors), 40% boxes (half extents
// U[0.1,0.5]) and 20% hulls of 8-16 points on ellipsoids (semi-axes U[0.1,0.5]);
// density 1000, v ~ N(0,1), uniform rotations, e = mu = 0.5.
void scene_batched(pkg_scene &s, uint64_t seed, uint32_t n_worlds)
{
    rng r(seed);
    double I[9];
    for (uint32_t w = 0; w < n_worlds; ++w)
    {
        uint32_t nb = static_cast<uint32_t>(16.0 * std::pow(256.0, r.u01()));
        double edge = 20.0 * std::cbrt(static_cast<double>(nb) / 1024.0);
        double gh = 2.5 * edge;
        uint32_t g = add_box_shape(s, gh, 0.5, gh);
        const double gp[3] = {0.0, -0.5, 0.0};
        diag(I, 1, 1, 1);
        add_body(s, true, g, gp, kZero3, kIdentityQ, 0.0, I, 0.5, 0.5);
        for (uint32_t i = 0; i < nb; ++i)
        {
            double kind = r.u01();
            double p[3] = {r.uni(-0.5 * edge, 0.5 * edge), r.uni(1.0, 1.0 + edge), r.uni(-0.5 * edge, 0.5 * edge)};
            double q[4];
            r.quat(q);
            double v[3] = {r.normal(), r.normal(), r.normal()};
            uint32_t sh;
            double m;
            if (kind < 0.4)
            {
                double rad = r.uni(0.05, 0.5);
                sh = add_sphere_shape(s, rad, 16, 32);
                m = 1000.0 * (4.0 / 3.0) * kPi * rad * rad * rad;
                diag(I, 0.4 * m * rad * rad, 0.4 * m * rad * rad, 0.4 * m * rad * rad);
            }
            else if (kind < 0.8)
            {
                double hx = r.uni(0.1, 0.5), hy = r.uni(0.1, 0.5), hz = r.uni(0.1, 0.5);
                sh = add_box_shape(s, hx, hy, hz);
                m = 1000.0 * 8.0 * hx * hy * hz;
                diag(I, m / 3.0 * (hy * hy + hz * hz), m / 3.0 * (hx * hx + hz * hz), m / 3.0 * (hx * hx + hy * hy));
            }
            else
            {
                uint32_t k = 8 + static_cast<uint32_t>(r.next() % 9);
                double a = r.uni(0.1, 0.5), b = r.uni(0.1, 0.5), c = r.uni(0.1, 0.5);
                sh = add_ellipsoid_hull(s, r, k, a, b, c);
                m = 1000.0 * (4.0 / 3.0) * kPi * a * b * c;
                diag(I, m / 5.0 * (b * b + c * c), m / 5.0 * (a * a + c * c), m / 5.0 * (a * a + b * b));
            }
            add_body(s, false, sh, p, v, q, m, I, 0.5, 0.5);
        }
        s.first_body.push_back(static_cast<uint32_t>(s.shape.size()));
    }
    s.info.dt = 1.0 / 60.0;
    s.info.iterations = 10;
    s.info.steps = 120;
    s.info.flags = 0;
}
// Static box at p with half extents h, rotated by q (w, x, y, z)
void add_static_box(pkg_scene &s, double hx, double hy, double hz, const double *p, const double *q)
{
    double I[9];
    diag(I, 1, 1, 1);
    uint32_t sh = add_box_shape(s, hx, hy, hz);
    add_body(s, true, sh, p, kZero3, q, 0.0, I, 0.5, 0.5);
}

// Granular pile: n bodies (70% UV spheres r U[0.01,0.03] with the given stacks and
// sectors, 30% boxes with half extents U[0.01,0.03]) on a jittered cubic lattice
// (spacing 0.07 m, jitter U[-0.004,0.004] per axis, uniform rotations, at rest),
// dropped into a hopper of 64 static boxes: a 2x2 floor of 5x5 m plates at y = 0
// and four walls sloping from half-width 1.2 m at y = 1 to 7 m at y = 9, each wall
// 15 plates (5 along the slope x 3 across). Density 1000, e = mu = 0.5. Every length
// is multiplied by L (L = 10: bodies of 10-30 cm).
void scene_granular(pkg_scene &s, uint64_t seed, uint32_t n, unsigned stacks, unsigned sectors, double L = 1.0)
{
    rng r(seed);
    double I[9];
    // floor
    for (int k = 0; k < 4; ++k)
    {
        double p[3] = {L * ((k & 1) ? 2.5 : -2.5), -0.05 * L, L * ((k & 2) ? 2.5 : -2.5)};
        add_static_box(s, 2.5 * L, 0.05 * L, 2.5 * L, p, kIdentityQ);
    }
    // four sloped walls: wall w faces the centre; a plate's local x runs across the
    // wall, local z down the slope, local y is the wall normal
    const double y0 = 1.0 * L, y1 = 9.0 * L, h0 = 1.2 * L, h1 = 7.0 * L;
    const double run = h1 - h0, rise = y1 - y0;
    const double len = std::sqrt(run * run + rise * rise);
    const double tilt = std::atan2(rise, run); // wall angle above the horizontal
    for (int w = 0; w < 4; ++w)
    {
        double yaw = 0.5 * kPi * w; // wall w sits on the +x side rotated by yaw about y
        for (int a = 0; a < 5; ++a)     // along the slope
            for (int b = 0; b < 3; ++b) // across
            {
                double u = (a + 0.5) / 5.0;
                double rad = h0 + u * run, y = y0 + u * rise;
                double across = (b - 1) * (2.0 * rad / 3.0);
                double lp[3] = {rad + 0.05 * L, y, across}; // in the wall frame (before yaw)
                double c = std::cos(yaw), sn = std::sin(yaw);
                double p[3] = {c * lp[0] + sn * lp[2], lp[1], -sn * lp[0] + c * lp[2]};
                // orientation: about z by tilt - pi/2, taking the plate's long axis
                // (local y) onto the slope (cos tilt, sin tilt, 0); then the yaw
                double qt[4] = {std::cos(0.5 * (tilt - 0.5 * kPi)), 0.0, 0.0, std::sin(0.5 * (tilt - 0.5 * kPi))};
                double qy[4] = {std::cos(0.5 * yaw), 0.0, std::sin(0.5 * yaw), 0.0};
                double q[4] = {qy[0] * qt[0] - qy[1] * qt[1] - qy[2] * qt[2] - qy[3] * qt[3],
                               qy[0] * qt[1] + qy[1] * qt[0] + qy[2] * qt[3] - qy[3] * qt[2],
                               qy[0] * qt[2] - qy[1] * qt[3] + qy[2] * qt[0] + qy[3] * qt[1],
                               qy[0] * qt[3] + qy[1] * qt[2] - qy[2] * qt[1] + qy[3] * qt[0]};
                add_static_box(s, 0.05 * L, len / 10.0, rad / 3.0 + 0.05 * L, p, q);
            }
    }
    // lattice of bodies above the hopper opening
    uint32_t side = 1;
    while (static_cast<uint64_t>(side) * side * side < n) ++side;
    const double sp = 0.07 * L, jit = 0.004 * L;
    const double base = -0.5 * sp * (side - 1);
    uint32_t placed = 0;
    for (uint32_t iy = 0; iy < side && placed < n; ++iy)
        for (uint32_t iz = 0; iz < side && placed < n; ++iz)
            for (uint32_t ix = 0; ix < side && placed < n; ++ix, ++placed)
            {
                double p[3] = {base + sp * ix + r.uni(-jit, jit), 10.0 * L + sp * iy + r.uni(-jit, jit),
                               base + sp * iz + r.uni(-jit, jit)};
                double q[4];
                r.quat(q);
                uint32_t sh;
                double m;
                if (r.u01() < 0.7)
                {
                    double rad = L * r.uni(0.01, 0.03);
                    sh = add_sphere_shape(s, rad, stacks, sectors);
                    m = 1000.0 * (4.0 / 3.0) * kPi * rad * rad * rad;
                    diag(I, 0.4 * m * rad * rad, 0.4 * m * rad * rad, 0.4 * m * rad * rad);
                }
                else
                {
                    double hx = L * r.uni(0.01, 0.03), hy = L * r.uni(0.01, 0.03), hz = L * r.uni(0.01, 0.03);
                    sh = add_box_shape(s, hx, hy, hz);
                    m = 1000.0 * 8.0 * hx * hy * hz;
                    diag(I, m / 3.0 * (hy * hy + hz * hz), m / 3.0 * (hx * hx + hz * hz), m / 3.0 * (hx * hx + hy * hy));
                }
                add_body(s, false, sh, p, kZero3, q, m, I, 0.5, 0.5);
            }
    s.info.dt = 1.0 / 240.0;
    s.info.iterations = 10;
    s.info.steps = 240;
    s.info.flags = 0;
}
} // namespace

extern "C"
{
pkg_status pkg_scene_create(pkg_scene **out, int kind, uint64_t seed, uint32_t n_override)
{
    if (!out) return PKG_INVALID_ARG;
    auto *s = new (std::nothrow) pkg_scene();
    if (!s) return PKG_OOM;
    switch (kind)
    {
    case 0: scene_boxes(*s, seed, n_override ? n_override : 1024); break;
    case 1: scene_hulls(*s, seed, n_override ? n_override : 262144); break;
    case 2: scene_pyramids(*s, seed, n_override ? n_override : 64); break;
    case 3: scene_batched(*s, seed, n_override ? n_override : 4096); break;
    case 4: scene_granular(*s, seed, n_override ? n_override : 2097152, 16, 32); break;
    case 5: scene_granular(*s, seed, n_override ? n_override : 2097152, 8, 12); break;
    case 6: scene_granular(*s, seed, n_override ? n_override : 2097152, 16, 32, 10.0); break;
    default: delete s; return PKG_INVALID_ARG;
    }
    if (kind != 3) s->first_body.push_back(static_cast<uint32_t>(s->shape.size()));
    s->info.n_worlds = static_cast<uint32_t>(s->first_body.size() - 1);
    s->info.n_shapes = static_cast<uint32_t>(s->shape_offsets.size() - 1);
    s->info.n_verts = static_cast<uint32_t>(s->verts.size() / 3);
    s->info.n_bodies = static_cast<uint32_t>(s->shape.size());
    uint32_t ns = 0;
    for (auto v : s->is_static) ns += v;
    s->info.n_static = ns;
    *out = s;
    return PKG_OK;
}

pkg_status pkg_scene_info_get(const pkg_scene *s, pkg_scene_info *info)
{
    if (!s || !info) return PKG_INVALID_ARG;
    *info = s->info;
    return PKG_OK;
}

pkg_status pkg_scene_copy(const pkg_scene *s, uint32_t *shape_offsets, double *verts,
                          uint8_t *is_static, uint32_t *shape, double *pos, double *vel,
                          double *angvel, double *quat, double *mass, double *inertia,
                          double *restitution, double *friction)
{
    if (!s) return PKG_INVALID_ARG;
    auto cp = [](auto *dst, const auto &v)
    {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(shape_offsets, s->shape_offsets);
    cp(verts, s->verts);
    cp(is_static, s->is_static);
    cp(shape, s->shape);
    cp(pos, s->pos);
    cp(vel, s->vel);
    cp(angvel, s->angvel);
    cp(quat, s->quat);
    cp(mass, s->mass);
    cp(inertia, s->inertia);
    cp(restitution, s->restitution);
    cp(friction, s->friction);
    return PKG_OK;
}

pkg_status pkg_scene_copy_worlds(const pkg_scene *s, uint32_t *first_body)
{
    if (!s || !first_body) return PKG_INVALID_ARG;
    std::memcpy(first_body, s->first_body.data(), s->first_body.size() * sizeof(uint32_t));
    return PKG_OK;
}

pkg_status pkg_scene_destroy(pkg_scene *s)
{
    delete s;
    return PKG_OK;
}
}
