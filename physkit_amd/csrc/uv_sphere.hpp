// Recognises meshes that are UV spheres as mesh::sphere(r, stacks, sectors) builds them
// (src/mesh.cpp:60-115): vertex 0 = (0, r, 0), then ring i = 1..stacks-1 with sector
// j = 0..sectors-1 at (r sin(phi) cos(th), r cos(phi), r sin(phi) sin(th)),
// phi = pi i / stacks, th = 2 pi j / sectors, then (0, -r, 0). The device support of
// such a mesh evaluates only the vertices around the analytic optimum and certifies
// the result (pk_gjk.hpp, uv_support), so a mesh is tagged only if every vertex
// equals that construction bit for bit. Host code.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

namespace pk
{
class uv_sphere_detector
{
public:
    // (stacks << 16) | sectors for a UV sphere with at most 1024 stacks and sectors
    // (the device bounds assume cells of at least pi/1024), 0 otherwise
    uint32_t tag(const double *v, uint32_t nv)
    {
        if (nv < 5) return 0;
        const double r = v[1];
        if (!(r > 0.0) || !std::isfinite(r) || v[0] != 0.0 || v[2] != 0.0) return 0;
        const double *last = v + 3 * static_cast<size_t>(nv - 1);
        if (last[0] != 0.0 || last[1] != -r || last[2] != 0.0) return 0;
        const uint32_t m = nv - 2;
        for (uint32_t t = 3; t <= m && t <= 1024u; ++t)
        {
            if (m % t) continue;
            uint32_t s = m / t + 1;
            if (s < 2 || s > 1024u) continue;
            if (matches(v, r, s, t)) return (s << 16) | t;
        }
        return 0;
    }

private:
    struct tables
    {
        std::vector<double> sphi, cphi, cth, sth;
    };
    std::map<std::pair<uint32_t, uint32_t>, tables> M_tab;

    const tables &tab(uint32_t s, uint32_t t)
    {
        auto it = M_tab.find({s, t});
        if (it != M_tab.end()) return it->second;
        const double pi = 3.141592653589793;
        tables tb;
        for (uint32_t i = 0; i < s; ++i)
        {
            double phi = pi * static_cast<double>(i) / static_cast<double>(s);
            tb.sphi.push_back(std::sin(phi));
            tb.cphi.push_back(std::cos(phi));
        }
        for (uint32_t j = 0; j < t; ++j)
        {
            double th = 2.0 * pi * static_cast<double>(j) / static_cast<double>(t);
            tb.cth.push_back(std::cos(th));
            tb.sth.push_back(std::sin(th));
        }
        return M_tab.emplace(std::make_pair(s, t), std::move(tb)).first->second;
    }

    bool matches(const double *v, double r, uint32_t s, uint32_t t)
    {
        const tables &tb = tab(s, t);
        for (uint32_t i = 1; i < s; ++i)
        {
            const double rs = r * tb.sphi[i], y = r * tb.cphi[i];
            for (uint32_t j = 0; j < t; ++j)
            {
                const double *p = v + 3 * (1 + static_cast<size_t>(i - 1) * t + j);
                const double w[3] = {rs * tb.cth[j], y, rs * tb.sth[j]};
                if (std::memcmp(p, w, sizeof(w)) != 0) return false;
            }
        }
        return true;
    }
};
} // namespace pk
