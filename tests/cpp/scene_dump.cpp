// Reads scene files with the C++ reader (include/physkit_amd/scene_json.hpp) and prints
// what it built, one line per world and per object, every double with 17 significant
// digits, for tests/test_scene_json.py to compare with the Python reader.
#include "physkit_amd/scene_json.hpp"

#include <cstdio>

using namespace physkit_amd;

int main(int argc, char **argv)
{
    for (int a = 1; a < argc; ++a)
    {
        try
        {
            scene_json::scene_file s = scene_json::read_scene(argv[a]);
            const vec3 g = s.desc.gravity();
            std::printf("world %.17g %.17g %.17g %u %zu\n", g.x, g.y, g.z, s.desc.solver_iterations(), s.objects.size());
            for (const object_desc &d : s.objects)
            {
                const vec3 p = d.pos(), v = d.vel(), w = d.ang_vel();
                const quat q = d.orientation();
                const mat3 &I = d.inertia_tensor();
                std::printf("object %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g "
                            "%.17g %.17g %.17g %.17g %.17g %.17g",
                            d.type() == body_type::stat ? 1 : 0, p.x, p.y, p.z, v.x, v.y, v.z, q.w, q.x, q.y, q.z, w.x,
                            w.y, w.z, I[0], I[4], I[8], d.type() == body_type::stat ? 0.0 : d.mass(), d.restitution(),
                            d.friction());
                const auto &vs = d.mesh_ptr()->vertices();
                std::printf(" %zu", vs.size());
                for (const vec3 &x : vs) std::printf(" %.17g %.17g %.17g", x.x, x.y, x.z);
                std::printf("\n");
            }
        }
        catch (const std::exception &e)
        {
            std::printf("error %s\n", e.what());
        }
    }
    return 0;
}
