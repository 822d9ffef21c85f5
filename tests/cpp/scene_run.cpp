// A scene file through the C++ reader (include/physkit_amd/scene_json.hpp) into the C++
// facade world, stepped on the GPU: argv = path, steps, dt. Prints every body's position
// and orientation after the last step, 17 significant digits, one body per line.
#include "physkit_amd/scene_json.hpp"

#include <cstdio>
#include <cstdlib>

using namespace physkit_amd;

int main(int argc, char **argv)
{
    if (argc != 4) return 2;
    try
    {
        scene_json::scene_file s = scene_json::read_scene(argv[1]);
        world w(s.desc);
        std::vector<handle> hs = scene_json::populate(w, s);
        const int steps = std::atoi(argv[2]);
        const double dt = std::atof(argv[3]);
        for (int k = 0; k < steps; ++k) w.step(dt);
        for (handle h : hs)
        {
            auto o = w.get_rigid(h);
            const vec3 p = o->pos();
            const quat q = o->orientation();
            std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", p.x, p.y, p.z, q.w, q.x, q.y, q.z);
        }
    }
    catch (const std::exception &e)
    {
        std::printf("error %s\n", e.what());
        return 1;
    }
    return 0;
}
