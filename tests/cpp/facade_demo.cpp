// Drives the physkit_amd C++ facade the way PhysKit user code drives physkit::world.
// Scenario 1 follows tests/co/co_tests.cpp:358-429: two boxes fly into each other with
// no gravity at 5 ms steps; the collision-begin callback fires once, then the
// collision-end callback after they separate. Scenario 2 drops a box on a static box
// and checks it comes to rest. Scenario 3 removes a body and reuses its slot.
// Exit code 0 on success.
#include "physkit_amd/world.hpp"

#include <cmath>
#include <cstdio>

using namespace physkit_amd;

static int fail(const char *what)
{
    std::printf("FAIL: %s\n", what);
    return 1;
}

int main()
{
    std::setvbuf(stdout, nullptr, _IONBF, 0);
    {
        world w(world_desc::make().with_gravity({0, 0, 0}));
        auto box = mesh::box({0.5, 0.5, 0.5});
        handle a = w.create_rigid(object_desc::dynam().with_mesh(box).with_pos({-2, 0, 0}).with_vel({2, 0, 0}));
        handle b = w.create_rigid(object_desc::dynam().with_mesh(box).with_pos({2, 0, 0}).with_vel({-2, 0, 0}));
        int begins = 0, ends = 0, begin_step = -1, end_step = -1, step = 0;
        w.on_collision = [&](handle x, handle y)
        {
            if ((x == a && y == b) || (x == b && y == a))
            {
                ++begins;
                if (begin_step < 0) begin_step = step;
            }
        };
        w.on_collision_exit = [&](handle, handle)
        {
            ++ends;
            if (end_step < 0) end_step = step;
        };
        for (step = 0; step < 400; ++step) w.step(0.005);
        std::printf("head-on: begins=%d (step %d) ends=%d (step %d)\n", begins, begin_step, ends, end_step);
        if (begins < 1) return fail("no collision-begin event");
        if (ends < 1 || end_step < begin_step) return fail("no collision-end event after the begin");
        auto sa = w.get_rigid(a), sb = w.get_rigid(b);
        if (!sa || !sb) return fail("handles invalid");
        if (!(sa->vel().x < 0 && sb->vel().x > 0)) return fail("boxes did not bounce apart");
    }
    {
        world w(world_desc::make().with_epa_guard(true));
        auto ground = mesh::box({5, 0.5, 5});
        auto box = mesh::box({0.5, 0.5, 0.5});
        w.create_rigid(object_desc::stat().with_mesh(ground).with_pos({0, -0.5, 0}));
        mat3 I = {1.0 / 6, 0, 0, 0, 1.0 / 6, 0, 0, 0, 1.0 / 6};
        handle h = w.create_rigid(object_desc::dynam().with_mesh(box).with_pos({0, 0.6, 0}).with_inertia_tensor(I)
                                      .with_restitution(0.0));
        for (int k = 0; k < 480; ++k) w.step(1.0 / 120.0);
        auto s = w.get_rigid(h);
        std::printf("resting box: y=%.4f |v|=%.2e\n", s->pos().y,
                    std::sqrt(s->vel().x * s->vel().x + s->vel().y * s->vel().y + s->vel().z * s->vel().z));
        if (!(s->pos().y > 0.45 && s->pos().y < 0.52)) return fail("box not resting on the ground");
    }
    {
        world w(world_desc::make());
        auto box = mesh::box({0.5, 0.5, 0.5});
        handle a = w.create_rigid(object_desc::dynam().with_mesh(box).with_pos({0, 10, 0}));
        handle b = w.create_rigid(object_desc::dynam().with_mesh(box).with_pos({5, 10, 0}));
        w.step(1.0 / 60.0);
        if (!w.remove_rigid(a)) return fail("remove_rigid failed");
        if (w.get_rigid(a)) return fail("stale handle still valid");
        handle c = w.create_rigid(object_desc::dynam().with_mesh(box).with_pos({-5, 10, 0}));
        if (c.index() != a.index() || c.generation() == a.generation()) return fail("slot not reused");
        for (int k = 0; k < 30; ++k) w.step(1.0 / 60.0);
        auto sb = w.get_rigid(b), sc = w.get_rigid(c);
        if (!sb || !sc || !(sc->pos().y < 10.0) || !(sb->pos().y < 10.0)) return fail("bodies did not fall");
        std::printf("slot reuse: ok (index %u, generation %u)\n", c.index(), c.generation());
    }
    std::printf("facade ok\n");
    return 0;
}
