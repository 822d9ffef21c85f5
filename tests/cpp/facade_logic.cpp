// Host logic of the C++ facade (handles, slot reuse, re-upload keeping the state
// reached so far, state download) against a CPU test double of the C ABI
// (stub_pkg.cpp). Exit code 0 on success.
#include "physkit_amd/world.hpp"

#include <cmath>
#include <cstdio>

using namespace physkit_amd;

#define REQUIRE(c)                                                                                 \
    do                                                                                             \
    {                                                                                              \
        if (!(c))                                                                                  \
        {                                                                                          \
            std::printf("FAIL line %d: %s\n", __LINE__, #c);                                       \
            return 1;                                                                              \
        }                                                                                          \
    } while (0)

int main()
{
    std::setvbuf(stdout, nullptr, _IONBF, 0);
    world w(world_desc::make().with_gravity({0, -10, 0}));
    auto box = mesh::box({0.5, 0.5, 0.5});
    handle a = w.create_rigid(object_desc::dynam().with_mesh(box).with_pos({0, 10, 0}));
    handle b = w.create_rigid(object_desc::dynam().with_mesh(box).with_pos({5, 10, 0}));
    handle g = w.create_rigid(object_desc::stat().with_mesh(box).with_pos({0, -1, 0}));
    REQUIRE(a.index() == 0 && b.index() == 1 && g.index() == 2);
    w.step(0.1);
    REQUIRE(w.counters().pairs == 1); // one upload so far
    REQUIRE(std::fabs(w.get_rigid(a)->pos().y - 9.9) < 1e-12);
    REQUIRE(std::fabs(w.get_rigid(b)->vel().y + 1.0) < 1e-12);
    REQUIRE(w.get_rigid(g)->pos().y == -1.0);

    // remove a; b keeps its state across the re-upload
    REQUIRE(w.remove_rigid(a));
    REQUIRE(!w.remove_rigid(a));
    REQUIRE(!w.get_rigid(a));
    w.step(0.1);
    REQUIRE(w.counters().pairs == 2);
    REQUIRE(std::fabs(w.get_rigid(b)->pos().y - (10.0 - 0.1 - 0.2)) < 1e-12);

    // the freed slot is reused with a new generation and starts from its own desc
    handle c = w.create_rigid(object_desc::dynam().with_mesh(box).with_pos({-5, 20, 0}));
    REQUIRE(c.index() == a.index() && c.generation() == a.generation() + 1);
    REQUIRE(!w.get_rigid(a));
    auto sc = w.get_rigid(c); // uploads, then downloads
    REQUIRE(sc && sc->pos().x == -5.0 && sc->pos().y == 20.0 && sc->vel().y == 0.0);
    REQUIRE(std::fabs(w.get_rigid(b)->pos().y - 9.7) < 1e-12);
    w.step(0.1);
    REQUIRE(std::fabs(w.get_rigid(c)->pos().y - 19.9) < 1e-12);
    REQUIRE(std::fabs(w.get_rigid(b)->pos().y - (9.7 - 0.3)) < 1e-12);

    // a new slot appended after upload
    handle d = w.create_rigid(object_desc::dynam().with_mesh(box).with_pos({1, 2, 3}));
    REQUIRE(d.index() == 3);
    w.step(0.1);
    REQUIRE(std::fabs(w.get_rigid(d)->pos().y - 1.9) < 1e-12);
    REQUIRE(std::fabs(w.get_rigid(c)->pos().y - (19.9 - 0.2)) < 1e-12);

    // joints: one arena per type, LIFO slot reuse with generations; built at the next step
    joint_handle j0 = w.add_constraint(ball_socket_constraint::desc::make(b, c).with_anchor({0, 5, 0}));
    joint_handle j1 = w.add_constraint(ball_socket_constraint::desc::make(c, d).with_anchor({0, 5, 0}));
    joint_handle j2 = w.add_constraint(distance_constraint::desc::make(b, d).with_distance(2.0));
    REQUIRE(j0.index == 0 && j1.index == 1 && j2.index == 0);
    REQUIRE(j0.key() == PKG_JOINT_KEY(PKG_JOINT_BALL_SOCKET, 0) && j2.key() == 0u);
    REQUIRE(!w.joint_impulses(j0)); // not built before a step
    w.step(0.1);
    REQUIRE(w.joint_impulses(j0) && w.joint_impulses(j2));
    REQUIRE(w.remove_constraint(j0));
    REQUIRE(!w.remove_constraint(j0));
    REQUIRE(!w.joint_impulses(j0));
    joint_handle j3 = w.add_constraint(ball_socket_constraint::desc::make(b, d).with_anchor({1, 1, 1}));
    REQUIRE(j3.index == 0 && j3.generation == 1); // the freed slot, next generation
    joint_handle j4 = w.add_constraint(weld_constraint::desc::make(b, c).with_anchor({1, 1, 1}));
    REQUIRE(w.remove_constraint(j4)); // removed before it was built: never reaches the device
    w.step(0.1);
    REQUIRE(w.joint_impulses(j3) && w.joint_impulses(j1) && !w.joint_impulses(j4));

    bool threw = false;
    try
    {
        w.add_constraint(distance_constraint::desc::make(a, b)); // a was removed
    }
    catch (const std::invalid_argument &)
    {
        threw = true;
    }
    REQUIRE(threw);
    threw = false;
    try
    {
        w.create_rigid(object_desc::dynam());
    }
    catch (const std::invalid_argument &)
    {
        threw = true;
    }
    REQUIRE(threw);

    // a step that fails part-way: the facade throws and refuses to step again
    threw = false;
    try
    {
        w.step(1234.5);
    }
    catch (const std::runtime_error &)
    {
        threw = true;
    }
    REQUIRE(threw);
    threw = false;
    try
    {
        w.step(0.1);
    }
    catch (const std::runtime_error &)
    {
        threw = true;
    }
    REQUIRE(threw);
    std::printf("facade logic ok\n");
    return 0;
}
