// The reference's physics tests (test/physics/*.cpp), restated against the
// C++ facade (titan_amd/cpp/Titan/sim.h) running on the MI355X engine, plus
// a parity scene whose final state is printed for comparison with the CPU
// oracle.  Usage: facade_tests <test name>; prints PASS/FAIL lines and exits
// non-zero on failure.
#include <Titan/sim.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <string>

using titan::Vec;

static int failures = 0;
#define EXPECT_NEAR(a, b, tol)                                                                       \
    do {                                                                                             \
        const double a_ = (a), b_ = (b), t_ = (tol);                                                 \
        if (!(std::fabs(a_ - b_) <= t_)) {                                                           \
            std::printf("FAIL %s:%d |%.17g - %.17g| > %.3g\n", __FILE__, __LINE__, a_, b_, t_);      \
            ++failures;                                                                              \
        }                                                                                            \
    } while (0)
#define EXPECT_TRUE(c)                                                     \
    do {                                                                   \
        if (!(c)) {                                                        \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);        \
            ++failures;                                                    \
        }                                                                  \
    } while (0)

static const double NORMAL = 20000.0;

// test/testutil/utils.h:12-30 with the gravity used by the scene, plus the
// penalty energy NORMAL disp^2 / 2 of masses below the plane z = 0, which
// the reference's helper leaves out (a sample inside a contact would
// otherwise drop by the stored penalty energy).
static double energy(titan::Simulation &sim, double g) {
    double potential_g = 0, potential_s = 0, kinetic = 0, penalty = 0;
    sim.getAll();
    for (titan::Mass *m : sim.masses) {
        potential_g += g * (m->pos)[2] * (m->m);
        kinetic += 0.5 * (m->m) * std::pow((m->vel).norm(), 2);
        if (m->pos[2] < 0) penalty += 0.5 * NORMAL * m->pos[2] * m->pos[2];
    }
    for (titan::Spring *s : sim.springs)
        potential_s += s->_k * std::pow((s->_left->pos - s->_right->pos).norm() - (s->_rest), 2) / 2;
    return potential_s + kinetic + potential_g + penalty;
}

static bool finite_state(titan::Simulation &sim) {
    sim.getAll();
    for (titan::Mass *m : sim.masses)
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(m->pos[k]) || !std::isfinite(m->vel[k])) return false;
    return true;
}

// simple_unittest.cpp:15-38
static void simple_test() {
    titan::Simulation sim;
    sim.createMass(Vec(1, 0, 1));
    sim.setTimeStep(0.0001);
    sim.setGlobalAcceleration(Vec(0, 0, -9.8));
    sim.createPlane(Vec(0, 0, 1), 0);
    sim.start();
    const double total = energy(sim, 9.8);
    double avg = total;
    const double alpha = 0.9;
    while (sim.time() < 5) {
        sim.wait(0.1);
        avg = (1 - alpha) * energy(sim, 9.8) + alpha * avg;
        EXPECT_NEAR(avg, total, total * 0.01);
        sim.resume();
    }
    sim.stop();
}

// energy_unittest.cpp:14-40
static void energy_test() {
    titan::Simulation sim;
    sim.createLattice(Vec(0, 0, 5), Vec(4, 4, 4), 20, 20, 20);
    sim.setAllSpringConstantValues(1000);
    sim.setTimeStep(0.0001);
    sim.setGlobalAcceleration(Vec(0, 0, -9.8));
    sim.defaultRestLengths();
    sim.createPlane(Vec(0, 0, 1), 0);
    sim.start();
    const double total = energy(sim, 9.8);
    double avg = total;
    const double alpha = 0.7;
    while (sim.time() < 5) {
        sim.wait(0.1);
        avg = (1 - alpha) * energy(sim, 9.8) + alpha * avg;
        EXPECT_NEAR(avg, total, total * 0.01);
        sim.resume();
    }
    sim.stop();
}

// external_unittest.cpp:17-57
static void external_test(bool via_gravity) {
    titan::Simulation sim;
    titan::Mass *m1 = sim.createMass(Vec(1, 0, 1));
    sim.setTimeStep(0.0001);
    if (via_gravity) {
        m1->setExternalForce(m1->m * Vec(0, 0, 0));
        sim.setGlobalAcceleration(Vec(0, 0, -9.8));
    } else {
        m1->setExternalForce(m1->m * Vec(0, 0, -9.8));
        sim.setGlobalAcceleration(Vec(0, 0, 0.0));
    }
    sim.start();
    const double horizon = via_gravity ? 1 : 5;
    while (sim.time() < horizon) {
        sim.wait(0.1);
        sim.getAll();
        EXPECT_NEAR(m1->pos[0], 1, 0.000001);
        EXPECT_NEAR(m1->pos[1], 0, 0.000001);
        EXPECT_NEAR(m1->pos[2], 1 + 0.5 * -9.8 * std::pow(sim.time(), 2), 0.1);
        sim.resume();
    }
    sim.stop();
}

// multiagent_unittest.cpp:21-71 (a smoke test in the reference)
static void multiagent_test() {
    const double SIZE = 5.0, SPACE = 3.0;
    const int NX = 10, NY = 10, DENSITY = 5;
    titan::Simulation sim;
    sim.setGlobalAcceleration(Vec(0, 0, -9.8));
    titan::Lattice *grid[NY][NX];
    for (int i = 0; i < NY; i++)
        for (int j = 0; j < NX; j++)
            grid[i][j] = sim.createLattice(Vec((SIZE + SPACE) * j + SIZE / 2, (SIZE + SPACE) * i + SIZE / 2, SIZE / 2),
                                           Vec(SIZE, SIZE, SIZE), DENSITY, DENSITY, DENSITY);
    const int right[4][2] = {{100, 0}, {120, 20}, {104, 4}, {124, 24}};
    const int up[4][2] = {{20, 0}, {120, 100}, {124, 104}, {24, 4}};
    for (int i = 0; i < NY - 1; i++)
        for (int j = 0; j < NX - 1; j++) {
            for (auto &p : right) {
                titan::Spring *s = sim.createSpring(grid[i][j]->masses[p[0]], grid[i][j + 1]->masses[p[1]]);
                s->_k = 0.01;
                s->defaultLength();
            }
            for (auto &p : up) {
                titan::Spring *s = sim.createSpring(grid[i][j]->masses[p[0]], grid[i + 1][j]->masses[p[1]]);
                s->_k = 0.01;
                s->defaultLength();
            }
        }
    sim.defaultRestLengths();
    sim.start();
    sim.pause(1.0);
    sim.getAll();
    EXPECT_TRUE(sim.time() >= 1.0);
    EXPECT_TRUE(finite_state(sim));
    sim.stop();
}

// rotate_unittest.cpp:11-41 (a smoke test in the reference): teleport a
// body through get/rotate/set every simulated second
static void rotate_test() {
    titan::Simulation sim;
    titan::Lattice *l2 = sim.createLattice(Vec(0, 0, 10), Vec(5, 5, 5), 10, 10, 10);
    sim.setAllSpringConstantValues(1E5);
    l2->rotate(Vec(0, 0, 1), -0.78);
    sim.createPlane(Vec(0, 0, 1), 0, 10, 10);
    sim.setGlobalAcceleration(Vec(0, 0, -9.8));
    sim.start();
    int rounds = 0;
    while (true) {
        sim.pause(sim.time() + 1);
        sim.get(l2);
        const Vec before = l2->masses[0]->pos;
        l2->rotate(Vec(0, 0, 1), 0.5);
        sim.set(l2);
        sim.get(l2);
        EXPECT_TRUE(!(l2->masses[0]->pos == before));  // the rotation reached the device
        ++rounds;
        if (sim.time() > 5.0) {
            EXPECT_TRUE(finite_state(sim));
            sim.stop();
            break;
        }
        sim.resume();
    }
    EXPECT_TRUE(rounds == 5);
}

// The README scene with its bottom face in contact, paused at t = 0.0105:
// the final state is printed as hex floats for the oracle comparison.
static void readme_parity() {
    titan::Simulation sim;
    titan::Lattice *l = sim.createLattice(Vec(0, 0, 2.45), Vec(5, 5, 5), 5, 5, 5);
    for (titan::Spring *s : l->springs) s->addDamping(0.1);
    sim.createPlane(Vec(0, 0, 1), 0, 0.8, 1.0);
    sim.start();
    sim.pause(0.0105);
    sim.getAll();
    std::printf("T %a\n", sim.time());
    for (titan::Mass *m : sim.masses)
        std::printf("M %a %a %a %a %a %a %a\n", m->pos[0], m->pos[1], m->pos[2], m->vel[0], m->vel[1], m->vel[2], m->T);
    sim.stop();
}

// Deletions mid-run: a 10x10x10 lattice and a 3x3x3 one on the ground plane
// step 60 steps; then every 37th spring of the first lattice, two of its
// masses and the whole second lattice (deleteContainer) are deleted, and the
// scene steps 90 more.  The live masses' final state is printed as hex
// floats for the oracle comparison.
static void delete_parity() {
    titan::Simulation sim;
    titan::Lattice *l = sim.createLattice(Vec(0, 0, 0.45), Vec(1, 1, 1), 10, 10, 10);
    for (titan::Spring *s : l->springs) s->addDamping(0.1);
    titan::Lattice *l2 = sim.createLattice(Vec(3, 0, 0.45), Vec(1, 1, 1), 3, 3, 3);
    for (titan::Spring *s : l2->springs) s->addDamping(0.1);
    sim.createPlane(Vec(0, 0, 1), 0, 0.8, 1.0);
    sim.start();
    sim.step(60);
    sim.pause(sim.time());  // the first step boundary with T >= T: no further step
    std::vector<titan::Spring *> del;
    for (size_t k = 0; k < l->springs.size(); k += 37) del.push_back(l->springs[k]);
    for (titan::Spring *s : del) sim.deleteSpring(s);
    sim.deleteMass(l->masses[123]);
    sim.deleteMass(l->masses[456]);
    sim.deleteContainer(l2);
    sim.resume();
    sim.step(90);
    sim.getAll();
    std::printf("T %a\n", sim.time());
    for (titan::Mass *m : sim.masses)
        std::printf("M %a %a %a %a %a %a %a\n", m->pos[0], m->pos[1], m->pos[2], m->vel[0], m->vel[1], m->vel[2], m->T);
    sim.stop();
}

// start / breakpoint / step / set-get round trips of the facade itself
static void control_test() {
    titan::Simulation sim;
    titan::Mass *a = sim.createMass(Vec(0, 0, 1));
    titan::Mass *b = sim.createMass(Vec(0, 0, 2));
    sim.createSpring(a, b);
    sim.setGlobalAcceleration(Vec(0, 0, 0));
    sim.start();
    sim.step(10);
    EXPECT_NEAR(sim.time(), 10 * 0.0001, 1e-15);
    sim.setBreakpoint(0.002);
    sim.setBreakpoint(0.003);
    sim.waitForEvent();
    EXPECT_TRUE(sim.time() >= 0.002 && sim.time() < 0.002 + 0.0001 + 1e-12);
    EXPECT_TRUE(!sim.running());
    sim.resume();
    sim.waitForEvent();
    EXPECT_TRUE(sim.time() >= 0.003);
    // a mass added while paused joins the body; like the reference's
    // createMass (src/sim.cu:236-252) it is copied at creation, so a force set
    // afterwards reaches the device through set()
    titan::Mass *c = sim.createMass(Vec(1, 0, 1));
    c->setExternalForce(Vec(0.1, 0, 0));
    sim.set(c);
    sim.resume();
    sim.wait(0.01);
    sim.get(c);
    EXPECT_TRUE(c->pos[0] > 1.0);
    EXPECT_TRUE(c->acceleration()[0] == 0.1 / c->m);  // extern force / m with g = 0
    // fix() followed by set() stops a mass
    sim.get(b);
    b->fix();
    sim.set(b);
    const Vec pb = b->pos;
    sim.step(100);
    sim.get(b);
    EXPECT_TRUE(b->pos == pb);
    sim.stop();
    bool threw = false;
    try {
        sim.step(1);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    EXPECT_TRUE(threw);
}

int main(int argc, char **argv) {
    std::map<std::string, std::function<void()>> tests = {
        {"simple", simple_test},
        {"energy", energy_test},
        {"external", [] { external_test(false); }},
        {"external_acc", [] { external_test(true); }},
        {"multiagent", multiagent_test},
        {"rotate", rotate_test},
        {"readme_parity", readme_parity},
        {"delete_parity", delete_parity},
        {"control", control_test},
    };
    if (argc >= 2 && std::string(argv[1]) == "all") {  // every test except the parity dump
        int bad = 0;
        for (auto &t : tests) {
            if (t.first == "readme_parity" || t.first == "delete_parity") continue;
            failures = 0;
            try {
                t.second();
            } catch (const std::exception &e) {
                std::printf("FAIL exception in %s: %s\n", t.first.c_str(), e.what());
                ++failures;
            }
            std::printf(failures ? "FAILED %s (%d)\n" : "PASS %s\n", t.first.c_str(), failures);
            std::fflush(stdout);
            bad += failures ? 1 : 0;
        }
        return bad ? 1 : 0;
    }
    if (argc < 2 || !tests.count(argv[1])) {
        std::printf("usage: %s <test>; tests:", argv[0]);
        for (auto &t : tests) std::printf(" %s", t.first.c_str());
        std::printf("\n");
        return 2;
    }
    try {
        tests[argv[1]]();
    } catch (const std::exception &e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 1;
    }
    std::printf(failures ? "FAILED %s (%d)\n" : "PASS %s\n", argv[1], failures);
    return failures ? 1 : 0;
}
