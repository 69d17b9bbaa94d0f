// titan::Simulation — C++ drop-in for the reference's Simulation
// (include/Titan/sim.h:37-220) over the MI355X engine's C ABI
// (include/titan_mi355x.h).
//
// The scene-building API, the breakpoint rule and the result conventions
// are the reference's.  What differs is the control model: the reference
// runs the step loop on a background thread and the user thread polls
// flags (src/sim.cu:1389,1503-1657); here every call that lets simulated
// time pass (pause, wait, waitUntil, waitForEvent, stop(t), and the added
// step(n)) advances the simulation on the calling thread through the
// engine's stream-ordered stepping and returns when the engine has reached
// the requested time.  get/getAll read the engine's state; set/setAll write
// host objects back, rebuilding the device body when topology, constraints
// or spring parameters changed.
#ifndef TITAN_MI355X_SIM_H
#define TITAN_MI355X_SIM_H

#include <array>
#include <cstdint>
#include <set>
#include <string>
#include <vector>

#include "mass.h"
#include "object.h"
#include "spring.h"
#include "vec.h"

struct titan_ctx;

namespace titan {

class Simulation {
   public:
    // Create
    Mass *createMass();
    Mass *createMass(const Vec &pos);
    Spring *createSpring();
    Spring *createSpring(Mass *m1, Mass *m2);

    // Delete
    void deleteMass(Mass *m);
    void deleteSpring(Spring *s);

    void get(Mass *m);
    void get(Spring *s);  // springs are never changed by the step
    void get(Container *c);
    void set(Mass *m);
    void set(Spring *s);
    void set(Container *c);
    void getAll();
    void setAll();

    // Global constraints
    void createPlane(const Vec &abc, double d);
    void createPlane(const Vec &abc, double d, double FRICTION_K, double FRICTION_S);
    void createBall(const Vec &center, double r);
    void clearConstraints();  // clears the host shapes only, as the reference does

    // Containers
    Container *createContainer();
    void deleteContainer(Container *c);
    Cube *createCube(const Vec &center, double side_length);
    Lattice *createLattice(const Vec &center, const Vec &dims, int nx = 10, int ny = 10, int nz = 10);
    Beam *createBeam(const Vec &center, const Vec &dims, int nx = 10, int ny = 10, int nz = 10);
    Container *importFromSTL(const std::string &path, double density = 10.0, int num_rays = 5);

    // Bulk modifications (host objects)
    void setAllSpringConstantValues(double k);
    void setAllMassValues(double m);  // adds m, as the reference does (src/sim.cu:788-796)
    void setTimeStep(double delta_t);
    double getTimeStep();
    void setGlobalAcceleration(const Vec &global_acc);
    void defaultRestLengths();

    // Control
    void start();
    void stop();
    void stop(double time);
    void pause(double t);
    void resume();
    void reset();
    void setBreakpoint(double time);
    void wait(double t);
    void waitUntil(double t);
    void waitForEvent();
    // Advance exactly n steps (not in the reference, whose only time control
    // is breakpoints).
    void step(int64_t n);

    double time();
    bool running();
    void printPositions();

    // The integrator follows the reference's build options as the program
    // sees them: RK2, else VERLET, else the Euler branch (CMakeLists.txt:13-14,
    // src/sim.cu:1141-1167).  Simulation(k) picks one explicitly (0 Euler,
    // 1 Verlet, 2 RK2).
#if defined(RK2)
    Simulation() : Simulation(2) {}
#elif defined(VERLET)
    Simulation() : Simulation(1) {}
#else
    Simulation() : Simulation(0) {}
#endif
    explicit Simulation(int integrator);
    ~Simulation();

    Spring *getSpringByIndex(int i);
    Mass *getMassByIndex(int i);
    Container *getContainerByIndex(int i);

    std::vector<Mass *> masses;
    std::vector<Spring *> springs;
    std::vector<Container *> containers;

   private:
    Mass *createMass(Mass *m);
    Spring *createSpring(Spring *s);

    void write_record(Mass *m);    // CUDA_MASS(Mass&) into m's device record
    void write_record(Spring *s);  // CUDA_SPRING(Spring&) into s's device record
    void pull_state();             // engine -> device records (pos, vel, acc, T)
    void push_state();             // device records -> engine (pos, vel, m, T, ext)
    void rebuild();                // re-create the engine body from the records
    // apply deletions to the built body in place; false: it must be rebuilt
    bool remove_in_place(const std::vector<Mass *> &ms, const std::vector<Spring *> &ss);
    void advance(int64_t n);       // rebuild if needed, then n engine steps
    int64_t steps_until_after(double t, bool strict, bool &hit_breakpoint);
    void check_alive(const char *what);

    double dt;
    double T;
    Vec _global_acc;
    bool RUNNING, STARTED, ENDED, FREED;
    std::vector<Constraint *> constraints;                 // host shapes
    std::vector<std::array<double, 6>> dplanes;            // what the engine applies
    std::vector<std::array<double, 4>> dballs;
    std::set<double> bpts;
    std::vector<Mass *> retired_masses;
    std::vector<Spring *> retired_springs;

    titan_ctx *ctx = nullptr;
    int32_t body = -1;
    int integrator_ = 0;
    void advance_clock(double &t) const;
    bool dirty = false;  // the engine body must be rebuilt before the next step
};

}  // namespace titan

#endif
