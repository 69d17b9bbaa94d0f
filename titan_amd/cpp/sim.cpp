// C++ facade: titan::Simulation and friends over the MI355X engine's C ABI.
//
// Scene building, containers and constraints restate the reference's host
// code (src/object.cu:26-360, src/mass.cu, src/spring.cu, src/sim.cu:77-340,
// 565-807,1319-1660,1826-2087); the step itself runs in libtitan_mi355x.so.
// Every Mass and Spring carries a copy of what was last written to the
// device (the reference's CUDA_MASS / CUDA_SPRING record), so get/set keep
// the reference's split between host objects and device state.
#include "Titan/sim.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <limits>
#include <stdexcept>

#include "titan_mi355x.h"

namespace titan {

namespace {

void check(titan_ctx *ctx, int rc) {
    if (rc != TITAN_OK) throw std::runtime_error(std::string("titan engine: ") + titan_last_error(ctx));
}

template <class T>
void erase_one(std::vector<T *> &v, T *x) {
    v.erase(std::remove(v.begin(), v.end(), x), v.end());
}

}  // namespace

// ----------------------------------------------------------------- Mass

Mass::Mass() : m(1.0), T(0) {}  // src/mass.cu:8-19

Mass::Mass(const Vec &position, double mass, bool /*fixed: ignored, src/mass.cu:42-55*/)
    : m(mass), T(0), pos(position) {}

// src/mass.cu:78-96 with the constructors of src/object.cu:58-63,108-131
void Mass::addConstraint(CONSTRAINT_TYPE type, const Vec &vec, double num) {
    Local l;
    l.kind = (int)type;
    l.num = num;
    if (type == BALL) {
        l.v = vec;
    } else {
        if (vec.norm() == 0.0) throw std::runtime_error("addConstraint: zero normal or tangent");
        l.v = vec / vec.norm();
    }
    locals.push_back(l);
}

void Mass::clearConstraints(CONSTRAINT_TYPE type) {
    locals.erase(std::remove_if(locals.begin(), locals.end(), [&](const Local &l) { return l.kind == (int)type; }),
                 locals.end());
}

void Mass::clearConstraints() { locals.clear(); }
void Mass::setDrag(double C) { drag_coefficient = C; }
void Mass::fix() { fixed = true; }
void Mass::unfix() { fixed = false; }

// ------------------------------------------------------------ Container

void Container::translate(const Vec &displ) {
    for (Mass *m : masses) m->pos += displ;
}

// src/object.cu:198-230
void Container::rotate(const Vec &axis, double angle) {
    Vec com(0, 0, 0);
    double total_mass = 0;
    for (Mass *m : masses) {
        com += m->m * m->pos;
        total_mass += m->m;
    }
    com = com / total_mass;
    Vec temp_axis = axis / axis.norm();
    for (Mass *m : masses) {
        Vec temp = m->pos - com;
        Vec y = temp - dot(temp, temp_axis) * temp_axis;
        if (y.norm() < 0.0001) continue;
        Vec planar(-std::sin(angle) * y.norm(), std::cos(angle) * y.norm(), 0);
        Vec spatial = planar[0] * cross(temp_axis, y / y.norm()) + y / y.norm() * planar[1] +
                      dot(temp, temp_axis) * temp_axis + com;
        m->pos = spatial;
    }
}

void Container::setMassValues(double m) {
    for (Mass *mass : masses) mass->m += m;
}
void Container::setSpringConstants(double k) {
    for (Spring *s : springs) s->_k = k;
}
void Container::setRestLengths(double len) {
    for (Spring *s : springs) s->_rest = len;
}
void Container::addConstraint(CONSTRAINT_TYPE type, const Vec &v, double d) {
    for (Mass *m : masses) m->addConstraint(type, v, d);
}
void Container::clearConstraints() {
    for (Mass *m : masses) m->clearConstraints();
}
void Container::fix() {
    for (Mass *m : masses) m->fixed = true;
}
void Container::add(Mass *m) { masses.push_back(m); }
void Container::add(Spring *s) { springs.push_back(s); }
void Container::add(Container *c) {
    for (Mass *m : c->masses) masses.push_back(m);
    for (Spring *s : c->springs) springs.push_back(s);
}

// src/object.cu:179-196
Cube::Cube(const Vec &center, double side_length) : _side_length(side_length), _center(center) {
    for (int i = 0; i < 8; i++)
        masses.push_back(new Mass(side_length * (Vec(i & 1, (i >> 1) & 1, (i >> 2) & 1) - Vec(0.5, 0.5, 0.5)) + center));
    for (int i = 0; i < 8; i++)
        for (int j = i + 1; j < 8; j++) springs.push_back(new Spring(masses[i], masses[j]));
    for (Spring *s : springs) s->setRestLength((s->_right->pos - s->_left->pos).norm());
}

namespace {

// Lattice masses and springs in the reference's order (src/object.cu:232-293),
// from the engine's generator.
void fill_lattice(Container &c, const Vec &center, const Vec &dims, int nx, int ny, int nz) {
    if (nx <= 0 || ny <= 0 || nz <= 0) throw std::runtime_error("Lattice: dimensions must be positive");
    const int64_t n = (int64_t)nx * ny * nz, S = titan_lattice_spring_count(nx, ny, nz);
    std::vector<double> pos(3 * n), rest(S);
    std::vector<int32_t> L(S), R(S);
    if (titan_lattice(center[0], center[1], center[2], dims[0], dims[1], dims[2], nx, ny, nz, pos.data(), L.data(),
                      R.data(), rest.data()) != TITAN_OK)
        throw std::runtime_error("Lattice: generator failed");
    for (int64_t i = 0; i < n; ++i) c.masses.push_back(new Mass(Vec(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2])));
    for (int64_t s = 0; s < S; ++s) {
        Spring *sp = new Spring(c.masses[L[s]], c.masses[R[s]]);
        sp->setRestLength(rest[s]);  // |right - left|, src/object.cu:290-292
        c.springs.push_back(sp);
    }
}

}  // namespace

Lattice::Lattice(const Vec &center, const Vec &dims, int nx_, int ny_, int nz_)
    : nx(nx_), ny(ny_), nz(nz_), _center(center), _dims(dims) {
    fill_lattice(*this, center, dims, nx, ny, nz);
}

// src/object.cu:296-360: a lattice whose i = 0 face is fixed
Beam::Beam(const Vec &center, const Vec &dims, int nx_, int ny_, int nz_)
    : nx(nx_), ny(ny_), nz(nz_), _center(center), _dims(dims) {
    fill_lattice(*this, center, dims, nx, ny, nz);
    for (int64_t i = 0; i < (int64_t)ny * nz; ++i) masses[i]->fixed = true;
}

// ----------------------------------------------------------- Simulation

Simulation::Simulation(int integrator)
    : dt(0.0001), T(0), _global_acc(0, 0, -9.81), RUNNING(false), STARTED(false), ENDED(false), FREED(false),
      integrator_(integrator) {}

Simulation::~Simulation() {
    if (ctx) titan_ctx_destroy(ctx);
    ctx = nullptr;
    std::set<Mass *> ms(masses.begin(), masses.end());
    std::set<Spring *> ss(springs.begin(), springs.end());
    ms.insert(retired_masses.begin(), retired_masses.end());
    ss.insert(retired_springs.begin(), retired_springs.end());
    for (Container *c : containers) {
        ms.insert(c->masses.begin(), c->masses.end());
        ss.insert(c->springs.begin(), c->springs.end());
    }
    for (Spring *s : ss) delete s;
    for (Mass *m : ms) delete m;
    for (Container *c : containers) delete c;
    for (Constraint *c : constraints) delete c;
}

void Simulation::check_alive(const char *what) {
    if (ENDED) throw std::runtime_error(std::string("The simulation has ended. ") + what);
}

Mass *Simulation::createMass(Mass *m) {
    check_alive("Cannot modify simulation after the end of the simulation.");
    if (STARTED && RUNNING) throw std::runtime_error("The simulation is running. Stop the simulation to make changes.");
    if (STARTED) pull_state();
    masses.push_back(m);
    if (STARTED) {  // joins the engine body at the next step (src/sim.cu:236-252)
        write_record(m);
        m->index = -1;
        dirty = true;
    }
    return m;
}

Mass *Simulation::createMass() { return createMass(new Mass()); }
Mass *Simulation::createMass(const Vec &pos) { return createMass(new Mass(pos)); }

Spring *Simulation::createSpring(Spring *s) {
    check_alive("Cannot modify simulation after the end of the simulation.");
    // the reference calls exit(1) here (src/sim.cu:303-305); this facade throws
    if (STARTED && RUNNING) throw std::runtime_error("The simulation is running. Stop the simulation to make changes.");
    if (STARTED) pull_state();
    springs.push_back(s);
    if (STARTED) {
        write_record(s);
        dirty = true;
    }
    return s;
}

Spring *Simulation::createSpring() { return createSpring(new Spring()); }
Spring *Simulation::createSpring(Mass *m1, Mass *m2) { return createSpring(new Spring(m1, m2)); }

// Deleted objects stay allocated until the simulation is destroyed: springs
// may still point at a deleted mass, which then counts as missing and the
// spring does nothing (src/sim.cu:1052).
// On a built body, a deletion is applied in place (titan_body_remove): the
// removed mass stays in the engine body, frozen, and only the incidence
// lists it touches are rewritten; otherwise the body is rebuilt before the
// next step.
bool Simulation::remove_in_place(const std::vector<Mass *> &ms, const std::vector<Spring *> &ss) {
    if (!STARTED || !ctx || body < 0 || dirty) return false;
    std::vector<int64_t> mi, si;
    for (Mass *m : ms)
        if (m->index >= 0) mi.push_back(m->index);
    for (Spring *s : ss)
        if (s->index >= 0) si.push_back(s->index);
    check(ctx, titan_body_remove(ctx, body, mi.data(), (int64_t)mi.size(), si.data(), (int64_t)si.size()));
    for (Spring *s : ss) s->index = -1;
    return true;
}

void Simulation::deleteMass(Mass *m) {
    check_alive("Cannot modify simulation after the end of the simulation.");
    if (STARTED && RUNNING) throw std::runtime_error("The simulation is running. Stop the simulation to make changes.");
    const bool in_place = remove_in_place({m}, {});
    if (STARTED && !in_place && m->index >= 0) pull_state();
    erase_one(masses, m);
    m->valid = false;
    retired_masses.push_back(m);
    if (STARTED && !in_place) dirty = true;
}

void Simulation::deleteSpring(Spring *s) {
    check_alive("Cannot modify simulation after the end of the simulation.");
    if (STARTED && RUNNING) throw std::runtime_error("The simulation is running. Stop the simulation to make changes.");
    const bool in_place = remove_in_place({}, {s});
    if (STARTED && !in_place) pull_state();
    erase_one(springs, s);
    retired_springs.push_back(s);
    if (STARTED && !in_place) dirty = true;
}

// CUDA_MASS(Mass&) (src/mass.cu:57-74): m, T, pos, vel, extern_force and the
// constraints (drag truncated to int, object.h:213); acc is not written.
void Simulation::write_record(Mass *m) {
    m->dev.m = m->m;
    m->dev.T = m->T;
    m->dev.pos = m->pos;
    m->dev.vel = m->vel;
    m->dev.ext = m->extern_force;
    m->dev.fixed = m->fixed;
    const double c = m->drag_coefficient;
    m->dev.drag = (c >= 2147483647.0) ? 2147483647 : (c <= -2147483648.0) ? (-2147483647 - 1) : (int32_t)c;
    m->dev.locals = m->locals;
}

// CUDA_SPRING(Spring&) (src/spring.cu:35-54)
void Simulation::write_record(Spring *s) {
    s->dev.left = s->_left;
    s->dev.right = s->_right;
    s->dev.k = s->_k;
    s->dev.rest = s->_rest;
    s->dev.damping = s->_damping;
    s->dev.omega = s->_omega;
    s->dev.type = (int)s->_type;
}

// While a rebuild is pending the records are the newest state (they were
// pulled just before the change that made the body dirty, and no step has
// run since), so there is nothing to read back.
void Simulation::pull_state() {
    if (!ctx || body < 0 || dirty) return;
    const int64_t n = (int64_t)masses.size();
    int64_t on = 0;
    for (Mass *m : masses) on = std::max(on, m->index + 1);
    for (Mass *m : retired_masses) on = std::max(on, m->index + 1);
    if (on == 0) return;
    std::vector<double> pos(3 * on), vel(3 * on), acc(3 * on), Tm(on);
    check(ctx, titan_body_download(ctx, body, pos.data(), vel.data(), acc.data(), Tm.data()));
    auto fill = [&](Mass *m) {
        if (m->index < 0) return;
        const int64_t i = m->index;
        m->dev.pos = Vec(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        m->dev.vel = Vec(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
        m->dev.acc = Vec(acc[3 * i], acc[3 * i + 1], acc[3 * i + 2]);
        m->dev.T = Tm[i];
    };
    for (Mass *m : masses) fill(m);
    for (Mass *m : retired_masses) fill(m);
    (void)n;
}

void Simulation::push_state() {
    if (!ctx || body < 0) return;
    int64_t on = 0;
    for (Mass *m : masses) on = std::max(on, m->index + 1);
    for (Mass *m : retired_masses) on = std::max(on, m->index + 1);
    std::vector<double> pos(3 * on), vel(3 * on), mm(on), Tm(on), ext(3 * on);
    auto put = [&](Mass *m) {
        if (m->index < 0) return;
        const int64_t i = m->index;
        for (int k = 0; k < 3; ++k) {
            pos[3 * i + k] = m->dev.pos[k];
            vel[3 * i + k] = m->dev.vel[k];
            ext[3 * i + k] = m->dev.ext[k];
        }
        mm[i] = m->dev.m;
        Tm[i] = m->dev.T;
    };
    for (Mass *m : retired_masses) put(m);
    for (Mass *m : masses) put(m);
    check(ctx, titan_body_upload(ctx, body, pos.data(), vel.data(), mm.data(), Tm.data(), ext.data()));
}

// Re-create the engine body from the device records of the current masses
// and springs, continuing the simulation clock.
void Simulation::rebuild() {
    pull_state();
    if (body >= 0) check(ctx, titan_body_destroy(ctx, body));
    body = -1;
    for (Mass *m : retired_masses) m->index = -1;
    const int64_t n = (int64_t)masses.size(), S = (int64_t)springs.size();
    std::vector<double> pos(3 * n), vel(3 * n), mm(n), Tm(n), ext(3 * n);
    std::vector<uint8_t> fixed(n);
    std::vector<int32_t> drag(n);
    std::vector<int64_t> lc_ptr(n + 1, 0);
    std::vector<int32_t> lc_type;
    std::vector<double> lc_vec, lc_num;
    bool any_lc = false;
    for (int64_t i = 0; i < n; ++i) {
        Mass *m = masses[i];
        m->index = i;
        for (int k = 0; k < 3; ++k) {
            pos[3 * i + k] = m->dev.pos[k];
            vel[3 * i + k] = m->dev.vel[k];
            ext[3 * i + k] = m->dev.ext[k];
        }
        mm[i] = m->dev.m;
        Tm[i] = m->dev.T;
        fixed[i] = m->dev.fixed ? 1 : 0;
        drag[i] = m->dev.drag;
        for (const auto &l : m->dev.locals) {
            lc_type.push_back(l.kind);
            lc_vec.insert(lc_vec.end(), {l.v[0], l.v[1], l.v[2]});
            lc_num.push_back(l.num);
            any_lc = true;
        }
        lc_ptr[i + 1] = (int64_t)lc_type.size();
    }
    std::vector<int32_t> L(S), R(S), type(S);
    std::vector<double> k(S), rest(S), damping(S), omega(S);
    for (Spring *sp : retired_springs) sp->index = -1;
    for (int64_t s = 0; s < S; ++s) {
        springs[s]->index = s;
        const Spring::Record &r = springs[s]->dev;
        // a missing or deleted endpoint makes the spring inert (src/sim.cu:1052)
        L[s] = (r.left && r.left->valid && r.left->index >= 0) ? (int32_t)r.left->index : -1;
        R[s] = (r.right && r.right->valid && r.right->index >= 0) ? (int32_t)r.right->index : -1;
        k[s] = r.k;
        rest[s] = r.rest;
        damping[s] = r.damping;
        omega[s] = r.omega;
        type[s] = r.type;
    }
    if (n == 0) return;
    titan_body_desc d{};
    d.n_masses = n;
    d.n_springs = S;
    d.pos = pos.data();
    d.vel = vel.data();
    d.m = mm.data();
    d.T = Tm.data();
    d.ext = ext.data();
    d.fixed = fixed.data();
    d.drag = drag.data();
    d.left = L.data();
    d.right = R.data();
    d.k = k.data();
    d.rest = rest.data();
    d.damping = damping.data();
    d.type = type.data();
    d.omega = omega.data();
    if (any_lc) {
        d.lc_ptr = lc_ptr.data();
        d.lc_type = lc_type.data();
        d.lc_vec = lc_vec.data();
        d.lc_num = lc_num.data();
    }
    d.t0 = T;
    d.layout = TITAN_LAYOUT_AUTO;
    check(ctx, titan_body_create(ctx, &d, &body));
    dirty = false;
}

void Simulation::advance(int64_t n) {
    if (n <= 0) return;
    if (dirty) rebuild();
    if (body >= 0) check(ctx, titan_step(ctx, n));
    else {  // no masses: only the clock moves
        for (int64_t i = 0; i < n; ++i) advance_clock(T);
    }
    if (body >= 0) check(ctx, titan_time(ctx, &T));
}

// One step of the driver's clock: T += dt (src/sim.cu:1590), or for RK2 two
// stages of T += 0.5 * dt (:1577,1582), as the engine counts it.
void Simulation::advance_clock(double &t) const {
    if (integrator_ == 2) {
        t += 0.5 * dt;
        t += 0.5 * dt;
    } else {
        t += dt;
    }
}

// Steps until the driver loop of src/sim.cu:1503-1590 would hand control
// back: at a breakpoint (the first step boundary with breakpoint <= T), or,
// for wait/waitUntil, once T > t (src/sim.cu:1628-1646).
int64_t Simulation::steps_until_after(double t, bool strict, bool &hit_breakpoint) {
    double TT = T;
    int64_t k = 0;
    hit_breakpoint = false;
    for (;;) {
        if (!bpts.empty() && *bpts.begin() <= TT) {
            hit_breakpoint = true;
            return k;
        }
        if (strict ? (TT > t) : (TT >= t)) return k;
        advance_clock(TT);
        if (++k > ((int64_t)1 << 40)) throw std::runtime_error("requested time is unreachable");
    }
}

void Simulation::get(Mass *m) {
    if (!STARTED) {
        std::cerr << "The simulation has not started. Get and set commands cannot be called before sim.start()" << std::endl;
        return;
    }
    check_alive("Cannot modify simulation after the end of the simulation.");
    pull_state();
    // Mass::operator=(CUDA_MASS&), src/mass.cu:20-40
    m->m = m->dev.m;
    m->T = m->dev.T;
    m->pos = m->dev.pos;
    m->vel = m->dev.vel;
    m->acc = m->dev.acc;
    m->extern_force = m->dev.ext;
}

void Simulation::get(Spring *) {
    if (!STARTED) {
        std::cerr << "The simulation has not started. Get and set commands cannot be called before sim.start()" << std::endl;
        return;
    }
    check_alive("Cannot modify simulation after the end of the simulation.");
}

void Simulation::get(Container *c) {
    check_alive("Cannot get updates from the GPU after the end of the simulation.");
    if (!STARTED) {
        std::cerr << "sim.get() does nothing if called before the simulation has started." << std::endl;
        return;
    }
    pull_state();
    for (Mass *m : c->masses) {
        m->m = m->dev.m;
        m->T = m->dev.T;
        m->pos = m->dev.pos;
        m->vel = m->dev.vel;
        m->acc = m->dev.acc;
        m->extern_force = m->dev.ext;
    }
}

void Simulation::getAll() {
    if (!STARTED) {
        std::cerr << "The simulation has not started. Get and set commands cannot be called before sim.start()" << std::endl;
        return;
    }
    check_alive("Cannot modify simulation after the end of the simulation.");
    pull_state();
    for (Mass *m : masses) {
        m->m = m->dev.m;
        m->T = m->dev.T;
        m->pos = m->dev.pos;
        m->vel = m->dev.vel;
        m->acc = m->dev.acc;
        m->extern_force = m->dev.ext;
    }
}

namespace {

bool same_locals(const std::vector<Mass::Local> &a, const std::vector<Mass::Local> &b);

}  // namespace

void Simulation::set(Mass *m) {
    if (!STARTED) {
        std::cerr << "The simulation has not started. Get and set commands cannot be called before sim.start()" << std::endl;
        return;
    }
    check_alive("Cannot modify simulation after the end of the simulation.");
    pull_state();
    const Mass::Record before = m->dev;
    write_record(m);
    if (m->index < 0 || before.fixed != m->dev.fixed || before.drag != m->dev.drag ||
        !same_locals(before.locals, m->dev.locals))
        dirty = true;
    else
        push_state();
}

void Simulation::set(Spring *s) {
    if (!STARTED) {
        std::cerr << "The simulation has not started. Get and set commands cannot be called before sim.start()" << std::endl;
        return;
    }
    pull_state();
    write_record(s);
    dirty = true;
}

void Simulation::set(Container *c) {
    check_alive("Cannot modify simulation after the end of the simulation.");
    if (!STARTED) {
        std::cerr << "The simulation has not started. Get and set commands cannot be called before sim.start()" << std::endl;
        return;
    }
    pull_state();
    bool rebuild_needed = !c->springs.empty();
    for (Mass *m : c->masses) {
        const Mass::Record before = m->dev;
        write_record(m);
        if (m->index < 0 || before.fixed != m->dev.fixed || before.drag != m->dev.drag ||
            !same_locals(before.locals, m->dev.locals))
            rebuild_needed = true;
    }
    for (Spring *s : c->springs) write_record(s);
    if (rebuild_needed) dirty = true;
    else push_state();
}

void Simulation::setAll() {
    check_alive("Cannot modify simulation after the end of the simulation.");
    if (!STARTED) {
        std::cerr << "The simulation has not started. Get and set commands cannot be called before sim.start()" << std::endl;
        return;
    }
    pull_state();
    for (Mass *m : masses) write_record(m);
    for (Spring *s : springs) write_record(s);
    dirty = true;
}

namespace {

bool same_locals(const std::vector<Mass::Local> &a, const std::vector<Mass::Local> &b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i].kind != b[i].kind || !(a[i].v == b[i].v) || a[i].num != b[i].num) return false;
    return true;
}

}  // namespace

void Simulation::createPlane(const Vec &abc, double d) { createPlane(abc, d, 0.0, 0.0); }

void Simulation::createPlane(const Vec &abc, double d, double FRICTION_K, double FRICTION_S) {
    check_alive("New objects cannot be created.");
    ContactPlane *p = new ContactPlane(abc, d);
    p->_FRICTION_K = FRICTION_K;
    p->_FRICTION_S = FRICTION_S;
    constraints.push_back(p);
    dplanes.push_back({p->_normal[0], p->_normal[1], p->_normal[2], d, FRICTION_K, FRICTION_S});
    if (ctx) check(ctx, titan_add_plane(ctx, p->_normal[0], p->_normal[1], p->_normal[2], d, FRICTION_K, FRICTION_S));
}

void Simulation::createBall(const Vec &center, double r) {
    check_alive("New constraints cannot be added.");
    constraints.push_back(new Ball(center, r));
    dballs.push_back({center[0], center[1], center[2], r});
    if (ctx) check(ctx, titan_add_ball(ctx, center[0], center[1], center[2], r));
}

// The reference clears only the host list of shapes; planes and balls
// already sent to the device keep acting (src/sim.cu:2037-2040).
void Simulation::clearConstraints() {
    for (Constraint *c : constraints) delete c;
    constraints.clear();
}

Container *Simulation::createContainer() {
    Container *c = new Container();
    containers.push_back(c);
    return c;
}

void Simulation::deleteContainer(Container *c) {
    check_alive("Cannot modify simulation after the end of the simulation.");
    if (RUNNING) throw std::runtime_error("The simulation is running. Stop the simulation to make changes.");
    const bool in_place = remove_in_place(c->masses, c->springs);
    if (STARTED && !in_place) pull_state();
    for (Mass *m : c->masses) {
        erase_one(masses, m);
        m->valid = false;
        retired_masses.push_back(m);
    }
    for (Spring *s : c->springs) {
        erase_one(springs, s);
        retired_springs.push_back(s);
    }
    erase_one(containers, c);
    c->masses.clear();
    c->springs.clear();
    delete c;
    if (STARTED && !in_place) dirty = true;
}

Cube *Simulation::createCube(const Vec &center, double side_length) {
    check_alive("Cannot create new objects");
    Cube *cube = new Cube(center, side_length);
    for (Mass *m : cube->masses) createMass(m);
    for (Spring *s : cube->springs) createSpring(s);
    containers.push_back(cube);
    return cube;
}

Lattice *Simulation::createLattice(const Vec &center, const Vec &dims, int nx, int ny, int nz) {
    check_alive("New objects cannot be created.");
    Lattice *l = new Lattice(center, dims, nx, ny, nz);
    masses.reserve(masses.size() + l->masses.size());
    springs.reserve(springs.size() + l->springs.size());
    for (Mass *m : l->masses) createMass(m);
    for (Spring *s : l->springs) createSpring(s);
    containers.push_back(l);
    return l;
}

Beam *Simulation::createBeam(const Vec &center, const Vec &dims, int nx, int ny, int nz) {
    check_alive("New objects cannot be created.");
    Beam *b = new Beam(center, dims, nx, ny, nz);
    for (Mass *m : b->masses) createMass(m);
    for (Spring *s : b->springs) createSpring(s);
    containers.push_back(b);
    return b;
}

Container *Simulation::importFromSTL(const std::string &, double, int) {
    throw std::runtime_error("importFromSTL: STL import is not part of this engine");
}

void Simulation::setAllSpringConstantValues(double k) {
    check_alive("Cannot modify simulation after the end of the simulation.");
    for (Spring *s : springs) s->_k = k;
}

void Simulation::setAllMassValues(double m) {
    check_alive("Cannot modify simulation after the end of the simulation.");
    for (Mass *mass : masses) mass->m += m;
}

void Simulation::setTimeStep(double delta_t) {
    check_alive("Cannot modify simulation after the end of the simulation.");
    if (delta_t <= 0) throw std::runtime_error("Cannot set time step to negative or zero value.");
    dt = delta_t;
    if (ctx) check(ctx, titan_set_dt(ctx, dt));
}

double Simulation::getTimeStep() { return dt; }

void Simulation::setGlobalAcceleration(const Vec &global_acc) {
    if (RUNNING) throw std::runtime_error("The simulation is running. The global force parameter cannot be changed during runtime");
    _global_acc = global_acc;
    if (ctx) check(ctx, titan_set_global_acc(ctx, global_acc[0], global_acc[1], global_acc[2]));
}

void Simulation::defaultRestLengths() {
    check_alive("Cannot modify simulation after the end of the simulation.");
    for (Spring *s : springs) s->defaultLength();
}

void Simulation::start() {  // src/sim.cu:1349-1390
    if (ENDED) throw std::runtime_error("The simulation has ended. Cannot call sim.start() after the end of the simulation.");
    if (masses.empty()) throw std::runtime_error("No masses have been added. Please add masses before starting the simulation.");
    std::cout << "Starting simulation with " << masses.size() << " masses and " << springs.size() << " springs." << std::endl;
    if (dt <= 0) throw std::runtime_error("Simultation timestep Simulation::dt is invalid. Please choose a positive non-zero value.");
    RUNNING = true;
    STARTED = true;
    T = 0;
    if (!ctx) {
        const char *dev = std::getenv("TITAN_DEVICE");
        check(nullptr, titan_ctx_create(dev ? std::atoi(dev) : 0, 0, &ctx));
        check(ctx, titan_set_integrator(ctx, integrator_));
    }
    check(ctx, titan_set_dt(ctx, dt));
    check(ctx, titan_set_global_acc(ctx, _global_acc[0], _global_acc[1], _global_acc[2]));
    for (const auto &p : dplanes) check(ctx, titan_add_plane(ctx, p[0], p[1], p[2], p[3], p[4], p[5]));
    for (const auto &b : dballs) check(ctx, titan_add_ball(ctx, b[0], b[1], b[2], b[3]));
    for (Mass *m : masses) {
        write_record(m);
        m->index = -1;
    }
    for (Spring *s : springs) write_record(s);
    rebuild();
}

void Simulation::stop() {
    if (RUNNING) {
        setBreakpoint(time());
        waitForEvent();
    }
    ENDED = true;
    if (ctx) titan_ctx_destroy(ctx);
    ctx = nullptr;
    body = -1;
    FREED = true;
}

void Simulation::stop(double t) {
    if (RUNNING) {
        setBreakpoint(t);
        waitForEvent();
    }
    ENDED = true;
    if (ctx) titan_ctx_destroy(ctx);
    ctx = nullptr;
    body = -1;
    FREED = true;
}

void Simulation::pause(double t) {
    check_alive("Control functions cannot be called.");
    setBreakpoint(t);
    waitForEvent();
}

void Simulation::resume() {  // src/sim.cu:1483-1501
    check_alive("Cannot resume the simulation.");
    if (!STARTED) throw std::runtime_error("The simulation has not started. You cannot resume a simulation before calling sim.start().");
    if (masses.empty()) throw std::runtime_error("No masses have been added. Please add masses before starting the simulation.");
    RUNNING = true;
}

void Simulation::reset() {  // src/sim.cu:98-129
    masses.clear();
    springs.clear();
    containers.clear();
    constraints.clear();
    RUNNING = STARTED = ENDED = FREED = false;
    _global_acc = Vec(0, 0, -9.81);
    if (ctx) titan_ctx_destroy(ctx);
    ctx = nullptr;
    body = -1;
    dirty = false;
}

void Simulation::setBreakpoint(double time) {
    check_alive("Cannot set breakpoints after the end of the simulation run.");
    bpts.insert(time);
}

// The reference waits while the driver thread runs until a breakpoint
// stops it (src/sim.cu:1649-1657); here the steps up to that breakpoint are
// taken on the calling thread.  With no breakpoint the reference would wait
// forever; this throws instead.
void Simulation::waitForEvent() {
    if (ENDED && !FREED) throw std::runtime_error("The simulation has ended. Control functions cannot be called.");
    if (!RUNNING) return;
    if (bpts.empty()) throw std::runtime_error("waitForEvent: no breakpoint is set, the simulation would never stop");
    bool hit = false;
    const int64_t k = steps_until_after(std::numeric_limits<double>::infinity(), true, hit);
    advance(k);
    bpts.erase(bpts.begin());
    RUNNING = false;
}

// Simulated time passes until T > time() + t (src/sim.cu:1628-1637) or a
// breakpoint stops the run first.
void Simulation::wait(double t) {
    check_alive("Control functions cannot be called.");
    if (!RUNNING) return;
    bool hit = false;
    const int64_t k = steps_until_after(time() + t, true, hit);
    advance(k);
    if (hit) {
        bpts.erase(bpts.begin());
        RUNNING = false;
    }
}

void Simulation::waitUntil(double t) {
    if (ENDED && !FREED) throw std::runtime_error("The simulation has ended. Control functions cannot be called.");
    if (!RUNNING) return;
    bool hit = false;
    const int64_t k = steps_until_after(t, true, hit);
    advance(k);
    if (hit) {
        bpts.erase(bpts.begin());
        RUNNING = false;
    }
}

void Simulation::step(int64_t n) {
    check_alive("Control functions cannot be called.");
    if (!STARTED) throw std::runtime_error("The simulation has not started. Call sim.start() before stepping.");
    if (n < 0) throw std::runtime_error("step: negative step count");
    advance(n);
}

double Simulation::time() { return T; }
bool Simulation::running() { return RUNNING; }

void Simulation::printPositions() {
    check_alive("You cannot view parameters of the simulation after it has been stopped.");
    if (STARTED) {
        std::cout << "\nDEVICE MASSES: " << std::endl;
        pull_state();
        int count = 0;
        for (Mass *m : masses) std::cout << count++ << ": " << m->dev.pos << std::endl;
    } else {
        std::cout << "\nHOST MASSES: " << std::endl;
        int count = 0;
        for (Mass *m : masses) std::cout << count++ << ": " << m->pos << std::endl;
    }
    std::cout << std::endl;
}

Spring *Simulation::getSpringByIndex(int i) {
    if (i < 0 || i >= (int)springs.size()) throw std::out_of_range("getSpringByIndex");
    return springs[i];
}

Mass *Simulation::getMassByIndex(int i) {
    if (i < 0 || i >= (int)masses.size()) throw std::out_of_range("getMassByIndex");
    return masses[i];
}

Container *Simulation::getContainerByIndex(int i) {
    if (i < 0 || i >= (int)containers.size()) throw std::out_of_range("getContainerByIndex");
    return containers[i];
}

}  // namespace titan
