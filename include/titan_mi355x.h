/*
 * titan_mi355x.h — C ABI of the MI355X FP64 mass-spring step engine.
 *
 * The step that replaces the reference's computeSpringForces +
 * massForcesAndUpdate pair (src/sim.cu:1046-1169) driven by
 * Simulation::execute (src/sim.cu:1503-1617).  Plain pointers and sizes
 * only; the library owns all device memory, the caller owns host buffers.
 * Every function returns 0 on success or a negative TITAN_E* code; the
 * message of the last failure is available from titan_last_error().  No
 * C++ exception crosses this boundary.  One context may be used from one
 * host thread at a time; there is no background thread in the library.
 *
 * Host vectors are interleaved x,y,z triples (the reference's Vec layout,
 * include/Titan/vec.h:162); on the device the state is kept as FP64 SoA.
 */
#ifndef TITAN_MI355X_H
#define TITAN_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TITAN_OK 0
#define TITAN_E_INVALID (-1)     /* bad argument or call order */
#define TITAN_E_HIP (-2)         /* HIP runtime failure */
#define TITAN_E_NOMEM (-3)       /* device memory or HBM budget exhausted */
#define TITAN_E_NOTFOUND (-4)    /* unknown body / ticket */
#define TITAN_E_INJECTED (-5)    /* failure forced by titan_debug_inject_failure */

#define TITAN_ALL_BODIES (-1)

/* Spring kinds. 0-3 are the reference's SpringType values
 * (include/Titan/spring.h:17); TITAN_SPRING_GENERAL uses the actuation law
 * rest * (1 + a * sin(omega * t + phi)). */
#define TITAN_PASSIVE_SOFT 0
#define TITAN_PASSIVE_STIFF 1
#define TITAN_ACTIVE_CONTRACT_THEN_EXPAND 2
#define TITAN_ACTIVE_EXPAND_THEN_CONTRACT 3
#define TITAN_SPRING_GENERAL 4

/* Local constraint kinds, the reference's CONSTRAINT_TYPE values
 * (include/Titan/object.h:231-233). */
#define TITAN_CONSTRAINT_PLANE 0
#define TITAN_CONTACT_PLANE 1
#define TITAN_BALL 2
#define TITAN_DIRECTION 3

#define TITAN_LAYOUT_AUTO 0
#define TITAN_LAYOUT_GATHER 1
#define TITAN_LAYOUT_TILED 2
#define TITAN_LAYOUT_RESIDENT 3

typedef struct titan_ctx titan_ctx;

/* One body: a set of masses and the springs among them, copied in by
 * titan_body_create.  Optional arrays may be NULL.  Spring endpoints are
 * body-local mass indices; a negative endpoint marks a spring with no mass,
 * which does nothing (src/sim.cu:1052). */
typedef struct {
    int64_t n_masses;
    int64_t n_springs;
    const double *pos;        /* 3*n_masses */
    const double *vel;        /* 3*n_masses, NULL = zero */
    const double *m;          /* n_masses (kg) */
    const double *T;          /* n_masses local time, NULL = 0 */
    const double *ext;        /* 3*n_masses extern force, NULL = none */
    const uint8_t *fixed;     /* n_masses, NULL = none fixed */
    const int32_t *drag;      /* n_masses integer drag coefficient, NULL = 0 */
    const int32_t *left;      /* n_springs */
    const int32_t *right;     /* n_springs */
    const double *k;          /* n_springs */
    const double *rest;       /* n_springs */
    const double *damping;    /* n_springs, NULL = 0 */
    const int32_t *type;      /* n_springs, NULL = passive */
    const double *omega;      /* n_springs, read for kinds 2-4 */
    const double *act_a;      /* n_springs, read for kind 4 */
    const double *act_phi;    /* n_springs, read for kind 4 */
    /* local constraints, CSR over masses in insertion order
     * (src/mass.cu:78-96): lc_ptr has n_masses+1 entries; lc_vec holds the
     * unit normal / unit tangent / ball centre (3 per entry) and lc_num the
     * offset / friction / radius.  NULL lc_ptr = none. */
    const int64_t *lc_ptr;
    const int32_t *lc_type;
    const double *lc_vec;
    const double *lc_num;
    /* body time at creation: the first step's actuation time */
    double t0;
    /* device layout: TITAN_LAYOUT_AUTO (the context's schedule decides, see
     * titan_schedule; tiled only when the body's tiles fit the kernel's
     * limits), TITAN_LAYOUT_GATHER, TITAN_LAYOUT_TILED (fails if a tile does
     * not fit) or TITAN_LAYOUT_RESIDENT (at most 32767 masses: one
     * workgroup runs all the body's steps of a stepping call, in one launch
     * for every resident body; a body of at most 1024 masses is held in LDS
     * when the launch has at most one body per compute unit, otherwise the
     * body steps in its own state buffers).  Results are identical either
     * way. */
    int32_t layout;
} titan_body_desc;

/* ---- context (replaces Simulation's constructor/stop/freeGPU,
 *      src/sim.cu:77-221,1319-1347) ---- */
int titan_ctx_create(int device, uint64_t hbm_budget_bytes, titan_ctx **out);
int titan_ctx_destroy(titan_ctx *ctx);
/* Message of the last failure on ctx (or of titan_ctx_create if ctx is
 * NULL); never NULL. */
const char *titan_last_error(const titan_ctx *ctx);
/* Grow the context's arenas once, now, to hold n_masses masses and
 * n_springs springs (n_actuated of them actuated) in n_bodies bodies, so that
 * creating bodies up to that scale grows nothing.  No counterpart in the
 * reference, whose every mass and spring is an allocation of its own: this
 * is for servers, which reserve their HBM at startup rather than grow it
 * under load (a growth waits for the steps in flight).  A capacity hint:
 * bodies beyond it still grow the arenas on demand (the tiled layout keeps
 * a copy of an actuator's record in each tile it touches, reserved for
 * two tiles per actuated spring).  Counts toward the HBM budget
 * (TITAN_E_NOMEM if it would exceed it; nothing is grown then). */
int titan_reserve(titan_ctx *ctx, int64_t n_masses, int64_t n_springs, int64_t n_actuated, int32_t n_bodies);

/* ---- globals (src/sim.cu:797-807,2081-2087,1998-2040) ---- */
int titan_set_dt(titan_ctx *ctx, double dt);                       /* setTimeStep */
int titan_get_dt(const titan_ctx *ctx, double *dt);                /* getTimeStep */
int titan_set_global_acc(titan_ctx *ctx, double gx, double gy, double gz); /* setGlobalAcceleration */
/* createPlane(abc, d, FRICTION_K, FRICTION_S); the normal must already be
 * unit length, as ContactPlane's constructor makes it (object.h:102-104). */
int titan_add_plane(titan_ctx *ctx, double nx, double ny, double nz, double d,
                    double mu_k, double mu_s);
int titan_add_ball(titan_ctx *ctx, double cx, double cy, double cz, double r); /* createBall */
int titan_clear_planes_and_balls(titan_ctx *ctx);

/* ---- bodies (replace toArray/massToArray/springToArray,
 *      src/sim.cu:829-930) ---- */
int titan_body_create(titan_ctx *ctx, const titan_body_desc *desc, int32_t *body_id);
int titan_body_destroy(titan_ctx *ctx, int32_t body_id);
int titan_body_count(const titan_ctx *ctx, int32_t *n);
/* Remove springs and masses of a live body mid-run (replaces
 * Simulation::deleteSpring / deleteMass after start, src/sim.cu:352-412):
 * a removed spring exerts no force from the next step on; a removed mass
 * stops moving (its state stays as it is, the reference no longer updates
 * it) and every spring touching it is removed, as an invalidated endpoint
 * makes a spring inert (src/sim.cu:1052).  Indices are the body's own:
 * masses in titan_body_desc order, springs in its spring order; removing
 * one twice changes nothing.  Only the incidence lists of the masses
 * touched are rewritten, on the device; the rest of the layout is kept.
 * Waits for the steps in flight. */
int titan_body_remove(titan_ctx *ctx, int32_t body_id, const int64_t *masses, int64_t n_masses,
                      const int64_t *springs, int64_t n_springs);

/* ---- stepping (replaces the loop body of Simulation::execute,
 *      src/sim.cu:1569-1590) ---- */
/* Advance every body by n_steps; each step runs at the body's time T, then
 * T += dt (src/sim.cu:1590).  Enqueued on the context's stream; returns
 * without waiting unless a later call needs the results. */
int titan_step(titan_ctx *ctx, int64_t n_steps);
/* titan_step that allocates nothing (no counterpart in the reference: for
 * callers that cannot allocate device memory, such as code under stream
 * capture or a tight serving loop).  Same steps and results as titan_step,
 * but every step is launched directly (no graph is captured), the ball grid
 * is not rebuilt (if the balls changed since the last titan_step, every
 * mass tests every ball, with the same result), and the step kernels take
 * their scratch buffers (actuation factors of actuated springs; the RK2
 * stage state) from a static device array of 64 MiB per device that the
 * library holds until the process ends.  Returns TITAN_E_INVALID, before any
 * work, when the context needs more than that: 16 B per actuator slot, plus
 * 48 B per mass slot under RK2.  Every call on one device shares the array,
 * so calls that overlap on one device (contexts stepped from different
 * threads) race; the array's contents at entry never matter. */
int titan_step_noalloc(titan_ctx *ctx, int64_t n_steps);
/* Step until the first step boundary with T >= t, the breakpoint rule of
 * src/sim.cu:1505; *steps_done (may be NULL) receives the count. */
int titan_step_until(titan_ctx *ctx, double t, int64_t *steps_done);
int titan_time(const titan_ctx *ctx, double *T);   /* Simulation::time, src/sim.cu:1082-1084 */
int titan_sync(titan_ctx *ctx);
/* Per-body stepping, for batched rollouts and serving: advance each listed
 * body by its own number of steps (n_steps[k] for body_ids[k]); the other
 * bodies do not move and the context clock (titan_time) does not advance.
 * Generalises the driver loop of Simulation::execute (src/sim.cu:1503-1617),
 * which has one clock for the whole simulation.  With count == 1 the
 * launches cover that body's tiles or chunks only (one body per launch);
 * otherwise one launch per step covers every body.  Returns once the steps
 * are enqueued. */
int titan_step_bodies(titan_ctx *ctx, const int32_t *body_ids, const int64_t *n_steps, int32_t count);
/* A body's own clock (T += dt per step it took) and step count. */
int titan_body_time(const titan_ctx *ctx, int32_t body_id, double *T, int64_t *steps);

/* ---- readback / write-back (replace getAll/get/setAll/set,
 *      src/sim.cu:588-764,940-1005) ---- */
/* Any output pointer may be NULL.  pos/vel/acc: 3*n_masses; T: n_masses. */
int titan_body_download(titan_ctx *ctx, int32_t body_id, double *pos, double *vel,
                        double *acc, double *T);
/* CUDA_MASS(Mass&) semantics (src/mass.cu:57-74): pos, vel, m, T, ext are
 * written (NULL leaves the field unchanged); acc is not. */
int titan_body_upload(titan_ctx *ctx, int32_t body_id, const double *pos, const double *vel,
                      const double *m, const double *T, const double *ext);

/* ---- timing on the context's stream (hipEvents) ---- */
int titan_timer_start(titan_ctx *ctx);
/* Stops the timer started by titan_timer_start; *ms = elapsed GPU time. */
int titan_timer_stop(titan_ctx *ctx, float *ms);

/* ---- introspection used by tests and the benchmark ---- */
/* Device bytes currently allocated by the context. */
int titan_mem_used(const titan_ctx *ctx, uint64_t *bytes);
/* Names of the step kernel variants the next titan_step will launch
 * ("+"-separated when both layouts are in use). */
int titan_kernel_name(const titan_ctx *ctx, char *buf, int64_t len);
/* Layout a body was given: TITAN_LAYOUT_GATHER, _TILED or _RESIDENT. */
int titan_body_layout(const titan_ctx *ctx, int32_t body_id, int32_t *layout);

/* ---- integrator (the reference picks one at build time: the default Euler
 *      branch, or VERLET / RK2, CMakeLists.txt:13-14, src/sim.cu:1141-1167,
 *      1572-1590) ---- */
#define TITAN_INTEGRATOR_EULER 0
#define TITAN_INTEGRATOR_VERLET 1
#define TITAN_INTEGRATOR_RK2 2
/* Set before the context holds any body.  Euler: acc = F/m, vel += acc dt,
 * pos += vel dt.  Verlet: vel += (acc_prev + F/m) dt / 2, acc = F/m, pos +=
 * vel dt + acc dt^2 / 2 (acc_prev starts at 0).  RK2: two force evaluations
 * per step (midpoint), the clock advancing dt / 2 after each.  Step counts
 * everywhere are whole steps. */
int titan_set_integrator(titan_ctx *ctx, int32_t kind);
int titan_get_integrator(const titan_ctx *ctx, int32_t *kind);

/* ---- schedule and autotuning (no counterpart in the reference, which has
 *      one fixed launch shape per kernel, src/sim.cu:1569-1590) ---- */
/* How bodies created with TITAN_LAYOUT_AUTO are laid out: bodies with at
 * most resident_max_masses masses (0 .. 32767; 0 = none) get the resident
 * layout; of the others, those with at least tiled_min_masses masses get the
 * tiled layout and smaller ones the gather layout; tile_rows bounds a tile's
 * LDS rows (masses + halo, 64 .. 2304; fewer rows allow more workgroups per
 * CU, more rows bigger tiles).  Default: resident_max_masses 256,
 * tiled_min_masses 0, tile_rows 2304 (a body of at most 256 masses is
 * resident: a few small scenes step fastest so, each in one launch per
 * stepping call with its spring terms spread over a workgroup; every larger
 * body is tiled; for a population titan_autotune measures the
 * alternatives).  Applies to bodies created afterwards. */
typedef struct titan_schedule {
    int64_t tiled_min_masses;
    int32_t tile_rows;
    int32_t resident_max_masses;
} titan_schedule;
int titan_set_schedule(titan_ctx *ctx, const titan_schedule *sched);
int titan_get_schedule(const titan_ctx *ctx, titan_schedule *sched);
/* Pick the schedule by measurement: for each candidate (resident_max_masses
 * in {0, 256, 1024, 32767}, tiled_min_masses in {0, 128, 256, 512, 1024,
 * never}, tile_rows in {682, 2304}; 48 in that order, the last index fastest) a scratch
 * context on the same device with this context's dt, gravity, planes and
 * balls holds `replicas` copies of the n_samples sample bodies and is timed
 * over `steps` steps (after a warm-up of the same length).  The fastest in
 * mass-steps/s becomes this context's schedule and is returned in *best
 * (may be NULL); *rate_out (may be NULL) receives each candidate's rate, in
 * candidate order (48 entries; candidates that lay out every sample alike
 * are timed once and share the rate). */
int titan_autotune(titan_ctx *ctx, const titan_body_desc *samples, int32_t n_samples, int32_t replicas,
                   int32_t steps, titan_schedule *best, double *rate_out);

/* ---- serving: continuous batching of rollout requests (no counterpart in
 *      the reference, whose one Simulation runs one scene on a driver
 *      thread, src/sim.cu:1389,1503-1617) ----
 * A request is a body and a horizon (steps).  titan_serve_submit copies the
 * body and queues it.  titan_serve_pump, called by the host loop (the
 * library starts worker threads for the layouts only), admits queued
 * requests in arrival order while the running set stays within max_running
 * requests and the serving HBM budget (estimated bytes per body; a body that
 * does not fit waits until others finish, and no later request overtakes
 * it), steps every running request together, up to `quantum`
 * steps each and never past its horizon, in one launch per step, and
 * retires the finished ones: their final positions and velocities go to the
 * request's out buffers and their device memory is released.  A request's
 * result does not depend on what runs beside it.  Latency is measured from
 * submit to completion on the host's steady clock: for a request with out
 * buffers, when they are written; for one without, when a pump sees that
 * the device has run its last step (until then it counts as running).
 * Layouts are built by worker threads from submission; while the device has
 * work, a pump admits only requests whose layouts are built (in arrival
 * order among those; one still being built is passed over and a later pump
 * admits it), and with nothing running it waits for the oldest request's.
 * Admission within a pump stops, between placements, once the device has
 * run the quantum enqueued before that pump's.  With TITAN_SERVE_ADMIT=wait
 * in the environment when the context first serves, admission waits for
 * every layout it picks and places every request that fits instead, so that
 * what a pump admits depends on neither the workers' nor the device's
 * timing.  TITAN_SERVE_WORKERS=<n> sets the number of layout threads
 * (default: one per host thread but four, at most 12). */
#define TITAN_REQ_QUEUED 0
#define TITAN_REQ_RUNNING 1
#define TITAN_REQ_DONE 2
#define TITAN_REQ_CANCELLED 3
#define TITAN_REQ_FAILED 4     /* its body was rejected by titan_body_create */

typedef struct titan_request {
    titan_body_desc body;   /* copied by titan_serve_submit */
    int64_t horizon;        /* steps to run, >= 0 */
    double *out_pos;        /* optional: 3*n_masses, final positions; must stay valid until done */
    double *out_vel;        /* optional: 3*n_masses, final velocities */
} titan_request;

typedef struct titan_serve_report {
    int64_t submitted, completed, cancelled, failed, running, queued;
    double p50_latency_s, p99_latency_s, max_latency_s;  /* over completed requests */
    double mass_steps;          /* of completed requests */
    double mass_steps_per_s;    /* mass_steps / (last completion - first submit) */
    double device_idle_s;       /* step stream idle between consecutive quanta (device clock) */
} titan_serve_report;

/* max_running: requests on the device at once (>= 1); hbm_bytes: serving
 * budget (0: the context's budget, or unlimited); quantum: steps per pump
 * (>= 1).  Defaults: 4096, 0, 256. */
int titan_serve_config(titan_ctx *ctx, int32_t max_running, uint64_t hbm_bytes, int64_t quantum);
int titan_serve_submit(titan_ctx *ctx, const titan_request *req, int64_t *ticket);
/* One round of admit / step / retire; *running and *queued (may be NULL)
 * receive the counts afterwards. */
int titan_serve_pump(titan_ctx *ctx, int32_t *running, int32_t *queued);
int titan_serve_poll(titan_ctx *ctx, int64_t ticket, int32_t *state, int64_t *steps_done);
/* Drops a queued or running request (its out buffers are not written). */
int titan_serve_cancel(titan_ctx *ctx, int64_t ticket);
int titan_serve_stats(titan_ctx *ctx, titan_serve_report *out);
/* Forget completed/cancelled requests and reset the statistics. */
int titan_serve_reset_stats(titan_ctx *ctx);

/* HBM calibration (no counterpart in the reference; a measurement aid for
 * the roofline): on `device`, a streaming copy between two buffers of
 * `bytes` bytes (16-byte loads, non-temporal 16-byte stores; *copy_gbps
 * counts the bytes read plus written) and a streaming read of one
 * (*read_gbps), each timed with HIP events over `iters` launches after a
 * warm-up.  Allocates and frees its own buffers; errors are reported by
 * titan_last_error(NULL). */
int titan_measure_hbm(int device, uint64_t bytes, int32_t iters, double *copy_gbps, double *read_gbps);

/* Host-side lattice generator with the reference's mass and spring order
 * (Lattice::Lattice, src/object.cu:232-293).  pos: 3*nx*ny*nz; left/right/
 * rest: titan_lattice_spring_count(nx, ny, nz) entries each. */
int64_t titan_lattice_spring_count(int nx, int ny, int nz);
int titan_lattice(double cx, double cy, double cz, double dx, double dy, double dz,
                  int nx, int ny, int nz, double *pos, int32_t *left, int32_t *right,
                  double *rest);

/* Fault injection for host-side tests of the error paths: the next `after`
 * successful calls proceed, then one call of titan_step returns
 * TITAN_E_INJECTED without touching the device.  Negative `after` clears it. */
int titan_debug_inject_failure(titan_ctx *ctx, int32_t after);
/* The same at another call site: TITAN_INJECT_STEP (titan_step and
 * titan_step_bodies, as above), TITAN_INJECT_DOWNLOAD (titan_body_download)
 * or TITAN_INJECT_DESTROY (titan_body_destroy); the failing call changes
 * nothing.  Used to test the serving loop's recovery from a failed
 * retirement. */
/* A body's device layout, for tests that compare two layouts byte for byte
 * (no counterpart in the reference): item `what` copied to out (at most cap
 * bytes; *len = its full size), arena offsets made relative to the body's
 * own ranges: TITAN_DEBUG_NAMES (uint32 words of its incidence span),
 * _REST, _K, _DAMPING (double per entry; empty when shared), _DEG, _MFLAGS
 * (per mass slot), _STATE (48 B per slot, buffer 0), _MASS, _TILES
 * (40-byte tile records: body, first chunk, masses, halo size, halo
 * offset, first-tile flag, actuators, first actuator record; body zeroed),
 * _HALO (int32, body slots), _CHUNKS (4 int64 per chunk: first entry,
 * groups, first actuator, actuators), _PERM (int32 per mass: device order
 * -> caller's order) and _TACT (the tiles' actuator records, 4 doubles
 * each: a, omega, phi, rest). */
#define TITAN_DEBUG_NAMES 0
#define TITAN_DEBUG_REST 1
#define TITAN_DEBUG_K 2
#define TITAN_DEBUG_DAMPING 3
#define TITAN_DEBUG_DEG 4
#define TITAN_DEBUG_MFLAGS 5
#define TITAN_DEBUG_STATE 6
#define TITAN_DEBUG_MASS 7
#define TITAN_DEBUG_TILES 8
#define TITAN_DEBUG_HALO 9
#define TITAN_DEBUG_CHUNKS 10
#define TITAN_DEBUG_PERM 11
#define TITAN_DEBUG_TACT 12
int titan_debug_body_bytes(titan_ctx *ctx, int32_t body_id, int32_t what, void *out, int64_t cap, int64_t *len);

#define TITAN_INJECT_STEP 0
#define TITAN_INJECT_DOWNLOAD 1
#define TITAN_INJECT_DESTROY 2
int titan_debug_inject_failure_at(titan_ctx *ctx, int32_t site, int32_t after);

#ifdef __cplusplus
}
#endif

#endif
