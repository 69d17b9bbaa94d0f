/*
 * CPU oracle for the FP64 mass-spring step: TEST INFRASTRUCTURE ONLY.
 *
 * This is a host C++17 restatement of the reference's step arithmetic: the
 * tests compare the HIP path's results with it, and the benchmark times it
 * beside the HIP path as the CPU baseline.  Nothing in the product (libtitan_mi355x.so, the C++
 * facade) links or calls it; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.
 *
 * Functions and the reference code each one restates:
 *   orc_vec_*            include/Titan/vec.h:109-159, src/vec.cu:39-45
 *   orc_spring_forces    src/sim.cu:1046-1080 (computeSpringForces)
 *   orc_step             src/sim.cu:1090-1139,1141-1167 (massForcesAndUpdate,
 *                        CONSTRAINTS on) after the spring pass, with forces
 *                        summed per mass in ascending spring index; the
 *                        Euler branch (:1161-1167), the VERLET branch
 *                        (:1156-1160) or one RK2 stage (:1141-1155)
 *   orc_run              src/sim.cu:1569-1590 (the step driver; RK2: two
 *                        stages per step, T += 0.5 dt after each)
 *   orc_contact_plane    src/object.cu:26,73-106
 *   orc_ball             src/object.cu:53-56
 *   orc_constraint_plane src/object.cu:115-124
 *   orc_direction        src/object.cu:133-141
 *   orc_lattice          src/object.cu:232-293 (Lattice::Lattice)
 *   orc_scatter_step     the reference's serial scatter order (used only to
 *                        show that the per-mass gather sums in the same order)
 *
 * Pinning: every function above is checked in tests/test_oracle.py against
 * closed forms, hand-computed special cases, brute force and the physical
 * checks of the reference's own tests (the files under test/physics).
 */
#ifndef TITAN_ORACLE_H
#define TITAN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Spring kinds, in the reference's enum order (include/Titan/spring.h:17),
 * plus ORC_GENERAL for the generalised actuation law
 * rest * (1 + a * sin(omega * t + phi)). */
enum { ORC_PASSIVE_SOFT = 0, ORC_PASSIVE_STIFF = 1, ORC_CONTRACT_THEN_EXPAND = 2,
       ORC_EXPAND_THEN_CONTRACT = 3, ORC_GENERAL = 4 };

/* Local constraint kinds, in the reference's enum order
 * (include/Titan/object.h:231-233). */
enum { ORC_CONSTRAINT_PLANE = 0, ORC_CONTACT_PLANE = 1, ORC_BALL = 2, ORC_DIRECTION = 3 };

/* Integrators: the default Euler branch and the reference's two
 * compile-time alternatives (CMakeLists.txt:13-14). */
enum { ORC_EULER = 0, ORC_VERLET = 1, ORC_RK2 = 2 };

typedef struct {
    int64_t n_masses;
    int64_t n_springs;
    /* mass state, SoA, updated in place */
    double *px, *py, *pz, *vx, *vy, *vz;
    double *ax, *ay, *az;      /* acceleration written by each step */
    double *T;                 /* per-mass local time */
    const double *m;
    const double *ex, *ey, *ez;    /* extern force, may be NULL */
    const uint8_t *fixed;          /* may be NULL */
    const int32_t *drag;           /* int drag coefficient, may be NULL */
    /* springs: endpoints (-1 = no mass), parameters */
    const int32_t *left, *right;
    const double *k, *rest, *damping;
    const int32_t *type;           /* may be NULL = all passive */
    const double *omega;           /* used by type 2/3/4 */
    const double *act_a, *act_phi; /* used by type 4 */
    /* global contact planes: normal (already unit length), offset, mu_k, mu_s */
    int32_t n_planes;
    const double *plane;           /* 6 doubles per plane: nx ny nz d mu_k mu_s */
    int32_t n_balls;
    const double *ball;            /* 4 doubles per ball: cx cy cz r */
    /* per-mass local constraints, CSR over masses, in insertion order;
     * lc_vec: unit normal/tangent or ball centre; lc_num: offset, friction
     * or radius (src/object.cu:29-33,517-560, src/mass.cu:78-96) */
    const int64_t *lc_ptr;         /* may be NULL */
    const int32_t *lc_type;
    const double *lc_vec;          /* 3 per entry */
    const double *lc_num;
    /* globals */
    double gx, gy, gz;
    double dt;
    int32_t n_threads;             /* OpenMP threads, <=0 = default */
    int32_t integrator;            /* ORC_EULER, ORC_VERLET or ORC_RK2 */
    /* RK2: the state saved by a first stage for the second (6 per mass:
     * pos, vel), n_masses * 6 doubles; unused otherwise */
    double *rk2_backup;
} orc_scene;

/* Incident-spring CSR (ascending spring index per mass). Returns 0 on
 * success.  row must hold n_masses+1 entries, sid/sgn 2*n_springs. */
int orc_build_incidence(const orc_scene *s, int64_t *row, int32_t *sid, int8_t *sgn);

/* One step (one RK2 stage: stage 0 or 1) at simulation time t (the host T
 * before it is advanced).  The incidence CSR comes from orc_build_incidence;
 * fbuf holds 3*n_springs doubles of scratch. */
int orc_step(const orc_scene *s, double t, const int64_t *row, const int32_t *sid,
             const int8_t *sgn, double *fbuf, int stage);

/* n steps; T is accumulated as T += dt from t0 (src/sim.cu:1590), or for
 * RK2 as T += 0.5 dt after each stage (:1577,1582); the final time is
 * returned in *t_out. */
int orc_run(const orc_scene *s, double t0, int64_t n_steps, double *t_out);

/* Serial reference-order scatter step: forces are added spring by spring in
 * ascending index to both endpoints (src/sim.cu:1067-1073 run serially). */
int orc_scatter_step(const orc_scene *s, double t);

/* Spring forces alone (force on the right endpoint, 3 per spring). */
int orc_spring_forces(const orc_scene *s, double t, double *f);

/* Mass count and spring count of an nx*ny*nz lattice. */
int64_t orc_lattice_springs(int nx, int ny, int nz);

/* Lattice::Lattice: positions (3 per mass, x y z interleaved) and springs
 * (left, right, rest) in the reference's creation order. */
int orc_lattice(double cx, double cy, double cz, double dx, double dy, double dz,
                int nx, int ny, int nz, double *pos, int32_t *left, int32_t *right,
                double *rest);

/* Energy as in test/testutil/utils.h:12-30 with gravity g_abs (the
 * reference hard-codes 9.8). */
double orc_energy(const orc_scene *s, double g_abs);

#ifdef __cplusplus
}
#endif

#endif
