// CPU oracle for the FP64 mass-spring step: TEST INFRASTRUCTURE ONLY.
//
// A host C++17 + OpenMP restatement of the reference's step arithmetic,
// operation by operation, with forces summed per mass in ascending spring
// index (the order a serial run of the reference's scatter would use).  The
// tests compare the HIP path's results with it and the benchmark times it as
// the CPU baseline; the product never links it.  Build: oracle/Makefile (g++ -O2 -fopenmp
// -ffp-contract=off; x86-64 has no FMA at the default -march, so no
// contraction can occur either way).
//
// Every expression keeps the reference's evaluation order and operand order
// (IEEE multiplication and addition are commutative, so only association
// matters; it is kept exactly).  pow(x, 2) is written x * x: both gcc and
// clang fold pow(x, 2.0) to a single multiplication, which is exact
// rounding of x squared (checked in tests/test_oracle.py).

#include "oracle.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// include/Titan/vec.h:33-163 — only the operations the step uses.
struct V {
    double x, y, z;
};
inline V vadd(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }   // vec.h:109-111
inline V vsub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }   // vec.h:113-115
inline V vneg(V a) { return {-a.x, -a.y, -a.z}; }                       // vec.h:85-87
// double * Vec -> Vec(v[i] * x) (vec.h:117-119); Vec * double is x * v (vec.h:121-123)
inline V vscale(V v, double x) { return {v.x * x, v.y * x, v.z * x}; }
inline V vdiv(V v, double x) { return {v.x / x, v.y / x, v.z / x}; }   // vec.h:133-135
// norm = sqrt(pow(x,2) + pow(y,2) + pow(z,2)) (vec.h:149-151)
inline double vnorm(V v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
// dot = (a * b).sum() = (a0 b0 + a1 b1) + a2 b2 (vec.cu:39-41, vec.h:153-155)
inline double vdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

const double NORMAL = 20000.0;  // src/object.cu:26

inline V load3(const double *x, const double *y, const double *z, int64_t i) {
    return {x[i], y[i], z[i]};
}

struct MassView {  // the fields of CUDA_MASS the constraints touch
    V pos, vel, force;
};

// CudaContactPlane::applyForce, src/object.cu:73-106.
inline void contact_plane(MassView &m, V n, double offset, double mu_k, double mu_s) {
    double disp = vdot(m.pos, n) - offset;
    V f_normal = vscale(n, vdot(m.force, n));
    if (disp < 0 && (mu_s > 0 || mu_k > 0)) {
        V v_perp = vsub(m.vel, vscale(n, vdot(m.vel, n)));
        double v_norm = vnorm(v_perp);
        if (v_norm > 1e-16) {
            double friction_mag = mu_k * vnorm(f_normal);
            m.force = vsub(m.force, vdiv(vscale(v_perp, friction_mag), v_norm));
        } else {
            V f_perp = vsub(m.force, f_normal);
            if (mu_s * vnorm(f_normal) > vnorm(f_perp)) {
                m.force = vsub(m.force, f_perp);
            }
        }
    }
    V contact = (disp < 0) ? vscale(n, -disp * NORMAL) : vscale(n, 0.0);
    m.force = vadd(m.force, contact);
}

// CudaBall::applyForce, src/object.cu:53-56.
inline void ball(MassView &m, V c, double r) {
    double dist = vnorm(vsub(m.pos, c));
    V add = (dist <= r) ? vdiv(vscale(vsub(m.pos, c), NORMAL), dist) : V{0, 0, 0};
    m.force = vadd(m.force, add);
}

// CudaConstraintPlane::applyForce, src/object.cu:115-124.
inline void constraint_plane(MassView &m, V n, double friction) {
    double normal_force = vdot(m.force, n);
    m.force = vadd(m.force, vscale(vneg(n), normal_force));
    double v_norm = vnorm(m.vel);
    if (v_norm >= 1e-16) {
        m.vel = vadd(m.vel, vscale(vneg(n), vdot(m.vel, n)));
        m.force = vadd(m.force, vdiv(vscale(m.vel, -friction * normal_force), v_norm));
    }
}

// CudaDirection::applyForce, src/object.cu:133-141.
inline void direction(MassView &m, V t, double friction) {
    V normal_force = vsub(m.force, vscale(t, vdot(m.force, t)));
    m.force = vadd(m.force, vneg(normal_force));
    if (vnorm(m.vel) >= 1e-16) {
        m.vel = vscale(t, vdot(m.vel, t));
        m.force = vadd(m.force, vscale(t, -vnorm(normal_force) * friction));
    }
}

// computeSpringForces body, src/sim.cu:1055-1065: force on the right
// endpoint; the left endpoint receives its negation (:1067-1073).
inline V spring_force(const orc_scene *s, int64_t i, double t) {
    const int64_t L = s->left[i], R = s->right[i];
    V lp = load3(s->px, s->py, s->pz, L), rp = load3(s->px, s->py, s->pz, R);
    V lv = load3(s->vx, s->vy, s->vz, L), rv = load3(s->vx, s->vy, s->vz, R);
    V temp = vsub(rp, lp);
    double scale = 1.0;
    int type = s->type ? s->type[i] : ORC_PASSIVE_SOFT;
    if (type == ORC_CONTRACT_THEN_EXPAND) {
        scale = (1 - 0.2 * std::sin(s->omega[i] * t));
    } else if (type == ORC_EXPAND_THEN_CONTRACT) {
        scale = (1 + 0.2 * std::sin(s->omega[i] * t));
    } else if (type == ORC_GENERAL) {
        // generalised law L0 (1 + a sin(omega t + phi)); with a = -/+0.2 and
        // phi = 0 it is bit-identical to the two reference kinds above
        scale = (1 + s->act_a[i] * std::sin(s->omega[i] * t + s->act_phi[i]));
    }
    V u = vdiv(temp, vnorm(temp));
    V force = vscale(u, s->k[i] * (s->rest[i] * scale - vnorm(temp)));
    force = vadd(force, vscale(u, vdot(vsub(lv, rv), u) * s->damping[i]));
    return force;
}

inline bool spring_live(const orc_scene *s, int64_t i) {
    // src/sim.cu:1052: a spring with a missing endpoint does nothing
    return s->left[i] >= 0 && s->right[i] >= 0;
}

inline bool is_fixed(const orc_scene *s, int64_t i) { return s->fixed && s->fixed[i]; }

// massForcesAndUpdate after the spring pass: src/sim.cu:1100-1139, then the
// integrator (:1141-1167).
inline void mass_update(const orc_scene *s, int64_t i, V force, int stage) {
    MassView m;
    m.pos = load3(s->px, s->py, s->pz, i);
    m.vel = load3(s->vx, s->vy, s->vz, i);
    m.force = force;
    const double mm = s->m[i];
    m.force = vadd(m.force, vscale(V{s->gx, s->gy, s->gz}, mm));      // :1104
    V ext = s->ex ? load3(s->ex, s->ey, s->ez, i) : V{0, 0, 0};
    m.force = vadd(m.force, ext);                                      // :1105
    for (int j = 0; j < s->n_planes; j++) {                            // :1109-1111
        const double *p = s->plane + 6 * j;
        contact_plane(m, V{p[0], p[1], p[2]}, p[3], p[4], p[5]);
    }
    for (int j = 0; j < s->n_balls; j++) {                             // :1113-1115
        const double *b = s->ball + 4 * j;
        ball(m, V{b[0], b[1], b[2]}, b[3]);
    }
    if (s->lc_ptr) {  // :1118-1132, each kind in insertion order
        const int64_t a = s->lc_ptr[i], e = s->lc_ptr[i + 1];
        const int order[4] = {ORC_CONTACT_PLANE, ORC_BALL, ORC_CONSTRAINT_PLANE, ORC_DIRECTION};
        for (int kind : order) {
            for (int64_t j = a; j < e; j++) {
                if (s->lc_type[j] != kind) continue;
                V v{s->lc_vec[3 * j], s->lc_vec[3 * j + 1], s->lc_vec[3 * j + 2]};
                double num = s->lc_num[j];
                // local contact planes carry no friction (src/object.cu:58-63)
                if (kind == ORC_CONTACT_PLANE) contact_plane(m, v, num, 0.0, 0.0);
                else if (kind == ORC_BALL) ball(m, v, num);
                else if (kind == ORC_CONSTRAINT_PLANE) constraint_plane(m, v, num);
                else direction(m, v, num);
            }
        }
    }
    if (vnorm(m.vel) != 0.0) {                                         // :1135-1138
        double norm = vnorm(m.vel);
        int c = s->drag ? s->drag[i] : 0;
        m.force = vadd(m.force, vdiv(vscale(m.vel, (double)(-c) * (norm * norm)), norm));
    }
    const double dt = s->dt;
    V acc, vel, pos;
    if (s->integrator == ORC_RK2) {
        double *bk = s->rk2_backup + 6 * i;
        acc = vdiv(m.force, mm);                                       // :1143, :1151
        if (stage == 0) {
            bk[0] = m.pos.x; bk[1] = m.pos.y; bk[2] = m.pos.z;         // :1144-1145
            bk[3] = m.vel.x; bk[4] = m.vel.y; bk[5] = m.vel.z;
            pos = vadd(m.pos, vscale(vscale(m.vel, 0.5), dt));         // :1147
            vel = vadd(m.vel, vscale(vscale(acc, 0.5), dt));           // :1148
        } else {
            pos = vadd(V{bk[0], bk[1], bk[2]}, vscale(m.vel, dt));     // :1152
            vel = vadd(V{bk[3], bk[4], bk[5]}, vscale(acc, dt));       // :1153
        }
        s->T[i] += 0.5 * dt;                                           // :1149, :1154
    } else if (s->integrator == ORC_VERLET) {
        const V prev{s->ax[i], s->ay[i], s->az[i]};
        vel = vadd(m.vel, vscale(vscale(vadd(prev, vdiv(m.force, mm)), 0.5), dt));  // :1157
        acc = vdiv(m.force, mm);                                       // :1158
        pos = vadd(m.pos, vadd(vscale(vel, dt), vscale(vscale(acc, 0.5), dt * dt)));  // :1159
        s->T[i] += dt;                                                 // :1160
    } else {
        acc = vdiv(m.force, mm);                                       // :1162
        vel = vadd(m.vel, vscale(acc, dt));                            // :1163
        pos = vadd(m.pos, vscale(vel, dt));                            // :1164
        s->T[i] += dt;                                                 // :1165
    }
    s->ax[i] = acc.x; s->ay[i] = acc.y; s->az[i] = acc.z;
    s->vx[i] = vel.x; s->vy[i] = vel.y; s->vz[i] = vel.z;
    s->px[i] = pos.x; s->py[i] = pos.y; s->pz[i] = pos.z;
}

int threads_of(const orc_scene *s) {
#ifdef _OPENMP
    return s->n_threads > 0 ? s->n_threads : omp_get_max_threads();
#else
    (void)s;
    return 1;
#endif
}

}  // namespace

extern "C" {

int orc_spring_forces(const orc_scene *s, double t, double *f) {
    const int nt = threads_of(s);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t i = 0; i < s->n_springs; i++) {
        V F = spring_live(s, i) ? spring_force(s, i, t) : V{0, 0, 0};
        f[3 * i] = F.x; f[3 * i + 1] = F.y; f[3 * i + 2] = F.z;
    }
    return 0;
}

int orc_build_incidence(const orc_scene *s, int64_t *row, int32_t *sid, int8_t *sgn) {
    const int64_t N = s->n_masses, S = s->n_springs;
    std::vector<int64_t> cnt(N + 1, 0);
    for (int64_t i = 0; i < S; i++) {
        if (!spring_live(s, i)) continue;
        if (s->left[i] >= N || s->right[i] >= N) return -1;
        cnt[s->left[i] + 1]++;
        cnt[s->right[i] + 1]++;
    }
    for (int64_t i = 0; i < N; i++) cnt[i + 1] += cnt[i];
    std::memcpy(row, cnt.data(), sizeof(int64_t) * (N + 1));
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    // ascending spring index per mass; a self-spring lists both roles,
    // right (+F) first, as the scatter adds to the right endpoint first
    for (int64_t i = 0; i < S; i++) {
        if (!spring_live(s, i)) continue;
        int64_t R = s->right[i], L = s->left[i];
        sid[fill[R]] = (int32_t)i; sgn[fill[R]++] = 1;
        sid[fill[L]] = (int32_t)i; sgn[fill[L]++] = -1;
    }
    return 0;
}

int orc_step(const orc_scene *s, double t, const int64_t *row, const int32_t *sid,
             const int8_t *sgn, double *fbuf, int stage) {
    orc_spring_forces(s, t, fbuf);
    const int nt = threads_of(s);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t i = 0; i < s->n_masses; i++) {
        if (is_fixed(s, i)) continue;                                  // :1100-1101
        V force{0, 0, 0};
        for (int64_t j = row[i]; j < row[i + 1]; j++) {
            const double *F = fbuf + 3 * (int64_t)sid[j];
            V Fv{F[0], F[1], F[2]};
            force = vadd(force, sgn[j] > 0 ? Fv : vneg(Fv));
        }
        mass_update(s, i, force, stage);
    }
    return 0;
}

int orc_run(const orc_scene *s, double t0, int64_t n_steps, double *t_out) {
    const int64_t N = s->n_masses, S = s->n_springs;
    std::vector<int64_t> row(N + 1);
    std::vector<int32_t> sid(2 * S + 1);
    std::vector<int8_t> sgn(2 * S + 1);
    std::vector<double> fbuf(3 * S + 3);
    if (orc_build_incidence(s, row.data(), sid.data(), sgn.data())) return -1;
    double T = t0;
    if (s->integrator == ORC_RK2) {
        orc_scene r = *s;
        std::vector<double> bk;
        if (!r.rk2_backup) {
            bk.assign(6 * N + 6, 0.0);
            r.rk2_backup = bk.data();
        }
        for (int64_t n = 0; n < n_steps; n++) {                         // src/sim.cu:1572-1582
            orc_step(&r, T, row.data(), sid.data(), sgn.data(), fbuf.data(), 0);
            T += 0.5 * s->dt;
            orc_step(&r, T, row.data(), sid.data(), sgn.data(), fbuf.data(), 1);
            T += 0.5 * s->dt;
        }
    } else {
        for (int64_t n = 0; n < n_steps; n++) {
            orc_step(s, T, row.data(), sid.data(), sgn.data(), fbuf.data(), 0);
            T += s->dt;                                                 // src/sim.cu:1590
        }
    }
    if (t_out) *t_out = T;
    return 0;
}

int orc_scatter_step(const orc_scene *s, double t) {
    const int64_t N = s->n_masses;
    std::vector<V> force(N, V{0, 0, 0});
    for (int64_t i = 0; i < s->n_springs; i++) {
        if (!spring_live(s, i)) continue;
        V F = spring_force(s, i, t);
        if (!is_fixed(s, s->right[i])) force[s->right[i]] = vadd(force[s->right[i]], F);
        if (!is_fixed(s, s->left[i])) force[s->left[i]] = vadd(force[s->left[i]], vneg(F));
    }
    for (int64_t i = 0; i < N; i++) {
        if (is_fixed(s, i)) continue;
        mass_update(s, i, force[i], 0);
    }
    return 0;
}

int64_t orc_lattice_springs(int nx, int ny, int nz) {
    int64_t c = 0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                int64_t a = (i != nx - 1) ? 2 : 1, b = (j != ny - 1) ? 2 : 1, d = (k != nz - 1) ? 2 : 1;
                c += a * b * d - 1;
                if (k != nz - 1) {
                    if (j != ny - 1) c++;
                    if (i != nx - 1) c++;
                    if (j != ny - 1 && i != nx - 1) c += 3;
                }
                if (j != ny - 1 && i != nx - 1) c++;
            }
    return c;
}

// Lattice::Lattice, src/object.cu:232-293.
int orc_lattice(double cx, double cy, double cz, double dx, double dy, double dz,
                int nx, int ny, int nz, double *pos, int32_t *left, int32_t *right,
                double *rest) {
    auto idx = [&](int i, int j, int k) -> int64_t { return k + (int64_t)j * nz + (int64_t)i * ny * nz; };
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                double fx = (nx > 1) ? (double)i / (nx - 1.0) - 0.5 : 0;
                double fy = (ny > 1) ? j / (ny - 1.0) - 0.5 : 0;
                double fz = (nz > 1) ? k / (nz - 1.0) - 0.5 : 0;
                int64_t a = idx(i, j, k);
                pos[3 * a] = fx * dx + cx;
                pos[3 * a + 1] = fy * dy + cy;
                pos[3 * a + 2] = fz * dz + cz;
            }
    int64_t c = 0;
    auto add = [&](int64_t L, int64_t R) { left[c] = (int32_t)L; right[c] = (int32_t)R; c++; };
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                for (int l = 0; l < ((i != nx - 1) ? 2 : 1); l++)
                    for (int m = 0; m < ((j != ny - 1) ? 2 : 1); m++)
                        for (int n = 0; n < ((k != nz - 1) ? 2 : 1); n++)
                            if (l != 0 || m != 0 || n != 0) add(idx(i, j, k), idx(i + l, j + m, k + n));
                if (k != nz - 1) {
                    if (j != ny - 1) add(idx(i, j, k + 1), idx(i, j + 1, k));
                    if (i != nx - 1) add(idx(i, j, k + 1), idx(i + 1, j, k));
                    if (j != ny - 1 && i != nx - 1) {
                        add(idx(i, j, k + 1), idx(i + 1, j + 1, k));
                        add(idx(i + 1, j, k + 1), idx(i, j + 1, k));
                        add(idx(i, j + 1, k + 1), idx(i + 1, j, k));
                    }
                }
                if (j != ny - 1 && i != nx - 1) add(idx(i, j + 1, k), idx(i + 1, j, k));
            }
    for (int64_t s = 0; s < c; s++) {  // object.cu:290-292: |right - left|
        V r{pos[3 * right[s]], pos[3 * right[s] + 1], pos[3 * right[s] + 2]};
        V l{pos[3 * left[s]], pos[3 * left[s] + 1], pos[3 * left[s] + 2]};
        rest[s] = vnorm(vsub(r, l));
    }
    return 0;
}

// test/testutil/utils.h:12-30, with the gravity constant as a parameter.
double orc_energy(const orc_scene *s, double g_abs) {
    double potential_g = 0, potential_s = 0, kinetic = 0;
    for (int64_t i = 0; i < s->n_masses; i++) {
        potential_g += g_abs * s->pz[i] * s->m[i];
        double vn = vnorm(load3(s->vx, s->vy, s->vz, i));
        kinetic += 0.5 * s->m[i] * (vn * vn);
    }
    for (int64_t i = 0; i < s->n_springs; i++) {
        if (!spring_live(s, i)) continue;
        double d = vnorm(vsub(load3(s->px, s->py, s->pz, s->left[i]), load3(s->px, s->py, s->pz, s->right[i]))) - s->rest[i];
        potential_s += s->k[i] * (d * d) / 2;
    }
    return potential_s + kinetic + potential_g;
}

}  // extern "C"
