// Step kernels for gfx950 (MI355X): the FP64 mass-spring step as one fused
// pass per step.  Replaces the reference's computeSpringForces (one thread
// per spring, FP64 CAS-loop scatter, src/sim.cu:1046-1080) and
// massForcesAndUpdate (one thread per mass over 200-byte AoS records,
// src/sim.cu:1090-1169) with a mass-centric gather: one lane per mass walks
// its incident springs in ascending spring index (SELL-64 incidence order),
// recomputes each spring force exactly as the reference writes it, sums the
// forces in that fixed order, then applies gravity, extern force, global
// planes and balls, local constraints, drag and the symplectic Euler update
// in registers and writes the other state buffer.
//
// Every FP64 expression keeps the reference's association and operand order
// and the library is built with -ffp-contract=off, so a passive spring
// network gives bit-identical states to a serial run of the reference's
// arithmetic summed in ascending spring index.  sqrt and division are the
// correctly rounded IEEE operations; sin (actuated springs only) is the
// device math library's.
#include <hip/hip_runtime.h>

#include "fastmath.h"
#include "kernels.h"
#include "layout.h"

namespace titan_mi355x {

namespace {

constexpr double kNormal = 20000.0;  // src/object.cu:26

// Blocks are dispatched round-robin over the 8 XCDs (block w runs on XCD
// w % 8).  Map them so that each XCD walks one contiguous range of chunks:
// lattice neighbours a plane apart are then read through the same L2.
__device__ __forceinline__ int64_t xcd_block(int64_t w, int64_t G) {
    const int64_t q = G >> 3, r = G & 7;
    const int64_t x = w & 7, k = w >> 3;
    return x * q + (x < r ? x : r) + k;
}

// The order a step launch walks its tiles or chunks in: forward on even
// launch slots, backward on odd ones, so that each launch starts on the
// incidence data and states the previous launch read last, which the
// device's 256 MB memory-side cache still holds (the tiles of a launch are
// independent, so the order changes no result).  Measured against walking
// forward every launch: 1M lattice 86.9 against 91.6 us per step, 2M kNN
// cloud 176.9 against 187.5 us, 4096 robots 303.4 against 316.6 us.
__device__ __forceinline__ int64_t step_block(const StepArgs &a) {
    const int64_t w = xcd_block(blockIdx.x, gridDim.x);
    return (a.slot & 1) ? (int64_t)gridDim.x - 1 - w : w;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
// Wave votes over the active lanes, from the lane mask of the predicate (a
// compare's mask is used as it is, without first making a VGPR of it).
__device__ __forceinline__ bool wave_all(bool p) { return __builtin_amdgcn_ballot_w64(p) == __builtin_amdgcn_read_exec(); }
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// Load base[byte_off / sizeof(T)] with a 32-bit byte offset from a
// wave-uniform base, so the access compiles to a scalar-base global load.
// Every per-component array is below 4 GiB (the arena caps a component at
// 2^29 doubles).
template <class T>
__device__ __forceinline__ T ldo(const T *__restrict__ base, uint32_t byte_off) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}

struct MassRegs {
    double px, py, pz, vx, vy, vz, fx, fy, fz;
};

// CudaContactPlane::applyForce, src/object.cu:73-106.  For disp >= 0 the
// reference adds 0 * normal, which changes nothing but the sign of an
// exactly zero force component; that addition is skipped.
__device__ __forceinline__ void contact_plane(MassRegs &r, double nx, double ny, double nz,
                                              double off, double mu_k, double mu_s) {
    const double disp = (r.px * nx + r.py * ny + r.pz * nz) - off;
    if (disp < 0) {
        const double dF = r.fx * nx + r.fy * ny + r.fz * nz;
        const double fnx = nx * dF, fny = ny * dF, fnz = nz * dF;
        if (mu_s > 0 || mu_k > 0) {
            const double dv = r.vx * nx + r.vy * ny + r.vz * nz;
            const double vpx = r.vx - nx * dv, vpy = r.vy - ny * dv, vpz = r.vz - nz * dv;
            const double vn = sqrt(vpx * vpx + vpy * vpy + vpz * vpz);
            if (vn > 1e-16) {
                const double fm = mu_k * sqrt(fnx * fnx + fny * fny + fnz * fnz);
                r.fx = r.fx - (vpx * fm) / vn;
                r.fy = r.fy - (vpy * fm) / vn;
                r.fz = r.fz - (vpz * fm) / vn;
            } else {
                const double fpx = r.fx - fnx, fpy = r.fy - fny, fpz = r.fz - fnz;
                if (mu_s * sqrt(fnx * fnx + fny * fny + fnz * fnz) >
                    sqrt(fpx * fpx + fpy * fpy + fpz * fpz)) {
                    r.fx = r.fx - fpx;
                    r.fy = r.fy - fpy;
                    r.fz = r.fz - fpz;
                }
            }
        }
        const double c = -disp * kNormal;
        r.fx = r.fx + nx * c;
        r.fy = r.fy + ny * c;
        r.fz = r.fz + nz * c;
    }
}

// CudaBall::applyForce, src/object.cu:53-56 (constant magnitude NORMAL).
__device__ __forceinline__ void ball(MassRegs &r, double cx, double cy, double cz, double rad) {
    const double dx = r.px - cx, dy = r.py - cy, dz = r.pz - cz;
    const double dist = sqrt(dx * dx + dy * dy + dz * dz);
    if (dist <= rad) {
        r.fx = r.fx + (dx * kNormal) / dist;
        r.fy = r.fy + (dy * kNormal) / dist;
        r.fz = r.fz + (dz * kNormal) / dist;
    }
}

// CudaConstraintPlane::applyForce, src/object.cu:115-124.
__device__ __forceinline__ void constraint_plane(MassRegs &r, double nx, double ny, double nz,
                                                 double fr) {
    const double nf = r.fx * nx + r.fy * ny + r.fz * nz;
    r.fx = r.fx + (-nx) * nf;
    r.fy = r.fy + (-ny) * nf;
    r.fz = r.fz + (-nz) * nf;
    const double vn = sqrt(r.vx * r.vx + r.vy * r.vy + r.vz * r.vz);
    if (vn >= 1e-16) {
        const double vd = r.vx * nx + r.vy * ny + r.vz * nz;
        r.vx = r.vx + (-nx) * vd;
        r.vy = r.vy + (-ny) * vd;
        r.vz = r.vz + (-nz) * vd;
        const double c = -fr * nf;
        r.fx = r.fx + (r.vx * c) / vn;
        r.fy = r.fy + (r.vy * c) / vn;
        r.fz = r.fz + (r.vz * c) / vn;
    }
}

// CudaDirection::applyForce, src/object.cu:133-141.
__device__ __forceinline__ void direction(MassRegs &r, double tx, double ty, double tz,
                                          double fr) {
    const double dft = r.fx * tx + r.fy * ty + r.fz * tz;
    const double nfx = r.fx - tx * dft, nfy = r.fy - ty * dft, nfz = r.fz - tz * dft;
    r.fx = r.fx + (-nfx);
    r.fy = r.fy + (-nfy);
    r.fz = r.fz + (-nfz);
    if (sqrt(r.vx * r.vx + r.vy * r.vy + r.vz * r.vz) >= 1e-16) {
        const double dvt = r.vx * tx + r.vy * ty + r.vz * tz;
        r.vx = tx * dvt;
        r.vy = ty * dvt;
        r.vz = tz * dvt;
        const double c = -sqrt(nfx * nfx + nfy * nfy + nfz * nfz) * fr;
        r.fx = r.fx + tx * c;
        r.fy = r.fy + ty * c;
        r.fz = r.fz + tz * c;
    }
}

__device__ __forceinline__ void apply_local(MassRegs &r, const LocalC &c) {
    switch (c.kind) {
        case 1:  // CONTACT_PLANE: local planes carry no friction (src/object.cu:58-63)
            contact_plane(r, c.v[0], c.v[1], c.v[2], c.num, 0.0, 0.0);
            break;
        case 2:
            ball(r, c.v[0], c.v[1], c.v[2], c.num);
            break;
        case 0:
            constraint_plane(r, c.v[0], c.v[1], c.v[2], c.num);
            break;
        default:
            direction(r, c.v[0], c.v[1], c.v[2], c.num);
            break;
    }
}

// What a launch's integrator reads besides the mass's state: Verlet the
// previous step's acceleration (p[0..2]), an RK2 second stage the state its
// first stage left (p[0..5]: pos, vel).  Loaded early, with the other loads.
struct IntegIn {
    double p[6];
};
__device__ __forceinline__ IntegIn load_integ(const StepArgs &a, int64_t i, int stage) {
    IntegIn q;
#pragma unroll
    for (int k = 0; k < 6; ++k) q.p[k] = 0.0;
    if (a.integ == kIntVerlet) {
        q.p[0] = a.acc[i];
        q.p[1] = a.acc[a.ncap + i];
        q.p[2] = a.acc[2 * a.ncap + i];
    } else if (a.integ == kIntRK2 && stage == 1) {
        const double2 *__restrict__ B = reinterpret_cast<const double2 *>(a.rk) + 3 * i;
        const double2 x = B[0], y = B[1], z = B[2];
        q.p[0] = x.x;
        q.p[1] = x.y;
        q.p[2] = y.x;
        q.p[3] = y.y;
        q.p[4] = z.x;
        q.p[5] = z.y;
    }
    return q;
}

// Everything of massForcesAndUpdate after the spring sum (src/sim.cu:1104-1167),
// applied to registers, with the context's integrator.  Returns the
// acceleration.
// Where the mass pass reads what is the same for every mass of a body and
// every step of a launch: the global planes (8 doubles each: the launch's
// array, or a copy in LDS) and the body's shared local-constraint list.
struct MassEnv {
    const double *__restrict__ planes;
    int64_t lb, ln;
};
__device__ __forceinline__ MassEnv mass_env(const StepArgs &a, int b) { return {a.planes, a.b_lc[b], a.b_lcn[b]}; }

__device__ __forceinline__ void mass_pass(const StepArgs &a, MassRegs &r, int64_t i, uint32_t fl,
                                          const MassEnv &env, double m, int stage, const IntegIn &in,
                                          double &ax, double &ay, double &az) {
    r.fx = r.fx + a.gx * m;                                    // :1104
    r.fy = r.fy + a.gy * m;
    r.fz = r.fz + a.gz * m;
    if (fl & kExt) {                                           // :1105
        r.fx = r.fx + a.ext[i];
        r.fy = r.fy + a.ext[a.ncap + i];
        r.fz = r.fz + a.ext[2 * a.ncap + i];
    }
    for (int q = 0; q < a.nplanes; ++q) {                      // :1109-1111
        const double *P = env.planes + 8 * q;
        contact_plane(r, P[0], P[1], P[2], P[3], P[4], P[5]);
    }
    if (a.bg_start) {                                          // :1113-1115
        // only the balls listed for the mass's grid cell, in creation order:
        // any other ball is farther than its radius from every point of the
        // cell (with margins for rounding), so it would add a zero vector
        const double fx = (r.px - a.bg_x0) * a.bg_inv_h;
        const double fy = (r.py - a.bg_y0) * a.bg_inv_h;
        const double fz = (r.pz - a.bg_z0) * a.bg_inv_h;
        if (fx >= 0.0 && fx < (double)a.bg_nx && fy >= 0.0 && fy < (double)a.bg_ny && fz >= 0.0 &&
            fz < (double)a.bg_nz) {
            const int64_t cell = ((int64_t)(int)fz * a.bg_ny + (int)fy) * a.bg_nx + (int)fx;
            // records are loaded four at a time, then tested in order, so that
            // a typical list (3.9 balls per mass on the 2M kNN cloud with 1000
            // balls, 4.9 iterations per wave) costs one round trip to the
            // cache rather than one per ball; a load past the list's end reads
            // the next list's first record or the spare one that ends them.
            // Eight at a time took the step to 132 VGPRs and was 7% slower.
            const double2 *__restrict__ R = reinterpret_cast<const double2 *>(a.bg_rec);
            const int32_t q0 = a.bg_start[cell], q1 = a.bg_start[cell + 1];
            constexpr int PB = 4;
            for (int32_t qb = q0; qb < q1; qb += PB) {
                double2 c[PB][2];  // (cx, cy), (cz, r)
#pragma unroll
                for (int j = 0; j < PB; ++j) {
                    const int32_t q = min(qb + j, q1);
                    c[j][0] = R[2 * q];
                    c[j][1] = R[2 * q + 1];
                }
#pragma unroll
                for (int j = 0; j < PB; ++j)
                    if (qb + j < q1) ball(r, c[j][0].x, c[j][0].y, c[j][1].x, c[j][1].y);
            }
        }
    } else {
        for (int q = 0; q < a.nballs; ++q) {
            const double *B = a.balls + 4 * q;
            ball(r, B[0], B[1], B[2], B[3]);
        }
    }
    int64_t lb, ln;                                            // :1118-1132
    if (fl & kLocalLC) {
        const int64_t v = a.mlc[i];
        lb = v >> 20;
        ln = v & 0xFFFFF;
    } else {
        lb = env.lb;
        ln = env.ln;
    }
    for (int64_t q = 0; q < ln; ++q) apply_local(r, a.lc[lb + q]);
    if (fl & kDrag) {                                          // :1135-1138
        const double vn = sqrt(r.vx * r.vx + r.vy * r.vy + r.vz * r.vz);
        if (vn != 0.0) {
            const double c = (double)(-a.drag[i]) * (vn * vn);
            r.fx = r.fx + (r.vx * c) / vn;
            r.fy = r.fy + (r.vy * c) / vn;
            r.fz = r.fz + (r.vz * c) / vn;
        }
    }
    const double dt = a.dt;
    ax = r.fx / m;                                             // :1143, :1151, :1158, :1162
    ay = r.fy / m;
    az = r.fz / m;
    if (a.integ == kIntRK2) {
        if (stage == 0) {                                      // :1142-1149
            double2 *__restrict__ B = reinterpret_cast<double2 *>(a.rk) + 3 * i;
            B[0] = make_double2(r.px, r.py);
            B[1] = make_double2(r.pz, r.vx);
            B[2] = make_double2(r.vy, r.vz);
            r.px = r.px + (0.5 * r.vx) * dt;
            r.py = r.py + (0.5 * r.vy) * dt;
            r.pz = r.pz + (0.5 * r.vz) * dt;
            r.vx = r.vx + (0.5 * ax) * dt;
            r.vy = r.vy + (0.5 * ay) * dt;
            r.vz = r.vz + (0.5 * az) * dt;
        } else {                                               // :1150-1155
            r.px = in.p[0] + r.vx * dt;
            r.py = in.p[1] + r.vy * dt;
            r.pz = in.p[2] + r.vz * dt;
            r.vx = in.p[3] + ax * dt;
            r.vy = in.p[4] + ay * dt;
            r.vz = in.p[5] + az * dt;
        }
    } else if (a.integ == kIntVerlet) {                        // :1156-1160
        const double dt2 = dt * dt;                            // pow(dt, 2)
        r.vx = r.vx + (0.5 * (in.p[0] + ax)) * dt;
        r.vy = r.vy + (0.5 * (in.p[1] + ay)) * dt;
        r.vz = r.vz + (0.5 * (in.p[2] + az)) * dt;
        r.px = r.px + (r.vx * dt + (0.5 * ax) * dt2);
        r.py = r.py + (r.vy * dt + (0.5 * ay) * dt2);
        r.pz = r.pz + (r.vz * dt + (0.5 * az) * dt2);
    } else {                                                   // :1161-1167
        r.vx = r.vx + ax * dt;
        r.vy = r.vy + ay * dt;
        r.vz = r.vz + az * dt;
        r.px = r.px + r.vx * dt;
        r.py = r.py + r.vy * dt;
        r.pz = r.pz + r.vz * dt;
    }
}

// mass_pass for a mass without extern force, drag or local constraints, in a
// launch without balls, under symplectic Euler: the same operations in the
// same order (src/sim.cu:1104-1111, :1143, :1161-1167), with none of the
// branches the other cases need.
__device__ __forceinline__ void mass_pass_lean(const StepArgs &a, MassRegs &r, const MassEnv &env, double m,
                                               double &ax, double &ay, double &az) {
    r.fx = r.fx + a.gx * m;
    r.fy = r.fy + a.gy * m;
    r.fz = r.fz + a.gz * m;
    for (int q = 0; q < a.nplanes; ++q) {
        const double *P = env.planes + 8 * q;
        contact_plane(r, P[0], P[1], P[2], P[3], P[4], P[5]);
    }
    const double dt = a.dt;
    ax = r.fx / m;
    ay = r.fy / m;
    az = r.fz / m;
    r.vx = r.vx + ax * dt;
    r.vy = r.vy + ay * dt;
    r.vz = r.vz + az * dt;
    r.px = r.px + r.vx * dt;
    r.py = r.py + r.vy * dt;
    r.pz = r.pz + r.vz * dt;
}

}  // namespace

// Spring forces on one mass, src/sim.cu:1055-1073, summed in ascending
// spring index: the lane walks its d incidences (SELL-64 groups of G
// starting at entry `ebase` of its chunk) and fetches the other endpoint's
// state (px py | pz vx | vy vz) with fetch(name, q0, q1, q2).  G = 4: four
// uint32 names per 16-byte load; G = 2: two uint16 names per 4-byte load.
//
// Each incident spring is evaluated from this mass's side: tt = p_other -
// p_self and dv = v_self - v_other.  If this mass is the left endpoint, tt
// and dv are the reference's temp and dv and G below is its F; this mass
// receives -F = -G.  If it is the right endpoint, the reference's temp and
// dv are -tt and -dv exactly (x - y == -(y - x) under round-to-nearest), so
// its F is -G exactly and this mass receives +F = -G.  Either way the
// contribution is -G, bit for bit (an exact zero may differ in sign).
//
// The next group's names, rest lengths and constants are loaded while the
// current group is evaluated.
template <int G>
struct Names;
template <>
struct Names<4> {
    uint4 v;
    __device__ static Names load(const uint32_t *base, int64_t ebase, int lane, int g) {
        return {reinterpret_cast<const uint4 *>(base + ebase)[(int64_t)g * kChunk + lane]};
    }
    __device__ void get(uint32_t (&o)[4]) const {
        o[0] = v.x & ~kActBit32;
        o[1] = v.y & ~kActBit32;
        o[2] = v.z & ~kActBit32;
        o[3] = v.w & ~kActBit32;
    }
    __device__ bool actuated(int q) const {
        const uint32_t w = q == 0 ? v.x : q == 1 ? v.y : q == 2 ? v.z : v.w;
        return (w & kActBit32) != 0;
    }
};
template <>
struct Names<2> {
    uint32_t v;
    __device__ static Names load(const uint32_t *base, int64_t ebase, int lane, int g) {
        return {base[ebase + (int64_t)g * kChunk + lane]};
    }
    __device__ void get(uint32_t (&o)[2]) const {
        o[0] = v & (0xffffu & ~kActBit16);
        o[1] = (v >> 16) & ~kActBit16;
    }
    __device__ bool actuated(int q) const { return ((v >> (16 * q)) & kActBit16) != 0; }
};

// The body's spring constant and damping: by value when every spring of
// the body shares them, else per incidence (kd bit 0: k, bit 1: damping;
// uniform over the wave, so the loads it selects need no divergent branch).
struct BodyKD {
    double k, d;
    uint32_t kd;
};

// Where a walk reads incidence data: the arena (global memory), addressed
// through the kernel arguments or through pointers held by the walk, or a
// copy in LDS for a resident body (entry indices then relative to the body's
// first).  Entry offsets are even, so rest lengths and constants load as
// 16-byte pairs.  Which form the compiler allocates fewer registers for
// differs by kernel; each kernel uses the one measured best.
struct ArenaSrc {
    const StepArgs &a;
    __device__ const uint32_t *nbr() const { return a.inc_nbr; }
    __device__ const double *rest() const { return a.inc_rest; }
    __device__ const double *k() const { return a.inc_k; }
    __device__ const double *damp() const { return a.inc_damp; }
};
struct PtrSrc {
    const uint32_t *n;
    const double *r, *kk, *d;
    __device__ const uint32_t *nbr() const { return n; }
    __device__ const double *rest() const { return r; }
    __device__ const double *k() const { return kk; }
    __device__ const double *damp() const { return d; }
};

// One group's incidence data in registers: the names, rest lengths (or
// actuator indices), per-spring constants (bodies that have them) and the
// actuated rest lengths (actuated only).  KD says where k and damping come
// from: bit 0 set, k per incidence; bit 1 set, damping per incidence; the
// other constant travels by value.  KD = kKdRuntime reads the choice from
// the body (bkd.kd) at run time; a compile-time KD issues the same loads
// for every group, so that the compiler's wait counts are exact and a walk
// can keep two groups' loads in flight behind the one it evaluates.
constexpr int kKdRuntime = 4;
template <bool ACT, int G, int KD = kKdRuntime>
struct Group {
    Names<G> n;
    double rest[G], k[G], damp[G], f[G];
    template <class Src>
    __device__ __forceinline__ void load(const Src &src, int64_t ebase, int lane, int g, const BodyKD &bkd) {
        constexpr int P = G / 2;  // 16-byte pairs per group and lane
        n = Names<G>::load(src.nbr(), ebase, lane, g);
        const int64_t e2 = ebase / 2 + ((int64_t)g * kChunk + lane) * P;  // double2 index
        const bool pk = KD == kKdRuntime ? (bkd.kd & 1u) != 0 : (KD & 1) != 0;
        const bool pd = KD == kKdRuntime ? (bkd.kd & 2u) != 0 : (KD & 2) != 0;
#pragma unroll
        for (int h = 0; h < P; ++h) {
            const double2 rr = reinterpret_cast<const double2 *>(src.rest())[e2 + h];
            rest[2 * h] = rr.x;
            rest[2 * h + 1] = rr.y;
            if (pk) {
                const double2 kk = reinterpret_cast<const double2 *>(src.k())[e2 + h];
                k[2 * h] = kk.x;
                k[2 * h + 1] = kk.y;
            } else {
                k[2 * h] = k[2 * h + 1] = bkd.k;
            }
            if (pd) {
                const double2 dd = reinterpret_cast<const double2 *>(src.damp())[e2 + h];
                damp[2 * h] = dd.x;
                damp[2 * h + 1] = dd.y;
            } else {
                damp[2 * h] = damp[2 * h + 1] = bkd.d;
            }
        }
    }
    // The actuated rest lengths of the group's actuated incidences (rc: this
    // step's), skipped for an incidence slot where no lane of the wave has
    // one.  Issued before the next group's loads, so that waiting for them
    // does not wait for those.
    __device__ __forceinline__ void fetch_f(const double *__restrict__ rc) {
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const bool on = n.actuated(q);
            f[q] = rest[q];
            if (wave_any(on)) {
                const double v = rc[on ? __double_as_longlong(rest[q]) : 0];
                f[q] = on ? v : rest[q];
            }
        }
    }
};

// Actuated rest lengths at time T of the n actuator records at rec into
// dst[0, n), one per lane in turn: rest * (1 + a sin(omega T + phi))
// (src/sim.cu:1058-1066, the reference's two kinds being a = -0.2 and 0.2
// with phi = 0), evaluated in the reference's order.
__device__ __forceinline__ void actuate_records(const double *rec, int n, double T, double *__restrict__ dst,
                                                int lane, int stride) {
    const double2 *__restrict__ A = reinterpret_cast<const double2 *>(rec);
    for (int q = lane; q < n; q += stride) {
        const double2 aw = A[2 * q], pr = A[2 * q + 1];  // (a, omega), (phi, rest)
        dst[q] = pr.y * (1 + aw.x * sin(aw.y * T + pr.x));
    }
}
// The same for actuators [first, first + n) of the arena, into dst[first, first + n).
__device__ __forceinline__ void actuate_slice(const StepArgs &a, int64_t first, int n, double T,
                                              double *__restrict__ dst, int lane, int stride) {
    actuate_records(a.act + 4 * first, n, T, dst + first, lane, stride);
}

// Add -G of one incident spring (axis nrm, u) to the mass's force.
template <bool ACT, class Grp>
__device__ __forceinline__ void add_spring(MassRegs &r, const Grp &cur, int q, const double2 (&nb)[3],
                                           double nrm, double ux, double uy, double uz) {
    const double dvx = r.vx - nb[1].y, dvy = r.vy - nb[2].x, dvz = r.vz - nb[2].y;
    const double rl = ACT ? cur.f[q] : cur.rest[q];  // ACT: the actuated rest length where marked
    const double k = cur.k[q];
    const double dmp = cur.damp[q];
    const double mag = k * (rl - nrm);
    const double dd = (dvx * ux + dvy * uy + dvz * uz) * dmp;
    r.fx = r.fx - (ux * mag + ux * dd);
    r.fy = r.fy - (uy * mag + uy * dd);
    r.fz = r.fz - (uz * mag + uz * dd);
}

// Evaluate the K groups of one register set (N = G K incidences, the first
// being incidence G g0 of the lane; the lane has d).  When every lane of the
// wave has all N and they are in the fast range, the N evaluations are
// straight-line code that interleaves; otherwise each incidence that exists
// takes the general path.
template <bool ACT, int G, int K, int KD, class Fetch>
__device__ __forceinline__ void eval_set(MassRegs &r, const Group<ACT, G, KD> (&set)[K], int g0, int d,
                                         Fetch &fetch) {
    constexpr int N = G * K;
    double2 nb[N][3];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        uint32_t oq[G];
        set[j].n.get(oq);
#pragma unroll
        for (int q = 0; q < G; ++q) fetch(oq[q], nb[j * G + q][0], nb[j * G + q][1], nb[j * G + q][2]);
    }
    double tx[N], ty[N], tz[N], n2[N];
    uint64_t fast = __builtin_amdgcn_ballot_w64(G * g0 + N <= d);  // lanes that may take the fast path
#pragma unroll
    for (int e = 0; e < N; ++e) {
        tx[e] = nb[e][0].x - r.px;
        ty[e] = nb[e][0].y - r.py;
        tz[e] = nb[e][1].x - r.pz;
        n2[e] = tx[e] * tx[e] + ty[e] * ty[e] + tz[e] * tz[e];
        fast &= fast_lanes_nz(n2[e], tx[e], ty[e], tz[e]);
    }
    if (fast == __builtin_amdgcn_read_exec()) {
#pragma unroll
        for (int e = 0; e < N; ++e) {
            const double nrm = fast_sqrt(n2[e]);
            const double y = fast_recip(nrm);
            const double ux = fast_div(tx[e], nrm, y), uy = fast_div(ty[e], nrm, y), uz = fast_div(tz[e], nrm, y);
            add_spring<ACT>(r, set[e / G], e % G, nb[e], nrm, ux, uy, uz);
        }
    } else {
#pragma unroll
        for (int e = 0; e < N; ++e) {
            if (G * g0 + e >= d) break;
            double nrm, ux, uy, uz;
            unit_axis(tx[e], ty[e], tz[e], nrm, ux, uy, uz);
            add_spring<ACT>(r, set[e / G], e % G, nb[e], nrm, ux, uy, uz);
        }
    }
}

// Spring forces of one lane, walking ng groups (uniform over the wave; a
// lane uses the first ceil(d / G)).  The group data streams from HBM, so two
// register sets of K consecutive groups alternate: one set's K groups are
// loaded while the other's are evaluated, keeping K groups of loads in
// flight per wave without copies (a refill index past the last group is
// clamped to it and re-reads a group already loaded).  prime() issues the
// first set's loads, so that a kernel can start them before it waits for
// anything else.  K = 1: a set of G incidences.  (Sets of four incidences
// with G = 2 took the tiled step to 164 VGPRs and 3 waves per SIMD; one
// group per set keeps it at 110 and 4 waves, measured 5-10% faster.)
template <bool ACT, int G, class Src>
struct SpringWalk {
    static_assert(G == 2 || G == 4, "groups of 2 or 4");
    static constexpr int K = 1;  // groups per register set
    Group<ACT, G> A[K], B[K];
    const Src &src;
    int64_t ebase;
    int ng, lane;
    BodyKD bkd;

    __device__ explicit SpringWalk(const Src &s) : src(s) {}

    __device__ __forceinline__ void prime(int64_t eb, int n, int ln, const BodyKD &kd) {
        ebase = eb;
        ng = n;
        lane = ln;
        bkd = kd;
        if (ng <= 0) return;
#pragma unroll
        for (int j = 0; j < K; ++j) A[j].load(src, ebase, lane, min(j, ng - 1), bkd);
    }

    // rc: this step's actuation factors (actuated only)
    template <class Fetch>
    __device__ __forceinline__ void run(MassRegs &r, int d, const double *rc, Fetch fetch) {
        r.fx = 0.0;
        r.fy = 0.0;
        r.fz = 0.0;
        const int last = ng - 1;
        for (int g = 0; g < ng; g += 2 * K) {
            if (ACT) {
#pragma unroll
                for (int j = 0; j < K; ++j) A[j].fetch_f(rc);
            }
#pragma unroll
            for (int j = 0; j < K; ++j) B[j].load(src, ebase, lane, min(g + K + j, last), bkd);
            eval_set<ACT, G, K>(r, A, g, d, fetch);
            if (g + K >= ng) break;
            if (ACT) {
#pragma unroll
                for (int j = 0; j < K; ++j) B[j].fetch_f(rc);
            }
#pragma unroll
            for (int j = 0; j < K; ++j) A[j].load(src, ebase, lane, min(g + 2 * K + j, last), bkd);
            eval_set<ACT, G, K>(r, B, g + K, d, fetch);
        }
    }
};

// The walk of the tiled step: D register sets of one group each in a ring,
// so that while one group is evaluated the next D - 1 groups' loads are in
// flight, with the body's constant mode KD fixed at compile time (each
// group issues the same loads, and the waits the compiler places allow the
// later groups to stay outstanding).  D = 3 on the 1M lattice, the 2M kNN
// cloud and the robot population was 1%, 4% and 6% faster than two
// register sets (measured on one box, alternating).  Deeper rings cost
// registers: D = 4 and 5 take the passive step from 122 to 134 and 148
// VGPRs (three waves per SIMD instead of four) and were 5-15% slower on all
// three workloads; D = 4 held to four waves spills and was no faster.
#ifndef TITAN_WALK_DEPTH
#define TITAN_WALK_DEPTH 3
#endif
template <bool ACT, int KD, class Src, int D = TITAN_WALK_DEPTH>
struct SpringWalkD {
    static_assert(D >= 2 && D <= 8, "ring of 2 to 8 register sets");
    static constexpr int G = kGroupTiled;
    typedef Group<ACT, G, KD> Grp;
    Grp R[D];
    const Src &src;
    int64_t ebase;
    int ng, lane, dl;
    BodyKD bkd;

    __device__ explicit SpringWalkD(const Src &s) : src(s) {}

    // The lane whose entries a load of group q reads.  With per-spring
    // constants (16 B more per entry), a lane with no incidence in the group
    // reads lane 0's entries instead of its padding, so that the padding's
    // lines are not fetched (the values are not used): on the 2M kNN cloud,
    // 11% of the entries, 2% of the step's time.  Without them the select
    // costs the lattice more than it saves, and the walk reads every lane.
    __device__ __forceinline__ int src_lane(int q) const {
        if constexpr ((KD & 1) != 0 && KD != kKdRuntime) return G * q < dl ? lane : 0;
        return lane;
    }

    // d: the lane's incidence count
    __device__ __forceinline__ void prime(int64_t eb, int n, int ln, const BodyKD &kd, int d) {
        ebase = eb;
        ng = n;
        lane = ln;
        bkd = kd;
        dl = d;
        if (ng <= 0) return;
#pragma unroll
        for (int j = 0; j < D - 1; ++j) {
            const int q = min(j, ng - 1);
            R[j].load(src, ebase, src_lane(q), q, bkd);
        }
    }

    // rc: this step's actuation factors (actuated only).  A refill index
    // past the last group is clamped to it.  The loop has one exit: a group
    // past the last is skipped by a uniform branch rather than a break, so
    // that the compiler's wait at the loop head leaves the refills issued in
    // the previous round in flight (with breaks it waited for every load
    // there).
    template <class Fetch>
    __device__ __forceinline__ void run(MassRegs &r, int d, const double *rc, Fetch fetch) {
        r.fx = 0.0;
        r.fy = 0.0;
        r.fz = 0.0;
        const int last = ng - 1;
        for (int g = 0; g < ng; g += D) {
#pragma unroll
            for (int j = 0; j < D; ++j) {
                const bool on = g + j < ng;  // uniform over the wave
                if (ACT && on) R[j].fetch_f(rc);
                const int q = min(g + j + D - 1, last);
                R[(j + D - 1) % D].load(src, ebase, src_lane(q), q, bkd);
                if (on) eval_set<ACT, G, 1>(r, *reinterpret_cast<const Grp(*)[1]>(&R[j]), g + j, d, fetch);
            }
        }
    }
};

// The body clock of launch slot s is advanced once per launch by the body's
// first lane: T += dt (src/sim.cu:1590; RK2: T += dt / 2 per stage,
// :1577,1582) while launches are left.
__device__ __forceinline__ void advance_body_clock(const StepArgs &a, int b, int s, int64_t left,
                                                   int64_t nst, double T) {
    const bool go = left > 0;
    a.bT[s ^ 1][b] = go ? T + a.hdt : T;
    a.bn[s ^ 1][b] = go ? nst + 1 : nst;
    a.bleft[s ^ 1][b] = go ? left - 1 : 0;
}

// The new state is stored non-temporally: it is read again only by the next
// launch, and lines left dirty in the XCDs' L2 caches would have to be
// written back at the end of this one (measured: about 5% of the step).
typedef double f64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_nt(double2 *p, double x, double y) {
    __builtin_nontemporal_store((f64x2){x, y}, reinterpret_cast<f64x2 *>(p));
}

// The acceleration is stored on the body's last step of a stepping call
// (left == 1), where the caller can read it.
__device__ __forceinline__ void store_mass(const StepArgs &a, double2 *__restrict__ nxt, int64_t i,
                                           const MassRegs &r, double ax, double ay, double az, bool last) {
    store_nt(&nxt[3 * i], r.px, r.py);
    store_nt(&nxt[3 * i + 1], r.pz, r.vx);
    store_nt(&nxt[3 * i + 2], r.vy, r.vz);
    if (last || a.integ == kIntVerlet) {  // Verlet reads it back next step
        a.acc[i] = ax;
        a.acc[a.ncap + i] = ay;
        a.acc[2 * a.ncap + i] = az;
    }
}

// Gather step: one lane per mass slot, one wavefront per 64-mass chunk, four
// chunks per workgroup; the other endpoints' states are read from the
// current state buffer (global memory).  ACT: some springs are actuated.
template <bool ACT>
__global__ __launch_bounds__(kBlock) void k_step_gather(const StepArgs a) {
    const int64_t blk = step_block(a);
    const int64_t chunk = a.chunk_base + blk * kWavesPerBlock + uniform(threadIdx.x >> 6);
    if (chunk >= a.nchunks) return;
    const int lane = threadIdx.x & (kChunk - 1);
    const int b = uniform(a.chunk_body[chunk]);
    if (b < 0 || a.b_mode[b] != kModeGather) return;
    // read before this kernel's first store: scalar loads
    const int64_t ebase = a.chunk_ptr[chunk];
    const int ng = a.chunk_ng[chunk];
    const BodyKD bkd{a.b_k[b], a.b_d[b], a.b_kd[b]};
    const int s = a.slot;
    const int64_t left = a.bleft[s][b];
    const int64_t nst = a.bn[s][b];
    const double T = a.bT[s][b];
    if (lane == 0 && a.b_chunk0[b] == chunk) advance_body_clock(a, b, s, left, nst, T);
    if (left <= 0) return;

    const int64_t n6 = a.ncap;
    const int64_t i = chunk * kChunk + lane;
    const int p = (int)(nst & 1);
    const double2 *__restrict__ cur = reinterpret_cast<const double2 *>(a.st + (int64_t)p * 6 * n6);
    double2 *__restrict__ nxt = reinterpret_cast<double2 *>(a.st + (int64_t)(p ^ 1) * 6 * n6);
    // every load that does not depend on another is issued here
    const uint32_t fl = a.mflags[i];
    const double2 o0 = cur[3 * i], o1 = cur[3 * i + 1], o2 = cur[3 * i + 2];
    const int d = a.deg[i];
    const double m = a.m[i];
    const PtrSrc src{a.inc_nbr, a.inc_rest, a.inc_k, a.inc_damp};
    SpringWalk<ACT, kGroupGather, PtrSrc> w(src);
    w.prime(ebase, ng, lane, bkd);
    // the chunk's share of its body's actuation factors for the next step
    // (the body's clock then reads T + dt), while the loads are in flight
    if (ACT && left > 1) actuate_slice(a, a.chunk_act0[chunk], a.chunk_nact[chunk], T + a.hdt, a.ract[s ^ 1], lane, kChunk);
    if (fl & kPad) return;
    if (fl & kFixed) {  // src/sim.cu:1100-1101: a fixed mass keeps its state
        store_nt(&nxt[3 * i], o0.x, o0.y);
        store_nt(&nxt[3 * i + 1], o1.x, o1.y);
        store_nt(&nxt[3 * i + 2], o2.x, o2.y);
        return;
    }
    MassRegs r;
    r.px = o0.x;
    r.py = o0.y;
    r.pz = o1.x;
    r.vx = o1.y;
    r.vy = o2.x;
    r.vz = o2.y;
    // 48 bytes per mass; the arena caps a buffer below 4 GiB
    w.run(r, d, a.ract[s], [&](uint32_t o, double2 &q0, double2 &q1, double2 &q2) {
        const uint32_t ob = o * 48u;
        q0 = ldo(cur, ob);
        q1 = ldo(cur, ob + 16u);
        q2 = ldo(cur, ob + 32u);
    });
    // read after the walk, so that nothing is held through it (Euler reads nothing)
    const int stage = (int)(nst & 1);
    const IntegIn in = load_integ(a, i, stage);
    double ax, ay, az;
    mass_pass(a, r, i, fl, mass_env(a, b), m, stage, in, ax, ay, az);
    store_mass(a, nxt, i, r, ax, ay, az, left == 1);
}

// Tiled step: one workgroup per tile of up to 256 masses of one body (masses
// in Morton order).  The tile's masses (rows 0..nmass-1) and its halo (the
// other endpoints of its springs outside the tile, rows nmass..) are staged
// in LDS, 48 bytes per row, after which each lane walks its incidences as
// the gather step does, reading the other endpoints from LDS rows instead
// of global memory.  ACT: the workgroup computes the actuated rest lengths
// of the tile's actuators at the body's clock into LDS after the halo rows,
// where the walk reads those of its actuated incidences (a gather from a
// buffer written by the previous launch made each group's evaluation wait
// for the next group's loads too: vector loads retire in order).
// The tiled step after the workgroup's uniform fields are read (see
// k_step_tiles), for a body whose constant mode is KD.
template <bool ACT, int KD>
__device__ __forceinline__ void tiles_body(const StepArgs &a, double2 *__restrict__ S, int t, int nmass, int nhalo,
                                           int64_t halo_off, int nact, int64_t act_off, int64_t chunk, bool live,
                                           int64_t ebase, int ng, int b, const BodyKD &bkd, int64_t left, int64_t nst,
                                           double T) {
    const int64_t n6 = a.ncap;
    const int p = (int)(nst & 1);
    const double2 *__restrict__ cur = reinterpret_cast<const double2 *>(a.st + (int64_t)p * 6 * n6);
    double2 *__restrict__ nxt = reinterpret_cast<double2 *>(a.st + (int64_t)(p ^ 1) * 6 * n6);
    const int64_t i = chunk * kChunk + (t & (kChunk - 1));
    const bool mine = t < nmass;
    // every load that does not depend on another is issued before the
    // staging waits: the mass's state and scalars and its first incidence
    // groups
    uint32_t fl = kPad;
    double2 o0 = make_double2(0.0, 0.0), o1 = o0, o2 = o0;
    int d = 0;
    double m = 1.0;
    const ArenaSrc src{a};
    SpringWalkD<ACT, KD, ArenaSrc> w(src);
    if (live) {
        fl = a.mflags[i];
        o0 = cur[3 * i];
        o1 = cur[3 * i + 1];
        o2 = cur[3 * i + 2];
        d = a.deg[i];
        m = (bkd.kd & kKdUniformM) ? a.b_m[b] : a.m[i];  // uniform over the workgroup
        w.prime(ebase, ng, t & (kChunk - 1), bkd, d);
    }
    // the tile's actuated rest lengths, at this step's clock
    double *Sact = reinterpret_cast<double *>(S + 3 * (nmass + nhalo));
    if (ACT && nact > 0) actuate_records(a.tact + 4 * act_off, nact, T, Sact, t, kTileMasses);
    // halo rows, two per thread per round: both indices, then both states.
    // Loads past the halo are clamped to its last entry and stored to the
    // spare row a.ldim, so that no load sits behind a branch.
    const int32_t *__restrict__ hl = a.halo + halo_off;
    for (int h0 = 0; h0 < nhalo; h0 += 2 * kTileMasses) {
        const int ha = h0 + t, hb = ha + kTileMasses;
        const uint32_t ia = (uint32_t)hl[min(ha, nhalo - 1)] * 48u;
        const uint32_t ib = (uint32_t)hl[min(hb, nhalo - 1)] * 48u;
        const double2 x0 = ldo(cur, ia), x1 = ldo(cur, ia + 16u), x2 = ldo(cur, ia + 32u);
        const double2 y0 = ldo(cur, ib), y1 = ldo(cur, ib + 16u), y2 = ldo(cur, ib + 32u);
        const int qa = 3 * (ha < nhalo ? nmass + ha : a.ldim);
        const int qb = 3 * (hb < nhalo ? nmass + hb : a.ldim);
        S[qa] = x0;
        S[qa + 1] = x1;
        S[qa + 2] = x2;
        S[qb] = y0;
        S[qb + 1] = y1;
        S[qb + 2] = y2;
    }
    if (mine) {
        S[3 * t] = o0;
        S[3 * t + 1] = o1;
        S[3 * t + 2] = o2;
    }
    __syncthreads();  // the last barrier: lanes may leave from here on
    if (fl & kPad) return;
    if (fl & kFixed) {  // src/sim.cu:1100-1101
        store_nt(&nxt[3 * i], o0.x, o0.y);
        store_nt(&nxt[3 * i + 1], o1.x, o1.y);
        store_nt(&nxt[3 * i + 2], o2.x, o2.y);
        return;
    }
    MassRegs r;
    r.px = o0.x;
    r.py = o0.y;
    r.pz = o1.x;
    r.vx = o1.y;
    r.vy = o2.x;
    r.vz = o2.y;
    w.run(r, d, Sact, [&](uint32_t o, double2 &q0, double2 &q1, double2 &q2) {
        q0 = S[3 * o];
        q1 = S[3 * o + 1];
        q2 = S[3 * o + 2];
    });
    const int stage = (int)(nst & 1);
    const IntegIn in = load_integ(a, i, stage);  // after the walk (see the gather step)
    double ax, ay, az;
    mass_pass(a, r, i, fl, mass_env(a, b), m, stage, in, ax, ay, az);
    store_mass(a, nxt, i, r, ax, ay, az, left == 1);
}

template <bool ACT>
__global__ __launch_bounds__(kTileMasses) void k_step_tiles(const StepArgs a) {
    extern __shared__ double2 S[];  // 3 per LDS row
    const int64_t tile = a.tile_base + step_block(a);
    if (tile >= a.ntiles) return;
    const int t = threadIdx.x;
    // uniform fields first: read before this kernel's first store, they
    // are scalar loads
    const TileDesc *__restrict__ td = a.tiles + tile;
    const int b = td->body;
    if (b < 0) return;
    const int nmass = td->nmass, nhalo = td->nhalo, nact = ACT ? td->nact : 0;
    const int64_t halo_off = td->halo_off, act_off = ACT ? td->act_off : 0;
    const int64_t chunk = (int64_t)td->chunk0 + uniform(t >> 6);
    // a wave that holds tile masses loads for all its lanes (the chunk's
    // padding slots are allocated)
    const bool live = uniform(t >> 6) * kChunk < nmass;
    int64_t ebase = 0;
    int ng = 0;
    if (live) {
        ebase = a.chunk_ptr[chunk];
        ng = a.chunk_ng[chunk];
    }
    const BodyKD bkd{a.b_k[b], a.b_d[b], a.b_kd[b]};
    const int s = a.slot;
    const int64_t left = a.bleft[s][b];
    const int64_t nst = a.bn[s][b];
    const double T = a.bT[s][b];
    if (t == 0 && td->first) advance_body_clock(a, b, s, left, nst, T);
    if (left <= 0) return;  // uniform over the workgroup

    switch (bkd.kd & 3u) {  // uniform over the workgroup
        case 0:
            tiles_body<ACT, 0>(a, S, t, nmass, nhalo, halo_off, nact, act_off, chunk, live, ebase, ng, b, bkd,
                               left, nst, T);
            break;
        case 1:
            tiles_body<ACT, 1>(a, S, t, nmass, nhalo, halo_off, nact, act_off, chunk, live, ebase, ng, b, bkd,
                               left, nst, T);
            break;
        case 2:
            tiles_body<ACT, 2>(a, S, t, nmass, nhalo, halo_off, nact, act_off, chunk, live, ebase, ng, b, bkd,
                               left, nst, T);
            break;
        default:
            tiles_body<ACT, 3>(a, S, t, nmass, nhalo, halo_off, nact, act_off, chunk, live, ebase, ng, b, bkd,
                               left, nst, T);
            break;
    }
}

// Actuation factors for the first step of a stepping call: each wave takes
// one chunk's share, at its body's clock, for bodies with steps left that
// the gather or tiled kernels step (a resident body computes its own).
__global__ __launch_bounds__(kBlock) void k_actuate(const StepArgs a) {
    const int64_t chunk = a.chunk_base + (int64_t)xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock +
                          uniform(threadIdx.x >> 6);
    if (chunk >= a.nchunks) return;
    const int b = uniform(a.chunk_body[chunk]);
    if (b < 0 || a.b_mode[b] == kModeResident) return;
    const int s = a.slot;
    if (a.bleft[s][b] <= 0) return;
    actuate_slice(a, a.chunk_act0[chunk], a.chunk_nact[chunk], a.bT[s][b], a.ract[s], threadIdx.x & (kChunk - 1),
                  kChunk);
}

// Resident rollout: one workgroup per body of at most kResidentMaxBody
// masses, which runs the body's whole budget of steps for this call in one
// launch.  The body's state is double-buffered, in LDS when the launch has
// LDS rows for it (then it never returns to HBM until the last step), else
// in its two arena state buffers: step k reads buffer k & 1 and writes the
// other, with one barrier per step.  Each wave walks the body's chunks w, w + 4, ...;
// every mass is computed exactly as the other kernels do (the same spring
// walk, mass pass and integrator, T += hdt per launch), so results are
// identical.  A body whose incidence data fits the launch's cache region
// copies it to LDS first and walks it there; a wave with one chunk keeps
// its masses' constants in registers across steps.  The body's clock
// fields are written to launch slot final_slot, where the stepping call
// ends.
struct MassConst {
    uint32_t fl;
    int deg;
    double m;
    int64_t ebase;
    int ng;
};

template <bool ACT, class Src>
__device__ __forceinline__ void resident_steps(const StepArgs &a, const Src &src, int64_t ebase_off, int b,
                                               const BodyKD &bkd, int chunk0, int nch, int64_t g0, int64_t K,
                                               int64_t nst, double &T, int nact, int64_t act0,
                                               double2 *__restrict__ B0, double2 *__restrict__ B1) {
    const int t = threadIdx.x, wave = uniform(t >> 6), lane = t & (kChunk - 1);
    const int64_t n6 = a.ncap;
    constexpr int kWaves = kResidentBlock / kChunk;
    const MassEnv env = mass_env(a, b);
    auto consts = [&](int c) {
        MassConst mc;
        const int64_t i = g0 + (int64_t)c * kChunk + lane;
        mc.fl = a.mflags[i];
        mc.deg = a.deg[i];
        mc.m = a.m[i];
        mc.ebase = a.chunk_ptr[chunk0 + c] - ebase_off;
        mc.ng = a.chunk_ng[chunk0 + c];
        return mc;
    };
    auto one_mass = [&](int c, const MassConst &mc, int64_t k, const double2 *__restrict__ cur,
                        double2 *__restrict__ nxt) {
        const int q = c * kChunk + lane;
        const int64_t i = g0 + q;
        if (mc.fl & kPad) return;
        const double2 o0 = cur[3 * q], o1 = cur[3 * q + 1], o2 = cur[3 * q + 2];
        if (mc.fl & kFixed) {  // src/sim.cu:1100-1101
            nxt[3 * q] = o0;
            nxt[3 * q + 1] = o1;
            nxt[3 * q + 2] = o2;
            return;
        }
        MassRegs r;
        r.px = o0.x;
        r.py = o0.y;
        r.pz = o1.x;
        r.vx = o1.y;
        r.vy = o2.x;
        r.vz = o2.y;
        SpringWalk<ACT, kGroupTiled, Src> w(src);
        w.prime(mc.ebase, mc.ng, lane, bkd);
        w.run(r, mc.deg, a.ract[k & 1], [&](uint32_t o, double2 &q0, double2 &q1, double2 &q2) {
            q0 = cur[3 * o];
            q1 = cur[3 * o + 1];
            q2 = cur[3 * o + 2];
        });
        double ax, ay, az;
        const int stage = (int)((nst + k) & 1);
        const IntegIn in = load_integ(a, i, stage);
        mass_pass(a, r, i, mc.fl, env, mc.m, stage, in, ax, ay, az);
        nxt[3 * q] = make_double2(r.px, r.py);
        nxt[3 * q + 1] = make_double2(r.pz, r.vx);
        nxt[3 * q + 2] = make_double2(r.vy, r.vz);
        if (k == K - 1 || a.integ == kIntVerlet) {
            a.acc[i] = ax;
            a.acc[n6 + i] = ay;
            a.acc[2 * n6 + i] = az;
        }
    };
    if (nch <= kWaves) {  // one chunk per wave: its constants stay in registers
        const bool mine = wave < nch;
        MassConst mc{};
        if (mine) mc = consts(wave);
        for (int64_t k = 0; k < K; ++k) {
            if (mine) one_mass(wave, mc, k, (k & 1) ? B1 : B0, (k & 1) ? B0 : B1);
            T = T + a.hdt;  // src/sim.cu:1590 (RK2: :1577,1582)
            if (ACT && k + 1 < K) actuate_slice(a, act0, nact, T, a.ract[(k + 1) & 1], t, kResidentBlock);
            __syncthreads();
        }
    } else {
        for (int64_t k = 0; k < K; ++k) {
            for (int c = wave; c < nch; c += kWaves) one_mass(c, consts(c), k, (k & 1) ? B1 : B0, (k & 1) ? B0 : B1);
            T = T + a.hdt;
            if (ACT && k + 1 < K) actuate_slice(a, act0, nact, T, a.ract[(k + 1) & 1], t, kResidentBlock);
            __syncthreads();
        }
    }
}

// One force component of a row of a split-form body: its incidences' terms
// (slots j * nmass + the row, +0 past the row's degree) added in ascending
// spring index, as the walk sums.  nd (the body's largest degree rounded up
// to a multiple of 4) terms, all read before the first add when nd <= 32,
// so that one wait covers them; else eight at a time.
template <int N>
__device__ __forceinline__ double row_sum_n(const double *__restrict__ f, int nmass) {
    double v[N];
#pragma unroll
    for (int u = 0; u < N; ++u) v[u] = f[u * nmass];
    double sum = 0.0;
#pragma unroll
    for (int u = 0; u < N; ++u) sum = sum + v[u];
    return sum;
}
__device__ __forceinline__ double row_sum(const double *__restrict__ f, int nmass, int nd) {
    switch (nd) {
        case 4: return row_sum_n<4>(f, nmass);
        case 8: return row_sum_n<8>(f, nmass);
        case 12: return row_sum_n<12>(f, nmass);
        case 16: return row_sum_n<16>(f, nmass);
        case 20: return row_sum_n<20>(f, nmass);
        case 24: return row_sum_n<24>(f, nmass);
        case 28: return row_sum_n<28>(f, nmass);
        case 32: return row_sum_n<32>(f, nmass);
        default: break;
    }
    const int maxdeg = nd;
    double sum = 0.0;
    int j = 0;
    for (; j + 8 <= maxdeg; j += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = f[(j + u) * nmass];
#pragma unroll
        for (int u = 0; u < 8; ++u) sum = sum + v[u];
    }
    for (; j < maxdeg; ++j) sum = sum + f[j * nmass];
    return sum;
}

// v, or -v when `neg` (a flip of the sign bit; x + (-v) is x - v exactly)
__device__ __forceinline__ double signed_term(double v, bool neg) {
    return __hiloint2double(__double2hiint(v) ^ (neg ? (int)0x80000000 : 0), __double2loint(v));
}

// The terms G of springs x, x + B, ..., x + (U - 1) B (B = BLOCK) of a split-form
// body (those below nsp), each from its endpoint a's side and formed exactly
// as the walk forms the term at a, stored as -G in the slot of a's incidence
// and +G in b's; U springs at a time so that their latencies overlap (a lane
// past the end repeats spring x and stores it once).
// KD: where k and damping come from (as the walk's Group: bit 0 k per
// spring, bit 1 damping per spring, kKdRuntime from bkd.kd).
template <bool ACT, int U, int BLOCK = kResidentBlock, int KD = kKdRuntime>
__device__ __forceinline__ void split_terms(int x, int nsp, const uint32_t *__restrict__ csw,
                                            const double *__restrict__ cr, const double *__restrict__ kp,
                                            const double *__restrict__ dp, const BodyKD &bkd,
                                            const double2 *__restrict__ cur, const double *__restrict__ rc,
                                            double *__restrict__ fx, double *__restrict__ fy,
                                            double *__restrict__ fz) {
    double n2[U], nrm[U], dvx[U], dvy[U], dvz[U], tx[U], ty[U], tz[U], rl[U], kk[U], dmp[U];
    uint32_t w[U], iw[U];
    int xi[U];
    const bool pk = KD == kKdRuntime ? (bkd.kd & 1u) != 0 : (KD & 1) != 0;
    const bool pd = KD == kKdRuntime ? (bkd.kd & 2u) != 0 : (KD & 2) != 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int xu = x + u * BLOCK;
        xi[u] = xu < nsp ? xu : x;
        w[u] = csw[xi[u]];
        iw[u] = csw[nsp + xi[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        rl[u] = cr[xi[u]];
        if (KD == kKdRuntime) {  // read in any case (kp, dp are valid arrays), then replaced
            kk[u] = kp[xi[u]];
            dmp[u] = dp[xi[u]];
        } else {
            kk[u] = pk ? kp[xi[u]] : bkd.k;
            dmp[u] = pd ? dp[xi[u]] : bkd.d;
        }
    }
    bool fast = true;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t ea = w[u] & 0xffffu, eb = (w[u] >> 16) & (0xffffu & ~kActBit16);
        if (ACT && (w[u] & (kActBit16 << 16))) rl[u] = rc[__double_as_longlong(rl[u])];
        if (KD == kKdRuntime && !pk) kk[u] = bkd.k;
        if (KD == kKdRuntime && !pd) dmp[u] = bkd.d;
        const double2 s0 = cur[3 * ea], s1 = cur[3 * ea + 1], s2 = cur[3 * ea + 2];
        const double2 n0 = cur[3 * eb], n1 = cur[3 * eb + 1], n2v = cur[3 * eb + 2];
        tx[u] = n0.x - s0.x;
        ty[u] = n0.y - s0.y;
        tz[u] = n1.x - s1.x;
        dvx[u] = s1.y - n1.y;
        dvy[u] = s2.x - n2v.x;
        dvz[u] = s2.y - n2v.y;
        n2[u] = tx[u] * tx[u] + ty[u] * ty[u] + tz[u] * tz[u];
        fast = fast & fast_range(n2[u], tx[u], ty[u], tz[u]);
    }
    if (wave_all(fast)) {  // the same operations as unit_axis's fast branch
        double y[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            nrm[u] = fast_sqrt(n2[u]);
            y[u] = fast_recip(nrm[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            tx[u] = fast_div(tx[u], nrm[u], y[u]);
            ty[u] = fast_div(ty[u], nrm[u], y[u]);
            tz[u] = fast_div(tz[u], nrm[u], y[u]);
        }
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            double ux, uy, uz;
            unit_axis(tx[u], ty[u], tz[u], nrm[u], ux, uy, uz);
            tx[u] = ux;
            ty[u] = uy;
            tz[u] = uz;
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const double mag = kk[u] * (rl[u] - nrm[u]);
        const double dd = (dvx[u] * tx[u] + dvy[u] * ty[u] + dvz[u] * tz[u]) * dmp[u];
        const double gx = tx[u] * mag + tx[u] * dd, gy = ty[u] * mag + ty[u] * dd, gz = tz[u] * mag + tz[u] * dd;
        if (x + u * BLOCK < nsp) {
            const int ia = (int)(iw[u] & 0xffffu), ib = (int)(iw[u] >> 16);
            fx[ia] = signed_term(gx, true);
            fy[ia] = signed_term(gy, true);
            fz[ia] = signed_term(gz, true);
            fx[ib] = gx;
            fy[ib] = gy;
            fz[ib] = gz;
        }
    }
}

// Every spring's term once, thread t taking springs t, t + BLOCK, ...: one
// spring per thread and pass, four waves per SIMD hiding the chain's
// latency; a wave with a lane that has a second spring evaluates both
// interleaved (every x < BLOCK < nsp there).
template <bool ACT, int BLOCK, int KD>
__device__ __forceinline__ void terms_pass(int t, int nsp, const uint32_t *__restrict__ csw,
                                           const double *__restrict__ cr, const double *__restrict__ kp,
                                           const double *__restrict__ dp, const BodyKD &bkd,
                                           const double2 *__restrict__ cur, const double *__restrict__ rc,
                                           double *__restrict__ fx, double *__restrict__ fy, double *__restrict__ fz) {
    int x = t;
    if (wave_any(x + BLOCK < nsp)) {
        split_terms<ACT, 2, BLOCK, KD>(x, nsp, csw, cr, kp, dp, bkd, cur, rc, fx, fy, fz);
        x += 2 * BLOCK;
    }
    for (; x < nsp; x += BLOCK) split_terms<ACT, 1, BLOCK, KD>(x, nsp, csw, cr, kp, dp, bkd, cur, rc, fx, fy, fz);
}

// The split form of resident_steps (a body of at most kSplitMaxChunks
// chunks with its split-form region in LDS).  Each step, all the
// workgroup's threads compute the terms G of the body's springs, each once,
// into the incidences of both endpoints; after a barrier each mass's lane
// adds up its incidences in ascending spring index as the walk does and
// runs the mass pass.  A body of one or two chunks then spreads its spring
// work over four waves, with half the evaluations of the walk.
template <bool ACT, int BLOCK = kResidentBlock, bool LEAN = false>
__device__ __forceinline__ void resident_steps_split(const StepArgs &a, const uint32_t *__restrict__ csw,
                                                     const double *__restrict__ cr, const double *__restrict__ kp,
                                                     const double *__restrict__ dp, double *__restrict__ fx,
                                                     double *__restrict__ fy, double *__restrict__ fz,
                                                     double *__restrict__ sums, int maxdeg, int nsp,
                                                     const MassEnv &env,
                                                     const BodyKD &bkd, int nch, int nmass, int64_t g0, int64_t K,
                                                     int64_t nst, double &T, int nact, int64_t act0,
                                                     double2 *__restrict__ B0, double2 *__restrict__ B1) {
    constexpr int U = 4;
    // a workgroup with a lane for each component of each row sums the
    // components on lanes of their own, after which the mass passes read
    // them (one more barrier per step, a third of the dependent adds per
    // lane)
    constexpr bool SPREAD = BLOCK >= 3 * kSplitMaxChunks * kChunk;
    const int t = threadIdx.x, wave = uniform(t >> 6), lane = t & (kChunk - 1);
    const int64_t n6 = a.ncap;
    // the rows go to the first four waves (one per SIMD), 16, 32 or 64 lanes
    // each, so that every SIMD runs a share of the sums and mass passes
    const int L = nmass > 2 * kChunk ? kChunk : nmass > kChunk ? kChunk / 2 : kChunk / 4;
    const int q = wave * L + lane;
    const bool mine = wave < kSplitMaxChunks && lane < L && q < nmass;
    const int64_t i = g0 + q;
    uint32_t fl = kPad;
    double m = 1.0;
    if (mine) {
        fl = a.mflags[i];
        m = a.m[i];
    }
    for (int64_t k = 0; k < K; ++k) {
        const double2 *__restrict__ cur = (k & 1) ? B1 : B0;
        double2 *__restrict__ nxt = (k & 1) ? B0 : B1;
        const double *__restrict__ rc = a.ract[k & 1];
        if (BLOCK == kResidentBlock) {
            int x = t;
            for (; x + (U - 1) * BLOCK < nsp; x += U * BLOCK)
                split_terms<ACT, U, BLOCK>(x, nsp, csw, cr, kp, dp, bkd, cur, rc, fx, fy, fz);
            for (; x < nsp; x += BLOCK) split_terms<ACT, 1, BLOCK>(x, nsp, csw, cr, kp, dp, bkd, cur, rc, fx, fy, fz);
        } else {
            switch (bkd.kd & 3u) {  // uniform: the body's constant mode fixed at compile time
                case 0: terms_pass<ACT, BLOCK, 0>(t, nsp, csw, cr, kp, dp, bkd, cur, rc, fx, fy, fz); break;
                case 1: terms_pass<ACT, BLOCK, 1>(t, nsp, csw, cr, kp, dp, bkd, cur, rc, fx, fy, fz); break;
                case 2: terms_pass<ACT, BLOCK, 2>(t, nsp, csw, cr, kp, dp, bkd, cur, rc, fx, fy, fz); break;
                default: terms_pass<ACT, BLOCK, 3>(t, nsp, csw, cr, kp, dp, bkd, cur, rc, fx, fy, fz); break;
            }
        }
        __syncthreads();
        if (SPREAD) {  // each component's sums on a lane of its own, into `sums`
            const int P = (nmass + kChunk - 1) & ~(kChunk - 1);
            const int c = uniform(t / P), qq = t - c * P;
            if (c < 3 && qq < nmass) sums[c * P + qq] = row_sum((c == 0 ? fx : c == 1 ? fy : fz) + qq, nmass, maxdeg);
            __syncthreads();
        }
        if (mine && !(fl & kPad)) {
            const double2 o0 = cur[3 * q], o1 = cur[3 * q + 1], o2 = cur[3 * q + 2];
            if (fl & kFixed) {  // src/sim.cu:1100-1101
                nxt[3 * q] = o0;
                nxt[3 * q + 1] = o1;
                nxt[3 * q + 2] = o2;
            } else {
                MassRegs r;
                r.px = o0.x;
                r.py = o0.y;
                r.pz = o1.x;
                r.vx = o1.y;
                r.vy = o2.x;
                r.vz = o2.y;
                if (SPREAD) {
                    const int P = (nmass + kChunk - 1) & ~(kChunk - 1);
                    r.fx = sums[q];
                    r.fy = sums[P + q];
                    r.fz = sums[2 * P + q];
                } else {
                    r.fx = row_sum(fx + q, nmass, maxdeg);
                    r.fy = row_sum(fy + q, nmass, maxdeg);
                    r.fz = row_sum(fz + q, nmass, maxdeg);
                }
                double ax, ay, az;
                if (LEAN) {
                    mass_pass_lean(a, r, env, m, ax, ay, az);
                } else {
                    const int stage = (int)((nst + k) & 1);
                    const IntegIn in = load_integ(a, i, stage);
                    mass_pass(a, r, i, fl, env, m, stage, in, ax, ay, az);
                }
                nxt[3 * q] = make_double2(r.px, r.py);
                nxt[3 * q + 1] = make_double2(r.pz, r.vx);
                nxt[3 * q + 2] = make_double2(r.vy, r.vz);
                if (k == K - 1 || (!LEAN && a.integ == kIntVerlet)) {
                    a.acc[i] = ax;
                    a.acc[n6 + i] = ay;
                    a.acc[2 * n6 + i] = az;
                }
            }
        }
        T = T + a.hdt;  // src/sim.cu:1590 (RK2: :1577,1582)
        if (ACT && k + 1 < K) actuate_slice(a, act0, nact, T, a.ract[(k + 1) & 1], t, BLOCK);
        __syncthreads();
    }
}

// Fill a body's split-form region: the body's largest degree rounded up to a
// multiple of 4 (returned), copies of the springs' endpoint and slot words, rest lengths, k
// and damping from the body's incidence span, and +0 in the term slots past
// each row's degree up to the returned value (terms at fx, fy, fz: that
// many x nmass slots each).
// The first kSplitMaxChunks waves hold one row per thread; a larger
// workgroup's other threads only help with the copies.
static_assert(kResidentBlock == kSplitMaxChunks * kChunk, "one thread per row of a split-form body");
template <int BLOCK = kResidentBlock>
__device__ __forceinline__ int split_fill(const StepArgs &a, char *__restrict__ base, int nch, int nmass,
                                          int64_t g0, int64_t e0, int nsp, int nd) {
    const int t = threadIdx.x, wave = t >> 6, lane = t & (kChunk - 1);
    int *wmax = reinterpret_cast<int *>(base);
    const bool row = wave < kSplitMaxChunks;
    const int dg = (row && t < nch * kChunk) ? a.deg[g0 + t] : 0;
    int mx = dg;  // maximum over the wave
#pragma unroll
    for (int off = 1; off < kChunk; off <<= 1) mx = max(mx, __shfl_xor(mx, off));
    if (row && lane == 0) wmax[wave] = mx;
    uint32_t *cw = reinterpret_cast<uint32_t *>(base + kSplitHead);  // endpoint words, then slot words
    double *cr = reinterpret_cast<double *>(base + ((kSplitHead + 8 * (int64_t)nsp + 15) & ~(int64_t)15));
    for (int j = t; j < 2 * nsp; j += BLOCK) cw[j] = a.inc_nbr[e0 + j];
    for (int j = t; j < nd; j += BLOCK) cr[j] = a.inc_rest[e0 + j];
    __syncthreads();
    const int maxdeg = (max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3])) + 3) & ~3;
    const int ns = maxdeg * nmass;
    double *fx = cr + nd;
    if (t < nmass) {
        for (int j = dg; j < maxdeg; ++j) {
            fx[j * nmass + t] = 0.0;
            fx[ns + j * nmass + t] = 0.0;
            fx[2 * ns + j * nmass + t] = 0.0;
        }
    }
    return maxdeg;
}

// A split-form body held in LDS, stepped by a workgroup of BLOCK threads:
// its state copied in, the split-form region filled, K steps, the state
// copied out to the buffer of its new step parity.
template <bool ACT, int BLOCK, bool LEAN>
__device__ __forceinline__ void split_in_lds(const StepArgs &a, double2 *__restrict__ S, double *__restrict__ sums,
                                             const ResidentDesc *__restrict__ rd, const MassEnv &env,
                                             const BodyKD &bkd, int nch, int nmass, int64_t inc0, int ninc, int64_t g0,
                                             int64_t K, int64_t nst, double &T, int nact, int64_t act0) {
    const int t = threadIdx.x;
    const int64_t n6 = a.ncap;
    double2 *__restrict__ B0 = S, *__restrict__ B1 = S + 3 * a.res_rows;
    {
        const double2 *__restrict__ src = reinterpret_cast<const double2 *>(a.st + (int64_t)(nst & 1) * 6 * n6);
        for (int q = t; q < 3 * nmass; q += BLOCK) B0[q] = src[3 * g0 + q];
    }
    char *base = reinterpret_cast<char *>(S + 6 * a.res_rows);
    const int nsp = rd->nsp;
    const int nd = nsp * (1 + ((bkd.kd & 1u) ? 1 : 0) + ((bkd.kd & 2u) ? 1 : 0));
    const int maxdeg = split_fill<BLOCK>(a, base, nch, nmass, g0, inc0 + ninc, nsp, nd);
    const int nslots = maxdeg * nmass;
    const uint32_t *csw = reinterpret_cast<const uint32_t *>(base + kSplitHead);
    const double *cr = reinterpret_cast<const double *>(base + ((kSplitHead + 8 * (int64_t)nsp + 15) & ~(int64_t)15));
    // per-spring k and damping where the body has them; else a valid array
    // read in their place, whose values are then replaced
    const double *ck = cr + nsp;
    const double *cd = ck + ((bkd.kd & 1u) ? nsp : 0);
    const double *kp = (bkd.kd & 1u) ? ck : cr;
    const double *dp = (bkd.kd & 2u) ? cd : cr;
    double *fx = const_cast<double *>(cr) + nd;
    double *fy = fx + nslots, *fz = fy + nslots;
    resident_steps_split<ACT, BLOCK, LEAN>(a, csw, cr, kp, dp, fx, fy, fz, sums, maxdeg, nsp, env, bkd, nch, nmass,
                                           g0, K, nst, T, nact, act0, B0, B1);
    const double2 *__restrict__ fin = (K & 1) ? B1 : B0;
    double2 *__restrict__ dst = reinterpret_cast<double2 *>(a.st + (int64_t)((nst + K) & 1) * 6 * n6);
    for (int q = t; q < 3 * nmass; q += BLOCK) store_nt(&dst[3 * g0 + q], fin[q].x, fin[q].y);
}

// The steps of a resident body held in the launch's LDS: its state is
// copied in, stepped in the two LDS buffers (in the split form, with a copy
// of its incidence data, or reading that from the arena) and copied out to
// the state buffer of its new step parity.
template <bool ACT, bool SPLIT>
__device__ __forceinline__ void resident_in_lds(const StepArgs &a, double2 *__restrict__ S,
                                                const ResidentDesc *__restrict__ rd, int b, const BodyKD &bkd,
                                                int chunk0, int nch, int nmass, int64_t inc0, int ninc, int cache,
                                                int64_t g0, int64_t K, int64_t nst, double &T, int nact,
                                                int64_t act0) {
    const int t = threadIdx.x;
    const int64_t n6 = a.ncap;
    double2 *__restrict__ B0 = S, *__restrict__ B1 = S + 3 * a.res_rows;
    {
        const double2 *__restrict__ src = reinterpret_cast<const double2 *>(a.st + (int64_t)(nst & 1) * 6 * n6);
        for (int q = t; q < nmass; q += kResidentBlock) {
            const int64_t i = g0 + q;
            B0[3 * q] = src[3 * i];
            B0[3 * q + 1] = src[3 * i + 1];
            B0[3 * q + 2] = src[3 * i + 2];
        }
    }
    // uniform over the block: the split form, else a copy of the incidence
    // data, when the launch's region holds it; else the arena
    const bool split = SPLIT && rd->scache > 0 && rd->scache <= a.res_cache;
    const bool cached = !split && cache > 0 && cache <= a.res_cache;
    char *base = reinterpret_cast<char *>(S + 6 * a.res_rows);
    if (split) {
        const int nsp = rd->nsp;
        const int nd = nsp * (1 + ((bkd.kd & 1u) ? 1 : 0) + ((bkd.kd & 2u) ? 1 : 0));
        const int maxdeg = split_fill(a, base, nch, nmass, g0, inc0 + ninc, nsp, nd);
        const int nslots = maxdeg * nmass;
        const uint32_t *csw = reinterpret_cast<const uint32_t *>(base + kSplitHead);
        const double *cr = reinterpret_cast<const double *>(base + ((kSplitHead + 8 * (int64_t)nsp + 15) & ~(int64_t)15));
        // per-spring k and damping where the body has them; else a valid
        // array read in their place, whose values are then replaced
        const double *ck = cr + nsp;
        const double *cd = ck + ((bkd.kd & 1u) ? nsp : 0);
        const double *kp = (bkd.kd & 1u) ? ck : cr;
        const double *dp = (bkd.kd & 2u) ? cd : cr;
        double *fx = const_cast<double *>(cr) + nd;
        double *fy = fx + nslots, *fz = fy + nslots;
        resident_steps_split<ACT>(a, csw, cr, kp, dp, fx, fy, fz, nullptr, maxdeg, nsp, mass_env(a, b), bkd, nch,
                                  nmass, g0, K, nst, T, nact, act0, B0, B1);
    } else if (cached) {
        // names (4 B per entry), then from a 16-byte boundary the rest
        // lengths and, where the body has them, k and damping (8 B each)
        uint32_t *cn = reinterpret_cast<uint32_t *>(base);
        double *cr = reinterpret_cast<double *>(base + ((4 * (int64_t)ninc + 15) & ~(int64_t)15));
        double *ck = cr + ninc;
        double *cd = ck + ((bkd.kd & 1u) ? ninc : 0);
        for (int e = t; e < ninc; e += kResidentBlock) {
            cn[e] = a.inc_nbr[inc0 + e];
            cr[e] = a.inc_rest[inc0 + e];
            if (bkd.kd & 1u) ck[e] = a.inc_k[inc0 + e];
            if (bkd.kd & 2u) cd[e] = a.inc_damp[inc0 + e];
        }
        __syncthreads();
        const PtrSrc src{cn, cr, ck, cd};
        resident_steps<ACT>(a, src, inc0, b, bkd, chunk0, nch, g0, K, nst, T, nact, act0, B0, B1);
    } else {
        __syncthreads();
        const ArenaSrc src{a};
        resident_steps<ACT>(a, src, 0, b, bkd, chunk0, nch, g0, K, nst, T, nact, act0, B0, B1);
    }
    // the final state to the state buffer of the body's new step parity
    {
        const double2 *__restrict__ fin = (K & 1) ? B1 : B0;
        double2 *__restrict__ dst = reinterpret_cast<double2 *>(a.st + (int64_t)((nst + K) & 1) * 6 * n6);
        for (int q = t; q < nmass; q += kResidentBlock) {
            const int64_t i = g0 + q;
            store_nt(&dst[3 * i], fin[3 * q].x, fin[3 * q].y);
            store_nt(&dst[3 * i + 1], fin[3 * q + 1].x, fin[3 * q + 1].y);
            store_nt(&dst[3 * i + 2], fin[3 * q + 2].x, fin[3 * q + 2].y);
        }
    }
}

// SPLIT: the launch uses the split form where a body has it.  The variant
// without it is launched otherwise: the split form's interleaved term
// evaluation needs more registers, which would cost the walk occupancy.
template <bool ACT, bool SPLIT>
__global__ __launch_bounds__(kResidentBlock) void k_rollout_resident(const StepArgs a, int final_slot) {
    extern __shared__ double2 S[];  // two buffers of 3 x res_rows pairs, then the incidence copy
    const int64_t rid = a.res_base + blockIdx.x;
    if (rid >= a.nres) return;
    const ResidentDesc *__restrict__ rd = a.rdesc + rid;
    const int b = rd->body;
    if (b < 0) return;
    const int nmass = rd->nmass, chunk0 = rd->chunk0, nact = rd->nact;
    const int64_t act0 = rd->act0, inc0 = rd->inc0;
    const int ninc = rd->ninc, cache = rd->cache;
    const int nch = (nmass + kChunk - 1) / kChunk;
    const BodyKD bkd{a.b_k[b], a.b_d[b], a.b_kd[b]};
    const int s0 = a.slot;
    const int64_t K = a.bleft[s0][b];
    const int64_t nst = a.bn[s0][b];
    double T = a.bT[s0][b];
    const int t = threadIdx.x;
    if (K <= 0) {  // not stepped in this call: only its clock fields move to final_slot
        if (t == 0) {
            a.bT[final_slot][b] = T;
            a.bn[final_slot][b] = nst;
            a.bleft[final_slot][b] = 0;
        }
        return;
    }
    const int64_t g0 = (int64_t)chunk0 * kChunk;
    const int64_t n6 = a.ncap;
    // actuation factors: step k reads ract[k & 1] (this block's actuators
    // only), written before the barrier that precedes the step
    if (ACT) actuate_slice(a, act0, nact, T, a.ract[0], t, kResidentBlock);
    if (nmass > a.res_rows) {
        // a body larger than the launch's LDS rows (every body, when the
        // launch has none) steps in the arena: the two state buffers of its
        // slots alternate as the LDS buffers do, and the barrier that ends
        // a step orders its stores before the next step's loads (the
        // workgroup shares one CU and its L1)
        double2 *__restrict__ G0 = reinterpret_cast<double2 *>(a.st + (int64_t)(nst & 1) * 6 * n6) + 3 * g0;
        double2 *__restrict__ G1 = reinterpret_cast<double2 *>(a.st + (int64_t)((nst + 1) & 1) * 6 * n6) + 3 * g0;
        __syncthreads();
        const ArenaSrc src{a};
        resident_steps<ACT>(a, src, 0, b, bkd, chunk0, nch, g0, K, nst, T, nact, act0, G0, G1);
    } else {
        resident_in_lds<ACT, SPLIT>(a, S, rd, b, bkd, chunk0, nch, nmass, inc0, ninc, cache, g0, K, nst, T, nact,
                                    act0);
    }
    if (t == 0) {
        a.bT[final_slot][b] = T;
        a.bn[final_slot][b] = nst + K;
        a.bleft[final_slot][b] = 0;
    }
}

// The resident launch when every body in it is in LDS in the split form:
// workgroups of kSplitBlock threads, so that a body's spring terms are spread
// over four times as many threads (the README scene's 1,036 springs: one per
// thread), with the same arithmetic and order as k_rollout_resident.
// LEAN: every mass of the launch's bodies takes mass_pass_lean's case (the
// launch checks this on the host).
template <bool ACT, bool LEAN>
__global__ __launch_bounds__(kSplitBlock) void k_rollout_split(const StepArgs a, int final_slot) {
    extern __shared__ double2 S[];  // two buffers of 3 x res_rows pairs, then the split-form region
    __shared__ double planes[8 * kSplitPlanes];  // the global planes (the launch has at most kSplitPlanes)
    __shared__ double sums[3 * kSplitMaxChunks * kChunk];  // each row's force components
    const int64_t rid = a.res_base + blockIdx.x;
    if (rid >= a.nres) return;
    const ResidentDesc *__restrict__ rd = a.rdesc + rid;
    const int b = rd->body;
    if (b < 0) return;
    const int nmass = rd->nmass, chunk0 = rd->chunk0, nact = rd->nact;
    const int64_t act0 = rd->act0, inc0 = rd->inc0;
    const int ninc = rd->ninc;
    const int nch = (nmass + kChunk - 1) / kChunk;
    const BodyKD bkd{a.b_k[b], a.b_d[b], a.b_kd[b]};
    const int s0 = a.slot;
    const int64_t K = a.bleft[s0][b];
    const int64_t nst = a.bn[s0][b];
    double T = a.bT[s0][b];
    const int t = threadIdx.x;
    if (K > 0) {
        if (ACT) actuate_slice(a, act0, nact, T, a.ract[0], t, kSplitBlock);
        for (int j = t; j < 8 * a.nplanes; j += kSplitBlock) planes[j] = a.planes[j];  // visible after split_fill's barriers
        const MassEnv env{planes, a.b_lc[b], a.b_lcn[b]};
        split_in_lds<ACT, kSplitBlock, LEAN>(a, S, sums, rd, env, bkd, nch, nmass, inc0, ninc, (int64_t)chunk0 * kChunk,
                                             K, nst, T, nact, act0);
    }
    if (t == 0) {
        a.bT[final_slot][b] = T;
        a.bn[final_slot][b] = nst + (K > 0 ? K : 0);
        a.bleft[final_slot][b] = 0;
    }
}

// Per-mass local time: T += dt once per step the body took since the last
// sync (src/sim.cu:1165), for every non-fixed mass.
__global__ void k_advance_T(double *Tm, const uint32_t *mflags, const int32_t *chunk_body,
                            const int64_t *bn, const int64_t *bnsync, int64_t chunk_begin,
                            int64_t chunk_end, double dt) {
    const int64_t i = chunk_begin * kChunk + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t chunk = i / kChunk;
    if (chunk >= chunk_end) return;
    const int b = chunk_body[chunk];
    if (b < 0) return;
    if (mflags[i] & (kFixed | kPad)) return;
    const int64_t cnt = bn[b] - bnsync[b];
    double t = Tm[i];
    for (int64_t q = 0; q < cnt; ++q) t += dt;
    Tm[i] = t;
}

__global__ void k_sync_bodies(const int64_t *bn, int64_t *bnsync, int64_t nb) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) bnsync[b] = bn[b];
}

__global__ void k_set_left(int64_t *bleft, const int32_t *b_chunk0, int64_t nb, int64_t n) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) bleft[b] = b_chunk0[b] >= 0 ? n : 0;
}

__global__ void k_set_i64(int64_t *p, int64_t v) { *p = v; }

// ---- launchers ----

hipError_t launch_step(const StepArgs &a, bool act, hipStream_t s) {
    const int64_t blocks = (a.nchunks - a.chunk_base + kWavesPerBlock - 1) / kWavesPerBlock;
    if (blocks <= 0) return hipSuccess;
    const dim3 g((unsigned)blocks), blk(kBlock);
    if (act) hipLaunchKernelGGL((k_step_gather<true>), g, blk, 0, s, a);
    else hipLaunchKernelGGL((k_step_gather<false>), g, blk, 0, s, a);
    return hipGetLastError();
}

// the tile rows plus one spare row (k_step_tiles stores unused halo loads there)
size_t tiles_lds_bytes(int ldim) { return (size_t)48 * (ldim + 1); }

hipError_t launch_step_tiles(const StepArgs &a, bool act, hipStream_t s) {
    if (a.ntiles <= a.tile_base) return hipSuccess;
    const size_t lds = tiles_lds_bytes(a.ldim);
    const dim3 g((unsigned)(a.ntiles - a.tile_base)), blk(kTileMasses);
    if (act) hipLaunchKernelGGL((k_step_tiles<true>), g, blk, lds, s, a);
    else hipLaunchKernelGGL((k_step_tiles<false>), g, blk, lds, s, a);
    return hipGetLastError();
}

hipError_t launch_actuate(const StepArgs &a, hipStream_t s) {
    const int64_t n = a.nchunks - a.chunk_base;
    if (n <= 0) return hipSuccess;
    const dim3 g((unsigned)((n + kWavesPerBlock - 1) / kWavesPerBlock)), blk(kBlock);
    hipLaunchKernelGGL(k_actuate, g, blk, 0, s, a);
    return hipGetLastError();
}

// A body joins the launches (on = 1: its clock fields, shared constants and
// mode, then its chunks, tiles and resident descriptor name it) or leaves
// them (on = 0: they name no body).  Its other arrays were uploaded before.
__global__ __launch_bounds__(kBlock) void k_set_body(const BodyTable t, const BodyInit bi) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t who = bi.on ? bi.bid : -1;
    if (q < bi.nchunks) t.chunk_body[bi.chunk0 + q] = who;
    if (q < bi.ntiles) t.tiles[bi.tile0 + q].body = who;
    if (q != 0) return;
    if (bi.res >= 0) t.rdesc[bi.res].body = who;
    const int b = bi.bid;
    if (bi.on) {
        for (int s = 0; s < 2; ++s) {
            t.bT[s][b] = bi.t0;
            t.bn[s][b] = 0;
            t.bleft[s][b] = 0;
        }
        t.bnsync[b] = 0;
        t.b_lc[b] = bi.lcb;
        t.b_lcn[b] = bi.lcn;
        t.b_k[b] = bi.k;
        t.b_d[b] = bi.d;
        t.b_m[b] = bi.m;
        t.b_kd[b] = bi.kd;
    }
    t.b_chunk0[b] = bi.on ? (int32_t)bi.chunk0 : -1;
    t.b_mode[b] = bi.on ? bi.mode : kModeGather;
}

hipError_t launch_set_body(const BodyTable &t, const BodyInit &bi, hipStream_t s) {
    const int64_t n = std::max<int64_t>(1, std::max(bi.nchunks, bi.ntiles));
    hipLaunchKernelGGL(k_set_body, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, t, bi);
    return hipGetLastError();
}

size_t resident_lds_bytes(int rows) { return (size_t)2 * 48 * rows; }

hipError_t launch_rollout_resident(const StepArgs &a, bool act, int final_slot, hipStream_t s) {
    if (a.nres <= a.res_base) return hipSuccess;
    const size_t lds = resident_lds_bytes(a.res_rows) + (size_t)a.res_cache;
    const dim3 g((unsigned)(a.nres - a.res_base)), blk(kResidentBlock);
    if (a.res_split >= 2) {
        const dim3 sb(kSplitBlock);
        if (a.res_split == 3) {
            if (act) hipLaunchKernelGGL((k_rollout_split<true, true>), g, sb, lds, s, a, final_slot);
            else hipLaunchKernelGGL((k_rollout_split<false, true>), g, sb, lds, s, a, final_slot);
        } else {
            if (act) hipLaunchKernelGGL((k_rollout_split<true, false>), g, sb, lds, s, a, final_slot);
            else hipLaunchKernelGGL((k_rollout_split<false, false>), g, sb, lds, s, a, final_slot);
        }
    } else if (a.res_split) {
        if (act) hipLaunchKernelGGL((k_rollout_resident<true, true>), g, blk, lds, s, a, final_slot);
        else hipLaunchKernelGGL((k_rollout_resident<false, true>), g, blk, lds, s, a, final_slot);
    } else {
        if (act) hipLaunchKernelGGL((k_rollout_resident<true, false>), g, blk, lds, s, a, final_slot);
        else hipLaunchKernelGGL((k_rollout_resident<false, false>), g, blk, lds, s, a, final_slot);
    }
    return hipGetLastError();
}

const char *resident_kernel_name(bool act) {
    return act ? "k_rollout_resident<actuated>" : "k_rollout_resident<passive>";
}

const char *split_kernel_name(bool act) { return act ? "k_rollout_split<actuated>" : "k_rollout_split<passive>"; }

hipError_t set_tiles_lds_limit(int *resident_max, int *split_max) {
    int dev = 0, lim = 0;
    hipError_t er = hipGetDevice(&dev);
    if (er != hipSuccess) return er;
    er = hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (er != hipSuccess) return er;
    const int rb = std::min(lim, (int)resident_lds_bytes(kResidentMax) + kResidentCacheMax);
    *resident_max = rb;
    const void *res[4] = {(const void *)k_rollout_resident<false, false>, (const void *)k_rollout_resident<true, false>,
                          (const void *)k_rollout_resident<false, true>, (const void *)k_rollout_resident<true, true>};
    for (const void *f : res) {
        er = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, rb);
        if (er != hipSuccess) return er;
    }
    // the split launch's own LDS (its planes) comes out of the same budget
    *split_max = std::min(rb, lim - kSplitStaticLds);
    const void *spl[4] = {(const void *)k_rollout_split<false, false>, (const void *)k_rollout_split<true, false>,
                          (const void *)k_rollout_split<false, true>, (const void *)k_rollout_split<true, true>};
    for (const void *f : spl) {
        er = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, *split_max);
        if (er != hipSuccess) return er;
    }
    const int bytes = (int)tiles_lds_bytes(kMaxTileRows);
    const hipError_t e =
        hipFuncSetAttribute((const void *)k_step_tiles<false>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void *)k_step_tiles<true>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

const char *step_tiles_kernel_name(bool act) { return act ? "k_step_tiles<actuated>" : "k_step_tiles<passive>"; }

const char *step_kernel_name(bool act) { return act ? "k_step_gather<actuated>" : "k_step_gather<passive>"; }

hipError_t launch_advance_T(double *Tm, const uint32_t *mflags, const int32_t *chunk_body,
                            const int64_t *bn, int64_t *bnsync, int64_t chunk_begin, int64_t chunk_end,
                            int64_t nb, double dt, hipStream_t s) {
    const int64_t n = (chunk_end - chunk_begin) * kChunk;
    if (n > 0) {
        hipLaunchKernelGGL(k_advance_T, dim3((n + 255) / 256), dim3(256), 0, s, Tm, mflags,
                           chunk_body, bn, (const int64_t *)bnsync, chunk_begin, chunk_end, dt);
    }
    if (nb > 0) hipLaunchKernelGGL(k_sync_bodies, dim3((nb + 255) / 256), dim3(256), 0, s, bn, bnsync, nb);
    return hipGetLastError();
}

hipError_t launch_set_i64(int64_t *p, int64_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(1), 0, s, p, v);
    return hipGetLastError();
}

// Streaming kernels for the HBM calibration: each lane moves two 16-byte
// words per round (non-temporal loads and stores), a wave's words
// contiguous, the whole grid sweeping one contiguous window through memory.
// Measured on MI355X over 1 GiB, the best of grids of 256 to 16384
// workgroups of 256 threads: 512 for the copy (6.2 TB/s read + written),
// 1024 for the read (6.9 TB/s); grids of 8192 or more fall to 5.0-5.5.
constexpr int kCopyBlocks = 512, kReadBlocks = 1024, kStreamThreads = 256, kStreamUnroll = 2;
__device__ __forceinline__ double2 load_nt(const double2 *p) {
    const f64x2 v = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(p));
    return make_double2(v.x, v.y);
}

__global__ __launch_bounds__(kStreamThreads) void k_stream_copy(const double2 *__restrict__ src,
                                                                double2 *__restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kStreamThreads;
    for (int64_t i = (int64_t)blockIdx.x * kStreamThreads + threadIdx.x; i < n; i += kStreamUnroll * stride) {
        double2 v[kStreamUnroll];
#pragma unroll
        for (int u = 0; u < kStreamUnroll; ++u) {
            const int64_t j = i + u * stride;
            v[u] = j < n ? load_nt(&src[j]) : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < kStreamUnroll; ++u) {
            const int64_t j = i + u * stride;
            if (j < n) store_nt(&dst[j], v[u].x, v[u].y);
        }
    }
}

__global__ __launch_bounds__(kStreamThreads) void k_stream_read(const double2 *__restrict__ src,
                                                                double *__restrict__ out, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * kStreamThreads;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kStreamThreads + threadIdx.x; i < n; i += kStreamUnroll * stride) {
        double2 v[kStreamUnroll];
#pragma unroll
        for (int u = 0; u < kStreamUnroll; ++u) {
            const int64_t j = i + u * stride;
            v[u] = j < n ? load_nt(&src[j]) : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < kStreamUnroll; ++u) acc += v[u].x + v[u].y;
    }
    out[(int64_t)blockIdx.x * kStreamThreads + threadIdx.x] = acc;
}

hipError_t launch_stream_copy(const void *src, void *dst, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_stream_copy, dim3(kCopyBlocks), dim3(kStreamThreads), 0, s,
                       static_cast<const double2 *>(src), static_cast<double2 *>(dst), n);
    return hipGetLastError();
}

hipError_t launch_stream_read(const void *src, double *out, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_stream_read, dim3(kReadBlocks), dim3(kStreamThreads), 0, s,
                       static_cast<const double2 *>(src), out, n);
    return hipGetLastError();
}

hipError_t launch_set_left(int64_t *bleft, const int32_t *b_chunk0, int64_t nb, int64_t n,
                           hipStream_t s) {
    if (nb > 0) hipLaunchKernelGGL(k_set_left, dim3((nb + 255) / 256), dim3(256), 0, s, bleft, b_chunk0, nb, n);
    return hipGetLastError();
}


}  // namespace titan_mi355x
