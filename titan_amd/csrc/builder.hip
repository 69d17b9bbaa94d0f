// Device-side layout builder (see builder.h).  Each step restates one rule
// of the host builders in tiles.cpp with device sorts, so that the result
// is the same bytes:
//   * morton_order: codes from the positions' bounding box, sorted stably
//     (ties keep the caller's order);
//   * build_tiles: tiles of four whole chunks (256 masses) whenever every
//     such tile's masses, halo and actuator rows fit the row budget (rows
//     only grow with a tile's chunks, so the host's greedy then takes four
//     chunks too); a halo is the ascending set of outside endpoints of the
//     tile's springs, and a tile's actuators the ascending set of actuated
//     springs with an endpoint in it, each with a copy of its record;
//   * degree_sort_tiles: masses of a tile by descending degree, stably;
//   * build_sell: each mass's incidences in ascending spring index (the
//     right end of a spring first, so a self-spring is listed twice), at
//     entry coff[c] + ((j / 2) * 64 + lane) * 2 + j % 2, padding naming the
//     mass itself (row 0 for slots past the body) with rest 1 and zero
//     constants, and an actuated incidence marked by bit 15 of its row with
//     its spring's position in its tile's actuator list in the rest slot
//     (mark_actuated).
#include "builder.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

namespace titan_mi355x {

namespace {

constexpr int kBT = 256;          // threads per block of the builder's kernels
constexpr int kBoxBlocks = 1024;  // partial bounding boxes
constexpr uint32_t kNone = 0xffffffffu;

inline unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kBT - 1) / kBT); }

__device__ __forceinline__ uint64_t spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

// per-block min and max of the finite coordinates: part[6 * block + c] (lo),
// part[6 * block + 3 + c] (hi)
__global__ __launch_bounds__(kBT) void k_box_part(const double *__restrict__ pos, int64_t n, double *__restrict__ part) {
    __shared__ double sh[6][kBT];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * kBT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBT)
        for (int c = 0; c < 3; ++c) {
            const double x = pos[3 * i + c];
            if (isfinite(x)) {
                lo[c] = fmin(lo[c], x);
                hi[c] = fmax(hi[c], x);
            }
        }
    for (int c = 0; c < 3; ++c) {
        sh[c][threadIdx.x] = lo[c];
        sh[3 + c][threadIdx.x] = hi[c];
    }
    __syncthreads();
    for (int w = kBT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 3; ++c) {
                sh[c][threadIdx.x] = fmin(sh[c][threadIdx.x], sh[c][threadIdx.x + w]);
                sh[3 + c][threadIdx.x] = fmax(sh[3 + c][threadIdx.x], sh[3 + c][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(kBT) void k_box_final(const double *__restrict__ part, int nparts, double *__restrict__ box) {
    if (threadIdx.x != 0) return;
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = 0; b < nparts; ++b)
        for (int c = 0; c < 3; ++c) {
            v[c] = fmin(v[c], part[6 * b + c]);
            v[3 + c] = fmax(v[3 + c], part[6 * b + 3 + c]);
        }
    for (int c = 0; c < 6; ++c) box[c] = v[c];
}

struct Box {
    double lo[3], scale[3];
};

__global__ __launch_bounds__(kBT) void k_codes(const double *__restrict__ pos, int64_t n, Box bx,
                                               uint64_t *__restrict__ code, uint32_t *__restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (i >= n) return;
    uint64_t cd = 0;
    for (int c = 0; c < 3; ++c) {
        double q = (pos[3 * i + c] - bx.lo[c]) * bx.scale[c];
        if (!(q >= 0)) q = 0;
        if (q > (double)((1 << 21) - 1)) q = (double)((1 << 21) - 1);
        cd |= spread3((uint64_t)q) << c;
    }
    code[i] = cd;
    idx[i] = (uint32_t)i;
}

__global__ __launch_bounds__(kBT) void k_invert(const uint32_t *__restrict__ perm, int64_t n, int32_t *__restrict__ inv) {
    const int64_t q = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (q < n) inv[perm[q]] = (int32_t)q;
}

// degrees, in the labels `inv` gives, and the largest
__global__ __launch_bounds__(kBT) void k_degrees(const int32_t *__restrict__ L, const int32_t *__restrict__ R, int64_t S,
                                                 const int32_t *__restrict__ inv, int32_t *__restrict__ deg) {
    const int64_t s = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (s >= S) return;
    const int32_t a = L[s], b = R[s];
    if (a < 0 || b < 0) return;
    atomicAdd(&deg[inv[a]], 1);
    atomicAdd(&deg[inv[b]], 1);
}

__global__ __launch_bounds__(kBT) void k_maxdeg(const int32_t *__restrict__ deg, int64_t n, int32_t *__restrict__ mx) {
    const int64_t q = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (q < n) atomicMax(mx, deg[q]);
}

// tile-major, descending degree: the stable sort of these keys over the
// Morton labels is degree_sort_tiles
__global__ __launch_bounds__(kBT) void k_degkey(const int32_t *__restrict__ deg, int64_t n, uint64_t *__restrict__ key,
                                                uint32_t *__restrict__ val) {
    const int64_t q = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (q >= n) return;
    key[q] = ((uint64_t)(q / kTileMasses) << 16) | (uint64_t)(65535 - deg[q]);
    val[q] = (uint32_t)q;
}

// final order: perm[new] = caller's index, inv[caller's] = new, deg in the
// final labels
__global__ __launch_bounds__(kBT) void k_compose(const uint32_t *__restrict__ perm_m, const uint32_t *__restrict__ order,
                                                 const int32_t *__restrict__ deg_m, int64_t n, int32_t *__restrict__ perm,
                                                 int32_t *__restrict__ inv, int32_t *__restrict__ deg) {
    const int64_t q = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (q >= n) return;
    const uint32_t o = order[q];
    const int32_t c = (int32_t)perm_m[o];
    perm[q] = c;
    inv[c] = (int32_t)q;
    deg[q] = deg_m[o];
}

__global__ __launch_bounds__(kBT) void k_relabel(const int32_t *__restrict__ L, const int32_t *__restrict__ R, int64_t S,
                                                 const int32_t *__restrict__ inv, int32_t *__restrict__ fl,
                                                 int32_t *__restrict__ fr, uint32_t *__restrict__ cross) {
    const int64_t s = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (s >= S) return;
    const int32_t a = L[s], b = R[s];
    const bool ok = a >= 0 && b >= 0;
    const int32_t x = ok ? inv[a] : -1, y = ok ? inv[b] : -1;
    fl[s] = x;
    fr[s] = y;
    cross[s] = (ok && x / kTileMasses != y / kTileMasses) ? 1u : 0u;
}

// halo pairs (tile << 32 | outside mass), two per crossing spring
__global__ __launch_bounds__(kBT) void k_halo_pairs(const int32_t *__restrict__ fl, const int32_t *__restrict__ fr,
                                                    const uint32_t *__restrict__ cross, const uint32_t *__restrict__ at,
                                                    int64_t S, uint64_t *__restrict__ hk) {
    const int64_t s = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (s >= S || !cross[s]) return;
    const int64_t p = 2 * (int64_t)at[s];
    const int32_t x = fl[s], y = fr[s];
    hk[p] = ((uint64_t)(x / kTileMasses) << 32) | (uint32_t)y;
    hk[p + 1] = ((uint64_t)(y / kTileMasses) << 32) | (uint32_t)x;
}

// tile actuator pairs (tile << 32 | spring): one per tile an actuated
// spring has an endpoint in, the slots of absent pairs kNone64 (sorted last)
constexpr uint64_t kNone64 = ~(uint64_t)0;
__global__ __launch_bounds__(kBT) void k_act_pairs(const int32_t *__restrict__ fl, const int32_t *__restrict__ fr,
                                                   const int32_t *__restrict__ act, int64_t S,
                                                   uint64_t *__restrict__ ak) {
    const int64_t s = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (s >= S) return;
    const int32_t x = fl[s], y = fr[s];
    uint64_t p = kNone64, q = kNone64;
    if (act[s] >= 0 && x >= 0 && y >= 0) {
        const uint64_t tx = (uint64_t)(x / kTileMasses), ty = (uint64_t)(y / kTileMasses);
        p = (tx << 32) | (uint64_t)s;
        if (ty != tx) q = (ty << 32) | (uint64_t)s;
    }
    ak[2 * s] = p;
    ak[2 * s + 1] = q;
}

// the first of n sorted keys that is at least key
__device__ __forceinline__ int64_t lower_bound64(const uint64_t *__restrict__ v, int64_t n, uint64_t key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (v[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// each tile's first actuator pair (ntiles + 1 offsets: the last is the count)
__global__ __launch_bounds__(kBT) void k_tile_acts(const uint64_t *__restrict__ aks, int64_t n, int64_t ntiles,
                                                   int64_t *__restrict__ aoff) {
    const int64_t t = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (t <= ntiles) aoff[t] = lower_bound64(aks, n, (uint64_t)t << 32);
}

// each tile's halo range in the sorted unique pairs, its rows (with its
// actuator rows: aoff, or null when the body has none), the largest
__global__ __launch_bounds__(kBT) void k_tile_halos(const uint64_t *__restrict__ hu, const uint64_t *__restrict__ nu,
                                                    const int64_t *__restrict__ aoff, int64_t ntiles, int64_t N,
                                                    int64_t *__restrict__ hoff, int32_t *__restrict__ rmax) {
    const int64_t t = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (t > ntiles) return;
    const int64_t n = (int64_t)*nu;
    const uint64_t key = (uint64_t)t << 32;
    int64_t lo = 0, hi = n;  // first pair with key >= t << 32
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (hu[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    hoff[t] = lo;
    if (t < ntiles) {
        int64_t lo2 = lo, hi2 = n;
        const uint64_t key2 = (uint64_t)(t + 1) << 32;
        while (lo2 < hi2) {
            const int64_t mid = (lo2 + hi2) / 2;
            if (hu[mid] < key2) lo2 = mid + 1;
            else hi2 = mid;
        }
        const int64_t nm = min((int64_t)kTileMasses, N - t * kTileMasses);
        const int64_t na = aoff ? aoff[t + 1] - aoff[t] : 0;
        atomicMax(rmax, (int32_t)(nm + (lo2 - lo) + tile_act_rows(na)));
    }
}

// entries of each chunk: 2 * 64 per group, ceil(max degree / 2) groups
__global__ __launch_bounds__(kBT) void k_chunk_entries(const int32_t *__restrict__ deg, int64_t N, int64_t nch,
                                                       int64_t *__restrict__ ent) {
    const int64_t c = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (c >= nch) return;
    int32_t w = 0;
    for (int l = 0; l < kChunk; ++l) {
        const int64_t q = c * kChunk + l;
        if (q < N) w = max(w, deg[q]);
    }
    ent[c] = (int64_t)((w + kGroupTiled - 1) / kGroupTiled) * kGroupTiled * kChunk;
}

// the chunk holding entry e (coff ascending, nch + 1 entries)
__device__ __forceinline__ int64_t chunk_of(const int64_t *__restrict__ coff, int64_t nch, int64_t e) {
    int64_t lo = 0, hi = nch;  // last c with coff[c] <= e
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) / 2;
        if (coff[mid] <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

// padding everywhere: rest 1, zero constants; names of the mass itself (its
// row; row 0 past the body), the unused second half of each chunk's name
// words zero
__global__ __launch_bounds__(kBT) void k_pad_entries(const int64_t *__restrict__ coff, int64_t nch, int64_t N,
                                                     DevBuildOut o, bool pk, bool pd) {
    const int64_t e = (int64_t)blockIdx.x * kBT + threadIdx.x;
    const int64_t ne = coff[nch];
    if (e >= ne) return;
    const int64_t c = chunk_of(coff, nch, e);
    const int64_t rel = e - coff[c], len = coff[c + 1] - coff[c];
    o.inc_rest[o.inc0 + e] = 1.0;
    if (pk) o.inc_k[o.inc0 + e] = 0.0;
    if (pd) o.inc_damp[o.inc0 + e] = 0.0;
    // word e of the span: the names of entries 2 rel and 2 rel + 1, or unused
    uint32_t w = 0;
    if (rel < len / 2) {
        const int64_t m0 = (c / (kTileMasses / kChunk)) * kTileMasses;
        for (int h = 0; h < 2; ++h) {
            const int64_t r = 2 * rel + h;
            const int64_t me = c * kChunk + (r / kGroupTiled) % kChunk;
            const uint32_t row = me < N ? (uint32_t)(me - m0) : 0u;
            w |= row << (16 * h);
        }
    }
    o.inc_nbr[o.inc0 + e] = w;
}

// incidence items: the right end then the left end of each spring, in
// spring order (invalid springs sort last)
__global__ __launch_bounds__(kBT) void k_items(const int32_t *__restrict__ fl, const int32_t *__restrict__ fr, int64_t S,
                                               uint32_t *__restrict__ key, uint32_t *__restrict__ val) {
    const int64_t s = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (s >= S) return;
    const int32_t x = fl[s], y = fr[s];
    const bool ok = x >= 0 && y >= 0;
    key[2 * s] = ok ? (uint32_t)y : kNone;
    key[2 * s + 1] = ok ? (uint32_t)x : kNone;
    val[2 * s] = (uint32_t)(2 * s);
    val[2 * s + 1] = (uint32_t)(2 * s + 1);
}

struct PlaceArgs {
    const uint32_t *key, *val;
    const int32_t *fl, *fr;
    const int64_t *start, *coff, *hoff, *aoff;
    const uint64_t *hu, *aks;
    const double *rest, *k, *damp;
    const int32_t *act;
    int64_t nvalid, N;
    bool pk, pd;
};

// each incidence into its SELL entry
__global__ __launch_bounds__(kBT) void k_place(PlaceArgs p, DevBuildOut o) {
    const int64_t x = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (x >= p.nvalid) return;
    const int32_t q = (int32_t)p.key[x];
    const uint32_t v = p.val[x];
    const int64_t s = v >> 1;
    const int32_t other = (v & 1u) ? p.fr[s] : p.fl[s];
    const int64_t j = x - p.start[q];
    const int64_t c = q / kChunk, lane = q % kChunk;
    const int64_t rel = ((j / kGroupTiled) * kChunk + lane) * kGroupTiled + j % kGroupTiled;
    const int64_t e = p.coff[c] + rel;
    // the other endpoint's LDS row in q's tile
    const int64_t t = q / kTileMasses, m0 = t * kTileMasses;
    const int64_t nm = min((int64_t)kTileMasses, p.N - m0);
    uint32_t row;
    if (other >= m0 && other < m0 + nm) {
        row = (uint32_t)(other - m0);
    } else {
        int64_t lo = p.hoff[t], hi = p.hoff[t + 1];
        while (lo < hi) {
            const int64_t mid = (lo + hi) / 2;
            if ((int32_t)(uint32_t)p.hu[mid] < other) lo = mid + 1;
            else hi = mid;
        }
        row = (uint32_t)(nm + (lo - p.hoff[t]));
    }
    const int32_t a = p.act ? p.act[s] : -1;
    if (a >= 0) row |= kActBit16;
    reinterpret_cast<uint16_t *>(o.inc_nbr)[2 * (o.inc0 + p.coff[c]) + rel] = (uint16_t)row;
    if (a >= 0) {  // the spring's position in the tile's actuator list
        const int64_t a0 = p.aoff[t];
        const int64_t ix = lower_bound64(p.aks + a0, p.aoff[t + 1] - a0, ((uint64_t)t << 32) | (uint64_t)s);
        o.inc_rest[o.inc0 + e] = __longlong_as_double(ix);
    } else {
        o.inc_rest[o.inc0 + e] = p.rest[s];
    }
    if (p.pk) o.inc_k[o.inc0 + e] = p.k[s];
    if (p.pd) o.inc_damp[o.inc0 + e] = p.damp ? p.damp[s] : 0.0;
}

struct MassArgs {
    const int32_t *perm, *deg;
    const double *pos, *vel, *m, *T, *ext;
    const uint8_t *fixed;
    const int32_t *drag;
    int64_t N, npad;
};

// the mass arrays in the final order; padding slots past the body
__global__ __launch_bounds__(kBT) void k_masses(MassArgs a, DevBuildOut o, uint32_t *__restrict__ flags_out) {
    const int64_t q = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (q >= a.npad) return;
    const int64_t i = o.ch0 * kChunk + q;
    double *st = o.st + 6 * i;
    if (q < a.N) {
        const int64_t c = a.perm[q];
        for (int k = 0; k < 3; ++k) {
            st[k] = a.pos[3 * c + k];
            st[3 + k] = a.vel ? a.vel[3 * c + k] : 0.0;
        }
        o.m[i] = a.m[c];
        o.Tm[i] = a.T ? a.T[c] : 0.0;
        uint32_t f = 0;
        if (a.fixed && a.fixed[c]) f |= kFixed;
        if (a.ext && (a.ext[3 * c] != 0 || a.ext[3 * c + 1] != 0 || a.ext[3 * c + 2] != 0)) f |= kExt;
        if (a.drag && a.drag[c] != 0) f |= kDrag;
        o.mflags[i] = f;
        flags_out[q] = f;
        o.deg[i] = a.deg[q];
        if (o.ext)
            for (int k = 0; k < 3; ++k) o.ext[k * o.ncap + i] = a.ext ? a.ext[3 * c + k] : 0.0;
        if (o.drag) o.drag[i] = a.drag ? a.drag[c] : 0;
    } else {
        for (int k = 0; k < 6; ++k) st[k] = 0.0;
        o.m[i] = 1.0;
        o.Tm[i] = 0.0;
        o.mflags[i] = kPad;
        o.deg[i] = 0;
        if (o.ext)
            for (int k = 0; k < 3; ++k) o.ext[k * o.ncap + i] = 0.0;
        if (o.drag) o.drag[i] = 0;
    }
    for (int k = 0; k < 3; ++k) o.acc[k * o.ncap + i] = 0.0;
    if (o.mlc) o.mlc[i] = 0;
}

__global__ __launch_bounds__(kBT) void k_chunks(const int64_t *__restrict__ coff, int64_t nch, DevBuildOut o) {
    const int64_t c = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (c >= nch) return;
    o.chunk_ptr[o.ch0 + c] = o.inc0 + coff[c];
    o.chunk_ng[o.ch0 + c] = (int32_t)((coff[c + 1] - coff[c]) / (kGroupTiled * kChunk));
    // a tiled body's tiles compute their actuators: no chunk share
    o.chunk_act0[o.ch0 + c] = o.act0;
    o.chunk_nact[o.ch0 + c] = 0;
}

// the tiles' copies of their actuators' records
__global__ __launch_bounds__(kBT) void k_tact(const uint64_t *__restrict__ aks, const int32_t *__restrict__ act, int64_t n,
                                              DevBuildOut o) {
    const int64_t x = (int64_t)blockIdx.x * kBT + threadIdx.x;
    if (x >= n) return;
    const int64_t a = act[(uint32_t)aks[x]];
    for (int k = 0; k < 4; ++k) o.tact[4 * (o.ta0 + x) + k] = o.act[4 * (o.act0 + a) + k];
}

__global__ __launch_bounds__(kBT) void k_tiles(const int64_t *__restrict__ hoff, const uint64_t *__restrict__ hu,
                                               const int64_t *__restrict__ aoff, int64_t ntiles, int64_t N,
                                               DevBuildOut o) {
    const int64_t x = (int64_t)blockIdx.x * kBT + threadIdx.x;
    const int64_t nh = hoff[ntiles];
    if (x < ntiles) {
        TileDesc d{};
        d.body = -1;  // set when the body joins the launches
        d.chunk0 = (int32_t)(o.ch0 + x * (kTileMasses / kChunk));
        d.nmass = (int32_t)min((int64_t)kTileMasses, N - x * kTileMasses);
        d.nhalo = (int32_t)(hoff[x + 1] - hoff[x]);
        d.halo_off = o.hl0 + hoff[x];
        d.first = x == 0 ? 1 : 0;
        d.nact = aoff ? (int32_t)(aoff[x + 1] - aoff[x]) : 0;
        d.act_off = o.ta0 + (aoff ? aoff[x] : 0);
        o.tiles[o.tl0 + x] = d;
    }
    if (x < nh) o.halo[o.hl0 + x] = (int32_t)(o.ch0 * kChunk + (int64_t)(uint32_t)hu[x]);
}

}  // namespace

struct DevBuilder::Bufs {
    // inputs
    double *pos, *vel, *m, *T, *ext, *rest, *k, *damp;
    uint8_t *fixed;
    int32_t *drag, *L, *R, *act;
    // orders
    double *part, *box;
    uint64_t *code, *code2, *key2, *key2s, *hk, *hks, *hu, *nu, *ak, *aks;
    uint32_t *idx, *perm_m, *val2, *order, *cross, *at, *ik, *iks, *iv, *ivs;
    int32_t *inv_m, *deg_m, *perm, *inv, *deg, *fl, *fr, *maxdeg, *rmax;
    int64_t *start, *hoff, *aoff, *ent, *coff;
    uint32_t *flags;
    void *tmp;
    size_t tmp_bytes;
};

DevBuilder::~DevBuilder() { release(); }

void DevBuilder::release() {
    if (pool_) (void)hipFree(pool_);
    pool_ = nullptr;
    cap_ = 0;
    delete b_;
    b_ = nullptr;
}

#define BCHK(x)                                \
    do {                                       \
        const hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_;       \
    } while (0)

hipError_t DevBuilder::plan(const DevBuildIn &in, hipStream_t s, DevBuildSizes *sizes) {
    in_ = in;
    const int64_t N = in.N, S = in.S;
    const int64_t npad = (N + kChunk - 1) / kChunk * kChunk, nch = npad / kChunk;
    const int64_t ntiles = (N + kTileMasses - 1) / kTileMasses;
    sz_ = DevBuildSizes{};
    sz_.nch = nch;
    sz_.ntiles = ntiles;

    // rocPRIM scratch: the largest any step needs
    size_t tb = 0, t;
    BCHK(rocprim::radix_sort_pairs(nullptr, t, (uint64_t *)nullptr, (uint64_t *)nullptr, (uint32_t *)nullptr,
                                   (uint32_t *)nullptr, (size_t)N, 0, 64, s));
    tb = std::max(tb, t);
    BCHK(rocprim::exclusive_scan(nullptr, t, (uint32_t *)nullptr, (uint32_t *)nullptr, 0u, (size_t)std::max<int64_t>(S, 1),
                                 rocprim::plus<uint32_t>(), s));
    tb = std::max(tb, t);
    BCHK(rocprim::radix_sort_keys(nullptr, t, (uint64_t *)nullptr, (uint64_t *)nullptr, (size_t)(2 * S), 0, 64, s));
    tb = std::max(tb, t);
    BCHK(rocprim::unique(nullptr, t, (uint64_t *)nullptr, (uint64_t *)nullptr, (uint64_t *)nullptr, (size_t)(2 * S),
                         rocprim::equal_to<uint64_t>(), s));
    tb = std::max(tb, t);
    BCHK(rocprim::exclusive_scan(nullptr, t, (int64_t *)nullptr, (int64_t *)nullptr, (int64_t)0, (size_t)(nch + 1),
                                 rocprim::plus<int64_t>(), s));
    tb = std::max(tb, t);
    BCHK(rocprim::exclusive_scan(nullptr, t, (int32_t *)nullptr, (int64_t *)nullptr, (int64_t)0, (size_t)(N + 1),
                                 rocprim::plus<int64_t>(), s));
    tb = std::max(tb, t);
    BCHK(rocprim::radix_sort_pairs(nullptr, t, (uint32_t *)nullptr, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                   (uint32_t *)nullptr, (size_t)(2 * S), 0, 32, s));
    tb = std::max(tb, t);

    // carve the pool
    struct Carve {
        size_t at = 0;
        std::vector<std::pair<void **, size_t>> want;
        void add(void **p, size_t bytes) { want.push_back({p, bytes}); }
    } cv;
    if (!b_) b_ = new Bufs{};
    Bufs &b = *b_;
    std::memset(&b, 0, sizeof(b));
    auto add = [&](auto *&p, int64_t n) { cv.add(reinterpret_cast<void **>(&p), sizeof(*p) * (size_t)std::max<int64_t>(n, 1)); };
    add(b.pos, 3 * N);
    if (in.vel) add(b.vel, 3 * N);
    add(b.m, N);
    if (in.T) add(b.T, N);
    if (in.any_ext && in.ext) add(b.ext, 3 * N);
    if (in.fixed) add(b.fixed, N);
    if (in.any_drag && in.drag) add(b.drag, N);
    add(b.L, S);
    add(b.R, S);
    add(b.rest, S);
    if (in.per_k) add(b.k, S);
    if (in.per_d && in.damping) add(b.damp, S);
    if (in.act_local) add(b.act, S);
    add(b.part, 6 * kBoxBlocks);
    add(b.box, 6);
    add(b.code, N);
    add(b.code2, N);
    add(b.idx, N);
    add(b.perm_m, N);
    add(b.inv_m, N);
    add(b.deg_m, N);
    add(b.maxdeg, 1);
    add(b.key2, N);
    add(b.key2s, N);
    add(b.val2, N);
    add(b.order, N);
    add(b.perm, N);
    add(b.inv, N);
    add(b.deg, N + 1);
    add(b.start, N + 1);
    add(b.fl, S);
    add(b.fr, S);
    add(b.cross, S);
    add(b.at, S);
    add(b.hk, 2 * S);
    add(b.hks, 2 * S);
    add(b.hu, 2 * S);
    add(b.nu, 1);
    add(b.hoff, ntiles + 1);
    if (in.act_local) {
        add(b.ak, 2 * S);
        add(b.aks, 2 * S);
        add(b.aoff, ntiles + 1);
    }
    add(b.rmax, 1);
    add(b.ent, nch + 1);
    add(b.coff, nch + 1);
    add(b.ik, 2 * S);
    add(b.iks, 2 * S);
    add(b.iv, 2 * S);
    add(b.ivs, 2 * S);
    add(b.flags, N);
    size_t total = 0;
    for (auto &w : cv.want) total += (w.second + 255) & ~(size_t)255;
    total += (tb + 255) & ~(size_t)255;
    if (total > cap_) {
        if (pool_) {
            BCHK(hipStreamSynchronize(s));
            (void)hipFree(pool_);
            pool_ = nullptr;
            cap_ = 0;
        }
        BCHK(hipMalloc((void **)&pool_, total));
        cap_ = total;
    }
    size_t at = 0;
    for (auto &w : cv.want) {
        *w.first = pool_ + at;
        at += (w.second + 255) & ~(size_t)255;
    }
    b.tmp = pool_ + at;
    b.tmp_bytes = tb;

    // uploads (the caller's arrays)
    auto up = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
        if (!src || bytes == 0) return hipSuccess;
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    };
    BCHK(up(b.pos, in.pos, 24 * (size_t)N));
    BCHK(up(b.vel, in.vel, 24 * (size_t)N));
    BCHK(up(b.m, in.m, 8 * (size_t)N));
    BCHK(up(b.T, in.T, 8 * (size_t)N));
    if (b.ext) BCHK(up(b.ext, in.ext, 24 * (size_t)N));
    BCHK(up(b.fixed, in.fixed, (size_t)N));
    if (b.drag) BCHK(up(b.drag, in.drag, 4 * (size_t)N));
    BCHK(up(b.L, in.left, 4 * (size_t)S));
    BCHK(up(b.R, in.right, 4 * (size_t)S));
    BCHK(up(b.rest, in.rest, 8 * (size_t)S));
    if (b.k) BCHK(up(b.k, in.k, 8 * (size_t)S));
    if (b.damp) BCHK(up(b.damp, in.damping, 8 * (size_t)S));
    if (b.act) BCHK(up(b.act, in.act_local, 4 * (size_t)S));

    // Morton order (morton_order): the bounding box of the finite
    // coordinates, then codes, sorted stably
    const int nparts = (int)std::min<int64_t>(kBoxBlocks, std::max<int64_t>(1, (N + kBT - 1) / kBT));
    hipLaunchKernelGGL(k_box_part, dim3(nparts), dim3(kBT), 0, s, b.pos, N, b.part);
    hipLaunchKernelGGL(k_box_final, dim3(1), dim3(kBT), 0, s, b.part, nparts, b.box);
    double box[6];
    BCHK(hipMemcpyAsync(box, b.box, sizeof box, hipMemcpyDeviceToHost, s));
    BCHK(hipStreamSynchronize(s));
    Box bx;
    for (int c = 0; c < 3; ++c) {  // as morton_order computes them
        double lo = box[c], hi = box[3 + c];
        const double ext = hi - lo;
        bx.scale[c] = (std::isfinite(ext) && ext > 0) ? ((double)((1 << 21) - 1) / ext) : 0.0;
        if (!std::isfinite(lo)) lo = 0.0;
        bx.lo[c] = lo;
    }
    hipLaunchKernelGGL(k_codes, dim3(blocks_for(N)), dim3(kBT), 0, s, b.pos, N, bx, b.code, b.idx);
    size_t tmpb = b.tmp_bytes;
    BCHK(rocprim::radix_sort_pairs(b.tmp, tmpb, b.code, b.code2, b.idx, b.perm_m, (size_t)N, 0, 64, s));
    hipLaunchKernelGGL(k_invert, dim3(blocks_for(N)), dim3(kBT), 0, s, b.perm_m, N, b.inv_m);
    // degrees (labels do not matter), then the degree order inside tiles
    BCHK(hipMemsetAsync(b.deg_m, 0, sizeof(int32_t) * N, s));
    BCHK(hipMemsetAsync(b.maxdeg, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_degrees, dim3(blocks_for(S)), dim3(kBT), 0, s, b.L, b.R, S, b.inv_m, b.deg_m);
    hipLaunchKernelGGL(k_maxdeg, dim3(blocks_for(N)), dim3(kBT), 0, s, b.deg_m, N, b.maxdeg);
    hipLaunchKernelGGL(k_degkey, dim3(blocks_for(N)), dim3(kBT), 0, s, b.deg_m, N, b.key2, b.val2);
    tmpb = b.tmp_bytes;
    BCHK(rocprim::radix_sort_pairs(b.tmp, tmpb, b.key2, b.key2s, b.val2, b.order, (size_t)N, 0, 64, s));
    hipLaunchKernelGGL(k_compose, dim3(blocks_for(N)), dim3(kBT), 0, s, b.perm_m, b.order, b.deg_m, N, b.perm, b.inv,
                       b.deg);
    // springs in the final labels; the crossing ones give the halos
    hipLaunchKernelGGL(k_relabel, dim3(blocks_for(S)), dim3(kBT), 0, s, b.L, b.R, S, b.inv, b.fl, b.fr, b.cross);
    int64_t ncross = 0;
    if (S > 0) {
        tmpb = b.tmp_bytes;
        BCHK(rocprim::exclusive_scan(b.tmp, tmpb, b.cross, b.at, 0u, (size_t)S, rocprim::plus<uint32_t>(), s));
        uint32_t last_at = 0, last_cross = 0;
        BCHK(hipMemcpyAsync(&last_at, b.at + (S - 1), 4, hipMemcpyDeviceToHost, s));
        BCHK(hipMemcpyAsync(&last_cross, b.cross + (S - 1), 4, hipMemcpyDeviceToHost, s));
        BCHK(hipStreamSynchronize(s));
        ncross = (int64_t)last_at + last_cross;
    }
    hipLaunchKernelGGL(k_halo_pairs, dim3(blocks_for(S)), dim3(kBT), 0, s, b.fl, b.fr, b.cross, b.at, S, b.hk);
    BCHK(hipMemsetAsync(b.nu, 0, sizeof(uint64_t), s));
    if (ncross > 0) {
        tmpb = b.tmp_bytes;
        BCHK(rocprim::radix_sort_keys(b.tmp, tmpb, b.hk, b.hks, (size_t)(2 * ncross), 0, 64, s));
        tmpb = b.tmp_bytes;
        BCHK(rocprim::unique(b.tmp, tmpb, b.hks, b.hu, b.nu, (size_t)(2 * ncross), rocprim::equal_to<uint64_t>(), s));
    }
    // each tile's actuators: the sorted tile actuator pairs
    if (b.act && S > 0) {
        hipLaunchKernelGGL(k_act_pairs, dim3(blocks_for(S)), dim3(kBT), 0, s, b.fl, b.fr, b.act, S, b.ak);
        tmpb = b.tmp_bytes;
        BCHK(rocprim::radix_sort_keys(b.tmp, tmpb, b.ak, b.aks, (size_t)(2 * S), 0, 64, s));
        hipLaunchKernelGGL(k_tile_acts, dim3(blocks_for(ntiles + 1)), dim3(kBT), 0, s, b.aks, 2 * S, ntiles, b.aoff);
    } else if (b.act) {
        BCHK(hipMemsetAsync(b.aoff, 0, sizeof(int64_t) * (ntiles + 1), s));
    }
    BCHK(hipMemsetAsync(b.rmax, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_tile_halos, dim3(blocks_for(ntiles + 1)), dim3(kBT), 0, s, b.hu, b.nu, b.aoff, ntiles, N,
                       b.hoff, b.rmax);
    // entries per chunk and their offsets
    hipLaunchKernelGGL(k_chunk_entries, dim3(blocks_for(nch)), dim3(kBT), 0, s, b.deg, N, nch, b.ent);
    tmpb = b.tmp_bytes;
    BCHK(rocprim::exclusive_scan(b.tmp, tmpb, b.ent, b.coff, (int64_t)0, (size_t)(nch + 1), rocprim::plus<int64_t>(), s));
    BCHK(hipMemsetAsync(b.deg + N, 0, sizeof(int32_t), s));
    tmpb = b.tmp_bytes;
    BCHK(rocprim::exclusive_scan(b.tmp, tmpb, b.deg, b.start, (int64_t)0, (size_t)(N + 1), rocprim::plus<int64_t>(), s));
    BCHK(hipGetLastError());
    int64_t nhalo = 0, ninc = 0, ntact = 0;
    int32_t rmax = 0, maxdeg = 0;
    BCHK(hipMemcpyAsync(&nhalo, b.nu, 8, hipMemcpyDeviceToHost, s));
    if (b.aoff) BCHK(hipMemcpyAsync(&ntact, b.aoff + ntiles, 8, hipMemcpyDeviceToHost, s));
    BCHK(hipMemcpyAsync(&ninc, b.coff + nch, 8, hipMemcpyDeviceToHost, s));
    BCHK(hipMemcpyAsync(&rmax, b.rmax, 4, hipMemcpyDeviceToHost, s));
    BCHK(hipMemcpyAsync(&maxdeg, b.maxdeg, 4, hipMemcpyDeviceToHost, s));
    BCHK(hipStreamSynchronize(s));
    sz_.n_halo = nhalo;
    sz_.n_inc = ninc;
    sz_.n_tact = ntact;
    sz_.rmax = rmax;
    sz_.fits = rmax <= in.tile_rows && maxdeg <= 65535;
    *sizes = sz_;
    return hipSuccess;
}

hipError_t DevBuilder::write(const DevBuildOut &o, const DevBuildMirror &mr, hipStream_t s) {
    if (!b_) return hipErrorInvalidValue;
    Bufs &b = *b_;
    const int64_t N = in_.N, S = in_.S, nch = sz_.nch, ntiles = sz_.ntiles;
    const int64_t npad = nch * kChunk;
    const bool pk = in_.per_k, pd = in_.per_d;
    if (sz_.n_inc > 0)
        hipLaunchKernelGGL(k_pad_entries, dim3(blocks_for(sz_.n_inc)), dim3(kBT), 0, s, b.coff, nch, N, o, pk, pd);
    int64_t nvalid = 0;
    if (S > 0) {
        hipLaunchKernelGGL(k_items, dim3(blocks_for(S)), dim3(kBT), 0, s, b.fl, b.fr, S, b.ik, b.iv);
        size_t tmpb = b.tmp_bytes;
        BCHK(rocprim::radix_sort_pairs(b.tmp, tmpb, b.ik, b.iks, b.iv, b.ivs, (size_t)(2 * S), 0, 32, s));
        BCHK(hipMemcpyAsync(&nvalid, b.start + N, 8, hipMemcpyDeviceToHost, s));
        BCHK(hipStreamSynchronize(s));
        PlaceArgs p{b.iks,  b.ivs, b.fl,   b.fr,  b.start, b.coff,  b.hoff, b.aoff, b.hu,
                    b.aks,  b.rest, b.k,   b.damp, b.act, nvalid, N,      pk,     pd};
        if (nvalid > 0) hipLaunchKernelGGL(k_place, dim3(blocks_for(nvalid)), dim3(kBT), 0, s, p, o);
    }
    MassArgs ma{b.perm, b.deg, b.pos, b.vel, b.m, b.T, b.ext, b.fixed, b.drag, N, npad};
    hipLaunchKernelGGL(k_masses, dim3(blocks_for(npad)), dim3(kBT), 0, s, ma, o, b.flags);
    hipLaunchKernelGGL(k_chunks, dim3(blocks_for(nch)), dim3(kBT), 0, s, b.coff, nch, o);
    hipLaunchKernelGGL(k_tiles, dim3(blocks_for(std::max(ntiles, sz_.n_halo))), dim3(kBT), 0, s, b.hoff, b.hu, b.aoff,
                       ntiles, N, o);
    if (sz_.n_tact > 0)
        hipLaunchKernelGGL(k_tact, dim3(blocks_for(sz_.n_tact)), dim3(kBT), 0, s, b.aks, b.act, sz_.n_tact, o);
    BCHK(hipGetLastError());
    // host mirror
    BCHK(hipMemcpyAsync(mr.perm, b.perm, 4 * (size_t)N, hipMemcpyDeviceToHost, s));
    BCHK(hipMemcpyAsync(mr.flags, b.flags, 4 * (size_t)N, hipMemcpyDeviceToHost, s));
    if (S > 0) {
        BCHK(hipMemcpyAsync(mr.sl, b.fl, 4 * (size_t)S, hipMemcpyDeviceToHost, s));
        BCHK(hipMemcpyAsync(mr.sr, b.fr, 4 * (size_t)S, hipMemcpyDeviceToHost, s));
    }
    BCHK(hipMemcpyAsync(mr.coff, b.coff, 8 * (size_t)(nch + 1), hipMemcpyDeviceToHost, s));
    BCHK(hipStreamSynchronize(s));
    return hipSuccess;
}

}  // namespace titan_mi355x
