// Host runtime of the MI355X step engine and its C ABI (include/titan_mi355x.h).
//
// Replaces the reference's per-object cudaMalloc staging (massToArray /
// springToArray, src/sim.cu:829-914), its pointer-table device vectors
// (include/Titan/sim.h:150-164) and its asynchronous driver thread
// (src/sim.cu:1389,1503-1617) with:
//   * arenas of 64-mass chunks and SELL-64 incidence slots (layout.h), grown
//     by doubling within an optional HBM budget;
//   * bodies uploaded in bulk, one H2D copy per array;
//   * synchronous, stream-ordered stepping: n steps are n launches of the
//     fused step kernel on the context's stream, the bulk of them replayed
//     from a captured hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "builder.h"
#include "engine_internal.h"
#include "kernels.h"
#include "layout.h"
#include "tiles.h"
#include "titan_mi355x.h"

using namespace titan_mi355x;

namespace {

struct Fail {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string &m) { throw Fail{code, m}; }

#define HIPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) fail(TITAN_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// First-fit allocator over [0, cap) used for chunks, incidence slots,
// actuators, local constraints and body ids.
class Ranges {
   public:
    int64_t cap() const { return cap_; }
    int64_t alloc(int64_t len) {
        if (len <= 0) return 0;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= len) {
                const int64_t b = it->first, rem = it->second - len;
                free_.erase(it);
                if (rem) free_[b + len] = rem;
                return b;
            }
        }
        return -1;
    }
    void release(int64_t b, int64_t len) {
        if (len <= 0) return;
        auto it = free_.emplace(b, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    void grow(int64_t newcap) {
        if (newcap <= cap_) return;
        release(cap_, newcap - cap_);
        cap_ = newcap;
    }
    // end of the highest allocated range
    int64_t high_water() const {
        if (free_.empty()) return cap_;
        auto last = std::prev(free_.end());
        return (last->first + last->second == cap_) ? last->first : cap_;
    }

   private:
    std::map<int64_t, int64_t> free_;
    int64_t cap_ = 0;
};

struct BodyH {
    bool live = false;
    bool tiled = false;
    bool resident = false;
    uint8_t kd = 0;    // its b_kd bits
    int64_t res = -1;  // resident descriptor (resident layout)
    int32_t res_cache = 0;  // LDS bytes of its incidence copy (resident layout; 0: none)
    int32_t res_scache = 0;  // LDS bytes of its split-form region (0: none)
    int64_t n_masses = 0, chunk0 = 0, nchunks = 0;
    int64_t inc0 = 0, inc_len = 0, act0 = 0, act_len = 0, lc0 = 0, lc_len = 0;
    int64_t tile0 = 0, ntiles = 0, halo0 = 0, nhalo = 0, tact0 = 0, tact_len = 0;
    int32_t rows = 0;  // largest LDS row count of a tile (tiled layout)
    double T = 0;   // host mirror of the body clock
    int64_t n = 0;  // steps taken (parity selects the state buffer)
    std::vector<uint32_t> flags;   // device order
    std::vector<int32_t> perm;     // device order -> caller's order (empty = identity)
    // topology, for removals (titan_body_remove): each spring's endpoints in
    // device order (-1: none) in the caller's spring order, removed springs,
    // each chunk's first incidence entry (body-relative, nch + 1 entries),
    // and the first mass of each chunk's tile (the LDS row base of the
    // tiled and resident layouts)
    std::vector<int32_t> sl, sr;
    std::vector<uint8_t> sdead;
    std::vector<int64_t> coff;
    std::vector<int32_t> row0;
    // per-mass spring lists in ascending spring index, built on the first
    // removal (mptr: N + 1 offsets into mspr)
    std::vector<int64_t> mptr;
    std::vector<int32_t> mspr;
};

}  // namespace

struct titan_ctx {
    int device = 0;
    hipStream_t stream = nullptr;  // every launch, in order
    hipStream_t up = nullptr;      // uploads of new bodies into free arena space
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timing = false;
    std::string err;
    uint64_t budget = 0, used = 0;

    double dt = 0.0001;                  // src/sim.cu:78
    int32_t integ = kIntEuler;           // the reference's build option (CMakeLists.txt:13-14)
    double *rk = nullptr;                // RK2: state saved by a first stage, 6 per mass slot
    double g[3] = {0.0, 0.0, -9.81};     // src/sim.cu:86
    double T = 0.0;                      // Simulation::T
    std::vector<double> planes_h, balls_h;

    Ranges chunks, incs, acts, lcs, bids, tiles, halos, tacts, rdescs;
    std::vector<BodyH> bodies;
    // bodies of at most four chunks resident (alone or among few, one launch
    // per stepping call with the split form), every larger body tiled
    titan_schedule sched{0, kTileRows, kSplitMaxChunks * kChunk};

    // mass arena (ncap = chunks.cap() * 64)
    int64_t ncap = 0;
    double *st = nullptr, *m = nullptr, *Tm = nullptr, *acc = nullptr, *ext = nullptr;
    uint32_t *mflags = nullptr;
    int32_t *deg = nullptr, *drag = nullptr;
    int64_t *mlc = nullptr;
    int32_t *chunk_body = nullptr;
    int64_t *chunk_ptr = nullptr;
    int32_t *chunk_ng = nullptr;   // incidence groups per lane (uniform over the chunk)
    int64_t *chunk_act0 = nullptr; // the chunk's share of its body's actuators
    int32_t *chunk_nact = nullptr;
    // incidence arena
    uint32_t *inc_nbr = nullptr;
    double *inc_rest = nullptr, *inc_k = nullptr, *inc_damp = nullptr;
    double *act = nullptr;
    double *ract[2] = {nullptr, nullptr};  // actuation factors by launch slot
    LocalC *lc = nullptr;
    // tiled layout arenas
    TileDesc *tile = nullptr;
    int32_t *halo = nullptr;
    double *tact = nullptr;  // 4 per tile actuator: a copy of its actuator's record
    // resident layout: one descriptor per body
    ResidentDesc *rdesc = nullptr;
    uint8_t *b_mode = nullptr;  // per body: kModeGather, kModeTiled, kModeResident
    // body table
    int64_t bcap = 0;
    double *bT[2] = {nullptr, nullptr};
    int64_t *bn[2] = {nullptr, nullptr}, *bleft[2] = {nullptr, nullptr}, *bnsync = nullptr;
    int32_t *b_chunk0 = nullptr, *b_lc = nullptr, *b_lcn = nullptr;
    double *b_k = nullptr, *b_d = nullptr;  // the body's k / damping when all its springs share it
    double *b_m = nullptr;                  // the body's mass value when all its masses share it
    uint8_t *b_kd = nullptr;                 // bit 0: k per incidence, bit 1: damping per incidence
    double *planes = nullptr, *balls = nullptr;
    int64_t planes_cap = 0, balls_cap = 0;
    // ball grid (rebuilt before the next launch after the balls change)
    bool ball_grid_dirty = false;
    int32_t *bg_start = nullptr;
    double *bg_rec = nullptr;  // 4 per listed ball, plus one spare record
    double bg_x0 = 0, bg_y0 = 0, bg_z0 = 0, bg_inv_h = 0;
    int32_t bg_n[3] = {0, 0, 0};

    bool has_act = false;  // some spring is actuated: launch the actuated variants
    int res_lds_max = 0;   // LDS a resident workgroup may have
    int split_lds_max = 0;  // dynamic LDS a split-form resident workgroup may have
    int n_cu = 0;          // compute units of the device
    int slot = 0;

    // titan_step_noalloc: the device's scratch array (its address on this
    // context's device) and, during such a call, where the step kernels take
    // their actuation factors and RK2 stage state from (null outside a call)
    double *scratch = nullptr;
    double *na_ract[2] = {nullptr, nullptr};
    double *na_rk = nullptr;
    bool in_noalloc = false;
    // whole-range stepping replays captured 64-step graphs; off when the
    // environment sets TITAN_GRAPHS=0, and by default under the rocprofv3
    // tracer (which sets ROCPROF_OUTPUT_PATH for the program it runs): a
    // traced run of several contexts in one process once died in the host
    // runtime inside hipGraphLaunch, with no fault found in the engine's use
    // of the graph, so traced runs launch each step directly instead
    bool use_graphs = true;
    // tiled bodies of at least this many masses are laid out on the device
    // (TITAN_DEVICE_BUILD: 0 never, a number: that threshold); its scratch
    DevBuilder *builder = nullptr;
    int64_t device_min = kDeviceBuildMin;
    // captured multi-step graphs, valid while `version` is unchanged, keyed
    // by the launch shape they were captured with (most recent last)
    uint64_t version = 1;
    struct Graph {
        hipGraphExec_t exec;
        uint64_t version;
        int slot;
        int64_t c1, t1;
        int32_t ldim;
        bool gather, tiled;
    };
    std::vector<Graph> graphs;
    // pinned staging of per-body step budgets, two buffers used in turn
    int64_t *left_pin[2] = {nullptr, nullptr};
    hipEvent_t left_ev[2] = {nullptr, nullptr};
    int64_t left_cap = 0;
    int left_k = 0;
    // pinned staging for body downloads and uploads (grown on demand)
    double *stage = nullptr;
    int64_t stage_cap = 0;
    int32_t inject[3] = {-1, -1, -1};  // by TITAN_INJECT_* site: calls left before a forced failure

    std::map<void *, size_t> allocs;
    // destroyed bodies whose slots are released once `ev` (recorded after
    // their deactivation on the step stream) has completed
    struct Pending {
        BodyH b;
        int32_t id;
        hipEvent_t ev;
    };
    std::vector<Pending> pending;
};

namespace {

// The scratch array of titan_step_noalloc: one per device, held until the
// process ends.  Host code uses only the address hipGetSymbolAddress gives
// for it on each device (device_scratch), never the array's name.
constexpr int64_t kScratchBytes = (int64_t)64 << 20;
static __device__ __align__(256) double scratch[kScratchBytes / 8];

constexpr int kGraphSteps = 64;  // even: a replay leaves the launch slot unchanged
constexpr size_t kGraphCache = 8;  // captured graphs kept per context
// Steps are graph-captured only when a step's launches have fewer
// workgroups than this, about one round of the device: a longer step hides
// the host's launch cost anyway, and captured launches of large steps
// measured 3-4% slower than the same launches issued directly (1M lattice
// 97.0-97.4 against 93.5-94.6 us per step, 2M kNN cloud 198 against 190 us,
// alternating on one box; a 4096-robot population equal either way).
constexpr int64_t kGraphMaxBlocks = 1024;
thread_local std::string g_create_err;

template <class T>
T *dalloc(titan_ctx *c, int64_t n) {
    if (n <= 0) return nullptr;
    const size_t bytes = sizeof(T) * (size_t)n;
    if (c->budget && c->used + bytes > c->budget)
        fail(TITAN_E_NOMEM, "HBM budget exhausted: need " + std::to_string(bytes) + " more bytes, " +
                                std::to_string(c->budget - c->used) + " left");
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) fail(TITAN_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->used += bytes;
    c->allocs[p] = bytes;
    return (T *)p;
}

void dfree(titan_ctx *c, void *p) {
    if (!p) return;
    auto it = c->allocs.find(p);
    if (it != c->allocs.end()) {
        c->used -= it->second;
        c->allocs.erase(it);
    }
    (void)hipFree(p);
}

// Fail before an arena grows several arrays unless all of the new arrays fit
// the budget together (the old ones are freed one by one as they are
// replaced), so that running out of budget never leaves an arena half-grown.
void check_growth(titan_ctx *c, size_t bytes) {
    if (c->budget && c->used + bytes > c->budget)
        fail(TITAN_E_NOMEM, "HBM budget exhausted: growing an arena needs " + std::to_string(bytes) +
                                " bytes, " + std::to_string(c->budget - c->used) + " left");
}

// Reallocate a K-component SoA array from cap_old to cap_new entries per
// component, keeping the first cap_old entries of each component.
template <class T>
void grow_soa(titan_ctx *c, T *&p, int K, int64_t cap_old, int64_t cap_new, bool keep = true) {
    T *q = dalloc<T>(c, (int64_t)K * cap_new);
    HIPCHK(hipMemsetAsync(q, 0, sizeof(T) * (size_t)K * cap_new, c->stream));
    if (p && keep && cap_old > 0) {
        for (int k = 0; k < K; ++k)
            HIPCHK(hipMemcpyAsync(q + (int64_t)k * cap_new, p + (int64_t)k * cap_old,
                                  sizeof(T) * cap_old, hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    dfree(c, p);
    p = q;
}

template <class T>
void h2d(titan_ctx *c, T *dst, const T *src, int64_t n) {
    if (n > 0) HIPCHK(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyHostToDevice, c->stream));
}
template <class T>
void d2h(titan_ctx *c, T *dst, const T *src, int64_t n) {
    if (n > 0) HIPCHK(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
}

void drop_graphs(titan_ctx *c) {
    if (!c->graphs.empty()) {
        (void)hipStreamSynchronize(c->stream);  // none of them may still be running
        for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);
        c->graphs.clear();
    }
    c->version++;
}

// The capacity an arena of capacity `old` grows to when it must hold `need`
// entries (at least doubling, at least `floor`), and the device bytes of the
// arrays each arena entry carries: what the growth checks count.
int64_t grown_cap(int64_t old, int64_t need, int64_t floor) {
    return old >= need ? old : std::max<int64_t>(need, std::max<int64_t>(2 * old, floor));
}
int64_t max_chunk_cap() { return kMaxMassSlots / kChunk; }  // the kernels' 32-bit state byte offsets
size_t chunk_bytes(const titan_ctx *c) {  // a chunk: 64 mass slots and the chunk table
    return (size_t)kChunk * (96 + 8 + 8 + 24 + 4 + 4 + (c->ext ? 24 : 0) + (c->drag ? 4 : 0) + (c->mlc ? 8 : 0) +
                             (c->integ == kIntRK2 ? 48 : 0)) +
           (4 + 8 + 4 + 8 + 4);
}
size_t inc_bytes(const titan_ctx *c) { return 4 + 8 + (c->inc_k ? 8 : 0) + (c->inc_damp ? 8 : 0); }
size_t act_bytes(const titan_ctx *c) { return 4 * 8 + (c->ract[0] ? 2 * 8 : 0); }
size_t body_bytes() { return 2 * (8 + 8 + 8) + 8 + 4 + 4 + 1 + 4 + 8 + 8 + 1; }

void ensure_chunks(titan_ctx *c, int64_t need) {
    if (c->chunks.cap() >= need) return;
    const int64_t old = c->chunks.cap();
    if (need > max_chunk_cap()) fail(TITAN_E_NOMEM, "mass arena limit: at most 2^26 mass slots");
    const int64_t nc = std::min(grown_cap(old, need, 16), max_chunk_cap());
    const int64_t on = old * kChunk, nn = nc * kChunk;
    check_growth(c, (size_t)nc * chunk_bytes(c));
    grow_soa(c, c->st, 2, 6 * on, 6 * nn);
    if (c->integ == kIntRK2) grow_soa(c, c->rk, 1, 6 * on, 6 * nn);
    grow_soa(c, c->m, 1, on, nn);
    grow_soa(c, c->Tm, 1, on, nn);
    grow_soa(c, c->acc, 3, on, nn);
    grow_soa(c, c->mflags, 1, on, nn);
    grow_soa(c, c->deg, 1, on, nn);
    if (c->ext) grow_soa(c, c->ext, 3, on, nn);
    if (c->drag) grow_soa(c, c->drag, 1, on, nn);
    if (c->mlc) grow_soa(c, c->mlc, 1, on, nn);
    // chunk_body of free chunks is -1
    int32_t *cb = dalloc<int32_t>(c, nc);
    HIPCHK(hipMemsetAsync(cb, 0xff, sizeof(int32_t) * nc, c->stream));
    if (c->chunk_body) HIPCHK(hipMemcpyAsync(cb, c->chunk_body, sizeof(int32_t) * old, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    dfree(c, c->chunk_body);
    c->chunk_body = cb;
    grow_soa(c, c->chunk_ptr, 1, old, nc);
    grow_soa(c, c->chunk_ng, 1, old, nc);
    grow_soa(c, c->chunk_act0, 1, old, nc);
    grow_soa(c, c->chunk_nact, 1, old, nc);
    c->chunks.grow(nc);
    c->ncap = nn;
    drop_graphs(c);
}

void ensure_incs(titan_ctx *c, int64_t need) {
    if (c->incs.cap() >= need) return;
    const int64_t old = c->incs.cap();
    const int64_t nc = grown_cap(old, need, 4096);
    check_growth(c, (size_t)nc * inc_bytes(c));
    grow_soa(c, c->inc_nbr, 1, old, nc);
    grow_soa(c, c->inc_rest, 1, old, nc);
    if (c->inc_k) grow_soa(c, c->inc_k, 1, old, nc);
    if (c->inc_damp) grow_soa(c, c->inc_damp, 1, old, nc);
    c->incs.grow(nc);
    drop_graphs(c);
}

void ensure_acts(titan_ctx *c, int64_t need) {
    if (c->acts.cap() >= need) return;
    const int64_t old = c->acts.cap();
    const int64_t nc = grown_cap(old, need, 1024);
    check_growth(c, (size_t)nc * act_bytes(c));
    grow_soa(c, c->act, 1, 4 * old, 4 * nc);
    if (c->ract[0]) {
        grow_soa(c, c->ract[0], 1, old, nc);
        grow_soa(c, c->ract[1], 1, old, nc);
    }
    c->acts.grow(nc);
    drop_graphs(c);
}

void ensure_lcs(titan_ctx *c, int64_t need) {
    if (c->lcs.cap() >= need) return;
    const int64_t old = c->lcs.cap();
    const int64_t nc = std::max<int64_t>(need, std::max<int64_t>(2 * old, 64));
    grow_soa(c, c->lc, 1, old, nc);
    c->lcs.grow(nc);
    drop_graphs(c);
}

void ensure_bodies(titan_ctx *c, int64_t need) {
    if (c->bcap >= need) return;
    const int64_t old = c->bcap;
    const int64_t nc = grown_cap(old, need, 64);
    check_growth(c, (size_t)nc * body_bytes());
    for (int s = 0; s < 2; ++s) {
        grow_soa(c, c->bT[s], 1, old, nc);
        grow_soa(c, c->bn[s], 1, old, nc);
        grow_soa(c, c->bleft[s], 1, old, nc);
    }
    grow_soa(c, c->bnsync, 1, old, nc);
    grow_soa(c, c->b_lc, 1, old, nc);
    grow_soa(c, c->b_lcn, 1, old, nc);
    grow_soa(c, c->b_mode, 1, old, nc);
    grow_soa(c, c->b_k, 1, old, nc);
    grow_soa(c, c->b_d, 1, old, nc);
    grow_soa(c, c->b_m, 1, old, nc);
    grow_soa(c, c->b_kd, 1, old, nc);
    int32_t *q = dalloc<int32_t>(c, nc);
    HIPCHK(hipMemsetAsync(q, 0xff, sizeof(int32_t) * nc, c->stream));
    if (c->b_chunk0) HIPCHK(hipMemcpyAsync(q, c->b_chunk0, sizeof(int32_t) * old, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    dfree(c, c->b_chunk0);
    c->b_chunk0 = q;
    c->bids.grow(nc);
    c->bcap = nc;
    c->bodies.resize(nc);
    drop_graphs(c);
}

void ensure_tiles(titan_ctx *c, int64_t need) {
    if (c->tiles.cap() >= need) return;
    const int64_t old = c->tiles.cap();
    const int64_t nc = grown_cap(old, need, 64);
    check_growth(c, (size_t)nc * sizeof(TileDesc));
    grow_soa(c, c->tile, 1, old, nc);
    // launches cover the whole arena: unused tiles are marked free
    std::vector<TileDesc> dead(nc - old);
    for (auto &t : dead) {
        t = TileDesc{};
        t.body = -1;
    }
    h2d(c, c->tile + old, dead.data(), nc - old);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->tiles.grow(nc);
    drop_graphs(c);
}

void ensure_rdescs(titan_ctx *c, int64_t need) {
    if (c->rdescs.cap() >= need) return;
    const int64_t old = c->rdescs.cap();
    const int64_t nc = grown_cap(old, need, 64);
    check_growth(c, (size_t)nc * sizeof(ResidentDesc));
    grow_soa(c, c->rdesc, 1, old, nc);
    std::vector<ResidentDesc> dead(nc - old, ResidentDesc{-1, 0, 0, 0, 0, 0, 0, 0, 0, 0});
    h2d(c, c->rdesc + old, dead.data(), nc - old);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->rdescs.grow(nc);
    drop_graphs(c);
}

void ensure_halos(titan_ctx *c, int64_t need) {
    if (c->halos.cap() >= need) return;
    const int64_t old = c->halos.cap();
    const int64_t nc = grown_cap(old, need, 1024);
    check_growth(c, (size_t)nc * sizeof(int32_t));
    grow_soa(c, c->halo, 1, old, nc);
    c->halos.grow(nc);
    drop_graphs(c);
}

void ensure_tacts(titan_ctx *c, int64_t need) {
    if (c->tacts.cap() >= need) return;
    const int64_t old = c->tacts.cap();
    const int64_t nc = grown_cap(old, need, 1024);
    check_growth(c, (size_t)nc * 4 * sizeof(double));
    grow_soa(c, c->tact, 1, 4 * old, 4 * nc);
    c->tacts.grow(nc);
    drop_graphs(c);
}

// Create the per-incidence spring constant (or damping) array the first time
// a body needs one.  At least one entry, so that it exists (and grows with
// the arena) even while the arena is empty.
void make_inc_array(titan_ctx *c, double *&p) {
    if (p) return;
    const int64_t n = std::max<int64_t>(c->incs.cap(), 1);
    p = dalloc<double>(c, n);
    drop_graphs(c);
}

void make_act(titan_ctx *c) {
    if (c->has_act) return;
    c->has_act = true;
    const int64_t na = std::max<int64_t>(c->acts.cap(), 1);
    for (int k = 0; k < 2; ++k) c->ract[k] = dalloc<double>(c, na);
    drop_graphs(c);
}

// Launches per step and the clock advance of one launch (RK2: two stages of
// dt / 2, src/sim.cu:1572-1582).
int64_t launches_per_step(const titan_ctx *c) { return c->integ == kIntRK2 ? 2 : 1; }
double launch_dt(const titan_ctx *c) { return c->integ == kIntRK2 ? 0.5 * c->dt : c->dt; }

// LDS rows of the largest live tile
int32_t tiled_rows(const titan_ctx *c) {
    int32_t h = 0;
    for (auto &b : c->bodies)
        if (b.live && b.tiled) h = std::max(h, b.rows);
    return h;
}

// Whether the resident launch has at most one body per compute unit.  Then
// each workgroup has a CU to itself and its step latency is what counts:
// bodies of at most kResidentMax masses step in LDS, in the split form where
// they have it (it spreads a small body's spring terms over all four waves).
// With more bodies every body steps in the arena, so that no workgroup
// holds LDS: LDS sized for the largest body (and a split-form region) would
// cost workgroups per CU.
bool resident_few(const titan_ctx *c) {
    int64_t n = 0;
    for (auto &b : c->bodies)
        if (b.live && b.resident) ++n;
    return n <= c->n_cu;
}

// LDS rows per buffer of the resident kernel: the largest live resident body
// of at most kResidentMax masses when the launch has few bodies, else none
int32_t resident_rows(const titan_ctx *c) {
    if (!resident_few(c)) return 0;
    int32_t h = 0;
    for (auto &b : c->bodies)
        if (b.live && b.resident && b.n_masses <= kResidentMax) h = std::max(h, (int32_t)b.n_masses);
    return h;
}

// The resident launch's LDS region after the state buffers: the largest
// region a live resident body in LDS uses that fits beside them (its
// split-form region when it has one, else its incidence copy).  A body
// whose region is larger reads the arena.  None when the launch has many
// bodies.
int32_t resident_cache(const titan_ctx *c, bool few) {
    if (!few) return 0;
    const int32_t rows = resident_rows(c);
    const int64_t avail = (int64_t)c->res_lds_max - (int64_t)resident_lds_bytes(rows);
    int32_t h = 0;
    for (auto &b : c->bodies) {
        if (!(b.live && b.resident) || b.n_masses > rows) continue;
        int32_t r = 0;
        if (b.res_scache > 0 && b.res_scache <= avail) r = b.res_scache;
        else if (b.res_cache > 0 && b.res_cache <= avail) r = b.res_cache;
        h = std::max(h, r);
    }
    return h;
}

uint8_t mode_of(const BodyH &b) { return b.resident ? kModeResident : b.tiled ? kModeTiled : kModeGather; }

struct Range;
// Whether every resident body a launch covers steps in LDS in the split form
// (then the launch uses k_rollout_split).
int32_t all_split(const titan_ctx *c, const Range &rg, const StepArgs &a);

bool any_body(const titan_ctx *c, uint8_t mode) {
    for (auto &b : c->bodies)
        if (b.live && mode_of(b) == mode) return true;
    return false;
}

bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

int kind_rank(int kind) {
    // application order of local constraints (src/sim.cu:1118-1132)
    switch (kind) {
        case TITAN_CONTACT_PLANE: return 0;
        case TITAN_BALL: return 1;
        case TITAN_CONSTRAINT_PLANE: return 2;
        default: return 3;
    }
}

// The chunks, tiles and resident descriptors one stepping call covers, and
// which kernels it needs.
struct Range {
    int64_t c0 = 0, c1 = 0, t0 = 0, t1 = 0, r0 = 0, r1 = 0;
    bool gather = false, tiled = false, resident = false, all = true;
};

// Every live chunk and tile (up to the arenas' high-water marks; free ones
// are marked and their blocks exit at once).
Range full_range(const titan_ctx *c) {
    Range r;
    r.c1 = c->chunks.high_water();
    r.t1 = c->tiles.high_water();
    r.r1 = c->rdescs.high_water();
    r.gather = any_body(c, kModeGather);
    r.tiled = any_body(c, kModeTiled);
    r.resident = any_body(c, kModeResident);
    return r;
}

Range body_range(const BodyH &b) {
    Range r;
    r.c0 = b.chunk0;
    r.c1 = b.chunk0 + b.nchunks;
    r.t0 = b.tile0;
    r.t1 = b.tile0 + b.ntiles;
    r.r0 = b.resident ? b.res : 0;
    r.r1 = b.resident ? b.res + 1 : 0;
    r.gather = mode_of(b) == kModeGather;
    r.tiled = mode_of(b) == kModeTiled;
    r.resident = b.resident;
    r.all = false;
    return r;
}

int32_t all_split(const titan_ctx *c, const Range &rg, const StepArgs &a) {
    // the split launch copies the planes into its static LDS, which its
    // dynamic region must leave room for
    if (a.nplanes > kSplitPlanes ||
        resident_lds_bytes(a.res_rows) + (int64_t)a.res_cache > (int64_t)c->split_lds_max)
        return 1;
    bool any = false;
    // the lean mass pass: Euler, no balls, no extern force, drag or local
    // constraints on any mass of the launch's bodies
    bool lean = a.integ == kIntEuler && a.nballs == 0;
    for (const auto &b : c->bodies) {
        if (!(b.live && b.resident) || b.res < rg.r0 || b.res >= rg.r1) continue;
        if (!(b.res_scache > 0 && b.res_scache <= a.res_cache && b.n_masses <= a.res_rows)) return 1;
        any = true;
        if (lean && b.lc_len > 0) lean = false;
        for (size_t q = 0; lean && q < b.flags.size(); ++q)
            if (b.flags[q] & (kExt | kDrag | kLocalLC)) lean = false;
    }
    return any ? (lean ? 3 : 2) : 1;
}

StepArgs make_args(titan_ctx *c, int slot, const Range &rg) {
    StepArgs a{};
    a.st = c->st;
    a.ncap = c->ncap;
    a.m = c->m;
    a.mflags = c->mflags;
    a.deg = c->deg;
    a.ext = c->ext;
    a.drag = c->drag;
    a.mlc = c->mlc;
    a.acc = c->acc;
    a.chunk_body = c->chunk_body;
    a.chunk_ptr = c->chunk_ptr;
    a.chunk_ng = c->chunk_ng;
    a.chunk_base = rg.c0;
    a.nchunks = rg.c1;
    a.inc_nbr = c->inc_nbr;
    a.inc_rest = c->inc_rest;
    a.inc_k = c->inc_k;
    a.inc_damp = c->inc_damp;
    a.b_k = c->b_k;
    a.b_m = c->b_m;
    a.b_d = c->b_d;
    a.b_kd = c->b_kd;
    a.act = c->act;
    a.tact = c->tact;
    a.integ = c->integ;
    a.hdt = launch_dt(c);
    a.rk = c->in_noalloc ? c->na_rk : c->rk;
    a.ract[0] = c->in_noalloc ? c->na_ract[0] : c->ract[0];
    a.ract[1] = c->in_noalloc ? c->na_ract[1] : c->ract[1];
    a.chunk_act0 = c->chunk_act0;
    a.chunk_nact = c->chunk_nact;
    for (int s = 0; s < 2; ++s) {
        a.bT[s] = c->bT[s];
        a.bn[s] = c->bn[s];
        a.bleft[s] = c->bleft[s];
    }
    a.b_chunk0 = c->b_chunk0;
    a.b_lc = c->b_lc;
    a.b_lcn = c->b_lcn;
    a.lc = c->lc;
    a.planes = c->planes;
    a.nplanes = (int32_t)(c->planes_h.size() / 8);
    a.balls = c->balls;
    a.nballs = (int32_t)(c->balls_h.size() / 4);
    // a ball grid not rebuilt since the balls changed (possible only in
    // titan_step_noalloc, which builds nothing): every mass tests every ball,
    // with the same result
    const bool grid = !c->ball_grid_dirty;
    a.bg_start = grid ? c->bg_start : nullptr;
    a.bg_rec = grid ? c->bg_rec : nullptr;
    a.bg_x0 = c->bg_x0;
    a.bg_y0 = c->bg_y0;
    a.bg_z0 = c->bg_z0;
    a.bg_inv_h = c->bg_inv_h;
    a.bg_nx = c->bg_n[0];
    a.bg_ny = c->bg_n[1];
    a.bg_nz = c->bg_n[2];
    a.gx = c->g[0];
    a.gy = c->g[1];
    a.gz = c->g[2];
    a.dt = c->dt;
    a.slot = slot;
    a.tiles = c->tile;
    a.tile_base = rg.t0;
    a.ntiles = rg.t1;
    a.halo = c->halo;
    a.ldim = tiled_rows(c);
    a.b_mode = c->b_mode;
    a.rdesc = c->rdesc;
    a.res_base = rg.r0;
    a.nres = rg.r1;
    a.res_rows = resident_rows(c);
    a.res_split = resident_few(c) ? 1 : 0;
    a.res_cache = resident_cache(c, a.res_split != 0);
    if (a.res_split) a.res_split = all_split(c, rg, a);
    return a;
}

// One step of the bodies in a range: the gather kernel over gather-layout
// bodies and the tiled kernel over tiled bodies (disjoint; both read launch
// slot `slot` and write slot ^ 1).
hipError_t launch_range(titan_ctx *c, int slot, const Range &rg) {
    StepArgs a = make_args(c, slot, rg);
    const bool act = c->has_act;
    if (rg.gather) {
        hipError_t e = launch_step(a, act, c->stream);
        if (e != hipSuccess) return e;
    }
    if (rg.tiled) return launch_step_tiles(a, act, c->stream);
    return hipSuccess;
}

hipGraphExec_t graph_for(titan_ctx *c, int slot) {
    const Range rg = full_range(c);
    const int32_t ldim = tiled_rows(c);
    for (auto &g : c->graphs)
        if (g.version == c->version && g.slot == slot && g.c1 == rg.c1 && g.t1 == rg.t1 && g.ldim == ldim &&
            g.gather == rg.gather && g.tiled == rg.tiled)
            return g.exec;
    if (c->graphs.size() >= kGraphCache) {  // evict the oldest
        HIPCHK(hipStreamSynchronize(c->stream));
        (void)hipGraphExecDestroy(c->graphs.front().exec);
        c->graphs.erase(c->graphs.begin());
    }
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int s = slot;
    for (int i = 0; i < kGraphSteps; ++i) {
        hipError_t e = launch_range(c, s, rg);
        if (e != hipSuccess) {
            hipGraph_t junk;
            (void)hipStreamEndCapture(c->stream, &junk);
            fail(TITAN_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        }
        s ^= 1;
    }
    HIPCHK(hipStreamEndCapture(c->stream, &graph));
    hipGraphExec_t exec;
    HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    c->graphs.push_back({exec, c->version, slot, rg.c1, rg.t1, ldim, rg.gather, rg.tiled});
    return exec;
}

int64_t live_body_hw(titan_ctx *c) { return c->bids.high_water(); }

// Build the ball grid: cells of side h over the balls' bounding box, each
// listing (ascending) the balls whose sphere, grown by a relative 1e-9, meets
// the cell grown by 1e-6 h.  A point's computed distance to a ball can fall
// within its radius only if the true distance is within r (1 + 1e-15), and
// the cell index computed from a point's coordinates is off by at most a
// rounding error, far below 1e-6 h; so a ball missing from a point's cell
// list contributes a zero vector.  Points outside the grid touch no ball.
void build_ball_grid(titan_ctx *c) {
    c->ball_grid_dirty = false;
    dfree(c, c->bg_start);
    dfree(c, c->bg_rec);
    c->bg_start = nullptr;
    c->bg_rec = nullptr;
    const int64_t nb = (int64_t)c->balls_h.size() / 4;
    if (nb < kBallGridMin) return;
    const double *B = c->balls_h.data();
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY}, rsum = 0;
    for (int64_t b = 0; b < nb; ++b) {
        const double rad = std::fabs(B[4 * b + 3]);
        if (!std::isfinite(rad) || !std::isfinite(B[4 * b]) || !std::isfinite(B[4 * b + 1]) ||
            !std::isfinite(B[4 * b + 2]))
            return;  // a non-finite ball: test every ball, as the reference does
        rsum += rad;
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], B[4 * b + k] - rad);
            hi[k] = std::max(hi[k], B[4 * b + k] + rad);
        }
    }
    double ext = 0;
    for (int k = 0; k < 3; ++k) ext = std::max(ext, hi[k] - lo[k]);
    if (!(ext > 0)) return;
    for (int k = 0; k < 3; ++k) {  // grow the box so that rounding never leaves it
        const double g = 1e-6 * ext + 1e-300;
        lo[k] -= g;
        hi[k] += g;
    }
    const double vol = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    double h = std::max(2.0 * rsum / (double)nb, std::cbrt(vol / (double)nb));
    int64_t n[3];
    for (;;) {
        for (int k = 0; k < 3; ++k) n[k] = std::max<int64_t>(1, (int64_t)std::ceil((hi[k] - lo[k]) / h));
        if (n[0] * n[1] * n[2] <= kBallGridMaxCells) break;
        h *= 1.26;
    }
    const int64_t ncell = n[0] * n[1] * n[2];
    const double delta = 1e-6 * h;
    std::vector<std::vector<int32_t>> cells(ncell);
    for (int64_t b = 0; b < nb; ++b) {
        const double cx[3] = {B[4 * b], B[4 * b + 1], B[4 * b + 2]};
        const double R = std::fabs(B[4 * b + 3]) * (1.0 + 1e-9) + delta;
        int64_t i0[3], i1[3];
        for (int k = 0; k < 3; ++k) {
            i0[k] = std::max<int64_t>(0, (int64_t)std::floor((cx[k] - R - lo[k]) / h) - 1);
            i1[k] = std::min<int64_t>(n[k] - 1, (int64_t)std::floor((cx[k] + R - lo[k]) / h) + 1);
        }
        for (int64_t iz = i0[2]; iz <= i1[2]; ++iz)
            for (int64_t iy = i0[1]; iy <= i1[1]; ++iy)
                for (int64_t ix = i0[0]; ix <= i1[0]; ++ix) {
                    const int64_t id[3] = {ix, iy, iz};
                    double d2 = 0;  // squared distance from the centre to the grown cell
                    for (int k = 0; k < 3; ++k) {
                        const double a0 = lo[k] + (double)id[k] * h - delta, a1 = lo[k] + (double)(id[k] + 1) * h + delta;
                        const double e = cx[k] < a0 ? a0 - cx[k] : (cx[k] > a1 ? cx[k] - a1 : 0.0);
                        d2 += e * e;
                    }
                    if (d2 <= R * R) cells[(iz * n[1] + iy) * n[0] + ix].push_back((int32_t)b);
                }
    }
    std::vector<int32_t> start(ncell + 1, 0), list;
    for (int64_t q = 0; q < ncell; ++q) {
        start[q + 1] = start[q] + (int32_t)cells[q].size();
        list.insert(list.end(), cells[q].begin(), cells[q].end());
    }
    c->bg_start = dalloc<int32_t>(c, ncell + 1);
    // the listed balls' records, and a spare one (zeros) that the walk over
    // a list may load one past its end
    std::vector<double> rec(4 * (list.size() + 1), 0.0);
    for (size_t q = 0; q < list.size(); ++q)
        for (int k = 0; k < 4; ++k) rec[4 * q + k] = B[4 * (size_t)list[q] + k];
    c->bg_rec = dalloc<double>(c, (int64_t)rec.size());
    h2d(c, c->bg_start, start.data(), ncell + 1);
    h2d(c, c->bg_rec, rec.data(), (int64_t)rec.size());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->bg_x0 = lo[0];
    c->bg_y0 = lo[1];
    c->bg_z0 = lo[2];
    c->bg_inv_h = 1.0 / h;
    for (int k = 0; k < 3; ++k) c->bg_n[k] = (int32_t)n[k];
}

void prepare_launches(titan_ctx *c) {
    if (c->ball_grid_dirty) {
        build_ball_grid(c);
        drop_graphs(c);
    }
}

void maybe_inject(titan_ctx *c, int site) {
    int32_t &k = c->inject[site];
    if (k >= 0) {
        if (k == 0) {
            k = -1;
            static const char *what[3] = {"step", "body_download", "body_destroy"};
            fail(TITAN_E_INJECTED, std::string(what[site]) + ": injected failure");
        }
        k--;
    }
}

// Run n launches over range rg; every body's steps-left for this call are
// already in bleft[c->slot].  Bodies whose budget ends early stop there
// (and store their acceleration on their last step).  A range of one body
// leaves the other bodies' launch-slot fields alone, so when n is odd they
// are first copied to the slot the call ends on.
void run_launches(titan_ctx *c, int64_t n, const Range &rg) {
    const int64_t nb = live_body_hw(c);
    int s = c->slot;
    if (!rg.all && (n & 1)) {
        HIPCHK(hipMemcpyAsync(c->bT[s ^ 1], c->bT[s], sizeof(double) * nb, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->bn[s ^ 1], c->bn[s], sizeof(int64_t) * nb, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->bleft[s ^ 1], c->bleft[s], sizeof(int64_t) * nb, hipMemcpyDeviceToDevice, c->stream));
    }
    if (rg.resident) {  // every step of the resident bodies, in one launch
        const StepArgs a = make_args(c, s, rg);
        HIPCHK(launch_rollout_resident(a, c->has_act, s ^ (int)(n & 1), c->stream));
    }
    int64_t rest = (rg.gather || rg.tiled) ? n : 0;
    if (rest > 0 && c->has_act) {  // the first step's actuation factors
        const StepArgs a = make_args(c, s, rg);
        HIPCHK(launch_actuate(a, c->stream));
    }
    const int64_t blocks = (rg.gather ? (rg.c1 - rg.c0 + kWavesPerBlock - 1) / kWavesPerBlock : 0) +
                           (rg.tiled ? rg.t1 - rg.t0 : 0);
    if (rg.all && rest >= kGraphSteps && c->use_graphs && !c->in_noalloc && blocks < kGraphMaxBlocks) {
        hipGraphExec_t g = graph_for(c, s);  // an even number of steps: ends on slot s
        for (int64_t r = 0; r < rest / kGraphSteps; ++r) HIPCHK(hipGraphLaunch(g, c->stream));
        rest %= kGraphSteps;
    }
    for (int64_t r = 0; r < rest; ++r) {
        HIPCHK(launch_range(c, s, rg));
        s ^= 1;
    }
    s = c->slot ^ (int)(n & 1);
    HIPCHK(launch_advance_T(c->Tm, c->mflags, c->chunk_body, c->bn[s], c->bnsync, rg.c0, rg.c1, nb, launch_dt(c),
                            c->stream));
    c->slot = s;
}

// For the length of a titan_step_noalloc call: the step kernels take their
// scratch buffers (actuation factors, RK2 stage state) from the device's
// static scratch array, and no graph is captured.
struct NoallocScope {
    titan_ctx *c;
    NoallocScope(titan_ctx *ctx, bool on) : c(on ? ctx : nullptr) {
        if (!c) return;
        const int64_t nact = c->has_act ? c->acts.high_water() : 0;
        const int64_t nrk = c->integ == kIntRK2 ? 6 * c->chunks.high_water() * kChunk : 0;
        const int64_t need = 8 * (2 * nact + nrk);
        if (need > kScratchBytes)
            fail(TITAN_E_INVALID, "step_noalloc: the context needs " + std::to_string(need) +
                                      " bytes of scratch (8 B x 2 per actuator slot, and 48 B per mass slot "
                                      "under RK2), more than the " + std::to_string(kScratchBytes) +
                                      " the library reserves per device");
        c->na_ract[0] = c->scratch;
        c->na_ract[1] = c->scratch + nact;
        c->na_rk = c->scratch + 2 * nact;
        c->in_noalloc = true;
    }
    ~NoallocScope() {
        if (!c) return;
        c->in_noalloc = false;
        c->na_ract[0] = c->na_ract[1] = c->na_rk = nullptr;
    }
};

// noalloc: titan_step_noalloc (nothing allocated: no ball grid rebuilt, no
// graph captured, scratch from the static array).
void do_step(titan_ctx *c, int64_t n, bool noalloc = false) {
    if (n < 0) fail(TITAN_E_INVALID, "step: negative step count");
    maybe_inject(c, TITAN_INJECT_STEP);
    if (n == 0) return;
    const int64_t nl = n * launches_per_step(c);
    const double h = launch_dt(c);
    const int64_t nb = live_body_hw(c);
    if (nb > 0 && c->chunks.high_water() > 0) {
        const NoallocScope na(c, noalloc);
        if (!noalloc) prepare_launches(c);
        HIPCHK(launch_set_left(c->bleft[c->slot], c->b_chunk0, nb, nl, c->stream));
        run_launches(c, nl, full_range(c));
    }
    for (auto &b : c->bodies) {
        if (!b.live) continue;
        for (int64_t r = 0; r < nl; ++r) b.T += h;
        b.n += n;
    }
    for (int64_t r = 0; r < nl; ++r) c->T += h;  // src/sim.cu:1590 (RK2: :1577,1582)
}

BodyH &body_of(titan_ctx *c, int32_t id);

// Advance each listed body by its own number of steps; the others do not
// move.  One body alone is launched over its own tiles or chunks only.
void do_step_bodies(titan_ctx *c, const int32_t *ids, const int64_t *steps, int32_t count) {
    if (count < 0 || (count > 0 && (!ids || !steps))) fail(TITAN_E_INVALID, "step_bodies: bad arguments");
    std::vector<int64_t> left(live_body_hw(c), 0);
    int64_t n = 0;
    for (int32_t k = 0; k < count; ++k) {
        body_of(c, ids[k]);  // validates
        if (steps[k] < 0) fail(TITAN_E_INVALID, "step_bodies: negative step count");
        if (left[ids[k]] != 0) fail(TITAN_E_INVALID, "step_bodies: a body is listed twice");
        left[ids[k]] = steps[k] * launches_per_step(c);
        n = std::max(n, left[ids[k]]);
    }
    maybe_inject(c, TITAN_INJECT_STEP);
    if (n == 0) return;
    prepare_launches(c);
    if (count == 1) {
        // the budget travels as a kernel argument: nothing to keep alive
        HIPCHK(launch_set_i64(c->bleft[c->slot] + ids[0], left[ids[0]], c->stream));
        run_launches(c, n, body_range(c->bodies[ids[0]]));
    } else {
        // budgets go through one of two pinned buffers, so the call returns
        // once the steps are enqueued; a buffer is rewritten only after the
        // copy that last read it has run
        const int64_t nb = (int64_t)left.size();
        if (nb > c->left_cap) {
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int k = 0; k < 2; ++k) {
                if (c->left_pin[k]) (void)hipHostFree(c->left_pin[k]);
                c->left_pin[k] = nullptr;
            }
            const int64_t cap = std::max<int64_t>(nb, 2 * c->left_cap);
            for (int k = 0; k < 2; ++k)
                HIPCHK(hipHostMalloc((void **)&c->left_pin[k], sizeof(int64_t) * cap, hipHostMallocDefault));
            c->left_cap = cap;
        }
        const int k = c->left_k;
        c->left_k ^= 1;
        HIPCHK(hipEventSynchronize(c->left_ev[k]));
        std::copy(left.begin(), left.end(), c->left_pin[k]);
        HIPCHK(hipMemcpyAsync(c->bleft[c->slot], c->left_pin[k], sizeof(int64_t) * nb, hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipEventRecord(c->left_ev[k], c->stream));
        run_launches(c, n, full_range(c));
    }
    const double h = launch_dt(c);
    for (int32_t k = 0; k < count; ++k) {
        BodyH &b = c->bodies[ids[k]];
        for (int64_t r = 0; r < left[ids[k]]; ++r) b.T += h;
        b.n += steps[k];
    }
}

BodyH &body_of(titan_ctx *c, int32_t id) {
    if (id < 0 || id >= (int64_t)c->bodies.size() || !c->bodies[id].live)
        fail(TITAN_E_NOTFOUND, "unknown body id " + std::to_string(id));
    return c->bodies[id];
}

// A body's layout and arrays built on the host, before any device work.
struct PreparedBody {
    int64_t N = 0, S = 0, nch = 0, npad = 0, n_act = 0, n_lc = 0;
    bool per_k = false, per_d = false, any_lc = false, shared = false, tiled = false, resident = false;
    bool any_ext = false, any_drag = false;
    double k0 = 0, d0 = 0, t0 = 0;
    bool m_same = false;  // every mass of the body has m = m0
    double m0 = 0;
    std::vector<double> acth;       // 4 per actuator: a, omega, phi, rest
    std::vector<int32_t> perm;      // device order -> caller's order (empty = identity)
    TileBuild tb;                   // tiled layout, body-relative
    SellBuild sb;                   // gather names body-relative (arena slot = chunk0 * 64 + name)
    std::vector<double> st, mm, Tm, ext;
    std::vector<uint32_t> fl;
    std::vector<int32_t> drag;
    std::vector<LocalC> lch;        // the shared list, or every per-mass list in device order
    std::vector<int64_t> mlc;       // per-mass lists: (offset in lch << 20) | count
    // split form (resident bodies of at most kSplitMaxChunks chunks): the
    // springs' endpoint words then their incidence words, the rest lengths
    // (then k and damping where per spring), and each spring's local
    // actuator or -1
    std::vector<uint32_t> spw;
    std::vector<double> spd;
    std::vector<int32_t> spa;
    int64_t nsp = 0, nslots = 0;
    std::vector<int32_t> sl, sr;    // spring endpoints in device order (-1: none)
    // built on the device at commit (DevBuilder): the caller's descriptor
    // (read during the commit), the schedule it was prepared under, each
    // spring's body-local actuator (empty: none)
    bool device = false;
    const titan_body_desc *desc = nullptr;
    titan_schedule sched{};
    std::vector<int32_t> act_local;
};

// The split form of a resident body (endpoints in device rows): its springs
// once each and the slots of their two incidences' terms (incidence j of
// row q: j * N + q, each mass's incidences in the order its SELL entries
// hold them; the LDS region has the largest degree rounded up to a multiple
// of 4 slots per row); nothing when the slots exceed 16 bits or its LDS
// region would not fit.
void build_split(PreparedBody &P, const SpringsIn &sp, const std::vector<int32_t> &act_local) {
    const int64_t S = sp.n_springs, N = P.N;
    int64_t nsp = 0, nreal = 0, maxdeg = 0;
    for (int64_t s = 0; s < S; ++s)
        if (sp.left[s] >= 0 && sp.right[s] >= 0) ++nsp;
    for (const int32_t dg : P.sb.deg) {
        nreal += dg;
        maxdeg = std::max<int64_t>(maxdeg, dg);
    }
    const int64_t nslots = ((maxdeg + 3) & ~(int64_t)3) * N;  // the kernel sums whole fours of slots
    if (nreal != 2 * nsp || nslots > 0x10000 || split_region_bytes(nslots, nsp, P.per_k, P.per_d) > kResidentCacheMax)
        return;
    std::vector<int64_t> at(N, 0);  // the next incidence of each mass
    P.spw.assign(2 * nsp, 0);
    P.spd.assign(nsp * (1 + (P.per_k ? 1 : 0) + (P.per_d ? 1 : 0)), 0.0);
    if (sp.act) P.spa.assign(nsp, -1);
    int64_t j = 0;
    for (int64_t s = 0; s < S; ++s) {
        const int32_t a = sp.left[s], b = sp.right[s];
        if (a < 0 || b < 0) continue;
        const bool act = sp.act && act_local[s] >= 0;
        P.spw[j] = (uint32_t)a | ((uint32_t)b << 16) | (act ? kActBit16 << 16 : 0u);
        P.spd[j] = sp.rest[s];
        if (P.per_k) P.spd[nsp + j] = sp.k[s];
        if (P.per_d) P.spd[nsp * (P.per_k ? 2 : 1) + j] = sp.damping ? sp.damping[s] : 0.0;
        if (act) P.spa[j] = act_local[s];
        const uint32_t sa = (uint32_t)(at[a]++ * N + a);  // as build_sell places them: a, then b
        const uint32_t sb = (uint32_t)(at[b]++ * N + b);
        P.spw[nsp + j] = sa | (sb << 16);
        ++j;
    }
    P.nsp = nsp;
    P.nslots = nslots;
}

// device_min: a tiled body of at least this many masses is laid out on the
// device at commit (0: never)
PreparedBody prepare_body(const titan_schedule &sched, const titan_body_desc *d, int64_t device_min) {
    if (!d) fail(TITAN_E_INVALID, "body_create: null descriptor");
    PreparedBody P;
    const int64_t N = d->n_masses, S = d->n_springs;
    if (N <= 0) fail(TITAN_E_INVALID, "body_create: a body needs at least one mass");
    if (N > kMaxMassSlots) fail(TITAN_E_INVALID, "body_create: too many masses in one body");
    if (S < 0 || S > (int64_t)0x7fffffff) fail(TITAN_E_INVALID, "body_create: bad spring count");
    if (!d->pos || !d->m) fail(TITAN_E_INVALID, "body_create: pos and m are required");
    if (S > 0 && (!d->left || !d->right || !d->k || !d->rest))
        fail(TITAN_E_INVALID, "body_create: springs need left, right, k and rest");
    if (d->layout < TITAN_LAYOUT_AUTO || d->layout > TITAN_LAYOUT_RESIDENT)
        fail(TITAN_E_INVALID, "body_create: bad layout");
    const int64_t nch = (N + kChunk - 1) / kChunk;
    const int64_t npad = nch * kChunk;
    P.N = N;
    P.S = S;
    P.nch = nch;
    P.npad = npad;
    P.t0 = d->t0;

    // ---- springs: validation, shared constants, actuators (caller's order)
    int64_t n_act = 0;
    bool k_same = true, d_same = true, first = true;
    double k0 = 0, d0 = 0;
    std::vector<int32_t> act_local(S > 0 ? S : 0, -1);
    std::vector<double> &acth = P.acth;
    for (int64_t s = 0; s < S; ++s) {
        const int64_t L = d->left[s], R = d->right[s];
        if (L >= N || R >= N) fail(TITAN_E_INVALID, "body_create: spring " + std::to_string(s) + " endpoint out of range");
        if (L < 0 || R < 0) continue;
        const int ty = d->type ? d->type[s] : 0;
        if (ty < 0 || ty > TITAN_SPRING_GENERAL) fail(TITAN_E_INVALID, "body_create: bad spring type");
        if (ty >= TITAN_ACTIVE_CONTRACT_THEN_EXPAND) {
            if (!d->omega) fail(TITAN_E_INVALID, "body_create: actuated springs need omega");
            if (ty == TITAN_SPRING_GENERAL && (!d->act_a || !d->act_phi))
                fail(TITAN_E_INVALID, "body_create: general actuation needs act_a and act_phi");
            // the reference's two kinds are the general law with a = -/+0.2 and
            // phi = 0 (src/sim.cu:1058-1062): 1 + (-0.2) s == 1 - 0.2 s and
            // x + 0 == x exactly
            const double aa = ty == TITAN_ACTIVE_CONTRACT_THEN_EXPAND ? -0.2
                              : ty == TITAN_ACTIVE_EXPAND_THEN_CONTRACT ? 0.2 : d->act_a[s];
            const double ph = ty == TITAN_SPRING_GENERAL ? d->act_phi[s] : 0.0;
            act_local[s] = (int32_t)n_act++;
            acth.insert(acth.end(), {aa, d->omega[s], ph, d->rest[s]});
        }
        const double ks = d->k[s], ds = d->damping ? d->damping[s] : 0.0;
        if (first) {
            k0 = ks;
            d0 = ds;
            first = false;
        } else {
            k_same = k_same && same_bits(ks, k0);
            d_same = d_same && same_bits(ds, d0);
        }
    }
    // constants every spring of the body shares travel by value
    const bool per_k = !k_same, per_d = !d_same;
    P.n_act = n_act;
    P.per_k = per_k;
    P.per_d = per_d;
    P.k0 = k0;
    P.d0 = d0;

    // ---- local constraints, sorted stably into application order
    std::vector<std::vector<LocalC>> lists;
    bool any_lc = false, shared = false;
    if (d->lc_ptr) {
        if (!d->lc_type || !d->lc_vec || !d->lc_num) fail(TITAN_E_INVALID, "body_create: incomplete local constraints");
        lists.resize(N);
        for (int64_t i = 0; i < N; ++i) {
            for (int64_t j = d->lc_ptr[i]; j < d->lc_ptr[i + 1]; ++j) {
                LocalC e{};
                e.v[0] = d->lc_vec[3 * j];
                e.v[1] = d->lc_vec[3 * j + 1];
                e.v[2] = d->lc_vec[3 * j + 2];
                e.num = d->lc_num[j];
                e.kind = d->lc_type[j];
                if (e.kind < 0 || e.kind > 3) fail(TITAN_E_INVALID, "body_create: bad local constraint kind");
                lists[i].push_back(e);
            }
            std::stable_sort(lists[i].begin(), lists[i].end(),
                             [](const LocalC &a, const LocalC &b) { return kind_rank(a.kind) < kind_rank(b.kind); });
            if (lists[i].size() > 0xFFFFF) fail(TITAN_E_INVALID, "body_create: too many local constraints on one mass");
            if (!lists[i].empty()) any_lc = true;
        }
        if (any_lc) {
            shared = true;
            for (int64_t i = 1; i < N && shared; ++i) {
                if (lists[i].size() != lists[0].size() ||
                    std::memcmp(lists[i].data(), lists[0].data(), sizeof(LocalC) * lists[0].size()) != 0)
                    shared = false;
            }
        }
    }
    P.any_lc = any_lc;
    P.shared = shared;

    // ---- layout: resident, tiled (masses along a Morton curve) or gather (caller's order)
    std::vector<int32_t> &perm = P.perm;
    TileBuild &tb = P.tb;
    bool tiled = false;
    std::vector<int32_t> nl, nr;  // endpoints in device order
    if (d->layout == TITAN_LAYOUT_RESIDENT && (N < 1 || N > kResidentMaxBody))
        fail(TITAN_E_INVALID, "body_create: the resident layout holds 1 to " + std::to_string(kResidentMaxBody) +
                                  " masses, not " + std::to_string(N));
    const bool resident = d->layout == TITAN_LAYOUT_RESIDENT ||
                          (d->layout == TITAN_LAYOUT_AUTO && N >= 1 && N <= sched.resident_max_masses);
    const bool want_tiles = !resident && (d->layout == TITAN_LAYOUT_TILED ||
                                          (d->layout == TITAN_LAYOUT_AUTO && N >= sched.tiled_min_masses));
    if (want_tiles && N > kTileMasses && device_min > 0 && N >= device_min && !(any_lc && !shared)) {
        // the layout is built on the device at commit (builder.hip); here only
        // what the host keeps: flags summaries, the shared local list
        P.device = true;
        P.desc = d;
        P.sched = sched;
        if (n_act > 0) P.act_local.swap(act_local);
        for (int64_t i = 0; i < N && d->ext && !P.any_ext; ++i)
            P.any_ext = d->ext[3 * i] != 0 || d->ext[3 * i + 1] != 0 || d->ext[3 * i + 2] != 0;
        for (int64_t i = 0; i < N && d->drag && !P.any_drag; ++i) P.any_drag = d->drag[i] != 0;
        P.m0 = d->m[0];
        P.m_same = true;
        for (int64_t i = 1; i < N && P.m_same; ++i) P.m_same = same_bits(d->m[i], P.m0);
        if (any_lc) P.lch = lists[0];
        P.n_lc = (int64_t)P.lch.size();
        P.tiled = true;
        return P;
    }
    if (want_tiles && N > kTileMasses) perm = morton_order(d->pos, N);
    {
        std::vector<int32_t> inv;
        if (!perm.empty()) {
            inv.resize(N);
            for (int64_t q = 0; q < N; ++q) inv[perm[q]] = (int32_t)q;
        }
        nl.resize(S);
        nr.resize(S);
        for (int64_t s = 0; s < S; ++s) {
            const bool ok = d->left[s] >= 0 && d->right[s] >= 0;
            nl[s] = !ok ? -1 : perm.empty() ? (int32_t)d->left[s] : inv[d->left[s]];
            nr[s] = !ok ? -1 : perm.empty() ? (int32_t)d->right[s] : inv[d->right[s]];
        }
    }
    SpringsIn sp;
    sp.n_springs = S;
    sp.left = nl.data();
    sp.right = nr.data();
    sp.k = d->k;
    sp.rest = d->rest;
    sp.damping = d->damping;
    sp.act = n_act > 0 ? act_local.data() : nullptr;
    if (want_tiles) {
        tiled = build_tiles(N, sp, tb, sched.tile_rows);
        if (!tiled && d->layout == TITAN_LAYOUT_TILED)
            fail(TITAN_E_INVALID, "body_create: tiled layout does not fit: " + tb.why);
        if (!tiled && !perm.empty()) {  // fall back to the caller's order
            perm.clear();
            for (int64_t s = 0; s < S; ++s) {
                const bool ok = d->left[s] >= 0 && d->right[s] >= 0;
                nl[s] = ok ? (int32_t)d->left[s] : -1;
                nr[s] = ok ? (int32_t)d->right[s] : -1;
            }
        }
    }
    if (!tiled) tb = TileBuild{};
    P.tiled = tiled;
    P.resident = resident;
    if (tiled || resident) {  // masses of a tile (or resident body) by descending degree
        const std::vector<int32_t> order = degree_sort_tiles(N, sp, tiled ? &tb : nullptr);
        if (!order.empty()) {
            std::vector<int32_t> np(N), inv(N);
            for (int64_t q = 0; q < N; ++q) {
                np[q] = perm.empty() ? order[q] : perm[order[q]];
                inv[order[q]] = (int32_t)q;
            }
            perm.swap(np);
            for (int64_t s = 0; s < S; ++s) {
                if (nl[s] >= 0) nl[s] = inv[nl[s]];
                if (nr[s] >= 0) nr[s] = inv[nr[s]];
            }
        }
    }
    auto orig = [&](int64_t q) -> int64_t { return perm.empty() ? q : perm[q]; };
    // incidences: LDS rows (tiled, resident) or body-relative slots (gather)
    if (tiled) {
        build_sell(N, sp, per_k, per_d, &tb, 0, P.sb);
    } else if (resident) {  // the whole body is one tile with no halo: rows = body indices
        TileBuild one;
        TileDesc t{};
        t.nmass = (int32_t)N;
        t.first = 1;
        one.tiles.push_back(t);
        one.tile_of_chunk.assign(nch, 0);
        one.rmax = (int32_t)N;
        build_sell(N, sp, per_k, per_d, &one, 0, P.sb);
        if (nch <= kSplitMaxChunks) build_split(P, sp, act_local);
    } else {
        build_sell(N, sp, per_k, per_d, nullptr, 0, P.sb);
    }

    // ---- mass arrays in device order
    P.st.assign(6 * npad, 0.0);  // 6 per mass: pos, vel
    P.mm.assign(npad, 1.0);
    P.Tm.assign(npad, 0.0);
    P.fl.assign(npad, kPad);
    for (int64_t q = 0; q < N; ++q) {
        const int64_t i = orig(q);
        for (int k = 0; k < 3; ++k) {
            P.st[6 * q + k] = d->pos[3 * i + k];
            P.st[6 * q + 3 + k] = d->vel ? d->vel[3 * i + k] : 0.0;
        }
        P.mm[q] = d->m[i];
        P.Tm[q] = d->T ? d->T[i] : 0.0;
        uint32_t f = 0;
        if (d->fixed && d->fixed[i]) f |= kFixed;
        if (d->ext && (d->ext[3 * i] != 0 || d->ext[3 * i + 1] != 0 || d->ext[3 * i + 2] != 0)) {
            f |= kExt;
            P.any_ext = true;
        }
        if (d->drag && d->drag[i] != 0) {
            f |= kDrag;
            P.any_drag = true;
        }
        P.fl[q] = f;
    }
    // a body whose masses all share m passes it by value to the tiled step
    P.m0 = d->m[0];
    P.m_same = true;
    for (int64_t i = 1; i < N && P.m_same; ++i) P.m_same = same_bits(d->m[i], P.m0);
    if (P.any_ext) {
        P.ext.assign(3 * npad, 0.0);
        for (int64_t q = 0; q < N; ++q)
            for (int k = 0; k < 3; ++k) P.ext[k * npad + q] = d->ext[3 * orig(q) + k];
    }
    if (P.any_drag) {
        P.drag.assign(npad, 0);
        for (int64_t q = 0; q < N; ++q) P.drag[q] = d->drag[orig(q)];
    }
    if (any_lc) {
        if (shared) {
            P.lch = lists[0];
        } else {
            P.mlc.assign(npad, 0);
            for (int64_t q = 0; q < N; ++q) {
                const auto &l = lists[orig(q)];
                if (l.empty()) continue;
                P.mlc[q] = ((int64_t)P.lch.size() << 20) | (int64_t)l.size();
                P.fl[q] |= kLocalLC;
                P.lch.insert(P.lch.end(), l.begin(), l.end());
            }
        }
        P.n_lc = (int64_t)P.lch.size();
    }
    if (tiled && n_act > 0) P.act_local.swap(act_local);  // the tiles' actuator records come from it
    P.sl.swap(nl);
    P.sr.swap(nr);
    return P;
}

// Index of the body table and arenas the activation kernel writes.
BodyTable body_table(titan_ctx *c) {
    BodyTable t{};
    t.chunk_body = c->chunk_body;
    t.tiles = c->tile;
    t.rdesc = c->rdesc;
    for (int s = 0; s < 2; ++s) {
        t.bT[s] = c->bT[s];
        t.bn[s] = c->bn[s];
        t.bleft[s] = c->bleft[s];
    }
    t.bnsync = c->bnsync;
    t.b_chunk0 = c->b_chunk0;
    t.b_lc = c->b_lc;
    t.b_lcn = c->b_lcn;
    t.b_mode = c->b_mode;
    t.b_k = c->b_k;
    t.b_d = c->b_d;
    t.b_m = c->b_m;
    t.b_kd = c->b_kd;
    return t;
}

// Ranges of a destroyed body, released once the launches enqueued before its
// deactivation have run (until then they may still read its slots).
void harvest_released(titan_ctx *c, bool wait) {
    std::vector<titan_ctx::Pending> keep;
    for (auto &p : c->pending) {
        if (!wait && hipEventQuery(p.ev) == hipErrorNotReady) {
            keep.push_back(p);
            continue;
        }
        if (wait) HIPCHK(hipEventSynchronize(p.ev));
        const BodyH &b = p.b;
        c->chunks.release(b.chunk0, b.nchunks);
        c->incs.release(b.inc0, b.inc_len);
        c->acts.release(b.act0, b.act_len);
        c->lcs.release(b.lc0, b.lc_len);
        c->tiles.release(b.tile0, b.ntiles);
        c->halos.release(b.halo0, b.nhalo);
        c->tacts.release(b.tact0, b.tact_len);
        if (b.resident) c->rdescs.release(b.res, 1);
        c->bids.release(p.id, 1);
        (void)hipEventDestroy(p.ev);
    }
    c->pending.swap(keep);
}

// Arena ranges of one body, and what a body needs reserved.
struct Slots {
    int64_t ch0 = -1, inc0 = -1, act0 = -1, lc0 = -1, tl0 = -1, hl0 = -1, ta0 = -1, rs0 = -1, bid = -1;
};
struct SlotNeeds {
    int64_t nch, n_inc, n_act, n_lc, n_tiles, n_halo, n_tact;
    bool resident;
};

// Reserve a body's arena ranges, growing arenas as needed (slots of
// destroyed bodies are reclaimed first), and create the side arrays it
// needs; on failure release whatever was taken.
Slots reserve_slots(titan_ctx *c, const SlotNeeds &n, const PreparedBody &P) {
    Slots r;
    auto take = [&](Ranges &rg, int64_t len, void (*grow)(titan_ctx *, int64_t)) -> int64_t {
        int64_t at = rg.alloc(len);
        if (at < 0 && !c->pending.empty()) {  // slots of destroyed bodies first
            harvest_released(c, true);
            at = rg.alloc(len);
        }
        if (at < 0) {
            grow(c, rg.high_water() + len);
            at = rg.alloc(len);
        }
        return at;
    };
    try {
        r.ch0 = take(c->chunks, n.nch, ensure_chunks);
        if ((r.ch0 + n.nch) * kChunk > kMaxMassSlots)
            fail(TITAN_E_NOMEM, "body_create: mass arena limit (2^26 slots) reached");
        r.inc0 = take(c->incs, n.n_inc, ensure_incs);
        r.act0 = take(c->acts, n.n_act, ensure_acts);
        r.lc0 = take(c->lcs, n.n_lc, ensure_lcs);
        r.tl0 = take(c->tiles, n.n_tiles, ensure_tiles);
        r.hl0 = take(c->halos, n.n_halo, ensure_halos);
        r.ta0 = take(c->tacts, n.n_tact, ensure_tacts);
        if (n.resident) r.rs0 = take(c->rdescs, 1, ensure_rdescs);
        r.bid = take(c->bids, 1, ensure_bodies);
        if (P.per_k) make_inc_array(c, c->inc_k);
        if (P.per_d) make_inc_array(c, c->inc_damp);
        if (n.n_act > 0) make_act(c);
        // side arrays created on first use, zeroed before any upload into them
        bool made = false;
        if (P.any_ext && !c->ext) {
            c->ext = dalloc<double>(c, 3 * c->ncap);
            HIPCHK(hipMemsetAsync(c->ext, 0, sizeof(double) * 3 * c->ncap, c->stream));
            made = true;
        }
        if (P.any_drag && !c->drag) {
            c->drag = dalloc<int32_t>(c, c->ncap);
            HIPCHK(hipMemsetAsync(c->drag, 0, sizeof(int32_t) * c->ncap, c->stream));
            made = true;
        }
        if (P.any_lc && !P.shared && !c->mlc) {
            c->mlc = dalloc<int64_t>(c, c->ncap);
            HIPCHK(hipMemsetAsync(c->mlc, 0, sizeof(int64_t) * c->ncap, c->stream));
            made = true;
        }
        if (made) {
            HIPCHK(hipStreamSynchronize(c->stream));
            drop_graphs(c);
        }
    } catch (...) {
        if (r.ch0 >= 0) c->chunks.release(r.ch0, n.nch);
        if (r.inc0 >= 0) c->incs.release(r.inc0, n.n_inc);
        if (r.act0 >= 0) c->acts.release(r.act0, n.n_act);
        if (r.lc0 >= 0) c->lcs.release(r.lc0, n.n_lc);
        if (r.tl0 >= 0) c->tiles.release(r.tl0, n.n_tiles);
        if (r.hl0 >= 0) c->halos.release(r.hl0, n.n_halo);
        if (r.ta0 >= 0) c->tacts.release(r.ta0, n.n_tact);
        if (r.rs0 >= 0) c->rdescs.release(r.rs0, 1);
        if (r.bid >= 0) c->bids.release(r.bid, 1);
        throw;
    }
    return r;
}

// Place a prepared body on the device.  Its arrays are copied on the upload
// stream into free arena space, which the launches in flight do not read;
// the body then joins the launches through a small kernel on the step
// stream, after the launches enqueued before it.  Nothing here waits for
// the step stream (only growing an arena does).
int32_t commit_device_built(titan_ctx *c, PreparedBody &P);

int32_t commit_body(titan_ctx *c, PreparedBody &P) {
    if (P.device) {
        const int32_t id = commit_device_built(c, P);
        if (id >= 0) return id;
        P = prepare_body(P.sched, P.desc, 0);  // a tile outgrew the row budget: the host's greedy decides
    }
    harvest_released(c, false);
    const int64_t N = P.N, nch = P.nch, npad = P.npad, n_act = P.n_act, n_lc = P.n_lc;
    const bool tiled = P.tiled, resident = P.resident;
    SellBuild &sb = P.sb;
    TileBuild &tb = P.tb;
    const int64_t n_tiles = (int64_t)tb.tiles.size(), n_halo = (int64_t)tb.halo.size(),
                  n_tact = (int64_t)tb.tact.size();
    const int64_t n_inc = sb.coff[nch];
    // a resident body's split-form data follows its SELL entries
    const int64_t n_extra = std::max<int64_t>((int64_t)P.spw.size(), (int64_t)P.spd.size());

    // ---- reserve space (released again on failure)
    const Slots sl = reserve_slots(c, {nch, n_inc + n_extra, n_act, n_lc, n_tiles, n_halo, n_tact, resident}, P);
    const int64_t ch0 = sl.ch0, inc0 = sl.inc0, act0 = sl.act0, lc0 = sl.lc0, tl0 = sl.tl0, hl0 = sl.hl0,
                  ta0 = sl.ta0, rs0 = sl.rs0, bid = sl.bid;
    const int64_t g0 = ch0 * kChunk;

    // ---- arena addresses: gather slots, actuated incidences (a tiled
    // body's name their tile's list positions), tiles, local lists
    if (!tiled && !resident)
        for (auto &x : sb.nbr) x += (uint32_t)g0;
    mark_actuated(sb, tiled || resident, tiled ? 0 : act0);
    for (auto &t : tb.tiles) {
        t.body = -1;  // set by the activation kernel
        t.chunk0 += (int32_t)ch0;
        t.halo_off += hl0;
        t.act_off += ta0;
    }
    // the tiles' copies of their actuators' records
    std::vector<double> trec(4 * (size_t)n_tact);
    for (int64_t x = 0; x < n_tact; ++x) {
        const int32_t a = P.act_local[tb.tact[x]];
        for (int k = 0; k < 4; ++k) trec[4 * x + k] = P.acth[4 * (size_t)a + k];
    }
    for (auto &h : tb.halo) h += (int32_t)g0;
    for (auto &v : P.mlc)
        if (v) v += lc0 << 20;
    std::vector<int64_t> cptr(nch), ca0(nch, 0);
    std::vector<int32_t> cng(nch), cna(nch, 0);
    {  // each chunk takes an equal share of the body's actuators (a tiled
       // body's tiles compute theirs: no share)
        const int64_t per = nch > 0 && !tiled ? (n_act + nch - 1) / nch : 0;
        for (int64_t ch = 0; ch < nch; ++ch) {
            ca0[ch] = act0 + ch * per;
            cna[ch] = (int32_t)std::max<int64_t>(0, std::min(per, n_act - ch * per));
        }
    }
    const int64_t gsz = (int64_t)(tiled || resident ? kGroupTiled : kGroupGather) * kChunk;
    for (int64_t ch = 0; ch < nch; ++ch) {
        cptr[ch] = inc0 + sb.coff[ch];
        cng[ch] = (int32_t)((sb.coff[ch + 1] - sb.coff[ch]) / gsz);
    }

    // ---- upload on the upload stream (buffer parity 0: a new body has taken no steps)
    auto up = [&](auto *dst, const auto *src, int64_t n) {
        if (n > 0) HIPCHK(hipMemcpyAsync(dst, src, sizeof(*src) * n, hipMemcpyHostToDevice, c->up));
    };
    up(c->st + 6 * g0, P.st.data(), 6 * npad);
    up(c->m + g0, P.mm.data(), npad);
    up(c->Tm + g0, P.Tm.data(), npad);
    up(c->mflags + g0, P.fl.data(), npad);
    up(c->deg + g0, sb.deg.data(), npad);
    std::vector<double> zeros(3 * npad, 0.0);
    for (int k = 0; k < 3; ++k) up(c->acc + k * c->ncap + g0, zeros.data(), npad);
    if (c->ext)
        for (int k = 0; k < 3; ++k)
            up(c->ext + k * c->ncap + g0, P.any_ext ? P.ext.data() + k * npad : zeros.data(), npad);
    std::vector<int32_t> zi(npad, 0);
    std::vector<int64_t> zl(npad, 0);
    if (c->drag) up(c->drag + g0, P.any_drag ? P.drag.data() : zi.data(), npad);
    if (c->mlc) up(c->mlc + g0, P.mlc.empty() ? zl.data() : P.mlc.data(), npad);
    up(c->chunk_ptr + ch0, cptr.data(), nch);
    up(c->chunk_ng + ch0, cng.data(), nch);
    up(c->chunk_act0 + ch0, ca0.data(), nch);
    up(c->chunk_nact + ch0, cna.data(), nch);
    if (n_inc > 0) {
        up(c->inc_nbr + inc0, sb.nbr.data(), n_inc);
        up(c->inc_rest + inc0, sb.rest.data(), n_inc);
        if (P.per_k) up(c->inc_k + inc0, sb.k.data(), n_inc);
        if (P.per_d) up(c->inc_damp + inc0, sb.damp.data(), n_inc);
    }
    if (!P.spw.empty()) {
        for (size_t j = 0; j < P.spa.size(); ++j) {  // actuated: the actuator's arena index
            if (P.spa[j] < 0) continue;
            const int64_t ix = act0 + P.spa[j];
            std::memcpy(&P.spd[j], &ix, sizeof ix);
        }
        up(c->inc_nbr + inc0 + n_inc, P.spw.data(), (int64_t)P.spw.size());
        up(c->inc_rest + inc0 + n_inc, P.spd.data(), (int64_t)P.spd.size());
    }
    up(c->tile + tl0, tb.tiles.data(), n_tiles);
    up(c->halo + hl0, tb.halo.data(), n_halo);
    up(c->tact + 4 * ta0, trec.data(), 4 * n_tact);
    // a resident body's incidence copy: names (4 B per entry, to a 16-byte
    // boundary), rest lengths and per-incidence constants (8 B each)
    int64_t cache = ((4 * n_inc + 15) & ~(int64_t)15) + 8 * n_inc * (1 + (P.per_k ? 1 : 0) + (P.per_d ? 1 : 0));
    if (cache > kResidentCacheMax) cache = 0;
    // and its split-form region, where it has that form
    const int64_t scache = P.spw.empty() ? 0 : split_region_bytes(P.nslots, P.nsp, P.per_k, P.per_d);
    const ResidentDesc rdh{-1, (int32_t)ch0, (int32_t)N, (int32_t)n_act, act0, inc0, (int32_t)n_inc, (int32_t)cache,
                           (int32_t)scache, (int32_t)P.nsp};
    if (resident) up(c->rdesc + rs0, &rdh, 1);
    up(c->act + 4 * act0, P.acth.data(), (int64_t)P.acth.size());
    up(c->lc + lc0, P.lch.data(), n_lc);
    HIPCHK(hipStreamSynchronize(c->up));  // the copies only: the host arrays may go

    // ---- join the launches
    BodyInit bi{};
    bi.bid = (int32_t)bid;
    bi.chunk0 = ch0;
    bi.nchunks = nch;
    bi.tile0 = tl0;
    bi.ntiles = n_tiles;
    bi.res = resident ? rs0 : -1;
    bi.t0 = P.t0;
    bi.k = P.k0;
    bi.m = P.m0;
    bi.d = P.d0;
    bi.lcb = (int32_t)(P.shared ? lc0 : 0);
    bi.lcn = (int32_t)(P.shared ? n_lc : 0);
    bi.mode = resident ? kModeResident : tiled ? kModeTiled : kModeGather;
    bi.kd = (uint8_t)((P.per_k ? 1 : 0) | (P.per_d ? 2 : 0) | (P.m_same ? kKdUniformM : 0));
    bi.on = 1;
    HIPCHK(launch_set_body(body_table(c), bi, c->stream));

    BodyH &b = c->bodies[bid];
    b = BodyH{};
    b.live = true;
    b.tiled = tiled;
    b.resident = resident;
    b.kd = bi.kd;
    b.res = rs0;
    b.res_cache = resident ? (int32_t)cache : 0;
    b.res_scache = resident ? (int32_t)scache : 0;
    b.n_masses = N;
    b.chunk0 = ch0;
    b.nchunks = nch;
    b.inc0 = inc0;
    b.inc_len = n_inc + n_extra;
    b.act0 = act0;
    b.act_len = n_act;
    b.lc0 = lc0;
    b.lc_len = n_lc;
    b.tile0 = tl0;
    b.ntiles = n_tiles;
    b.halo0 = hl0;
    b.nhalo = n_halo;
    b.tact0 = ta0;
    b.tact_len = n_tact;
    b.rows = tb.rmax;
    b.T = P.t0;
    b.n = 0;
    b.flags.assign(P.fl.begin(), P.fl.begin() + N);
    b.perm = std::move(P.perm);
    b.sl = std::move(P.sl);
    b.sr = std::move(P.sr);
    b.coff = sb.coff;
    b.row0.assign(nch, 0);
    if (tiled)
        for (int64_t ch = 0; ch < nch; ++ch)
            b.row0[ch] = (int32_t)((int64_t)tb.tiles[tb.tile_of_chunk[ch]].chunk0 - ch0) * kChunk;
    return (int32_t)bid;
}

// Place a body built on the device (DevBuilder): its layout is computed by
// the builder's first phase, the arena ranges reserved for the sizes it
// found, and its second phase writes the body into them; the body then joins
// the launches as a host-built one does.  Returns -1 (nothing reserved) when
// a tile would not fit the row budget, for the host build to decide.
int32_t commit_device_built(titan_ctx *c, PreparedBody &P) {
    harvest_released(c, false);
    const titan_body_desc *d = P.desc;
    if (!c->builder) c->builder = new DevBuilder;
    DevBuildIn in;
    in.N = P.N;
    in.S = P.S;
    in.pos = d->pos;
    in.vel = d->vel;
    in.m = d->m;
    in.T = d->T;
    in.ext = P.any_ext ? d->ext : nullptr;
    in.fixed = d->fixed;
    in.drag = P.any_drag ? d->drag : nullptr;
    in.left = d->left;
    in.right = d->right;
    in.rest = d->rest;
    in.k = P.per_k ? d->k : nullptr;
    in.damping = P.per_d ? d->damping : nullptr;
    in.act_local = P.act_local.empty() ? nullptr : P.act_local.data();
    in.per_k = P.per_k;
    in.per_d = P.per_d;
    in.any_ext = P.any_ext;
    in.any_drag = P.any_drag;
    in.tile_rows = P.sched.tile_rows;
    DevBuildSizes sz;
    HIPCHK(c->builder->plan(in, c->up, &sz));
    if (!sz.fits) return -1;
    const int64_t N = P.N, nch = sz.nch, n_act = P.n_act, n_lc = P.n_lc;
    const Slots sl = reserve_slots(c, {nch, sz.n_inc, n_act, n_lc, sz.ntiles, sz.n_halo, sz.n_tact, false}, P);
    BodyH b;
    b.perm.resize(N);
    b.flags.resize(N);
    b.sl.resize(P.S);
    b.sr.resize(P.S);
    b.coff.resize(nch + 1);
    DevBuildOut o;
    o.ncap = c->ncap;
    o.st = c->st;
    o.m = c->m;
    o.Tm = c->Tm;
    o.acc = c->acc;
    o.ext = c->ext;
    o.mflags = c->mflags;
    o.deg = c->deg;
    o.drag = c->drag;
    o.mlc = c->mlc;
    o.chunk_ptr = c->chunk_ptr;
    o.chunk_act0 = c->chunk_act0;
    o.chunk_ng = c->chunk_ng;
    o.chunk_nact = c->chunk_nact;
    o.inc_nbr = c->inc_nbr;
    o.inc_rest = c->inc_rest;
    o.inc_k = P.per_k ? c->inc_k : nullptr;
    o.inc_damp = P.per_d ? c->inc_damp : nullptr;
    o.tiles = c->tile;
    o.halo = c->halo;
    o.act = c->act;
    o.tact = c->tact;
    o.ch0 = sl.ch0;
    o.inc0 = sl.inc0;
    o.tl0 = sl.tl0;
    o.hl0 = sl.hl0;
    o.act0 = sl.act0;
    o.ta0 = sl.ta0;
    o.n_act = n_act;
    const DevBuildMirror mr{b.perm.data(), b.flags.data(), b.sl.data(), b.sr.data(), b.coff.data()};
    // the actuator records first: the tiles' copies are taken from them
    hipError_t e = hipSuccess;
    if (!P.acth.empty())
        e = hipMemcpyAsync(c->act + 4 * sl.act0, P.acth.data(), sizeof(double) * P.acth.size(), hipMemcpyHostToDevice,
                           c->up);
    if (e == hipSuccess) e = c->builder->write(o, mr, c->up);
    if (e == hipSuccess && n_lc > 0)
        e = hipMemcpyAsync(c->lc + sl.lc0, P.lch.data(), sizeof(LocalC) * n_lc, hipMemcpyHostToDevice, c->up);
    if (e == hipSuccess) e = hipStreamSynchronize(c->up);
    if (e != hipSuccess) {
        c->chunks.release(sl.ch0, nch);
        c->incs.release(sl.inc0, sz.n_inc);
        c->acts.release(sl.act0, n_act);
        c->lcs.release(sl.lc0, n_lc);
        c->tiles.release(sl.tl0, sz.ntiles);
        c->halos.release(sl.hl0, sz.n_halo);
        c->tacts.release(sl.ta0, sz.n_tact);
        c->bids.release(sl.bid, 1);
        fail(TITAN_E_HIP, std::string("body_create (device build): ") + hipGetErrorString(e));
    }
    // ---- join the launches
    BodyInit bi{};
    bi.bid = (int32_t)sl.bid;
    bi.chunk0 = sl.ch0;
    bi.nchunks = nch;
    bi.tile0 = sl.tl0;
    bi.ntiles = sz.ntiles;
    bi.res = -1;
    bi.t0 = P.t0;
    bi.k = P.k0;
    bi.m = P.m0;
    bi.d = P.d0;
    bi.lcb = (int32_t)(P.shared ? sl.lc0 : 0);
    bi.lcn = (int32_t)(P.shared ? n_lc : 0);
    bi.mode = kModeTiled;
    bi.kd = (uint8_t)((P.per_k ? 1 : 0) | (P.per_d ? 2 : 0) | (P.m_same ? kKdUniformM : 0));
    bi.on = 1;
    HIPCHK(launch_set_body(body_table(c), bi, c->stream));
    BodyH &h = c->bodies[sl.bid];
    h = std::move(b);
    h.live = true;
    h.tiled = true;
    h.resident = false;
    h.kd = bi.kd;
    h.n_masses = N;
    h.chunk0 = sl.ch0;
    h.nchunks = nch;
    h.inc0 = sl.inc0;
    h.inc_len = sz.n_inc;
    h.act0 = sl.act0;
    h.act_len = n_act;
    h.lc0 = sl.lc0;
    h.lc_len = n_lc;
    h.tile0 = sl.tl0;
    h.ntiles = sz.ntiles;
    h.halo0 = sl.hl0;
    h.nhalo = sz.n_halo;
    h.tact0 = sl.ta0;
    h.tact_len = sz.n_tact;
    h.rows = sz.rmax;
    h.T = P.t0;
    h.n = 0;
    h.row0.assign(nch, 0);
    for (int64_t ch = 0; ch < nch; ++ch) h.row0[ch] = (int32_t)(ch / (kTileMasses / kChunk)) * kTileMasses;
    return (int32_t)sl.bid;
}

int32_t body_create(titan_ctx *c, const titan_body_desc *d) {
    PreparedBody P = prepare_body(c->sched, d, c->device_min);
    const int32_t id = commit_body(c, P);
    // a large build's scratch is not kept
    if (c->builder && c->builder->pool_bytes() > ((size_t)256 << 20)) c->builder->release();
    return id;
}

// The body leaves the launches through a kernel on the step stream; its
// slots are released once that kernel has run (harvest_released).
void body_destroy(titan_ctx *c, int32_t id) {
    BodyH &b = body_of(c, id);
    maybe_inject(c, TITAN_INJECT_DESTROY);
    BodyInit bi{};
    bi.bid = id;
    bi.chunk0 = b.chunk0;
    bi.nchunks = b.nchunks;
    bi.tile0 = b.tile0;
    bi.ntiles = b.ntiles;
    bi.res = b.resident ? b.res : -1;
    bi.on = 0;
    HIPCHK(launch_set_body(body_table(c), bi, c->stream));
    hipEvent_t ev;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(ev, c->stream));
    c->pending.push_back({b, id, ev});
    b = BodyH{};
}

// Remove springs and masses of a live body (Simulation::deleteSpring /
// deleteMass after start, src/sim.cu:352-412).  A removed spring leaves the
// incidence lists of both its endpoints; a removed mass becomes fixed (the
// reference stops updating it: it leaves d_masses) and every spring
// touching it is removed, as an invalidated endpoint makes a spring inert
// (src/sim.cu:1052).  Only the chunks holding a touched mass are rewritten:
// each touched lane's remaining incidences move up in order, the freed
// positions become padding, and its degree shrinks; group counts, tiles and
// halos are kept (a halo row no spring names any more is loaded but unused),
// so the body's layout, captured graphs and results of untouched masses are
// as before.  A resident body stops using its split form (its walk gives the
// same results).
void body_remove(titan_ctx *c, int32_t id, const int64_t *masses, int64_t n_m, const int64_t *springs,
                 int64_t n_s) {
    BodyH &b = body_of(c, id);
    const int64_t N = b.n_masses, S = (int64_t)b.sl.size();
    if (n_m < 0 || n_s < 0 || (n_m > 0 && !masses) || (n_s > 0 && !springs))
        fail(TITAN_E_INVALID, "body_remove: bad arguments");
    for (int64_t q = 0; q < n_m; ++q)
        if (masses[q] < 0 || masses[q] >= N) fail(TITAN_E_INVALID, "body_remove: mass index out of range");
    for (int64_t q = 0; q < n_s; ++q)
        if (springs[q] < 0 || springs[q] >= S) fail(TITAN_E_INVALID, "body_remove: spring index out of range");
    if (b.sdead.empty()) b.sdead.assign(S, 0);
    if (b.mptr.empty()) {  // per-mass lists, in the order build_sell lists them
        b.mptr.assign(N + 1, 0);
        for (int64_t k = 0; k < S; ++k) {
            if (b.sl[k] < 0 || b.sr[k] < 0) continue;
            b.mptr[b.sr[k] + 1]++;
            b.mptr[b.sl[k] + 1]++;
        }
        for (int64_t q = 0; q < N; ++q) b.mptr[q + 1] += b.mptr[q];
        b.mspr.assign(b.mptr[N], 0);
        std::vector<int64_t> at(b.mptr.begin(), b.mptr.end() - 1);
        for (int64_t k = 0; k < S; ++k) {
            if (b.sl[k] < 0 || b.sr[k] < 0) continue;
            b.mspr[at[b.sr[k]]++] = (int32_t)k;  // the right end first (a self-spring is listed twice)
            b.mspr[at[b.sl[k]]++] = (int32_t)k;
        }
    }
    std::vector<int32_t> dev_of(N);  // caller's mass index -> device order
    for (int64_t q = 0; q < N; ++q) dev_of[b.perm.empty() ? q : b.perm[q]] = (int32_t)q;
    // springs removed now, and the masses whose lists change
    std::vector<int32_t> gone;
    std::vector<uint8_t> fix(N, 0);
    for (int64_t q = 0; q < n_m; ++q) {
        const int32_t m = dev_of[masses[q]];
        fix[m] = 1;
        for (int64_t e = b.mptr[m]; e < b.mptr[m + 1]; ++e) gone.push_back(b.mspr[e]);
    }
    for (int64_t q = 0; q < n_s; ++q) gone.push_back((int32_t)springs[q]);
    std::sort(gone.begin(), gone.end());
    gone.erase(std::unique(gone.begin(), gone.end()), gone.end());
    std::vector<uint8_t> now(S, 0);
    std::vector<int32_t> touched;
    for (int32_t k : gone) {
        if (b.sdead[k] || b.sl[k] < 0 || b.sr[k] < 0) continue;
        now[k] = 1;
        touched.push_back(b.sl[k]);
        touched.push_back(b.sr[k]);
    }
    std::vector<int32_t> fixed_now;
    for (int64_t m = 0; m < N; ++m)
        if (fix[m] && !(b.flags[m] & kFixed)) fixed_now.push_back((int32_t)m);
    std::sort(touched.begin(), touched.end());
    touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
    if (touched.empty() && fixed_now.empty()) return;

    HIPCHK(hipStreamSynchronize(c->stream));  // no launch in flight reads what is rewritten
    const bool packed = b.tiled || b.resident;
    const int G = packed ? kGroupTiled : kGroupGather;
    const int64_t g0 = b.chunk0 * kChunk;
    const bool pk = (b.kd & 1u) != 0, pd = (b.kd & 2u) != 0;
    size_t t = 0;
    while (t < touched.size()) {
        const int64_t ch = touched[t] / kChunk, e0 = b.coff[ch], len = b.coff[ch + 1] - e0;
        std::vector<uint32_t> nbr(len);
        std::vector<double> rest(len), kk(pk ? len : 0), dd(pd ? len : 0);
        std::vector<int32_t> deg(kChunk);
        d2h(c, nbr.data(), c->inc_nbr + b.inc0 + e0, len);
        d2h(c, rest.data(), c->inc_rest + b.inc0 + e0, len);
        if (pk) d2h(c, kk.data(), c->inc_k + b.inc0 + e0, len);
        if (pd) d2h(c, dd.data(), c->inc_damp + b.inc0 + e0, len);
        d2h(c, deg.data(), c->deg + g0 + ch * kChunk, kChunk);
        HIPCHK(hipStreamSynchronize(c->stream));
        // the name of entry e (relative to the chunk), read and written
        auto name = [&](int64_t e) -> uint32_t {
            if (!packed) return nbr[e];
            return (nbr[e / 2] >> (16 * (int)(e % 2))) & 0xffffu;
        };
        auto set_name = [&](int64_t e, uint32_t v) {
            if (!packed) {
                nbr[e] = v;
                return;
            }
            uint32_t &w = nbr[e / 2];
            const int sh = 16 * (int)(e % 2);
            w = (w & ~(0xffffu << sh)) | ((v & 0xffffu) << sh);
        };
        auto entry = [&](int lane, int64_t j) -> int64_t { return ((j / G) * kChunk + lane) * G + j % G; };
        for (; t < touched.size() && touched[t] / kChunk == ch; ++t) {
            const int32_t m = touched[t];
            const int lane = (int)(m % kChunk);
            struct E {
                uint32_t n;
                double r, k, d;
            };
            std::vector<E> keep;
            int64_t j = 0;  // position in the list as it stands on the device
            for (int64_t e = b.mptr[m]; e < b.mptr[m + 1]; ++e) {
                const int32_t k = b.mspr[e];
                if (b.sdead[k]) continue;
                if (!now[k]) {
                    const int64_t x = entry(lane, j);
                    keep.push_back({name(x), rest[x], pk ? kk[x] : 0.0, pd ? dd[x] : 0.0});
                }
                ++j;
            }
            if (j != deg[lane]) fail(TITAN_E_HIP, "body_remove: incidence list out of step with the device");
            // padding names the mass itself (its row, or its arena slot)
            const uint32_t self = packed ? (uint32_t)(m - b.row0[ch]) : (uint32_t)(g0 + m);
            for (int64_t q = 0; q < j; ++q) {
                const int64_t x = entry(lane, q);
                if (q < (int64_t)keep.size()) {
                    set_name(x, keep[q].n);
                    rest[x] = keep[q].r;
                    if (pk) kk[x] = keep[q].k;
                    if (pd) dd[x] = keep[q].d;
                } else {
                    set_name(x, self);
                    rest[x] = 1.0;
                    if (pk) kk[x] = 0.0;
                    if (pd) dd[x] = 0.0;
                }
            }
            deg[lane] = (int32_t)keep.size();
        }
        h2d(c, c->inc_nbr + b.inc0 + e0, nbr.data(), len);
        h2d(c, c->inc_rest + b.inc0 + e0, rest.data(), len);
        if (pk) h2d(c, c->inc_k + b.inc0 + e0, kk.data(), len);
        if (pd) h2d(c, c->inc_damp + b.inc0 + e0, dd.data(), len);
        h2d(c, c->deg + g0 + ch * kChunk, deg.data(), kChunk);
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    for (int64_t k = 0; k < S; ++k)
        if (now[k]) b.sdead[k] = 1;
    for (int32_t m : fixed_now) {
        b.flags[m] |= kFixed;
        h2d(c, c->mflags + g0 + m, &b.flags[m], 1);
    }
    if (b.resident && b.res_scache > 0) {  // the split form lists every spring: walk instead
        b.res_scache = 0;
        const int32_t zero = 0;
        h2d(c, reinterpret_cast<int32_t *>(reinterpret_cast<char *>(c->rdesc + b.res) + offsetof(ResidentDesc, scache)),
            &zero, 1);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
}

// device order -> caller's order for a (N, K) host array read from K SoA rows
void unpermute(const BodyH &b, const double *soa, int K, double *out) {
    const int64_t N = b.n_masses;
    for (int64_t q = 0; q < N; ++q) {
        const int64_t i = b.perm.empty() ? q : b.perm[q];
        for (int k = 0; k < K; ++k) out[K * i + k] = soa[k * N + q];
    }
}

// A pinned host buffer of at least n doubles: copies between it and the
// device run at the link's rate, without the runtime's pageable staging.
double *staging(titan_ctx *c, int64_t n) {
    if (n > c->stage_cap) {
        if (c->stage) (void)hipHostFree(c->stage);
        c->stage = nullptr;
        c->stage_cap = 0;
        const int64_t cap = std::max<int64_t>(n, 2 * c->stage_cap);
        HIPCHK(hipHostMalloc((void **)&c->stage, sizeof(double) * cap, hipHostMallocDefault));
        c->stage_cap = cap;
    }
    return c->stage;
}

void body_download(titan_ctx *c, int32_t id, double *pos, double *vel, double *acc, double *T) {
    BodyH &b = body_of(c, id);
    maybe_inject(c, TITAN_INJECT_DOWNLOAD);
    const int64_t N = b.n_masses, g0 = b.chunk0 * kChunk, n6 = c->ncap;
    const int p = (int)((b.n * launches_per_step(c)) & 1);
    auto host = [&](int64_t q) -> int64_t { return b.perm.empty() ? q : b.perm[q]; };
    HIPCHK(hipStreamSynchronize(c->stream));
    if (pos || vel) {
        double *st = staging(c, 6 * N);
        d2h(c, st, c->st + ((int64_t)p * n6 + g0) * 6, 6 * N);
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int64_t q = 0; q < N; ++q) {
            const int64_t i = host(q);
            for (int k = 0; k < 3; ++k) {
                if (pos) pos[3 * i + k] = st[6 * q + k];
                if (vel) vel[3 * i + k] = st[6 * q + 3 + k];
            }
        }
    }
    auto fetch = [&](const double *base, int K, double *out) {
        double *tmp = staging(c, 3 * N);
        for (int k = 0; k < K; ++k) d2h(c, tmp + k * N, base + k * n6 + g0, N);
        HIPCHK(hipStreamSynchronize(c->stream));
        unpermute(b, tmp, K, out);
    };
    if (acc) fetch(c->acc, 3, acc);
    if (T) fetch(c->Tm, 1, T);
}

void body_upload(titan_ctx *c, int32_t id, const double *pos, const double *vel, const double *m,
                 const double *T, const double *ext) {
    BodyH &b = body_of(c, id);
    const int64_t N = b.n_masses, g0 = b.chunk0 * kChunk, n6 = c->ncap;
    const int p = (int)((b.n * launches_per_step(c)) & 1);
    HIPCHK(hipStreamSynchronize(c->stream));
    auto host = [&](int64_t q) -> int64_t { return b.perm.empty() ? q : b.perm[q]; };
    if (pos || vel) {  // read-modify-write of the 48-byte records
        double *st = staging(c, 6 * N);
        double *base = c->st + ((int64_t)p * n6 + g0) * 6;
        d2h(c, st, base, 6 * N);
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int64_t q = 0; q < N; ++q) {
            const int64_t i = host(q);
            for (int k = 0; k < 3; ++k) {
                if (pos) st[6 * q + k] = pos[3 * i + k];
                if (vel) st[6 * q + 3 + k] = vel[3 * i + k];
            }
        }
        h2d(c, base, st, 6 * N);
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    auto put = [&](double *base, int K, const double *in) {
        double *tmp = staging(c, 3 * N);
        for (int64_t q = 0; q < N; ++q) {
            const int64_t i = host(q);
            for (int k = 0; k < K; ++k) tmp[k * N + q] = in[K * i + k];
        }
        for (int k = 0; k < K; ++k) h2d(c, base + k * n6 + g0, tmp + k * N, N);
        HIPCHK(hipStreamSynchronize(c->stream));
    };
    if (m) {
        put(c->m, 1, m);
        // the body's masses may no longer share one value (or may now)
        bool same = true;
        for (int64_t i = 1; i < N && same; ++i) same = same_bits(m[i], m[0]);
        const uint8_t kd = same ? (uint8_t)(b.kd | kKdUniformM) : (uint8_t)(b.kd & ~kKdUniformM);
        h2d(c, c->b_m + id, &m[0], 1);
        h2d(c, c->b_kd + id, &kd, 1);
        HIPCHK(hipStreamSynchronize(c->stream));
        b.kd = kd;
    }
    if (T) put(c->Tm, 1, T);
    if (ext) {
        bool changed = false;
        for (int64_t q = 0; q < N; ++q) {
            const int64_t i = host(q);
            const bool nz = ext[3 * i] != 0 || ext[3 * i + 1] != 0 || ext[3 * i + 2] != 0;
            const uint32_t f = nz ? (b.flags[q] | kExt) : (b.flags[q] & ~kExt);
            if (f != b.flags[q]) {
                b.flags[q] = f;
                changed = true;
            }
        }
        if (!c->ext) {  // launches captured so far have no extern-force array: drop them
            c->ext = dalloc<double>(c, 3 * c->ncap);
            HIPCHK(hipMemsetAsync(c->ext, 0, sizeof(double) * 3 * c->ncap, c->stream));
            drop_graphs(c);
        }
        put(c->ext, 3, ext);
        if (changed) {
            h2d(c, c->mflags + g0, b.flags.data(), N);
            HIPCHK(hipStreamSynchronize(c->stream));
            drop_graphs(c);  // a captured split-form launch may have chosen the lean mass pass
        }
    }
}

// The address of the scratch array on `device` (current), looked up once per
// device and kept for the process.
double *device_scratch(int device) {
    static std::mutex mu;
    static std::map<int, double *> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(device);
    if (it != addr.end()) return it->second;
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(TITAN_E_HIP, "hipGetSymbolAddress(scratch) gave a null address");
    addr[device] = static_cast<double *>(p);
    return static_cast<double *>(p);
}

void set_err(const titan_ctx *c, const std::string &m) {
    if (c) const_cast<titan_ctx *>(c)->err = m;
    else g_create_err = m;
}

}  // namespace

#define API_BEGIN try {
#define API_END(C)                                                  \
    }                                                               \
    catch (const Fail &f) {                                         \
        set_err(C, f.msg);                                          \
        return f.code;                                              \
    }                                                               \
    catch (const std::bad_alloc &) {                                \
        set_err(C, "host allocation failed");                       \
        return TITAN_E_NOMEM;                                       \
    }                                                               \
    catch (const std::exception &e) {                               \
        set_err(C, e.what());                                       \
        return TITAN_E_INVALID;                                     \
    }                                                               \
    return TITAN_OK;

#define NEED_CTX(C) \
    if (!(C)) {     \
        g_create_err = "null context"; \
        return TITAN_E_INVALID; \
    }

extern "C" {

int titan_ctx_create(int device, uint64_t hbm_budget_bytes, titan_ctx **out) {
    if (!out) {
        g_create_err = "ctx_create: null output pointer";
        return TITAN_E_INVALID;
    }
    *out = nullptr;
    titan_ctx *c = nullptr;
    API_BEGIN
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) fail(TITAN_E_INVALID, "ctx_create: no HIP device " + std::to_string(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        fail(TITAN_E_INVALID, std::string("ctx_create: this library is built for gfx950, device is ") + prop.gcnArchName);
    HIPCHK(hipSetDevice(device));
    c = new titan_ctx();
    c->device = device;
    c->budget = hbm_budget_bytes;
    {
        const char *g = std::getenv("TITAN_GRAPHS");
        c->use_graphs = g ? std::strcmp(g, "0") != 0 : std::getenv("ROCPROF_OUTPUT_PATH") == nullptr;
        const char *db = std::getenv("TITAN_DEVICE_BUILD");
        if (db) c->device_min = std::max<int64_t>(0, std::atoll(db));
    }
    HIPCHK(set_tiles_lds_limit(&c->res_lds_max, &c->split_lds_max));
    c->scratch = device_scratch(device);
    HIPCHK(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&c->up, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&c->ev0));
    HIPCHK(hipEventCreate(&c->ev1));
    for (int k = 0; k < 2; ++k) HIPCHK(hipEventCreateWithFlags(&c->left_ev[k], hipEventDisableTiming));
    ensure_bodies(c, 64);
    c->planes = dalloc<double>(c, 8 * 16);
    c->planes_cap = 16;
    c->balls = dalloc<double>(c, 4 * 16);
    c->balls_cap = 16;
    *out = c;
    c = nullptr;
    }
    catch (const Fail &f) {
        g_create_err = f.msg;
        if (c) titan_ctx_destroy(c);
        return f.code;
    }
    catch (const std::exception &e) {
        g_create_err = e.what();
        if (c) titan_ctx_destroy(c);
        return TITAN_E_INVALID;
    }
    return TITAN_OK;
}

// serve.cpp: drop the context's request table
extern "C" __attribute__((visibility("hidden"))) void titan_internal_serve_forget(const titan_ctx *ctx);

int titan_ctx_destroy(titan_ctx *c) {
    if (!c) return TITAN_OK;
    titan_internal_serve_forget(c);
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->up) (void)hipStreamSynchronize(c->up);
    for (auto &p : c->pending) (void)hipEventDestroy(p.ev);
    for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    for (auto &a : c->allocs) (void)hipFree(a.first);
    delete c->builder;
    for (int k = 0; k < 2; ++k) {
        if (c->left_ev[k]) (void)hipEventDestroy(c->left_ev[k]);
        if (c->left_pin[k]) (void)hipHostFree(c->left_pin[k]);
    }
    if (c->stage) (void)hipHostFree(c->stage);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->up) (void)hipStreamDestroy(c->up);
    delete c;
    return TITAN_OK;
}

const char *titan_last_error(const titan_ctx *c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int titan_set_dt(titan_ctx *c, double dt) {
    NEED_CTX(c);
    API_BEGIN
    if (!(dt > 0)) fail(TITAN_E_INVALID, "Cannot set time step to negative or zero value.");  // src/sim.cu:802-804
    c->dt = dt;
    drop_graphs(c);
    API_END(c)
}

int titan_get_dt(const titan_ctx *c, double *dt) {
    NEED_CTX(c);
    if (!dt) return TITAN_E_INVALID;
    *dt = c->dt;
    return TITAN_OK;
}

int titan_set_global_acc(titan_ctx *c, double gx, double gy, double gz) {
    NEED_CTX(c);
    API_BEGIN
    c->g[0] = gx;
    c->g[1] = gy;
    c->g[2] = gz;
    drop_graphs(c);
    API_END(c)
}

int titan_add_plane(titan_ctx *c, double nx, double ny, double nz, double d, double mu_k, double mu_s) {
    NEED_CTX(c);
    API_BEGIN
    c->planes_h.insert(c->planes_h.end(), {nx, ny, nz, d, mu_k, mu_s, 0.0, 0.0});
    const int64_t n = (int64_t)c->planes_h.size() / 8;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (n > c->planes_cap) {
        dfree(c, c->planes);
        c->planes_cap = 2 * n;
        c->planes = dalloc<double>(c, 8 * c->planes_cap);
    }
    h2d(c, c->planes, c->planes_h.data(), (int64_t)c->planes_h.size());
    HIPCHK(hipStreamSynchronize(c->stream));
    drop_graphs(c);
    API_END(c)
}

int titan_add_ball(titan_ctx *c, double cx, double cy, double cz, double r) {
    NEED_CTX(c);
    API_BEGIN
    c->balls_h.insert(c->balls_h.end(), {cx, cy, cz, r});
    const int64_t n = (int64_t)c->balls_h.size() / 4;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (n > c->balls_cap) {
        dfree(c, c->balls);
        c->balls_cap = 2 * n;
        c->balls = dalloc<double>(c, 4 * c->balls_cap);
    }
    h2d(c, c->balls, c->balls_h.data(), (int64_t)c->balls_h.size());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->ball_grid_dirty = true;
    drop_graphs(c);
    API_END(c)
}

int titan_clear_planes_and_balls(titan_ctx *c) {
    NEED_CTX(c);
    API_BEGIN
    c->planes_h.clear();
    c->balls_h.clear();
    c->ball_grid_dirty = true;
    drop_graphs(c);
    API_END(c)
}

int titan_body_create(titan_ctx *c, const titan_body_desc *desc, int32_t *body_id) {
    NEED_CTX(c);
    API_BEGIN
    int32_t id = body_create(c, desc);
    if (body_id) *body_id = id;
    API_END(c)
}

int titan_body_destroy(titan_ctx *c, int32_t body_id) {
    NEED_CTX(c);
    API_BEGIN
    body_destroy(c, body_id);
    API_END(c)
}

int titan_body_remove(titan_ctx *c, int32_t body_id, const int64_t *masses, int64_t n_masses,
                      const int64_t *springs, int64_t n_springs) {
    NEED_CTX(c);
    API_BEGIN
    body_remove(c, body_id, masses, n_masses, springs, n_springs);
    API_END(c)
}

int titan_body_count(const titan_ctx *c, int32_t *n) {
    NEED_CTX(c);
    if (!n) return TITAN_E_INVALID;
    int32_t k = 0;
    for (auto &b : c->bodies) k += b.live ? 1 : 0;
    *n = k;
    return TITAN_OK;
}

int titan_step(titan_ctx *c, int64_t n_steps) {
    NEED_CTX(c);
    API_BEGIN
    do_step(c, n_steps);
    API_END(c)
}

int titan_step_noalloc(titan_ctx *c, int64_t n_steps) {
    NEED_CTX(c);
    API_BEGIN
    do_step(c, n_steps, true);
    API_END(c)
}

int titan_step_until(titan_ctx *c, double t, int64_t *steps_done) {
    NEED_CTX(c);
    API_BEGIN
    if (std::isnan(t)) fail(TITAN_E_INVALID, "step_until: time is NaN");
    // the driver stops at the first loop iteration with breakpoint <= T
    // (src/sim.cu:1505); count the steps with the same T += dt recurrence
    double T = c->T;
    int64_t k = 0;
    const int64_t lps = launches_per_step(c);
    const double h = launch_dt(c);
    while (!(t <= T)) {
        for (int64_t q = 0; q < lps; ++q) T += h;
        if (++k > (int64_t)1 << 40) fail(TITAN_E_INVALID, "step_until: target time too far ahead");
    }
    do_step(c, k);
    if (steps_done) *steps_done = k;
    API_END(c)
}

int titan_time(const titan_ctx *c, double *T) {
    NEED_CTX(c);
    if (!T) return TITAN_E_INVALID;
    *T = c->T;
    return TITAN_OK;
}

int titan_reserve(titan_ctx *c, int64_t n_masses, int64_t n_springs, int64_t n_actuated, int32_t n_bodies) {
    NEED_CTX(c);
    API_BEGIN
    if (n_masses < 0 || n_springs < 0 || n_actuated < 0 || n_actuated > n_springs || n_bodies < 0)
        fail(TITAN_E_INVALID, "reserve: negative count, or more actuated springs than springs");
    if (n_masses == 0 && n_bodies == 0) return TITAN_OK;
    const int64_t nb = std::max<int64_t>(n_bodies, 1);
    // every body starts on a chunk boundary; a lane's incidence list is
    // padded to its chunk's group count (half as many entries again, plus
    // eight per mass slot, covers the padding of every layout on lattices
    // and kNN clouds) and a tile's halo lists up to about two outside masses
    // per tile mass
    const int64_t chunks = (n_masses + kChunk - 1) / kChunk + nb;
    const int64_t incs = 3 * n_springs + chunks * kChunk * 8;
    const int64_t tiles = (chunks * kChunk + kTileMasses - 1) / kTileMasses + nb;
    if (chunks > max_chunk_cap()) fail(TITAN_E_NOMEM, "reserve: mass arena limit: at most 2^26 mass slots");
    // the whole reservation must fit the budget before anything grows: the
    // bytes of every arena that grows, by the rule each growth applies
    auto grows = [](int64_t old, int64_t need, int64_t floor, size_t per) -> size_t {
        const int64_t nc = grown_cap(old, need, floor);
        return nc > old ? (size_t)nc * per : 0;
    };
    const int64_t old_ch = c->chunks.cap();
    size_t bytes = old_ch >= chunks ? 0 : (size_t)std::min(grown_cap(old_ch, chunks, 16), max_chunk_cap()) * chunk_bytes(c);
    bytes += grows(c->incs.cap(), incs, 4096, inc_bytes(c));
    if (n_actuated > 0) {
        bytes += grows(c->acts.cap(), n_actuated, 1024, act_bytes(c));
        bytes += grows(c->tacts.cap(), 2 * n_actuated, 1024, 4 * sizeof(double));
    }
    bytes += grows(c->tiles.cap(), tiles, 64, sizeof(TileDesc));
    bytes += grows(c->halos.cap(), 2 * n_masses, 1024, sizeof(int32_t));
    bytes += grows(c->rdescs.cap(), nb, 64, sizeof(ResidentDesc));
    bytes += grows(c->bcap, nb, 64, body_bytes());
    check_growth(c, bytes);
    ensure_chunks(c, chunks);
    ensure_incs(c, incs);
    if (n_actuated > 0) {
        ensure_acts(c, n_actuated);
        ensure_tacts(c, 2 * n_actuated);  // a spring is in the lists of at most two tiles
    }
    ensure_tiles(c, tiles);
    ensure_halos(c, 2 * n_masses);
    ensure_rdescs(c, nb);
    ensure_bodies(c, nb);
    API_END(c)
}

int titan_set_schedule(titan_ctx *c, const titan_schedule *s) {
    NEED_CTX(c);
    API_BEGIN
    if (!s) fail(TITAN_E_INVALID, "set_schedule: null schedule");
    if (s->tile_rows < kChunk || s->tile_rows > kMaxTileRows)
        fail(TITAN_E_INVALID, "set_schedule: tile_rows must be in [64, " + std::to_string(kMaxTileRows) + "]");
    if (s->tiled_min_masses < 0) fail(TITAN_E_INVALID, "set_schedule: negative tiled_min_masses");
    if (s->resident_max_masses < 0 || s->resident_max_masses > kResidentMaxBody)
        fail(TITAN_E_INVALID, "set_schedule: resident_max_masses must be in [0, " + std::to_string(kResidentMaxBody) +
                                  "]");
    c->sched = *s;
    API_END(c)
}

int titan_set_integrator(titan_ctx *c, int32_t kind) {
    NEED_CTX(c);
    API_BEGIN
    if (kind < TITAN_INTEGRATOR_EULER || kind > TITAN_INTEGRATOR_RK2) fail(TITAN_E_INVALID, "set_integrator: bad kind");
    for (const auto &b : c->bodies)
        if (b.live) fail(TITAN_E_INVALID, "set_integrator: choose the integrator before creating bodies");
    if (kind == c->integ) return TITAN_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->integ = kind;
    if (kind == kIntRK2 && !c->rk && c->ncap > 0) {
        c->rk = dalloc<double>(c, 6 * c->ncap);
        HIPCHK(hipMemsetAsync(c->rk, 0, sizeof(double) * 6 * c->ncap, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    drop_graphs(c);
    API_END(c)
}

int titan_get_integrator(const titan_ctx *c, int32_t *kind) {
    NEED_CTX(c);
    if (!kind) return TITAN_E_INVALID;
    *kind = c->integ;
    return TITAN_OK;
}

int titan_get_schedule(const titan_ctx *c, titan_schedule *s) {
    NEED_CTX(c);
    if (!s) return TITAN_E_INVALID;
    *s = c->sched;
    return TITAN_OK;
}

namespace {

// One autotuning trial: a scratch context with c's globals, `replicas`
// copies of the samples under schedule `sc`, timed over `steps` steps.
double time_schedule(titan_ctx *c, const titan_body_desc *samples, int32_t n_samples, int32_t replicas,
                     int32_t steps, const titan_schedule &sc) {
    titan_ctx *t = nullptr;
    auto chk = [&](int rc, const char *what) {
        if (rc != TITAN_OK) {
            std::string msg = std::string("autotune: ") + what + ": " + titan_last_error(t);
            titan_ctx_destroy(t);
            fail(rc, msg);
        }
    };
    if (titan_ctx_create(c->device, c->budget, &t) != TITAN_OK)
        fail(TITAN_E_HIP, std::string("autotune: scratch context: ") + titan_last_error(nullptr));
    chk(titan_set_dt(t, c->dt), "set_dt");
    chk(titan_set_integrator(t, c->integ), "set_integrator");
    chk(titan_set_global_acc(t, c->g[0], c->g[1], c->g[2]), "set_global_acc");
    for (size_t q = 0; q < c->planes_h.size(); q += 8) {
        const double *P = c->planes_h.data() + q;
        chk(titan_add_plane(t, P[0], P[1], P[2], P[3], P[4], P[5]), "add_plane");
    }
    for (size_t q = 0; q < c->balls_h.size(); q += 4) {
        const double *B = c->balls_h.data() + q;
        chk(titan_add_ball(t, B[0], B[1], B[2], B[3]), "add_ball");
    }
    chk(titan_set_schedule(t, &sc), "set_schedule");
    int64_t masses = 0;
    for (int32_t r = 0; r < replicas; ++r)
        for (int32_t k = 0; k < n_samples; ++k) {
            titan_body_desc d = samples[k];
            d.layout = TITAN_LAYOUT_AUTO;
            int32_t id = -1;
            chk(titan_body_create(t, &d, &id), "body_create");
            masses += d.n_masses;
        }
    chk(titan_step(t, steps), "warm-up");
    chk(titan_sync(t), "sync");
    float ms = 0;
    chk(titan_timer_start(t), "timer");
    chk(titan_step(t, steps), "step");
    chk(titan_timer_stop(t, &ms), "timer");
    titan_ctx_destroy(t);
    return ms > 0 ? (double)masses * steps / (ms * 1e-3) : 0.0;
}

}  // namespace

int titan_autotune(titan_ctx *c, const titan_body_desc *samples, int32_t n_samples, int32_t replicas,
                   int32_t steps, titan_schedule *best, double *rate_out) {
    NEED_CTX(c);
    API_BEGIN
    if (!samples || n_samples <= 0 || replicas <= 0 || steps <= 0)
        fail(TITAN_E_INVALID, "autotune: need samples, replicas > 0 and steps > 0");
    const int32_t res[] = {0, kSplitMaxChunks * kChunk, kResidentMax, kResidentMaxBody};
    const int64_t mins[] = {0, 128, 256, 512, 1024, INT64_MAX};
    const int32_t rows[] = {682, kMaxTileRows};
    titan_schedule win{0, kTileRows, 0};
    double win_rate = -1.0;
    // candidates that lay out every sample alike are timed once
    std::map<std::string, double> seen;
    int q = 0;
    for (int32_t rs : res)
        for (int64_t mn : mins)
            for (int32_t rw : rows) {
                const titan_schedule sc{mn, rw, rs};
                std::string sig;
                bool any_tiled = false;
                for (int32_t k = 0; k < n_samples; ++k) {
                    const int64_t n = samples[k].n_masses;
                    const char m = (n >= 1 && n <= rs) ? 'r' : (n >= mn ? 't' : 'g');
                    any_tiled |= m == 't';
                    sig += m;
                }
                if (any_tiled) sig += std::to_string(rw);
                double rate;
                auto it = seen.find(sig);
                if (it != seen.end()) {
                    rate = it->second;
                } else {
                    rate = time_schedule(c, samples, n_samples, replicas, steps, sc);
                    seen[sig] = rate;
                }
                if (rate_out) rate_out[q] = rate;
                ++q;
                if (rate > win_rate) {
                    win_rate = rate;
                    win = sc;
                }
            }
    c->sched = win;
    if (best) *best = win;
    API_END(c)
}

int titan_step_bodies(titan_ctx *c, const int32_t *body_ids, const int64_t *n_steps, int32_t count) {
    NEED_CTX(c);
    API_BEGIN
    do_step_bodies(c, body_ids, n_steps, count);
    API_END(c)
}

int titan_body_time(const titan_ctx *c, int32_t body_id, double *T, int64_t *steps) {
    NEED_CTX(c);
    titan_ctx *m = const_cast<titan_ctx *>(c);
    API_BEGIN
    const BodyH &b = body_of(m, body_id);
    if (T) *T = b.T;
    if (steps) *steps = b.n;
    API_END(m)
}

int titan_sync(titan_ctx *c) {
    NEED_CTX(c);
    API_BEGIN
    HIPCHK(hipStreamSynchronize(c->stream));
    API_END(c)
}

int titan_body_download(titan_ctx *c, int32_t body_id, double *pos, double *vel, double *acc, double *T) {
    NEED_CTX(c);
    API_BEGIN
    body_download(c, body_id, pos, vel, acc, T);
    API_END(c)
}

int titan_body_upload(titan_ctx *c, int32_t body_id, const double *pos, const double *vel,
                      const double *m, const double *T, const double *ext) {
    NEED_CTX(c);
    API_BEGIN
    body_upload(c, body_id, pos, vel, m, T, ext);
    API_END(c)
}

int titan_timer_start(titan_ctx *c) {
    NEED_CTX(c);
    API_BEGIN
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    c->timing = true;
    API_END(c)
}

int titan_timer_stop(titan_ctx *c, float *ms) {
    NEED_CTX(c);
    API_BEGIN
    if (!c->timing) fail(TITAN_E_INVALID, "timer_stop: timer not started");
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float v = 0;
    HIPCHK(hipEventElapsedTime(&v, c->ev0, c->ev1));
    c->timing = false;
    if (ms) *ms = v;
    API_END(c)
}

int titan_mem_used(const titan_ctx *c, uint64_t *bytes) {
    NEED_CTX(c);
    if (!bytes) return TITAN_E_INVALID;
    *bytes = c->used;
    return TITAN_OK;
}

int titan_kernel_name(const titan_ctx *c, char *buf, int64_t len) {
    NEED_CTX(c);
    if (!buf || len <= 0) return TITAN_E_INVALID;
    const bool act = c->has_act;
    const char *parts[3] = {any_body(c, kModeGather) ? step_kernel_name(act) : nullptr,
                            any_body(c, kModeTiled) ? step_tiles_kernel_name(act) : nullptr,
                            any_body(c, kModeResident) ? resident_kernel_name(act) : nullptr};
    if (parts[2]) {
        const StepArgs a = make_args(const_cast<titan_ctx *>(c), c->slot, full_range(c));
        if (a.res_split >= 2) parts[2] = split_kernel_name(act);
    }
    int64_t at = 0;
    buf[0] = 0;
    for (const char *p : parts) {
        if (!p || at >= len) continue;
        at += std::snprintf(buf + at, (size_t)(len - at), "%s%s", at ? "+" : "", p);
    }
    return TITAN_OK;
}

int titan_body_layout(const titan_ctx *c, int32_t body_id, int32_t *layout) {
    NEED_CTX(c);
    if (!layout) return TITAN_E_INVALID;
    if (body_id < 0 || body_id >= (int64_t)c->bodies.size() || !c->bodies[body_id].live) return TITAN_E_NOTFOUND;
    const BodyH &b = c->bodies[body_id];
    *layout = b.resident ? TITAN_LAYOUT_RESIDENT : b.tiled ? TITAN_LAYOUT_TILED : TITAN_LAYOUT_GATHER;
    return TITAN_OK;
}

int titan_measure_hbm(int device, uint64_t bytes, int32_t iters, double *copy_gbps, double *read_gbps) {
    if (bytes < 16 || iters < 1) {
        g_create_err = "measure_hbm: need bytes >= 16 and iters >= 1";
        return TITAN_E_INVALID;
    }
    void *src = nullptr, *dst = nullptr;
    double *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = TITAN_OK;
    try {
        HIPCHK(hipSetDevice(device));
        const int64_t n = (int64_t)(bytes / 16);
        HIPCHK(hipMalloc(&src, 16 * (size_t)n));
        HIPCHK(hipMalloc(&dst, 16 * (size_t)n));
        HIPCHK(hipMalloc((void **)&out, sizeof(double) * 1024 * 256));  // one per thread of the read grid
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipMemsetAsync(src, 0, 16 * (size_t)n, st));
        float ms = 0;
        HIPCHK(launch_stream_copy(src, dst, n, st));  // warm-up
        HIPCHK(hipEventRecord(e0, st));
        for (int32_t q = 0; q < iters; ++q) HIPCHK(launch_stream_copy(src, dst, n, st));
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (copy_gbps) *copy_gbps = 2.0 * 16.0 * (double)n * iters / ((double)ms * 1e6);
        HIPCHK(launch_stream_read(src, out, n, st));
        HIPCHK(hipEventRecord(e0, st));
        for (int32_t q = 0; q < iters; ++q) HIPCHK(launch_stream_read(src, out, n, st));
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (read_gbps) *read_gbps = 16.0 * (double)n * iters / ((double)ms * 1e6);
    } catch (const Fail &f) {
        g_create_err = f.msg;
        rc = f.code;
    }
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(src);
    (void)hipFree(dst);
    (void)hipFree(out);
    return rc;
}

int titan_debug_body_bytes(titan_ctx *c, int32_t body_id, int32_t what, void *out, int64_t cap, int64_t *len) {
    NEED_CTX(c);
    API_BEGIN
    const BodyH &b = body_of(c, body_id);
    const int64_t g0 = b.chunk0 * kChunk, npad = b.nchunks * kChunk;
    const int64_t ninc = b.coff.empty() ? 0 : b.coff.back();
    std::vector<char> buf;
    auto dev = [&](const void *src, size_t bytes) {
        buf.resize(bytes);
        if (bytes) HIPCHK(hipMemcpy(buf.data(), src, bytes, hipMemcpyDeviceToHost));
    };
    HIPCHK(hipStreamSynchronize(c->stream));
    switch (what) {
        case TITAN_DEBUG_NAMES: dev(c->inc_nbr + b.inc0, 4 * (size_t)ninc); break;
        case TITAN_DEBUG_REST: dev(c->inc_rest + b.inc0, 8 * (size_t)ninc); break;
        case TITAN_DEBUG_K: dev(b.kd & 1u ? c->inc_k + b.inc0 : nullptr, b.kd & 1u ? 8 * (size_t)ninc : 0); break;
        case TITAN_DEBUG_DAMPING: dev(b.kd & 2u ? c->inc_damp + b.inc0 : nullptr, b.kd & 2u ? 8 * (size_t)ninc : 0); break;
        case TITAN_DEBUG_DEG: dev(c->deg + g0, 4 * (size_t)npad); break;
        case TITAN_DEBUG_MFLAGS: dev(c->mflags + g0, 4 * (size_t)npad); break;
        case TITAN_DEBUG_STATE: dev(c->st + 6 * g0, 48 * (size_t)npad); break;
        case TITAN_DEBUG_MASS: dev(c->m + g0, 8 * (size_t)npad); break;
        case TITAN_DEBUG_TILES: {
            dev(c->tile + b.tile0, sizeof(TileDesc) * (size_t)b.ntiles);
            TileDesc *t = reinterpret_cast<TileDesc *>(buf.data());
            for (int64_t q = 0; q < b.ntiles; ++q) {  // relative to the body's ranges
                t[q].body = 0;
                t[q].chunk0 -= (int32_t)b.chunk0;
                t[q].halo_off -= b.halo0;
                t[q].act_off -= b.tact0;
            }
            break;
        }
        case TITAN_DEBUG_TACT: dev(c->tact + 4 * b.tact0, 32 * (size_t)b.tact_len); break;
        case TITAN_DEBUG_HALO: {
            dev(c->halo + b.halo0, 4 * (size_t)b.nhalo);
            int32_t *h = reinterpret_cast<int32_t *>(buf.data());
            for (int64_t q = 0; q < b.nhalo; ++q) h[q] -= (int32_t)g0;
            break;
        }
        case TITAN_DEBUG_CHUNKS: {  // chunk_ptr - inc0, chunk_ng, chunk_act0 - act0, chunk_nact: 4 int64 per chunk
            std::vector<int64_t> p(b.nchunks), a(b.nchunks);
            std::vector<int32_t> ng(b.nchunks), na(b.nchunks);
            HIPCHK(hipMemcpy(p.data(), c->chunk_ptr + b.chunk0, 8 * b.nchunks, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(a.data(), c->chunk_act0 + b.chunk0, 8 * b.nchunks, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(ng.data(), c->chunk_ng + b.chunk0, 4 * b.nchunks, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(na.data(), c->chunk_nact + b.chunk0, 4 * b.nchunks, hipMemcpyDeviceToHost));
            buf.resize(32 * (size_t)b.nchunks);
            int64_t *o = reinterpret_cast<int64_t *>(buf.data());
            for (int64_t q = 0; q < b.nchunks; ++q) {
                o[4 * q] = p[q] - b.inc0;
                o[4 * q + 1] = ng[q];
                o[4 * q + 2] = a[q] - b.act0;
                o[4 * q + 3] = na[q];
            }
            break;
        }
        case TITAN_DEBUG_PERM: {
            buf.resize(4 * (size_t)b.n_masses);
            for (int64_t q = 0; q < b.n_masses; ++q) {
                const int32_t v = b.perm.empty() ? (int32_t)q : b.perm[q];
                std::memcpy(buf.data() + 4 * q, &v, 4);
            }
            break;
        }
        default: fail(TITAN_E_INVALID, "debug_body_bytes: unknown item");
    }
    if (len) *len = (int64_t)buf.size();
    if (out && cap > 0) std::memcpy(out, buf.data(), (size_t)std::min<int64_t>(cap, (int64_t)buf.size()));
    API_END(c)
}

int titan_debug_inject_failure(titan_ctx *c, int32_t after) { return titan_debug_inject_failure_at(c, TITAN_INJECT_STEP, after); }

int titan_debug_inject_failure_at(titan_ctx *c, int32_t site, int32_t after) {
    NEED_CTX(c);
    if (site < TITAN_INJECT_STEP || site > TITAN_INJECT_DESTROY) return TITAN_E_INVALID;
    c->inject[site] = after < 0 ? -1 : after;
    return TITAN_OK;
}

}  // extern "C"

namespace titan_mi355x {

struct Prepared {
    PreparedBody body;
};

Prepared *prepare_body_sched(const titan_schedule &sched, const titan_body_desc *d, int *rc, std::string *err) {
    try {
        std::unique_ptr<Prepared> p(new Prepared);
        p->body = prepare_body(sched, d, 0);
        *rc = TITAN_OK;
        return p.release();
    } catch (const Fail &f) {
        *rc = f.code;
        *err = f.msg;
    } catch (const std::bad_alloc &) {
        *rc = TITAN_E_NOMEM;
        *err = "host allocation failed";
    } catch (const std::exception &e) {
        *rc = TITAN_E_INVALID;
        *err = e.what();
    }
    return nullptr;
}

int commit_body_device(titan_ctx *c, Prepared *p, int32_t *id) {
    API_BEGIN
    *id = commit_body(c, p->body);
    API_END(c)
}

void free_prepared(Prepared *p) { delete p; }

int record_step_event(titan_ctx *c, hipEvent_t ev) {
    API_BEGIN
    HIPCHK(hipEventRecord(ev, c->stream));
    API_END(c)
}

}  // namespace titan_mi355x
