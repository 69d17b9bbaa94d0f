// Device-side layout builder for large tiled bodies: the Morton order, the
// degree order inside tiles, the tiles' halos and the SELL-64 incidence
// lists, built on the GPU with the same rules as the host builders in
// tiles.cpp (morton_order, build_tiles, degree_sort_tiles, build_sell), so
// that the arena holds the same bytes either way.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.h"

namespace titan_mi355x {

// Tiled bodies of at least this many masses are laid out on the device by
// default: below it the host builders are as quick as the device's fixed
// costs (a dozen launches and three synchronisations).
constexpr int64_t kDeviceBuildMin = 16384;

// A body to build, as the caller gave it (host arrays, read during the
// build) and as the host validation found it.
struct DevBuildIn {
    int64_t N = 0, S = 0;
    const double *pos = nullptr, *vel = nullptr, *m = nullptr, *T = nullptr, *ext = nullptr;
    const uint8_t *fixed = nullptr;
    const int32_t *drag = nullptr;
    const int32_t *left = nullptr, *right = nullptr;
    const double *rest = nullptr, *k = nullptr, *damping = nullptr;
    const int32_t *act_local = nullptr;  // per spring: body-local actuator index or -1; null: none
    bool per_k = false, per_d = false, any_ext = false, any_drag = false;
    int32_t tile_rows = kTileRows;
};

// Where the body goes: the context's arena arrays and the ranges reserved
// for it.
struct DevBuildOut {
    int64_t ncap = 0;
    double *st = nullptr, *m = nullptr, *Tm = nullptr, *acc = nullptr, *ext = nullptr;
    uint32_t *mflags = nullptr;
    int32_t *deg = nullptr, *drag = nullptr;
    int64_t *mlc = nullptr;  // zeroed for the body (shared lists only)
    int64_t *chunk_ptr = nullptr, *chunk_act0 = nullptr;
    int32_t *chunk_ng = nullptr, *chunk_nact = nullptr;
    uint32_t *inc_nbr = nullptr;
    double *inc_rest = nullptr, *inc_k = nullptr, *inc_damp = nullptr;
    TileDesc *tiles = nullptr;
    int32_t *halo = nullptr;
    const double *act = nullptr;  // the actuator records, already written at act0
    double *tact = nullptr;       // the tiles' copies of them go to [ta0, ta0 + n_tact)
    int64_t ch0 = 0, inc0 = 0, tl0 = 0, hl0 = 0, act0 = 0, n_act = 0, ta0 = 0;
};

// The sizes a build needs reserved, known after its first phase.
struct DevBuildSizes {
    int64_t n_inc = 0, n_halo = 0, ntiles = 0, nch = 0;
    int64_t n_tact = 0;     // the tiles' actuators (entries of the tile actuator arena)
    int32_t rmax = 0;       // largest tile rows (masses + halo + actuator rows)
    bool fits = true;       // false: a tile exceeds tile_rows, or a degree exceeds 65535 (build on the host)
};

// Host copies of what the engine keeps for a body.
struct DevBuildMirror {
    int32_t *perm = nullptr;     // N: device order -> caller's order
    uint32_t *flags = nullptr;   // N
    int32_t *sl = nullptr, *sr = nullptr;  // S: endpoints in device order (-1: none)
    int64_t *coff = nullptr;     // nch + 1
};

class DevBuilder {
   public:
    DevBuilder() = default;
    ~DevBuilder();
    DevBuilder(const DevBuilder &) = delete;
    DevBuilder &operator=(const DevBuilder &) = delete;

    // First phase, on stream s: uploads the body and computes the layout up
    // to the sizes the arena must reserve.  Device scratch comes from this
    // builder's pool (grown on demand, kept for the next build).  Returns a
    // HIP error, or hipSuccess with *sizes filled (the stream synchronised).
    hipError_t plan(const DevBuildIn &in, hipStream_t s, DevBuildSizes *sizes);
    // Second phase: writes the body into the reserved arena ranges and the
    // host mirror (the stream synchronised on return).
    hipError_t write(const DevBuildOut &out, const DevBuildMirror &mirror, hipStream_t s);
    // Bytes of device scratch held.
    size_t pool_bytes() const { return cap_; }
    void release();

   private:
    DevBuildIn in_{};
    DevBuildSizes sz_{};
    char *pool_ = nullptr;
    size_t cap_ = 0;
    // carved from the pool by plan()
    struct Bufs;
    Bufs *b_ = nullptr;
};

}  // namespace titan_mi355x
