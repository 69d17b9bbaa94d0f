// Host-side builder of the tiled layout used by k_step_tiles.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "layout.h"

namespace titan_mi355x {

// Spring view handed to the builders: endpoints already in the body's
// device order (-1 = no mass), parameters in the caller's spring order.
struct SpringsIn {
    int64_t n_springs = 0;
    const int32_t *left = nullptr, *right = nullptr;
    const double *k = nullptr, *rest = nullptr, *damping = nullptr;
    const int32_t *act = nullptr;  // body-local actuator index or -1, may be null
};

struct TileBuild {
    std::vector<TileDesc> tiles;   // offsets relative to this body's ranges
    std::vector<int32_t> halo;     // body-local mass indices (device order)
    std::vector<uint32_t> ends;    // (left << 16) | right, tile-local
    std::vector<double> rest, k, damp;
    std::vector<int32_t> act;
    std::vector<uint64_t> inc;     // 4 incidences (uint16) per lane per group
    std::vector<int32_t> sid;      // spring index of each tile-spring slot (host only)
    int32_t hmax = 0;
    std::string why;               // reason when the layout does not apply
};

// Order of the masses along a Morton curve of their positions (3 doubles
// each).  Returns perm with perm[new] = old; ties keep the original order.
std::vector<int32_t> morton_order(const double *pos, int64_t n);

// Build the tiles of a body with n masses (device order) and the given
// springs.  deg_out receives each mass's incident-spring count.  Returns
// false (with tb.why) if a tile exceeds the kernel's limits.
bool build_tiles(int64_t n, const SpringsIn &sp, bool per_spring_kd, TileBuild &tb,
                 std::vector<int32_t> &deg_out);

}  // namespace titan_mi355x
