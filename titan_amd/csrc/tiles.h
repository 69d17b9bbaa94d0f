// Host-side builders of the device layouts: the Morton order and tiles of
// the tiled layout, and the SELL-64 incidence lists shared by both layouts.
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
    std::vector<TileDesc> tiles;   // chunk0, halo_off and act_off relative to this body
    std::vector<int32_t> halo;     // body-local mass indices (device order), ascending per tile
    // the tiles' actuator lists (acts: the body has actuated springs): the
    // actuated springs with an endpoint in the tile, ascending, at act_off
    std::vector<int64_t> tact;
    bool acts = false;
    std::vector<int32_t> tile_of_chunk;
    int32_t hmax = 0;              // largest halo
    int32_t rmax = 0;              // largest nmass + nhalo + tile_act_rows(nact) (LDS rows)
    std::string why;               // reason when the layout does not apply
};

// Incidences of a body in SELL-64 order (see layout.h): for the masses of
// chunk c, incidence j of lane l is entry coff[c] + ((j / G) * 64 + l) * G +
// j % G, with groups of G = 4 (gather) or 2 (tiled).  Every mass lists its
// springs in ascending spring index (a spring joining a mass to itself is
// listed twice).  Padding entries name the mass itself with rest 1 and are
// never evaluated.
struct SellBuild {
    std::vector<int64_t> coff;     // nch + 1 entry offsets, body-relative
    std::vector<int32_t> deg;      // per mass slot, nch * 64
    // gather: nbr[e] = arena slot g0 + other endpoint.  tiled: the uint32 at
    // coff[c] + (e - coff[c]) / 2 holds the LDS rows of one group's two
    // entries (low half: the even entry); the chunk's other half is unused
    std::vector<uint32_t> nbr;
    std::vector<double> rest, k, damp;  // k, damp: only when built per incidence
    // only if sp.act: per entry -1, or the spring's position in its tile's
    // actuator list (a tile build that lists actuators), else the body-local
    // actuator index
    std::vector<int32_t> act;
};

// Mark the actuated incidences (sb.act[e] >= 0) for the kernels: the name's
// top bit is set (bit 31 of a gather slot; bit 15 of a packed 16-bit row
// when `packed`) and the rest-length entry takes act0 + sb.act[e] as an
// int64 in its 8 bytes: the actuator's arena index (act0 = the body's first
// actuator), or its position in the tile's list (act0 = 0).
void mark_actuated(SellBuild &sb, bool packed, int64_t act0);

// Order of the masses along a Morton curve of their positions (3 doubles
// each).  Returns perm with perm[new] = old; ties keep the original order.
std::vector<int32_t> morton_order(const double *pos, int64_t n);

// Cut a body of n masses (device order) into tiles of whole chunks (see
// layout.h): each tile takes the most chunks (up to 4) whose masses, halo
// and actuator rows fit tile_rows LDS rows.  Lists each tile's halo and,
// when the body has actuated springs, its actuators.  Returns false (with
// tb.why) if one chunk with its halo and actuators exceeds kMaxTileRows.
bool build_tiles(int64_t n, const SpringsIn &sp, TileBuild &tb, int tile_rows = kTileRows);

// Reorder the masses inside each tile (tb; null: the whole body as one tile)
// by descending degree, ties in their current order.  A tile's masses are all
// in LDS, so their order inside it changes no locality, and chunks of even
// degree waste fewer SELL padding entries.  Returns the new order (new ->
// current index; empty when nothing moves) and rewrites tb's halo lists in
// the new indices (each still ascending); the tiles keep their chunks.
std::vector<int32_t> degree_sort_tiles(int64_t n, const SpringsIn &sp, TileBuild *tb);

// Build the incidence lists.  With tb == nullptr (gather layout) an
// incidence names the other endpoint's arena slot, g0 + its body index;
// with a tile build it names the endpoint's LDS row in the mass's tile (and
// an actuated one its spring's position in the tile's actuator list, when
// the tiles list them).
// per_k / per_d: also list each incidence's spring constant / damping.
void build_sell(int64_t n, const SpringsIn &sp, bool per_k, bool per_d, const TileBuild *tb, int64_t g0,
                SellBuild &out);

}  // namespace titan_mi355x
