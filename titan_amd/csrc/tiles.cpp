// Host-side builders of the tiled layout (see layout.h and k_step_tiles).
#include "tiles.h"

#include <algorithm>
#include <cstring>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace titan_mi355x {

namespace {

// spread the low 21 bits of v so that there are two zero bits between each
uint64_t spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

}  // namespace

std::vector<int32_t> morton_order(const double *pos, int64_t n) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const double x = pos[3 * i + c];
            if (std::isfinite(x)) {
                lo[c] = std::min(lo[c], x);
                hi[c] = std::max(hi[c], x);
            }
        }
    double scale[3];
    for (int c = 0; c < 3; ++c) {
        const double ext = hi[c] - lo[c];
        scale[c] = (std::isfinite(ext) && ext > 0) ? ((double)((1 << 21) - 1) / ext) : 0.0;
        if (!std::isfinite(lo[c])) lo[c] = 0.0;
    }
    std::vector<std::pair<uint64_t, int32_t>> key(n);
    for (int64_t i = 0; i < n; ++i) {
        uint64_t code = 0;
        for (int c = 0; c < 3; ++c) {
            double q = (pos[3 * i + c] - lo[c]) * scale[c];
            if (!(q >= 0)) q = 0;  // NaN or below
            if (q > (double)((1 << 21) - 1)) q = (double)((1 << 21) - 1);
            code |= spread3((uint64_t)q) << c;
        }
        key[i] = {code, (int32_t)i};
    }
    std::sort(key.begin(), key.end());
    std::vector<int32_t> perm(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = key[i].second;
    return perm;
}

namespace {

// Each mass's incident springs in ascending spring index, as CSR: for every
// live spring s (ascending), its right end lists (L, s), then its left end
// lists (R, s), so that a self-spring appears twice, right end first.
// mark_act: the other endpoint's top bit marks an actuated spring (kMarkAct),
// which costs the fill pass no array of its own.
struct Incidence {
    std::vector<int64_t> row;    // n + 1 offsets
    std::vector<int32_t> other;  // the other endpoint
    std::vector<int64_t> spring;
};
constexpr uint32_t kMarkAct = 0x80000000u;

void incidence_lists(int64_t n, const SpringsIn &sp, Incidence &inc, bool with_springs, bool mark_act = false) {
    const int64_t S = sp.n_springs;
    inc.row.assign(n + 1, 0);
    for (int64_t s = 0; s < S; ++s) {
        if (sp.left[s] < 0 || sp.right[s] < 0) continue;
        inc.row[sp.right[s] + 1]++;
        inc.row[sp.left[s] + 1]++;
    }
    for (int64_t i = 0; i < n; ++i) inc.row[i + 1] += inc.row[i];
    inc.other.resize(inc.row[n]);
    if (with_springs) inc.spring.resize(inc.row[n]);
    std::vector<int64_t> fill(inc.row.begin(), inc.row.end() - 1);
    const bool marks = mark_act && sp.act;
    for (int64_t s = 0; s < S; ++s) {
        const int32_t L = sp.left[s], R = sp.right[s];
        if (L < 0 || R < 0) continue;
        const uint32_t mk = marks && sp.act[s] >= 0 ? kMarkAct : 0u;
        const int64_t a = fill[R]++;
        inc.other[a] = (int32_t)((uint32_t)L | mk);
        const int64_t b = fill[L]++;
        inc.other[b] = (int32_t)((uint32_t)R | mk);
        if (with_springs) {
            inc.spring[a] = s;
            inc.spring[b] = s;
        }
    }
}

}  // namespace

bool build_tiles(int64_t n, const SpringsIn &sp, TileBuild &tb, int tile_rows) {
    const bool acts = sp.act != nullptr;
    Incidence inc;
    incidence_lists(n, sp, inc, false, acts);
    const std::vector<int64_t> &row = inc.row;
    const std::vector<int32_t> &other = inc.other;
    tb.tiles.clear();
    tb.halo.clear();
    tb.tact.clear();
    tb.acts = acts;
    tb.hmax = 0;
    tb.rmax = 0;
    std::vector<int32_t> best;
    std::vector<int64_t> stamp(n, -1);
    int64_t mark = 0, nact_all = 0;
    for (int64_t m0 = 0; m0 < n;) {
        // grow the tile a chunk at a time (up to 4, at least one) while its
        // masses, halo and actuator rows fit tile_rows, counting the halo
        // incrementally (a joining mass leaves it, the joining masses'
        // outside neighbours enter it) and the actuators too: a joining
        // mass's actuated spring is new to the tile unless its other end
        // joined earlier, and one between two joining masses counts at its
        // lower end (a self-spring's two listings count half each; acnt2
        // counts halves)
        const int64_t mk = ++mark;
        int64_t cnt = 0, acnt2 = 0, e_prev = m0, m1 = m0;
        for (int k = 1; k <= kTileMasses / kChunk; ++k) {
            if (k > 1 && m0 + (int64_t)(k - 1) * kChunk >= n) break;
            const int64_t e = std::min(n, m0 + (int64_t)k * kChunk);
            for (int64_t m = e_prev; m < e; ++m)
                if (stamp[m] == mk) {
                    stamp[m] = -1;
                    --cnt;
                }
            for (int64_t m = e_prev; m < e; ++m)
                for (int64_t j = row[m]; j < row[m + 1]; ++j) {
                    const int32_t o = other[j] & (int32_t)~kMarkAct;
                    const bool out = o < m0 || o >= e;
                    if (out && stamp[o] != mk) {
                        stamp[o] = mk;
                        ++cnt;
                    }
                    if (other[j] < 0) {  // actuated
                        if (out) acnt2 += 2;
                        else if (o >= e_prev) acnt2 += o > m ? 2 : o == m ? 1 : 0;
                        // else its other end joined earlier and counted it
                    }
                }
            if (k > 1 && e - m0 + cnt + tile_act_rows(acnt2 / 2) > tile_rows) break;
            m1 = e;
            e_prev = e;
        }
        // the chosen range's halo, ascending, and its actuators (listed
        // below), counted once each: from outside, or at the lower end
        const int64_t mh = ++mark;
        best.clear();
        int64_t nact2 = 0;
        for (int64_t m = m0; m < m1; ++m)
            for (int64_t j = row[m]; j < row[m + 1]; ++j) {
                const int32_t o = other[j] & (int32_t)~kMarkAct;
                const bool out = o < m0 || o >= m1;
                if (out && stamp[o] != mh) {
                    stamp[o] = mh;
                    best.push_back(o);
                }
                if (other[j] < 0) nact2 += (out || o > m) ? 2 : o == m ? 1 : 0;
            }
        std::sort(best.begin(), best.end());
        const int64_t nact = nact2 / 2;
        const int64_t rows = m1 - m0 + (int64_t)best.size() + tile_act_rows(nact);
        if (rows > kMaxTileRows) {
            tb.why = "a chunk with its halo and actuators needs more than " + std::to_string(kMaxTileRows) +
                     " LDS rows";
            return false;
        }
        tb.hmax = std::max<int32_t>(tb.hmax, (int32_t)best.size());
        tb.rmax = std::max<int32_t>(tb.rmax, (int32_t)rows);
        TileDesc d{};
        d.body = -1;
        d.chunk0 = (int32_t)(m0 / kChunk);
        d.nmass = (int32_t)(m1 - m0);
        d.nhalo = (int32_t)best.size();
        d.halo_off = (int64_t)tb.halo.size();
        d.first = m0 == 0;
        d.nact = (int32_t)nact;
        d.act_off = nact_all;
        nact_all += nact;
        tb.halo.insert(tb.halo.end(), best.begin(), best.end());
        tb.tiles.push_back(d);
        m0 = m1;
    }
    // the tile of every chunk
    tb.tile_of_chunk.assign((n + kChunk - 1) / kChunk, -1);
    for (size_t q = 0; q < tb.tiles.size(); ++q) {
        const TileDesc &d = tb.tiles[q];
        for (int64_t c = 0; c < (d.nmass + kChunk - 1) / kChunk; ++c) tb.tile_of_chunk[d.chunk0 + c] = (int32_t)q;
    }
    // the actuator lists: the springs in ascending order, each appended to
    // the lists of its endpoints' tiles (so every list comes out ascending)
    if (acts) {
        tb.tact.assign(nact_all, -1);
        std::vector<int64_t> at(tb.tiles.size());
        for (size_t q = 0; q < tb.tiles.size(); ++q) at[q] = tb.tiles[q].act_off;
        for (int64_t s = 0; s < sp.n_springs; ++s) {
            const int32_t L = sp.left[s], R = sp.right[s];
            if (sp.act[s] < 0 || L < 0 || R < 0) continue;
            const int32_t a = tb.tile_of_chunk[L / kChunk], b = tb.tile_of_chunk[R / kChunk];
            tb.tact[at[a]++] = s;
            if (b != a) tb.tact[at[b]++] = s;
        }
        for (size_t q = 0; q < tb.tiles.size(); ++q)
            if (at[q] != tb.tiles[q].act_off + tb.tiles[q].nact) throw std::logic_error("build_tiles: tile actuator count");
    }
    return true;
}

void build_sell(int64_t n, const SpringsIn &sp, bool per_k, bool per_d, const TileBuild *tb, int64_t g0,
                SellBuild &out) {
    const int64_t nch = (n + kChunk - 1) / kChunk;
    const int G = tb ? kGroupTiled : kGroupGather;
    Incidence inc;
    incidence_lists(n, sp, inc, true);
    out.deg.assign(nch * kChunk, 0);
    for (int64_t m = 0; m < n; ++m) out.deg[m] = (int32_t)(inc.row[m + 1] - inc.row[m]);
    out.coff.assign(nch + 1, 0);
    for (int64_t c = 0; c < nch; ++c) {
        int32_t w = 0;
        for (int l = 0; l < kChunk; ++l) w = std::max(w, out.deg[c * kChunk + l]);
        out.coff[c + 1] = out.coff[c] + (int64_t)((w + G - 1) / G) * G * kChunk;
    }
    const int64_t ne = out.coff[nch];
    out.nbr.assign(ne, 0);
    out.rest.assign(ne, 1.0);
    out.k.clear();
    out.damp.clear();
    out.act.clear();
    if (per_k) out.k.assign(ne, 0.0);
    if (per_d) out.damp.assign(ne, 0.0);
    if (sp.act) out.act.assign(ne, -1);
    // tiled: the LDS row of every mass the current tile names (its masses,
    // then its halo), and the list position of each of its actuators, set
    // when the walk enters the tile
    std::vector<uint32_t> rowmap(tb ? n : 0);
    const bool tile_acts = tb && tb->acts && sp.act;
    std::vector<int32_t> aloc(tile_acts ? sp.n_springs : 0);
    int32_t cur_tile = -1;
    for (int64_t c = 0; c < nch; ++c) {
        const int64_t c0 = out.coff[c];
        const int64_t W = (out.coff[c + 1] - c0) / kChunk;  // entries per lane
        int64_t m0 = 0;
        if (tb) {
            const int32_t q = tb->tile_of_chunk[c];
            const TileDesc &d = tb->tiles[q];
            m0 = (int64_t)d.chunk0 * kChunk;
            if (q != cur_tile) {
                cur_tile = q;
                for (int32_t r = 0; r < d.nmass; ++r) rowmap[m0 + r] = (uint32_t)r;
                const int32_t *h = tb->halo.data() + d.halo_off;
                for (int32_t r = 0; r < d.nhalo; ++r) rowmap[h[r]] = (uint32_t)(d.nmass + r);
                if (tile_acts)
                    for (int32_t r = 0; r < d.nact; ++r) aloc[tb->tact[d.act_off + r]] = r;
            }
        }
        for (int l = 0; l < kChunk; ++l) {
            const int64_t me = c * kChunk + l;
            const int64_t d = me < n ? inc.row[me + 1] - inc.row[me] : 0;
            const int64_t i0 = me < n ? inc.row[me] : 0;
            // padding entries name the mass itself (tiled: row 0 past the body)
            const uint32_t pad = tb ? (uint32_t)(me < n ? me - m0 : 0) : (uint32_t)(g0 + me);
            // entry of incidence j: c0 + ((j / G) * 64 + l) * G + j % G
            for (int64_t j = 0; j < W; ++j) {
                const int64_t e = c0 + ((j / G) * kChunk + l) * G + j % G;
                uint32_t name = pad;
                if (j < d) {
                    const int32_t o = inc.other[i0 + j];
                    const int64_t s = inc.spring[i0 + j];
                    name = tb ? rowmap[o] : (uint32_t)(g0 + o);
                    out.rest[e] = sp.rest[s];
                    if (per_k) out.k[e] = sp.k[s];
                    if (per_d) out.damp[e] = sp.damping ? sp.damping[s] : 0.0;
                    if (sp.act) out.act[e] = (tile_acts && sp.act[s] >= 0) ? aloc[s] : sp.act[s];
                }
                if (tb) {  // the LDS rows of a group's two entries packed at c0 + (e - c0) / 2
                    const int64_t rel = e - c0;
                    out.nbr[c0 + rel / 2] |= name << (16 * (int)(rel % 2));
                } else {
                    out.nbr[e] = name;
                }
            }
        }
    }
}

std::vector<int32_t> degree_sort_tiles(int64_t n, const SpringsIn &sp, TileBuild *tb) {
    std::vector<int32_t> deg(n, 0);
    for (int64_t s = 0; s < sp.n_springs; ++s) {
        if (sp.left[s] < 0 || sp.right[s] < 0) continue;
        deg[sp.left[s]]++;
        deg[sp.right[s]]++;
    }
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    auto sort_range = [&](int64_t m0, int64_t m1) {
        std::stable_sort(order.begin() + m0, order.begin() + m1,
                         [&](int32_t a, int32_t b) { return deg[a] > deg[b]; });
    };
    if (tb) {
        for (const TileDesc &d : tb->tiles) sort_range((int64_t)d.chunk0 * kChunk, (int64_t)d.chunk0 * kChunk + d.nmass);
    } else {
        sort_range(0, n);
    }
    bool moved = false;
    for (int64_t q = 0; q < n && !moved; ++q) moved = order[q] != q;
    if (!moved) return {};
    if (tb) {
        std::vector<int32_t> inv(n);
        for (int64_t q = 0; q < n; ++q) inv[order[q]] = (int32_t)q;
        for (const TileDesc &d : tb->tiles) {
            int32_t *h = tb->halo.data() + d.halo_off;
            for (int32_t j = 0; j < d.nhalo; ++j) h[j] = inv[h[j]];
            std::sort(h, h + d.nhalo);
        }
    }
    return order;
}

void mark_actuated(SellBuild &sb, bool packed, int64_t act0) {
    if (sb.act.empty()) return;
    const int64_t nch = (int64_t)sb.coff.size() - 1;
    for (int64_t c = 0; c < nch; ++c)
        for (int64_t e = sb.coff[c]; e < sb.coff[c + 1]; ++e) {
            if (sb.act[e] < 0) continue;
            const int64_t idx = act0 + sb.act[e];
            std::memcpy(&sb.rest[e], &idx, sizeof(idx));
            if (packed) {
                const int64_t rel = e - sb.coff[c];
                sb.nbr[sb.coff[c] + rel / 2] |= kActBit16 << (16 * (int)(rel % 2));
            } else {
                sb.nbr[e] |= kActBit32;
            }
        }
}

}  // namespace titan_mi355x
