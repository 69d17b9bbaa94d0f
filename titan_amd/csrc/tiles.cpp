// Host-side builders of the tiled layout (see layout.h and k_step_tiles).
#include "tiles.h"

#include <algorithm>
#include <cmath>
#include <numeric>

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

bool build_tiles(int64_t n, const SpringsIn &sp, bool per_spring_kd, TileBuild &tb,
                 std::vector<int32_t> &deg_out) {
    const int64_t S = sp.n_springs;
    // incident springs per mass, ascending spring index; the right endpoint
    // is listed first for a spring that joins a mass to itself
    std::vector<int64_t> row(n + 1, 0);
    for (int64_t s = 0; s < S; ++s) {
        if (sp.left[s] < 0 || sp.right[s] < 0) continue;
        row[sp.right[s] + 1]++;
        row[sp.left[s] + 1]++;
    }
    for (int64_t i = 0; i < n; ++i) row[i + 1] += row[i];
    std::vector<int32_t> inc_s(row[n]);
    std::vector<int8_t> inc_left(row[n]);
    {
        std::vector<int64_t> fill(row.begin(), row.end() - 1);
        for (int64_t s = 0; s < S; ++s) {
            const int32_t L = sp.left[s], R = sp.right[s];
            if (L < 0 || R < 0) continue;
            inc_s[fill[R]] = (int32_t)s;
            inc_left[fill[R]++] = 0;
            inc_s[fill[L]] = (int32_t)s;
            inc_left[fill[L]++] = 1;
        }
    }
    deg_out.assign(n, 0);
    for (int64_t i = 0; i < n; ++i) deg_out[i] = (int32_t)(row[i + 1] - row[i]);

    const int64_t ntiles = (n + kTileMasses - 1) / kTileMasses;
    std::vector<int32_t> mark(S, -1), slot_of(S, 0);
    std::vector<int32_t> list, halo;
    tb.tiles.clear();
    tb.halo.clear();
    tb.ends.clear();
    tb.rest.clear();
    tb.k.clear();
    tb.damp.clear();
    tb.act.clear();
    tb.inc.clear();
    tb.sid.clear();
    tb.hmax = 0;
    for (int64_t t = 0; t < ntiles; ++t) {
        const int64_t m0 = t * kTileMasses, m1 = std::min(n, m0 + kTileMasses);
        list.clear();
        for (int64_t m = m0; m < m1; ++m)
            for (int64_t j = row[m]; j < row[m + 1]; ++j) {
                const int32_t s = inc_s[j];
                if (mark[s] != (int32_t)t) {
                    mark[s] = (int32_t)t;
                    list.push_back(s);
                }
            }
        std::sort(list.begin(), list.end());
        if ((int64_t)list.size() > kMaxTileSprings) {
            tb.why = "a tile touches more than " + std::to_string(kMaxTileSprings) + " springs";
            return false;
        }
        halo.clear();
        for (size_t q = 0; q < list.size(); ++q) {
            const int32_t s = list[q];
            slot_of[s] = (int32_t)q;
            for (int32_t e : {sp.left[s], sp.right[s]})
                if (e < m0 || e >= m1) halo.push_back(e);
        }
        std::sort(halo.begin(), halo.end());
        halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
        if ((int64_t)halo.size() > kMaxHalo) {
            tb.why = "a tile has more than " + std::to_string(kMaxHalo) + " halo masses";
            return false;
        }
        tb.hmax = std::max<int32_t>(tb.hmax, (int32_t)halo.size());

        TileDesc d{};
        d.body = -1;
        d.chunk0 = (int32_t)(m0 / kChunk);
        d.nmass = (int32_t)(m1 - m0);
        d.nhalo = (int32_t)halo.size();
        d.nspr = (int32_t)list.size();
        d.first = t == 0;
        d.halo_off = (int64_t)tb.halo.size();
        d.spr_off = (int64_t)tb.ends.size();
        d.inc_off = (int64_t)tb.inc.size();
        auto local = [&](int32_t e) -> uint32_t {
            if (e >= m0 && e < m1) return (uint32_t)(e - m0);
            return (uint32_t)(kTileMasses + (std::lower_bound(halo.begin(), halo.end(), e) - halo.begin()));
        };
        tb.sid.insert(tb.sid.end(), list.begin(), list.end());
        for (int32_t s : list) {
            tb.ends.push_back((local(sp.left[s]) << 16) | local(sp.right[s]));
            tb.rest.push_back(sp.rest[s]);
            if (per_spring_kd) {
                tb.k.push_back(sp.k[s]);
                tb.damp.push_back(sp.damping ? sp.damping[s] : 0.0);
            }
            if (sp.act) tb.act.push_back(sp.act[s]);
        }
        tb.halo.insert(tb.halo.end(), halo.begin(), halo.end());
        // incidence groups, chunk by chunk
        int32_t g = 0;
        const int64_t nch = (m1 - m0 + kChunk - 1) / kChunk;
        for (int64_t c = 0; c < 4; ++c) {
            d.gstart[c] = g;
            if (c >= nch) continue;
            int32_t w = 0;
            for (int l = 0; l < kChunk; ++l) {
                const int64_t m = m0 + c * kChunk + l;
                if (m < m1) w = std::max(w, deg_out[m]);
            }
            const int32_t groups = (w + 3) / 4;
            const size_t base = tb.inc.size();
            tb.inc.resize(base + (size_t)groups * kChunk, 0);
            for (int l = 0; l < kChunk; ++l) {
                const int64_t m = m0 + c * kChunk + l;
                if (m >= m1) continue;
                for (int64_t j = row[m]; j < row[m + 1]; ++j) {
                    const int64_t q = j - row[m];
                    const uint64_t ent = (uint64_t)slot_of[inc_s[j]] | (inc_left[j] ? kIncLeft : 0);
                    tb.inc[base + (size_t)(q / 4) * kChunk + l] |= ent << (16 * (q % 4));
                }
            }
            g += groups;
        }
        tb.tiles.push_back(d);
    }
    return true;
}

}  // namespace titan_mi355x
