// Host check of the layout builders (titan_amd/csrc/tiles.cpp): for random
// graphs and a Morton-ordered lattice, every tile's halo and actuator list
// and every mass's incidence list (other endpoint, spring parameters,
// actuator, ascending spring order, padding) in both layouts are compared
// with brute force.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <vector>

#include "tiles.h"

using namespace titan_mi355x;

static int fails = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);       \
            if (++fails > 20) std::exit(1);                               \
        }                                                                 \
    } while (0)

// Decode mass m's incidence list from a SELL build and compare it with the
// brute-force list of (other endpoint, spring) in ascending spring index.
static void check_sell(int64_t n, const std::vector<int32_t> &L, const std::vector<int32_t> &R,
                       const std::vector<double> &k, const std::vector<double> &rest,
                       const std::vector<double> &damp, const std::vector<int32_t> &act, bool kd,
                       const TileBuild *tb, int64_t g0, const SellBuild &sb) {
    const int64_t S = (int64_t)L.size();
    const int64_t nch = (n + kChunk - 1) / kChunk;
    CHECK((int64_t)sb.coff.size() == nch + 1 && (int64_t)sb.deg.size() == nch * kChunk);
    CHECK((int64_t)sb.nbr.size() == sb.coff[nch] && (int64_t)sb.rest.size() == sb.coff[nch]);
    CHECK(kd ? (int64_t)sb.k.size() == sb.coff[nch] : sb.k.empty());
    std::vector<std::vector<std::pair<int32_t, int32_t>>> want(n);  // (other, spring)
    for (int64_t s = 0; s < S; ++s) {
        if (L[s] < 0 || R[s] < 0) continue;
        want[R[s]].push_back({L[s], (int32_t)s});
        want[L[s]].push_back({R[s], (int32_t)s});
    }
    const int G = tb ? kGroupTiled : kGroupGather;
    // the name stored for entry e of chunk c
    auto name = [&](int64_t c, int64_t e) -> uint32_t {
        if (!tb) return sb.nbr[e];
        const int64_t rel = e - sb.coff[c];
        return (sb.nbr[sb.coff[c] + rel / 2] >> (16 * (rel % 2))) & 0xffffu;
    };
    auto decode = [&](int64_t me, uint32_t v) -> int64_t {
        if (!tb) return (int64_t)v - g0;
        const TileDesc &d = tb->tiles[tb->tile_of_chunk[me / kChunk]];
        const int64_t m0 = (int64_t)d.chunk0 * kChunk;
        CHECK(me >= m0 && me < m0 + d.nmass && me < n);
        if (v < (uint32_t)d.nmass) return m0 + v;
        CHECK(v - d.nmass < (uint32_t)d.nhalo);
        return tb->halo[d.halo_off + v - d.nmass];
    };
    for (int64_t c = 0; c < nch; ++c) {
        int32_t w = 0;
        for (int l = 0; l < kChunk; ++l) w = std::max(w, sb.deg[c * kChunk + l]);
        CHECK(sb.coff[c + 1] - sb.coff[c] == (int64_t)((w + G - 1) / G) * G * kChunk);
        for (int l = 0; l < kChunk; ++l) {
            const int64_t m = c * kChunk + l;
            const int32_t dm = m < n ? (int32_t)want[m].size() : 0;
            CHECK(sb.deg[m] == dm);
            for (int64_t j = 0; j < (sb.coff[c + 1] - sb.coff[c]) / kChunk; ++j) {
                const int64_t e = sb.coff[c] + ((j / G) * kChunk + l) * G + j % G;
                if (j >= dm) {  // padding: the mass itself (row 0 past the body), rest 1
                    if (m < n) CHECK(decode(m, name(c, e)) == m);
                    else if (tb) CHECK(name(c, e) == 0);
                    CHECK(sb.rest[e] == 1.0);
                    continue;
                }
                const int32_t s = want[m][j].second;
                CHECK(decode(m, name(c, e)) == want[m][j].first);
                CHECK(sb.rest[e] == rest[s]);
                if (kd) CHECK(sb.k[e] == k[s] && sb.damp[e] == damp[s]);
                if (!act.empty() && tb && tb->acts) {  // the spring's position in its tile's list
                    const TileDesc &d = tb->tiles[tb->tile_of_chunk[c]];
                    if (act[s] < 0) {
                        CHECK(sb.act[e] == -1);
                    } else {
                        CHECK(sb.act[e] >= 0 && sb.act[e] < d.nact);
                        if (sb.act[e] >= 0 && sb.act[e] < d.nact) CHECK(tb->tact[d.act_off + sb.act[e]] == s);
                    }
                } else if (!act.empty()) {
                    CHECK(sb.act[e] == act[s]);
                }
            }
        }
    }
}

static void check(int64_t n, const std::vector<int32_t> &L, const std::vector<int32_t> &R, bool kd) {
    const int64_t S = (int64_t)L.size();
    std::vector<double> k(S), rest(S), damp(S);
    std::vector<int32_t> act(S);
    for (int64_t s = 0; s < S; ++s) {
        k[s] = 100.0 + s;
        rest[s] = 0.5 + 0.001 * s;
        damp[s] = 0.01 * (s % 7);
        act[s] = s % 3 == 0 ? (int32_t)(s / 3) : -1;
    }
    SpringsIn sp;
    sp.n_springs = S;
    sp.left = L.data();
    sp.right = R.data();
    sp.k = k.data();
    sp.rest = rest.data();
    sp.damping = damp.data();
    sp.act = kd ? act.data() : nullptr;
    const std::vector<int32_t> noact;
    // gather layout: arena slots
    SellBuild sg;
    build_sell(n, sp, kd, kd, nullptr, 4096, sg);
    check_sell(n, L, R, k, rest, damp, kd ? act : noact, kd, nullptr, 4096, sg);
    // tiled layout: tiles, halos and LDS rows, under two row budgets, then
    // with the masses of each tile reordered by degree
    auto verify_tiles = [&](const std::vector<int32_t> &L, const std::vector<int32_t> &R, const TileBuild &tb,
                            int rows) {
    int32_t hmax = 0, rmax = 0;
    int64_t next = 0;  // tiles cover the masses in order, chunk-aligned
    for (size_t t = 0; t < tb.tiles.size(); ++t) {
        const TileDesc &d = tb.tiles[t];
        const int64_t m0 = (int64_t)d.chunk0 * kChunk, m1 = m0 + d.nmass;
        CHECK(m0 == next && m1 <= n && d.nmass > 0 && d.nmass <= kTileMasses && d.first == (t == 0));
        CHECK(m1 == n || d.nmass % kChunk == 0);
        CHECK(d.act_off == (t == 0 ? 0 : tb.tiles[t - 1].act_off + tb.tiles[t - 1].nact));
        next = m1;
        std::set<int32_t> halo;
        std::vector<int64_t> acts;  // ascending
        for (int64_t s = 0; s < S; ++s) {
            if (L[s] < 0 || R[s] < 0) continue;
            const bool in = (L[s] >= m0 && L[s] < m1) || (R[s] >= m0 && R[s] < m1);
            if (!in) continue;
            for (int32_t e : {L[s], R[s]})
                if (e < m0 || e >= m1) halo.insert(e);
            if (sp.act && sp.act[s] >= 0) acts.push_back(s);
        }
        CHECK(d.nhalo == (int32_t)halo.size());
        CHECK(tb.acts == (sp.act != nullptr));
        CHECK(d.nact == (int32_t)acts.size());
        if (!tb.tact.empty() && d.nact == (int32_t)acts.size())  // listed (after build_sell)
            for (int32_t q = 0; q < d.nact; ++q) CHECK(tb.tact[d.act_off + q] == acts[q]);
        const int64_t r = d.nmass + d.nhalo + tile_act_rows(d.nact);
        // within the row budget, unless the tile is a single chunk
        CHECK(r <= rows || d.nmass <= kChunk);
        CHECK(r <= kMaxTileRows);
        hmax = std::max<int32_t>(hmax, d.nhalo);
        rmax = std::max<int32_t>(rmax, (int32_t)r);
        std::vector<int32_t> hv(halo.begin(), halo.end());
        for (int32_t h = 0; h < d.nhalo; ++h) CHECK(tb.halo[d.halo_off + h] == hv[h]);
        for (int64_t q = m0; q < m1; ++q) CHECK(tb.tile_of_chunk[q / kChunk] == (int32_t)t);
    }
    CHECK(next == n);
    if (!tb.tact.empty())
        CHECK((int64_t)tb.tact.size() == tb.tiles.back().act_off + tb.tiles.back().nact);
    CHECK(tb.rmax == rmax);
    CHECK(tb.hmax == hmax);
    };
    for (int rows : {kTileRows, 682}) {
    TileBuild tb;
    CHECK(build_tiles(n, sp, tb, rows));
    verify_tiles(L, R, tb, rows);
    SellBuild st;
    build_sell(n, sp, kd, kd, &tb, 0, st);
    verify_tiles(L, R, tb, rows);  // now with the actuator lists
    if (kd) {
        int64_t want = 0;
        for (const TileDesc &d : tb.tiles) want += d.nact;
        CHECK((int64_t)tb.tact.size() == want);
    }
    check_sell(n, L, R, k, rest, damp, kd ? act : noact, kd, &tb, 0, st);
    // reorder each tile's masses by degree: a permutation inside each tile,
    // degrees non-increasing there, halos rewritten in the new indices
    std::vector<int32_t> order = degree_sort_tiles(n, sp, &tb);
    if (order.empty()) continue;
    std::vector<int32_t> inv(n, -1), deg(n, 0), L2(S), R2(S);
    for (int64_t s = 0; s < S; ++s)
        if (L[s] >= 0 && R[s] >= 0) {
            deg[L[s]]++;
            deg[R[s]]++;
        }
    for (int64_t q = 0; q < n; ++q) {
        CHECK(order[q] >= 0 && order[q] < n && inv[order[q]] < 0);
        CHECK(tb.tile_of_chunk[q / kChunk] == tb.tile_of_chunk[order[q] / kChunk]);
        inv[order[q]] = (int32_t)q;
        if (q > 0 && tb.tile_of_chunk[q / kChunk] == tb.tile_of_chunk[(q - 1) / kChunk])
            CHECK(deg[order[q - 1]] >= deg[order[q]]);
    }
    for (int64_t s = 0; s < S; ++s) {
        L2[s] = L[s] < 0 ? -1 : inv[L[s]];
        R2[s] = R[s] < 0 ? -1 : inv[R[s]];
    }
    verify_tiles(L2, R2, tb, rows);
    SpringsIn sp2 = sp;
    sp2.left = L2.data();
    sp2.right = R2.data();
    SellBuild s2;
    build_sell(n, sp2, kd, kd, &tb, 0, s2);
    verify_tiles(L2, R2, tb, rows);
    check_sell(n, L2, R2, k, rest, damp, kd ? act : noact, kd, &tb, 0, s2);
    }
}

int main() {
    std::mt19937_64 rng(42);
    // random graphs with duplicates, self springs and dead springs
    for (int rep = 0; rep < 6; ++rep) {
        const int64_t n = 1 + (int64_t)(rng() % 900);
        const int64_t S = (int64_t)(rng() % (8 * n + 1));
        std::vector<int32_t> L(S), R(S);
        for (int64_t s = 0; s < S; ++s) {
            L[s] = (int32_t)(rng() % n);
            R[s] = (rng() % 11 == 0) ? L[s] : (int32_t)(rng() % n);
            if (rng() % 37 == 0) L[s] = -1;
        }
        check(n, L, R, rep % 2 == 0);
    }
    // a 9x10x11 lattice-like graph in Morton order
    {
        const int nx = 9, ny = 10, nz = 11;
        std::vector<double> pos;
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) pos.insert(pos.end(), {(double)i, (double)j, (double)k});
        const int64_t n = (int64_t)nx * ny * nz;
        std::vector<int32_t> perm = morton_order(pos.data(), n), inv(n);
        std::vector<bool> seen(n, false);
        for (int64_t q = 0; q < n; ++q) {
            CHECK(perm[q] >= 0 && perm[q] < n && !seen[perm[q]]);
            seen[perm[q]] = true;
            inv[perm[q]] = (int32_t)q;
        }
        std::vector<int32_t> L, R;
        auto at = [&](int i, int j, int k) { return k + j * nz + i * ny * nz; };
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k)
                    for (int di = 0; di <= 1; ++di)
                        for (int dj = -1; dj <= 1; ++dj)
                            for (int dk = -1; dk <= 1; ++dk) {
                                if (di == 0 && (dj < 0 || (dj == 0 && dk <= 0))) continue;
                                const int a = i + di, b = j + dj, cc = k + dk;
                                if (a >= nx || b < 0 || b >= ny || cc < 0 || cc >= nz) continue;
                                L.push_back(inv[at(i, j, k)]);
                                R.push_back(inv[at(a, b, cc)]);
                            }
        check(n, L, R, false);
        // a brick-shaped tile keeps the halo small
        SpringsIn sp;
        std::vector<double> ones(L.size(), 1.0);
        sp.n_springs = (int64_t)L.size();
        sp.left = L.data();
        sp.right = R.data();
        sp.k = ones.data();
        sp.rest = ones.data();
        TileBuild tb;
        CHECK(build_tiles(n, sp, tb));
        CHECK(tb.rmax <= kTileRows);
    }
    std::printf(fails ? "FAILED %d\n" : "OK\n", fails);
    return fails ? 1 : 0;
}
