// Host check of the tiled-layout builder (titan_amd/csrc/tiles.cpp): for
// random graphs and lattices, every tile's spring list, halo, endpoint
// encoding and per-mass incidence order are compared with brute force.
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

static void check(int64_t n, const std::vector<int32_t> &L, const std::vector<int32_t> &R, bool kd) {
    const int64_t S = (int64_t)L.size();
    std::vector<double> k(S), rest(S), damp(S);
    for (int64_t s = 0; s < S; ++s) {
        k[s] = 100.0 + s;
        rest[s] = 0.5 + 0.001 * s;
        damp[s] = 0.01 * (s % 7);
    }
    SpringsIn sp;
    sp.n_springs = S;
    sp.left = L.data();
    sp.right = R.data();
    sp.k = k.data();
    sp.rest = rest.data();
    sp.damping = damp.data();
    TileBuild tb;
    std::vector<int32_t> deg;
    CHECK(build_tiles(n, sp, kd, tb, deg));
    const int64_t ntiles = (n + kTileMasses - 1) / kTileMasses;
    CHECK((int64_t)tb.tiles.size() == ntiles);
    for (int64_t t = 0; t < ntiles; ++t) {
        const TileDesc &d = tb.tiles[t];
        const int64_t m0 = t * kTileMasses, m1 = std::min<int64_t>(n, m0 + kTileMasses);
        CHECK(d.nmass == m1 - m0 && d.chunk0 == m0 / kChunk && d.first == (t == 0));
        std::vector<int32_t> want;
        std::set<int32_t> halo;
        for (int64_t s = 0; s < S; ++s) {
            if (L[s] < 0 || R[s] < 0) continue;
            const bool in = (L[s] >= m0 && L[s] < m1) || (R[s] >= m0 && R[s] < m1);
            if (!in) continue;
            want.push_back((int32_t)s);
            for (int32_t e : {L[s], R[s]})
                if (e < m0 || e >= m1) halo.insert(e);
        }
        CHECK(d.nspr == (int32_t)want.size() && d.nhalo == (int32_t)halo.size());
        std::vector<int32_t> hv(halo.begin(), halo.end());
        for (int32_t h = 0; h < d.nhalo; ++h) CHECK(tb.halo[d.halo_off + h] == hv[h]);
        auto decode = [&](uint32_t l) -> int64_t { return l < (uint32_t)kTileMasses ? m0 + l : tb.halo[d.halo_off + l - kTileMasses]; };
        for (int32_t q = 0; q < d.nspr; ++q) {
            const int32_t s = tb.sid[d.spr_off + q];
            CHECK(s == want[q]);
            const uint32_t e = tb.ends[d.spr_off + q];
            CHECK(decode(e >> 16) == L[s] && decode(e & 0xffff) == R[s]);
            CHECK(tb.rest[d.spr_off + q] == rest[s]);
            if (kd) CHECK(tb.k[d.spr_off + q] == k[s] && tb.damp[d.spr_off + q] == damp[s]);
        }
        for (int64_t m = m0; m < m1; ++m) {
            std::vector<std::pair<int32_t, int>> inc;  // (spring, +1 right / -1 left)
            for (int64_t s = 0; s < S; ++s) {
                if (L[s] < 0 || R[s] < 0) continue;
                if (R[s] == m) inc.push_back({(int32_t)s, 1});
                if (L[s] == m) inc.push_back({(int32_t)s, -1});
            }
            CHECK(deg[m] == (int32_t)inc.size());
            const int64_t c = (m - m0) / kChunk, lane = (m - m0) % kChunk;
            int prev = -1;
            for (size_t j = 0; j < inc.size(); ++j) {
                const uint64_t g = tb.inc[d.inc_off + (int64_t)(d.gstart[c] + j / 4) * kChunk + lane];
                const uint32_t ent = (uint32_t)(g >> (16 * (j % 4))) & 0xffff;
                const int slot = (int)(ent & 0x7fff);
                CHECK(slot < d.nspr && slot >= prev);
                prev = slot;
                CHECK(tb.sid[d.spr_off + slot] == inc[j].first);
                CHECK(((ent & kIncLeft) != 0) == (inc[j].second < 0));
            }
        }
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
        std::vector<int32_t> deg;
        CHECK(build_tiles(n, sp, false, tb, deg));
        CHECK(tb.hmax < 2 * kTileMasses);
    }
    std::printf(fails ? "FAILED %d\n" : "OK\n", fails);
    return fails ? 1 : 0;
}
