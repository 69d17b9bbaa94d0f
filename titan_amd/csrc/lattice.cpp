// Host-side lattice generator (C ABI: titan_lattice, titan_lattice_spring_count).
//
// Reproduces the mass and spring order of the reference's Lattice
// constructor (src/object.cu:232-293): mass (i, j, k) has index
// k + j*nz + i*ny*nz and sits at (i/(nx-1) - 0.5, ...) * dims + center; each
// mass emits its 7 forward-corner springs, then the 6 extra diagonals that
// complete the 26-neighbourhood, and every rest length is the initial
// |right - left|.  Used by the C++ facade's createLattice and by the
// benchmark's scene set-up, which generates SoA arrays directly instead of
// one heap object per mass and spring.
#include <cmath>
#include <cstdint>

#include "titan_mi355x.h"

namespace {

struct Emit {
    int32_t *left, *right;
    int64_t n = 0;
    void operator()(int64_t l, int64_t r) {
        if (left) {
            left[n] = (int32_t)l;
            right[n] = (int32_t)r;
        }
        n++;
    }
};

template <class F>
void lattice_springs(int nx, int ny, int nz, F &&emit) {
    auto at = [&](int i, int j, int k) -> int64_t { return k + (int64_t)j * nz + (int64_t)i * ny * nz; };
    for (int i = 0; i < nx; i++) {
        for (int j = 0; j < ny; j++) {
            for (int k = 0; k < nz; k++) {
                const int li = (i != nx - 1) ? 2 : 1, mj = (j != ny - 1) ? 2 : 1, nk = (k != nz - 1) ? 2 : 1;
                for (int l = 0; l < li; l++)
                    for (int m = 0; m < mj; m++)
                        for (int n = 0; n < nk; n++)
                            if (l || m || n) emit(at(i, j, k), at(i + l, j + m, k + n));
                const bool ei = i != nx - 1, ej = j != ny - 1, ek = k != nz - 1;
                if (ek) {
                    if (ej) emit(at(i, j, k + 1), at(i, j + 1, k));
                    if (ei) emit(at(i, j, k + 1), at(i + 1, j, k));
                    if (ej && ei) {
                        emit(at(i, j, k + 1), at(i + 1, j + 1, k));
                        emit(at(i + 1, j, k + 1), at(i, j + 1, k));
                        emit(at(i, j + 1, k + 1), at(i + 1, j, k));
                    }
                }
                if (ej && ei) emit(at(i, j + 1, k), at(i + 1, j, k));
            }
        }
    }
}

}  // namespace

extern "C" {

int64_t titan_lattice_spring_count(int nx, int ny, int nz) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
    Emit e{nullptr, nullptr};
    lattice_springs(nx, ny, nz, e);
    return e.n;
}

int titan_lattice(double cx, double cy, double cz, double dx, double dy, double dz, int nx, int ny,
                  int nz, double *pos, int32_t *left, int32_t *right, double *rest) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !pos) return TITAN_E_INVALID;
    if ((int64_t)nx * ny * nz > (int64_t)0x7fffffff) return TITAN_E_INVALID;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                const int64_t a = k + (int64_t)j * nz + (int64_t)i * ny * nz;
                // Vec(...) * dims + center, element-wise (include/Titan/vec.h:129-131)
                const double fx = (nx > 1) ? (double)i / (nx - 1.0) - 0.5 : 0;
                const double fy = (ny > 1) ? j / (ny - 1.0) - 0.5 : 0;
                const double fz = (nz > 1) ? k / (nz - 1.0) - 0.5 : 0;
                pos[3 * a] = fx * dx + cx;
                pos[3 * a + 1] = fy * dy + cy;
                pos[3 * a + 2] = fz * dz + cz;
            }
    if (!left || !right) return TITAN_OK;
    Emit e{left, right};
    lattice_springs(nx, ny, nz, e);
    if (rest) {
        for (int64_t s = 0; s < e.n; s++) {
            const double *r = pos + 3 * (int64_t)right[s], *l = pos + 3 * (int64_t)left[s];
            const double x = r[0] - l[0], y = r[1] - l[1], z = r[2] - l[2];
            rest[s] = std::sqrt(x * x + y * y + z * z);
        }
    }
    return TITAN_OK;
}

}  // extern "C"
