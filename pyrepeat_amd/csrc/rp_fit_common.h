// Device-side Ewald data and helpers shared by the fp64 and split accumulation paths.
#pragma once
#include <cmath>

// Grid coordinates and distances follow the reference roundings (no FMA contraction).
#pragma clang fp contract(off)

#include "rp_internal.h"

namespace rp {

// Cube geometry: flat index -> (i, j, k) -> Cartesian point ((i*a0 + j*a1) + k*a2).
struct FitGeom {
    int nx, ny, nz;
    double ax[9];
    __device__ __forceinline__ void ijk(int f, int &i, int &j, int &k) const {
        const int layer = nx * ny;
        k = f / layer;
        const int rem = f - k * layer;
        j = rem / nx;
        i = rem - j * nx;
    }
    __device__ __forceinline__ void xyz(int f, double &x, double &y, double &z) const {
        int i, j, k;
        ijk(f, i, j, k);
        x = ((double)i * ax[0] + (double)j * ax[3]) + (double)k * ax[6];
        y = ((double)i * ax[1] + (double)j * ax[4]) + (double)k * ax[7];
        z = ((double)i * ax[2] + (double)j * ax[5]) + (double)k * ax[8];
    }
};

// Real-space sum and the augmented columns, evaluated per basis element.
struct EpiArgs {
    FitGeom geom;
    const int *idx;
    const double *v;
    int64_t row0, rows_valid;
    int N;
    int fit_type;
    const double *atoms;   // N x 3
    const double *shifts;  // S x 3 (pruned)
    int S;
    double R2, sqrt_alpha;

    // sum_s erfc(sqrt_alpha r)/r over r^2 <= Rc^2, zero shift skipped at r^2 <= 1e-20
    // (repeat.py:849-872); fit-type 0: 1/|d| for |d| > 1e-10 (repeat.py:991-994).
    __device__ __forceinline__ double real_space(double gx, double gy, double gz, int a) const {
        const double ax = atoms[3 * a], ay = atoms[3 * a + 1], az = atoms[3 * a + 2];
        if (fit_type == 0) {
            const double dx = gx - ax, dy = gy - ay, dz = gz - az;
            const double d = sqrt((dx * dx + dy * dy) + dz * dz);
            return d > 1e-10 ? 1.0 / d : 0.0;
        }
        double sum = 0.0;
        for (int s = 0; s < S; ++s) {
            const double sx = shifts[3 * s], sy = shifts[3 * s + 1], sz = shifts[3 * s + 2];
            const double dx = gx - (ax + sx), dy = gy - (ay + sy), dz = gz - (az + sz);
            const double r2 = (dx * dx + dy * dy) + dz * dz;
            if (r2 <= R2) {
                if (sx == 0.0 && sy == 0.0 && sz == 0.0 && r2 <= 1e-20) continue;
                const double r = sqrt(r2);
                sum += erfc(sqrt_alpha * r) / r;
            }
        }
        return sum;
    }
};

// Device copies of the Ewald setup and the chunking plan of one fit.
struct FitDevice {
    int K = 0, KP = 0, NP = 0, S = 0;
    int64_t rows = 0;      // points per chunk (multiple of 128)
    int splits = 1;        // Gram row splits per chunk
    FitGeom geom{};
    EpiArgs epi{};
    DBuf khkl, kcoef, kvecs, ex, ey, ez, shifts, d_tiles;
    std::vector<int2> tiles;
};

int prepare_device(rp_fit *fit, FitDevice &dv, int precision);
int accumulate_f64(rp_fit *fit, FitDevice &dv);
int accumulate_split(rp_fit *fit, FitDevice &dv);
__global__ void k_symmetrize(double *gram, int NP, int tile);

}  // namespace rp
