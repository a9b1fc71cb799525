/*
 * CPU oracle kernels for the REPEAT fit path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.
 * The product (pyrepeat_amd) never links or calls it.
 *
 * Everything here is a plain fp64 restatement of the reference's CPU
 * algorithm, compiled with -ffp-contract=off and without -ffast-math so the
 * arithmetic is IEEE round-to-nearest in the order written.
 *
 *   oracle_grid_point      repeat.py:2139-2151  (base_xy + k*z_vec, per component)
 *   oracle_filter_mask     repeat.py:2115-2180, 2081  (cdist predicate: keep = all(d > rmin) and
 *                          any(d <= rmax))
 *   oracle_phi             repeat.py:831-872, 1486-1515 (per-atom column: blocked cos sum
 *                          over k in blocks of block_k, then erfc real-space sum over the
 *                          shift list; Phi = recip + real)
 *   oracle_phi fit_flag 0  repeat.py:991-994   (1/|d| with d > 1e-10, else 0)
 *
 * Distances follow scipy.spatial.distance.cdist('euclidean') which evaluates
 * sqrt((dx*dx + dy*dy) + dz*dz) left to right; a test pins this bitwise
 * against scipy.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* Grid point (i,j,k): (i*a0 + j*a1) + k*a2 per component (repeat.py:2139-2151, 2096-2108). */
static inline void grid_point(const double *axis, int64_t i, int64_t j, int64_t k, double *g) {
    for (int c = 0; c < 3; ++c) {
        double base = (double)i * axis[0 * 3 + c] + (double)j * axis[1 * 3 + c];
        g[c] = base + (double)k * axis[2 * 3 + c];
    }
}

void oracle_grid_point(const double *axis, int64_t i, int64_t j, int64_t k, double *g) {
    grid_point(axis, i, j, k, g);
}

/*
 * Keep mask over the full cube in flat order f = i + j*nx + k*nx*ny:
 * keep[f] = all(d > rmin) and any(d <= rmax) over the effective atoms.
 * Brute force over every effective atom (the reference's CPU filter at the
 * default vdw-max tests all of them; its z-slab prefilter only drops atoms
 * that cannot satisfy either predicate when the first two step vectors have
 * no z component).
 */
void oracle_filter_mask(int nx, int ny, int nz, const double *axis,
                        int64_t n_eff, const double *eff_xyz,
                        const double *eff_rmin, const double *eff_rmax,
                        uint8_t *keep_out, int nthreads) {
    set_threads(nthreads);
    const int64_t layer = (int64_t)nx * ny;
    const int64_t n_total = layer * nz;
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t f = 0; f < n_total; ++f) {
        int64_t k = f / layer;
        int64_t rem = f - k * layer;
        int64_t j = rem / nx;
        int64_t i = rem - j * nx;
        double g[3];
        grid_point(axis, i, j, k, g);
        int valid = 1, near = 0;
        for (int64_t a = 0; a < n_eff; ++a) {
            double dx = g[0] - eff_xyz[3 * a + 0];
            double dy = g[1] - eff_xyz[3 * a + 1];
            double dz = g[2] - eff_xyz[3 * a + 2];
            double s = dx * dx + dy * dy;
            s = s + dz * dz;
            double d = sqrt(s);
            if (!(d > eff_rmin[a])) { valid = 0; break; }
            if (d <= eff_rmax[a]) near = 1;
        }
        keep_out[f] = (uint8_t)(valid && near);
    }
}

/* Pairwise Euclidean distances, cdist order, for pinning against scipy. */
void oracle_cdist(int64_t m, const double *x, int64_t n, const double *y, double *out) {
    for (int64_t i = 0; i < m; ++i)
        for (int64_t a = 0; a < n; ++a) {
            double dx = x[3 * i + 0] - y[3 * a + 0];
            double dy = x[3 * i + 1] - y[3 * a + 1];
            double dz = x[3 * i + 2] - y[3 * a + 2];
            double s = dx * dx + dy * dy;
            s = s + dz * dz;
            out[i * n + a] = sqrt(s);
        }
}

/*
 * Basis matrix phi[m*N + a] (row-major, M x N) for grid rows [0, M).
 * Reciprocal part: for each block of block_k k-vectors the partial sum
 * acc = sum_j cos(k_j . (g - R)) * w_j is added to phi (repeat.py:831-846).
 * Real part: sum over shifts s of erfc(sqrt_alpha r)/r for r^2 <= Rc^2, the
 * zero shift skipped when r^2 <= 1e-20 (repeat.py:849-872).
 * The column is recip + real (repeat.py:1512).
 */
void oracle_phi(int64_t M, const double *grid_pos, int64_t N, const double *atom_pos,
                int64_t K, const double *kvecs, const double *kcoefs,
                int64_t S, const double *shifts, double r_cutoff, double sqrt_alpha,
                int fit_flag, int block_k, double *phi, int nthreads) {
    set_threads(nthreads);
    if (block_k <= 0) block_k = 64;
    const double R2 = r_cutoff * r_cutoff;
    for (int64_t a = 0; a < N; ++a) {
        const double ax = atom_pos[3 * a + 0], ay = atom_pos[3 * a + 1], az = atom_pos[3 * a + 2];
        if (fit_flag == 0) {
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < M; ++i) {
                double dx = grid_pos[3 * i + 0] - ax;
                double dy = grid_pos[3 * i + 1] - ay;
                double dz = grid_pos[3 * i + 2] - az;
                double s = dx * dx + dy * dy;
                s = s + dz * dz;
                double d = sqrt(s);
                phi[i * N + a] = (d > 1e-10) ? 1.0 / d : 0.0;
            }
            continue;
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < M; ++i) {
            double dx = grid_pos[3 * i + 0] - ax;
            double dy = grid_pos[3 * i + 1] - ay;
            double dz = grid_pos[3 * i + 2] - az;
            double rec = 0.0;
            for (int64_t start = 0; start < K; start += block_k) {
                int64_t end = start + block_k < K ? start + block_k : K;
                double acc = 0.0;
                for (int64_t j = start; j < end; ++j) {
                    double kr = dx * kvecs[3 * j + 0] + dy * kvecs[3 * j + 1];
                    kr = kr + dz * kvecs[3 * j + 2];
                    acc += cos(kr) * kcoefs[j];
                }
                rec += acc;
            }
            double real = 0.0;
            for (int64_t s = 0; s < S; ++s) {
                double sh0 = shifts[3 * s + 0], sh1 = shifts[3 * s + 1], sh2 = shifts[3 * s + 2];
                int is_self = (sh0 == 0.0 && sh1 == 0.0 && sh2 == 0.0);
                double ex = grid_pos[3 * i + 0] - (ax + sh0);
                double ey = grid_pos[3 * i + 1] - (ay + sh1);
                double ez = grid_pos[3 * i + 2] - (az + sh2);
                double r2 = ex * ex + ey * ey;
                r2 = r2 + ez * ez;
                if (r2 <= R2) {
                    if (is_self && r2 <= 1e-20) continue;
                    double r = sqrt(r2);
                    real += erfc(sqrt_alpha * r) / r;
                }
            }
            phi[i * N + a] = rec + real;
        }
    }
}

/*
 * Phi . q for grid rows [0, M) without storing Phi (compute_stats, repeat.py:2406),
 * evaluated with the same per-element arithmetic as oracle_phi.
 */
void oracle_phi_matvec(int64_t M, const double *grid_pos, int64_t N, const double *atom_pos,
                       int64_t K, const double *kvecs, const double *kcoefs,
                       int64_t S, const double *shifts, double r_cutoff, double sqrt_alpha,
                       int fit_flag, int block_k, const double *q, double *out, int nthreads) {
    const int64_t rows = 256;
    double *tile = (double *)malloc(sizeof(double) * rows * (N > 0 ? N : 1));
    for (int64_t r0 = 0; r0 < M; r0 += rows) {
        int64_t mr = (M - r0) < rows ? (M - r0) : rows;
        oracle_phi(mr, grid_pos + 3 * r0, N, atom_pos, K, kvecs, kcoefs, S, shifts, r_cutoff,
                   sqrt_alpha, fit_flag, block_k, tile, nthreads);
        for (int64_t i = 0; i < mr; ++i) {
            double acc = 0.0;
            for (int64_t a = 0; a < N; ++a) acc += tile[i * N + a] * q[a];
            out[r0 + i] = acc;
        }
    }
    free(tile);
}
