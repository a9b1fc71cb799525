/*
 * Real-space part of the exact Ewald potential of point charges on a cube
 * grid, used only to synthesise ESP inputs for tests and benchmarks:
 *
 *   out[f] += sum_images q * erfc(eta * r) / r,  r = |g(i,j,k) - P| <= rcut
 *
 * with g(i,j,k) = (i*a0 + j*a1) + k*a2 and f = i + j*nx + k*nx*ny.
 * Images P are passed in with their fractional coordinates; the loop over
 * grid indices is bounded per image by the fractional half-widths
 * dfrac[c] = rcut * |b_c| / (2 pi), b_c the reciprocal vectors.
 * Parallel over z planes, so no two threads write the same element.
 */
#include <math.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void synth_esp_real(int nx, int ny, int nz, const double *axis, int64_t n_img,
                    const double *img_xyz, const double *img_frac, const double *img_q,
                    const double *dfrac, double eta, double rcut, double *out, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const double rc2 = rcut * rcut;
    const int64_t layer = (int64_t)nx * ny;
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < nz; ++k) {
        const double fz = (double)k / nz;
        for (int64_t p = 0; p < n_img; ++p) {
            const double *fr = img_frac + 3 * p;
            if (fz < fr[2] - dfrac[2] || fz > fr[2] + dfrac[2]) continue;
            int j0 = (int)floor((fr[1] - dfrac[1]) * ny), j1 = (int)ceil((fr[1] + dfrac[1]) * ny);
            int i0 = (int)floor((fr[0] - dfrac[0]) * nx), i1 = (int)ceil((fr[0] + dfrac[0]) * nx);
            if (j0 < 0) j0 = 0;
            if (i0 < 0) i0 = 0;
            if (j1 > ny - 1) j1 = ny - 1;
            if (i1 > nx - 1) i1 = nx - 1;
            const double px = img_xyz[3 * p], py = img_xyz[3 * p + 1], pz = img_xyz[3 * p + 2];
            const double q = img_q[p];
            for (int j = j0; j <= j1; ++j) {
                for (int i = i0; i <= i1; ++i) {
                    double gx = (i * axis[0] + j * axis[3]) + k * axis[6];
                    double gy = (i * axis[1] + j * axis[4]) + k * axis[7];
                    double gz = (i * axis[2] + j * axis[5]) + k * axis[8];
                    double dx = gx - px, dy = gy - py, dz = gz - pz;
                    double r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 > rc2 || r2 < 1e-24) continue;
                    double r = sqrt(r2);
                    out[i + j * (int64_t)nx + k * layer] += q * erfc(eta * r) / r;
                }
            }
        }
    }
}
