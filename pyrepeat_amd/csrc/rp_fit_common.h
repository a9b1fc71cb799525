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
    const double *afrac;   // N x 3 fractional coordinates of the atoms
    const double *shifts;  // S x 3 (pruned)
    int S;
    double R2, sqrt_alpha, r_cut;
    double box[9];         // cell rows
    double h[3];           // interplanar spacings V / |a_j x a_k|, shrunk by a relative 1e-12
    double inv_n[3];       // 1 / n_grid
    int nmax[3];           // extent of the full shift list, |n_i| <= nmax[i] (repeat.py:815-822)
    // |w . cell|^2 = q22 w2^2 + q11 (w1 + a12 w2)^2 + q00 (w0 + a01 w1 + a02 w2)^2 for fractional w: the
    // cell's metric G = B B^T in triangular form (q00 = G00, a0j = G0j / G00, q11 = G11 - G01 a01,
    // a12 = (G12 - G02 a01) / q11, q22 = G22 - q00 a02^2 - q11 a12^2), inv_q00 = 1 / q00,
    // inv_q11 = 1 / q11, reach2 = Rc / sqrt(q22) + 1e-9.
    double q00, q11, q22, a01, a02, a12, inv_q00, inv_q11, reach2;

    // The shift-list sum for one (point, atom) pair, visiting only the shifts that can lie within Rc:
    // image n is at distance >= |d_i - n_i| h_i from the point along plane family i (d = g_f - a_f, the
    // fractional displacement), so n_i runs over [d_i - Rc/h_i, d_i + Rc/h_i] within the list's extent.
    // The candidates are visited in the list's order (n_0 outer, n_2 inner), each shift vector rounded as
    // the list's ((n0 b0 + n1 b1) + n2 b2), so the terms inside Rc and their order are those of the loop
    // over the list: the same sum.  Every shift the list's AABB pruning drops is farther than Rc from
    // every kept point, so it adds nothing either way.
    template <class Term>
    __device__ __forceinline__ double image_sum(double d0, double d1, double d2, double gx, double gy, double gz,
                                                double ax, double ay, double az, Term term) const {
        int lo[3], hi[3];
        const double dd[3] = {d0, d1, d2};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double reach = r_cut / h[i] + 1e-9;
            lo[i] = max(-nmax[i], (int)ceil(dd[i] - reach));
            hi[i] = min(nmax[i], (int)floor(dd[i] + reach));
        }
        double sum = 0.0;
        for (int n0 = lo[0]; n0 <= hi[0]; ++n0)
            for (int n1 = lo[1]; n1 <= hi[1]; ++n1)
                for (int n2 = lo[2]; n2 <= hi[2]; ++n2) {
                    const double sx = ((double)n0 * box[0] + (double)n1 * box[3]) + (double)n2 * box[6];
                    const double sy = ((double)n0 * box[1] + (double)n1 * box[4]) + (double)n2 * box[7];
                    const double sz = ((double)n0 * box[2] + (double)n1 * box[5]) + (double)n2 * box[8];
                    const double dx = gx - (ax + sx), dy = gy - (ay + sy), dz = gz - (az + sz);
                    const double r2 = (dx * dx + dy * dy) + dz * dz;
                    if (r2 <= R2) {
                        if (sx == 0.0 && sy == 0.0 && sz == 0.0 && r2 <= 1e-20) continue;
                        sum += term(r2);
                    }
                }
        return sum;
    }

    // The same sum over the lattice points of the cutoff sphere only: with w = d - n the metric form
    // above bounds n2 by q22 w2^2 <= Rc^2, then n1 for each n2 and n0 for each (n2, n1) by what is left
    // of Rc^2 (the enumeration of Fincke and Pohst), each range widened by 1e-9 and clipped to the list's
    // extent.  Every candidate is tested exactly as in the list, its shift vector rounded as the list's,
    // so the terms are those of the list; they are added n2-major, so the sum agrees with the list's
    // to rounding.  For a wave whose lanes hold neighbouring points of one atom the ranges nearly agree
    // across lanes; for lanes of unrelated pairs the box above diverges less.
    template <class Term>
    __device__ __forceinline__ double image_sum_sphere(double d0, double d1, double d2, double gx, double gy,
                                                       double gz, double ax, double ay, double az, Term term) const {
        // sqrt(x) for x >= 0 within a relative 1e-6 (the v_rsq estimate is good to about 2^-23), widened:
        // only the ranges use it
        auto root_up = [](double x) { return x > 0.0 ? x * __builtin_amdgcn_rsq(x) * (1.0 + 1e-6) + 1e-9 : 1e-9; };
        const double R2s = R2 * (1.0 + 1e-12);
        double sum = 0.0;
        const int lo2 = max(-nmax[2], (int)ceil(d2 - reach2)), hi2 = min(nmax[2], (int)floor(d2 + reach2));
        for (int n2 = lo2; n2 <= hi2; ++n2) {
            const double w2 = d2 - (double)n2;
            const double rem2 = R2s - q22 * w2 * w2;
            if (rem2 < 0.0) continue;
            const double c1 = d1 + a12 * w2, s1 = root_up(rem2 * inv_q11);
            const int lo1 = max(-nmax[1], (int)ceil(c1 - s1)), hi1 = min(nmax[1], (int)floor(c1 + s1));
            for (int n1 = lo1; n1 <= hi1; ++n1) {
                const double w1 = d1 - (double)n1;
                const double t1 = w1 + a12 * w2;
                const double rem1 = rem2 - q11 * t1 * t1;
                if (rem1 < 0.0) continue;
                const double c0 = d0 + a01 * w1 + a02 * w2, s0 = root_up(rem1 * inv_q00);
                const int lo0 = max(-nmax[0], (int)ceil(c0 - s0)), hi0 = min(nmax[0], (int)floor(c0 + s0));
                for (int n0 = lo0; n0 <= hi0; ++n0) {
                    const double sx = ((double)n0 * box[0] + (double)n1 * box[3]) + (double)n2 * box[6];
                    const double sy = ((double)n0 * box[1] + (double)n1 * box[4]) + (double)n2 * box[7];
                    const double sz = ((double)n0 * box[2] + (double)n1 * box[5]) + (double)n2 * box[8];
                    const double dx = gx - (ax + sx), dy = gy - (ay + sy), dz = gz - (az + sz);
                    const double r2 = (dx * dx + dy * dy) + dz * dz;
                    if (r2 <= R2) {
                        if (sx == 0.0 && sy == 0.0 && sz == 0.0 && r2 <= 1e-20) continue;
                        sum += term(r2);
                    }
                }
            }
        }
        return sum;
    }

    // Nearest image of atom a seen from grid point (i, j, k): shift index n0 = rint(g_f - a_f) and the
    // lower bound min_i (1 - |w_i|) h_i on the distance to every other image (w = g_f - a_f - n0).
    // Returns true when that bound exceeds Rc, i.e. only the shift n0 . box can be within the cutoff.
    __device__ __forceinline__ bool nearest_image(int f, int a, double &sx, double &sy, double &sz) const {
        int i, j, k;
        geom.ijk(f, i, j, k);
        const double df0 = (double)i * inv_n[0] - afrac[3 * a];
        const double df1 = (double)j * inv_n[1] - afrac[3 * a + 1];
        const double df2 = (double)k * inv_n[2] - afrac[3 * a + 2];
        const double n0 = rint(df0), n1 = rint(df1), n2 = rint(df2);
        const double b = fmin(fmin((1.0 - fabs(df0 - n0)) * h[0], (1.0 - fabs(df1 - n1)) * h[1]),
                              (1.0 - fabs(df2 - n2)) * h[2]);
        // shift vector with the rounding of the shift list: ((n0*b0 + n1*b1) + n2*b2)
        sx = (n0 * box[0] + n1 * box[3]) + n2 * box[6];
        sy = (n0 * box[1] + n1 * box[4]) + n2 * box[7];
        sz = (n0 * box[2] + n1 * box[5]) + n2 * box[8];
        return b > r_cut;
    }

    // sum_s erfc(sqrt_alpha r)/r over r^2 <= Rc^2, zero shift skipped at r^2 <= 1e-20
    // (repeat.py:849-872); fit-type 0: 1/|d| for |d| > 1e-10 (repeat.py:991-994).
    __device__ __forceinline__ double real_space(double gx, double gy, double gz, int a, int f = -1) const {
        const double ax = atoms[3 * a], ay = atoms[3 * a + 1], az = atoms[3 * a + 2];
        if (fit_type == 0) {
            const double dx = gx - ax, dy = gy - ay, dz = gz - az;
            const double d = sqrt((dx * dx + dy * dy) + dz * dz);
            return d > 1e-10 ? 1.0 / d : 0.0;
        }
        double sx, sy, sz;
        if (f >= 0 && nearest_image(f, a, sx, sy, sz)) {
            // only this image can be within Rc: the shift-list sum reduces to (at most) this one term
            const double dx = gx - (ax + sx), dy = gy - (ay + sy), dz = gz - (az + sz);
            const double r2 = (dx * dx + dy * dy) + dz * dz;
            if (!(r2 <= R2)) return 0.0;
            if (sx == 0.0 && sy == 0.0 && sz == 0.0 && r2 <= 1e-20) return 0.0;
            const double r = sqrt(r2);
            return erfc(sqrt_alpha * r) / r;
        }
        if (f >= 0) {
            // several images may lie within Rc: the candidate shifts only
            int i, j, k;
            geom.ijk(f, i, j, k);
            const double d0 = (double)i * inv_n[0] - afrac[3 * a], d1 = (double)j * inv_n[1] - afrac[3 * a + 1],
                         d2 = (double)k * inv_n[2] - afrac[3 * a + 2];
            const double sa = sqrt_alpha;
            return image_sum(d0, d1, d2, gx, gy, gz, ax, ay, az, [sa](double r2) {
                const double r = sqrt(r2);
                return erfc(sa * r) / r;
            });
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

// Factorized reciprocal sum of the split form (rp_kspace.hip).
struct SepPlan {
    std::vector<int> first_line, end_line;  // per chunk: lines [first, end) holding its rows
};

// Device copies of the Ewald setup and the chunking plan of one fit.
struct FitDevice {
    int K = 0, KP = 0, NP = 0, S = 0;
    bool ready = false;
    std::vector<double> shifts_host;   // pruned shifts the device copy was built from
    DBuf afrac;
    int64_t rows = 0;      // points per chunk (multiple of 128)
    int splits = 1;        // Gram row splits per chunk
    FitGeom geom{};
    EpiArgs epi{};
    DBuf khkl, kcoef, kvecs, ex, ey, ez, shifts, d_tiles;
    std::vector<int2> tiles;
    // split form: factorized reciprocal sum (k-index box, per-plane partial sums) and the
    // table of grid lines holding kept points
    DBuf sep_kbox, sep_t1, sep_coef, line_row, line_key, line_cnt, line_off, chunk_lines, multi_flag;
    bool sep_kbox_ok = false;   // sep_kbox holds the index box of the current k-set
    SepPlan sep_plan;           // chunk line ranges of the last split accumulate (the streamed refinement's)
    hipEvent_t ev_plan = nullptr;   // marks the chunk line ranges on the stream
    FitDevice() = default;
    FitDevice(const FitDevice &) = delete;
    FitDevice &operator=(const FitDevice &) = delete;
    ~FitDevice() {
        if (ev_plan) (void)hipEventDestroy(ev_plan);
    }
};

int prepare_device(rp_fit *fit, FitDevice &dv, int precision);
// Persistent device setup of a fit (built on first use, rebuilt when the pruned shifts change).
int device_setup(rp_fit *fit, int precision, FitDevice *&dv);
int accumulate_f64(rp_fit *fit, FitDevice &dv);
int accumulate_split(rp_fit *fit, FitDevice &dv);
int accumulate_f64_factorized(rp_fit *fit, FitDevice &dv);

// Per accumulate: the k-index box, T1 for every plane, and the line table of the kept points.
int sep_prepare(rp_fit *fit, FitDevice &dv, int64_t rows, SepPlan &plan);
// Phi rows [r0, r0 + valid) of one chunk (k-space + real space + augmented columns), column-major
// [NP][rows] fp64, with the column maxima of |Phi| (bit patterns) into colmax.  The padding columns
// [N + 2, NP) are written as zeros only when zero_pad is set (a consumer that reads them); otherwise they
// are left as they are and no consumer reads them.
// gate (nullable): the launch does nothing unless gate[0] != 0 and gate_flags[0] == gate_flags[1] == 0 (a
// refinement step the device has decided to skip).
void sep_phi_chunk(rp_fit *fit, FitDevice &dv, const SepPlan &plan, int chunk, int64_t r0, int64_t valid,
                   int64_t rows, double *phi, unsigned long long *colmax, hipStream_t st, const int *gate = nullptr,
                   const int *gate_flags = nullptr, bool zero_pad = false);
// Column-major Phi chunk -> int8 slices [P][NP][rows] (P = 3 or 4) with per-column power-of-two scales,
// rows [0, span) written (span a multiple of 16, at most rows; rows past `valid` as zeros); the columns
// from ncols on (the padding) are zero slices, their Phi not read.
void sep_slice_chunk(const double *phi, int NP, int ncols, int64_t rows, int64_t valid, int64_t span,
                     const unsigned long long *colmax, int8_t *slices, double *scale_out, int P, hipStream_t st);
__global__ void k_symmetrize(double *gram, int NP, int tile);

}  // namespace rp
