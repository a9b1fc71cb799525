// Host-side setup of a fit: cell, Ewald parameters, k-set, shift lattice,
// effective atoms of the vdW filter and their exact distance thresholds.
//
// The arithmetic that feeds the keep mask (effective atom positions, radii)
// follows the reference's NumPy expressions operation by operation so the
// device mask is bit-identical:
//   radii            repeat.py:1936-1939   r_min = f*R(Z-1), r_max = v*R(Z-1)
//   27 images        repeat.py:807-813     ((kx*b0 + ky*b1) + kz*b2), kz outer, kx inner
//   clipping box     repeat.py:1941-1964   AABB of the first voxel's corners +- (max r_max + 0.1)
//   Ewald            repeat.py:1077-1143   alpha = (pi/Rc)^2, recip = 2 pi (b x c)/V, signed V
//   k-set            repeat.py:1126-1143   kx 0..7, ky,kz -7..7, 0 < k^2 < 49, no 2x on kx > 0
//   shifts           repeat.py:815-822     nmax = floor(Rc/|a_i|)+1, loop ix, iy, iz
//   pruning          repeat.py:1017-1069   AABB distance <= Rc^2, order kept
#include <algorithm>
#include <cmath>
#include <limits>

#include "rp_internal.h"

namespace rp {

namespace {

// UFF van der Waals radii in Bohr, indexed by Z-1 (repeat.py:564-579).
const double kVdw[107] = {
    2.72687, 2.23177, 2.31586, 2.59365, 3.85788, 3.63867, 3.45820, 3.30702, 3.17852, 3.06419, 2.81853,
    2.85443, 4.25094, 4.05819, 3.91835, 3.81252, 3.72937, 3.65473, 3.60182, 3.21159, 3.11332, 2.99994,
    2.97065, 2.85632, 2.79774, 2.75144, 2.71365, 2.67774, 3.30230, 2.61066, 4.14133, 4.04401, 3.99677,
    3.97315, 3.95803, 3.91268, 3.88717, 3.44025, 3.16057, 2.95175, 2.99049, 2.88372, 2.83270, 2.79963,
    2.76750, 2.73916, 2.97443, 2.69097, 4.21692, 4.14984, 4.17629, 4.22354, 4.25188, 4.16118, 4.26795,
    3.49883, 3.32781, 3.35993, 3.40718, 3.37789, 3.35143, 3.32592, 3.30041, 3.18230, 3.26072, 3.23899,
    3.22104, 3.20403, 3.18797, 3.17002, 3.43930, 2.96781, 2.99522, 2.89978, 2.79113, 2.94797, 2.68341,
    2.60215, 3.11143, 2.55585, 4.10732, 4.06008, 4.12905, 4.44936, 4.48810, 4.50227, 4.62983, 3.47426,
    3.28623, 3.20875, 3.23521, 3.20781, 3.23521, 3.23521, 3.19458, 3.14261, 3.15490, 3.13033, 3.11710,
    3.10482, 3.09348, 3.06892, 3.05758, 3.04513, 3.03268, 3.02023, 3.00778};

const double kPi = 3.141592653589793;

inline void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
inline double dot3(const double *a, const double *b) {
    double s = a[0] * b[0] + a[1] * b[1];
    return s + a[2] * b[2];
}

}  // namespace

double sqrt_threshold(double r) {
    // Largest s >= 0 with fl(sqrt(s)) <= r, so that  sqrt(s) > r  <=>  s > T(r)
    // and  sqrt(s) <= r  <=>  s <= T(r)  for every double s (sqrt is correctly
    // rounded and monotone).
    if (!(r >= 0.0)) return -1.0;
    if (std::isinf(r)) return std::numeric_limits<double>::infinity();
    double s = r * r;
    const double inf = std::numeric_limits<double>::infinity();
    while (std::sqrt(s) > r) s = std::nextafter(s, -inf);
    for (;;) {
        double n = std::nextafter(s, inf);
        if (std::isinf(n) || std::sqrt(n) > r) break;
        s = n;
    }
    return s;
}

std::vector<double> prune_shifts(const std::vector<double> &shifts, const double gmin[3], const double gmax[3],
                                 const std::vector<double> &atom_xyz, double r_cutoff) {
    size_t n = atom_xyz.size() / 3;
    if (n == 0) return shifts;
    double amin[3], amax[3];
    for (int c = 0; c < 3; ++c) {
        amin[c] = amax[c] = atom_xyz[c];
        for (size_t a = 1; a < n; ++a) {
            amin[c] = std::min(amin[c], atom_xyz[3 * a + c]);
            amax[c] = std::max(amax[c], atom_xyz[3 * a + c]);
        }
    }
    const double r2 = r_cutoff * r_cutoff;
    std::vector<double> kept;
    for (size_t s = 0; s < shifts.size() / 3; ++s) {
        double d2 = 0.0;
        for (int c = 0; c < 3; ++c) {
            double smin = amin[c] + shifts[3 * s + c], smax = amax[c] + shifts[3 * s + c];
            if (gmax[c] < smin) {
                double d = smin - gmax[c];
                d2 += d * d;
            } else if (smax < gmin[c]) {
                double d = gmin[c] - smax;
                d2 += d * d;
            }
        }
        if (d2 <= r2) kept.insert(kept.end(), shifts.begin() + 3 * s, shifts.begin() + 3 * s + 3);
    }
    if (kept.empty()) kept = {0.0, 0.0, 0.0};
    return kept;
}

int build_setup(const rp_system *sys, const rp_options *opt, Setup &s, std::string &err) {
    s.nx = sys->n_grid[0];
    s.ny = sys->n_grid[1];
    s.nz = sys->n_grid[2];
    if (s.nx <= 0 || s.ny <= 0 || s.nz <= 0) {
        err = "grid dimensions must be positive";
        return RP_EINVAL;
    }
    if ((int64_t)s.nx * s.ny * s.nz > (int64_t)INT32_MAX) {
        err = "grid larger than 2^31-1 points";
        return RP_EINVAL;
    }
    if (opt->fit_type != 0 && opt->fit_type != 1) {
        err = "fit_type must be 0 or 1";
        return RP_EINVAL;
    }
    if (!(opt->r_cutoff > 0.0)) {
        err = "cutoff must be positive";
        return RP_EINVAL;
    }
    std::copy(sys->axis, sys->axis + 9, s.axis);
    const int ng[3] = {s.nx, s.ny, s.nz};
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 3; ++c) s.box[3 * i + c] = (double)ng[i] * s.axis[3 * i + c];
    double bc[3];
    cross3(s.box + 3, s.box + 6, bc);
    s.volume = dot3(s.box, bc);
    s.fit_type = opt->fit_type;
    s.r_cutoff = opt->r_cutoff;
    s.alpha = (kPi / s.r_cutoff) * (kPi / s.r_cutoff);
    s.sqrt_alpha = std::sqrt(s.alpha);

    s.khkl.clear();
    s.kvecs.clear();
    s.kcoefs.clear();
    s.shifts.clear();
    if (s.fit_type == 1) {
        if (s.volume == 0.0) {
            err = "degenerate cell (zero volume)";
            return RP_EINVAL;
        }
        double cr[3][3];
        cross3(s.box + 3, s.box + 6, cr[0]);
        cross3(s.box + 6, s.box + 0, cr[1]);
        cross3(s.box + 0, s.box + 3, cr[2]);
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 3; ++c) s.recip[3 * i + c] = 2.0 * kPi * cr[i][c] / s.volume;
        for (int i = 0; i < 3; ++i) s.nmax[i] = (int)std::floor(s.r_cutoff / std::sqrt(dot3(s.box + 3 * i, s.box + 3 * i))) + 1;
        const double beta = 1.0 / (4.0 * s.alpha);
        for (int kx = 0; kx <= 7; ++kx)
            for (int ky = -7; ky <= 7; ++ky)
                for (int kz = -7; kz <= 7; ++kz) {
                    int ksq = kx * kx + ky * ky + kz * kz;
                    if (ksq >= 49 || ksq == 0) continue;
                    double kv[3];
                    for (int c = 0; c < 3; ++c) {
                        double t = (double)kx * s.recip[c] + (double)ky * s.recip[3 + c];
                        kv[c] = t + (double)kz * s.recip[6 + c];
                    }
                    double k2 = dot3(kv, kv);
                    s.khkl.insert(s.khkl.end(), {kx, ky, kz});
                    s.kvecs.insert(s.kvecs.end(), kv, kv + 3);
                    s.kcoefs.push_back((4.0 * kPi / s.volume) * std::exp(-beta * k2) / k2);
                }
        for (int ix = -s.nmax[0]; ix <= s.nmax[0]; ++ix)
            for (int iy = -s.nmax[1]; iy <= s.nmax[1]; ++iy)
                for (int iz = -s.nmax[2]; iz <= s.nmax[2]; ++iz)
                    for (int c = 0; c < 3; ++c) {
                        double t = (double)ix * s.box[c] + (double)iy * s.box[3 + c];
                        s.shifts.push_back(t + (double)iz * s.box[6 + c]);
                    }
    } else {
        s.shifts = {0.0, 0.0, 0.0};
    }

    // ---- effective atoms of the filter
    const int n = sys->n_atoms;
    std::vector<double> rmin(n), rmax(n);
    for (int a = 0; a < n; ++a) {
        int z = sys->atomic_number[a] - 1;
        if (z < -107 || z >= 107) {
            err = "atomic number outside the vdW radius table (1..107)";
            return RP_EINVAL;
        }
        double raw = kVdw[z < 0 ? z + 107 : z];  // Z = 0 selects the last entry (Python index -1)
        rmin[a] = opt->vdw_factor * raw;
        rmax[a] = opt->vdw_max * raw;
    }
    double global_max_r = 5.0;
    if (n > 0) {
        global_max_r = rmax[0];
        for (int a = 1; a < n; ++a) global_max_r = std::max(global_max_r, rmax[a]);
    }
    const double margin = global_max_r + 0.1;
    double lo[3], hi[3];
    for (int c = 0; c < 3; ++c) {
        double cmin = 0.0, cmax = 0.0;
        bool first = true;
        for (int b0 = 0; b0 <= 1; ++b0)
            for (int b1 = 0; b1 <= 1; ++b1)
                for (int b2 = 0; b2 <= 1; ++b2) {
                    double t = (double)b0 * s.axis[c] + (double)b1 * s.axis[3 + c];
                    double v = t + (double)b2 * s.axis[6 + c];
                    if (first || v < cmin) cmin = v;
                    if (first || v > cmax) cmax = v;
                    first = false;
                }
        lo[c] = cmin - margin;
        hi[c] = cmax + margin;
    }
    double nsh[27][3];
    int t = 0;
    for (int kz = -1; kz <= 1; ++kz)
        for (int ky = -1; ky <= 1; ++ky)
            for (int kx = -1; kx <= 1; ++kx, ++t)
                for (int c = 0; c < 3; ++c) {
                    double u = (double)kx * s.box[c] + (double)ky * s.box[3 + c];
                    nsh[t][c] = u + (double)kz * s.box[6 + c];
                }
    s.eff_xyz.clear();
    s.eff_smin.clear();
    s.eff_smax.clear();
    s.max_rmin = 0.0;
    s.max_rmax = 0.0;
    for (int a = 0; a < n; ++a) {
        const double smin = sqrt_threshold(rmin[a]);
        const double smax = sqrt_threshold(rmax[a]);
        for (int q = 0; q < 27; ++q) {
            double p[3];
            for (int c = 0; c < 3; ++c) p[c] = sys->atom_xyz[3 * a + c] + nsh[q][c];
            if (p[0] >= lo[0] && p[0] <= hi[0] && p[1] >= lo[1] && p[1] <= hi[1] && p[2] >= lo[2] && p[2] <= hi[2]) {
                s.eff_xyz.insert(s.eff_xyz.end(), p, p + 3);
                s.eff_smin.push_back(smin);
                s.eff_smax.push_back(smax);
                s.max_rmin = std::max(s.max_rmin, rmin[a]);
                s.max_rmax = std::max(s.max_rmax, rmax[a]);
            }
        }
    }
    return RP_OK;
}

}  // namespace rp
