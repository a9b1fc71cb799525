// vdW grid filter on the device: keep mask, stream compaction, mean-subtracted ESP.
//
// Predicate (repeat.py:2175-2177, 2081): keep a grid point iff every effective
// atom is farther than r_min and at least one is within r_max, with
// d = fl(sqrt((dx*dx + dy*dy) + dz*dz)) exactly as scipy's cdist computes it.
// The host turns each radius r into the squared-distance threshold
// T(r) = max{s : fl(sqrt(s)) <= r}, so the kernel compares the rounded s
// directly:  d > r_min  <=>  s > T(r_min),  d <= r_max  <=>  s <= T(r_max).
// Grid points are ((i*a0 + j*a1) + k*a2) per component (repeat.py:2139-2151).
// Contraction into FMA would change roundings, so it is switched off here.
//
// Work layout: one thread per grid point in flat order f = i + j*nx + k*nx*ny
// (consecutive lanes are neighbouring x points, so they scan the same cell-list
// bins).  Effective atoms within reach of the grid are binned on the host in
// cubes of edge h > max r_min: the 27 bins around a point hold every atom that
// can fail the r_min test.  The r_max test is usually satisfied by a binned
// atom; otherwise (r_max > h) the full effective list is scanned until a hit.
// Compaction: wave ballot + mbcnt prefix inside a block, a one-workgroup scan of
// per-block counts, then an ordered scatter of indices and ESP values.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "rp_internal.h"

#pragma clang fp contract(off)

namespace rp {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ void grid_xyz(const double *ax, int i, int j, int k, double &x, double &y, double &z) {
    x = ((double)i * ax[0] + (double)j * ax[3]) + (double)k * ax[6];
    y = ((double)i * ax[1] + (double)j * ax[4]) + (double)k * ax[7];
    z = ((double)i * ax[2] + (double)j * ax[5]) + (double)k * ax[8];
}

__device__ __forceinline__ double dist2(double gx, double gy, double gz, double4 a) {
    double dx = gx - a.x, dy = gy - a.y, dz = gz - a.z;
    double s = dx * dx + dy * dy;
    return s + dz * dz;
}

// bin_atoms: binned atoms as double4 {x, y, z, T(r_min)}, bin_smax: T(r_max) in the same order.
__global__ __launch_bounds__(kBlock) void k_filter_mask(FilterArgs a, const int *__restrict__ bin_start,
                                                        const double4 *__restrict__ bin_atoms,
                                                        const double *__restrict__ bin_smax,
                                                        const double4 *__restrict__ all_atoms,
                                                        const double *__restrict__ all_smax,
                                                        uint8_t *__restrict__ mask, int *__restrict__ block_count) {
    const int64_t n_total = (int64_t)a.nx * a.ny * a.nz;
    const int64_t f = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    bool keep = false;
    if (f < n_total) {
        const int64_t layer = (int64_t)a.nx * a.ny;
        const int k = (int)(f / layer);
        const int rem = (int)(f - (int64_t)k * layer);
        const int j = rem / a.nx;
        const int i = rem - j * a.nx;
        double gx, gy, gz;
        grid_xyz(a.ax, i, j, k, gx, gy, gz);
        const int bx = min(max((int)floor((gx - a.bmin[0]) * a.binv), 0), a.bn[0] - 1);
        const int by = min(max((int)floor((gy - a.bmin[1]) * a.binv), 0), a.bn[1] - 1);
        const int bz = min(max((int)floor((gz - a.bmin[2]) * a.binv), 0), a.bn[2] - 1);
        bool valid = true, near = false;
        // the 9 bin rows around the point, its own row first: a point inside an atom's r_min sphere (most of
        // the grid) usually finds that atom there and stops (the predicate does not depend on the order)
        for (int q = 0; q < 9 && valid; ++q) {
            const int dz = q == 0 ? 0 : (q <= 4 ? (q <= 2 ? 0 : (q == 3 ? -1 : 1)) : (q <= 6 ? -1 : 1));
            const int dy = q == 0 ? 0 : (q <= 4 ? (q <= 2 ? (q == 1 ? -1 : 1) : 0) : ((q & 1) ? -1 : 1));
            const int z2 = bz + dz, y2 = by + dy;
            if (z2 < 0 || z2 >= a.bn[2] || y2 < 0 || y2 >= a.bn[1]) continue;
            const int row = (z2 * a.bn[1] + y2) * a.bn[0];
            const int x0 = max(bx - 1, 0), x1 = min(bx + 1, a.bn[0] - 1);
            const int beg = bin_start[row + x0], end = bin_start[row + x1 + 1];
            for (int t = beg; t < end; ++t) {
                const double4 at = bin_atoms[t];
                const double s = dist2(gx, gy, gz, at);
                if (s <= at.w) {
                    valid = false;
                    break;
                }
                if (s <= bin_smax[t]) near = true;
            }
        }
        if (valid && !near && a.fallback) {
            for (int64_t t = 0; t < a.n_all; ++t) {
                if (dist2(gx, gy, gz, all_atoms[t]) <= all_smax[t]) {
                    near = true;
                    break;
                }
            }
        }
        keep = valid && near;
        mask[f] = keep ? 1 : 0;
    }
    const unsigned long long ballot = __ballot(keep);
    __shared__ int wave_cnt[kBlock / 64];
    if ((threadIdx.x & 63) == 0) wave_cnt[threadIdx.x >> 6] = __popcll(ballot);
    __syncthreads();
    if (threadIdx.x == 0) {
        int c = 0;
        for (int w = 0; w < kBlock / 64; ++w) c += wave_cnt[w];
        block_count[blockIdx.x] = c;
    }
}

// Exclusive scan of per-block counts in one workgroup; out[n_blocks] = total.
__global__ __launch_bounds__(1024) void k_scan_counts(const int *__restrict__ cnt, int64_t n, int64_t *__restrict__ out) {
    block_scan_counts_1024(cnt, n, out);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_compact(int nx, int ny, int nz, int esp_order, const uint8_t *__restrict__ mask,
                                                    const int64_t *__restrict__ block_off, const T *__restrict__ esp,
                                                    int *__restrict__ idx_out, double *__restrict__ v_out,
                                                    double *__restrict__ block_vsum) {
    const int64_t n_total = (int64_t)nx * ny * nz;
    const int64_t f = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool keep = f < n_total && mask[f];
    const unsigned long long ballot = __ballot(keep);
    __shared__ int wave_base[kBlock / 64];
    __shared__ double wave_sum[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wave_base[wave] = __popcll(ballot);
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            int v = wave_base[w];
            wave_base[w] = run;
            run += v;
        }
    }
    __syncthreads();
    double v = 0.0;
    if (keep) {
        const int rank = __popcll(ballot & ((1ull << lane) - 1ull));
        const int64_t pos = block_off[blockIdx.x] + wave_base[wave] + rank;
        int64_t src = f;
        if (esp_order == RP_ORDER_CUBE) {
            const int64_t layer = (int64_t)nx * ny;
            const int64_t k = f / layer;
            const int64_t rem = f - k * layer;
            const int64_t j = rem / nx;
            const int64_t i = rem - j * nx;
            src = (i * ny + j) * nz + k;
        }
        v = (double)esp[src];
        idx_out[pos] = (int)f;
        v_out[pos] = v;
    }
    // fixed-order block sum of the kept values (deterministic mean)
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) wave_sum[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) s += wave_sum[w];
        block_vsum[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(1024) void k_mean(const double *__restrict__ block_vsum, int64_t n_blocks,
                                               const int64_t *__restrict__ total, double *__restrict__ mean_out) {
    __shared__ double part[1024];
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < n_blocks; b += 1024) s += block_vsum[b];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) mean_out[0] = total[0] > 0 ? part[0] / (double)total[0] : 0.0;
}

// V -= mean; also the AABB of the kept grid points (for shift pruning) via per-block partials.
// A fixed grid strides over the kept points (the count m is read on the device), one bounds partial
// per block.
constexpr int kBoundBlocks = 1024;
__global__ __launch_bounds__(kBlock) void k_center_and_bounds(int nx, int ny, const double *__restrict__ ax_,
                                                               const int *__restrict__ idx, double *__restrict__ v,
                                                               const int64_t *__restrict__ m_dev,
                                                               const double *__restrict__ mean,
                                                               double *__restrict__ bounds_part) {
    __shared__ double ax[9];
    if (threadIdx.x < 9) ax[threadIdx.x] = ax_[threadIdx.x];
    __syncthreads();
    const int64_t m = *m_dev;
    const double mu = mean[0];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < m; t += (int64_t)gridDim.x * kBlock) {
        v[t] = v[t] - mu;
        const int f = idx[t];
        const int layer = nx * ny;
        const int k = f / layer;
        const int rem = f - k * layer;
        const int j = rem / nx;
        const int i = rem - j * nx;
        double g[3];
        grid_xyz(ax, i, j, k, g[0], g[1], g[2]);
        for (int c = 0; c < 3; ++c) {
            lo[c] = fmin(lo[c], g[c]);
            hi[c] = fmax(hi[c], g[c]);
        }
    }
    for (int c = 0; c < 3; ++c) {
        for (int off = 32; off > 0; off >>= 1) {
            lo[c] = fmin(lo[c], __shfl_down(lo[c], off, 64));
            hi[c] = fmax(hi[c], __shfl_down(hi[c], off, 64));
        }
    }
    __shared__ double wl[kBlock / 64][6];
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 3; ++c) {
            wl[threadIdx.x >> 6][c] = lo[c];
            wl[threadIdx.x >> 6][3 + c] = hi[c];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int c = 0; c < 3; ++c) {
            double a = wl[0][c], b = wl[0][3 + c];
            for (int w = 1; w < kBlock / 64; ++w) {
                a = fmin(a, wl[w][c]);
                b = fmax(b, wl[w][3 + c]);
            }
            bounds_part[6 * blockIdx.x + c] = a;
            bounds_part[6 * blockIdx.x + 3 + c] = b;
        }
    }
}

// The filter's results for the host, written into mapped pinned memory so one synchronisation
// returns them: out[0] the kept count, out[1] mean(V), out[2..4] / out[5..7] the AABB of the kept points.
__global__ __launch_bounds__(1024) void k_filter_publish(const double *__restrict__ bounds_part, int64_t n_blocks,
                                                         const int64_t *__restrict__ total,
                                                         const double *__restrict__ mean, double *__restrict__ out) {
    __shared__ double red[16][6];
    double r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int64_t b = threadIdx.x; b < n_blocks; b += 1024)
        for (int c = 0; c < 3; ++c) {
            r[c] = fmin(r[c], bounds_part[6 * b + c]);
            r[3 + c] = fmax(r[3 + c], bounds_part[6 * b + 3 + c]);
        }
    for (int c = 0; c < 3; ++c)
        for (int off = 32; off > 0; off >>= 1) {
            r[c] = fmin(r[c], __shfl_xor(r[c], off, 64));
            r[3 + c] = fmax(r[3 + c], __shfl_xor(r[3 + c], off, 64));
        }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 6; ++c) red[w][c] = r[c];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int c = threadIdx.x;
        double a = red[0][c];
        for (int g = 1; g < 16; ++g) a = c < 3 ? fmin(a, red[g][c]) : fmax(a, red[g][c]);
        out[2 + c] = a;
    }
    if (threadIdx.x == 0) {
        out[0] = (double)total[0];
        out[1] = mean[0];
    }
}

}  // namespace

// Cell-list geometry over the grid's bounding box: bin edge h >= max r_min, bn bins per axis
// starting at bmin, at most 2^22 bins.
void filter_bins(const Setup &s, double glo[3], double ghi[3], double &h, int bn[3], double bmin[3]) {
    for (int c = 0; c < 3; ++c) {
        glo[c] = INFINITY;
        ghi[c] = -INFINITY;
    }
    for (int ci = 0; ci <= 1; ++ci)
        for (int cj = 0; cj <= 1; ++cj)
            for (int ck = 0; ck <= 1; ++ck) {
                const int i = ci ? s.nx - 1 : 0, j = cj ? s.ny - 1 : 0, k = ck ? s.nz - 1 : 0;
                for (int c = 0; c < 3; ++c) {
                    double g = ((double)i * s.axis[c] + (double)j * s.axis[3 + c]) + (double)k * s.axis[6 + c];
                    glo[c] = std::min(glo[c], g);
                    ghi[c] = std::max(ghi[c], g);
                }
            }
    h = std::max(s.max_rmin * 1.001 + 1e-9, 1e-3);
    for (int c = 0; c < 3; ++c) {
        bmin[c] = glo[c] - h;
        bn[c] = (int)std::ceil((ghi[c] - glo[c] + 2.0 * h) / h) + 1;
    }
    // keep the bin count bounded for very fine radii
    while ((int64_t)bn[0] * bn[1] * bn[2] > (int64_t)1 << 22) {
        h *= 2.0;
        for (int c = 0; c < 3; ++c) {
            bmin[c] = glo[c] - h;
            bn[c] = (int)std::ceil((ghi[c] - glo[c] + 2.0 * h) / h) + 1;
        }
    }
}

uint64_t filter_bytes_bound(const Setup &s, int esp_dtype) {
    double glo[3], ghi[3], h, bmin[3];
    int bn[3];
    filter_bins(s, glo, ghi, h, bn, bmin);
    const int64_t n_total = (int64_t)s.nx * s.ny * s.nz;
    const int64_t nb = ceil_div(n_total, kBlock);
    const int64_t E = std::max<int64_t>((int64_t)s.eff_smin.size(), 1);
    const int64_t nbins = (int64_t)bn[0] * bn[1] * bn[2];
    uint64_t b = 0;
    b += 4 * (nbins + 1) + 2 * (32 + 8) * E;                 // bins, binned and unbinned images
    b += 4 * nb + 8 * (nb + 1) + 8 * nb + 8 + 72;             // block counts, offsets, V sums, mean, axis
    b += n_total * (1 + (esp_dtype == RP_F32 ? 4 : 8));       // mask, cube
    b += 12 * n_total + 48 * ceil_div(n_total, kBlock);       // kept indices and V (M <= n_total), bounds
    return b + 16 * 16;                                      // 16-byte minimum per buffer
}

int upload_filter(rp_fit *fit, const rp_system *sys) {
    Ctx *ctx = fit->ctx;
    const Setup &s = fit->setup;
    FilterState &fs = fit->filt;
    const int64_t n_total = (int64_t)s.nx * s.ny * s.nz;
    fit->n_total = n_total;
    fs.n_blocks = ceil_div(n_total, kBlock);
    fs.esp_dtype = sys->esp_dtype;
    fs.esp_order = sys->esp_order;

    // ---- cell list over the grid's bounding box (host)
    double glo[3], ghi[3], h, bmin[3];
    int bn[3];
    filter_bins(s, glo, ghi, h, bn, bmin);
    const int64_t n_eff = (int64_t)s.eff_smin.size();
    const double binv = 1.0 / h;
    const int64_t nbins = (int64_t)bn[0] * bn[1] * bn[2];
    std::vector<int> bin_of(n_eff, -1);
    std::vector<int> bin_start(nbins + 1, 0);
    for (int64_t t = 0; t < n_eff; ++t) {
        const double *p = &s.eff_xyz[3 * t];
        bool inside = true;
        int b[3];
        for (int c = 0; c < 3; ++c) {
            if (p[c] < glo[c] - h || p[c] > ghi[c] + h) inside = false;
            b[c] = std::min(std::max((int)std::floor((p[c] - bmin[c]) * binv), 0), bn[c] - 1);
        }
        if (!inside) continue;
        bin_of[t] = (b[2] * bn[1] + b[1]) * bn[0] + b[0];
        bin_start[bin_of[t] + 1]++;
    }
    for (int64_t b = 0; b < nbins; ++b) bin_start[b + 1] += bin_start[b];
    const int n_binned = bin_start[nbins];
    std::vector<double4> bin_atoms(std::max(n_binned, 1));
    std::vector<double> bin_smax(std::max(n_binned, 1));
    std::vector<double4> all_atoms(std::max<int64_t>(n_eff, 1));
    std::vector<double> all_smax(std::max<int64_t>(n_eff, 1));
    {
        std::vector<int> cur(bin_start.begin(), bin_start.end() - 1);
        for (int64_t t = 0; t < n_eff; ++t) {
            const double *p = &s.eff_xyz[3 * t];
            all_atoms[t] = make_double4(p[0], p[1], p[2], s.eff_smin[t]);
            all_smax[t] = s.eff_smax[t];
            if (bin_of[t] < 0) continue;
            const int pos = cur[bin_of[t]]++;
            bin_atoms[pos] = all_atoms[t];
            bin_smax[pos] = s.eff_smax[t];
        }
    }
    FilterArgs &fa = fs.fa;
    fa.nx = s.nx;
    fa.ny = s.ny;
    fa.nz = s.nz;
    std::copy(s.axis, s.axis + 9, fa.ax);
    std::copy(bmin, bmin + 3, fa.bmin);
    fa.binv = binv;
    std::copy(bn, bn + 3, fa.bn);
    fa.fallback = s.max_rmax > h ? 1 : 0;
    fa.n_all = n_eff;

    const int64_t nb = fs.n_blocks;
    RP_CHECK(fs.bin_start.grow(ctx, sizeof(int) * (nbins + 1)));
    RP_CHECK(fs.bin_atoms.grow(ctx, sizeof(double4) * bin_atoms.size()));
    RP_CHECK(fs.bin_smax.grow(ctx, sizeof(double) * bin_smax.size()));
    RP_CHECK(fs.all.grow(ctx, sizeof(double4) * all_atoms.size()));
    RP_CHECK(fs.all_smax.grow(ctx, sizeof(double) * all_smax.size()));
    RP_CHECK(fs.cnt.grow(ctx, sizeof(int) * nb));
    RP_CHECK(fs.off.grow(ctx, sizeof(int64_t) * (nb + 1)));
    RP_CHECK(fs.vsum.grow(ctx, sizeof(double) * nb));
    RP_CHECK(fs.mean.grow(ctx, sizeof(double)));
    RP_CHECK(fs.axis.grow(ctx, sizeof(double) * 9));
    RP_CHECK(fit->d_mask.grow(ctx, (size_t)n_total));
    RP_CHECK(fit->d_idx.grow(ctx, sizeof(int) * n_total));
    RP_CHECK(fit->d_v.grow(ctx, sizeof(double) * n_total));
    RP_CHECK(fs.bounds.grow(ctx, sizeof(double) * 6 * nb));
    const size_t esp_bytes = (size_t)n_total * (sys->esp_dtype == RP_F32 ? 4 : 8);
    RP_CHECK(fs.esp.grow(ctx, esp_bytes));
    hipStream_t st = ctx->stream;
    RP_HIP(ctx, hipMemcpyAsync(fs.bin_start.p, bin_start.data(), sizeof(int) * (nbins + 1), hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipMemcpyAsync(fs.bin_atoms.p, bin_atoms.data(), sizeof(double4) * bin_atoms.size(), hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipMemcpyAsync(fs.bin_smax.p, bin_smax.data(), sizeof(double) * bin_smax.size(), hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipMemcpyAsync(fs.all.p, all_atoms.data(), sizeof(double4) * all_atoms.size(), hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipMemcpyAsync(fs.all_smax.p, all_smax.data(), sizeof(double) * all_smax.size(), hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipMemcpyAsync(fs.axis.p, s.axis, sizeof(double) * 9, hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipMemcpyAsync(fs.esp.p, sys->esp, esp_bytes, hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipStreamSynchronize(st));  // host staging vectors go out of scope
    return RP_OK;
}

int run_filter(rp_fit *fit) {
    Ctx *ctx = fit->ctx;
    const Setup &s = fit->setup;
    FilterState &fs = fit->filt;
    hipStream_t st = ctx->stream;
    const int64_t n_blocks = fs.n_blocks;
    hipEvent_t e0, e1;
    RP_HIP(ctx, hipEventCreate(&e0));
    RP_HIP(ctx, hipEventCreate(&e1));
    RP_HIP(ctx, hipEventRecord(e0, st));
    {
        KTimer kt(ctx, KT_FILTER);
        k_filter_mask<<<n_blocks, kBlock, 0, st>>>(fs.fa, fs.bin_start.as<int>(), fs.bin_atoms.as<double4>(),
                                                   fs.bin_smax.as<double>(), fs.all.as<double4>(), fs.all_smax.as<double>(),
                                                   fit->d_mask.as<uint8_t>(), fs.cnt.as<int>());
    }
    k_scan_counts<<<1, 1024, 0, st>>>(fs.cnt.as<int>(), n_blocks, fs.off.as<int64_t>());
    // compaction, mean and bounds run before the host knows the count: the kept-point buffers hold
    // n_total entries and the grids cover every filter block
    const int64_t *d_count = fs.off.as<int64_t>() + n_blocks;
    if (fs.esp_dtype == RP_F32)
        k_compact<float><<<n_blocks, kBlock, 0, st>>>(s.nx, s.ny, s.nz, fs.esp_order, fit->d_mask.as<uint8_t>(),
                                                      fs.off.as<int64_t>(), fs.esp.as<float>(), fit->d_idx.as<int>(),
                                                      fit->d_v.as<double>(), fs.vsum.as<double>());
    else
        k_compact<double><<<n_blocks, kBlock, 0, st>>>(s.nx, s.ny, s.nz, fs.esp_order, fit->d_mask.as<uint8_t>(),
                                                       fs.off.as<int64_t>(), fs.esp.as<double>(), fit->d_idx.as<int>(),
                                                       fit->d_v.as<double>(), fs.vsum.as<double>());
    k_mean<<<1, 1024, 0, st>>>(fs.vsum.as<double>(), n_blocks, d_count, fs.mean.as<double>());
    const int bound_blocks = (int)std::min<int64_t>(n_blocks, kBoundBlocks);
    k_center_and_bounds<<<bound_blocks, kBlock, 0, st>>>(s.nx, s.ny, fs.axis.as<double>(), fit->d_idx.as<int>(),
                                                      fit->d_v.as<double>(), d_count, fs.mean.as<double>(),
                                                      fs.bounds.as<double>());
    RP_CHECK(fit->host_slots(ctx));
    k_filter_publish<<<1, 1024, 0, st>>>(fs.bounds.as<double>(), bound_blocks, d_count, fs.mean.as<double>(),
                                         fit->h_small_dev);
    RP_HIP(ctx, hipGetLastError());
    RP_HIP(ctx, hipEventRecord(e1, st));
    RP_HIP(ctx, hipStreamSynchronize(st));
    ctx->collect();
    float ms = 0.f;
    RP_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    fit->t_filter_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double *res = fit->h_small;
    const int64_t n_keep = (int64_t)res[0];
    fit->n_keep = n_keep;
    if (n_keep == 0) return ctx->fail(RP_EEMPTY, "the vdW filter kept no grid point");
    fit->v_mean = res[1];
    const double gmin[3] = {res[2], res[3], res[4]}, gmax[3] = {res[5], res[6], res[7]};
    if (s.fit_type == 1)
        fit->shifts_pruned = prune_shifts(s.shifts, gmin, gmax, fit->atom_xyz, s.r_cutoff);
    else
        fit->shifts_pruned = s.shifts;
    fit->accumulated = false;
    return RP_OK;
}

}  // namespace rp
