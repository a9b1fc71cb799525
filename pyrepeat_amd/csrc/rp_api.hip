// C ABI entry points of librepeat_hip.so and the fit lifecycle.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "rp_fit_common.h"

namespace rp {

namespace {
std::mutex g_err_mu;
std::string g_err;
}  // namespace

void set_global_error(const std::string &msg) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = msg;
}

DBuf::~DBuf() { release(); }

void DBuf::release() {
    if (p) {
        (void)hipFree(p);
        if (ctx) ctx->in_use -= bytes;
    }
    p = nullptr;
    bytes = 0;
}

int DBuf::alloc(Ctx *c, size_t n) {
    release();
    ctx = c;
    if (n == 0) n = 16;
    if (c->budget && c->in_use + n > c->budget)
        return c->fail(RP_EBUDGET, "HBM budget exceeded: need " + std::to_string(n) + " bytes, " +
                                       std::to_string(c->budget - c->in_use) + " left");
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        p = nullptr;
        return c->fail(RP_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    bytes = n;
    c->in_use += n;
    return RP_OK;
}

namespace {

// 2 pi r / n for integer-reduced r, as (cos, sin).
void phase_table(int hmin, int hmax, int n, std::vector<double2> &out) {
    const double two_pi = 6.283185307179586476925286766559;
    out.resize((size_t)(hmax - hmin + 1) * n);
    for (int h = hmin; h <= hmax; ++h)
        for (int i = 0; i < n; ++i) {
            long long r = ((long long)h * i) % n;
            if (r < 0) r += n;
            const double ang = two_pi * (double)r / (double)n;
            out[(size_t)(h - hmin) * n + i] = make_double2(std::cos(ang), std::sin(ang));
        }
}

// Basis at arbitrary points: direct fp64 cos(k.(g - R)) and the real-space sum (repeat.py:831-872).
__global__ __launch_bounds__(256) void k_phi_points(int64_t m, const double *__restrict__ g, int N,
                                                    const double *__restrict__ atoms, int K,
                                                    const double *__restrict__ kvecs, const double *__restrict__ kcoef,
                                                    EpiArgs e, double *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= m * N) return;
    const int64_t r = t / N;
    const int a = (int)(t % N);
    const double gx = g[3 * r], gy = g[3 * r + 1], gz = g[3 * r + 2];
    const double dx = gx - atoms[3 * a], dy = gy - atoms[3 * a + 1], dz = gz - atoms[3 * a + 2];
    double rec = 0.0;
    for (int k = 0; k < K; ++k) {
        const double kr = (dx * kvecs[3 * k] + dy * kvecs[3 * k + 1]) + dz * kvecs[3 * k + 2];
        rec += cos(kr) * kcoef[k];
    }
    out[t] = rec + e.real_space(gx, gy, gz, a);
}

__global__ __launch_bounds__(256) void k_zero(double *p, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

}  // namespace

}  // namespace rp

// Working memory of rp_fit_phi_noalloc: one copy per GPU, held until the process ends, so that entry
// point allocates nothing.  Host code never passes this name to a kernel or a copy (on the host it is
// a stand-in, not a device address): it uses the address hipGetSymbolAddress returns for the device.
static __device__ __align__(256) double scratch[RP_PHI_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace rp {

namespace {
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;   // per device: the address of `scratch` there (nullptr: not looked up)
std::vector<std::string> g_scratch_err;
}  // namespace

// Looks up the device address of the no-alloc scratch on `device` (which must be current), once per device.
void noalloc_scratch_init(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) {
        g_scratch_ptr.resize(device + 1, nullptr);
        g_scratch_err.resize(device + 1);
    }
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
        return;
    }
    g_scratch_ptr[device] = static_cast<double *>(p);
    g_scratch_err[device].clear();
}

// The scratch address on `device`, or nullptr with the reason in err.
double *noalloc_scratch(int device, std::string &err) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= (int)g_scratch_ptr.size() || !g_scratch_ptr[device]) {
        err = device >= 0 && device < (int)g_scratch_err.size() && !g_scratch_err[device].empty()
                  ? g_scratch_err[device]
                  : std::string("no-alloc scratch not set up on this device");
        return nullptr;
    }
    return g_scratch_ptr[device];
}

int prepare_device(rp_fit *fit, FitDevice &dv, int precision) {
    Ctx *ctx = fit->ctx;
    hipStream_t st = ctx->stream;
    const Setup &s = fit->setup;
    const int N = fit->n_atoms;
    dv.K = (int)s.kcoefs.size();
    dv.KP = dv.K > 0 ? (int)round_up(2 * dv.K, precision == RP_PREC_FP64 ? 16 : 64) : 0;
    dv.NP = (int)round_up(N + 2, 128);
    dv.S = (int)(fit->shifts_pruned.size() / 3);
    const bool rebuild = !dv.ready || dv.shifts_host != fit->shifts_pruned;
    if (!dv.ready) dv.sep_kbox_ok = false;
    dv.geom.nx = s.nx;
    dv.geom.ny = s.ny;
    dv.geom.nz = s.nz;
    std::copy(s.axis, s.axis + 9, dv.geom.ax);
    if (rebuild && dv.K > 0) {
        std::vector<double2> ex, ey, ez;
        phase_table(0, 7, s.nx, ex);
        phase_table(-7, 7, s.ny, ey);
        phase_table(-7, 7, s.nz, ez);
        RP_CHECK(dv.ex.grow(ctx, sizeof(double2) * ex.size()));
        RP_CHECK(dv.ey.grow(ctx, sizeof(double2) * ey.size()));
        RP_CHECK(dv.ez.grow(ctx, sizeof(double2) * ez.size()));
        RP_HIP(ctx, hipMemcpyAsync(dv.ex.p, ex.data(), sizeof(double2) * ex.size(), hipMemcpyHostToDevice, st));
        RP_HIP(ctx, hipMemcpyAsync(dv.ey.p, ey.data(), sizeof(double2) * ey.size(), hipMemcpyHostToDevice, st));
        RP_HIP(ctx, hipMemcpyAsync(dv.ez.p, ez.data(), sizeof(double2) * ez.size(), hipMemcpyHostToDevice, st));
        RP_CHECK(dv.khkl.grow(ctx, sizeof(int) * s.khkl.size()));
        RP_CHECK(dv.kcoef.grow(ctx, sizeof(double) * s.kcoefs.size()));
        RP_CHECK(dv.kvecs.grow(ctx, sizeof(double) * s.kvecs.size()));
        RP_HIP(ctx, hipMemcpyAsync(dv.khkl.p, s.khkl.data(), sizeof(int) * s.khkl.size(), hipMemcpyHostToDevice, st));
        RP_HIP(ctx, hipMemcpyAsync(dv.kcoef.p, s.kcoefs.data(), sizeof(double) * s.kcoefs.size(), hipMemcpyHostToDevice, st));
        RP_HIP(ctx, hipMemcpyAsync(dv.kvecs.p, s.kvecs.data(), sizeof(double) * s.kvecs.size(), hipMemcpyHostToDevice, st));
    }
    if (rebuild) {
        // fractional atom coordinates R . inv(box) and interplanar spacings, for the nearest-image test
        const double *B = s.box;
        double cof[9];
        cof[0] = B[4] * B[8] - B[5] * B[7];
        cof[1] = B[5] * B[6] - B[3] * B[8];
        cof[2] = B[3] * B[7] - B[4] * B[6];
        cof[3] = B[2] * B[7] - B[1] * B[8];
        cof[4] = B[0] * B[8] - B[2] * B[6];
        cof[5] = B[1] * B[6] - B[0] * B[7];
        cof[6] = B[1] * B[5] - B[2] * B[4];
        cof[7] = B[2] * B[3] - B[0] * B[5];
        cof[8] = B[0] * B[4] - B[1] * B[3];
        const double det = B[0] * cof[0] + B[1] * cof[1] + B[2] * cof[2];
        // inv(box)[c][i] = cof[3*i + c] / det ; frac_i = sum_c R_c inv[c][i]
        std::vector<double> af(3 * (size_t)N);
        for (int a = 0; a < N; ++a)
            for (int i = 0; i < 3; ++i) {
                double acc = 0.0;
                for (int c = 0; c < 3; ++c) acc += fit->atom_xyz[3 * a + c] * cof[3 * i + c] / det;
                af[3 * a + i] = acc;
            }
        RP_CHECK(dv.afrac.grow(ctx, sizeof(double) * af.size()));
        RP_HIP(ctx, hipMemcpyAsync(dv.afrac.p, af.data(), sizeof(double) * af.size(), hipMemcpyHostToDevice, st));
        RP_HIP(ctx, hipStreamSynchronize(st));
        for (int i = 0; i < 3; ++i) {
            const double *cr = cof + 3 * i;  // |a_j x a_k| for the plane family of axis i
            // shrink by a relative 1e-12 so rounding can only make the test more conservative
            dv.epi.h[i] = std::fabs(det) / std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) * (1.0 - 1e-12);
        }
        std::copy(B, B + 9, dv.epi.box);
        for (int i = 0; i < 3; ++i) dv.epi.nmax[i] = s.nmax[i];
        // the cell's metric in triangular form, for the sphere enumeration of the multi-image sum
        {
            double G[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    G[i][j] = B[3 * i] * B[3 * j] + B[3 * i + 1] * B[3 * j + 1] + B[3 * i + 2] * B[3 * j + 2];
            EpiArgs &e = dv.epi;
            e.q00 = G[0][0];
            e.a01 = G[0][1] / G[0][0];
            e.a02 = G[0][2] / G[0][0];
            e.q11 = G[1][1] - G[0][1] * e.a01;
            e.a12 = (G[1][2] - G[0][2] * e.a01) / e.q11;
            e.q22 = G[2][2] - e.q00 * e.a02 * e.a02 - e.q11 * e.a12 * e.a12;
            if (!(e.q00 > 0.0 && e.q11 > 0.0 && e.q22 > 0.0)) return ctx->fail(RP_EINVAL, "degenerate cell");
            e.inv_q00 = 1.0 / e.q00;
            e.inv_q11 = 1.0 / e.q11;
            e.reach2 = s.r_cutoff / std::sqrt(e.q22) + 1e-9;
        }
        dv.epi.inv_n[0] = 1.0 / s.nx;
        dv.epi.inv_n[1] = 1.0 / s.ny;
        dv.epi.inv_n[2] = 1.0 / s.nz;
    }
    if (rebuild) {
        RP_CHECK(dv.shifts.grow(ctx, sizeof(double) * fit->shifts_pruned.size()));
        RP_HIP(ctx, hipMemcpyAsync(dv.shifts.p, fit->shifts_pruned.data(), sizeof(double) * fit->shifts_pruned.size(),
                                   hipMemcpyHostToDevice, st));
        dv.shifts_host = fit->shifts_pruned;
    }
    dv.epi.geom = dv.geom;
    dv.epi.idx = fit->d_idx.as<int>();
    dv.epi.v = fit->d_v.as<double>();
    dv.epi.N = N;
    dv.epi.fit_type = s.fit_type;
    dv.epi.atoms = fit->d_atoms.as<double>();
    dv.epi.afrac = dv.afrac.as<double>();
    dv.epi.r_cut = s.r_cutoff;
    dv.epi.shifts = dv.shifts.as<double>();
    dv.epi.S = dv.S;
    dv.epi.R2 = s.r_cutoff * s.r_cutoff;
    dv.epi.sqrt_alpha = s.sqrt_alpha;

    // fp64 chunking (also used by the fp64 matvec of a split-precision fit): at most 65536 points
    // per chunk, at most ~2 GiB of chunk buffers
    const ChunkPlan cp = plan_chunks(fit->n_keep, N, dv.K, RP_PREC_FP64);
    dv.rows = cp.rows;
    const int nt = dv.NP / 128;
    dv.tiles.clear();
    for (int bi = 0; bi < nt; ++bi)
        for (int bj = bi; bj < nt; ++bj) dv.tiles.push_back(make_int2(bi, bj));
    dv.splits = cp.splits;
    if (rebuild || dv.d_tiles.bytes < sizeof(int2) * dv.tiles.size()) {
        RP_CHECK(dv.d_tiles.grow(ctx, sizeof(int2) * dv.tiles.size()));
        RP_HIP(ctx, hipMemcpyAsync(dv.d_tiles.p, dv.tiles.data(), sizeof(int2) * dv.tiles.size(), hipMemcpyHostToDevice, st));
    }
    dv.ready = true;
    return RP_OK;
}

int device_setup(rp_fit *fit, int precision, FitDevice *&dv) {
    if (!fit->dev) fit->dev = new FitDevice();
    dv = fit->dev;
    return prepare_device(fit, *dv, precision);
}

int run_accumulate(rp_fit *fit) {
    Ctx *ctx = fit->ctx;
    hipStream_t st = ctx->stream;
    FitDevice *dvp = nullptr;
    RP_CHECK(device_setup(fit, fit->opt.precision, dvp));
    FitDevice &dv = *dvp;
    const int64_t g2 = (int64_t)dv.NP * dv.NP;
    RP_CHECK(fit->d_gram.grow(ctx, sizeof(double) * g2));
    if (!fit->ev_acc0) RP_HIP(ctx, hipEventCreate(&fit->ev_acc0));
    if (!fit->ev_acc1) RP_HIP(ctx, hipEventCreate(&fit->ev_acc1));
    fit->accumulated = false;
    fit->phi_kept = false;
    fit->acc_timing_pending = false;
    RP_HIP(ctx, hipEventRecord(fit->ev_acc0, st));
    k_zero<<<ceil_div(g2, 256), 256, 0, st>>>(fit->d_gram.as<double>(), g2);
    const int prec = fit->opt.precision;
    int rc = prec == RP_PREC_FP64        ? accumulate_f64(fit, dv)
             : prec == RP_PREC_FP64_FACT ? accumulate_f64_factorized(fit, dv)
                                         : accumulate_split(fit, dv);
    if (rc != RP_OK) return rc;
    k_symmetrize<<<ceil_div(g2, 256), 256, 0, st>>>(fit->d_gram.as<double>(), dv.NP, 128);
    RP_HIP(ctx, hipGetLastError());
    RP_HIP(ctx, hipEventRecord(fit->ev_acc1, st));
    // no synchronisation: the Gram stays on the stream for the solve (or the next read-back), and
    // the phase time is read when that synchronises
    fit->acc_timing_pending = true;
    fit->na = fit->n_atoms + 2;
    fit->accumulated = true;
    return RP_OK;
}

int run_phi_points(rp_fit *fit, int64_t m, const double *grid_xyz, double *phi_out) {
    Ctx *ctx = fit->ctx;
    hipStream_t st = ctx->stream;
    FitDevice *dvp = nullptr;
    RP_CHECK(device_setup(fit, fit->opt.precision, dvp));
    FitDevice &dv = *dvp;
    const int N = fit->n_atoms;
    DBuf &d_g = fit->ws[20], &d_out = fit->ws[21];
    RP_CHECK(d_g.grow(ctx, sizeof(double) * 3 * std::max<int64_t>(m, 1)));
    RP_CHECK(d_out.grow(ctx, sizeof(double) * std::max<int64_t>(m * N, 1)));
    RP_HIP(ctx, hipMemcpyAsync(d_g.p, grid_xyz, sizeof(double) * 3 * m, hipMemcpyHostToDevice, st));
    if (m * N > 0)
        k_phi_points<<<ceil_div(m * N, 256), 256, 0, st>>>(m, d_g.as<double>(), N, fit->d_atoms.as<double>(),
                                                           fit->setup.fit_type == 1 ? dv.K : 0, dv.kvecs.as<double>(),
                                                           dv.kcoef.as<double>(), dv.epi, d_out.as<double>());
    RP_HIP(ctx, hipGetLastError());
    RP_HIP(ctx, hipMemcpyAsync(phi_out, d_out.p, sizeof(double) * m * N, hipMemcpyDeviceToHost, st));
    RP_HIP(ctx, hipStreamSynchronize(st));
    return RP_OK;
}

// rp_fit_phi without device allocation: the k-set, the pruned shifts and the atoms are uploaded into the
// static scratch, then the points go through it in blocks (points in, basis rows out), each block the
// same k_phi_points launch as rp_fit_phi's, so the rows are bitwise the same.
int run_phi_points_noalloc(rp_fit *fit, int64_t m, const double *grid_xyz, double *phi_out) {
    Ctx *ctx = fit->ctx;
    hipStream_t st = ctx->stream;
    const Setup &s = fit->setup;
    const int N = fit->n_atoms;
    const int K = s.fit_type == 1 ? (int)s.kcoefs.size() : 0;
    const int S = (int)(fit->shifts_pruned.size() / 3);
    if (N > RP_PHI_NOALLOC_MAX_ATOMS || K > RP_PHI_NOALLOC_MAX_K || S > RP_PHI_NOALLOC_MAX_SHIFTS)
        return ctx->fail(RP_EINVAL, "rp_fit_phi_noalloc: " + std::to_string(N) + " atoms, " + std::to_string(K) +
                                        " k-vectors, " + std::to_string(S) + " shifts exceed the static scratch (at most " +
                                        std::to_string(RP_PHI_NOALLOC_MAX_ATOMS) + ", " + std::to_string(RP_PHI_NOALLOC_MAX_K) +
                                        ", " + std::to_string(RP_PHI_NOALLOC_MAX_SHIFTS) + ")");
    std::string err;
    double *base = noalloc_scratch(ctx->device, err);
    if (!base) return ctx->fail(RP_ESTATE, "rp_fit_phi_noalloc: " + err);
    // layout in doubles, every region on a 256-byte boundary: kvecs 3K | kcoef K | shifts 3S | atoms 3N |
    // points 3B | basis rows B N
    const int64_t cap = RP_PHI_NOALLOC_SCRATCH_BYTES / sizeof(double);
    const int64_t o_kc = round_up(3 * (int64_t)K, 32), o_sh = o_kc + round_up(K, 32);
    const int64_t o_at = o_sh + round_up(3 * (int64_t)S, 32), o_pt = o_at + round_up(3 * (int64_t)N, 32);
    const int64_t B = (cap - o_pt - 32) / (3 + N);   // round_up(3B, 32) + B N <= cap - o_pt
    const int64_t o_out = o_pt + round_up(3 * B, 32);
    if (B < 1 || o_out + B * N > cap) return ctx->fail(RP_EINVAL, "rp_fit_phi_noalloc: static scratch too small");
    if (K > 0) {
        RP_HIP(ctx, hipMemcpyAsync(base, s.kvecs.data(), sizeof(double) * 3 * K, hipMemcpyHostToDevice, st));
        RP_HIP(ctx, hipMemcpyAsync(base + o_kc, s.kcoefs.data(), sizeof(double) * K, hipMemcpyHostToDevice, st));
    }
    if (S > 0)
        RP_HIP(ctx, hipMemcpyAsync(base + o_sh, fit->shifts_pruned.data(), sizeof(double) * 3 * S, hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipMemcpyAsync(base + o_at, fit->atom_xyz.data(), sizeof(double) * 3 * N, hipMemcpyHostToDevice, st));
    EpiArgs e{};
    e.N = N;
    e.fit_type = s.fit_type;
    e.atoms = base + o_at;
    e.shifts = base + o_sh;
    e.S = S;
    e.R2 = s.r_cutoff * s.r_cutoff;
    e.sqrt_alpha = s.sqrt_alpha;
    e.r_cut = s.r_cutoff;
    for (int64_t p0 = 0; p0 < m; p0 += B) {
        const int64_t nb = std::min<int64_t>(B, m - p0);
        // stream order: this block's upload follows the previous block's download of the rows
        RP_HIP(ctx, hipMemcpyAsync(base + o_pt, grid_xyz + 3 * p0, sizeof(double) * 3 * nb, hipMemcpyHostToDevice, st));
        k_phi_points<<<ceil_div(nb * N, 256), 256, 0, st>>>(nb, base + o_pt, N, base + o_at, K, base, base + o_kc, e,
                                                             base + o_out);
        RP_HIP(ctx, hipGetLastError());
        RP_HIP(ctx, hipMemcpyAsync(phi_out + p0 * N, base + o_out, sizeof(double) * nb * N, hipMemcpyDeviceToHost, st));
    }
    RP_HIP(ctx, hipStreamSynchronize(st));
    return RP_OK;
}

// Validates a system, derives the host setup, uploads atoms and the cube, and runs the vdW filter.
// Reuses the fit's device buffers (grow-only), so reloading a fit with a system of equal or smaller
// size allocates nothing.
int load_system(rp_fit *fit, const rp_system *sys, const rp_options *opt) {
    Ctx *ctx = fit->ctx;
    if (sys->n_atoms < 1 || !sys->atomic_number || !sys->atom_xyz || !sys->esp)
        return ctx->fail(RP_EINVAL, "rp_fit_create: system needs atoms, positions and ESP");
    if (sys->esp_dtype != RP_F64 && sys->esp_dtype != RP_F32) return ctx->fail(RP_EINVAL, "bad esp_dtype");
    if (sys->esp_order != RP_ORDER_CUBE && sys->esp_order != RP_ORDER_FLAT) return ctx->fail(RP_EINVAL, "bad esp_order");
    if (opt->precision != RP_PREC_FP64 && opt->precision != RP_PREC_OZAKI && opt->precision != RP_PREC_OZAKI_REFINED &&
        opt->precision != RP_PREC_FP64_FACT)
        return ctx->fail(RP_EINVAL, "bad precision");
    fit->opt = *opt;
    fit->n_atoms = sys->n_atoms;
    fit->Z.assign(sys->atomic_number, sys->atomic_number + sys->n_atoms);
    fit->atom_xyz.assign(sys->atom_xyz, sys->atom_xyz + 3 * sys->n_atoms);
    fit->setup = Setup();
    fit->n_keep = 0;
    fit->v_mean = 0.0;
    fit->shifts_pruned.clear();
    fit->accumulated = false;
    fit->lu_fallback = false;
    fit->phi_kept = false;
    fit->refine_stream = false;
    fit->refine_steps_done = 0;
    fit->refine_last_dq = 0.0;
    fit->refine_bound = -1.0;
    fit->sensitivity = -1.0;
    fit->solver = RP_SOLVER_CHOLESKY;
    fit->na = 0;
    if (fit->dev) fit->dev->ready = false;
    RP_CHECK(build_setup(sys, opt, fit->setup, ctx->err));
    RP_CHECK(fit->d_atoms.grow(ctx, sizeof(double) * 3 * sys->n_atoms));
    RP_HIP(ctx, hipMemcpyAsync(fit->d_atoms.p, sys->atom_xyz, sizeof(double) * 3 * sys->n_atoms, hipMemcpyHostToDevice,
                               ctx->stream));
    RP_CHECK(upload_filter(fit, sys));
    return run_filter(fit);
}

}  // namespace rp

using namespace rp;

rp_fit::~rp_fit() {
    delete dev;
    if (ev_acc0) (void)hipEventDestroy(ev_acc0);
    if (ev_acc1) (void)hipEventDestroy(ev_acc1);
    if (h_stage) (void)hipHostFree(h_stage);
    if (h_small) (void)hipHostFree(h_small);
}

static void destroy_ctx(rp_ctx *ctx) {
    (void)hipSetDevice(ctx->device);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    delete ctx;
}

extern "C" {

int rp_ctx_create(int device, uint64_t hbm_budget_bytes, rp_ctx **out) {
    if (!out) {
        set_global_error("rp_ctx_create: out is NULL");
        return RP_EINVAL;
    }
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || device < 0 || device >= n) {
        set_global_error("rp_ctx_create: no such HIP device");
        return RP_EINVAL;
    }
    e = hipSetDevice(device);
    if (e != hipSuccess) {
        set_global_error(std::string("hipSetDevice failed: ") + hipGetErrorString(e));
        return RP_EHIP;
    }
    noalloc_scratch_init(device);   // the device address of rp_fit_phi_noalloc's static scratch
    rp_ctx *c = new rp_ctx();
    c->device = device;
    c->budget = hbm_budget_bytes;
    (void)hipGetDeviceProperties(&c->prop, device);
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        set_global_error(std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
        return RP_EHIP;
    }
    *out = c;
    return RP_OK;
}

int rp_ctx_destroy(rp_ctx *ctx) {
    if (!ctx) return RP_OK;
    if (ctx->n_fits > 0) {
        ctx->closing = true;  // the last rp_fit_destroy finishes the job
        return RP_OK;
    }
    destroy_ctx(ctx);
    return RP_OK;
}

const char *rp_last_error(const rp_ctx *ctx) {
    if (ctx) return ctx->err.c_str();
    return g_err.c_str();
}

int rp_ctx_hbm_in_use(const rp_ctx *ctx, uint64_t *bytes) {
    if (!ctx || !bytes) return RP_EINVAL;
    *bytes = ctx->in_use;
    return RP_OK;
}

int rp_ctx_profile(rp_ctx *ctx, int enable) { return rp_ctx_profile_kinds(ctx, enable ? (1u << KT_N) - 1u : 0u); }

int rp_ctx_profile_kinds(rp_ctx *ctx, uint32_t kinds) {
    if (!ctx) return RP_EINVAL;
    ctx->profile = kinds & ((1u << KT_N) - 1u);
    for (int k = 0; k < KT_N; ++k) {
        ctx->kt_ms[k] = 0.0;
        ctx->kt_n[k] = 0;
    }
    return RP_OK;
}

int rp_ctx_kernel_stats(const rp_ctx *ctx, int kind, double *total_ms, int64_t *launches) {
    if (!ctx || kind < 0 || kind >= KT_N || !total_ms || !launches) return RP_EINVAL;
    *total_ms = ctx->kt_ms[kind];
    *launches = ctx->kt_n[kind];
    return RP_OK;
}

int rp_fit_refit(rp_fit *fit, const rp_constraints *con, double *q_out) {
    if (!fit || !con || !q_out) return fit ? fit->ctx->fail(RP_EINVAL, "rp_fit_refit: NULL argument") : RP_EINVAL;
    RP_HIP(fit->ctx, hipSetDevice(fit->ctx->device));
    RP_CHECK(run_filter(fit));
    RP_CHECK(run_accumulate(fit));
    return run_solve(fit, con, q_out, nullptr, nullptr);
}

int rp_fit_create(rp_ctx *ctx, const rp_system *sys, const rp_options *opt, rp_fit **out) {
    if (!ctx || !sys || !opt || !out) return ctx ? ctx->fail(RP_EINVAL, "rp_fit_create: NULL argument") : RP_EINVAL;
    *out = nullptr;
    RP_HIP(ctx, hipSetDevice(ctx->device));
    rp_fit *fit = new rp_fit();
    fit->ctx = ctx;
    const int rc = load_system(fit, sys, opt);
    if (rc != RP_OK) {
        delete fit;
        return rc;
    }
    ++ctx->n_fits;
    *out = fit;
    return RP_OK;
}

int rp_fit_reset(rp_fit *fit, const rp_system *sys, const rp_options *opt) {
    if (!fit || !sys || !opt) return fit ? fit->ctx->fail(RP_EINVAL, "rp_fit_reset: NULL argument") : RP_EINVAL;
    RP_HIP(fit->ctx, hipSetDevice(fit->ctx->device));
    RP_HIP(fit->ctx, hipStreamSynchronize(fit->ctx->stream));
    return load_system(fit, sys, opt);
}

int rp_fit_destroy(rp_fit *fit) {
    if (!fit) return RP_OK;
    rp_ctx *ctx = static_cast<rp_ctx *>(fit->ctx);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete fit;
    if (--ctx->n_fits == 0 && ctx->closing) destroy_ctx(ctx);
    return RP_OK;
}

int rp_fit_info_get(const rp_fit *fit, rp_fit_info *o) {
    if (!fit || !o) return RP_EINVAL;
    const Setup &s = fit->setup;
    o->n_total = fit->n_total;
    o->n_keep = fit->n_keep;
    o->n_eff = (int64_t)s.eff_smin.size();
    o->n_atoms = fit->n_atoms;
    o->n_k = (int32_t)s.kcoefs.size();
    o->n_shift = (int32_t)(s.shifts.size() / 3);
    o->n_shift_pruned = (int32_t)(fit->shifts_pruned.size() / 3);
    o->volume = s.volume;
    o->alpha = s.alpha;
    o->sqrt_alpha = s.sqrt_alpha;
    o->v_mean = fit->v_mean;
    return RP_OK;
}

int rp_fit_grid(const rp_fit *fit, uint8_t *mask, int64_t *flat_idx, double *v, double *grid_xyz) {
    if (!fit) return RP_EINVAL;
    Ctx *ctx = fit->ctx;
    hipStream_t st = ctx->stream;
    RP_HIP(ctx, hipSetDevice(ctx->device));
    if (mask) RP_HIP(ctx, hipMemcpyAsync(mask, fit->d_mask.p, (size_t)fit->n_total, hipMemcpyDeviceToHost, st));
    std::vector<int> idx;
    if (flat_idx || grid_xyz) {
        idx.resize(fit->n_keep);
        RP_HIP(ctx, hipMemcpyAsync(idx.data(), fit->d_idx.p, sizeof(int) * fit->n_keep, hipMemcpyDeviceToHost, st));
    }
    if (v) RP_HIP(ctx, hipMemcpyAsync(v, fit->d_v.p, sizeof(double) * fit->n_keep, hipMemcpyDeviceToHost, st));
    RP_HIP(ctx, hipStreamSynchronize(st));
    const Setup &s = fit->setup;
    for (int64_t t = 0; t < (int64_t)idx.size(); ++t) {
        if (flat_idx) flat_idx[t] = idx[t];
        if (grid_xyz) {
            const int64_t f = idx[t], layer = (int64_t)s.nx * s.ny;
            const int64_t k = f / layer, rem = f - k * layer, j = rem / s.nx, i = rem - j * s.nx;
            for (int c = 0; c < 3; ++c) {
                double u = (double)i * s.axis[c];
                u = u + (double)j * s.axis[3 + c];
                grid_xyz[3 * t + c] = u + (double)k * s.axis[6 + c];
            }
        }
    }
    return RP_OK;
}

int rp_fit_ewald(const rp_fit *fit, double *kvecs, double *kcoefs, double *shifts, double *shifts_pruned) {
    if (!fit) return RP_EINVAL;
    const Setup &s = fit->setup;
    if (kvecs) std::copy(s.kvecs.begin(), s.kvecs.end(), kvecs);
    if (kcoefs) std::copy(s.kcoefs.begin(), s.kcoefs.end(), kcoefs);
    if (shifts) std::copy(s.shifts.begin(), s.shifts.end(), shifts);
    if (shifts_pruned) std::copy(fit->shifts_pruned.begin(), fit->shifts_pruned.end(), shifts_pruned);
    return RP_OK;
}

int rp_fit_accumulate(rp_fit *fit) {
    if (!fit) return RP_EINVAL;
    RP_HIP(fit->ctx, hipSetDevice(fit->ctx->device));
    return run_accumulate(fit);
}

int rp_fit_normal_equations(const rp_fit *fit, double *ata, double *atb, double *sum_phi, double *sum_v) {
    if (!fit) return RP_EINVAL;
    Ctx *ctx = fit->ctx;
    if (!fit->accumulated) return ctx->fail(RP_ESTATE, "rp_fit_normal_equations before rp_fit_accumulate");
    RP_HIP(ctx, hipSetDevice(ctx->device));
    const int N = fit->n_atoms;
    const int NP = (int)round_up(N + 2, 128);
    std::vector<double> g((size_t)NP * NP);
    RP_HIP(ctx, hipMemcpyAsync(g.data(), fit->d_gram.p, sizeof(double) * g.size(), hipMemcpyDeviceToHost, ctx->stream));
    RP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int a = 0; a < N; ++a) {
        if (ata)
            for (int b = 0; b < N; ++b) ata[(int64_t)a * N + b] = g[(size_t)a * NP + b];
        if (atb) atb[a] = g[(size_t)a * NP + N];
        if (sum_phi) sum_phi[a] = g[(size_t)a * NP + N + 1];
    }
    if (sum_v) {
        sum_v[0] = g[(size_t)N * NP + N + 1];
    }
    return RP_OK;
}

int rp_fit_solve(rp_fit *fit, const rp_constraints *con, double *q_out, double *solution_out, int32_t *used_fallback) {
    if (!fit || !con || !q_out) return fit ? fit->ctx->fail(RP_EINVAL, "rp_fit_solve: NULL argument") : RP_EINVAL;
    RP_HIP(fit->ctx, hipSetDevice(fit->ctx->device));
    return run_solve(fit, con, q_out, solution_out, used_fallback);
}

int rp_fit_phi_matvec(rp_fit *fit, const double *q, double *v_coul) {
    if (!fit || !q || !v_coul) return RP_EINVAL;
    RP_HIP(fit->ctx, hipSetDevice(fit->ctx->device));
    return run_phi_matvec(fit, q, v_coul);
}

int rp_fit_phi(rp_fit *fit, int64_t m, const double *grid_xyz, double *phi_out) {
    if (!fit || m < 0 || (m > 0 && (!grid_xyz || !phi_out))) return RP_EINVAL;
    RP_HIP(fit->ctx, hipSetDevice(fit->ctx->device));
    return run_phi_points(fit, m, grid_xyz, phi_out);
}

int rp_fit_phi_noalloc(rp_fit *fit, int64_t m, const double *grid_xyz, double *phi_out) {
    if (!fit || m < 0 || (m > 0 && (!grid_xyz || !phi_out))) return RP_EINVAL;
    RP_HIP(fit->ctx, hipSetDevice(fit->ctx->device));
    return run_phi_points_noalloc(fit, m, grid_xyz, phi_out);
}

int rp_fit_refine_stats(const rp_fit *fit, int32_t *steps, double *last_correction) {
    if (!fit || !steps || !last_correction) return RP_EINVAL;
    *steps = fit->refine_steps_done;
    *last_correction = fit->refine_last_dq;
    return RP_OK;
}

int rp_fit_solve_report(const rp_fit *fit, rp_solve_report *out) {
    if (!fit || !out) return RP_EINVAL;
    const bool refined = fit->opt.precision == RP_PREC_OZAKI_REFINED;
    out->solver = fit->solver;
    out->refine_steps = fit->refine_steps_done;
    out->last_correction = fit->refine_last_dq;
    out->error_bound = refined ? fit->refine_bound : -1.0;
    out->within_target = refined ? (fit->refine_bound >= 0.0 && fit->refine_bound <= RP_REFINE_TARGET_E) : 1;
    out->sensitivity = fit->sensitivity;
    out->well_conditioned = fit->sensitivity >= 0.0 && fit->sensitivity <= RP_SENSITIVITY_LIMIT_E;
    return RP_OK;
}

int rp_fit_timings(const rp_fit *fit, double *t3) {
    if (!fit || !t3) return RP_EINVAL;
    const_cast<rp_fit *>(fit)->resolve_timings(true);
    t3[0] = fit->t_filter_ms;
    t3[1] = fit->t_accum_ms;
    t3[2] = fit->t_solve_ms;
    return RP_OK;
}

int rp_chunk_plan(int64_t n_keep, int32_t n_atoms, int32_t n_k, int32_t precision, rp_chunk_plan_info *out) {
    if (!out || n_keep < 1 || n_atoms < 1 || n_k < 0 || precision < RP_PREC_FP64 || precision > RP_PREC_FP64_FACT)
        return RP_EINVAL;
    const ChunkPlan c = plan_chunks(n_keep, n_atoms, n_k, precision);
    out->rows = c.rows;
    out->n_chunks = ceil_div(n_keep, c.rows);
    out->np = c.NP;
    out->kp = c.KP;
    out->splits = c.splits;
    out->n_tiles = c.n_tiles;
    const bool f64_gram = precision == RP_PREC_FP64 || precision == RP_PREC_FP64_FACT;
    out->slices = f64_gram ? 0 : c.slices;
    out->stage_rows = f64_gram ? f64_gram_stage_rows : gram_stage_rows(c.slices);
    return RP_OK;
}

}  // extern "C"
