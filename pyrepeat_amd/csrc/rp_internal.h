// Internal declarations shared by the librepeat_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/repeat_hip.h"

namespace rp {

// Last error of calls made without a context (e.g. rp_ctx_create failures).
void set_global_error(const std::string &msg);

struct Ctx;

// Kernel classes timed with HIP events when profiling is on (rp_ctx_profile).
enum { KT_KSPACE = 0, KT_GRAM = 1, KT_FILTER = 2, KT_SOLVE = 3, KT_PHASE = 4, KT_SLICE = 5, KT_REDUCE = 6, KT_REFINE = 7, KT_N = 8 };

// Device allocation charged against the owning context's HBM budget.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    Ctx *ctx = nullptr;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf();
    int alloc(Ctx *c, size_t n);   // frees any previous allocation first
    // Keeps the allocation when it is already large enough (contents are not preserved on growth).
    int grow(Ctx *c, size_t n) { return (p && bytes >= n && ctx == c) ? 0 : alloc(c, n); }
    void release();
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // a second stream (default priority) for the solve's sensitivity estimate, which then runs beside
    // the refinement; made on first use, with the events that fork it from and join it to `stream`
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // Creates the side stream and its fork and join events if they are not there yet.
    int ensure_side() {
        if (!side && hipStreamCreateWithFlags(&side, hipStreamNonBlocking) != hipSuccess)
            return fail(RP_EHIP, "hipStreamCreateWithFlags failed for the side stream");
        if (!ev_fork && hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming) != hipSuccess)
            return fail(RP_EHIP, "hipEventCreateWithFlags failed");
        if (!ev_join && hipEventCreateWithFlags(&ev_join, hipEventDisableTiming) != hipSuccess)
            return fail(RP_EHIP, "hipEventCreateWithFlags failed");
        return RP_OK;
    }
    uint64_t budget = 0;
    uint64_t in_use = 0;
    std::string err;
    hipDeviceProp_t prop{};
    uint32_t profile = 0;  // bit k: time launches of kind k
    int n_fits = 0;        // live fits on this context
    bool closing = false;  // rp_ctx_destroy was called while fits were alive
    double kt_ms[KT_N] = {};
    int64_t kt_n[KT_N] = {};
    struct Pending {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    // Asked for more budget when a fit could use it (the refined form's kept basis); returns whether the
    // budget grew by that much.  Unset: the budget is fixed.
    std::function<bool(uint64_t)> grow_budget;
    int fail(int code, const std::string &msg) {
        err = msg;
        return code;
    }
    // Adds the elapsed times of recorded launches (call after the stream is synchronised).
    void collect() {
        size_t keep = 0;
        for (auto &p : pending) {
            if (hipEventQuery(p.b) == hipErrorNotReady) {
                pending[keep++] = p;
                continue;
            }
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                kt_ms[p.kind] += ms;
                kt_n[p.kind] += 1;
            }
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        pending.resize(keep);
    }
};

// Scoped event pair around kernel launches on ctx->stream (no-op unless profiling).
struct KTimer {
    Ctx *c;
    int kind;
    hipEvent_t a = nullptr, b = nullptr;
    KTimer(Ctx *ctx, int k) : c(ctx), kind(k) {
        if ((c->profile >> kind & 1u) && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
            (void)hipEventRecord(a, c->stream);
        else
            a = b = nullptr;
    }
    ~KTimer() {
        if (a && b) {
            (void)hipEventRecord(b, c->stream);
            c->pending.push_back({kind, a, b});
        }
    }
};

// Device-resident inputs of the vdW filter (kept so a fit can be re-run from HBM).
struct FilterArgs {
    int nx, ny, nz;
    double ax[9];
    double bmin[3];
    double binv;
    int bn[3];
    int fallback;          // 1 if some r_max exceeds the bin edge
    int64_t n_all;
};

struct FilterState {
    FilterArgs fa{};
    int64_t n_blocks = 0;
    int esp_dtype = 0, esp_order = 0;
    DBuf bin_start, bin_atoms, bin_smax, all, all_smax, cnt, off, vsum, mean, esp, axis, bounds;
};

// Checks a HIP call; records the error on ctx and returns RP_EHIP from the caller.
#define RP_HIP(ctx, call)                                                                        \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return (ctx)->fail(RP_EHIP, std::string(#call " failed: ") + hipGetErrorString(e_)); \
    } while (0)

#define RP_CHECK(expr)            \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != RP_OK) return rc_; \
    } while (0)

// Exclusive scan of n per-block counts by one work-group of 1024 threads: out[b] = cnt[0] + ... + cnt[b-1],
// out[n] = the total.  Thread t sums its run of ceil(n / 1024) counts; the 1024 run sums are scanned by
// wave shuffles and a scan of the 16 wave totals; each thread then writes its run's offsets.  Integer
// sums, so the result does not depend on the order.
__device__ inline void block_scan_counts_1024(const int *__restrict__ cnt, int64_t n, int64_t *__restrict__ out) {
    __shared__ int64_t wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t per = (n + 1023) / 1024;
    const int64_t b0 = min((int64_t)t * per, n), b1 = min(b0 + per, n);
    int64_t s = 0;
    for (int64_t b = b0; b < b1; ++b) s += cnt[b];
    int64_t x = s;   // inclusive scan within the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (t < 64) {
        int64_t v = t < 16 ? wsum[t] : 0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const int64_t y = __shfl_up(v, o, 64);
            if (lane >= o) v += y;
        }
        if (t < 16) wsum[t] = v;
    }
    __syncthreads();
    const int64_t base = w ? wsum[w - 1] : 0;
    if (t == 1023) out[n] = base + x;
    int64_t run = base + x - s;
    for (int64_t b = b0; b < b1; ++b) {
        out[b] = run;
        run += cnt[b];
    }
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

// Chunk geometry of the basis streaming for M kept points, N atoms and K k-vectors: the accumulate
// paths allocate from it, and the serving loop's memory bound uses the same numbers.
struct ChunkPlan {
    int KP = 0;          // padded [cos | sin] width of the k-space operands
    int NP = 0;          // padded atom width (N + V column + ones column)
    int64_t rows = 0;    // kept points per chunk
    int n_tiles = 0;     // Gram tiles touching the upper triangle
    int splits = 1;      // row splits of the Gram per chunk
    int slices = 4;      // split form: int8 slices per Phi element
};

// Rows of one stage of the split Gram (k_oz_gram) and of the fp64 Gram (k_syrk_f64): a row split
// must hold a whole number of them.
constexpr int gram_stage_rows(int slices) { return slices == 3 ? 128 : 64; }
constexpr int f64_gram_stage_rows = 16;

inline ChunkPlan plan_chunks(int64_t M, int N, int K, int precision) {
    ChunkPlan c;
    c.NP = (int)round_up(N + 2, 128);
    if (precision == RP_PREC_FP64) {
        c.KP = K > 0 ? (int)round_up(2 * K, 16) : 0;
        const int64_t per_row = 8 * (int64_t)(c.KP + c.NP);
        int64_t rows = std::min<int64_t>(65536, round_up(M, 128));
        while (rows > 128 && rows * per_row > ((int64_t)1 << 31)) rows /= 2;
        c.rows = std::max<int64_t>(128, round_up(rows, 128));
        const int nt = c.NP / 128;
        c.n_tiles = nt * (nt + 1) / 2;
        while ((int64_t)c.n_tiles * c.splits < 512 && c.rows / (c.splits * 2) >= 1024 && (c.rows / (c.splits * 2)) % 64 == 0)
            c.splits *= 2;
    } else if (precision == RP_PREC_FP64_FACT) {
        // factorized basis (chunks of whole 512-row blocks, column-major) and the fp64 Gram's tiles
        c.slices = 0;
        int64_t rows = std::max<int64_t>(512, std::min<int64_t>(65536, round_up(M, 512)));
        while (rows > 512 && rows * 8 * c.NP > ((int64_t)1 << 31)) rows /= 2;
        c.rows = std::max<int64_t>(512, rows - rows % 512);
        const int nt = c.NP / 128;
        c.n_tiles = nt * (nt + 1) / 2;
        while ((int64_t)c.n_tiles * c.splits < 512 && c.rows / (c.splits * 2) >= 1024 &&
               (c.rows / (c.splits * 2)) % f64_gram_stage_rows == 0)
            c.splits *= 2;
    } else {
        c.KP = 0;  // the factorized reciprocal sum needs no k-space operand
        // refined: 3 slices (the refinement removes their error); otherwise 4
        c.slices = precision == RP_PREC_OZAKI_REFINED ? 3 : 4;
        int64_t rows = std::min<int64_t>(65536, round_up(M, 512));
        rows = std::max<int64_t>(512, rows);
        // bound the chunk buffers (Phi fp64, Phi slices) to about 2 GiB
        while (rows > 512 && rows * ((8LL + c.slices) * c.NP) > ((int64_t)1 << 31)) rows /= 2;
        c.rows = std::max<int64_t>(512, rows - rows % 512);
        c.n_tiles = 0;
        for (int bi = 0; bi < c.NP / 64; ++bi)
            for (int bj = 0; bj < c.NP / 128; ++bj)
                if (bi * 64 <= bj * 128 + 127) ++c.n_tiles;
        // every row split must be whole stages of the Gram: 128 rows deep with 3 slices, 64 with 4
        const int64_t kd = gram_stage_rows(c.slices);
        auto row_splits = [&](int64_t r) {
            int sp = 1;
            while ((int64_t)c.n_tiles * sp < 1024 && r / (sp * 2) >= 1024 && (r / (sp * 2)) % kd == 0) sp *= 2;
            return sp;
        };
        c.splits = row_splits(c.rows);
        // more than one chunk: equal chunks of whole 2048-row blocks where that keeps the row splits.  The
        // chunk length is then no power of two, so the residual pass's reads across all columns at one row
        // offset (columns one chunk length apart) spread over the memory channels.
        if (M > c.rows) {
            const int64_t nch = (M + c.rows - 1) / c.rows;
            const int64_t bal = round_up((M + nch - 1) / nch, 2048);
            if (bal < c.rows && row_splits(bal) == c.splits) c.rows = bal;
        }
    }
    return c;
}

// ---------------------------------------------------------------- geometry / Ewald (host)

struct Vec3 {
    double x, y, z;
};

// Everything the host derives from the cube and the options before the kernels run.
struct Setup {
    int nx = 0, ny = 0, nz = 0;
    double axis[9] = {};
    double box[9] = {};
    double volume = 0.0;        // signed, repeat.py:1763
    double recip[9] = {};
    double alpha = 0.0, sqrt_alpha = 0.0, r_cutoff = 20.0;
    int fit_type = 1;
    int nmax[3] = {0, 0, 0};
    // k-set in loop order (repeat.py:1127-1140)
    std::vector<int32_t> khkl;  // K x 3 integer indices
    std::vector<double> kvecs;  // K x 3
    std::vector<double> kcoefs; // K
    std::vector<double> shifts; // S x 3, full list (repeat.py:815-822)
    // filter
    std::vector<double> eff_xyz;    // E x 3
    std::vector<double> eff_smin;   // squared-distance thresholds (see filter_thresholds)
    std::vector<double> eff_smax;
    double max_rmin = 0.0, max_rmax = 0.0;
};

int build_setup(const rp_system *sys, const rp_options *opt, Setup &s, std::string &err);
double sqrt_threshold(double r);  // largest double s with fl(sqrt(s)) <= r
// Device bytes the vdW filter of this setup allocates at most (any kept count).
uint64_t filter_bytes_bound(const Setup &s, int esp_dtype);
std::vector<double> prune_shifts(const std::vector<double> &shifts, const double gmin[3], const double gmax[3],
                                 const std::vector<double> &atom_xyz, double r_cutoff);

}  // namespace rp

namespace rp {
struct FitDevice;
}

// The fit object behind the opaque rp_fit handle.
struct rp_fit {
    rp::Ctx *ctx = nullptr;
    rp::Setup setup;
    int32_t n_atoms = 0;
    std::vector<int32_t> Z;
    std::vector<double> atom_xyz;
    rp_options opt{};
    int64_t n_total = 0;
    int64_t n_keep = 0;
    double v_mean = 0.0;
    std::vector<double> shifts_pruned;
    bool accumulated = false;
    bool lu_fallback = false;     // last solve used LU (A not positive definite)
    int64_t kspace_products = 0;  // int8 MFMA products of a k-space GEMM (0: the split form factorizes it)
    double t_filter_ms = 0, t_accum_ms = 0, t_solve_ms = 0;
    // the accumulate phase is timed by two events that are read at the next synchronisation
    hipEvent_t ev_acc0 = nullptr, ev_acc1 = nullptr;
    bool acc_timing_pending = false;
    void resolve_timings(bool wait) {
        if (!acc_timing_pending) return;
        if (wait) (void)hipEventSynchronize(ev_acc1);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev_acc0, ev_acc1) == hipSuccess) {
            t_accum_ms = ms;
            acc_timing_pending = false;
        }
    }

    // device state
    rp::FilterState filt;
    rp::DBuf d_mask;      // n_total bytes
    rp::DBuf d_idx;       // n_keep int32 flat indices
    rp::DBuf d_v;         // n_keep doubles (mean subtracted)
    rp::DBuf d_atoms;     // N x 3 doubles
    rp::DBuf d_gram;      // NA x NA doubles, NA = N + 2 (augmented with V and 1)
    int32_t na = 0;       // augmented width N + 2
    // grow-only workspaces reused by every accumulate / solve / matvec of this fit
    rp::DBuf ws[32];
    void *h_stage = nullptr;        // mapped pinned block: the solve's constraint arrays, then its results
    void *h_stage_dev = nullptr;    // its device address
    // mapped pinned slots that kernels write results into for the host (filter count, mean and
    // bounds; chunk line ranges): a synchronisation returns them without a copy
    double *h_small = nullptr, *h_small_dev = nullptr;
    size_t h_small_bytes = 0;
    int host_slots(rp::Ctx *ctx, size_t bytes = 4096) {
        if (h_small_bytes >= bytes) return 0;
        if (h_small) (void)hipHostFree(h_small);
        h_small = h_small_dev = nullptr;
        h_small_bytes = 0;
        void *p = nullptr, *d = nullptr;
        if (hipHostMalloc(&p, bytes, hipHostMallocMapped) != hipSuccess) return ctx->fail(RP_ENOMEM, "hipHostMalloc failed");
        h_small = static_cast<double *>(p);
        h_small_bytes = bytes;
        if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) return ctx->fail(RP_EHIP, "hipHostGetDevicePointer failed");
        h_small_dev = static_cast<double *>(d);
        return 0;
    }
    size_t h_stage_bytes = 0;
    int split_tiles_np = 0;         // NP the split Gram's tile list in ws[9] was built for
    uint32_t bs_base = 0, bs_epoch = 0;   // triangular solves: tickets taken so far, launch epoch
    uint32_t bs2_base = 0, bs2_epoch = 0; // the same for the solves on the side stream (own counter, flags)
    // refinement (RP_PREC_OZAKI_REFINED): the accumulate keeps every chunk of the fp64 basis in ws[7]
    // ([chunk][NP][phi_rows] doubles); phi_kept says it matches the fit's current system
    bool phi_kept = false;
    bool refine_stream = false;     // refined form without the kept basis: each step recomputes it
    int64_t phi_rows = 0;
    int32_t refine_steps_done = 0;
    double refine_last_dq = 0.0;
    double refine_bound = -1.0;     // estimated error left by the refinement (e), -1: none
    double sensitivity = -1.0;      // max |dq| (e) from a perturbation of the system by 8 ulps, -1: none
    int32_t solver = RP_SOLVER_CHOLESKY;   // how the last solve went
    rp::FitDevice *dev = nullptr;   // persistent device copy of the Ewald setup
    ~rp_fit();
};

struct rp_ctx : rp::Ctx {};

namespace rp {
struct FitDevice;
// Phases implemented in the .hip files.
int load_system(rp_fit *fit, const rp_system *sys, const rp_options *opt);
int upload_filter(rp_fit *fit, const rp_system *sys);
int run_filter(rp_fit *fit);
int run_accumulate(rp_fit *fit);
int run_solve(rp_fit *fit, const rp_constraints *con, double *q_out, double *solution_out, int32_t *used_fallback);
int run_phi_matvec(rp_fit *fit, const double *q, double *v_coul);
int run_phi_points(rp_fit *fit, int64_t m, const double *grid_xyz, double *phi_out);
int run_qeq(rp_fit *fit, const double *chi, const double *j_self, double q_tot, double *q_out, double *j_out);
// Solves the n x n system held row-major in K2 (leading dimension ld, right-hand side in column n) by
// LU with partial pivoting into x; RP_ESINGULAR on an exactly zero pivot.
int lu_solve_augmented(Ctx *ctx, double *K2, int ld, int n, int *piv, int *info, double *x, hipStream_t st);
}  // namespace rp
