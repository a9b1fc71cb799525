// Normal-equation accumulation, plain fp64 form (fit_charges steps 1-3, repeat.py:2195-2262).
//
// The kept points are streamed in chunks of `rows` points.  Per chunk:
//   1. k_grid_phase  G[m][2k], G[m][2k+1] = w_k (cos, sin)(k . g_m)      (rows x KP, fp64)
//      k . g = 2 pi (h i/nx + l j/ny + n k/nz) exactly, so the phase is the
//      product of three 1-D tables built on the host from integer-reduced angles.
//   2. k_phi_gemm    Phi = G . B,  B[2k][a], B[2k+1][a] = (cos, sin)(k . R_a)   (fp64 MFMA)
//      Its epilogue adds the real-space sum over the pruned shifts,
//      erfc(sqrt(alpha) r)/r for r^2 <= Rc^2 (self image skipped at r^2 <= 1e-20),
//      and fills the two augmented columns V and 1, so that one Gram of
//      [Phi | V | 1] carries Phi^T Phi, Phi^T V, sum(Phi), sum(V) and the count.
//      cos(k.(g - R)) = cos(k.g)cos(k.R) + sin(k.g)sin(k.R) is the identity used
//      (reference: direct cos of k.(g-R), repeat.py:831-846; real space repeat.py:849-872).
//   3. k_syrk        upper-triangular 128x128 tiles of [Phi|V|1]^T [Phi|V|1], split over
//      rows for occupancy, partials reduced in a fixed order (deterministic).
#include <algorithm>
#include <cmath>

#include "rp_fit_common.h"
#include "rp_mfma.h"

namespace rp {

namespace {

constexpr int BM = 128, BN = 128, BK = 16;
static_assert(BK == f64_gram_stage_rows, "row splits of the fp64 Gram are planned in BK-row stages");

// A tile (BM x BK) from row-major src (ld), transposed into As[BK][BM]; B tile (BK x BN) row-major into Bs[BK][BN].
struct TileLoader {
    // A: thread t loads row t/2, k-segment (t%2)*8 .. +8
    static __device__ __forceinline__ void load_a(const double *src, int64_t ld, int64_t row0, int k0, double (&r)[8]) {
        const int t = threadIdx.x;
        const double2 *p = reinterpret_cast<const double2 *>(src + (row0 + (t >> 1)) * ld + k0 + (t & 1) * 8);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double2 v = p[q];
            r[2 * q] = v.x;
            r[2 * q + 1] = v.y;
        }
    }
    static __device__ __forceinline__ void store_a(double (*As)[BM], const double (&r)[8]) {
        const int t = threadIdx.x;
#pragma unroll
        for (int q = 0; q < 8; ++q) As[(t & 1) * 8 + q][t >> 1] = r[q];
    }
    // row-major (BK x BN) block: thread t loads k-row t/16, columns (t%16)*8 .. +8
    static __device__ __forceinline__ void load_b(const double *src, int64_t ld, int64_t krow0, int64_t col0, double (&r)[8]) {
        const int t = threadIdx.x;
        const double2 *p = reinterpret_cast<const double2 *>(src + (krow0 + (t >> 4)) * ld + col0 + (t & 15) * 8);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double2 v = p[q];
            r[2 * q] = v.x;
            r[2 * q + 1] = v.y;
        }
    }
    static __device__ __forceinline__ void store_b(double (*Bs)[BN], const double (&r)[8]) {
        const int t = threadIdx.x;
#pragma unroll
        for (int q = 0; q < 8; ++q) Bs[t >> 4][(t & 15) * 8 + q] = r[q];
    }
};

// 128x128 block of C = A . B accumulated in acc (4 waves, 2x2, each 64x64 = 4x4 MFMA tiles).
__device__ __forceinline__ void mma_tile(const double (*As)[BM], const double (*Bs)[BN], dvec4 (&acc)[4][4]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
        double af[4], bf[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) af[mt] = As[kk * 4 + (lane >> 4)][wm * 64 + mt * 16 + (lane & 15)];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) bf[nt] = Bs[kk * 4 + (lane >> 4)][wn * 64 + nt * 16 + (lane & 15)];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = F64x16::mma(af[mt], bf[nt], acc[mt][nt]);
    }
}

__global__ __launch_bounds__(256) void k_grid_phase_f64(FitGeom g, const int *__restrict__ idx, int64_t row0,
                                                        int64_t rows_valid, const int *__restrict__ khkl,
                                                        const double *__restrict__ kcoef, int K,
                                                        const double2 *__restrict__ ex, const double2 *__restrict__ ey,
                                                        const double2 *__restrict__ ez, double *__restrict__ G, int KP) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    const int64_t r = blockIdx.y;
    if (2 * p >= KP) return;
    double c = 0.0, s = 0.0;
    if (r < rows_valid && p < K) {
        int i, j, k;
        g.ijk(idx[row0 + r], i, j, k);
        const double2 a = ex[khkl[3 * p] * g.nx + i];
        const double2 b = ey[(khkl[3 * p + 1] + 7) * g.ny + j];
        const double2 d = ez[(khkl[3 * p + 2] + 7) * g.nz + k];
        const double re = a.x * b.x - a.y * b.y, im = a.x * b.y + a.y * b.x;
        const double w = kcoef[p];
        c = (re * d.x - im * d.y) * w;
        s = (re * d.y + im * d.x) * w;
    }
    reinterpret_cast<double2 *>(G + r * KP)[p] = make_double2(c, s);
}

__global__ __launch_bounds__(256) void k_atom_table_f64(int K, const double *__restrict__ kvecs, int N,
                                                        const double *__restrict__ atoms, double *__restrict__ B,
                                                        int NP) {
    const int a = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    if (a >= NP) return;
    double c = 0.0, s = 0.0;
    if (p < K && a < N) {
        const double th = (kvecs[3 * p] * atoms[3 * a] + kvecs[3 * p + 1] * atoms[3 * a + 1]) + kvecs[3 * p + 2] * atoms[3 * a + 2];
        sincos(th, &s, &c);
    }
    B[(int64_t)(2 * p) * NP + a] = c;
    B[(int64_t)(2 * p + 1) * NP + a] = s;
}

__global__ __launch_bounds__(256) void k_phi_gemm_f64(const double *__restrict__ G, int KP, const double *__restrict__ B,
                                                      int NP, double *__restrict__ Phi, EpiArgs e) {
    __shared__ double As[2][BK][BM];
    __shared__ double Bs[2][BK][BN];
    const int64_t m0 = (int64_t)blockIdx.y * BM;
    const int n0 = blockIdx.x * BN;
    dvec4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = dvec4{0.0, 0.0, 0.0, 0.0};
    const int nk = KP / BK;
    if (nk > 0) {
        double ra[8], rb[8];
        TileLoader::load_a(G, KP, m0, 0, ra);
        TileLoader::load_b(B, NP, 0, n0, rb);
        TileLoader::store_a(As[0], ra);
        TileLoader::store_b(Bs[0], rb);
        __syncthreads();
        for (int kt = 0; kt < nk; ++kt) {
            const bool more = kt + 1 < nk;
            if (more) {
                TileLoader::load_a(G, KP, m0, (kt + 1) * BK, ra);
                TileLoader::load_b(B, NP, (int64_t)(kt + 1) * BK, n0, rb);
            }
            mma_tile(As[kt & 1], Bs[kt & 1], acc);
            if (more) {
                TileLoader::store_a(As[(kt + 1) & 1], ra);
                TileLoader::store_b(Bs[(kt + 1) & 1], rb);
            }
            __syncthreads();
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t row = m0 + wm * 64 + mt * 16 + F64x16::d_row(lane, v);
            double gx = 0, gy = 0, gz = 0;
            const bool valid_row = row < e.rows_valid;
            if (valid_row) e.geom.xyz(e.idx[e.row0 + row], gx, gy, gz);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const int col = n0 + wn * 64 + nt * 16 + F64x16::d_col(lane);
                double val = 0.0;
                if (valid_row) {
                    if (col < e.N)
                        val = acc[mt][nt][v] + e.real_space(gx, gy, gz, col, e.idx[e.row0 + row]);
                    else if (col == e.N)
                        val = e.v[e.row0 + row];
                    else if (col == e.N + 1)
                        val = 1.0;
                }
                Phi[row * NP + col] = val;
            }
        }
}

// Upper-triangular tiles of Phi^T Phi over rows [split*rps, ...), one 128x128 tile per block.  Phi is a
// chunk of `rows` rows, row-major [rows][NP] (the fp64 GEMM's output) or column-major [NP][rows] (the
// factorized basis kernel's output, ColMajor).
template <bool ColMajor>
__device__ __forceinline__ void syrk_load(const double *Phi, int NP, int64_t rows, int64_t r, int col0, double (&v)[8]) {
    if constexpr (ColMajor)
        TileLoader::load_a(Phi, rows, col0, (int)r, v);
    else
        TileLoader::load_b(Phi, NP, r, col0, v);
}
template <bool ColMajor>
__device__ __forceinline__ void syrk_store(double (*S)[BM], const double (&v)[8]) {
    if constexpr (ColMajor)
        TileLoader::store_a(S, v);
    else
        TileLoader::store_b(S, v);
}

template <bool ColMajor>
__global__ __launch_bounds__(256) void k_syrk_f64(const double *__restrict__ Phi, int NP, int64_t rows, int rps,
                                                  const int2 *__restrict__ tiles, int n_tiles, double *__restrict__ out,
                                                  int direct) {
    __shared__ double As[2][BK][BM];
    __shared__ double Bs[2][BK][BN];
    const int2 tile = tiles[blockIdx.x];
    const int64_t r0 = (int64_t)blockIdx.y * rps;
    const int64_t r1 = min((int64_t)rows, r0 + rps);
    dvec4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = dvec4{0.0, 0.0, 0.0, 0.0};
    const int nk = (int)((r1 - r0) / BK);
    static_assert(BM == BN, "square tiles: one loader serves both operands");
    if (nk > 0) {
        double ra[8], rb[8];
        syrk_load<ColMajor>(Phi, NP, rows, r0, tile.x * BM, ra);
        syrk_load<ColMajor>(Phi, NP, rows, r0, tile.y * BN, rb);
        syrk_store<ColMajor>(As[0], ra);
        syrk_store<ColMajor>(Bs[0], rb);
        __syncthreads();
        for (int kt = 0; kt < nk; ++kt) {
            const bool more = kt + 1 < nk;
            if (more) {
                syrk_load<ColMajor>(Phi, NP, rows, r0 + (int64_t)(kt + 1) * BK, tile.x * BM, ra);
                syrk_load<ColMajor>(Phi, NP, rows, r0 + (int64_t)(kt + 1) * BK, tile.y * BN, rb);
            }
            mma_tile(As[kt & 1], Bs[kt & 1], acc);
            if (more) {
                syrk_store<ColMajor>(As[(kt + 1) & 1], ra);
                syrk_store<ColMajor>(Bs[(kt + 1) & 1], rb);
            }
            __syncthreads();
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = wm * 64 + mt * 16 + F64x16::d_row(lane, v);
                const int c = wn * 64 + nt * 16 + F64x16::d_col(lane);
                if (direct) {
                    double *dst = out + (int64_t)(tile.x * BM + r) * NP + tile.y * BN + c;
                    *dst += acc[mt][nt][v];
                } else {
                    // each (split, tile) slot belongs to one block: partials accumulate across chunks
                    out[((int64_t)blockIdx.y * n_tiles + blockIdx.x) * (BM * BN) + r * BN + c] += acc[mt][nt][v];
                }
            }
}

__global__ __launch_bounds__(256) void k_reduce_splits(const double *__restrict__ part, int n_splits, int n_tiles,
                                                       const int2 *__restrict__ tiles, double *__restrict__ gram, int NP) {
    const int t = blockIdx.x;
    const int2 tile = tiles[t];
    for (int e = threadIdx.x; e < BM * BN; e += 256) {
        double s = 0.0;
        for (int sp = 0; sp < n_splits; ++sp) s += part[((int64_t)sp * n_tiles + t) * (BM * BN) + e];
        gram[(int64_t)(tile.x * BM + e / BN) * NP + tile.y * BN + (e % BN)] += s;
    }
}

// v[r] = sum_a Phi[r][a] q[a] over the N basis columns; one wave per row.
__global__ __launch_bounds__(256) void k_rowdot(const double *__restrict__ Phi, int NP, int N, const double *__restrict__ q,
                                                int64_t rows_valid, double *__restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows_valid) return;
    double s = 0.0;
    for (int a = lane; a < N; a += 64) s += Phi[r * NP + a] * q[a];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) out[r] = s;
}

// v = Phi q over the kept points, from the basis the refined form keeps ([chunk][NP][rows], column-major
// per chunk, chunk c holding rows [c rows, (c + 1) rows)); two rows per thread.
__global__ __launch_bounds__(256) void k_kept_matvec(const double *__restrict__ phi, int64_t rows, int64_t plane, int N,
                                                     int64_t M, const double *__restrict__ q, double *__restrict__ out) {
    const int64_t R = 2 * ((int64_t)blockIdx.x * 256 + threadIdx.x);
    if (R >= M) return;
    const int64_t c = R / rows, r = R - c * rows;   // rows is even: R and R + 1 share the chunk
    const double *col = phi + c * plane + r;
    double a0[4] = {0.0, 0.0, 0.0, 0.0}, a1[4] = {0.0, 0.0, 0.0, 0.0};
    int j = 0;
    for (; j + 4 <= N; j += 4) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const double2 p = *reinterpret_cast<const double2 *>(col + (int64_t)(j + t) * rows);
            a0[t] = fma(p.x, q[j + t], a0[t]);
            a1[t] = fma(p.y, q[j + t], a1[t]);
        }
    }
    for (; j < N; ++j) {
        const double2 p = *reinterpret_cast<const double2 *>(col + (int64_t)j * rows);
        a0[0] = fma(p.x, q[j], a0[0]);
        a1[0] = fma(p.y, q[j], a1[0]);
    }
    out[R] = (a0[0] + a0[1]) + (a0[2] + a0[3]);
    if (R + 1 < M) out[R + 1] = (a1[0] + a1[1]) + (a1[2] + a1[3]);
}

// Zero fill (a kernel rather than a memset node, so the launch sequence can be captured in a graph).
__global__ __launch_bounds__(256) void k_zero_f64(double *__restrict__ p, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

}  // namespace

// Lower triangle from the upper one (every element with i <= j has been accumulated).
__global__ void k_symmetrize(double *gram, int NP, int tile) {
    (void)tile;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)NP * NP) return;
    const int i = (int)(e / NP), j = (int)(e % NP);
    if (i > j) gram[e] = gram[(int64_t)j * NP + i];
}

int accumulate_f64(rp_fit *fit, FitDevice &dv) {
    Ctx *ctx = fit->ctx;
    hipStream_t st = ctx->stream;
    const int K = dv.K;
    const int KP = dv.KP;
    const int NP = dv.NP;
    const int64_t M = fit->n_keep;
    const int64_t rows = dv.rows;     // chunk rows, multiple of BM
    DBuf &d_G = fit->ws[0], &d_B = fit->ws[1], &d_phi = fit->ws[2], &d_part = fit->ws[3];
    if (KP > 0) {
        RP_CHECK(d_G.grow(ctx, sizeof(double) * rows * KP));
        RP_CHECK(d_B.grow(ctx, sizeof(double) * (size_t)KP * NP));
        dim3 gb(ceil_div(NP, 256), KP / 2);
        k_atom_table_f64<<<gb, 256, 0, st>>>(K, dv.kvecs.as<double>(), fit->n_atoms, fit->d_atoms.as<double>(), d_B.as<double>(), NP);
    }
    RP_CHECK(d_phi.grow(ctx, sizeof(double) * rows * NP));
    const int n_tiles = (int)dv.tiles.size();
    const int n_splits = dv.splits;
    const int64_t part_n = (int64_t)n_splits * n_tiles * BM * BN;
    if (n_splits > 1) {
        RP_CHECK(d_part.grow(ctx, sizeof(double) * (size_t)part_n));
        k_zero_f64<<<ceil_div(part_n, 256), 256, 0, st>>>(d_part.as<double>(), part_n);
    }
    const int rps = (int)(rows / n_splits);
    for (int64_t r0 = 0; r0 < M; r0 += rows) {
        const int64_t valid = std::min<int64_t>(rows, M - r0);
        if (KP > 0) {
            dim3 gg(ceil_div(KP / 2, 256), rows);
            KTimer kt(ctx, KT_PHASE);
            k_grid_phase_f64<<<gg, 256, 0, st>>>(dv.geom, fit->d_idx.as<int>(), r0, valid, dv.khkl.as<int>(), dv.kcoef.as<double>(), K,
                                                 dv.ex.as<double2>(), dv.ey.as<double2>(), dv.ez.as<double2>(), d_G.as<double>(), KP);
        }
        EpiArgs e = dv.epi;
        e.row0 = r0;
        e.rows_valid = valid;
        dim3 gp(NP / BN, rows / BM);
        {
            KTimer kt(ctx, KT_KSPACE);
            k_phi_gemm_f64<<<gp, 256, 0, st>>>(d_G.as<double>(), KP, d_B.as<double>(), NP, d_phi.as<double>(), e);
        }
        dim3 gs(n_tiles, n_splits);
        KTimer kt(ctx, KT_GRAM);
        k_syrk_f64<false><<<gs, 256, 0, st>>>(d_phi.as<double>(), NP, rows, rps, dv.d_tiles.as<int2>(), n_tiles,
                                       n_splits > 1 ? d_part.as<double>() : fit->d_gram.as<double>(), n_splits > 1 ? 0 : 1);
    }
    if (n_splits > 1) {
        KTimer kt(ctx, KT_REDUCE);
        k_reduce_splits<<<n_tiles, 256, 0, st>>>(d_part.as<double>(), n_splits, n_tiles, dv.d_tiles.as<int2>(),
                                                 fit->d_gram.as<double>(), NP);
    }
    RP_HIP(ctx, hipGetLastError());
    RP_HIP(ctx, hipStreamSynchronize(st));  // chunk buffers are released on return
    ctx->collect();
    return RP_OK;
}

namespace {

// Rows [valid, rows) of every column of a column-major chunk: the factorized basis kernel writes only the
// chunk's kept points, and the Gram reads whole stages.
__global__ __launch_bounds__(256) void k_zero_tail(double *__restrict__ phi, int64_t rows, int64_t valid) {
    const int64_t r = valid + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r < rows) phi[(int64_t)blockIdx.y * rows + r] = 0.0;
}


}  // namespace

// RP_PREC_FP64_FACT: the basis from the factorized reciprocal sum (k_sep_phi, fp64, column-major chunks)
// and the Gram of [Phi | V | 1] on fp64 MFMA: fp64 throughout, like the plain path, without its
// K-wide trig tables and GEMM.
int accumulate_f64_factorized(rp_fit *fit, FitDevice &dv) {
    Ctx *ctx = fit->ctx;
    hipStream_t st = ctx->stream;
    const int N = fit->n_atoms, NP = dv.NP;
    const int64_t M = fit->n_keep;
    const ChunkPlan cp = plan_chunks(M, N, dv.K, RP_PREC_FP64_FACT);
    const int64_t rows = cp.rows;
    const int n_tiles = (int)dv.tiles.size();
    const int n_splits = cp.splits;
    if (cp.NP != NP || cp.n_tiles != n_tiles || rows % (n_splits * BK) != 0)
        return ctx->fail(RP_EINVAL, "fp64 factorized path: chunk plan does not match the Gram tiling");
    DBuf &d_phi = fit->ws[2], &d_part = fit->ws[3], &d_colmax = fit->ws[5];
    SepPlan &plan = dv.sep_plan;
    {
        KTimer kt(ctx, KT_PHASE);
        RP_CHECK(sep_prepare(fit, dv, rows, plan));
    }
    RP_CHECK(d_phi.grow(ctx, sizeof(double) * (size_t)NP * rows));
    RP_CHECK(d_colmax.grow(ctx, sizeof(unsigned long long) * NP));
    const int64_t part_n = (int64_t)n_splits * n_tiles * BM * BN;
    if (n_splits > 1) {
        RP_CHECK(d_part.grow(ctx, sizeof(double) * (size_t)part_n));
        k_zero_f64<<<ceil_div(part_n, 256), 256, 0, st>>>(d_part.as<double>(), part_n);
    }
    const int rps = (int)(rows / n_splits);
    int chunk = 0;
    for (int64_t r0 = 0; r0 < M; r0 += rows, ++chunk) {
        const int64_t valid = std::min<int64_t>(rows, M - r0);
        {
            KTimer kt(ctx, KT_KSPACE);
            // the fp64 Gram reads the padding columns: they are written as zeros
            sep_phi_chunk(fit, dv, plan, chunk, r0, valid, rows, d_phi.as<double>(),
                          d_colmax.as<unsigned long long>(), st, nullptr, nullptr, true);
        }
        if (valid < rows) {
            dim3 gz((unsigned)ceil_div(rows - valid, 256), (unsigned)NP);
            k_zero_tail<<<gz, 256, 0, st>>>(d_phi.as<double>(), rows, valid);
        }
        dim3 gs(n_tiles, n_splits);
        KTimer kt(ctx, KT_GRAM);
        k_syrk_f64<true><<<gs, 256, 0, st>>>(d_phi.as<double>(), NP, rows, rps, dv.d_tiles.as<int2>(), n_tiles,
                                             n_splits > 1 ? d_part.as<double>() : fit->d_gram.as<double>(),
                                             n_splits > 1 ? 0 : 1);
    }
    if (n_splits > 1) {
        KTimer kt(ctx, KT_REDUCE);
        k_reduce_splits<<<n_tiles, 256, 0, st>>>(d_part.as<double>(), n_splits, n_tiles, dv.d_tiles.as<int2>(),
                                                 fit->d_gram.as<double>(), NP);
    }
    RP_HIP(ctx, hipGetLastError());
    fit->phi_kept = false;
    fit->phi_rows = rows;
    return RP_OK;
}

int run_phi_matvec(rp_fit *fit, const double *q, double *v_coul) {
    Ctx *ctx = fit->ctx;
    hipStream_t st = ctx->stream;
    if (fit->phi_kept) {
        // the refined form keeps the fit's fp64 basis of every kept point: one streaming pass over it
        const int N = fit->n_atoms;
        const int64_t M = fit->n_keep, rows = fit->phi_rows, NP = round_up(N + 2, 128);
        DBuf &d_q = fit->ws[22], &d_out = fit->ws[23];
        RP_CHECK(d_q.grow(ctx, sizeof(double) * N));
        RP_CHECK(d_out.grow(ctx, sizeof(double) * M));
        RP_HIP(ctx, hipMemcpyAsync(d_q.p, q, sizeof(double) * N, hipMemcpyHostToDevice, st));
        k_kept_matvec<<<ceil_div(M, 512), 256, 0, st>>>(fit->ws[7].as<double>(), rows, NP * rows, N, M,
                                                         d_q.as<double>(), d_out.as<double>());
        RP_HIP(ctx, hipGetLastError());
        RP_HIP(ctx, hipMemcpyAsync(v_coul, d_out.p, sizeof(double) * M, hipMemcpyDeviceToHost, st));
        RP_HIP(ctx, hipStreamSynchronize(st));
        return RP_OK;
    }
    const int prec = fit->opt.precision;
    if (fit->accumulated && fit->dev && fit->phi_rows > 0 &&
        (prec == RP_PREC_OZAKI || prec == RP_PREC_OZAKI_REFINED || prec == RP_PREC_FP64_FACT)) {
        // after an accumulate of the factorized forms: their fp64 basis again, chunk by chunk into the
        // chunk buffer, each chunk's rows times q
        FitDevice &dv = *fit->dev;
        const int N = fit->n_atoms, NP = dv.NP;
        const int64_t M = fit->n_keep, rows = fit->phi_rows;
        DBuf &d_phi = fit->ws[2], &d_colmax = fit->ws[5], &d_q = fit->ws[22], &d_out = fit->ws[23];
        RP_CHECK(d_phi.grow(ctx, sizeof(double) * (size_t)NP * rows));
        RP_CHECK(d_colmax.grow(ctx, sizeof(unsigned long long) * NP));
        RP_CHECK(d_q.grow(ctx, sizeof(double) * N));
        RP_CHECK(d_out.grow(ctx, sizeof(double) * M));
        RP_HIP(ctx, hipMemcpyAsync(d_q.p, q, sizeof(double) * N, hipMemcpyHostToDevice, st));
        int chunk = 0;
        for (int64_t r0 = 0; r0 < M; r0 += rows, ++chunk) {
            const int64_t valid = std::min<int64_t>(rows, M - r0);
            sep_phi_chunk(fit, dv, dv.sep_plan, chunk, r0, valid, rows, d_phi.as<double>(),
                          d_colmax.as<unsigned long long>(), st);
            k_kept_matvec<<<ceil_div(valid, 512), 256, 0, st>>>(d_phi.as<double>(), rows, (int64_t)NP * rows, N, valid,
                                                                d_q.as<double>(), d_out.as<double>() + r0);
        }
        RP_HIP(ctx, hipGetLastError());
        RP_HIP(ctx, hipMemcpyAsync(v_coul, d_out.p, sizeof(double) * M, hipMemcpyDeviceToHost, st));
        RP_HIP(ctx, hipStreamSynchronize(st));
        return RP_OK;
    }
    // otherwise the basis in fp64 by the plain path's GEMM (KP padded for the fp64 GEMM)
    FitDevice *dvp = nullptr;
    RP_CHECK(device_setup(fit, RP_PREC_FP64, dvp));
    FitDevice &dv = *dvp;
    const int K = dv.K, NP = dv.NP, N = fit->n_atoms;
    const int KP = K > 0 ? (int)round_up(2 * K, 16) : 0;
    const int64_t M = fit->n_keep, rows = dv.rows;
    DBuf &d_G = fit->ws[0], &d_B = fit->ws[1], &d_phi = fit->ws[2], &d_q = fit->ws[22], &d_out = fit->ws[23];
    if (KP > 0) {
        RP_CHECK(d_G.grow(ctx, sizeof(double) * rows * KP));
        RP_CHECK(d_B.grow(ctx, sizeof(double) * (size_t)KP * NP));
        dim3 gb(ceil_div(NP, 256), KP / 2);
        k_atom_table_f64<<<gb, 256, 0, st>>>(K, dv.kvecs.as<double>(), N, fit->d_atoms.as<double>(), d_B.as<double>(), NP);
    }
    RP_CHECK(d_phi.grow(ctx, sizeof(double) * rows * NP));
    RP_CHECK(d_q.grow(ctx, sizeof(double) * N));
    RP_CHECK(d_out.grow(ctx, sizeof(double) * M));
    RP_HIP(ctx, hipMemcpyAsync(d_q.p, q, sizeof(double) * N, hipMemcpyHostToDevice, st));
    for (int64_t r0 = 0; r0 < M; r0 += rows) {
        const int64_t valid = std::min<int64_t>(rows, M - r0);
        if (KP > 0) {
            dim3 gg(ceil_div(KP / 2, 256), rows);
            k_grid_phase_f64<<<gg, 256, 0, st>>>(dv.geom, fit->d_idx.as<int>(), r0, valid, dv.khkl.as<int>(), dv.kcoef.as<double>(), K,
                                                 dv.ex.as<double2>(), dv.ey.as<double2>(), dv.ez.as<double2>(), d_G.as<double>(), KP);
        }
        EpiArgs e = dv.epi;
        e.row0 = r0;
        e.rows_valid = valid;
        dim3 gp(NP / BN, rows / BM);
        k_phi_gemm_f64<<<gp, 256, 0, st>>>(d_G.as<double>(), KP, d_B.as<double>(), NP, d_phi.as<double>(), e);
        k_rowdot<<<ceil_div(valid, 4), 256, 0, st>>>(d_phi.as<double>(), NP, N, d_q.as<double>(), valid, d_out.as<double>() + r0);
    }
    RP_HIP(ctx, hipGetLastError());
    RP_HIP(ctx, hipMemcpyAsync(v_coul, d_out.p, sizeof(double) * M, hipMemcpyDeviceToHost, st));
    RP_HIP(ctx, hipStreamSynchronize(st));
    return RP_OK;
}

}  // namespace rp
