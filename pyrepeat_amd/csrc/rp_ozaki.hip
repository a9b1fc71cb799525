// Split-precision accumulation: int8 fixed-point slices on MFMA with exact int32 accumulation.
//
// A value x with scale S (a power of two, |x|/S <= 127/128) is held as
//     x ~= S * sum_{p=1..P} c_p 2^{-7p},   c_p integer in [-127, 127]  (int8 "slices"),
// and a product of two such matrices as
//     A B ~= S_A S_B sum_{s} 2^{-7s} I_s,   I_s = sum_{p+q=s} A_p B_q   (int8 x int8 -> int32 MFMA).
// Each I_s is an exact integer sum (bounded well inside int32 for the block lengths used),
// so the only error is the slice representation itself: relative 2^{-7P} of the row or
// column maximum, i.e. a perturbed basis, not a perturbed Gram.  Group partials are
// converted to fp64 and accumulated in fp64 across k-blocks and row chunks.
//
// Two uses (fit_charges steps 1-3, repeat.py:2195-2262):
//   k-space   Phi_rec = G B, G[m][2k], G[m][2k+1] = w_k (cos, sin)(k.g_m), B = (cos, sin)(k.R_a).
//             G is sliced per k-block with the block's largest weight as scale, B with scale 2.
//             The epilogue adds the real-space sum (fp32 erfc with compensated summation; the
//             displacements and the cutoff test stay fp64) and the augmented columns V and 1,
//             writes Phi^T (fp64) and the per-column maxima.
//   Gram      [Phi|V|1]^T [Phi|V|1] from per-column sliced Phi^T, upper tiles, 4 slices.
//
// MFMA: v_mfma_i32_32x32x32_i8, one 32x32 tile per wave, 8 waves per block (2 x 4)
// -> 64x128 block tile, k-step 32, slices staged through LDS, double buffered.
// Seven exact int32 group accumulators per wave tile live in AGPRs.
#include <algorithm>
#include <cmath>

#include "rp_fit_common.h"
#include "rp_mfma.h"

namespace rp {

namespace {

constexpr int PMAX = 4;
constexpr int NG = 2 * PMAX - 1;   // groups s = p+q, p,q in [0, P)
constexpr int TM = 64, TN = 128, TK = 64;
constexpr int TKP = TK + 16;   // padded LDS row (bytes)
constexpr int NTHR = 512;

struct OzBlock {
    int k0, k1, P, PB;   // slices of A and of B used in this block
    double scale;        // S_A * S_B for the block (k-space); 1 for the Gram
};

struct OzOperands {
    const int8_t *A;   // A[p][row][k]
    int64_t a_plane;
    int64_t lda;
    const int8_t *B;   // B[p][col][k]
    int64_t b_plane;
    int64_t ldb;
    int nblocks;
    OzBlock blk[8];
};

__device__ __forceinline__ void slice_value(double x, double inv_scale, int P, int8_t *out, int64_t plane) {
    double r = x * inv_scale;
#pragma unroll
    for (int p = 0; p < PMAX; ++p) {
        if (p < P) {
            const double t = r * 128.0;
            const double c = rint(t);
            out[p * plane] = (int8_t)(int)c;
            r = t - c;
        }
    }
}

// Main loop: accumulates the block tile (rows m0.., cols n0..) of sum_b scale_b sum_s 2^{-7(s+2)} I_s into facc.
// K ranges of the blocks are multiples of TK; rows/cols are in range by construction.
// LDS rows are padded to TKP bytes so the 16-byte fragment reads of 16 consecutive lanes
// fall on distinct banks.  Each stage holds TK = 64 k (two MFMA k-steps); the next stage is
// loaded into registers while the current one is consumed.
__device__ __forceinline__ void oz_mainloop(const OzOperands &op, int64_t m0, int64_t n0, double (&facc)[16],
                                            int8_t *lds) {
    int8_t(*As)[PMAX][TM][TKP] = reinterpret_cast<int8_t(*)[PMAX][TM][TKP]>(lds);
    int8_t(*Bs)[PMAX][TN][TKP] = reinterpret_cast<int8_t(*)[PMAX][TN][TKP]>(lds + 2 * PMAX * TM * TKP);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave >> 2, wn = wave & 3;
#pragma unroll
    for (int r = 0; r < 16; ++r) facc[r] = 0.0;

    for (int b = 0; b < op.nblocks; ++b) {
        const OzBlock blk = op.blk[b];
        const int P = blk.P, PB = blk.PB;
        ivec16 acc[NG];
#pragma unroll
        for (int s = 0; s < NG; ++s)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[s][r] = 0;
        const int nk = (blk.k1 - blk.k0) / TK;
        // staging, 16-byte chunks: A c in [0, 256P): slice c/256, row (c%256)/4, quarter c%4;
        //                          B c in [0, 512PB): slice c/512, row (c%512)/4, quarter c%4
        ivec4 ra[2], rb[4];
        auto gload = [&](int kbase) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int c = t + NTHR * u;
                if (c < 256 * P) {
                    const int p = c >> 8, row = (c & 255) >> 2, qq = c & 3;
                    ra[u] = *reinterpret_cast<const ivec4 *>(op.A + p * op.a_plane + (m0 + row) * op.lda + kbase + 16 * qq);
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int c = t + NTHR * u;
                if (c < 512 * PB) {
                    const int p = c >> 9, row = (c & 511) >> 2, qq = c & 3;
                    rb[u] = *reinterpret_cast<const ivec4 *>(op.B + p * op.b_plane + (n0 + row) * op.ldb + kbase + 16 * qq);
                }
            }
        };
        auto lstore = [&](int buf) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int c = t + NTHR * u;
                if (c < 256 * P) {
                    const int p = c >> 8, row = (c & 255) >> 2, qq = c & 3;
                    *reinterpret_cast<ivec4 *>(&As[buf][p][row][16 * qq]) = ra[u];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int c = t + NTHR * u;
                if (c < 512 * PB) {
                    const int p = c >> 9, row = (c & 511) >> 2, qq = c & 3;
                    *reinterpret_cast<ivec4 *>(&Bs[buf][p][row][16 * qq]) = rb[u];
                }
            }
        };
        if (nk > 0) {
            gload(blk.k0);
            lstore(0);
            __syncthreads();
            for (int kt = 0; kt < nk; ++kt) {
                const int cur = kt & 1;
                const bool more = kt + 1 < nk;
                if (more) gload(blk.k0 + (kt + 1) * TK);
#pragma unroll
                for (int h = 0; h < TK / 32; ++h) {
                    ivec4 af[PMAX], bf[PMAX];
#pragma unroll
                    for (int p = 0; p < PMAX; ++p) {
                        if (p < P) af[p] = *reinterpret_cast<const ivec4 *>(&As[cur][p][wm * 32 + (lane & 31)][32 * h + 16 * (lane >> 5)]);
                        if (p < PB) bf[p] = *reinterpret_cast<const ivec4 *>(&Bs[cur][p][wn * 32 + (lane & 31)][32 * h + 16 * (lane >> 5)]);
                    }
#pragma unroll
                    for (int p = 0; p < PMAX; ++p)
#pragma unroll
                        for (int q = 0; q < PMAX; ++q) {
                            if (p < P && q < PB) acc[p + q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[p], bf[q], acc[p + q], 0, 0, 0);
                        }
                }
                if (more) lstore(cur ^ 1);
                __syncthreads();
            }
        }
        // exact integer groups -> fp64
#pragma unroll
        for (int s = 0; s < NG; ++s) {
            const double w = blk.scale * ldexp(1.0, -7 * (s + 2));
#pragma unroll
            for (int r = 0; r < 16; ++r) facc[r] += (double)acc[s][r] * w;
        }
    }
}

// ---------------------------------------------------------------- slicing kernels

// G slices, planes [P][rows][KP]: column 2k / 2k+1 of row m = w_k (cos, sin)(k . g_m) / S_block(k).
__global__ __launch_bounds__(256) void k_grid_phase_i8(FitGeom g, const int *__restrict__ idx, int64_t row0,
                                                       int64_t rows_valid, const int *__restrict__ khkl,
                                                       const double *__restrict__ kcoef_over_scale, int K,
                                                       const int *__restrict__ kslices, const double2 *__restrict__ ex,
                                                       const double2 *__restrict__ ey, const double2 *__restrict__ ez,
                                                       int8_t *__restrict__ G, int KP, int64_t plane) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    const int64_t r = blockIdx.y;
    if (2 * p >= KP) return;
    double c = 0.0, s = 0.0;
    int P = PMAX;
    if (p < K) P = kslices[p];
    if (r < rows_valid && p < K) {
        int i, j, k;
        g.ijk(idx[row0 + r], i, j, k);
        const double2 a = ex[khkl[3 * p] * g.nx + i];
        const double2 b = ey[(khkl[3 * p + 1] + 7) * g.ny + j];
        const double2 d = ez[(khkl[3 * p + 2] + 7) * g.nz + k];
        const double re = a.x * b.x - a.y * b.y, im = a.x * b.y + a.y * b.x;
        const double w = kcoef_over_scale[p];
        c = (re * d.x - im * d.y) * w;
        s = (re * d.y + im * d.x) * w;
    }
    int8_t *dst = G + r * KP + 2 * p;
    // zero every plane first so unused slices of short blocks stay 0
#pragma unroll
    for (int q = 0; q < PMAX; ++q) {
        dst[q * plane] = 0;
        dst[q * plane + 1] = 0;
    }
    slice_value(c, 1.0, P, dst, plane);
    slice_value(s, 1.0, P, dst + 1, plane);
}

// Atom table slices, planes [PMAX][NP][KP]: (cos, sin)(k . R_a) with scale 2.
__global__ __launch_bounds__(256) void k_atom_table_i8(int K, const double *__restrict__ kvecs, int N,
                                                       const double *__restrict__ atoms, int8_t *__restrict__ B, int KP,
                                                       int64_t plane) {
    const int p = blockIdx.x * 256 + threadIdx.x;   // k pair
    const int a = blockIdx.y;
    if (2 * p >= KP) return;
    double c = 0.0, s = 0.0;
    if (p < K && a < N) {
        const double th = (kvecs[3 * p] * atoms[3 * a] + kvecs[3 * p + 1] * atoms[3 * a + 1]) + kvecs[3 * p + 2] * atoms[3 * a + 2];
        sincos(th, &s, &c);
    }
    int8_t *dst = B + (int64_t)a * KP + 2 * p;
    slice_value(c, 0.5, PMAX, dst, plane);
    slice_value(s, 0.5, PMAX, dst + 1, plane);
}

// Phi^T (fp64, [NP][rows]) -> slices [PMAX][NP][rows] with per-column scale from the column maxima.
__global__ __launch_bounds__(256) void k_slice_phi(const double *__restrict__ phiT, int64_t rows,
                                                   const unsigned long long *__restrict__ colmax, int NP,
                                                   int8_t *__restrict__ out, int64_t plane, double *__restrict__ scale_out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)NP * rows) return;
    const int a = (int)(e / rows);
    const double mx = __longlong_as_double((long long)colmax[a]);
    double S = 1.0;
    if (mx > 0.0) S = ldexp(1.0, (int)ceil(log2(mx * (128.0 / 127.0))));
    // guard the ceil/log2 rounding: the scaled maximum must stay within 127/128
    while (mx / S > 127.0 / 128.0) S *= 2.0;
    if (e % rows == 0) scale_out[a] = S;
    slice_value(phiT[e], 1.0 / S, PMAX, out + e, plane);
}

// ---------------------------------------------------------------- GEMM kernels

struct KspaceEpi {
    EpiArgs e;
    double *phiT;                       // [NP][rows]
    int64_t rows;
    unsigned long long *colmax;         // [NP], bit patterns of max |Phi|
};

// fp32 real-space sum with compensated (Kahan) accumulation; displacements and cutoff test in fp64.
__device__ __forceinline__ double real_space_f32(const EpiArgs &e, double gx, double gy, double gz, int a, int f) {
    const double ax = e.atoms[3 * a], ay = e.atoms[3 * a + 1], az = e.atoms[3 * a + 2];
    if (e.fit_type == 0) {
        const double dx = gx - ax, dy = gy - ay, dz = gz - az;
        const double d = sqrt((dx * dx + dy * dy) + dz * dz);
        return d > 1e-10 ? 1.0 / d : 0.0;
    }
    const float sa = (float)e.sqrt_alpha;
    double sx0, sy0, sz0;
    if (e.nearest_image(f, a, sx0, sy0, sz0)) {
        const double dx = gx - (ax + sx0), dy = gy - (ay + sy0), dz = gz - (az + sz0);
        const double r2 = (dx * dx + dy * dy) + dz * dz;
        if (!(r2 <= e.R2)) return 0.0;
        if (sx0 == 0.0 && sy0 == 0.0 && sz0 == 0.0 && r2 <= 1e-20) return 0.0;
        const float r = sqrtf((float)r2);
        return (double)(erfcf(sa * r) / r);
    }
    float sum = 0.f, comp = 0.f;
    for (int s = 0; s < e.S; ++s) {
        const double sx = e.shifts[3 * s], sy = e.shifts[3 * s + 1], sz = e.shifts[3 * s + 2];
        const double dx = gx - (ax + sx), dy = gy - (ay + sy), dz = gz - (az + sz);
        const double r2 = (dx * dx + dy * dy) + dz * dz;
        if (r2 <= e.R2) {
            if (sx == 0.0 && sy == 0.0 && sz == 0.0 && r2 <= 1e-20) continue;
            const float r = sqrtf((float)r2);
            const float term = erfcf(sa * r) / r;
            const float y = term - comp;
            const float tt = sum + y;
            comp = (tt - sum) - y;
            sum = tt;
        }
    }
    return (double)sum - (double)comp;
}

__global__ __launch_bounds__(NTHR) void k_oz_kspace(OzOperands op, KspaceEpi ke, int n_col_tiles) {
    extern __shared__ int8_t lds[];
    // Blocks are dispatched round-robin over the 8 XCDs.  With the column tile fastest, an XCD
    // keeps seeing the same column tiles (all of them when there are 8), so its L2 holds their
    // atom-side slices while the grid-side slices stream through.
    const int L = blockIdx.x;
    const int col_tile = L % n_col_tiles;
    const int row_tile = L / n_col_tiles;
    const int64_t m0 = (int64_t)row_tile * TM;
    const int64_t n0 = (int64_t)col_tile * TN;
    double facc[16];
    oz_mainloop(op, m0, n0, facc, lds);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 2, wn = wave & 3;
    const EpiArgs &e = ke.e;
    const int col = (int)n0 + wn * 32 + M32::d_col(lane);
    double cmax = 0.0;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int64_t row = m0 + wm * 32 + M32::d_row(lane, r);
        double val = 0.0;
        if (row < e.rows_valid) {
            if (col < e.N) {
                double gx, gy, gz;
                const int f = e.idx[e.row0 + row];
                e.geom.xyz(f, gx, gy, gz);
                val = facc[r] + real_space_f32(e, gx, gy, gz, col, f);
            } else if (col == e.N) {
                val = e.v[e.row0 + row];
            } else if (col == e.N + 1) {
                val = 1.0;
            }
        }
        ke.phiT[(int64_t)col * ke.rows + row] = val;
        cmax = fmax(cmax, fabs(val));
    }
    // lanes l and l^32 hold the same column
    cmax = fmax(cmax, __shfl_xor(cmax, 32, 64));
    if (lane < 32 && cmax > 0.0) atomicMax(&ke.colmax[col], (unsigned long long)__double_as_longlong(cmax));
}

struct GramEpi {
    const double *scale;   // per column of Phi (atom)
    const int2 *tiles;     // (tile row in units of TM, tile col in units of TN)
    int n_tiles;
    double *part;          // [split][tile][TM*TN]
};

__global__ __launch_bounds__(NTHR) void k_oz_gram(OzOperands op, GramEpi ge, int rows_per_split) {
    extern __shared__ int8_t lds[];
    const int2 tile = ge.tiles[blockIdx.x];
    const int64_t m0 = (int64_t)tile.x * TM, n0 = (int64_t)tile.y * TN;
    // this split covers k (= grid rows) [split * rps, (split+1) * rps)
    OzOperands o = op;
    o.nblocks = 1;
    o.blk[0].k0 = blockIdx.y * rows_per_split;
    o.blk[0].k1 = o.blk[0].k0 + rows_per_split;
    o.blk[0].P = PMAX;
    o.blk[0].PB = PMAX;
    o.blk[0].scale = 1.0;
    double facc[16];
    oz_mainloop(o, m0, n0, facc, lds);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 2, wn = wave & 3;
    double *dst = ge.part + ((int64_t)blockIdx.y * ge.n_tiles + blockIdx.x) * (TM * TN);
    const int c = wn * 32 + M32::d_col(lane);
    const double sc = ge.scale[n0 + c];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int rr = wm * 32 + M32::d_row(lane, r);
        dst[rr * TN + c] += facc[r] * (ge.scale[m0 + rr] * sc);   // slot owned by this block across chunks
    }
}

__global__ __launch_bounds__(256) void k_oz_reduce(const double *__restrict__ part, int n_splits, int n_tiles,
                                                   const int2 *__restrict__ tiles, double *__restrict__ gram, int NP) {
    const int2 tile = tiles[blockIdx.x];
    for (int e = threadIdx.x; e < TM * TN; e += 256) {
        double s = 0.0;
        for (int sp = 0; sp < n_splits; ++sp) s += part[((int64_t)sp * n_tiles + blockIdx.x) * (TM * TN) + e];
        gram[(int64_t)(tile.x * TM + e / TN) * NP + tile.y * TN + (e % TN)] += s;
    }
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) p[t] = 0ull;
}

}  // namespace

// k-space plan.  The k-vectors are ordered by decreasing weight and cut into blocks of
// 16 (32 G columns, one MFMA k-step).  Every G entry must be represented to the
// absolute precision of the uniform 4-slice form, S_max 2^-28 (S_max = scale of the
// largest weight): a block whose weights are below S_max by 2^-e needs only
// P = ceil((28 - e) / 7) slices of G.  B (the atom phases, shared by every grid point)
// always keeps all 4 slices: truncating it per block perturbs each atom's basis function
// systematically, which the ill-conditioned normal equations amplify (measured 6.7e-6 e
// against 1.3e-6 e with 4 B slices on the 1000-atom cell).  Consecutive blocks with the
// same P form one range with its own power-of-two scale; group sums stay exact in int32.
static double pow2_scale(double vmax) {
    if (!(vmax > 0.0)) return 1.0;
    double S = std::ldexp(1.0, (int)std::ceil(std::log2(vmax * (128.0 / 127.0))));
    while (vmax / S > 127.0 / 128.0) S *= 2.0;
    return S;
}

struct KPlan {
    std::vector<int> order;        // sorted position -> original k index
    std::vector<OzBlock> ranges;
    std::vector<int> kslices;      // per sorted position
    std::vector<double> wscaled;   // w_k / S_range, per sorted position
    std::vector<int32_t> khkl;     // sorted (h, k, l)
    std::vector<double> kvecs;     // sorted k vectors
    int64_t products = 0;          // sum over ranges of (columns / 32) * P^2
};

static void plan_kspace(const Setup &s, int KP, KPlan &pl) {
    const int K = (int)s.kcoefs.size();
    pl.order.resize(K);
    for (int k = 0; k < K; ++k) pl.order[k] = k;
    std::stable_sort(pl.order.begin(), pl.order.end(),
                     [&](int a, int b) { return std::fabs(s.kcoefs[a]) > std::fabs(s.kcoefs[b]); });
    const double smax = pow2_scale(K ? std::fabs(s.kcoefs[pl.order[0]]) : 0.0);
    const int kpb = TK / 2;  // k-vectors per block (two G columns each)
    const int nblk = KP / TK;
    std::vector<int> pblk(nblk, 1);
    for (int b = 0; b < nblk; ++b) {
        const int k0 = b * kpb;
        if (k0 >= K) {
            pblk[b] = 1;
            continue;
        }
        const double sb = pow2_scale(std::fabs(s.kcoefs[pl.order[k0]]));
        const double e = std::log2(smax / sb);
        pblk[b] = std::min(PMAX, std::max(1, (int)std::ceil((28.0 - e) / 7.0)));
        if (b > 0) pblk[b] = std::min(pblk[b], pblk[b - 1]);
    }
    pl.ranges.clear();
    pl.kslices.assign(K, 1);
    pl.wscaled.assign(K, 0.0);
    pl.khkl.resize(3 * (size_t)K);
    pl.kvecs.resize(3 * (size_t)K);
    pl.products = 0;
    for (int b = 0; b < nblk;) {
        int e = b;
        while (e < nblk && pblk[e] == pblk[b]) ++e;
        const int k0 = b * kpb, k1 = std::min(K, e * kpb);
        const double S = k0 < K ? pow2_scale(std::fabs(s.kcoefs[pl.order[k0]])) : 1.0;
        OzBlock r{b * TK, e * TK, pblk[b], PMAX, S * 2.0};
        pl.ranges.push_back(r);
        pl.products += (int64_t)(e - b) * pblk[b] * PMAX;
        for (int p = k0; p < k1; ++p) {
            pl.kslices[p] = pblk[b];
            pl.wscaled[p] = s.kcoefs[pl.order[p]] / S;
        }
        b = e;
    }
    for (int p = 0; p < K; ++p)
        for (int c = 0; c < 3; ++c) {
            pl.khkl[3 * p + c] = s.khkl[3 * pl.order[p] + c];
            pl.kvecs[3 * p + c] = s.kvecs[3 * pl.order[p] + c];
        }
}

int accumulate_split(rp_fit *fit, FitDevice &dv) {
    Ctx *ctx = fit->ctx;
    hipStream_t st = ctx->stream;
    const int K = dv.K, NP = dv.NP, N = fit->n_atoms;
    const int64_t M = fit->n_keep;
    const ChunkPlan cp = plan_chunks(M, N, K, RP_PREC_OZAKI);
    const int KP = cp.KP;
    const int64_t rows = cp.rows;
    if (cp.NP != NP || KP % TK != 0 || PMAX != 4 || TM != 64 || TN != 128)
        return ctx->fail(RP_EINVAL, "split path: chunk plan does not match the kernel tiling");
    // operands
    DBuf &d_G = fit->ws[0], &d_B = fit->ws[1], &d_phiT = fit->ws[2], &d_part = fit->ws[3], &d_slices = fit->ws[4],
         &d_colmax = fit->ws[5], &d_scale = fit->ws[6], &d_wsc = fit->ws[7], &d_ksl = fit->ws[8], &d_tiles = fit->ws[9],
         &d_khkl = fit->ws[10], &d_kvec = fit->ws[11];
    KPlan pl;
    if (K > 0) {
        plan_kspace(fit->setup, KP, pl);
        if ((int)pl.ranges.size() > 8) return ctx->fail(RP_EINVAL, "k-space plan has more than 8 ranges");
        RP_CHECK(d_G.grow(ctx, (size_t)PMAX * rows * KP));
        RP_CHECK(d_B.grow(ctx, (size_t)PMAX * NP * KP));
        RP_CHECK(d_wsc.grow(ctx, sizeof(double) * K));
        RP_CHECK(d_ksl.grow(ctx, sizeof(int) * K));
        RP_CHECK(d_khkl.grow(ctx, sizeof(int) * 3 * K));
        RP_CHECK(d_kvec.grow(ctx, sizeof(double) * 3 * K));
        RP_HIP(ctx, hipMemcpyAsync(d_wsc.p, pl.wscaled.data(), sizeof(double) * K, hipMemcpyHostToDevice, st));
        RP_HIP(ctx, hipMemcpyAsync(d_ksl.p, pl.kslices.data(), sizeof(int) * K, hipMemcpyHostToDevice, st));
        RP_HIP(ctx, hipMemcpyAsync(d_khkl.p, pl.khkl.data(), sizeof(int) * 3 * K, hipMemcpyHostToDevice, st));
        RP_HIP(ctx, hipMemcpyAsync(d_kvec.p, pl.kvecs.data(), sizeof(double) * 3 * K, hipMemcpyHostToDevice, st));
        dim3 gb(ceil_div(KP / 2, 256), NP);
        k_atom_table_i8<<<gb, 256, 0, st>>>(K, d_kvec.as<double>(), N, fit->d_atoms.as<double>(), d_B.as<int8_t>(), KP,
                                            (int64_t)NP * KP);
    }
    RP_CHECK(d_phiT.grow(ctx, sizeof(double) * NP * rows));
    RP_CHECK(d_slices.grow(ctx, (size_t)PMAX * NP * rows));
    RP_CHECK(d_colmax.grow(ctx, sizeof(unsigned long long) * NP));
    RP_CHECK(d_scale.grow(ctx, sizeof(double) * NP));
    // Gram tiles (TM x TN) touching the upper triangle, and row splits
    std::vector<int2> tiles;
    for (int bi = 0; bi < NP / TM; ++bi)
        for (int bj = 0; bj < NP / TN; ++bj)
            if (bi * TM <= bj * TN + TN - 1) tiles.push_back(make_int2(bi, bj));
    const int n_tiles = (int)tiles.size();
    const int splits = cp.splits;
    if (n_tiles != cp.n_tiles || rows % (splits * TK) != 0)
        return ctx->fail(RP_EINVAL, "split path: Gram tiling does not match the chunk plan");
    const int rps = (int)(rows / splits);
    RP_CHECK(d_tiles.grow(ctx, sizeof(int2) * n_tiles));
    RP_HIP(ctx, hipMemcpyAsync(d_tiles.p, tiles.data(), sizeof(int2) * n_tiles, hipMemcpyHostToDevice, st));
    RP_CHECK(d_part.grow(ctx, sizeof(double) * (size_t)splits * n_tiles * TM * TN));
    RP_HIP(ctx, hipMemsetAsync(d_part.p, 0, sizeof(double) * (size_t)splits * n_tiles * TM * TN, st));

    const size_t lds_bytes = 2 * PMAX * (TM + TN) * TKP;  // 120 KiB
    RP_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(&k_oz_kspace),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
    RP_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(&k_oz_gram),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
    OzOperands ko{};
    ko.A = d_G.as<int8_t>();
    ko.a_plane = rows * KP;
    ko.lda = KP;
    ko.B = d_B.as<int8_t>();
    ko.b_plane = (int64_t)NP * KP;
    ko.ldb = KP;
    ko.nblocks = (int)pl.ranges.size();
    for (int b = 0; b < ko.nblocks; ++b) ko.blk[b] = pl.ranges[b];
    fit->kspace_products = pl.products;
    OzOperands go{};
    go.A = d_slices.as<int8_t>();
    go.a_plane = (int64_t)NP * rows;
    go.lda = rows;
    go.B = go.A;
    go.b_plane = go.a_plane;
    go.ldb = rows;
    GramEpi ge{d_scale.as<double>(), d_tiles.as<int2>(), n_tiles, d_part.as<double>()};

    for (int64_t r0 = 0; r0 < M; r0 += rows) {
        const int64_t valid = std::min<int64_t>(rows, M - r0);
        if (K > 0) {
            dim3 gg(ceil_div(KP / 2, 256), rows);
            KTimer kt(ctx, KT_PHASE);
            k_grid_phase_i8<<<gg, 256, 0, st>>>(dv.geom, fit->d_idx.as<int>(), r0, valid, d_khkl.as<int>(), d_wsc.as<double>(),
                                                K, d_ksl.as<int>(), dv.ex.as<double2>(), dv.ey.as<double2>(),
                                                dv.ez.as<double2>(), d_G.as<int8_t>(), KP, rows * KP);
        }
        k_zero_u64<<<ceil_div(NP, 256), 256, 0, st>>>(d_colmax.as<unsigned long long>(), NP);
        KspaceEpi ke{dv.epi, d_phiT.as<double>(), rows, d_colmax.as<unsigned long long>()};
        ke.e.row0 = r0;
        ke.e.rows_valid = valid;
        OzOperands kop = ko;
        if (K == 0) kop.nblocks = 0;
        // rows is a multiple of 128 = 2 * TM, so the row-tile count is even; pad to a multiple of 8
        const int row_tiles = (int)(rows / TM), col_tiles = NP / TN;
        if (row_tiles % 8 != 0) return ctx->fail(RP_EINVAL, "chunk rows must be a multiple of 512");
        {
            KTimer kt(ctx, KT_KSPACE);
            k_oz_kspace<<<row_tiles * col_tiles, NTHR, lds_bytes, st>>>(kop, ke, col_tiles);
        }
        {
            KTimer kts(ctx, KT_SLICE);
            k_slice_phi<<<ceil_div((int64_t)NP * rows, 256), 256, 0, st>>>(d_phiT.as<double>(), rows,
                                                                             d_colmax.as<unsigned long long>(), NP,
                                                                             d_slices.as<int8_t>(), (int64_t)NP * rows,
                                                                             d_scale.as<double>());
        }
        dim3 gs(n_tiles, splits);
        KTimer ktg(ctx, KT_GRAM);
        k_oz_gram<<<gs, NTHR, lds_bytes, st>>>(go, ge, rps);
    }
    {
        KTimer ktr(ctx, KT_GRAM);
        k_oz_reduce<<<n_tiles, 256, 0, st>>>(d_part.as<double>(), splits, n_tiles, d_tiles.as<int2>(),
                                              fit->d_gram.as<double>(), NP);
    }
    RP_HIP(ctx, hipGetLastError());
    RP_HIP(ctx, hipStreamSynchronize(st));
    ctx->collect();
    return RP_OK;
}

}  // namespace rp
