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
// Use (fit_charges steps 1-3, repeat.py:2195-2262): the Gram [Phi|V|1]^T [Phi|V|1] of each chunk of
// kept points, from Phi sliced per column (scale from the column maximum over the chunk), upper
// tiles, 4 slices of each operand (16 products, 7 exact groups) in the split form, 3 (9 products,
// 5 groups) in the refined form, whose solve is then refined against the fp64 Phi kept in HBM
// (rp_solve.hip).  Phi itself comes from the factorized reciprocal sum and the fp64 real-space sum
// (rp_kspace.hip).
//
// MFMA: v_mfma_i32_32x32x32_i8, one 32x32 tile per wave, 8 waves per block (2 x 4)
// -> 64x128 block tile, k-step 32, slices loaded straight into LDS, three stage buffers.
// 2P - 1 exact int32 group accumulators per wave tile.
#include <algorithm>
#include <cmath>

#include "rp_fit_common.h"
#include "rp_mfma.h"

namespace rp {

namespace {

constexpr int TM = 64, TN = 128;
constexpr int NTHR = 512;

// Slices [P][NP][rows] of the chunk's Phi (column-major per slice: rows contiguous).
struct GramOperands {
    const int8_t *S;
    int64_t plane;   // NP * rows
    int64_t ld;      // rows
};

// Main loop: the exact int32 group sums I_s = sum_{p+q=s} A_p^T B_q (s = 0 .. 2P-2) of the block
// tile (columns m0.. of A, n0.. of B) over k (= grid rows) [k0, k1), a multiple of the stage depth KD.
// Eight waves as 2 x 4, one 32 x 32 tile each.  The operands come by direct global-to-LDS loads
// (global_load_lds_dwordx4): no register staging, NS stage buffers, loads issued NS - 1 stages ahead.
// A stage holds, per slice, TM (A) and TN (B) rows of KD bytes, unpadded, its 16-byte units swizzled
// by row (gram_swz) so the fragment reads of 16 consecutive rows fall on distinct banks.  One load
// instruction fills 1 KB (1024 / KD rows): lane l writes base + 16 l, i.e. row l / (KD / 16), stored
// unit l % (KD / 16), and so fetches logical unit (l % (KD / 16)) ^ swz(row).
typedef __attribute__((address_space(3))) void lds_void;

// Stage depth (bytes of k per row) and stage buffers: 3 slices take 128-deep stages in two buffers
// (72 KB each: half the barriers), 4 slices 64-deep stages in three (48 KB each).
template <int P>
constexpr int gram_kd() { return gram_stage_rows(P); }
template <int P>
constexpr int gram_ns() { return P == 3 ? 2 : 3; }
template <int P>
constexpr size_t gram_stage_bytes() { return (size_t)P * (TM + TN) * gram_kd<P>(); }
template <int P>
constexpr size_t gram_lds_bytes_dma() { return gram_ns<P>() * gram_stage_bytes<P>(); }   // 144 KiB
// 16-byte unit q of row r of a stage is stored at unit q ^ swz(r), so the fragment reads of 16
// consecutive rows fall on distinct banks: (r >> 2) & 3 for 64-byte rows, (r >> 1) & 7 for 128-byte rows.
template <int KD>
__device__ __forceinline__ int gram_swz(int r) { return KD == 64 ? ((r >> 2) & 3) : ((r >> 1) & 7); }

// s_waitcnt with only the vector-memory count (gfx9 encoding: vmcnt bits [3:0] and [15:14],
// expcnt [6:4] and lgkmcnt [11:8] left at their maxima).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// Work-group barrier for LDS traffic only: this wave's LDS reads and writes are complete, then
// s_barrier.  __syncthreads() would also wait for every vector-memory operation in flight, i.e. for
// the direct loads of the stages ahead.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

template <int P>
__device__ __forceinline__ void gram_mainloop_dma(const GramOperands &op, int64_t m0, int64_t n0, int k0, int k1,
                                                  ivec16 (&acc)[2 * P - 1], int8_t *lds) {
    constexpr int KD = gram_kd<P>(), NS = gram_ns<P>();
    constexpr int RPI = 1024 / KD;   // rows per load instruction (1 KB)
    constexpr int LPR = KD / 16;     // lanes per row
    constexpr int WAVES = NTHR / 64;
    constexpr int CHUNKS = P * (TM + TN) / RPI;
    // load instructions per wave and stage: PER_WAVE_HI for waves below CHUNKS % WAVES, else PER_WAVE_LO
    constexpr int PER_WAVE_LO = CHUNKS / WAVES;
    constexpr int PER_WAVE_HI = (CHUNKS + WAVES - 1) / WAVES;
    constexpr size_t STAGE = gram_stage_bytes<P>();
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave >> 2, wn = wave & 3;
#pragma unroll
    for (int g = 0; g < 2 * P - 1; ++g)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[g][r] = 0;
    const int nk = (k1 - k0) / KD;
    if (nk <= 0) return;
    const int lrow = lane / LPR, qst = lane % LPR;
    auto issue = [&](int kt) {
        int8_t *buf = lds + (size_t)(kt % NS) * STAGE;
        const int kbase = k0 + kt * KD;
#pragma unroll
        for (int u = 0; u < PER_WAVE_HI; ++u) {
            const int ii = wave + WAVES * u;   // chunk of this load instruction (wave-uniform)
            if (ii >= CHUNKS) break;
            int p, row;
            int64_t col0;
            int8_t *dst;
            if (ii < P * TM / RPI) {
                p = ii / (TM / RPI);
                const int rc = ii % (TM / RPI);
                row = RPI * rc + lrow;
                col0 = m0;
                dst = buf + (size_t)(p * TM + RPI * rc) * KD;
            } else {
                const int jj = ii - P * TM / RPI;
                p = jj / (TN / RPI);
                const int rc = jj % (TN / RPI);
                row = RPI * rc + lrow;
                col0 = n0;
                dst = buf + (size_t)P * TM * KD + (size_t)(p * TN + RPI * rc) * KD;
            }
            const int qlog = qst ^ gram_swz<KD>(row);
            const int8_t *src = op.S + p * op.plane + (col0 + row) * op.ld + kbase + 16 * qlog;
            __builtin_amdgcn_global_load_lds(src, (lds_void *)dst, 16, 0, 0);
        }
    };
    for (int q = 0; q < NS - 1 && q < nk; ++q) issue(q);
    const int ra = wm * 32 + (lane & 31), rb = wn * 32 + (lane & 31);
    for (int kt = 0; kt < nk; ++kt) {
        // this wave's loads of stage kt are done when at most those of the stages issued after it (up
        // to NS - 2) are outstanding; the barrier makes every wave's visible and frees the buffer the
        // next issue overwrites (stage kt - 1's)
        if (NS > 2 && kt + 1 < nk) {
            if (wave < CHUNKS % WAVES) wait_vmcnt<PER_WAVE_HI>();
            else wait_vmcnt<PER_WAVE_LO>();
        } else {
            wait_vmcnt<0>();
        }
        lds_barrier();
        if (kt + NS - 1 < nk) issue(kt + NS - 1);
        const int8_t *As = lds + (size_t)(kt % NS) * STAGE;
        const int8_t *Bs = As + (size_t)P * TM * KD;
        // fragments of both 32-deep halves up front: the second half's reads overlap the first
        // half's MFMAs
        ivec4 af[KD / 32][P], bf[KD / 32][P];
#pragma unroll
        for (int h = 0; h < KD / 32; ++h) {
            const int qa = (2 * h + (lane >> 5)) ^ gram_swz<KD>(ra);
            const int qb = (2 * h + (lane >> 5)) ^ gram_swz<KD>(rb);
#pragma unroll
            for (int p = 0; p < P; ++p) {
                af[h][p] = *reinterpret_cast<const ivec4 *>(As + (size_t)(p * TM + ra) * KD + 16 * qa);
                bf[h][p] = *reinterpret_cast<const ivec4 *>(Bs + (size_t)(p * TN + rb) * KD + 16 * qb);
            }
        }
#pragma unroll
        for (int h = 0; h < KD / 32; ++h)
#pragma unroll
            for (int p = 0; p < P; ++p)
#pragma unroll
                for (int q = 0; q < P; ++q)
                    acc[p + q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[h][p], bf[h][q], acc[p + q], 0, 0, 0);
        // scheduling hint: interleave the fragment reads with the MFMAs (measured 4 % on this loop)
        __builtin_amdgcn_iglp_opt(1);
    }
}

// ---------------------------------------------------------------- Gram kernels

struct GramEpi {
    const double *scale;   // per column of Phi (atom)
    const int2 *tiles;     // (tile row in units of TM, tile col in units of TN)
    int n_tiles;
    double *part;          // [split][tile][TM*TN], then the second halves of halved items [8][halve][TM*TN]
    int accumulate;        // 0: the chunk's partial overwrites the slot (first chunk), 1: adds to it
    int halve;             // items per XCD run as two half-k items (the last ones of the XCD's range)
};

// All P^2 products of the P-slice operands, 2P - 1 exact groups s = p + q.
template <int P>
__global__ __launch_bounds__(NTHR) void k_oz_gram(GramOperands op, GramEpi ge, int rows_per_split, int n_splits) {
    extern __shared__ int8_t lds[];
    // Work-groups are dispatched round-robin over the 8 XCDs (linear id L -> XCD L % 8).  When the
    // row splits divide by 8, XCD x takes splits x, x + 8, ..., and runs all tiles of one split
    // before the next, so the tiles in flight on an XCD share the k-window its L2 holds.
    // Otherwise (few row splits: a large Gram) XCD x takes a contiguous range of the (tile, split) items,
    // tile-fastest, and the tile list is ordered in blocks of 8 x 4 tiles (512 x 512 entries), so the tiles in
    // flight on an XCD read few distinct columns; the grid is then 8 ranges, the last one short.
    // Balanced last round: when the XCD's items leave a partial last round of at most half its CUs, its
    // last `halve` items run as two half-k items each (the second half into a slot of its own), so that
    // round is whole.
    const int L = blockIdx.x;
    int tix, split, half = -1, item = 0;
    if (n_splits % 8 == 0) {
        const int n_items = ge.n_tiles * (n_splits / 8);   // per XCD
        const int j = L >> 3;
        item = j;
        if (ge.halve > 0 && j >= n_items - ge.halve) {
            const int h = j - (n_items - ge.halve);
            item = n_items - ge.halve + (h >> 1);
            half = h & 1;
        }
        tix = item % ge.n_tiles;
        split = (item / ge.n_tiles) * 8 + (L & 7);
    } else {
        const int items = ge.n_tiles * n_splits, ipx = (items + 7) / 8;
        const int w = (L & 7) * ipx + (L >> 3);
        if (w >= items) return;
        tix = w % ge.n_tiles;
        split = w / ge.n_tiles;
    }
    const int2 tile = ge.tiles[tix];
    const int64_t m0 = (int64_t)tile.x * TM, n0 = (int64_t)tile.y * TN;
    // this split covers k (= grid rows) [split * rps, (split+1) * rps); a halved item's first half the
    // whole stages below the middle, its second half the rest
    int k0 = split * rows_per_split, k1 = k0 + rows_per_split;
    if (half >= 0) {
        const int mid = k0 + (rows_per_split / 2) / gram_kd<P>() * gram_kd<P>();
        if (half == 0) k1 = mid; else k0 = mid;
    }
    ivec16 acc[2 * P - 1];
    gram_mainloop_dma<P>(op, m0, n0, k0, k1, acc, lds);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 2, wn = wave & 3;
    double *dst = half == 1 ? ge.part + ((int64_t)n_splits * ge.n_tiles + (L & 7) * ge.halve +
                                         (item - (ge.n_tiles * (n_splits / 8) - ge.halve))) * (TM * TN)
                            : ge.part + ((int64_t)split * ge.n_tiles + tix) * (TM * TN);
    const int c = wn * 32 + M32::d_col(lane);
    const double sc = ge.scale[n0 + c];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        // exact integer groups -> fp64, weights 2^{-7(s+2)}
        double v = 0.0;
#pragma unroll
        for (int g = 0; g < 2 * P - 1; ++g) v += (double)acc[g][r] * ldexp(1.0, -7 * (g + 2));
        const int rr = wm * 32 + M32::d_row(lane, r);
        // slot owned by this block across chunks; the first chunk overwrites it
        const double add = v * (ge.scale[m0 + rr] * sc);
        dst[rr * TN + c] = ge.accumulate ? dst[rr * TN + c] + add : add;
    }
}

// Sum of the row-split partials into the Gram: block (tile, part of REDUCE_PARTS), one entry per
// thread, the splits summed in order.
// The halved items are the last `halve` tiles (halve <= n_tiles) of the last 8 row splits, one per XCD;
// their second halves (slots after the split partials, [XCD][halve]) are summed after the splits.
// One entry per thread; the splits' loads are issued 8 at a time ahead of their (in-order) additions.
constexpr int REDUCE_PARTS = TM * TN / 256;
__global__ __launch_bounds__(256) void k_oz_reduce(const double *__restrict__ part, int n_splits, int n_tiles,
                                                   const int2 *__restrict__ tiles, double *__restrict__ gram, int NP,
                                                   int halve) {
    const int tix = blockIdx.x / REDUCE_PARTS;
    const int2 tile = tiles[tix];
    const int e = (blockIdx.x % REDUCE_PARTS) * 256 + threadIdx.x;
    const int hx = tix - (n_tiles - halve);   // this tile's index among the halved ones, if it is one
    const double *src = part + (int64_t)tix * (TM * TN) + e;
    const int64_t stride = (int64_t)n_tiles * (TM * TN);
    double s = 0.0;
    int sp = 0;
    for (; sp + 8 <= n_splits; sp += 8) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = src[(sp + q) * stride];
#pragma unroll
        for (int q = 0; q < 8; ++q) s += v[q];
    }
    for (; sp < n_splits; ++sp) s += src[sp * stride];
    if (halve > 0 && hx >= 0) {
        double v[8];
#pragma unroll
        for (int x = 0; x < 8; ++x) v[x] = part[((int64_t)n_splits * n_tiles + x * halve + hx) * (TM * TN) + e];
#pragma unroll
        for (int x = 0; x < 8; ++x) s += v[x];
    }
    gram[(int64_t)(tile.x * TM + e / TN) * NP + tile.y * TN + (e % TN)] += s;
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) p[t] = 0ull;
}

}  // namespace

namespace {

// Whether the refined form can keep its basis (`need` bytes in ws[7]), decided after every other buffer of
// the accumulate is allocated.  The device's free memory must hold it with 1 GiB to spare.  Under a budget
// with a grow hook (a serving slot, whose reservation already covers everything but the kept basis) the
// hook must grant its size; under a plain budget it must fit beside what the solve still allocates
// (solve_reserve).  An existing ws[7] is freed before a larger one is taken.
bool keep_basis_fits(rp_fit *fit, uint64_t need, uint64_t solve_reserve) {
    Ctx *ctx = fit->ctx;
    if (fit->opt.flags & RP_OPT_STREAM_REFINE) return false;
    DBuf &kept = fit->ws[7];
    if (kept.p && kept.bytes >= need && kept.ctx == ctx) return true;
    const uint64_t held = kept.p ? kept.bytes : 0;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && need + (1ull << 30) > (uint64_t)free_b + held) return false;
    if (ctx->budget) {
        if (ctx->grow_budget) return ctx->grow_budget(need - held);
        if (ctx->in_use - held + need + solve_reserve > ctx->budget) return false;
    }
    return true;
}

// Device bytes the solve allocates at most after an accumulate with N atoms and M kept points (N' <= N):
// the bordered system and its copy, the diagonal-block inverses, the least-squares eigenvectors, the
// refinement vectors and the residual over the kept points, with 1 MiB for the small arrays.
uint64_t solve_reserve_bytes(int N, int64_t M) {
    const uint64_t n = (uint64_t)N + 1, ld = (uint64_t)round_up((int64_t)n + 1, 4);
    return 8 * (2 * n * ld + n * n + 64 * 64 * (uint64_t)ceil_div(N, 64) + 4 * 256 * (uint64_t)N) +
           8 * ((uint64_t)M + 2 + (uint64_t)ceil_div(M, 512)) + (1u << 20);
}

// P slices per element; keep_phi: every chunk of the fp64 basis stays in ws[7] ([chunk][NP][rows]) for
// the refinement of the solve, otherwise one chunk buffer (ws[2]) is reused.  The refined form asks for
// the kept basis and streams its refinement instead when the basis does not fit.
template <int P>
int accumulate_split_p(rp_fit *fit, FitDevice &dv, bool keep_phi) {
    Ctx *ctx = fit->ctx;
    hipStream_t st = ctx->stream;
    const int K = dv.K, NP = dv.NP, N = fit->n_atoms;
    const int64_t M = fit->n_keep;
    const ChunkPlan cp = plan_chunks(M, N, K, fit->opt.precision);
    const int64_t rows = cp.rows;
    if (cp.NP != NP || cp.slices != P || TM != 64 || TN != 128 || rows % 512 != 0)
        return ctx->fail(RP_EINVAL, "split path: chunk plan does not match the kernel tiling");
    const int64_t n_chunks = ceil_div(M, rows);
    DBuf &d_part = fit->ws[3], &d_slices = fit->ws[4], &d_colmax = fit->ws[5], &d_scale = fit->ws[6],
         &d_tiles = fit->ws[9];
    SepPlan &plan = dv.sep_plan;
    {
        KTimer kt(ctx, KT_PHASE);
        RP_CHECK(sep_prepare(fit, dv, rows, plan));
    }
    RP_CHECK(d_slices.grow(ctx, (size_t)P * NP * rows));
    RP_CHECK(d_colmax.grow(ctx, sizeof(unsigned long long) * NP));
    RP_CHECK(d_scale.grow(ctx, sizeof(double) * NP));
    // Gram tiles (TM x TN) touching the upper triangle, and row splits
    // upper tiles in blocks of 8 x 4 tiles (512 x 512 entries), row-major inside a block and across blocks
    std::vector<int2> tiles;
    for (int BI = 0; BI < NP / TM; BI += 8)
        for (int BJ = 0; BJ < NP / TN; BJ += 4)
            for (int bi = BI; bi < std::min(BI + 8, NP / TM); ++bi)
                for (int bj = BJ; bj < std::min(BJ + 4, NP / TN); ++bj)
                    if (bi * TM <= bj * TN + TN - 1) tiles.push_back(make_int2(bi, bj));
    const int n_tiles = (int)tiles.size();
    const int splits = cp.splits;
    if (n_tiles != cp.n_tiles || rows % (splits * gram_kd<P>()) != 0)
        return ctx->fail(RP_EINVAL, "split path: Gram tiling does not match the chunk plan");
    const int rps = (int)(rows / splits);
    // exactness of the int32 groups: a slice digit is at most 127 (first slice) or 64 (the rest), so
    // a group's per-row sum is at most 2 * 127 * 64 + (P - 2) * 64 * 64; rps rows of it must fit
    const int64_t group_row_max = 2LL * 127 * 64 + (int64_t)(P - 2) * 64 * 64;
    if (group_row_max * rps > 2147483647LL)
        return ctx->fail(RP_EINVAL, "split path: row split too long for exact int32 slice groups");
    if (fit->split_tiles_np != NP || !d_tiles.p) {
        fit->split_tiles_np = 0;
        RP_CHECK(d_tiles.grow(ctx, sizeof(int2) * n_tiles));
        RP_HIP(ctx, hipMemcpyAsync(d_tiles.p, tiles.data(), sizeof(int2) * n_tiles, hipMemcpyHostToDevice, st));
        RP_HIP(ctx, hipStreamSynchronize(st));   // the host vector goes out of scope
        fit->split_tiles_np = NP;
    }
    // items per XCD and its CUs: a partial last round of at most half the CUs is made whole by halving
    // its items (their second halves in 8 * halve more partial slots)
    int halve = 0;
    if (splits % 8 == 0) {
        const int cu_xcd = std::max(1, ctx->prop.multiProcessorCount / 8);
        const int n_items = n_tiles * (splits / 8);
        const int rem = n_items % cu_xcd;
        if (rem > 0 && 2 * rem <= cu_xcd && rem <= 64 && rem <= n_tiles && rps / 2 >= gram_kd<P>()) halve = rem;
    }
    RP_CHECK(d_part.grow(ctx, sizeof(double) * ((size_t)splits * n_tiles + 8 * (size_t)halve) * TM * TN));
    // the basis: every chunk kept (refined, when it fits) or one chunk buffer
    const bool refined = keep_phi;
    if (refined) {
        keep_phi = keep_basis_fits(fit, sizeof(double) * (uint64_t)NP * rows * n_chunks, solve_reserve_bytes(N, M));
        if (!keep_phi) fit->ws[7].release();   // the memory is short: give back a basis kept before
    }
    fit->refine_stream = refined && !keep_phi;
    DBuf &d_phi = fit->ws[keep_phi ? 7 : 2];
    RP_CHECK(d_phi.grow(ctx, sizeof(double) * NP * rows * (keep_phi ? n_chunks : 1)));

    const size_t lds_bytes = gram_lds_bytes_dma<P>();
    RP_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(&k_oz_gram<P>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
    GramOperands go{d_slices.as<int8_t>(), (int64_t)NP * rows, rows};
    GramEpi ge{d_scale.as<double>(), d_tiles.as<int2>(), n_tiles, d_part.as<double>(), 0, halve};
    fit->kspace_products = 0;

    int chunk = 0;
    for (int64_t r0 = 0; r0 < M; r0 += rows, ++chunk) {
        const int64_t valid = std::min<int64_t>(rows, M - r0);
        // a short last chunk: each split shrinks to whole stages covering `valid` (the rows it takes
        // past `valid` are zero slices), so the Gram does not stream the chunk's empty tail
        const int rps_c = valid == rows ? rps : (int)round_up(ceil_div(valid, splits), gram_kd<P>());
        double *phi = d_phi.as<double>() + (keep_phi ? (int64_t)chunk * NP * rows : 0);
        k_zero_u64<<<ceil_div(NP, 256), 256, 0, st>>>(d_colmax.as<unsigned long long>(), NP);
        {
            KTimer kt(ctx, KT_KSPACE);
            sep_phi_chunk(fit, dv, plan, chunk, r0, valid, rows, phi, d_colmax.as<unsigned long long>(), st);
        }
        {
            KTimer kts(ctx, KT_SLICE);
            sep_slice_chunk(phi, NP, N + 2, rows, valid, (int64_t)rps_c * splits, d_colmax.as<unsigned long long>(),
                            d_slices.as<int8_t>(), d_scale.as<double>(), P, st);
        }
        KTimer ktg(ctx, KT_GRAM);
        ge.accumulate = chunk > 0;
        k_oz_gram<P><<<splits % 8 == 0 ? n_tiles * splits + 8 * halve : 8 * (int)ceil_div((int64_t)n_tiles * splits, 8),
                       NTHR, lds_bytes, st>>>(go, ge, rps_c, splits);
    }
    {
        KTimer ktr(ctx, KT_REDUCE);
        k_oz_reduce<<<n_tiles * REDUCE_PARTS, 256, 0, st>>>(d_part.as<double>(), splits, n_tiles, d_tiles.as<int2>(),
                                              fit->d_gram.as<double>(), NP, halve);
    }
    RP_HIP(ctx, hipGetLastError());
    fit->phi_kept = keep_phi;
    fit->phi_rows = rows;
    return RP_OK;
}

}  // namespace

int accumulate_split(rp_fit *fit, FitDevice &dv) {
    if (fit->opt.precision == RP_PREC_OZAKI_REFINED) return accumulate_split_p<3>(fit, dv, true);
    return accumulate_split_p<4>(fit, dv, false);
}

}  // namespace rp
