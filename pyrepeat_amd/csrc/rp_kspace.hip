// Reciprocal-space part of the basis, factorized over the lattice axes (split form).
//
// The reference sums phi_rec(g, a) = sum_k w_k cos(k . (g - R_a)) over all K k-vectors for every
// (point, atom) pair (repeat.py:831-846, 887-904; k-set and weights repeat.py:1077-1143).  A grid
// point is g = (i a0 + j a1) + m a2 and a k-vector has integer indices (h, k, l) over the
// reciprocal cell, so k . g = 2 pi (h i / nx + k j / ny + l m / nz) and
//
//     phi_rec(g, a) = Re sum_{h,k,l} c_a(h,k,l) X(h,i) Y(k,j) Z(l,m),
//     c_a(h,k,l) = w_hkl e^{-i k . R_a},  X(h,i) = e^{2 pi i h i / nx}  (Y, Z likewise).
//
// The triple sum is done as three short sums:
//     T1[m][a][h][k] = sum_l c_a(h,k,l) Z(l,m)        once per plane m and atom      (13 terms)
//     T2[a][h]       = sum_k T1[m][a][h][k] Y(k,j)    once per grid line (j, m)       (13 terms)
//     phi_rec        = Re sum_h T2[a][h] X(h,i)       once per kept point on the line (7 terms)
// all in fp64, which replaces the K = 754-term sum per (point, atom) by about 20 multiply-adds
// per (point, atom) plus 24 per (line, atom, h): the l and k sums are folded by conjugate symmetry.
// The phase factors come from the integer-reduced tables of the fp64 path, so the result equals the
// direct sum to fp64 rounding.
//
// Phi rows are produced per chunk of kept points (rows in ascending flat index, so a chunk is a
// run of whole or partial grid lines): one work-group takes up to SEP_LINES consecutive lines and
// 32 atoms; each wave forms T2 of those lines for 8 atoms (one lane per (atom, h)), then evaluates
// the points (one lane per (atom, point slot)) as a Horner sum in X(1, i), adds the real-space term
// where only the nearest image can be within the cutoff (fp64, erfc as one polynomial on the argument
// range [0, pi]) and writes Phi column-major.  The pairs with several images within the cutoff (cells
// narrower than twice the cutoff) get their real-space sum from k_sep_images afterwards, whose waves
// take one atom at 64 neighbouring points so that the image ranges agree across lanes.
#include "rp_fit_common.h"
#include "rp_erf_poly.h"

namespace rp {

namespace {

// A coefficient as an SGPR pair set up where it is used, and the FMA that takes it as its addend in
// the three-operand form.  Left to itself the compiler either keeps every coefficient in VGPRs for the
// life of the kernel (two per coefficient, a third of the register file here, which cuts the waves a
// SIMD holds) or copies each one into VGPRs for the two-operand FMA (three VALU issues per term).
template <uint64_t Bits>
__device__ __forceinline__ double sgpr_const() {
    uint32_t lo, hi;
    asm volatile("s_mov_b32 %0, %2\n\ts_mov_b32 %1, %3" : "=s"(lo), "=s"(hi)
                 : "i"((int32_t)(uint32_t)(Bits & 0xffffffffu)), "i"((int32_t)(uint32_t)(Bits >> 32)));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// a * b + c with c in SGPRs
__device__ __forceinline__ double fma_sgpr_addend(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
}

template <int D>
__device__ __forceinline__ double erf_poly_horner(double u, double p) {
    if constexpr (D < 0) {
        return p;
    } else {
        constexpr uint64_t bits = __builtin_bit_cast(uint64_t, kErfPoly[D]);
        return erf_poly_horner<D - 1>(u, fma_sgpr_addend(p, u, sgpr_const<bits>()));
    }
}

// 1/sqrt(x) for a positive normal x: the hardware estimate and one second-order Newton step,
// y = y0 + y0 e (1/2 + 3/8 e) with e = 1 - x y0^2, the steps the library's rsqrt takes; its select for
// zero, infinite and NaN inputs is left out (bitwise the same result otherwise; a pair at distance 0
// that is not the excluded self term gives NaN here instead of the library's infinity).
__device__ __forceinline__ double rsqrt_pos(double x) {
    const double y0 = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y0, y0, 1.0);
    return fma(y0 * e, fma(e, 0.375, 0.5), y0);
}

// erfc(sqrt(alpha) r) / r for 0 < r <= Rc (r^2 > 1e-20).  alpha = (pi / Rc)^2 (repeat.py:1087), so
// u = alpha r^2 lies in [0, pi^2] and erfc(sqrt(alpha) r) / r = 1/r - sqrt(alpha) P(u) with
// erf(x) = x P(x^2); P is one polynomial (max abs error of the erfc 7.5e-16), so the term needs one
// rsqrt and no table.
__device__ __forceinline__ double erfc_over_r(double r2, double alpha, double sqrt_alpha) {
    const double rinv = rsqrt_pos(r2);
    constexpr uint64_t mu0 = __builtin_bit_cast(uint64_t, -ERF_U0);
    const double u = fma_sgpr_addend(alpha, r2, sgpr_const<mu0>());
    constexpr uint64_t top = __builtin_bit_cast(uint64_t, kErfPoly[ERF_DEG]);
    const double p = erf_poly_horner<ERF_DEG - 1>(u, sgpr_const<top>());
    return fma(-sqrt_alpha, p, rinv);
}

constexpr int SH = 8;           // h slots per atom: lanes of a wave are (8 atoms) x (8 h or 8 point slots)
constexpr int HMAX = 6;         // the k-set has h^2 + k^2 + l^2 < 49, so 0 <= h <= 6
constexpr int SK = 13;          // k, l = -6..6
constexpr int SOFF = 6;
constexpr int SHT = HMAX + 1;   // h rows of T1 (h = 0..6)
constexpr int SBOX = SHT * SK;  // T1 entries per (plane, atom)
constexpr int SEP_WAVES = 4;
constexpr int SEP_ATOMS = 8 * SEP_WAVES;     // atoms per work-group, 8 per wave
constexpr int SEP_THREADS = 64 * SEP_WAVES;
constexpr int SEP_LINES = 5;   // lines per work-group: 23 KB of T2 rows, 6 work-groups per CU
constexpr int T1_PLANES = 16;   // planes per T1 work-group
constexpr int SEP_XMAX = 2048;  // longest axis 0 whose phase row X(1, .) the basis kernel keeps in LDS

// Z(-l, m) = conj Z(l, m), so a sum over l = -6..6 folds into
//     sum_l c(l) Z(l, m) = c(0) + sum_{l=1..6} [P(l) Re Z(l, m) + Q(l) Im Z(l, m)],
//     P(l) = c(l) + c(-l),  Q(l) = i (c(l) - c(-l)),
// 4 multiply-adds per l instead of 8 for the pair.  T1 is stored folded the same way over k, so that
// T2 = T1[0] + sum_{k=1..6} [P_k Re Y(k, j) + Q_k Im Y(k, j)]: slot 0 of row h holds T1[h][0], slot k
// holds T1[h][k] + T1[h][-k] and slot KF + k holds i (T1[h][k] - T1[h][-k]).  Only the phase rows of
// l, k >= 0 are read; the table's row of -l equals the conjugate of the row of l to an ulp.
constexpr int KF = SK / 2;   // folded pairs per row

// c_a(p) = w_p e^{-i k_p . R_a} for every atom and k-vector, once per fit: [atom][p].
__global__ __launch_bounds__(256) void k_sep_coef(int N, int K, const double *__restrict__ kvecs,
                                                  const double *__restrict__ kcoef, const double *__restrict__ atoms,
                                                  double2 *__restrict__ C) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (int64_t)N * K) return;
    const int a = (int)(t / K), p = (int)(t - (int64_t)a * K);
    const double ax = atoms[3 * a], ay = atoms[3 * a + 1], az = atoms[3 * a + 2];
    // the angle with the rounding of the atom tables: (kx x + ky y) + kz z
    const double th = (kvecs[3 * p] * ax + kvecs[3 * p + 1] * ay) + kvecs[3 * p + 2] * az;
    double s, co;
    sincos(th, &s, &co);
    const double w = kcoef[p];
    C[t] = make_double2(w * co, -(w * s));
}

// c_a(h, k, l) folded over l into f[0] = c(0), f[l] = P(l), f[KF + l] = Q(l) for one (h, k), from the
// atom's row of k_sep_coef's table.
__device__ __forceinline__ void t1_fold_coef(const int *__restrict__ kbox_row, const double2 *__restrict__ crow,
                                             double2 (&f)[SK]) {
    auto coef = [&](int ll) {
        const int p = kbox_row[ll];
        return p >= 0 ? crow[p] : make_double2(0.0, 0.0);
    };
    f[0] = coef(SOFF);
#pragma unroll
    for (int l = 1; l <= KF; ++l) {
        const double2 cp = coef(SOFF + l), cm = coef(SOFF - l);
        f[l] = make_double2(cp.x + cm.x, cp.y + cm.y);
        f[KF + l] = make_double2(cm.y - cp.y, cp.x - cm.x);
    }
}

__device__ __forceinline__ double2 t1_fold_sum(const double2 (&f)[SK], const double2 *__restrict__ ez, int nz, int m) {
    double2 acc = f[0];
#pragma unroll
    for (int l = 1; l <= KF; ++l) {
        const double2 z = ez[(l + 7) * nz + m];
        acc.x = fma(f[l].x, z.x, fma(f[KF + l].x, z.y, acc.x));
        acc.y = fma(f[l].y, z.x, fma(f[KF + l].y, z.y, acc.y));
    }
    return acc;
}

// Folded T1[m][a][h][.] for planes [blockIdx.y * T1_PLANES, +T1_PLANES); block = atom, lane = (h, k >= 0).
constexpr int T1_LANES = SHT * (KF + 1);
__global__ __launch_bounds__(64) void k_sep_t1(int N, int K, int nz, const int *__restrict__ kbox,
                                               const double2 *__restrict__ C, const double2 *__restrict__ ez,
                                               double2 *__restrict__ T1) {
    const int a = blockIdx.x, t = threadIdx.x;
    if (t >= T1_LANES) return;
    const int h = t / (KF + 1), k = t - h * (KF + 1);
    const double2 *crow = C + (int64_t)a * K;
    double2 fp[SK], fm[SK];
    t1_fold_coef(kbox + (h * SK + SOFF + k) * SK, crow, fp);
    if (k > 0) t1_fold_coef(kbox + (h * SK + SOFF - k) * SK, crow, fm);
    const int m0 = blockIdx.y * T1_PLANES, m1 = min(m0 + T1_PLANES, nz);
    for (int m = m0; m < m1; ++m) {
        double2 *row = T1 + ((int64_t)m * N + a) * SBOX + h * SK;
        const double2 tp = t1_fold_sum(fp, ez, nz, m);
        if (k == 0) {
            row[0] = tp;
        } else {
            const double2 tm = t1_fold_sum(fm, ez, nz, m);
            row[k] = make_double2(tp.x + tm.x, tp.y + tm.y);
            row[KF + k] = make_double2(tm.y - tp.y, tp.x - tm.x);
        }
    }
}

// Line starts among the kept points: row r starts a line when its line key f / nx differs from row r-1's.
__device__ __forceinline__ bool line_start(const int *__restrict__ idx, int64_t r, int64_t M, int nx) {
    return r < M && (r == 0 || idx[r] / nx != idx[r - 1] / nx);
}

__global__ __launch_bounds__(256) void k_line_count(const int *__restrict__ idx, int64_t M, int nx, int *__restrict__ cnt) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const unsigned long long b = __ballot(line_start(idx, r, M, nx));
    __shared__ int wsum[4];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// Exclusive offsets of the per-block counts, total in out[n].
__global__ __launch_bounds__(1024) void k_line_scan(const int *__restrict__ cnt, int64_t n, int64_t *__restrict__ out) {
    block_scan_counts_1024(cnt, n, out);
}

// line_row[L] = first row of line L (line_row[n_lines] = M), line_key[L] = j + m ny.
__global__ __launch_bounds__(256) void k_line_emit(const int *__restrict__ idx, int64_t M, int nx,
                                                   const int64_t *__restrict__ off, int64_t nblocks,
                                                   int *__restrict__ line_row, int *__restrict__ line_key) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool st = line_start(idx, r, M, nx);
    const unsigned long long b = __ballot(st);
    __shared__ int wbase[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wbase[wave] = __popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int w = 0; w < 4; ++w) {
            const int v = wbase[w];
            wbase[w] = run;
            run += v;
        }
    }
    __syncthreads();
    if (st) {
        const int64_t pos = off[blockIdx.x] + wbase[wave] + __popcll(b & ((1ull << lane) - 1ull));
        line_row[pos] = (int)r;
        line_key[pos] = idx[r] / nx;
    }
    if (r == M - 1) line_row[off[nblocks]] = (int)M;
}

// Per chunk c: the lines [first, end) that hold rows [c rows, min(M, (c + 1) rows)).
__global__ __launch_bounds__(256) void k_chunk_lines(const int *__restrict__ line_row, const int64_t *__restrict__ n_lines_p,
                                                     int64_t M, int64_t rows, int n_chunks, int *__restrict__ out) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n_chunks) return;
    const int64_t L = *n_lines_p;
    const int64_t r0 = (int64_t)c * rows, r1 = min(M, r0 + rows) - 1;
    // last line whose start is <= r
    auto line_of = [&](int64_t r) {
        int64_t lo = 0, hi = L - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (line_row[mid] <= r) lo = mid; else hi = mid - 1;
        }
        return lo;
    };
    out[2 * c] = (int)line_of(r0);
    out[2 * c + 1] = (int)line_of(r1) + 1;
}

// Fractional displacement d = g_f - a_f of a grid point (i, j, m) from an atom, its nearest image
// n = rint(d) and w = d - n; true when only that image can be within Rc, i.e. the bound
// min_i (1 - |w_i|) h_i on the distance to every other image exceeds Rc.  Written with explicit fma so
// that the basis kernel and the multi-image kernel classify every (point, atom) pair alike.
__device__ __forceinline__ bool nearest_only(const EpiArgs &e, double id, double jd, double md, double Fx, double Fy,
                                             double Fz, double &d0, double &d1, double &d2, double &n0, double &n1,
                                             double &n2) {
    d0 = fma(id, e.inv_n[0], -Fx);
    d1 = fma(jd, e.inv_n[1], -Fy);
    d2 = fma(md, e.inv_n[2], -Fz);
    n0 = rint(d0);
    n1 = rint(d1);
    n2 = rint(d2);
    const double b0 = (1.0 - fabs(d0 - n0)) * e.h[0], b1 = (1.0 - fabs(d1 - n1)) * e.h[1],
                 b2 = (1.0 - fabs(d2 - n2)) * e.h[2];
    return fmin(fmin(b0, b1), b2) > e.r_cut;
}

struct SepArgs {
    const int *line_row, *line_key;
    int l_first, l_end;
    int64_t r0, valid;
    const double2 *T1;          // null when there is no reciprocal part (fit-type 0)
    int N, NP;
    int ncols;                  // columns written: N + 2 (atoms, V, ones), or NP with the zero padding
    const double2 *ex, *ey;
    EpiArgs e;
    double *phi;                // [NP][rows]: column-major, chunk-local rows
    int64_t rows;               // leading dimension of phi
    unsigned long long *colmax;
    int *multi_flag;            // set when a pair of the chunk has several images within Rc
    int items, items_per_xcd;   // line groups of the chunk; line groups per XCD
    const int *gate, *gate_flags;   // nullable: run only while gate[0] != 0 and both flags are 0
};

// One wave takes 8 atoms of the work-group's lines; lane = (atom a_l = lane / 8, s = lane % 8).
// First the wave forms T2 of every line of the work-group (s = h), into a wave-private LDS block;
// then it streams the work-group's rows, which are contiguous across its lines (s = point slot),
// each lane following the line its rows fall in.  Waves never wait on each other after the line
// table is staged.
__global__ __launch_bounds__(SEP_THREADS) void k_sep_phi(SepArgs p) {
    // T2 rows of the wave's 8 atoms, padded to SH + 1 entries so the 8 atoms' reads of one h fall on
    // disjoint LDS banks
    __shared__ double2 t2w[SEP_WAVES][SEP_LINES][8][SH + 1];
    __shared__ int lrow[SEP_LINES + 1], lkey[SEP_LINES], lj[SEP_LINES], lm[SEP_LINES];
    extern __shared__ double2 xs[];   // X(1, i) for i < nx (when the reciprocal part is there)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int a_l = lane >> 3, s = lane & 7;
    // Work-groups go round-robin to the 8 XCDs (linear id L -> XCD L % 8): XCD x takes line groups
    // [x ipx, (x + 1) ipx) and walks the atom blocks in order, so its L2 keeps the T1 rows of the
    // few planes those lines lie in.
    const int L = blockIdx.x, xcd = L & 7, jj = L >> 3;
    const int item = xcd * p.items_per_xcd + jj % p.items_per_xcd;
    const int ablk = jj / p.items_per_xcd;
    if (item >= p.items) return;
    if (p.gate && (p.gate[0] == 0 || p.gate_flags[0] != 0 || p.gate_flags[1] != 0)) return;
    const int a = ablk * SEP_ATOMS + wave * 8 + a_l;
    const EpiArgs &e = p.e;
    const int nx = e.geom.nx, ny = e.geom.ny;
    const bool atom = a < p.N;
    const bool recip = atom && p.T1 != nullptr;
    const int l_begin = p.l_first + item * SEP_LINES;
    const int n_l = min(SEP_LINES, p.l_end - l_begin);
    // loads whose latency runs beside the staging: the atom coordinates of the line-term pass and of the
    // point step, and the grid indices of the first block of rows
    const int gl = lane >> 3, ga = ablk * SEP_ATOMS + wave * 8 + (lane & 7);
    const bool geo = e.fit_type != 0 && gl < n_l && ga < p.N;
    const double Fx = atom ? e.afrac[3 * a] : 0.0;
    double gFy = 0.0, gFz = 0.0;
    if (geo) {
        gFy = e.afrac[3 * ga + 1];
        gFz = e.afrac[3 * ga + 2];
    }
    // rows of the work-group (absolute; they fit in 32 bits)
    const int rs = (int)max<int64_t>(__builtin_amdgcn_readfirstlane(p.line_row[l_begin]), p.r0);
    const int rend = (int)min<int64_t>(__builtin_amdgcn_readfirstlane(p.line_row[l_begin + n_l]), p.r0 + p.valid);
    int f_blk = e.idx[min(rs + lane, rend - 1)];
    if ((int)threadIdx.x <= n_l) lrow[threadIdx.x] = p.line_row[l_begin + threadIdx.x];
    if ((int)threadIdx.x < n_l) {
        const int key = p.line_key[l_begin + threadIdx.x];
        lkey[threadIdx.x] = key;
        lm[threadIdx.x] = key / ny;
        lj[threadIdx.x] = key - (key / ny) * ny;
    }
    if (p.T1 != nullptr)
        for (int i = threadIdx.x; i < nx; i += SEP_THREADS) xs[i] = p.ex[nx + i];
    __syncthreads();
    // ---- per (line, atom), one lane each: the parts of the nearest-image test and of the displacement
    // that depend on the line (j, m) only, into entries 7 and 8 of the atom's row (entries the T2 pass
    // below leaves alone: h <= HMAX = 6):
    //     (t_x, t_y), (t_z, +-b12),  t = w1 b1 + w2 b2 (rows of the cell, fma as the point step uses),
    //     b12 = min((1 - |w1|) h1, (1 - |w2|) h2), + when the image (n1, n2) is (0, 0), - otherwise,
    // so that the point step is left with the terms of grid axis 0 (nearest_only() split in two).
    {
        if (geo) {
            const double d1 = fma((double)lj[gl], e.inv_n[1], -gFy);
            const double d2 = fma((double)lm[gl], e.inv_n[2], -gFz);
            const double n1 = rint(d1), n2 = rint(d2);
            const double w1 = d1 - n1, w2 = d2 - n2;
            const double b12 = fmin((1.0 - fabs(w1)) * e.h[1], (1.0 - fabs(w2)) * e.h[2]);
            t2w[wave][gl][lane & 7][SH - 1] =
                make_double2(fma(w1, e.box[3], w2 * e.box[6]), fma(w1, e.box[4], w2 * e.box[7]));
            t2w[wave][gl][lane & 7][SH] =
                make_double2(fma(w1, e.box[5], w2 * e.box[8]), (n1 == 0.0 && n2 == 0.0) ? b12 : -b12);
        }
    }
    // ---- T2 of every line (lane s = h < 7); the line key is uniform, so the Y row comes through the
    // scalar cache
    if (recip && s < SHT) {
        double2 t1[SK];
        int cur_m = -1;
        for (int li = 0; li < n_l; ++li) {
            const int m = __builtin_amdgcn_readfirstlane(lm[li]);
            const int j = __builtin_amdgcn_readfirstlane(lj[li]);
            if (m != cur_m) {
                cur_m = m;
                const double2 *src = p.T1 + ((int64_t)m * p.N + a) * SBOX + s * SK;
#pragma unroll
                for (int kk = 0; kk < SK; ++kk) t1[kk] = src[kk];
            }
            double2 t2 = t1[0];
#pragma unroll
            for (int k = 1; k <= KF; ++k) {
                const double2 y = p.ey[(k + 7) * ny + j];
                t2.x = fma(t1[k].x, y.x, fma(t1[KF + k].x, y.y, t2.x));
                t2.y = fma(t1[k].y, y.x, fma(t1[KF + k].y, y.y, t2.y));
            }
            t2w[wave][li][a_l][s] = t2;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // ---- points (lane s = point slot), in blocks of 64 rows: lane L of the wave loads the grid index of
    // row rb + L (one coalesced load per block, issued a block ahead) and packs its column i and line
    // slot li as (i << 3) | li (i < 2048); a step fetches its row's pack from the lane that holds it
    // (ds_bpermute), so it waits neither on a global load nor on the line table.  Rows are absolute
    // and fit in 32 bits.
    const double alpha = e.sqrt_alpha * e.sqrt_alpha;
    double cmax = 0.0;
    bool multi = false;   // a pair of this lane has several images within Rc
    {
        // line boundaries (uniform); slots past the last line never match
        int bnd[SEP_LINES - 1];
#pragma unroll
        for (int q = 0; q < SEP_LINES - 1; ++q)
            bnd[q] = q + 1 < n_l ? __builtin_amdgcn_readfirstlane(lrow[q + 1]) : 0x7fffffff;
        double *const out0 = p.phi + (int64_t)a * p.rows - p.r0;
        // every lane runs the block loop (the packs are exchanged across the whole wave); the atom
        // columns' lanes evaluate the pairs
        for (int rb = rs; rb < rend; rb += 64) {
            int pk;
            {
                const int rr = rb + lane;
                int li = 0;
#pragma unroll
                for (int q = 0; q < SEP_LINES - 1; ++q) li += rr >= bnd[q] ? 1 : 0;
                pk = ((f_blk - lkey[li] * nx) << 3) | li;
            }
            f_blk = e.idx[min(rb + 64 + lane, rend - 1)];
            // the next step's pack is fetched a step ahead
            int pnext = __shfl(pk, s, 64);
            for (int k = 0; k < 8 && rb + 8 * k < rend; ++k) {
                const int r = rb + 8 * k + s;
                const int prow = pnext;
                if (k < 7) pnext = __shfl(pk, 8 * (k + 1) + s, 64);
                if (r >= rend || !atom) continue;
                const int i = prow >> 3, li = prow & 7;
                double *out = out0 + r;
                // (entries 7 and 8 are read whatever the fit type: fit-type 0 does not use them)
                const double2 g01 = t2w[wave][li][a_l][SH - 1], g2 = t2w[wave][li][a_l][SH];
                double acc = 0.0;
                if (p.T1 != nullptr) {
                    // Re sum_h T2[h] X(1, i)^h by Horner from h = HMAX
                    const double2 x = xs[i];
                    double2 t = t2w[wave][li][a_l][HMAX];
#pragma unroll
                    for (int q = HMAX - 1; q >= 1; --q) {
                        const double2 c = t2w[wave][li][a_l][q];
                        const double tx = fma(t.x, x.x, fma(-t.y, x.y, c.x));
                        const double ty = fma(t.x, x.y, fma(t.y, x.x, c.y));
                        t = make_double2(tx, ty);
                    }
                    acc = fma(t.x, x.x, fma(-t.y, x.y, t2w[wave][li][a_l][0].x));
                }
                const double id = (double)i;
                double real = 0.0;
                bool mine = true;   // val is the pair's final value (not left to the multi-image kernel)
                if (e.fit_type == 0) {
                    const int key = lkey[li];
                    const double jd = (double)(key % ny), md = (double)(key / ny);
                    const double gx = (id * e.geom.ax[0] + jd * e.geom.ax[3]) + md * e.geom.ax[6];
                    const double gy = (id * e.geom.ax[1] + jd * e.geom.ax[4]) + md * e.geom.ax[7];
                    const double gz = (id * e.geom.ax[2] + jd * e.geom.ax[5]) + md * e.geom.ax[8];
                    const double dx = gx - e.atoms[3 * a], dy = gy - e.atoms[3 * a + 1], dz = gz - e.atoms[3 * a + 2];
                    const double d = sqrt((dx * dx + dy * dy) + dz * dz);
                    real = d > 1e-10 ? 1.0 / d : 0.0;
                } else {
                    // nearest_only() with the line's terms from entries 7 and 8
                    const double d0 = fma(id, e.inv_n[0], -Fx);
                    const double n0 = rint(d0);
                    const double w0 = d0 - n0;
                    const double b0 = (1.0 - fabs(w0)) * e.h[0];
                    if (fmin(b0, fabs(g2.y)) > e.r_cut) {
                        // only the nearest image can be within Rc: displacement w . cell to it
                        const double dx = fma(w0, e.box[0], g01.x);
                        const double dy = fma(w0, e.box[1], g01.y);
                        const double dz = fma(w0, e.box[2], g2.x);
                        const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                        if (r2 <= e.R2 && !(n0 == 0.0 && g2.y > 0.0 && r2 <= 1e-20))
                            real = erfc_over_r(r2, alpha, e.sqrt_alpha);
                    } else {
                        // several images may lie within Rc: k_sep_images adds that sum afterwards
                        mine = false;
                        multi = true;
                    }
                }
                const double val = acc + real;
                *out = val;
                if (mine) cmax = fmax(cmax, fabs(val));
            }
        }
    }
    if (!atom && a < p.ncols) {
        // the V column, the ones column and (when asked for) the zero padding of the atom columns
        const int64_t r1 = rs + s;
        double *out = p.phi + (int64_t)a * p.rows + (r1 - p.r0);
        for (int64_t r = r1; r < rend; r += SH, out += SH) {
            const double val = a == p.N ? e.v[r] : (a == p.N + 1 ? 1.0 : 0.0);
            *out = val;
            cmax = fmax(cmax, fabs(val));
        }
    }
    if (__ballot(multi) != 0 && lane == 0) *p.multi_flag = 1;
    // column maximum over the 8 point slots of the atom
    cmax = fmax(cmax, __shfl_xor(cmax, 1, 64));
    cmax = fmax(cmax, __shfl_xor(cmax, 2, 64));
    cmax = fmax(cmax, __shfl_xor(cmax, 4, 64));
    if (s == 0 && cmax > 0.0 && a < p.NP) atomicMax(p.colmax + a, (unsigned long long)__double_as_longlong(cmax));
}

// The real-space sum of the (point, atom) pairs with several images within Rc, added to their Phi
// entries after k_sep_phi (which leaves those sums out and sets *flag).  Work-group = 64 consecutive kept
// points (lane = point) x 32 atoms, 8 per wave taken one after another: a wave's lanes hold neighbouring
// points of one atom, so their image ranges (EpiArgs::image_sum_sphere) nearly agree.  The column maxima
// take the completed entries.
struct ImageArgs {
    EpiArgs e;
    int64_t r0, valid, rows;
    int N;
    double *phi;
    unsigned long long *colmax;
    const int *flag;
    const int *gate, *gate_flags;
};
__global__ __launch_bounds__(256) void k_sep_images(ImageArgs p) {
    __shared__ double cmx[32][65];   // per atom of the work-group and lane, the largest completed |entry|
    if (p.flag[0] == 0) return;
    if (p.gate && (p.gate[0] == 0 || p.gate_flags[0] != 0 || p.gate_flags[1] != 0)) return;
    const EpiArgs &e = p.e;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t rl = (int64_t)blockIdx.x * 64 + lane;   // chunk-local row
    const bool row = rl < p.valid;
    int i = 0, j = 0, m = 0;
    if (row) e.geom.ijk(e.idx[p.r0 + rl], i, j, m);
    const double id = (double)i, jd = (double)j, md = (double)m;
    const double gx = (id * e.geom.ax[0] + jd * e.geom.ax[3]) + md * e.geom.ax[6];
    const double gy = (id * e.geom.ax[1] + jd * e.geom.ax[4]) + md * e.geom.ax[7];
    const double gz = (id * e.geom.ax[2] + jd * e.geom.ax[5]) + md * e.geom.ax[8];
    const double alpha = e.sqrt_alpha * e.sqrt_alpha, sa = e.sqrt_alpha;
    const int a0 = (blockIdx.y * 4 + wave) * 8;
#pragma unroll 1
    for (int q = 0; q < 8; ++q) {
        const int a = a0 + q;   // uniform
        if (a >= p.N) {
            cmx[wave * 8 + q][lane] = 0.0;
            continue;
        }
        double *o = p.phi + (int64_t)a * p.rows + rl;
        const double before = row ? *o : 0.0;   // in flight while the sum runs
        const double Fx = e.afrac[3 * a], Fy = e.afrac[3 * a + 1], Fz = e.afrac[3 * a + 2];
        double d0, d1, d2, n0, n1, n2;
        const bool multi = row && !nearest_only(e, id, jd, md, Fx, Fy, Fz, d0, d1, d2, n0, n1, n2);
        double cm = 0.0;
        if (multi) {
            const double real = e.image_sum_sphere(d0, d1, d2, gx, gy, gz, e.atoms[3 * a], e.atoms[3 * a + 1],
                                                   e.atoms[3 * a + 2],
                                                   [alpha, sa](double r2) { return erfc_over_r(r2, alpha, sa); });
            const double v = before + real;
            *o = v;
            cm = fabs(v);
        }
        cmx[wave * 8 + q][lane] = cm;
    }
    // column maxima: thread (atom t / 8, part t % 8) takes 8 lanes' values, then 3 exchanges
    __syncthreads();
    const int t = threadIdx.x, ca = t >> 3, part = t & 7;
    double cm = 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) cm = fmax(cm, cmx[ca][part * 8 + u]);
    cm = fmax(cm, __shfl_xor(cm, 1, 64));
    cm = fmax(cm, __shfl_xor(cm, 2, 64));
    cm = fmax(cm, __shfl_xor(cm, 4, 64));
    const int a = blockIdx.y * 32 + ca;
    if (part == 0 && cm > 0.0 && a < p.N) atomicMax(p.colmax + a, (unsigned long long)__double_as_longlong(cm));
}

__global__ void k_zero_flag(int *f) { *f = 0; }

// Column-major Phi -> slices [P][NP][rows] of x / S_col, S_col the power of two with
// max |Phi_col| / S_col <= 127/128; rows [valid, span) are zero, rows from span on are not written.
// Both sides are streams: a thread takes 16 consecutive rows of one column (128-byte read, one
// 16-byte write per slice).
template <int P>
__global__ __launch_bounds__(256) void k_slice_cols(const double *__restrict__ phi, int64_t rows, int64_t valid,
                                                    int64_t span, int ncols,
                                                    const unsigned long long *__restrict__ colmax,
                                                    int8_t *__restrict__ out, double *__restrict__ scale_out) {
    const int a = blockIdx.y;
    const int64_t r0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 16;
    const double mx = __longlong_as_double((long long)colmax[a]);
    double S = 1.0;
    if (mx > 0.0) S = ldexp(1.0, (int)ceil(log2(mx * (128.0 / 127.0))));
    while (mx / S > 127.0 / 128.0) S *= 2.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) scale_out[a] = S;
    if (r0 >= span) return;
    const double inv = 1.0 / S;
    const double *src = phi + (int64_t)a * rows + r0;
    double v[16];
    if (a >= ncols) {   // padding column: zero slices, its Phi entries are not written
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = 0.0;
    } else if (r0 + 16 <= valid) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const double2 t = reinterpret_cast<const double2 *>(src)[q];
            v[2 * q] = t.x;
            v[2 * q + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = r0 + q < valid ? src[q] : 0.0;
    }
    uint32_t w[P][4];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int s = 0; s < P; ++s) w[s][b] = 0u;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        double r = v[q] * inv;
#pragma unroll
        for (int s = 0; s < P; ++s) {
            const double tv = r * 128.0;
            const double c = rint(tv);
            r = tv - c;
            w[s][q >> 2] |= (uint32_t)(uint8_t)(int8_t)(int)c << (8 * (q & 3));
        }
    }
    const int64_t plane = (int64_t)gridDim.y * rows;
#pragma unroll
    for (int s = 0; s < P; ++s)
        *reinterpret_cast<uint4 *>(out + s * plane + (int64_t)a * rows + r0) = make_uint4(w[s][0], w[s][1], w[s][2], w[s][3]);
}

}  // namespace

__global__ __launch_bounds__(256) void k_ints_to_host(const int *__restrict__ src, int n, int *__restrict__ dst) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

int sep_prepare(rp_fit *fit, FitDevice &dv, int64_t rows, SepPlan &plan) {
    Ctx *ctx = fit->ctx;
    hipStream_t st = ctx->stream;
    const Setup &s = fit->setup;
    const int N = fit->n_atoms, K = dv.K;
    const int64_t M = fit->n_keep;
    if (K > 0 && s.nx > SEP_XMAX)
        return ctx->fail(RP_EINVAL, "grid axis 0 longer than the basis kernel's phase table (2048 points)");
    if (K > 0) {
        if (!dv.sep_kbox_ok) {   // the k-set's index box, once per Ewald setup
            std::vector<int> kbox(SH * SK * SK, -1);
            for (int p = 0; p < K; ++p) {
                const int h = s.khkl[3 * p], k = s.khkl[3 * p + 1] + SOFF, l = s.khkl[3 * p + 2] + SOFF;
                if (h < 0 || h > HMAX || k < 0 || k >= SK || l < 0 || l >= SK)
                    return ctx->fail(RP_EINVAL, "k-vector index outside the factorized box");
                kbox[(h * SK + k) * SK + l] = p;
            }
            RP_CHECK(dv.sep_kbox.grow(ctx, sizeof(int) * kbox.size()));
            RP_HIP(ctx, hipMemcpyAsync(dv.sep_kbox.p, kbox.data(), sizeof(int) * kbox.size(), hipMemcpyHostToDevice, st));
            RP_HIP(ctx, hipStreamSynchronize(st));   // the host vector goes out of scope
            dv.sep_kbox_ok = true;
        }
        RP_CHECK(dv.sep_t1.grow(ctx, sizeof(double2) * (size_t)s.nz * N * SBOX));
        RP_CHECK(dv.sep_coef.grow(ctx, sizeof(double2) * (size_t)N * K));
    }
    RP_CHECK(dv.multi_flag.grow(ctx, sizeof(int)));
    // line table of the kept points
    const int64_t nb = ceil_div(M, 256);
    RP_CHECK(dv.line_cnt.grow(ctx, sizeof(int) * nb));
    RP_CHECK(dv.line_off.grow(ctx, sizeof(int64_t) * (nb + 1)));
    RP_CHECK(dv.line_row.grow(ctx, sizeof(int) * (M + 1)));
    RP_CHECK(dv.line_key.grow(ctx, sizeof(int) * M));
    k_line_count<<<nb, 256, 0, st>>>(fit->d_idx.as<int>(), M, s.nx, dv.line_cnt.as<int>());
    k_line_scan<<<1, 1024, 0, st>>>(dv.line_cnt.as<int>(), nb, dv.line_off.as<int64_t>());
    k_line_emit<<<nb, 256, 0, st>>>(fit->d_idx.as<int>(), M, s.nx, dv.line_off.as<int64_t>(), nb, dv.line_row.as<int>(),
                                    dv.line_key.as<int>());
    const int n_chunks = (int)ceil_div(M, rows);
    RP_CHECK(dv.chunk_lines.grow(ctx, sizeof(int) * 2 * n_chunks));
    k_chunk_lines<<<ceil_div(n_chunks, 256), 256, 0, st>>>(dv.line_row.as<int>(), dv.line_off.as<int64_t>() + nb, M, rows,
                                                           n_chunks, dv.chunk_lines.as<int>());
    RP_HIP(ctx, hipGetLastError());
    // the line ranges into the mapped host slots (after the filter's 8 doubles)
    RP_CHECK(fit->host_slots(ctx, std::max<size_t>(4096, 64 + sizeof(int) * 2 * (size_t)n_chunks)));
    k_ints_to_host<<<ceil_div(2 * n_chunks, 256), 256, 0, st>>>(dv.chunk_lines.as<int>(), 2 * n_chunks,
                                                               reinterpret_cast<int *>(fit->h_small_dev + 8));
    RP_HIP(ctx, hipGetLastError());
    // the T1 tables go on the stream behind the line ranges, so the GPU computes them while the
    // host waits for the ranges and plans the chunks
    if (!dv.ev_plan) RP_HIP(ctx, hipEventCreateWithFlags(&dv.ev_plan, hipEventDisableTiming));
    RP_HIP(ctx, hipEventRecord(dv.ev_plan, st));
    if (K > 0) {
        k_sep_coef<<<ceil_div((int64_t)N * K, 256), 256, 0, st>>>(N, K, dv.kvecs.as<double>(), dv.kcoef.as<double>(),
                                                                 fit->d_atoms.as<double>(), dv.sep_coef.as<double2>());
        dim3 g(N, (unsigned)ceil_div(s.nz, T1_PLANES));
        k_sep_t1<<<g, 64, 0, st>>>(N, K, s.nz, dv.sep_kbox.as<int>(), dv.sep_coef.as<double2>(), dv.ez.as<double2>(),
                                    dv.sep_t1.as<double2>());
        RP_HIP(ctx, hipGetLastError());
    }
    RP_HIP(ctx, hipEventSynchronize(dv.ev_plan));
    const int *cl = reinterpret_cast<const int *>(fit->h_small + 8);
    plan.first_line.resize(n_chunks);
    plan.end_line.resize(n_chunks);
    for (int c = 0; c < n_chunks; ++c) {
        plan.first_line[c] = cl[2 * c];
        plan.end_line[c] = cl[2 * c + 1];
        if (plan.end_line[c] <= plan.first_line[c]) return ctx->fail(RP_EINVAL, "empty chunk in the line table");
    }
    return RP_OK;
}

void sep_phi_chunk(rp_fit *fit, FitDevice &dv, const SepPlan &plan, int chunk, int64_t r0, int64_t valid, int64_t rows,
                   double *phi, unsigned long long *colmax, hipStream_t st, const int *gate, const int *gate_flags,
                   bool zero_pad) {
    SepArgs a{};
    a.gate = gate;
    a.gate_flags = gate_flags;
    a.rows = rows;
    a.line_row = dv.line_row.as<int>();
    a.line_key = dv.line_key.as<int>();
    a.l_first = plan.first_line[chunk];
    a.l_end = plan.end_line[chunk];
    a.r0 = r0;
    a.valid = valid;
    a.T1 = dv.K > 0 ? dv.sep_t1.as<double2>() : nullptr;
    a.N = fit->n_atoms;
    a.NP = dv.NP;
    a.ncols = zero_pad ? dv.NP : fit->n_atoms + 2;
    a.ex = dv.ex.as<double2>();
    a.ey = dv.ey.as<double2>();
    a.e = dv.epi;
    a.phi = phi;
    a.colmax = colmax;
    a.items = (int)ceil_div(a.l_end - a.l_first, SEP_LINES);
    a.items_per_xcd = (int)ceil_div(a.items, 8);
    const int blocks = 8 * a.items_per_xcd * (int)ceil_div(a.ncols, SEP_ATOMS);
    const bool images = dv.epi.fit_type != 0;
    if (images) k_zero_flag<<<1, 1, 0, st>>>(dv.multi_flag.as<int>());
    a.multi_flag = dv.multi_flag.as<int>();
    k_sep_phi<<<blocks, SEP_THREADS, dv.K > 0 ? sizeof(double2) * fit->setup.nx : 0, st>>>(a);
    if (images) {
        ImageArgs b{};
        b.e = dv.epi;
        b.r0 = r0;
        b.valid = valid;
        b.rows = rows;
        b.N = fit->n_atoms;
        b.phi = phi;
        b.colmax = colmax;
        b.flag = dv.multi_flag.as<int>();
        b.gate = gate;
        b.gate_flags = gate_flags;
        dim3 g((unsigned)ceil_div(valid, 64), (unsigned)ceil_div(fit->n_atoms, 32));
        k_sep_images<<<g, 256, 0, st>>>(b);
    }
}

void sep_slice_chunk(const double *phi, int NP, int ncols, int64_t rows, int64_t valid, int64_t span,
                     const unsigned long long *colmax, int8_t *slices, double *scale_out, int P, hipStream_t st) {
    dim3 g((unsigned)ceil_div(span, 256 * 16), (unsigned)NP);
    if (P == 3)
        k_slice_cols<3><<<g, 256, 0, st>>>(phi, rows, valid, span, ncols, colmax, slices, scale_out);
    else
        k_slice_cols<4><<<g, 256, 0, st>>>(phi, rows, valid, span, ncols, colmax, slices, scale_out);
}

}  // namespace rp
