// Constrained least squares on the device (fit_charges steps 4-5, repeat.py:2266-2313).
//
//   A = ATA - n phibar phibar^T,  b = ATb - n phibar Vbar           (repeat.py:2266-2276)
//   symmetry merge of rows/columns and b, group sizes as weights    (repeat.py:2445-2470)
//   RESP / QEq diagonal terms                                       (repeat.py:2511-2535)
//   [[A, w], [w^T, 0]] [q; lambda] = [b; q_tot]                     (repeat.py:2293-2304)
//   The reference solves the bordered system with LAPACK LU (scipy.linalg.solve,
//   repeat.py:2307).  Here A (symmetric, positive definite for any fit with more kept
//   points than unknowns) is factored by a blocked Cholesky and the Lagrange row is
//   eliminated by its Schur complement:
//       y_b = A^-1 b,  y_w = A^-1 w,  lambda = (w.y_b - q_tot) / (w.y_w),  q = y_b - lambda y_w,
//   the same solution of the same system.  If a Cholesky pivot is not positive
//   (indefinite or singular A) the bordered system is solved by LU with partial
//   pivoting instead; an exactly zero LU pivot is reported as RP_ESINGULAR, where the
//   reference falls back to lstsq.
//
// Cholesky: right-looking, 64-column blocks: the diagonal block is factored in LDS (8 x 8
// sub-blocks) together with its inverse; the panel below is the product with that inverse, the
// trailing lower triangle takes a rank-64 update.  The forward substitution of w and b rides
// along with the factorisation; the multiplier and one backward substitution finish the solve.
//
// LU fallback: the augmented matrix [K | rhs] (n x (n+1), n = N'+1) is factored in place,
// right-looking and blocked by NB columns: one-workgroup panel factorisation,
// row interchanges applied to the other columns, a triangular solve for the U
// block row and a rank-NB update of the trailing matrix.  The rhs column rides
// along, so it holds L^-1 P b at the end; a back substitution finishes.
//
// Refined split form: the Cholesky solution of the split Gram's system is then corrected by iterative
// refinement against the fp64 basis that the accumulate kept (the refinement kernels below), so the
// result solves the fp64 normal equations.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "rp_fit_common.h"
#include "rp_mfma.h"

namespace rp {

namespace {

constexpr int NB = 32;

// Refinement state of a solve (device): go = the next step runs, steps run so far, max |dq| of the last
// correction.
struct RefineState {
    int go;
    int steps;
    double last_dq;
    double bound;   // estimated max |q - q_converged| left after the last step (e); -1 before the first
};

// Contraction assumed for the bound after a single step, where no ratio of corrections exists yet:
// twice the largest per-step contraction measured (5.3e-3, a 5000-atom cell with cond(A) = 4.9e10).
constexpr double RHO_PRIOR = 1e-2;
// Refinement steps queued with the solve; more only after a synchronisation shows they are needed.
constexpr int REFINE_EAGER = 1;
// Relative perturbation of the sensitivity estimate: 8 units in the last place of an fp64 entry.
constexpr double SENS_EPS = 8.0 * 2.220446049250313e-16;

__global__ __launch_bounds__(256) void k_assemble(const double *__restrict__ gram, int NP, int N,
                                                  const int *__restrict__ red_ptr, const int *__restrict__ red_src,
                                                  const double *__restrict__ red_mult, const double *__restrict__ w,
                                                  const double *__restrict__ diag_add, const double *__restrict__ rhs_add,
                                                  double q_tot, int nr, double *__restrict__ K, double *__restrict__ K2,
                                                  int ld, double *__restrict__ wc, double *__restrict__ bc,
                                                  int *__restrict__ cinfo, RefineState *__restrict__ rs) {
    const int n = nr + 1;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e == 0) {
        cinfo[0] = 0;
        cinfo[1] = 0;
        rs->go = 1;
        rs->steps = 0;
        rs->last_dq = 0.0;
        rs->bound = -1.0;
    }
    if (e >= (int64_t)n * (n + 1)) return;
    const int I = (int)(e / (n + 1)), J = (int)(e % (n + 1));
    const double ng = gram[(int64_t)(N + 1) * NP + N + 1];
    const double vbar = gram[(int64_t)N * NP + N + 1] / ng;
    double val = 0.0;
    if (J < nr && I < nr) {
        for (int p = red_ptr[I]; p < red_ptr[I + 1]; ++p) {
            const int a = red_src[p];
            const double pa = gram[(int64_t)a * NP + N + 1] / ng;
            for (int q = red_ptr[J]; q < red_ptr[J + 1]; ++q) {
                const int b = red_src[q];
                const double pb = gram[(int64_t)b * NP + N + 1] / ng;
                const double c = gram[(int64_t)a * NP + b] - ng * (pa * pb);
                val += red_mult[p] * red_mult[q] * c;
            }
        }
        if (I == J && diag_add) val += diag_add[I];
    } else if (J < nr && I == nr) {
        val = w[J];
    } else if (J == nr && I < nr) {
        val = w[I];
    } else if (J == nr && I == nr) {
        val = 0.0;
    } else if (J == n) {
        if (I < nr) {
            for (int p = red_ptr[I]; p < red_ptr[I + 1]; ++p) {
                const int a = red_src[p];
                const double pa = gram[(int64_t)a * NP + N + 1] / ng;
                val += red_mult[p] * (gram[(int64_t)a * NP + N] - (ng * pa) * vbar);
            }
            if (rhs_add) val += rhs_add[I];
        } else {
            val = q_tot;
        }
    }
    K[(int64_t)I * ld + J] = val;
    K2[(int64_t)I * ld + J] = val;   // kept for the LU fallback
    if (I < nr && J == nr) wc[I] = val;
    if (I < nr && J == n) bc[I] = val;
}

// Factor columns [k0, k0+nb) of rows [k0, n): pivot search, in-panel swap, scale, rank-1 update.
__global__ __launch_bounds__(1024) void k_panel(double *__restrict__ A, int ld, int n, int k0, int nb,
                                                int *__restrict__ piv, int *__restrict__ info) {
    __shared__ double sval[1024];
    __shared__ int sidx[1024];
    const int t = threadIdx.x;
    for (int j = k0; j < k0 + nb; ++j) {
        double best = -1.0;
        int bi = n;
        for (int r = j + t; r < n; r += 1024) {
            const double v = fabs(A[(int64_t)r * ld + j]);
            if (v > best) {
                best = v;
                bi = r;
            }
        }
        sval[t] = best;
        sidx[t] = bi;
        __syncthreads();
        for (int s = 512; s > 0; s >>= 1) {
            if (t < s) {
                const double o = sval[t + s];
                const int oi = sidx[t + s];
                if (o > sval[t] || (o == sval[t] && oi < sidx[t])) {
                    sval[t] = o;
                    sidx[t] = oi;
                }
            }
            __syncthreads();
        }
        const int p = sidx[0];
        const bool zero = !(sval[0] > 0.0);
        __syncthreads();
        if (t == 0) {
            piv[j] = p;
            if (zero && info[0] == 0) info[0] = j + 1;
        }
        if (p != j && t < nb) {
            double *rj = A + (int64_t)j * ld + k0 + t;
            double *rp = A + (int64_t)p * ld + k0 + t;
            const double tmp = *rj;
            *rj = *rp;
            *rp = tmp;
        }
        __syncthreads();
        const double pv = A[(int64_t)j * ld + j];
        if (!zero)
            for (int r = j + 1 + t; r < n; r += 1024) A[(int64_t)r * ld + j] /= pv;
        __syncthreads();
        const int ncol = k0 + nb - j - 1;
        if (ncol > 0) {
            const int64_t tot = (int64_t)(n - j - 1) * ncol;
            for (int64_t e = t; e < tot; e += 1024) {
                const int r = j + 1 + (int)(e / ncol);
                const int c = j + 1 + (int)(e % ncol);
                A[(int64_t)r * ld + c] -= A[(int64_t)r * ld + j] * A[(int64_t)j * ld + c];
            }
        }
        __syncthreads();
    }
}

// Apply the panel's interchanges (in order) to every column outside the panel.
__global__ __launch_bounds__(256) void k_swap_rest(double *__restrict__ A, int ld, int ncols, int k0, int nb,
                                                   const int *__restrict__ piv) {
    int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncols - nb) return;
    if (c >= k0) c += nb;
    for (int j = k0; j < k0 + nb; ++j) {
        const int p = piv[j];
        if (p != j) {
            const double tmp = A[(int64_t)j * ld + c];
            A[(int64_t)j * ld + c] = A[(int64_t)p * ld + c];
            A[(int64_t)p * ld + c] = tmp;
        }
    }
}

// U12 = L11^{-1} A12 for the block row [k0, k0+nb), columns [k0+nb, ncols).
__global__ __launch_bounds__(256) void k_trsm(double *__restrict__ A, int ld, int ncols, int k0, int nb) {
    __shared__ double L[NB][NB];
    for (int e = threadIdx.x; e < nb * nb; e += 256) L[e / nb][e % nb] = A[(int64_t)(k0 + e / nb) * ld + k0 + e % nb];
    __syncthreads();
    const int c = k0 + nb + blockIdx.x * 256 + threadIdx.x;
    if (c >= ncols) return;
    double x[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        if (i < nb) {
            double v = A[(int64_t)(k0 + i) * ld + c];
            for (int q = 0; q < i; ++q) v -= L[i][q] * x[q];
            x[i] = v;
            A[(int64_t)(k0 + i) * ld + c] = v;
        }
    }
}

// A22 -= L21 U12 over rows [r0, n) and columns [r0, ncols), r0 = k0 + nb; 64x64 tiles.
__global__ __launch_bounds__(256) void k_update(double *__restrict__ A, int ld, int n, int ncols, int k0, int nb) {
    __shared__ double Ls[64][NB + 1];
    __shared__ double Us[NB][64];
    const int r0 = k0 + nb;
    const int tr = r0 + blockIdx.y * 64, tc = r0 + blockIdx.x * 64;
    const int t = threadIdx.x;
    for (int e = t; e < 64 * NB; e += 256) {
        const int r = e / NB, q = e % NB;
        Ls[r][q] = (tr + r < n && q < nb) ? A[(int64_t)(tr + r) * ld + k0 + q] : 0.0;
        const int qq = e / 64, c = e % 64;
        Us[qq][c] = (tc + c < ncols && qq < nb) ? A[(int64_t)(k0 + qq) * ld + tc + c] : 0.0;
    }
    __syncthreads();
    const int ty = t / 16, tx = t % 16;
    double acc[4][4] = {};
    for (int q = 0; q < nb; ++q) {
        double l[4], u[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) l[a] = Ls[ty * 4 + a][q];
#pragma unroll
        for (int b = 0; b < 4; ++b) u[b] = Us[q][tx + 16 * b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += l[a] * u[b];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int r = tr + ty * 4 + a;
        if (r >= n) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int c = tc + tx + 16 * b;
            if (c < ncols) A[(int64_t)r * ld + c] -= acc[a][b];
        }
    }
}

// ---------------------------------------------------------------- Cholesky of the leading nr x nr block

constexpr int CB = 64;

// A double from one lane of the wave (lane uniform).
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

constexpr int PB = 8;   // pivots per step of the diagonal block

// 1/sqrt(d) for d > 0: the hardware estimate (v_rsq_f64, about 2^-23 relative) and two Newton
// steps y <- y (3 - d y^2) / 2, each of which squares the relative error, so the result is within a
// few ulp.  It replaces a correctly rounded sqrt and a division on the serial pivot chain.
__device__ __forceinline__ double rsqrt_newton(double d) {
    double y = __builtin_amdgcn_rsq(d);
    const double h = 0.5 * d;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

// Diagonal block of the blocked Cholesky (64 x 64, 256 threads): factor L, its inverse X = L^-1
// (written to Dinv, row-major, identity past the block) and the forward-substitution step of the two
// right-hand sides, z_k = X r_k (rows k0.. of rw, rb, in place).
// The block lives in registers: thread (tr, tc) holds rows 4 tr.. and columns 4 tc.. of the Schur
// complement S and of W, the right-looking inverse (W = I at the start).  Step J takes the 8 pivots
// j = 8J .. 8J+7:
//   A (wave 0): from the 8 columns j.. of S and the 8 rows j.. of W (one lane per row, resp. per
//     column), L[r][j..j+7] (the pivots one after another, the pivot rows broadcast by readlane) and
//     X[j..j+7][c] = L_jj^-1 W[j..j+7][c], into LDS;
//   B (the owners): the rank-8 update S -= L L^T, W -= L X of their blocks below and right of the
//     pivot block, and the finished L and X entries into place.
// Lookahead: wave 0 does not wait for B of step J to start A of step J + 1.  The owners publish the
// columns and rows of step J + 2 as they stand after their update of step J (two steps ahead), and
// wave 0 brings those of step J + 1 up to date itself with the rank-8 update of step J, so A(J + 1)
// runs while the owners are still in B(J): one barrier per step.  (Wave 0 owns blocks of the trailing
// part only in steps 0 and 1; it does those updates after A.)  The panels, W rows, L and X go through
// double buffers by step parity.  Ownership is aligned with the 8-pivot blocks (two 4-wide thread
// blocks), so no entry is masked.  A pivot at or below 1e-13 of the original diagonal entry counts as
// a failure: the system is then left to the pivoted LU, which reports exact singularity as the
// reference's LAPACK call does.
struct PotrfLds {
    union {
        struct {
            double colS[2][CB][PB + 1];   // S[r][j + q] of a step, as published
            double xr[2][PB][CB];         // W[j + q][c] of a step, as published
            double lbT[2][PB][CB];        // L[r][j + q], transposed
            double xb[2][PB][CB];         // X[j + q][c]
        } loop;
        double Xs[CB][CB + 1];            // the inverse, for the outputs
    } u;
    double od[CB];              // original diagonal entries for the pivot test
    double zw[CB], zb[CB];
};

__device__ __forceinline__ void potrf_block(double *__restrict__ A, int ld, int n, int k0, const double *__restrict__ orig,
                                            int *__restrict__ info, double *__restrict__ Dinv, double *__restrict__ rw,
                                            double *__restrict__ rb, PotrfLds &sm) {
    auto &colS = sm.u.loop.colS;
    auto &xr = sm.u.loop.xr;
    auto &lbT = sm.u.loop.lbT;
    auto &xb = sm.u.loop.xb;
    auto &od = sm.od;
    auto &Xs = sm.u.Xs;
    auto &zw = sm.zw;
    auto &zb = sm.zb;
    const int nb = min(CB, n - k0);
    const int t = threadIdx.x, tr = t >> 4, tc = t & 15, lane = t & 63, wave = t >> 6;
    const bool lower = tc <= tr;
    double a[4][4], x[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 4 * tr + u, c = 4 * tc + v;
            a[u][v] = (i < nb && c <= i) ? A[(int64_t)(k0 + i) * ld + k0 + c] : (i == c ? 1.0 : 0.0);
            x[u][v] = i == c ? 1.0 : 0.0;
        }
    if (t < CB) {
        zw[t] = t < nb ? rw[k0 + t] : 0.0;
        zb[t] = t < nb ? rb[k0 + t] : 0.0;
        od[t] = t < nb ? orig[(int64_t)(k0 + t) * ld + k0 + t] : 1.0;
    }
    // W is lower triangular: row entries right of the owners' blocks stay zero
    for (int e = t; e < 2 * PB * CB; e += 256) (&xr[0][0][0])[e] = 0.0;
    __syncthreads();
    // columns and W rows of steps 0 and 1 (S and W as loaded)
    if (lower && tc < 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) colS[tc >> 1][4 * tr + u][4 * (tc & 1) + v] = a[u][v];
    }
    if (lower && tr < 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) xr[tr >> 1][4 * (tr & 1) + u][4 * tc + v] = x[u][v];
    }
    __syncthreads();
    int bad = 0;
    // wave 0's panel of the next step: c[q] = S[lane][j + q], w[q] = W[j + q][lane]
    double c[PB], w[PB];
    if (wave == 0) {
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            c[q] = colS[0][lane][q];
            w[q] = xr[0][q][lane];
        }
    }
    // ---- phase A of step J from c, w (wave 0): L and X of the step into buffer J & 1
    auto phase_a = [&](int J) {
        const int j = J * PB, bf = J & 1;
        double l[PB], Ljj[PB][PB], inv[PB];
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const double d = readlane_f64(c[q], j + q);
            const bool live = j + q < nb;
            const bool ok = d > 1e-13 * fabs(od[j + q]);
            bad = (live && !ok && !bad) ? k0 + j + q + 1 : bad;   // a select: no branch in the chain
            const double dd = live && ok ? d : 1.0;   // dead or failed pivots stay 1
            const double ip = rsqrt_newton(dd);
            const double piv = dd * ip;
            inv[q] = ip;
            l[q] = lane > j + q ? c[q] * ip : (lane == j + q ? piv : 0.0);
#pragma unroll
            for (int p = q + 1; p < PB; ++p) {
                Ljj[p][q] = readlane_f64(l[q], j + p);
                c[p] = fma(-l[q], Ljj[p][q], c[p]);
            }
        }
#pragma unroll
        for (int q = 0; q < PB; ++q) lbT[bf][q][lane] = l[q];
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const double xv = w[q] * inv[q];
            xb[bf][q][lane] = xv;
#pragma unroll
            for (int p = q + 1; p < PB; ++p) w[p] = fma(-Ljj[p][q], xv, w[p]);
        }
    };
    if (wave == 0) phase_a(0);
    for (int J = 0; J < CB / PB; ++J) {
        const int j = J * PB, bf = J & 1;
        __syncthreads();   // L, X of step J and the step-(J+1) publications are in LDS
        // ---- wave 0: step J + 1's panel brought up to date with step J's rank-8 update, then its A
        if (wave == 0 && J + 1 < CB / PB) {
            const int jn = j + PB, bn = (J + 1) & 1;
            double lr[PB];
#pragma unroll
            for (int p = 0; p < PB; ++p) lr[p] = lbT[bf][p][lane];
#pragma unroll
            for (int q = 0; q < PB; ++q) {
                double cv = colS[bn][lane][q], wv = xr[bn][q][lane];
#pragma unroll
                for (int p = 0; p < PB; ++p) {
                    cv = fma(-lr[p], lbT[bf][p][jn + q], cv);
                    wv = fma(-lbT[bf][p][jn + q], xb[bf][p][lane], wv);
                }
                c[q] = cv;
                w[q] = wv;
            }
            phase_a(J + 1);
        }
        // ---- phase B of step J (owners)
        if (lower && tr >= 2 * J) {
            if (tr >= 2 * J + 2) {
                // rows below the pivot block: W -= L X, S -= L L^T right of it, L columns into place
                const bool upd = tc >= 2 * J + 2, put = tc >= 2 * J && tc < 2 * J + 2;
#pragma unroll
                for (int q = 0; q < PB; ++q) {
                    const double2 l01 = *reinterpret_cast<const double2 *>(&lbT[bf][q][4 * tr]);
                    const double2 l23 = *reinterpret_cast<const double2 *>(&lbT[bf][q][4 * tr + 2]);
                    const double lr[4] = {l01.x, l01.y, l23.x, l23.y};
                    const double2 x01 = *reinterpret_cast<const double2 *>(&xb[bf][q][4 * tc]);
                    const double2 x23 = *reinterpret_cast<const double2 *>(&xb[bf][q][4 * tc + 2]);
                    const double xj[4] = {x01.x, x01.y, x23.x, x23.y};
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) x[u][v] = fma(-lr[u], xj[v], x[u][v]);
                    if (upd) {
                        const double2 c01 = *reinterpret_cast<const double2 *>(&lbT[bf][q][4 * tc]);
                        const double2 c23 = *reinterpret_cast<const double2 *>(&lbT[bf][q][4 * tc + 2]);
                        const double lc[4] = {c01.x, c01.y, c23.x, c23.y};
#pragma unroll
                        for (int u = 0; u < 4; ++u)
#pragma unroll
                            for (int v = 0; v < 4; ++v) a[u][v] = fma(-lr[u], lc[v], a[u][v]);
                    }
                }
                if (put) {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) a[u][v] = lbT[bf][4 * tc + v - j][4 * tr + u];
                }
            } else {
                // rows of the pivot block: their X rows and the pivot block of L
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = 4 * tr + u, cc = 4 * tc + v;
                        x[u][v] = xb[bf][r - j][cc];
                        if (tc >= 2 * J && cc <= r) a[u][v] = lbT[bf][cc - j][r];
                    }
            }
            // publish the columns of S and the rows of W of step J + 2, as they stand now
            if (J + 2 < CB / PB) {
                const int j2 = j + 2 * PB;
                if (tc >= 2 * J + 4 && tc < 2 * J + 6 && tr >= 2 * J + 4) {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) colS[bf][4 * tr + u][4 * tc + v - j2] = a[u][v];
                }
                if (tr >= 2 * J + 4 && tr < 2 * J + 6) {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) xr[bf][4 * tr + u - j2][4 * tc + v] = x[u][v];
                }
            }
        }
    }
    __syncthreads();   // every owner is past its last update; the LDS is reused for the inverse
    // ---- outputs: factor, inverse, z_k = X r_k
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 4 * tr + u, c = 4 * tc + v;
            const double xv = lower && c <= i ? x[u][v] : 0.0;
            Xs[i][c] = xv;
            Dinv[i * CB + c] = xv;
            if (lower && i < nb && c <= i) A[(int64_t)(k0 + i) * ld + k0 + c] = a[u][v];
        }
    __syncthreads();
    {
        // z = X r for both right-hand sides: four lanes per row, 16 terms each (X is lower
        // triangular, zero above the diagonal)
        const int row = t >> 2, part = t & 3;
        double sa = 0.0, sb = 0.0;
#pragma unroll
        for (int c = part * 16; c < part * 16 + 16; ++c) {
            sa = fma(Xs[row][c], zw[c], sa);
            sb = fma(Xs[row][c], zb[c], sb);
        }
        sa += __shfl_xor(sa, 1);
        sb += __shfl_xor(sb, 1);
        sa += __shfl_xor(sa, 2);
        sb += __shfl_xor(sb, 2);
        if (part == 0 && row < nb) {
            rw[k0 + row] = sa;
            rb[k0 + row] = sb;
        }
    }
    if (t == 0 && bad && info[0] == 0) info[0] = bad;
}

__global__ __launch_bounds__(256) void k_potrf_inv(double *__restrict__ A, int ld, int n, int k0,
                                                   const double *__restrict__ orig, int *__restrict__ info,
                                                   double *__restrict__ Dinv, double *__restrict__ rw,
                                                   double *__restrict__ rb) {
    __shared__ PotrfLds sm;
    potrf_block(A, ld, n, k0, orig, info, Dinv, rw, rb, sm);
}

// Panel below the diagonal block: L21 = A21 L11^{-T} = A21 Dinv^T, 32 rows per workgroup, and the
// forward-substitution update of those rows of the right-hand sides, r_i -= L21[i] . z_k.
// L21 goes in place (the trailing update reads it there) and, transposed, into the upper triangle
// (A[k0 + c][r] = L21[r][c]), where the backward substitution reads it.
__global__ __launch_bounds__(256) void k_chol_trsm_inv(double *__restrict__ A, int ld, int n, int k0,
                                                       const double *__restrict__ Dinv, double *__restrict__ rw,
                                                       double *__restrict__ rb) {
    __shared__ double Xs[CB][CB + 1];
    __shared__ double As[32][CB + 1];
    __shared__ double z[2][CB];
    const int nb = min(CB, n - k0);
    const int r0 = k0 + nb + blockIdx.x * 32;
    const int t = threadIdx.x;
    #pragma unroll
    for (int e = t; e < CB * CB; e += 256) Xs[e / CB][e % CB] = Dinv[e];
    #pragma unroll
    for (int e = t; e < 32 * CB; e += 256) {
        const int i = e / CB, c = e % CB;
        As[i][c] = (r0 + i < n && c < nb) ? A[(int64_t)(r0 + i) * ld + k0 + c] : 0.0;
    }
    if (t < CB) {
        z[0][t] = t < nb ? rw[k0 + t] : 0.0;
        z[1][t] = t < nb ? rb[k0 + t] : 0.0;
    }
    __syncthreads();
    const int ty = t / 16, tx = t % 16;
    double acc[2][4] = {};
    for (int q = 0; q < CB; ++q) {
        double l[2], u[4];
#pragma unroll
        for (int a = 0; a < 2; ++a) l[a] = As[ty * 2 + a][q];
#pragma unroll
        for (int b = 0; b < 4; ++b) u[b] = Xs[tx + 16 * b][q];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += l[a] * u[b];
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int r = r0 + ty * 2 + a;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int c = tx + 16 * b;
            As[ty * 2 + a][c] = c < nb ? acc[a][b] : 0.0;
            if (r < n && c < nb) A[(int64_t)r * ld + k0 + c] = acc[a][b];
        }
    }
    __syncthreads();
    for (int e = t; e < 32 * CB; e += 256) {
        const int c = e / 32, i = e % 32;
        if (r0 + i < n && c < nb) A[(int64_t)(k0 + c) * ld + r0 + i] = As[i][c];
    }
    if (t < 64) {
        const int i = t & 31, which = t >> 5;
        const int r = r0 + i;
        if (r < n) {
            double s = 0.0;
            for (int c = 0; c < nb; ++c) s += As[i][c] * z[which][c];
            double *v = which ? rb : rw;
            v[r] -= s;
        }
    }
}

// A22 -= L21 L21^T on the lower triangle (64x64 tiles with tile row >= tile col).
struct UpdateLds {
    double Ls[64][CB + 1];
    double Us[CB][64];
};

// Trailing update of one 64 x 64 tile (ti, tj) below the panel of block column k0.
__device__ __forceinline__ void update_tile(double *__restrict__ A, int ld, int n, int k0, int ti, int tj, UpdateLds &sm) {
    const int nb = min(CB, n - k0);
    const int r0 = k0 + nb;
    auto &Ls = sm.Ls;
    auto &Us = sm.Us;
    const int tr = r0 + ti * 64, tc = r0 + tj * 64;
    const int t = threadIdx.x;
    #pragma unroll
    for (int e = t; e < 64 * CB; e += 256) {
        const int r = e / CB, q = e % CB;
        Ls[r][q] = (tr + r < n && q < nb) ? A[(int64_t)(tr + r) * ld + k0 + q] : 0.0;
        Us[q][r] = (tc + r < n && q < nb) ? A[(int64_t)(tc + r) * ld + k0 + q] : 0.0;
    }
    __syncthreads();
    const int ty = t / 16, tx = t % 16;
    double acc[4][4] = {};
    for (int q = 0; q < nb; ++q) {
        double l[4], u[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) l[a] = Ls[ty * 4 + a][q];
#pragma unroll
        for (int b = 0; b < 4; ++b) u[b] = Us[q][tx + 16 * b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] += l[a] * u[b];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int r = tr + ty * 4 + a;
        if (r >= n) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int c = tc + tx + 16 * b;
            if (c < n && c <= r) A[(int64_t)r * ld + c] -= acc[a][b];
        }
    }
}

// Trailing update after block column k0; the work-group of tile (0, 0), which is the next diagonal
// block, then factors that block (potrf_block), so the serial factorisation of block k0 + 64 runs
// while the other tiles of this update are still in flight.
__global__ __launch_bounds__(256) void k_chol_update_potrf(double *__restrict__ A, int ld, int n, int k0,
                                                           const double *__restrict__ orig, int *__restrict__ info,
                                                           double *__restrict__ Dinv_next, double *__restrict__ rw,
                                                           double *__restrict__ rb) {
    __shared__ union {
        UpdateLds up;
        PotrfLds pf;
    } sm;
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (tj > ti) return;
    update_tile(A, ld, n, k0, ti, tj, sm.up);
    if (ti == 0 && tj == 0) {
        __syncthreads();   // the tile's update is in A; the LDS is reused
        potrf_block(A, ld, n, k0 + CB, orig, info, Dinv_next, rw, rb, sm.pf);
    }
}

// One step of the blocked Cholesky after diagonal block k0 was factored (its inverse X in Dinv),
// with the panel fused into the trailing update.  Work-group (ti, tj), tj <= ti, owns tile (ti, tj)
// of the trailing lower triangle:
//   1. L_i = A21_i X^T and L_j = A21_j X^T, the panel rows of tiles ti and tj, from the lower-triangle
//      panel A21, which every work-group of the step reads and none writes;
//   2. A22_ij -= L_i L_j^T on the lower triangle;
//   3. a diagonal work-group (ti == tj) stores its L rows transposed in the upper triangle,
//      A[k0 + c][r] = L[r][c], and applies the forward substitution r_i -= L_i z_k to its rows of
//      the right-hand sides (z_k: rows k0.. of rw, rb, left there by the block's factorisation);
//   4. work-group (0, 0) then factors the next diagonal block.
// Each panel row block is computed by every work-group of its tile row and column.  That costs
// nothing while the grid is at most one work-group per CU, which is when the host picks this kernel;
// it saves the separate panel launch on the serial chain of diagonal blocks.
struct StepLds {
    double XT[CB][CB + 1];    // X^T: XT[q][c] = X[c][q]
    double LiT[CB][CB + 1];   // A21 rows of tile ti, then their L rows, transposed: LiT[q][r]
    double LjT[CB][CB + 1];
    double z[2][CB];
};
constexpr size_t STEP_LDS = sizeof(StepLds) > sizeof(PotrfLds) ? sizeof(StepLds) : sizeof(PotrfLds);

// Both 64 x 64 x 64 products run on fp64 MFMA (16 x 16 x 4): wave w owns the 32 x 32 quadrant
// (w / 2, w % 2) as 2 x 2 tiles; the operands sit k-major in LDS, so each k group of 16 lanes reads
// 128 contiguous bytes.
__global__ __launch_bounds__(256) void k_chol_step(double *__restrict__ A, int ld, int n, int k0,
                                                   const double *__restrict__ Dinv, const double *__restrict__ orig,
                                                   int *__restrict__ info, double *__restrict__ Dinv_next,
                                                   double *__restrict__ rw, double *__restrict__ rb) {
    extern __shared__ __align__(16) char step_smem[];
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (tj > ti) return;
    StepLds &sm = *reinterpret_cast<StepLds *>(step_smem);
    const int nb = min(CB, n - k0);
    const int r0 = k0 + nb;
    const int ri = r0 + ti * 64, rj = r0 + tj * 64;
    const bool diag = ti == tj;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wr = wave >> 1, wc = wave & 1, lr = lane & 15, lk = lane >> 4;
    // every global read of the step is issued up front, into registers: the block inverse, the panel
    // rows of both tiles and this lane's entries of the tile to update (no other work-group writes
    // them in this step), then the LDS is filled k-major
    double xv[16], av[16], bv[16], a22[2][2][4];
#pragma unroll
    for (int i = 0; i < 16; ++i) xv[i] = Dinv[t + 256 * i];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int e = t + 256 * i, r = e / CB, q = e % CB;
        av[i] = (ri + r < n && q < nb) ? A[(int64_t)(ri + r) * ld + k0 + q] : 0.0;
        bv[i] = (!diag && rj + r < n && q < nb) ? A[(int64_t)(rj + r) * ld + k0 + q] : 0.0;
    }
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int nj = 0; nj < 2; ++nj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = ri + wr * 32 + mi * 16 + F64x16::d_row(lane, v);
                const int c = rj + wc * 32 + nj * 16 + F64x16::d_col(lane);
                a22[mi][nj][v] = (r < n && c < n && c <= r) ? A[(int64_t)r * ld + c] : 0.0;
            }
    double zv = 0.0;
    if (diag && t < 2 * CB) zv = (t & 63) < nb ? ((t < CB) ? rw : rb)[k0 + (t & 63)] : 0.0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int e = t + 256 * i, r = e / CB, q = e % CB;
        sm.XT[q][r] = xv[i];
        sm.LiT[q][r] = av[i];
        if (!diag) sm.LjT[q][r] = bv[i];
    }
    if (diag && t < 2 * CB) sm.z[t >> 6][t & 63] = zv;
    __syncthreads();
    // 1. panel rows: L[r][c] = sum_q A21[r][q] X[c][q]
    dvec4 ci[2][2], cj[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int nj = 0; nj < 2; ++nj) {
            ci[mi][nj] = dvec4{0.0, 0.0, 0.0, 0.0};
            cj[mi][nj] = dvec4{0.0, 0.0, 0.0, 0.0};
        }
    for (int q0 = 0; q0 < CB; q0 += 4) {
        const int q = q0 + lk;
        const double b0 = sm.XT[q][wc * 32 + lr], b1 = sm.XT[q][wc * 32 + 16 + lr];
#pragma unroll
        for (int mi = 0; mi < 2; ++mi) {
            const double a = sm.LiT[q][wr * 32 + mi * 16 + lr];
            ci[mi][0] = F64x16::mma(a, b0, ci[mi][0]);
            ci[mi][1] = F64x16::mma(a, b1, ci[mi][1]);
        }
        if (!diag) {
#pragma unroll
            for (int mi = 0; mi < 2; ++mi) {
                const double a = sm.LjT[q][wr * 32 + mi * 16 + lr];
                cj[mi][0] = F64x16::mma(a, b0, cj[mi][0]);
                cj[mi][1] = F64x16::mma(a, b1, cj[mi][1]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int nj = 0; nj < 2; ++nj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = wr * 32 + mi * 16 + F64x16::d_row(lane, v), c = wc * 32 + nj * 16 + F64x16::d_col(lane);
                sm.LiT[c][r] = c < nb ? ci[mi][nj][v] : 0.0;
                if (!diag) sm.LjT[c][r] = c < nb ? cj[mi][nj][v] : 0.0;
            }
    __syncthreads();
    // 2. trailing update of the tile: A22[r][c] -= sum_q L_i[r][q] L_j[c][q]
    double (*Lc)[CB + 1] = diag ? sm.LiT : sm.LjT;
    dvec4 cu[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int nj = 0; nj < 2; ++nj) cu[mi][nj] = dvec4{0.0, 0.0, 0.0, 0.0};
    for (int q0 = 0; q0 < CB; q0 += 4) {
        const int q = q0 + lk;
        const double b0 = Lc[q][wc * 32 + lr], b1 = Lc[q][wc * 32 + 16 + lr];
#pragma unroll
        for (int mi = 0; mi < 2; ++mi) {
            const double a = sm.LiT[q][wr * 32 + mi * 16 + lr];
            cu[mi][0] = F64x16::mma(a, b0, cu[mi][0]);
            cu[mi][1] = F64x16::mma(a, b1, cu[mi][1]);
        }
    }
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int nj = 0; nj < 2; ++nj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = ri + wr * 32 + mi * 16 + F64x16::d_row(lane, v);
                const int c = rj + wc * 32 + nj * 16 + F64x16::d_col(lane);
                if (r < n && c < n && c <= r) A[(int64_t)r * ld + c] = a22[mi][nj][v] - cu[mi][nj][v];
            }
    if (!diag) return;
    // 3. the forward substitution of the tile row's right-hand-side entries, r_i -= L_i z_k: four
    // lanes per dot product, 16 terms each
    {
        const int row = t >> 2, part = t & 3;
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            double s = 0.0;
#pragma unroll
            for (int c = part * 16; c < part * 16 + 16; ++c) s = fma(sm.LiT[c][row], sm.z[w][c], s);
            s += __shfl_xor(s, 1);
            s += __shfl_xor(s, 2);
            if (part == 0 && ri + row < n) (w ? rb : rw)[ri + row] -= s;
        }
    }
    // 4. the next diagonal block (the rows it reads are in memory after the barrier); the panel,
    // transposed, goes out afterwards: only the backward substitution reads it
    if (ti == 0) {
        __syncthreads();
        // the LDS is reused by the factorisation: keep this tile's panel rows in registers
        double lt[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int e = t + 256 * i, c = e / 64, r = e % 64;
            lt[i] = sm.LiT[c][r];
        }
        __syncthreads();
        potrf_block(A, ld, n, k0 + CB, orig, info, Dinv_next, rw, rb, *reinterpret_cast<PotrfLds *>(step_smem));
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int e = t + 256 * i, c = e / 64, r = e % 64;
            if (ri + r < n && c < nb) A[(int64_t)(k0 + c) * ld + ri + r] = lt[i];
        }
        return;
    }
    for (int e = t; e < 64 * CB; e += 256) {
        const int c = e / 64, r = e % 64;
        if (ri + r < n && c < nb) A[(int64_t)(k0 + c) * ld + ri + r] = sm.LiT[c][r];
    }
}

// After the factorisation (which leaves z_w = L^-1 w, z_b = L^-1 b in zw, zb): the Lagrange
// multiplier lambda = (w.A^-1 b - q_tot) / (w.A^-1 w) = (z_w.z_b - q_tot) / (z_w.z_w), and the
// backward substitution L^T y = u with u = z_b - lambda z_w; q = y, x[n] = lambda.
//
// One launch, one work-group per diagonal block.  A work-group takes a ticket when it starts; ticket
// i owns block kb = nblk - 1 - i, so every block it needs (the ones after kb) belongs to a work-group
// that started before it, and each of those depends only on earlier starters again: the chain always
// progresses, whatever the dispatch order.  Work-group kb
//   - forms lambda itself (two dot products of length n) and u_kb;
//   - for each later block k, oldest first, loads its part of the transposed panel (the upper-triangle
//     block in the rows of block kb and the columns of block k, contiguous rows) before it waits for
//     block k's flag, then subtracts U_{kb,k} y_k;
//   - forms y_kb = D_kb^T u_kb (D_kb = L_kb,kb^-1) and publishes it: the values, a release fence at
//     agent scope, then the flag set to this solve's epoch.
// A wait that sees no progress for spin_limit polls records the failure in *status and the
// work-group leaves, so the kernel always ends; the host then solves the system by the LU.
constexpr unsigned BACKSUB_SPIN = 1u << 22;   // ~0.3 s of polling; RP_BACKSUB_SPIN overrides it

__global__ __launch_bounds__(256) void k_backsub(const double *__restrict__ A, int ld, int n,
                                                 const double *__restrict__ Dinv, const double *__restrict__ zw,
                                                 const double *__restrict__ zb, double q_tot, double *__restrict__ x,
                                                 unsigned *__restrict__ sync, unsigned base, unsigned epoch,
                                                 unsigned spin_limit, int *__restrict__ status,
                                                 const int *__restrict__ cinfo,
                                                 const double *__restrict__ q_tot_dev = nullptr,
                                                 const int *__restrict__ go = nullptr) {
    __shared__ double Dk[CB][CB + 1];
    __shared__ double uk[CB], yk[CB];
    __shared__ double red[2][4];
    __shared__ unsigned sh_ticket;
    __shared__ int sh_abort;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    unsigned *counter = sync, *flags = sync + 1;
    if (t == 0) {
        sh_ticket = atomicAdd(counter, 1u) - base;
        sh_abort = 0;
    }
    __syncthreads();
    if (*cinfo != 0) return;   // a pivot failed: the host reruns the system through the LU
    if (go && (*go == 0 || *status != 0)) return;   // a refinement step that is not needed
    if (q_tot_dev) q_tot = *q_tot_dev;
    const int nblk = (n + CB - 1) / CB;
    if (sh_ticket >= (unsigned)nblk) {   // tickets out of step with the host's count: give up, LU answers
        if (t == 0) atomicExch(status, 1);
        return;
    }
    const int kb = nblk - 1 - (int)sh_ticket;
    const int j0 = kb * CB, nb = min(CB, n - j0);
    for (int e = t; e < CB * CB; e += 256) Dk[e / CB][e % CB] = Dinv[(int64_t)kb * CB * CB + e];
    // lambda
    double pw = 0.0, pb = 0.0;
    for (int r = t; r < n; r += 256) {
        pw = fma(zw[r], zw[r], pw);
        pb = fma(zw[r], zb[r], pb);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        pw += __shfl_xor(pw, o);
        pb += __shfl_xor(pb, o);
    }
    if (lane == 0) {
        red[0][wave] = pw;
        red[1][wave] = pb;
    }
    __syncthreads();
    const double lambda = (((red[1][0] + red[1][1]) + (red[1][2] + red[1][3])) - q_tot) /
                          ((red[0][0] + red[0][1]) + (red[0][2] + red[0][3]));
    if (t < CB) uk[t] = t < nb ? zb[j0 + t] - lambda * zw[j0 + t] : 0.0;
    if (kb == nblk - 1 && t == 0) x[n] = lambda;
    // contributions of the later blocks: thread (row c = t / 4, columns q .. q + 15 of block k)
    const int c = t >> 2, q = (t & 3) * 16;
    const bool crow = c < nb;
    for (int k = nblk - 1; k > kb; --k) {
        const int k0 = k * CB, kn = min(CB, n - k0);
        double ub[16];
        const double *row = A + (int64_t)(j0 + c) * ld + k0 + q;
#pragma unroll
        for (int i = 0; i < 16; ++i) ub[i] = crow && q + i < kn ? row[i] : 0.0;
        if (t == 0) {
            // at most spin_limit polls (none when it is 0); the polls are relaxed and one acquire
            // fence follows the flag, so the wait does not invalidate the cache on every poll
            for (unsigned spins = 0;; ++spins) {
                if (spins >= spin_limit) {
                    sh_abort = 1;
                    atomicExch(status, 1);
                    break;
                }
                if (__hip_atomic_load(flags + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __syncthreads();
        if (sh_abort) return;
        if (t < CB) yk[t] = t < kn ? __hip_atomic_load(x + k0 + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
        __syncthreads();
        double sacc = 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) sacc = fma(ub[i], yk[q + i], sacc);
        sacc += __shfl_xor(sacc, 1);
        sacc += __shfl_xor(sacc, 2);
        if ((t & 3) == 0 && crow) uk[c] -= sacc;
        __syncthreads();
    }
    // y_kb and its flag
    __syncthreads();
    if (t < nb) {
        double y = 0.0;
        for (int i = t; i < nb; ++i) y = fma(Dk[i][t], uk[i], y);
        __hip_atomic_store(x + j0 + t, y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
    __syncthreads();
    if (t == 0) __hip_atomic_store(flags + kb, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_fill_u32(unsigned *p, int n, unsigned v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- iterative refinement (split form)
//
// The split Gram is the Gram of a perturbed basis (int8 slices of Phi).  A refinement step forms the
// residual of the fp64 normal equations at the current solution x (reduced charges, multiplier x[nr])
// from the fp64 basis kept by the accumulate:
//   u = P x (expanded with the merge multiplicities),  e = V - Phi u over the kept points,
//   r = Phi_c^T (V - Phi_c u) = Phi^T e - phibar sum(e)      (Phi_c the centred basis, phibar its mean),
//   r1 = P^T r + rhs_add - diag_add x - w lambda,  r2 = q_tot - w . x,
// then solves the bordered system of the split Gram for the correction with the same factor: a forward
// substitution z_r = L^-1 r1 and k_backsub with (z_w, z_r, r2), which forms the multiplier's correction
// (z_w . z_r - r2) / (z_w . z_w) and the charges' L^-T (z_r - dlambda z_w).
//
// State (RefineState, device): go = the step runs; a step runs while the previous correction exceeded
// the stop threshold, the pivots were positive and no wait gave up.  Every kernel of a skipped step
// returns at once.
__device__ __forceinline__ bool refine_active(const int *cinfo, const RefineState *rs) {
    return cinfo[0] == 0 && cinfo[1] == 0 && rs->go != 0;
}

// u[b] = sum over the merge sets holding atom b of multiplicity * x[set]
__global__ __launch_bounds__(256) void k_ref_expand(const double *__restrict__ x, const int *__restrict__ inv_ptr,
                                                    const int *__restrict__ inv_set, const double *__restrict__ inv_mult,
                                                    int N, double *__restrict__ u, const int *__restrict__ cinfo,
                                                    const RefineState *__restrict__ rs) {
    if (!refine_active(cinfo, rs)) return;
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= N) return;
    double s = 0.0;
    for (int p = inv_ptr[b]; p < inv_ptr[b + 1]; ++p) s = fma(inv_mult[p], x[inv_set[p]], s);
    u[b] = s;
}

// Block sum of one value per thread (256 threads), returned to every thread.
__device__ __forceinline__ double block_sum_256(double v, double *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// e = V - Phi u over the kept points, two rows per thread (the basis chunk c holds global rows
// [c rows, (c + 1) rows), column-major [NP][rows]); per-block sums of e into esum.
__global__ __launch_bounds__(256) void k_ref_rows(const double *__restrict__ phi, int64_t rows, int64_t plane, int N,
                                                  int64_t M, const double *__restrict__ v, const double *__restrict__ u,
                                                  double *__restrict__ e, double *__restrict__ esum,
                                                  const int *__restrict__ cinfo, const RefineState *__restrict__ rs) {
    __shared__ double red[4];
    if (!refine_active(cinfo, rs)) return;
    const int64_t R = 2 * ((int64_t)blockIdx.x * 256 + threadIdx.x);
    double e0 = 0.0, e1 = 0.0;
    if (R < M) {
        const int64_t c = R / rows, r = R - c * rows;   // rows is even: R and R + 1 share the chunk
        const double *col = phi + c * plane + r;
        double a0[4] = {0.0, 0.0, 0.0, 0.0}, a1[4] = {0.0, 0.0, 0.0, 0.0};
        int j = 0;
        for (; j + 4 <= N; j += 4) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const double2 p = *reinterpret_cast<const double2 *>(col + (int64_t)(j + t) * rows);
                a0[t] = fma(p.x, u[j + t], a0[t]);
                a1[t] = fma(p.y, u[j + t], a1[t]);
            }
        }
        for (; j < N; ++j) {
            const double2 p = *reinterpret_cast<const double2 *>(col + (int64_t)j * rows);
            a0[0] = fma(p.x, u[j], a0[0]);
            a1[0] = fma(p.y, u[j], a1[0]);
        }
        e0 = v[R] - ((a0[0] + a0[1]) + (a0[2] + a0[3]));
        e[R] = e0;
        if (R + 1 < M) {   // the second row of the last pair may lie past the kept points
            e1 = v[R + 1] - ((a1[0] + a1[1]) + (a1[2] + a1[3]));
            e[R + 1] = e1;
        }
    }
    const double sum = block_sum_256(e0 + e1, red);
    if (threadIdx.x == 0) esum[blockIdx.x] = sum;
}

// One work-group per atom column j: g_j = sum_R Phi[j][R] e_R, s_j = sum_R Phi[j][R] over the kept
// points, and r_j = g_j - s_j E / M with E = sum(e) (the block sums of k_ref_rows, in order).
__global__ __launch_bounds__(256) void k_ref_cols(const double *__restrict__ phi, int64_t rows, int64_t plane,
                                                  int64_t M, const double *__restrict__ e,
                                                  const double *__restrict__ esum, int n_esum, double *__restrict__ r,
                                                  const int *__restrict__ cinfo, const RefineState *__restrict__ rs) {
    __shared__ double red[4];
    if (!refine_active(cinfo, rs)) return;
    const int j = blockIdx.x, t = threadIdx.x;
    double g0 = 0.0, g1 = 0.0, s0 = 0.0, s1 = 0.0;
    for (int64_t c0 = 0, c = 0; c0 < M; c0 += rows, ++c) {
        const int64_t valid = min(rows, M - c0);
        const double *col = phi + c * plane + (int64_t)j * rows;
        const double *ec = e + c0;
        int64_t q = 2 * t;
        for (; q + 1 < valid; q += 512) {
            const double2 p = *reinterpret_cast<const double2 *>(col + q);
            g0 = fma(p.x, ec[q], g0);
            g1 = fma(p.y, ec[q + 1], g1);
            s0 += p.x;
            s1 += p.y;
        }
        if (q < valid) {   // an odd row count leaves one row
            g0 = fma(col[q], ec[q], g0);
            s0 += col[q];
        }
    }
    double E = 0.0;
    for (int i = t; i < n_esum; i += 256) E += esum[i];
    const double G = block_sum_256(g0 + g1, red);
    const double S = block_sum_256(s0 + s1, red);
    const double Et = block_sum_256(E, red);
    if (t == 0) r[j] = G - S * (Et / (double)M);
}

// k_ref_cols over one chunk of the basis at a time (the streamed refinement, which recomputes the chunks):
// each thread's four partial sums carry over between chunks in acc ([N][256][4]), so the sums run in
// k_ref_cols's order and the residual is bitwise the same.  The last chunk finishes r_j.
__global__ __launch_bounds__(256) void k_ref_cols_chunk(const double *__restrict__ col0, int64_t rows, int64_t valid,
                                                        const double *__restrict__ ec, double *__restrict__ acc,
                                                        int first, int last, const double *__restrict__ esum,
                                                        int n_esum, int64_t M, double *__restrict__ r,
                                                        const int *__restrict__ cinfo,
                                                        const RefineState *__restrict__ rs) {
    __shared__ double red[4];
    if (!refine_active(cinfo, rs)) return;
    const int j = blockIdx.x, t = threadIdx.x;
    double *a = acc + ((int64_t)j * 256 + t) * 4;
    double g0 = 0.0, g1 = 0.0, s0 = 0.0, s1 = 0.0;
    if (!first) {
        g0 = a[0];
        g1 = a[1];
        s0 = a[2];
        s1 = a[3];
    }
    const double *col = col0 + (int64_t)j * rows;
    int64_t q = 2 * t;
    for (; q + 1 < valid; q += 512) {
        const double2 p = *reinterpret_cast<const double2 *>(col + q);
        g0 = fma(p.x, ec[q], g0);
        g1 = fma(p.y, ec[q + 1], g1);
        s0 += p.x;
        s1 += p.y;
    }
    if (q < valid) {
        g0 = fma(col[q], ec[q], g0);
        s0 += col[q];
    }
    if (!last) {
        a[0] = g0;
        a[1] = g1;
        a[2] = s0;
        a[3] = s1;
        return;
    }
    double E = 0.0;
    for (int i = t; i < n_esum; i += 256) E += esum[i];
    const double G = block_sum_256(g0 + g1, red);
    const double S = block_sum_256(s0 + s1, red);
    const double Et = block_sum_256(E, red);
    if (t == 0) r[j] = G - S * (Et / (double)M);
}

// The residual pass over one chunk of the basis in a single read of it (N <= RESID_MAX_N atoms):
// e = V - Phi u on the chunk's rows, and the sums g_j = sum_R Phi[j][R] e_R, s_j = sum_R Phi[j][R],
// E = sum_R e_R that r_j = g_j - s_j E / M needs (k_ref_finish forms r).  Work-group w takes the passes
// w, w + G, ... of the chunk (G work-groups; a pass is one or more 16-row blocks, below); each block is
// loaded once into registers: lane (row quad rq = lane & 3, column lane cl = (lane >> 2) + 16 wave) holds
// rows 4 rq .. 4 rq + 3 of the columns cl + 128 k, so a wave reads 16 whole 128-byte column segments per
// k.  The row dot products are summed over the column lanes (shuffles, then the 8 waves' partials in order
// through LDS), e comes back through LDS, and each lane adds its rows' share to its columns' g and s.  The
// next pass's loads are in flight meanwhile.  At the chunk's end a column's blocks and its four row-quad
// lanes' sums are combined and added to the work-group's running sums in acc ([G][2N + 1]: g, s, E),
// which carry over between chunks, so the kept and the recomputed (streamed) basis give the same sums
// bitwise.
constexpr int RESID_THREADS = 512;
constexpr int RESID_KC = 8;                    // column slots per lane
constexpr int RESID_MAX_N = 128 * RESID_KC;    // 1024
constexpr int RESID_MAX_G = 256;               // work-groups

// More than 512 atoms: one 16-row block per pass, eight column groups.
__global__ __launch_bounds__(RESID_THREADS) void k_ref_resid_wide(const double *__restrict__ phi, int64_t rows,
                                                             int64_t valid, int N, const double *__restrict__ v,
                                                             const double *__restrict__ u, double *__restrict__ acc,
                                                             int first, const int *__restrict__ cinfo,
                                                             const RefineState *__restrict__ rs) {
    __shared__ double red[RESID_THREADS / 64][16];   // [wave][row]: the wave's partial of the row's dot product
    __shared__ double es[16];
    if (!refine_active(cinfo, rs)) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int rq = lane & 3, cl = (lane >> 2) + 16 * wave;
    const int G = gridDim.x, w = blockIdx.x;
    const int64_t n_blk = (valid + 15) / 16;
    // a lane's columns past the atoms read the last atom's column with weight 0 and are not stored
    double uk[RESID_KC];
    const double *colp[RESID_KC];
#pragma unroll
    for (int k = 0; k < RESID_KC; ++k) {
        const int j = cl + 128 * k;
        uk[k] = j < N ? u[j] : 0.0;
        colp[k] = phi + (int64_t)min(j, N - 1) * rows + 4 * rq;
    }
    double g[RESID_KC], sc[RESID_KC];
#pragma unroll
    for (int k = 0; k < RESID_KC; ++k) g[k] = sc[k] = 0.0;
    double esum = 0.0;   // threads 0..15: running sum of e over their row of every block
    // column slots k < kc hold atoms (the last one partly); the others stay zero and are skipped
    const int kc = (N + 127) / 128;
    double nx[RESID_KC][4], vn = 0.0;
#pragma unroll
    for (int k = 0; k < RESID_KC; ++k) nx[k][0] = nx[k][1] = nx[k][2] = nx[k][3] = 0.0;
    // the block's rows of this lane's columns (and, threads 0..15, V of its row), none past the block count
    auto load = [&](int64_t b) {
        if (b < n_blk) {
#pragma unroll
            for (int k = 0; k < RESID_KC; ++k) {
                if (k >= kc) break;
                const double2 *p = reinterpret_cast<const double2 *>(colp[k] + 16 * b);
                const double2 a = p[0], c = p[1];
                nx[k][0] = a.x;
                nx[k][1] = a.y;
                nx[k][2] = c.x;
                nx[k][3] = c.y;
            }
            if (t < 16 && 16 * b + t < valid) vn = v[16 * b + t];
        }
    };
    load(w);
    for (int64_t b = w; b < n_blk; b += G) {
        // rows past the chunk's kept points count as zero
        const int64_t r = 16 * b + 4 * rq;
        double cu[RESID_KC][4];
#pragma unroll
        for (int k = 0; k < RESID_KC; ++k)
#pragma unroll
            for (int i = 0; i < 4; ++i) cu[k][i] = r + i < valid ? nx[k][i] : 0.0;
        const double vc = vn;
        load(b + G);
        double p[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < RESID_KC; ++k) {
            if (k >= kc) break;
#pragma unroll
            for (int i = 0; i < 4; ++i) p[i] = fma(cu[k][i], uk[k], p[i]);
        }
#pragma unroll
        for (int o = 4; o < 64; o <<= 1)
#pragma unroll
            for (int i = 0; i < 4; ++i) p[i] += __shfl_xor(p[i], o, 64);
        if (lane < 4) {
#pragma unroll
            for (int i = 0; i < 4; ++i) red[wave][4 * lane + i] = p[i];
        }
        __syncthreads();
        if (t < 16) {
            double d = 0.0;
#pragma unroll
            for (int q = 0; q < RESID_THREADS / 64; ++q) d += red[q][t];
            const double e = 16 * b + t < valid ? vc - d : 0.0;
            es[t] = e;
            esum += e;
        }
        __syncthreads();
        double ev[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) ev[i] = es[4 * rq + i];
#pragma unroll
        for (int k = 0; k < RESID_KC; ++k) {
            if (k >= kc) break;
            g[k] = fma(cu[k][0], ev[0], fma(cu[k][1], ev[1], fma(cu[k][2], ev[2], fma(cu[k][3], ev[3], g[k]))));
            sc[k] += (cu[k][0] + cu[k][1]) + (cu[k][2] + cu[k][3]);
        }
    }
    // the chunk's sums of the work-group: the four row quads of each column, then the running sums
#pragma unroll
    for (int k = 0; k < RESID_KC; ++k) {
        g[k] += __shfl_xor(g[k], 1, 64);
        sc[k] += __shfl_xor(sc[k], 1, 64);
        g[k] += __shfl_xor(g[k], 2, 64);
        sc[k] += __shfl_xor(sc[k], 2, 64);
    }
    double *aw = acc + (int64_t)w * (2 * N + 1);
    if (rq == 0) {
#pragma unroll
        for (int k = 0; k < RESID_KC; ++k) {
            const int j = cl + 128 * k;
            if (j >= N) continue;
            aw[j] = first ? g[k] : aw[j] + g[k];
            aw[N + j] = first ? sc[k] : aw[N + j] + sc[k];
        }
    }
    if (wave == 0) {
        double E = t < 16 ? esum : 0.0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) E += __shfl_xor(E, o, 64);
        if (t == 0) aw[2 * N] = first ? E : aw[2 * N] + E;
    }
}

// Up to 512 atoms, KCL column groups of 128 atoms (a power of two, 128 KCL >= N): a lane's 8 slots hold its KCL columns of
// RB = 8 / KCL consecutive 16-row blocks (slot k: block k / KCL, column group k % KCL), so with few atoms
// one pass of the loop covers 16 RB rows and its two barriers serve them all.
template <int KCL>
__global__ __launch_bounds__(RESID_THREADS) void k_ref_resid(const double *__restrict__ phi, int64_t rows,
                                                             int64_t valid, int N, const double *__restrict__ v,
                                                             const double *__restrict__ u, double *__restrict__ acc,
                                                             int first, const int *__restrict__ cinfo,
                                                             const RefineState *__restrict__ rs) {
    constexpr int RB = RESID_KC / KCL;   // row blocks per pass
    __shared__ double red[RESID_THREADS / 64][16 * RB];   // [wave][row]: the wave's partial of the row's product
    __shared__ double es[16 * RB];
    if (!refine_active(cinfo, rs)) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int rq = lane & 3, cl = (lane >> 2) + 16 * wave;
    const int G = gridDim.x, w = blockIdx.x;
    const int64_t n_blk = (valid + 15) / 16, n_pass = (n_blk + RB - 1) / RB;
    // column groups past the atoms (gg >= kc) are neither loaded nor summed; a lane's column past the atoms
    // reads the last atom's column with weight 0 and is not stored
    const int kc = (N + 127) / 128;
    double uk[KCL];
    const double *colp[KCL];
#pragma unroll
    for (int gg = 0; gg < KCL; ++gg) {
        const int j = cl + 128 * gg;
        uk[gg] = j < N ? u[j] : 0.0;
        colp[gg] = phi + (int64_t)min(j, N - 1) * rows + 4 * rq;
    }
    double g[RESID_KC], sc[RESID_KC];
#pragma unroll
    for (int k = 0; k < RESID_KC; ++k) g[k] = sc[k] = 0.0;
    double esum = 0.0;   // threads below 16 RB: running sum of e over their row of every pass
    double nx[RESID_KC][4], vn = 0.0;
#pragma unroll
    for (int k = 0; k < RESID_KC; ++k) nx[k][0] = nx[k][1] = nx[k][2] = nx[k][3] = 0.0;
    // a pass's rows of this lane's columns (and, threads below 16 RB, V of their row)
    auto load = [&](int64_t ps) {
        if (ps < n_pass) {
#pragma unroll
            for (int k = 0; k < RESID_KC; ++k) {
                const int b = k / KCL, gg = k % KCL;
                const int64_t blk = RB * ps + b;
                if (gg >= kc || blk >= n_blk) continue;
                const double2 *p = reinterpret_cast<const double2 *>(colp[gg] + 16 * blk);
                const double2 a = p[0], c = p[1];
                nx[k][0] = a.x;
                nx[k][1] = a.y;
                nx[k][2] = c.x;
                nx[k][3] = c.y;
            }
            if (t < 16 * RB && 16 * RB * ps + t < valid) vn = v[16 * RB * ps + t];
        }
    };
    load(w);
    for (int64_t ps = w; ps < n_pass; ps += G) {
        // rows past the chunk's kept points count as zero
        double cu[RESID_KC][4];
#pragma unroll
        for (int k = 0; k < RESID_KC; ++k) {
            const int64_t r = 16 * (RB * ps + k / KCL) + 4 * rq;
#pragma unroll
            for (int i = 0; i < 4; ++i) cu[k][i] = r + i < valid ? nx[k][i] : 0.0;
        }
        const double vc = vn;
        load(ps + G);
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            double p[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int gg = 0; gg < KCL; ++gg) {
                if (gg >= kc) break;
#pragma unroll
                for (int i = 0; i < 4; ++i) p[i] = fma(cu[b * KCL + gg][i], uk[gg], p[i]);
            }
#pragma unroll
            for (int o = 4; o < 64; o <<= 1)
#pragma unroll
                for (int i = 0; i < 4; ++i) p[i] += __shfl_xor(p[i], o, 64);
            if (lane < 4) {
#pragma unroll
                for (int i = 0; i < 4; ++i) red[wave][16 * b + 4 * lane + i] = p[i];
            }
        }
        __syncthreads();
        if (t < 16 * RB) {
            double d = 0.0;
#pragma unroll
            for (int q = 0; q < RESID_THREADS / 64; ++q) d += red[q][t];
            const double e = 16 * RB * ps + t < valid ? vc - d : 0.0;
            es[t] = e;
            esum += e;
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            double ev[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) ev[i] = es[16 * b + 4 * rq + i];
#pragma unroll
            for (int gg = 0; gg < KCL; ++gg) {
                if (gg >= kc) break;
                const int k = b * KCL + gg;
                g[k] = fma(cu[k][0], ev[0], fma(cu[k][1], ev[1], fma(cu[k][2], ev[2], fma(cu[k][3], ev[3], g[k]))));
                sc[k] += (cu[k][0] + cu[k][1]) + (cu[k][2] + cu[k][3]);
            }
        }
    }
    // the chunk's sums of the work-group: a column's blocks in order, its four row quads, then the running sums
    double gs[KCL], ss[KCL];
#pragma unroll
    for (int gg = 0; gg < KCL; ++gg) {
        gs[gg] = g[gg];
        ss[gg] = sc[gg];
#pragma unroll
        for (int b = 1; b < RB; ++b) {
            gs[gg] += g[b * KCL + gg];
            ss[gg] += sc[b * KCL + gg];
        }
        gs[gg] += __shfl_xor(gs[gg], 1, 64);
        ss[gg] += __shfl_xor(ss[gg], 1, 64);
        gs[gg] += __shfl_xor(gs[gg], 2, 64);
        ss[gg] += __shfl_xor(ss[gg], 2, 64);
    }
    double *aw = acc + (int64_t)w * (2 * N + 1);
    if (rq == 0) {
#pragma unroll
        for (int gg = 0; gg < KCL; ++gg) {
            const int j = cl + 128 * gg;
            if (j >= N) continue;
            aw[j] = first ? gs[gg] : aw[j] + gs[gg];
            aw[N + j] = first ? ss[gg] : aw[N + j] + ss[gg];
        }
    }
    // the e sums of the rows, in order (the red array is free after the loop's second barrier)
    __syncthreads();
    if (t < 16 * RB) red[0][t] = esum;
    __syncthreads();
    if (t == 0) {
        double E = 0.0;
        for (int q = 0; q < 16 * RB; ++q) E += red[0][q];
        aw[2 * N] = first ? E : aw[2 * N] + E;
    }
}

// The residual kernel for N atoms: up to 512, column groups rounded up to a power of two.
inline void launch_ref_resid(int N, int G, hipStream_t st, const double *phi, int64_t rows, int64_t valid,
                             const double *v, const double *u, double *acc, int first, const int *cinfo,
                             const RefineState *rs) {
    const int kc = (N + 127) / 128;
    if (kc <= 1)
        k_ref_resid<1><<<G, RESID_THREADS, 0, st>>>(phi, rows, valid, N, v, u, acc, first, cinfo, rs);
    else if (kc <= 2)
        k_ref_resid<2><<<G, RESID_THREADS, 0, st>>>(phi, rows, valid, N, v, u, acc, first, cinfo, rs);
    else if (kc <= 4)
        k_ref_resid<4><<<G, RESID_THREADS, 0, st>>>(phi, rows, valid, N, v, u, acc, first, cinfo, rs);
    else
        k_ref_resid_wide<<<G, RESID_THREADS, 0, st>>>(phi, rows, valid, N, v, u, acc, first, cinfo, rs);
}

// r_j = g_j - s_j E / M from the work-groups' sums of k_ref_resid: 32 columns per work-group, eight
// threads per column, thread q summing the work-groups w = q mod 8 in order, then the eight in order.
__global__ __launch_bounds__(256) void k_ref_finish(const double *__restrict__ acc, int G, int N, int64_t M,
                                                    double *__restrict__ r, const int *__restrict__ cinfo,
                                                    const RefineState *__restrict__ rs) {
    __shared__ double gp[8][32], sp[8][32], ep[8];
    if (!refine_active(cinfo, rs)) return;
    const int t = threadIdx.x, q = t >> 5, jj = t & 31, j = blockIdx.x * 32 + jj;
    const int64_t ld = 2 * N + 1;
    double gs = 0.0, ss = 0.0, E = 0.0;
    for (int w = q; w < G; w += 8) {
        if (j < N) {
            gs += acc[w * ld + j];
            ss += acc[w * ld + N + j];
        }
        if (jj == 0) E += acc[w * ld + 2 * N];
    }
    gp[q][jj] = gs;
    sp[q][jj] = ss;
    if (jj == 0) ep[q] = E;
    __syncthreads();
    if (t < 32 && j < N) {
        double g = 0.0, s = 0.0, e = 0.0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            g += gp[k][t];
            s += sp[k][t];
            e += ep[k];
        }
        r[j] = g - s * (e / (double)M);
    }
}

// r1 = P^T r + rhs_add - diag_add x - w lambda (into r1), r2 = q_tot - w . x; one work-group.
__global__ __launch_bounds__(256) void k_ref_rhs(const double *__restrict__ r, const int *__restrict__ red_ptr,
                                                 const int *__restrict__ red_src, const double *__restrict__ red_mult,
                                                 const double *__restrict__ w, const double *__restrict__ diag_add,
                                                 const double *__restrict__ rhs_add, double q_tot, int nr,
                                                 const double *__restrict__ x, double *__restrict__ r1,
                                                 double *__restrict__ r2, const int *__restrict__ cinfo,
                                                 const RefineState *__restrict__ rs) {
    __shared__ double red[4];
    if (!refine_active(cinfo, rs)) return;
    const double lambda = x[nr];
    double pw = 0.0;
    for (int I = threadIdx.x; I < nr; I += 256) {
        double s = 0.0;
        for (int p = red_ptr[I]; p < red_ptr[I + 1]; ++p) s = fma(red_mult[p], r[red_src[p]], s);
        if (rhs_add) s += rhs_add[I];
        if (diag_add) s -= diag_add[I] * x[I];
        r1[I] = s - w[I] * lambda;
        pw = fma(w[I], x[I], pw);
    }
    const double wx = block_sum_256(pw, red);
    if (threadIdx.x == 0) *r2 = q_tot - wx;
}

// z = L^-1 r1 in one launch: the order of k_backsub reversed.  Ticket i owns diagonal block i and waits
// for the blocks before it, oldest first; L_{kb,m} (m < kb) is read from the transposed panels in the
// upper triangle (row m0 + j, column j0 + i holds L[j0 + i][m0 + j]), loaded before each wait.
// Work-group kb publishes z_kb = D_kb u_kb (D = the diagonal block's inverse) with a release fence and
// its flag; a wait that sees no progress for spin_limit polls sets *status and the work-group leaves.
__global__ __launch_bounds__(256) void k_fwdsub(const double *__restrict__ A, int ld, int n,
                                                const double *__restrict__ Dinv, const double *__restrict__ r1,
                                                double *__restrict__ z, unsigned *__restrict__ sync, unsigned base,
                                                unsigned epoch, unsigned spin_limit, int *__restrict__ status,
                                                const int *__restrict__ cinfo, const RefineState *__restrict__ rs) {
    __shared__ double Dk[CB][CB + 1];
    __shared__ double uk[CB], zm[CB];
    __shared__ double part[4][CB];
    __shared__ unsigned sh_ticket;
    __shared__ int sh_abort;
    const int t = threadIdx.x;
    unsigned *counter = sync, *flags = sync + 1;
    if (t == 0) {
        sh_ticket = atomicAdd(counter, 1u) - base;
        sh_abort = 0;
    }
    __syncthreads();
    if (!refine_active(cinfo, rs)) return;
    const int nblk = (n + CB - 1) / CB;
    if (sh_ticket >= (unsigned)nblk) {
        if (t == 0) atomicExch(status, 1);
        return;
    }
    const int kb = (int)sh_ticket;
    const int j0 = kb * CB, nb = min(CB, n - j0);
    for (int q = t; q < CB * CB; q += 256) Dk[q / CB][q % CB] = Dinv[(int64_t)kb * CB * CB + q];
    if (t < CB) uk[t] = t < nb ? r1[j0 + t] : 0.0;
    // thread (i = t % 64: row of block kb, g = t / 64: 16 columns of block m)
    const int i = t & 63, g = t >> 6;
    const bool irow = i < nb;
    for (int m = 0; m < kb; ++m) {
        const int m0 = m * CB;   // blocks before kb are full
        double lv[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) lv[q] = irow ? A[(int64_t)(m0 + g * 16 + q) * ld + j0 + i] : 0.0;
        if (t == 0) {
            for (unsigned spins = 0;; ++spins) {
                if (spins >= spin_limit) {
                    sh_abort = 1;
                    atomicExch(status, 1);
                    break;
                }
                if (__hip_atomic_load(flags + m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __syncthreads();
        if (sh_abort) return;
        if (t < CB) zm[t] = __hip_atomic_load(z + m0 + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 16; ++q) s = fma(lv[q], zm[g * 16 + q], s);
        part[g][i] = s;
        __syncthreads();
        if (t < CB) uk[t] -= (part[0][t] + part[1][t]) + (part[2][t] + part[3][t]);
        __syncthreads();
    }
    __syncthreads();
    if (t < nb) {
        double y = 0.0;
        for (int q = 0; q <= t; ++q) y = fma(Dk[t][q], uk[q], y);
        __hip_atomic_store(z + j0 + t, y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
    __syncthreads();
    if (t == 0) __hip_atomic_store(flags + kb, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// x += dx (charges and multiplier).  The correction's max |dq| gives the error bound left: the steps
// contract the error by rho per step, estimated as the ratio of the last two corrections (RHO_PRIOR
// after the first), so the remaining error is at most rho / (1 - rho) |dq|.  The next step runs while
// that bound exceeds the target.
__global__ __launch_bounds__(256) void k_ref_update(double *__restrict__ x, const double *__restrict__ dx, int nr,
                                                    double target, const int *__restrict__ cinfo,
                                                    RefineState *__restrict__ rs) {
    __shared__ double red[4];
    if (!refine_active(cinfo, rs)) return;
    double mx = 0.0;
    for (int I = threadIdx.x; I <= nr; I += 256) {
        const double d = dx[I];
        x[I] += d;
        if (I < nr) mx = fmax(mx, fabs(d));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        const double prev = rs->last_dq;
        rs->steps += 1;
        rs->last_dq = m;
        const double rho = (rs->steps >= 2 && prev > 0.0) ? m / prev : RHO_PRIOR;
        const double bound = m == 0.0 ? 0.0 : (rho < 1.0 ? rho / (1.0 - rho) * m : INFINITY);
        rs->bound = bound;
        rs->go = bound > target ? 1 : 0;
    }
}

// Sensitivity of the charges to the rounding of the system in fp64: the right-hand side
// z_i = +-eps sum_j |K_ij| |x_j| (eps = 8 units in the last place, signs pseudo-random), whose solution
// with the same factor is the charge change that a perturbation of every matrix entry by eps relative
// causes; one work-group per row.  Also arms the state that gates that solve.
__global__ __launch_bounds__(256) void k_sens_rhs(const double *__restrict__ K, int ld, int nr, const double *__restrict__ x,
                                                  double eps, double *__restrict__ z, const int *__restrict__ cinfo,
                                                  RefineState *__restrict__ sens) {
    __shared__ double red[4];
    if (cinfo[0] != 0 || cinfo[1] != 0) return;
    const int i = blockIdx.x;
    double acc = 0.0;
    for (int j = threadIdx.x; j < nr; j += 256) acc = fma(fabs(K[(int64_t)i * ld + j]), fabs(x[j]), acc);
    const double sum = block_sum_256(acc, red);
    if (threadIdx.x == 0) {
        const unsigned h = (unsigned)i * 2654435761u;
        z[i] = ((h >> 15) & 1u ? eps : -eps) * sum;
        if (i == 0) {
            sens->go = 1;
            sens->steps = 0;
            sens->last_dq = 0.0;
            sens->bound = -1.0;
        }
    }
}

// *out = max |dq| of the sensitivity solve (-1 when the solve did not run); one work-group.
__global__ __launch_bounds__(256) void k_sens_max(const double *__restrict__ dq, int nr, const int *__restrict__ cinfo,
                                                  double *__restrict__ out) {
    __shared__ double red[4];
    double m = 0.0;
    for (int i = threadIdx.x; i < nr; i += 256) m = fmax(m, fabs(dq[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        *out = (cinfo[0] != 0 || cinfo[1] != 0) ? -1.0 : fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Restarts the refinement after a fallback solve: clears the pivot / wait flags and the step state.
__global__ void k_ref_reset(int *__restrict__ cinfo, RefineState *__restrict__ rs) {
    cinfo[0] = 0;
    cinfo[1] = 0;
    rs->go = 1;
    rs->steps = 0;
    rs->last_dq = 0.0;
    rs->bound = -1.0;
}

// dx = F^-1 [r1; r2] with the LU factors F (unit lower L below the diagonal, U on and above it) and the
// interchanges piv of lu_solve_augmented (LAPACK getrs order: swaps, L, U); one work-group.  Used for the
// refinement's corrections after the Cholesky fell back to LU.
__global__ __launch_bounds__(1024) void k_lu_getrs(const double *__restrict__ F, int ld, int n, const int *__restrict__ piv,
                                                   const double *__restrict__ r1, const double *__restrict__ r2,
                                                   double *__restrict__ dx, const int *__restrict__ cinfo,
                                                   const RefineState *__restrict__ rs) {
    extern __shared__ double y[];
    if (!refine_active(cinfo, rs)) return;
    const int t = threadIdx.x;
    for (int i = t; i < n; i += 1024) y[i] = i < n - 1 ? r1[i] : *r2;
    __syncthreads();
    if (t == 0)
        for (int j = 0; j < n; ++j) {
            const int p = piv[j];
            if (p != j) {
                const double tmp = y[j];
                y[j] = y[p];
                y[p] = tmp;
            }
        }
    __syncthreads();
    for (int j = 0; j < n - 1; ++j) {
        const double yj = y[j];
        for (int r = j + 1 + t; r < n; r += 1024) y[r] -= F[(int64_t)r * ld + j] * yj;
        __syncthreads();
    }
    for (int i = n - 1; i >= 0; --i) {
        const double xi = y[i] / F[(int64_t)i * ld + i];
        __syncthreads();
        if (t == 0) y[i] = xi;
        for (int r = t; r < i; r += 1024) y[r] -= F[(int64_t)r * ld + i] * xi;
        __syncthreads();
    }
    for (int i = t; i < n; i += 1024) dx[i] = y[i];
}

// ---- minimum-norm least squares of the bordered system (numpy.linalg.lstsq, rcond=None): the
// symmetric matrix S = V diag(l) V^T by parallel cyclic Jacobi, then x = sum over |l_i| > eps n max|l|
// of v_i (v_i . y) / l_i.  For symmetric S the singular values are |l_i| and the right singular
// vectors are the v_i, so this is the SVD solution lstsq computes.
//
// A sweep visits every pair (p, q) once in ne - 1 rounds of ne / 2 disjoint pairs (ne = n rounded up to
// even; index n is a dummy when n is odd): round-robin pairing, slot 0 fixed and slot s >= 1 holding
// index 1 + (s - 1 + r) mod (ne - 1); pair l joins slots l and ne - 1 - l.  The rotations of a round act on
// disjoint planes, so S <- G^T S G for all of them at once: work-group l rewrites rows p_l and q_l, each
// entry from the 2 x 2 block of its own rows and the column pair of that entry.  Rows of W are the
// eigenvectors (W <- G^T W).
struct JacPair {
    double c, s, t, apq;
    int p, q, rot, pad_;
};

__device__ __forceinline__ int jac_index(int slot, int r, int ne) { return slot == 0 ? 0 : 1 + (slot - 1 + r) % (ne - 1); }
__device__ __forceinline__ int jac_slot(int idx, int r, int ne) {
    return idx == 0 ? 0 : 1 + ((idx - 1 - r) % (ne - 1) + (ne - 1)) % (ne - 1);
}

// Rotation of each pair of round r: zero S[p][q] (Rutishauser's tangent), unless |S[p][q]| is at most
// rel sqrt(|S[p][p] S[q][q]|) or the absolute floor; counts the rotations of the sweep.
__global__ __launch_bounds__(256) void k_jac_angles(const double *__restrict__ S, int ld, int n, int ne, int r,
                                                    const double *__restrict__ floor_abs, double rel,
                                                    JacPair *__restrict__ pairs, int *__restrict__ n_rot) {
    const int l = blockIdx.x * 256 + threadIdx.x;
    if (l >= ne / 2) return;
    JacPair P{1.0, 0.0, 0.0, 0.0, jac_index(l, r, ne), jac_index(ne - 1 - l, r, ne), 0, 0};
    if (P.p < n && P.q < n) {
        const double app = S[(int64_t)P.p * ld + P.p], aqq = S[(int64_t)P.q * ld + P.q];
        const double apq = 0.5 * (S[(int64_t)P.p * ld + P.q] + S[(int64_t)P.q * ld + P.p]);
        if (fabs(apq) > fmax(rel * sqrt(fabs(app) * fabs(aqq)), *floor_abs)) {
            const double th = (aqq - app) / (2.0 * apq);
            const double t = fabs(th) > 1e150 ? 0.5 / th : copysign(1.0, th) / (fabs(th) + sqrt(fma(th, th, 1.0)));
            const double c = 1.0 / sqrt(fma(t, t, 1.0));
            P.c = c;
            P.s = t * c;
            P.t = t;
            P.apq = apq;
            P.rot = 1;
            atomicAdd(n_rot, 1);
        }
    }
    pairs[l] = P;
}

// One round: work-group l rewrites rows p_l, q_l of S (both held in LDS first) and of W.
__global__ __launch_bounds__(256) void k_jac_round(double *__restrict__ S, int ld, int n, int ne, int r,
                                                   const JacPair *__restrict__ pairs, double *__restrict__ W) {
    extern __shared__ double rows[];   // [2][n]: old rows p and q
    const JacPair P = pairs[blockIdx.x];
    const bool hp = P.p < n, hq = P.q < n;
    double *sp = rows, *sq = rows + n;
    const int t = threadIdx.x;
    for (int j = t; j < n; j += 256) {
        sp[j] = hp ? S[(int64_t)P.p * ld + j] : 0.0;
        sq[j] = hq ? S[(int64_t)P.q * ld + j] : 0.0;
    }
    __syncthreads();
    for (int j = t; j < n; j += 256) {
        const int slot = jac_slot(j, r, ne), pslot = ne - 1 - slot;
        const int pj = jac_index(pslot, r, ne);
        double vp = sp[j], vq = sq[j];
        if (pj < n) {
            const JacPair Q = pairs[min(slot, pslot)];
            // column j of S G: c S[., j] - s S[., partner] for the first of a pair, s S[., partner] + c S[., j]
            // for the second
            if (slot < pslot) {
                vp = Q.c * sp[j] - Q.s * sp[pj];
                vq = Q.c * sq[j] - Q.s * sq[pj];
            } else {
                vp = Q.s * sp[pj] + Q.c * sp[j];
                vq = Q.s * sq[pj] + Q.c * sq[j];
            }
        }
        double np_ = P.c * vp - P.s * vq, nq = P.s * vp + P.c * vq;
        if (P.rot && (j == P.p || j == P.q)) {
            // the 2 x 2 block of the pair itself, diagonalised exactly
            if (j == P.p) {
                np_ = sp[P.p] - P.t * P.apq;
                nq = 0.0;
            } else {
                np_ = 0.0;
                nq = sq[P.q] + P.t * P.apq;
            }
        }
        if (hp) S[(int64_t)P.p * ld + j] = np_;
        if (hq) S[(int64_t)P.q * ld + j] = nq;
    }
    if (P.rot)
        for (int i = t; i < n; i += 256) {
            const double wp = W[(int64_t)P.p * n + i], wq = W[(int64_t)P.q * n + i];
            W[(int64_t)P.p * n + i] = P.c * wp - P.s * wq;
            W[(int64_t)P.q * n + i] = P.s * wp + P.c * wq;
        }
}

__global__ __launch_bounds__(256) void k_eye(double *__restrict__ W, int n) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < (int64_t)n * n) W[e] = (e / n == e % n) ? 1.0 : 0.0;
}

__global__ void k_fill_i32(int *p, int v) { *p = v; }

// scal[0] = the rotation floor (||S||_F times 2^-52 / 100); one work-group.
__global__ __launch_bounds__(1024) void k_jac_floor(const double *__restrict__ S, int ld, int n, double *__restrict__ scal) {
    __shared__ double red[16];
    double acc = 0.0;
    for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += 1024) {
        const double v = S[(e / n) * ld + e % n];
        acc = fma(v, v, acc);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < 16; ++w) s += red[w];
        scal[0] = sqrt(s) * 2.220446049250313e-16 * 1e-2;
    }
}

// scal[1] = the cut-off eps * n * max |l| of numpy.linalg.lstsq (rcond = None); one work-group.
__global__ __launch_bounds__(1024) void k_jac_cut(const double *__restrict__ S, int ld, int n, double *__restrict__ scal) {
    __shared__ double red[16];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) m = fmax(m, fabs(S[(int64_t)i * ld + i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < 16; ++w) s = fmax(s, red[w]);
        scal[1] = 2.220446049250313e-16 * (double)n * s;
    }
}

// y = [a[0 .. n-1) ; *b] (a strided by stride), into one contiguous vector.
__global__ __launch_bounds__(256) void k_gather_rhs(const double *__restrict__ a, int64_t stride, const double *__restrict__ b,
                                                    int n, double *__restrict__ y, const int *__restrict__ cinfo,
                                                    const RefineState *__restrict__ rs) {
    if (rs && !refine_active(cinfo, rs)) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = i < n - 1 ? a[(int64_t)i * stride] : *b;
}

// coef_i = (w_i . y) / l_i for |l_i| above the cut-off, else 0; one work-group per i.
__global__ __launch_bounds__(256) void k_pinv_coef(const double *__restrict__ W, const double *__restrict__ S, int ld, int n,
                                                   const double *__restrict__ y, const double *__restrict__ scal,
                                                   double *__restrict__ coef, const int *__restrict__ cinfo,
                                                   const RefineState *__restrict__ rs) {
    __shared__ double red[4];
    if (rs && !refine_active(cinfo, rs)) return;
    const int i = blockIdx.x;
    double acc = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) acc = fma(W[(int64_t)i * n + j], y[j], acc);
    const double d = block_sum_256(acc, red);
    if (threadIdx.x == 0) {
        const double l = S[(int64_t)i * ld + i];
        coef[i] = fabs(l) > scal[1] ? d / l : 0.0;
    }
}

// x_j = sum_i coef_i w_i[j]
__global__ __launch_bounds__(256) void k_pinv_combine(const double *__restrict__ W, int n, const double *__restrict__ coef,
                                                      double *__restrict__ x, const int *__restrict__ cinfo,
                                                      const RefineState *__restrict__ rs) {
    if (rs && !refine_active(cinfo, rs)) return;
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    double acc = 0.0;
    for (int i = 0; i < n; ++i) acc = fma(coef[i], W[(int64_t)i * n + j], acc);
    x[j] = acc;
}

// Back substitution U x = y (y = column n of the factored augmented matrix); x -> out.
__global__ __launch_bounds__(1024) void k_backsub(const double *__restrict__ A, int ld, int n, double *__restrict__ out) {
    extern __shared__ double y[];
    for (int r = threadIdx.x; r < n; r += 1024) y[r] = A[(int64_t)r * ld + n];
    __syncthreads();
    for (int i = n - 1; i >= 0; --i) {
        const double xi = y[i] / A[(int64_t)i * ld + i];
        __syncthreads();
        if (threadIdx.x == 0) y[i] = xi;
        for (int r = threadIdx.x; r < i; r += 1024) y[r] -= A[(int64_t)r * ld + i] * xi;
        __syncthreads();
    }
    for (int r = threadIdx.x; r < n; r += 1024) out[r] = y[r];
}

// Copy of the constraint block from mapped pinned host memory (a few KB; the kernel's loads cross
// the bus once, which starts sooner than a copy-engine transfer).
__global__ __launch_bounds__(256) void k_fetch_host(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

// The results into mapped pinned host memory: out[0] the pivot flag (-1: a triangular-solve wait gave
// up, which the host also answers with the LU), out[1..N] the expanded charges, out[1+N..1+N+n) the
// reduced solution and the multiplier, then the refinement steps run and the last correction; one
// synchronisation then returns them.
__global__ __launch_bounds__(256) void k_publish(const double *__restrict__ x, const int *__restrict__ src, int N, int n,
                                                 const int *__restrict__ cinfo, const RefineState *__restrict__ rs,
                                                 double *__restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0) {
        out[0] = cinfo[0] != 0 ? (double)cinfo[0] : (cinfo[1] != 0 ? -1.0 : 0.0);
        out[1 + N + n] = (double)rs->steps;
        out[2 + N + n] = rs->last_dq;
        out[3 + N + n] = rs->bound;
        out[5 + N + n] = (double)rs->go;
    }
    if (i < N) out[1 + i] = src[i] >= 0 ? x[src[i]] : 0.0;
    if (i < n) out[1 + N + i] = x[i];
}

}  // namespace

int lu_solve_augmented(Ctx *ctx, double *K2, int ld, int n, int *piv, int *info, double *x, hipStream_t st) {
    const int ncols = n + 1;
    k_fill_i32<<<1, 1, 0, st>>>(info, 0);
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = std::min(NB, n - k0);
        k_panel<<<1, 1024, 0, st>>>(K2, ld, n, k0, nb, piv, info);
        k_swap_rest<<<ceil_div(ncols, 256), 256, 0, st>>>(K2, ld, ncols, k0, nb, piv);
        const int right = ncols - (k0 + nb);
        if (right > 0) {
            k_trsm<<<ceil_div(right, 256), 256, 0, st>>>(K2, ld, ncols, k0, nb);
            const int down = n - (k0 + nb);
            if (down > 0) {
                dim3 g(ceil_div(right, 64), ceil_div(down, 64));
                k_update<<<g, 256, 0, st>>>(K2, ld, n, ncols, k0, nb);
            }
        }
    }
    int h_info = 0;
    RP_HIP(ctx, hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, st));
    RP_HIP(ctx, hipStreamSynchronize(st));
    if (h_info != 0)
        return ctx->fail(RP_ESINGULAR, "system matrix is exactly singular (zero pivot at row " + std::to_string(h_info) + ")");
    k_backsub<<<1, 1024, sizeof(double) * n, st>>>(K2, ld, n, x);
    RP_HIP(ctx, hipGetLastError());
    return RP_OK;
}

namespace {

// Eigendecomposition of the symmetric n x n block of S (in place: S -> diag(l), rows of W = eigenvectors)
// by parallel cyclic Jacobi; scal: [0] rotation floor, [1] lstsq cut-off (set here).  Sweeps until one
// makes no rotation (at most max_sweeps); one synchronisation per sweep reads the count.
int jacobi_eigen(Ctx *ctx, double *S, int ld, int n, double *W, JacPair *pairs, int *d_rot, int *h_rot,
                 double *scal, hipStream_t st, int max_sweeps = 40) {
    const int ne = n + (n & 1), m = ne / 2;
    const size_t lds = 2 * sizeof(double) * (size_t)n;
    if (lds > 160 * 1024) return ctx->fail(RP_EINVAL, "least-squares fallback: more than 10240 unknowns");
    if (lds > 64 * 1024)
        RP_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(&k_jac_round),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    k_eye<<<ceil_div((int64_t)n * n, 256), 256, 0, st>>>(W, n);
    k_jac_floor<<<1, 1024, 0, st>>>(S, ld, n, scal);
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        k_fill_i32<<<1, 1, 0, st>>>(d_rot, 0);
        for (int r = 0; r < ne - 1; ++r) {
            k_jac_angles<<<ceil_div(m, 256), 256, 0, st>>>(S, ld, n, ne, r, scal, 1e-15, pairs, d_rot);
            k_jac_round<<<m, 256, lds, st>>>(S, ld, n, ne, r, pairs, W);
        }
        RP_HIP(ctx, hipGetLastError());
        RP_HIP(ctx, hipMemcpyAsync(h_rot, d_rot, sizeof(int), hipMemcpyDeviceToHost, st));
        RP_HIP(ctx, hipStreamSynchronize(st));
        if (*h_rot == 0) {
            k_jac_cut<<<1, 1024, 0, st>>>(S, ld, n, scal);
            RP_HIP(ctx, hipGetLastError());
            return RP_OK;
        }
    }
    return ctx->fail(RP_ESINGULAR, "system matrix is exactly singular and the least-squares fallback (Jacobi "
                                   "eigendecomposition) did not converge in " + std::to_string(max_sweeps) + " sweeps");
}

}  // namespace

int run_solve(rp_fit *fit, const rp_constraints *con, double *q_out, double *solution_out, int32_t *used_fallback) {
    Ctx *ctx = fit->ctx;
    hipStream_t st = ctx->stream;
    if (!fit->accumulated) return ctx->fail(RP_ESTATE, "rp_fit_solve called before rp_fit_accumulate");
    const int N = fit->n_atoms;
    const int nr = con->n_reduced;
    if (nr <= 0 || nr > N) return ctx->fail(RP_EINVAL, "n_reduced out of range");
    const int nsrc = con->red_ptr[nr];
    for (int i = 0; i < nsrc; ++i)
        if (con->red_src[i] < 0 || con->red_src[i] >= N) return ctx->fail(RP_EINVAL, "red_src index out of range");
    for (int a = 0; a < N; ++a)
        if (con->expand_src[a] >= nr) return ctx->fail(RP_EINVAL, "expand_src index out of range");
    const int n = nr + 1;
    const int ncols = n + 1;
    const int ld = (int)round_up(ncols, 4);
    // the constraint arrays in one pinned block, one upload: ints (red_ptr, red_src, expand_src, and the
    // merge sets by atom for the refinement: inv_ptr, inv_set), then doubles (red_mult, weights,
    // diag_add, rhs_add, inv_mult)
    // The same mapped pinned block then receives the results (k_publish).
    const size_t n_int = (size_t)(nr + 1) + nsrc + N + (size_t)(N + 1) + nsrc;
    const size_t off_d = round_up((int64_t)(4 * n_int), 8);
    const size_t stage = round_up((int64_t)(off_d + 8 * (2 * (size_t)nsrc + 3 * (size_t)nr)), 16);
    const size_t out_bytes = sizeof(double) * (6 + (size_t)N + n);
    if (fit->h_stage_bytes < stage + out_bytes) {
        if (fit->h_stage) (void)hipHostFree(fit->h_stage);
        fit->h_stage = nullptr;
        fit->h_stage_dev = nullptr;
        fit->h_stage_bytes = 0;
        RP_HIP(ctx, hipHostMalloc(&fit->h_stage, stage + out_bytes, hipHostMallocMapped));
        fit->h_stage_bytes = stage + out_bytes;
        RP_HIP(ctx, hipHostGetDevicePointer(&fit->h_stage_dev, fit->h_stage, 0));
    }
    {
        int *hi = static_cast<int *>(fit->h_stage);
        std::copy(con->red_ptr, con->red_ptr + nr + 1, hi);
        std::copy(con->red_src, con->red_src + nsrc, hi + nr + 1);
        std::copy(con->expand_src, con->expand_src + N, hi + nr + 1 + nsrc);
        double *hd = reinterpret_cast<double *>(static_cast<char *>(fit->h_stage) + off_d);
        std::copy(con->red_mult, con->red_mult + nsrc, hd);
        std::copy(con->weights, con->weights + nr, hd + nsrc);
        if (con->diag_add) std::copy(con->diag_add, con->diag_add + nr, hd + nsrc + nr);
        if (con->rhs_add) std::copy(con->rhs_add, con->rhs_add + nr, hd + nsrc + 2 * nr);
        // the merge sets by atom (entries p of red_src in set order), for u = P x in the refinement
        int *inv_ptr = hi + nr + 1 + nsrc + N, *inv_set = inv_ptr + N + 1;
        double *inv_mult = hd + nsrc + 3 * nr;
        std::fill(inv_ptr, inv_ptr + N + 1, 0);
        for (int p = 0; p < nsrc; ++p) ++inv_ptr[con->red_src[p] + 1];
        for (int a = 0; a < N; ++a) inv_ptr[a + 1] += inv_ptr[a];
        std::vector<int> fill(inv_ptr, inv_ptr + N);
        for (int I = 0; I < nr; ++I)
            for (int p = con->red_ptr[I]; p < con->red_ptr[I + 1]; ++p) {
                const int slot = fill[con->red_src[p]]++;
                inv_set[slot] = I;
                inv_mult[slot] = con->red_mult[p];
            }
    }
    DBuf &d_K = fit->ws[12], &d_con = fit->ws[13], &d_piv = fit->ws[19];
    DBuf &d_info = fit->ws[24], &d_x = fit->ws[25], &d_q = fit->ws[27];
    RP_CHECK(d_K.grow(ctx, sizeof(double) * (size_t)n * ld));
    RP_CHECK(d_con.grow(ctx, stage));
    RP_CHECK(d_piv.grow(ctx, sizeof(int) * n));
    RP_CHECK(d_info.grow(ctx, sizeof(int)));
    RP_CHECK(d_x.grow(ctx, sizeof(double) * n));
    RP_CHECK(d_q.grow(ctx, sizeof(double) * N));
    k_fetch_host<<<(unsigned)std::min<int64_t>(ceil_div((int64_t)(stage / 16), 256), 64), 256, 0, st>>>(
        static_cast<const uint4 *>(fit->h_stage_dev), d_con.as<uint4>(), (int64_t)(stage / 16));
    double *h_out = reinterpret_cast<double *>(static_cast<char *>(fit->h_stage) + stage);
    double *h_out_dev = reinterpret_cast<double *>(static_cast<char *>(fit->h_stage_dev) + stage);
    const int *dc_ptr = d_con.as<int>(), *dc_src = dc_ptr + nr + 1, *dc_exp = dc_src + nsrc;
    const int *dc_inv_ptr = dc_exp + N, *dc_inv_set = dc_inv_ptr + N + 1;
    const double *dc_d = reinterpret_cast<const double *>(static_cast<const char *>(d_con.p) + off_d);
    const double *dc_mult = dc_d, *dc_w = dc_d + nsrc, *dc_diag = dc_w + nr, *dc_rhs = dc_diag + nr;
    const double *dc_inv_mult = dc_rhs + nr;

    const int NPg = (int)round_up(N + 2, 128);
    DBuf &d_K2 = fit->ws[20], &d_wc = fit->ws[23];
    DBuf &d_cinfo = fit->ws[28], &d_bc = fit->ws[29], &d_dinv = fit->ws[30];
    RP_CHECK(d_K2.grow(ctx, sizeof(double) * (size_t)n * ld));
    RP_CHECK(d_cinfo.grow(ctx, 2 * sizeof(int)));   // pivot flag, backward-substitution wait status
    RP_CHECK(d_wc.grow(ctx, sizeof(double) * nr));
    RP_CHECK(d_bc.grow(ctx, sizeof(double) * nr));
    RP_CHECK(d_dinv.grow(ctx, sizeof(double) * CB * CB * (size_t)ceil_div(nr, CB)));
    // refinement vectors: u (N), r (N), r1 (nr), z (nr), dx (nr + 1), r2, the sensitivity solve's
    // sz (nr), szz (nr), sdx (nr + 1), then the two states; e over the kept points and its block sums
    const bool refine = fit->opt.precision == RP_PREC_OZAKI_REFINED && (fit->phi_kept || fit->refine_stream);
    DBuf &d_rv = fit->ws[10], &d_e = fit->ws[8];
    const size_t rv_doubles = 2 * (size_t)N + 6 * (size_t)nr + 3;
    RP_CHECK(d_rv.grow(ctx, sizeof(double) * round_up((int64_t)rv_doubles, 2) + 2 * sizeof(RefineState)));
    RefineState *d_rs = reinterpret_cast<RefineState *>(d_rv.as<double>() + round_up((int64_t)rv_doubles, 2));
    RefineState *d_sens = d_rs + 1;   // gates the sensitivity solve
    const int n_esum = (int)ceil_div(fit->n_keep, 512);
    // up to RESID_MAX_N atoms the residual pass is one fused read of each chunk (k_ref_resid), with its
    // work-groups' running sums in ws[15]; above, e over the kept points and a second read for Phi^T e
    const bool fused = N <= RESID_MAX_N;
    const int g_res = (int)std::min<int64_t>(RESID_MAX_G, std::max<int64_t>(fit->phi_rows / 16, 1));
    if (refine && !fused) RP_CHECK(d_e.grow(ctx, sizeof(double) * ((size_t)fit->n_keep + 2 + (size_t)n_esum)));
    DBuf &d_cacc = fit->ws[15];   // fused: the work-groups' sums; streamed: the column pass's partial sums
    if (refine && fused) RP_CHECK(d_cacc.grow(ctx, sizeof(double) * (size_t)g_res * (2 * (size_t)N + 1)));
    if (refine && !fused && fit->refine_stream) RP_CHECK(d_cacc.grow(ctx, sizeof(double) * 4 * 256 * (size_t)N));
    // One residual pass of the refinement: u = P x, e = V - Phi u over the kept points and r = Phi_c^T e,
    // from the kept basis, or (streamed) from each chunk recomputed by the basis kernel into ws[2].
    auto residual_pass = [&]() {
        const int64_t M = fit->n_keep, rows = fit->phi_rows, plane = (int64_t)NPg * rows;
        double *e = fused ? nullptr : d_e.as<double>(), *esum = fused ? nullptr : e + M + 2;
        double *u = d_rv.as<double>(), *r = u + N;
        const int *ci = d_cinfo.as<int>();
        k_ref_expand<<<ceil_div(N, 256), 256, 0, st>>>(d_x.as<double>(), dc_inv_ptr, dc_inv_set, dc_inv_mult, N, u,
                                                       ci, d_rs);
        const int n_chunks = (int)ceil_div(M, rows);
        if (!fit->refine_stream) {
            const double *phi = fit->ws[7].as<double>();
            if (fused) {
                for (int c = 0; c < n_chunks; ++c) {
                    const int64_t r0 = (int64_t)c * rows, valid = std::min<int64_t>(rows, M - r0);
                    launch_ref_resid(N, g_res, st, phi + c * plane, rows, valid, fit->d_v.as<double>() + r0, u,
                                     d_cacc.as<double>(), c == 0, ci, d_rs);
                }
                k_ref_finish<<<ceil_div(N, 32), 256, 0, st>>>(d_cacc.as<double>(), g_res, N, M, r, ci, d_rs);
                return;
            }
            k_ref_rows<<<n_esum, 256, 0, st>>>(phi, rows, plane, N, M, fit->d_v.as<double>(), u, e, esum, ci, d_rs);
            k_ref_cols<<<N, 256, 0, st>>>(phi, rows, plane, M, e, esum, n_esum, r, ci, d_rs);
            return;
        }
        FitDevice &dv = *fit->dev;
        double *phi = fit->ws[2].as<double>();
        for (int c = 0; c < n_chunks; ++c) {
            const int64_t r0 = (int64_t)c * rows, valid = std::min<int64_t>(rows, M - r0);
            sep_phi_chunk(fit, dv, dv.sep_plan, c, r0, valid, rows, phi, fit->ws[5].as<unsigned long long>(), st,
                          &d_rs->go, ci);
            if (fused) {
                launch_ref_resid(N, g_res, st, phi, rows, valid, fit->d_v.as<double>() + r0, u, d_cacc.as<double>(),
                                 c == 0, ci, d_rs);
                continue;
            }
            k_ref_rows<<<ceil_div(valid, 512), 256, 0, st>>>(phi, rows, plane, N, valid, fit->d_v.as<double>() + r0, u,
                                                             e + r0, esum + r0 / 512, ci, d_rs);
            k_ref_cols_chunk<<<N, 256, 0, st>>>(phi, rows, valid, e + r0, d_cacc.as<double>(), c == 0,
                                                c == n_chunks - 1, esum, n_esum, M, r, ci, d_rs);
        }
        if (fused) k_ref_finish<<<ceil_div(N, 32), 256, 0, st>>>(d_cacc.as<double>(), g_res, N, M, r, ci, d_rs);
    };

    hipEvent_t e0, e1;
    RP_HIP(ctx, hipEventCreate(&e0));
    RP_HIP(ctx, hipEventCreate(&e1));
    RP_HIP(ctx, hipEventRecord(e0, st));
    KTimer *kts = new KTimer(ctx, KT_SOLVE);
    // the bordered system, its copy for the LU fallback and the w and b columns of the Cholesky
    k_assemble<<<ceil_div((int64_t)n * ncols, 256), 256, 0, st>>>(fit->d_gram.as<double>(), NPg, N, dc_ptr, dc_src, dc_mult,
                                                                  dc_w, con->diag_add ? dc_diag : nullptr,
                                                                  con->rhs_add ? dc_rhs : nullptr, con->q_tot, nr,
                                                                  d_K.as<double>(), d_K2.as<double>(), ld, d_wc.as<double>(),
                                                                  d_bc.as<double>(), d_cinfo.as<int>(), d_rs);
    // ---- Cholesky of A (leading nr x nr block) + Schur step
    // block 0 on its own; every later diagonal block is factored by the preceding trailing update
    k_potrf_inv<<<1, 256, 0, st>>>(d_K.as<double>(), ld, nr, 0, d_K2.as<double>(), d_cinfo.as<int>(), d_dinv.as<double>(),
                                   d_wc.as<double>(), d_bc.as<double>());
    // steps whose trailing triangle has at most one 64 x 64 tile per CU fuse the panel into the update
    const int n_cu = std::max(1, ctx->prop.multiProcessorCount);
    RP_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(&k_chol_step),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)STEP_LDS));
    for (int k0 = 0; k0 < nr; k0 += CB) {
        const int nb = std::min(CB, nr - k0);
        double *dinv = d_dinv.as<double>() + (int64_t)(k0 / CB) * CB * CB;
        const int below = nr - (k0 + nb);
        if (below > 0) {
            const int nt = (int)ceil_div(below, 64);
            dim3 g(nt, nt);
            if ((int64_t)nt * (nt + 1) / 2 <= n_cu) {
                k_chol_step<<<g, 256, STEP_LDS, st>>>(d_K.as<double>(), ld, nr, k0, dinv, d_K2.as<double>(),
                                                      d_cinfo.as<int>(), dinv + CB * CB, d_wc.as<double>(),
                                                      d_bc.as<double>());
            } else {
                k_chol_trsm_inv<<<ceil_div(below, 32), 256, 0, st>>>(d_K.as<double>(), ld, nr, k0, dinv,
                                                                     d_wc.as<double>(), d_bc.as<double>());
                k_chol_update_potrf<<<g, 256, 0, st>>>(d_K.as<double>(), ld, nr, k0, d_K2.as<double>(),
                                                       d_cinfo.as<int>(), dinv + CB * CB, d_wc.as<double>(),
                                                       d_bc.as<double>());
            }
        }
    }
    // the multiplier and the backward solve, one launch per diagonal block (skipped on the device
    // when a pivot failed); then expansion, results and the pivot flag in one round trip
    // ws[31]: the ticket counter and one flag per block, cleared when (re)allocated; flags compare
    // with a per-solve epoch, so they are never cleared between solves
    // Two such regions, each half of the buffer (its offset changes only when the buffer is reallocated,
    // which clears both): the second serves the solves queued on the side stream.
    DBuf &d_sync = fit->ws[31];
    const int nblk = (int)ceil_div(nr, CB);
    const void *sync_before = d_sync.p;
    const size_t sync_bytes_before = d_sync.bytes;
    RP_CHECK(d_sync.grow(ctx, 2 * sizeof(unsigned) * (size_t)(1 + nblk)));
    if (d_sync.p != sync_before || d_sync.bytes != sync_bytes_before) {
        const int words = (int)(d_sync.bytes / sizeof(unsigned));
        k_fill_u32<<<ceil_div(words, 256), 256, 0, st>>>(d_sync.as<unsigned>(), words, 0u);
        fit->bs_base = 0;
        fit->bs_epoch = 0;
        fit->bs2_base = 0;
        fit->bs2_epoch = 0;
    }
    unsigned *sync2 = d_sync.as<unsigned>() + d_sync.bytes / sizeof(unsigned) / 2;
    static const unsigned spin_limit = [] {
        const char *v = std::getenv("RP_BACKSUB_SPIN");
        return v ? (unsigned)std::strtoul(v, nullptr, 10) : BACKSUB_SPIN;
    }();
    // every ticketed launch takes nblk tickets and a fresh epoch for its flags
    k_backsub<<<nblk, 256, 0, st>>>(d_K.as<double>(), ld, nr, d_dinv.as<double>(), d_wc.as<double>(),
                                    d_bc.as<double>(), con->q_tot, d_x.as<double>(), d_sync.as<unsigned>(),
                                    fit->bs_base, ++fit->bs_epoch, spin_limit, d_cinfo.as<int>() + 1,
                                    d_cinfo.as<int>());
    fit->bs_base += (unsigned)nblk;
    delete kts;
    kts = nullptr;
    // refinement: fp64 residual against the kept basis, correction through the same factor; a step whose
    // predecessor left an error bound of at most RP_REFINE_TARGET_E returns at once on the device.  The
    // first REFINE_EAGER steps are queued with the solve; further ones (up to the cap) only when the
    // published state after them still asks for a step.
    const int refine_cap = fit->opt.refine_steps > 0 ? fit->opt.refine_steps : RP_REFINE_STEPS_DEFAULT;
    auto refine_step = [&]() {
        double *u = d_rv.as<double>(), *r = u + N, *r1 = r + N, *z = r1 + nr, *dx = z + nr, *r2 = dx + nr + 1;
        residual_pass();
        k_ref_rhs<<<1, 256, 0, st>>>(r, dc_ptr, dc_src, dc_mult, dc_w, con->diag_add ? dc_diag : nullptr,
                                     con->rhs_add ? dc_rhs : nullptr, con->q_tot, nr, d_x.as<double>(), r1, r2,
                                     d_cinfo.as<int>(), d_rs);
        k_fwdsub<<<nblk, 256, 0, st>>>(d_K.as<double>(), ld, nr, d_dinv.as<double>(), r1, z, d_sync.as<unsigned>(),
                                       fit->bs_base, ++fit->bs_epoch, spin_limit, d_cinfo.as<int>() + 1,
                                       d_cinfo.as<int>(), d_rs);
        fit->bs_base += (unsigned)nblk;
        k_backsub<<<nblk, 256, 0, st>>>(d_K.as<double>(), ld, nr, d_dinv.as<double>(), d_wc.as<double>(), z, 0.0,
                                        dx, d_sync.as<unsigned>(), fit->bs_base, ++fit->bs_epoch, spin_limit,
                                        d_cinfo.as<int>() + 1, d_cinfo.as<int>(), r2, &d_rs->go);
        fit->bs_base += (unsigned)nblk;
        k_ref_update<<<1, 256, 0, st>>>(d_x.as<double>(), dx, nr, RP_REFINE_TARGET_E, d_cinfo.as<int>(), d_rs);
    };
    // sensitivity to the fp64 rounding of the system: one more solve with the same factor, from the
    // solution before refinement (the estimate needs only its magnitudes).  With refinement it runs on
    // the side stream, beside the refinement's passes over the kept basis, with its own vectors and
    // ticket region; the main stream waits for it before publishing.
    {
        double *sz = d_rv.as<double>() + 2 * (size_t)N + 3 * (size_t)nr + 2, *szz = sz + nr, *sdx = szz + nr;
        {
            KTimer kts2(ctx, KT_SOLVE);
            k_sens_rhs<<<nr, 256, 0, st>>>(d_K2.as<double>(), ld, nr, d_x.as<double>(), SENS_EPS, sz,
                                           d_cinfo.as<int>(), d_sens);
        }
        hipStream_t ss = st;
        unsigned *ssync = d_sync.as<unsigned>();
        uint32_t *sbase = &fit->bs_base, *sepoch = &fit->bs_epoch;
        if (refine) {
            RP_CHECK(ctx->ensure_side());
            RP_HIP(ctx, hipEventRecord(ctx->ev_fork, st));
            RP_HIP(ctx, hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
            ss = ctx->side;
            ssync = sync2;
            sbase = &fit->bs2_base;
            sepoch = &fit->bs2_epoch;
        }
        k_fwdsub<<<nblk, 256, 0, ss>>>(d_K.as<double>(), ld, nr, d_dinv.as<double>(), sz, szz, ssync, *sbase,
                                       ++*sepoch, spin_limit, d_cinfo.as<int>() + 1, d_cinfo.as<int>(), d_sens);
        *sbase += (unsigned)nblk;
        k_backsub<<<nblk, 256, 0, ss>>>(d_K.as<double>(), ld, nr, d_dinv.as<double>(), d_wc.as<double>(), szz, 0.0, sdx,
                                        ssync, *sbase, ++*sepoch, spin_limit, d_cinfo.as<int>() + 1,
                                        d_cinfo.as<int>(), nullptr, &d_sens->go);
        *sbase += (unsigned)nblk;
        k_sens_max<<<1, 256, 0, ss>>>(sdx, nr, d_cinfo.as<int>(), h_out_dev + 4 + N + n);
        if (refine) RP_HIP(ctx, hipEventRecord(ctx->ev_join, ctx->side));
    }
    int refine_queued = 0;
    if (refine) {
        {
            KTimer ktr(ctx, KT_REFINE);
            for (; refine_queued < std::min(refine_cap, REFINE_EAGER); ++refine_queued) refine_step();
        }
        RP_HIP(ctx, hipStreamWaitEvent(st, ctx->ev_join, 0));
    }
    k_publish<<<ceil_div(N + 1, 256), 256, 0, st>>>(d_x.as<double>(), dc_exp, N, n, d_cinfo.as<int>(), d_rs, h_out_dev);
    RP_HIP(ctx, hipGetLastError());
    RP_HIP(ctx, hipEventRecord(e1, st));
    RP_HIP(ctx, hipStreamSynchronize(st));
    if (refine && h_out[0] == 0.0 && h_out[5 + N + n] != 0.0 && refine_queued < refine_cap) {
        // the eager steps left the bound above the target: the rest of the allowed steps
        {
            KTimer ktr(ctx, KT_REFINE);
            for (; refine_queued < refine_cap; ++refine_queued) refine_step();
        }
        k_publish<<<ceil_div(N + 1, 256), 256, 0, st>>>(d_x.as<double>(), dc_exp, N, n, d_cinfo.as<int>(), d_rs,
                                                        h_out_dev);
        RP_HIP(ctx, hipGetLastError());
        RP_HIP(ctx, hipEventRecord(e1, st));
        RP_HIP(ctx, hipStreamSynchronize(st));
    }
    const int cinfo = (int)h_out[0];
    fit->solver = RP_SOLVER_CHOLESKY;
    if (cinfo == 0) {
        std::copy(h_out + 1, h_out + 1 + N, q_out);
        if (solution_out) std::copy(h_out + 1 + N, h_out + 1 + N + n, solution_out);
        fit->refine_steps_done = refine ? (int32_t)h_out[1 + N + n] : 0;
        fit->refine_last_dq = refine ? h_out[2 + N + n] : 0.0;
        fit->refine_bound = refine ? h_out[3 + N + n] : -1.0;
        fit->sensitivity = h_out[4 + N + n];
    } else {
        fit->sensitivity = -1.0;   // not estimated after a fallback
    }
    fit->lu_fallback = cinfo != 0;
    if (cinfo < 0) {
        // a backward-substitution wait gave up: start the tickets and flags afresh for the next solve
        const int words = (int)(d_sync.bytes / sizeof(unsigned));
        k_fill_u32<<<ceil_div(words, 256), 256, 0, st>>>(d_sync.as<unsigned>(), words, 0u);
        fit->bs_base = 0;
        fit->bs_epoch = 0;
        fit->bs2_base = 0;
        fit->bs2_epoch = 0;
    }
    if (cinfo != 0) {
        // indefinite or singular A (or a triangular-solve wait that gave up): LU with partial pivoting on
        // the bordered system; on an exactly zero pivot the minimum-norm least-squares solution, as the
        // reference's lstsq branch (repeat.py:2306-2310).  The refined form then refines that solution
        // with the same corrections against the kept fp64 basis, each solved with the fallback's factors.
        int rc = lu_solve_augmented(ctx, d_K2.as<double>(), ld, n, d_piv.as<int>(), d_info.as<int>(),
                                    d_x.as<double>(), st);
        fit->solver = RP_SOLVER_LU;
        DBuf &d_W = fit->ws[11], &d_jw = fit->ws[14];
        double *jy = nullptr, *jcoef = nullptr, *jscal = nullptr;
        if (rc == RP_ESINGULAR) {
            fit->solver = RP_SOLVER_LSTSQ;
            const int ne = n + (n & 1);
            const size_t pair_bytes = round_up((int64_t)sizeof(JacPair) * (ne / 2), 256);
            RP_CHECK(d_W.grow(ctx, sizeof(double) * (size_t)n * n));
            RP_CHECK(d_jw.grow(ctx, pair_bytes + sizeof(double) * (2 * (size_t)n + 4) + 256));
            JacPair *pairs = d_jw.as<JacPair>();
            jy = reinterpret_cast<double *>(d_jw.as<char>() + pair_bytes);
            jcoef = jy + n;
            jscal = jcoef + n;
            int *d_rot = reinterpret_cast<int *>(jscal + 2);
            // the bordered system again (the LU overwrote its copy)
            k_assemble<<<ceil_div((int64_t)n * ncols, 256), 256, 0, st>>>(
                fit->d_gram.as<double>(), NPg, N, dc_ptr, dc_src, dc_mult, dc_w, con->diag_add ? dc_diag : nullptr,
                con->rhs_add ? dc_rhs : nullptr, con->q_tot, nr, d_K.as<double>(), d_K2.as<double>(), ld,
                d_wc.as<double>(), d_bc.as<double>(), d_cinfo.as<int>(), d_rs);
            int h_rot = 0;
            rc = jacobi_eigen(ctx, d_K2.as<double>(), ld, n, d_W.as<double>(), pairs, d_rot, &h_rot, jscal, st);
            if (rc == RP_OK) {
                // right-hand side: column n of the bordered system, its last row q_tot
                k_gather_rhs<<<ceil_div(n, 256), 256, 0, st>>>(d_K2.as<double>() + n, ld,
                                                              d_K2.as<double>() + (int64_t)(n - 1) * ld + n, n, jy,
                                                              nullptr, nullptr);
                k_pinv_coef<<<n, 256, 0, st>>>(d_W.as<double>(), d_K2.as<double>(), ld, n, jy, jscal, jcoef, nullptr,
                                               nullptr);
                k_pinv_combine<<<ceil_div(n, 256), 256, 0, st>>>(d_W.as<double>(), n, jcoef, d_x.as<double>(), nullptr,
                                                                 nullptr);
                RP_HIP(ctx, hipGetLastError());
                std::fprintf(stderr, "Warning: Matrix singular, using least squares fallback.\n");
            }
        }
        if (rc != RP_OK) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return rc;
        }
        fit->refine_steps_done = 0;
        fit->refine_last_dq = 0.0;
        fit->refine_bound = -1.0;
        if (refine) {
            KTimer ktr(ctx, KT_REFINE);
            double *u = d_rv.as<double>(), *r = u + N, *r1 = r + N, *z = r1 + nr, *dx = z + nr, *r2 = dx + nr + 1;
            (void)z;
            const int *ci = d_cinfo.as<int>();
            const int steps = fit->opt.refine_steps > 0 ? fit->opt.refine_steps : RP_REFINE_STEPS_DEFAULT;
            k_ref_reset<<<1, 1, 0, st>>>(d_cinfo.as<int>(), d_rs);
            for (int it = 0; it < steps; ++it) {
                residual_pass();
                k_ref_rhs<<<1, 256, 0, st>>>(r, dc_ptr, dc_src, dc_mult, dc_w, con->diag_add ? dc_diag : nullptr,
                                             con->rhs_add ? dc_rhs : nullptr, con->q_tot, nr, d_x.as<double>(), r1, r2,
                                             ci, d_rs);
                if (fit->solver == RP_SOLVER_LU) {
                    k_lu_getrs<<<1, 1024, sizeof(double) * n, st>>>(d_K2.as<double>(), ld, n, d_piv.as<int>(), r1, r2,
                                                                     dx, ci, d_rs);
                } else {
                    k_gather_rhs<<<ceil_div(n, 256), 256, 0, st>>>(r1, 1, r2, n, jy, ci, d_rs);
                    k_pinv_coef<<<n, 256, 0, st>>>(d_W.as<double>(), d_K2.as<double>(), ld, n, jy, jscal, jcoef, ci,
                                                   d_rs);
                    k_pinv_combine<<<ceil_div(n, 256), 256, 0, st>>>(d_W.as<double>(), n, jcoef, dx, ci, d_rs);
                }
                k_ref_update<<<1, 256, 0, st>>>(d_x.as<double>(), dx, nr, RP_REFINE_TARGET_E, ci, d_rs);
            }
        }
        k_publish<<<ceil_div(N + 1, 256), 256, 0, st>>>(d_x.as<double>(), dc_exp, N, n, d_cinfo.as<int>(), d_rs,
                                                        h_out_dev);
        RP_HIP(ctx, hipGetLastError());
        RP_HIP(ctx, hipEventRecord(e1, st));
        RP_HIP(ctx, hipStreamSynchronize(st));
        std::copy(h_out + 1, h_out + 1 + N, q_out);
        if (solution_out) std::copy(h_out + 1 + N, h_out + 1 + N + n, solution_out);
        if (refine) {
            fit->refine_steps_done = (int32_t)h_out[1 + N + n];
            fit->refine_last_dq = h_out[2 + N + n];
            fit->refine_bound = h_out[3 + N + n];
        }
    }
    if (used_fallback) *used_fallback = fit->solver;
    ctx->collect();
    fit->resolve_timings(false);
    float ms = 0.f;
    RP_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    fit->t_solve_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return RP_OK;
}

}  // namespace rp
