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
// Cholesky: right-looking, 64-column blocks: diagonal block factored in LDS by one
// workgroup, the panel below by a triangular solve, the trailing lower triangle by a
// rank-64 update.  The two right-hand sides are solved by one workgroup, blocked.
//
// LU fallback: the augmented matrix [K | rhs] (n x (n+1), n = N'+1) is factored in place,
// right-looking and blocked by NB columns: one-workgroup panel factorisation,
// row interchanges applied to the other columns, a triangular solve for the U
// block row and a rank-NB update of the trailing matrix.  The rhs column rides
// along, so it holds L^-1 P b at the end; a back substitution finishes.
#include <algorithm>
#include <cmath>

#include "rp_fit_common.h"

namespace rp {

namespace {

constexpr int NB = 32;

__global__ __launch_bounds__(256) void k_assemble(const double *__restrict__ gram, int NP, int N,
                                                  const int *__restrict__ red_ptr, const int *__restrict__ red_src,
                                                  const double *__restrict__ red_mult, const double *__restrict__ w,
                                                  const double *__restrict__ diag_add, const double *__restrict__ rhs_add,
                                                  double q_tot, int nr, double *__restrict__ K, int ld) {
    const int n = nr + 1;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
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

// Uniform broadcast of lane l's double (l a compile-time constant after unrolling: v_readlane).
__device__ __forceinline__ double readlane_d(double x, int l) {
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Cholesky of one 64x64 diagonal block by a single wave, lane t holding row t in registers.
// A pivot at or below 1e-13 of the original diagonal entry counts as a failure: the system is then
// left to the pivoted LU, which reports exact singularity as the reference's LAPACK call does.
__global__ __launch_bounds__(64) void k_potrf_diag(double *__restrict__ A, int ld, int n, int k0,
                                                   const double *__restrict__ orig, int *__restrict__ info) {
    const int nb = min(CB, n - k0);
    const int t = threadIdx.x;
    double r[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) r[c] = (t < nb && c <= t) ? A[(int64_t)(k0 + t) * ld + k0 + c] : 0.0;
    int bad = 0;
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        if (j < nb) {
            const double d = readlane_d(r[j], j);
            const double od = orig[(int64_t)(k0 + j) * ld + k0 + j];
            double piv;
            if (!(d > 1e-13 * fabs(od))) {
                if (!bad) bad = k0 + j + 1;
                piv = 1.0;
            } else {
                piv = sqrt(d);
            }
            if (t == j)
                r[j] = piv;
            else if (t > j)
                r[j] = r[j] / piv;
            const double ltj = r[j];
#pragma unroll
            for (int c = j + 1; c < CB; ++c) {
                if (c < nb) {
                    const double lcj = readlane_d(r[j], c);
                    if (t >= c) r[c] -= ltj * lcj;
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < CB; ++c)
        if (t < nb && c <= t) A[(int64_t)(k0 + t) * ld + k0 + c] = r[c];
    if (t == 0 && bad && info[0] == 0) info[0] = bad;
}

// L21 = A21 L11^{-T}: one thread per row below the block.
__global__ __launch_bounds__(256) void k_chol_trsm(double *__restrict__ A, int ld, int n, int k0) {
    __shared__ double L[CB][CB + 1];
    const int nb = min(CB, n - k0);
    for (int e = threadIdx.x; e < nb * nb; e += 256) {
        const int i = e / nb, j = e % nb;
        L[i][j] = j <= i ? A[(int64_t)(k0 + i) * ld + k0 + j] : 0.0;
    }
    __syncthreads();
    const int r = k0 + nb + blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    double x[CB];
    double *row = A + (int64_t)r * ld + k0;
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        if (j < nb) {
            double v = row[j];
            for (int q = 0; q < j; ++q) v -= x[q] * L[j][q];
            x[j] = v / L[j][j];
            row[j] = x[j];
        }
    }
}

// A22 -= L21 L21^T on the lower triangle (64x64 tiles with tile row >= tile col).
__global__ __launch_bounds__(256) void k_chol_update(double *__restrict__ A, int ld, int n, int k0) {
    const int nb = min(CB, n - k0);
    const int r0 = k0 + nb;
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (tj > ti) return;
    __shared__ double Ls[64][CB + 1];
    __shared__ double Us[CB][64];
    const int tr = r0 + ti * 64, tc = r0 + tj * 64;
    const int t = threadIdx.x;
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

// Solves L L^T Y = [w b] (columns cw, cb of A), then the Schur step for the Lagrange row.
// One workgroup; the right-hand sides live in global memory (L1/L2 coherent within the block).
__global__ __launch_bounds__(256) void k_chol_solve(const double *__restrict__ A, int ld, int n, const double *__restrict__ w,
                                                      const double *__restrict__ b, double q_tot, double *__restrict__ yw,
                                                      double *__restrict__ yb, double *__restrict__ x) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int r = t; r < n; r += 256) {
        yw[r] = w[r];
        yb[r] = b[r];
    }
    __syncthreads();
    // forward: L z = rhs, 64-row blocks
    for (int j0 = 0; j0 < n; j0 += CB) {
        const int nb = min(CB, n - j0);
        if (wave == 0) {
            double zw = lane < nb ? yw[j0 + lane] : 0.0, zb = lane < nb ? yb[j0 + lane] : 0.0;
            // lane's row of the diagonal block in registers (fully unrolled loop keeps it there)
            double lr[CB];
#pragma unroll
            for (int j = 0; j < CB; ++j) lr[j] = (lane < nb && j < nb) ? A[(int64_t)(j0 + lane) * ld + j0 + j] : 0.0;
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                if (j < nb) {
                    const double d = readlane_d(lr[j], j);
                    const double xw = readlane_d(zw, j) / d, xb = readlane_d(zb, j) / d;
                    if (lane == j) {
                        zw = xw;
                        zb = xb;
                    } else if (lane > j) {
                        zw -= lr[j] * xw;
                        zb -= lr[j] * xb;
                    }
                }
            }
            if (lane < nb) {
                yw[j0 + lane] = zw;
                yb[j0 + lane] = zb;
            }
        }
        __syncthreads();
        for (int r = j0 + nb + t; r < n; r += 256) {
            double sw = 0.0, sb = 0.0;
            const double *row = A + (int64_t)r * ld + j0;
            for (int q = 0; q < nb; ++q) {
                sw += row[q] * yw[j0 + q];
                sb += row[q] * yb[j0 + q];
            }
            yw[r] -= sw;
            yb[r] -= sb;
        }
        __syncthreads();
    }
    // backward: L^T y = z
    for (int j1 = n; j1 > 0; j1 -= CB) {
        const int j0 = max(0, j1 - CB), nb = j1 - j0;
        if (wave == 0) {
            double zw = lane < nb ? yw[j0 + lane] : 0.0, zb = lane < nb ? yb[j0 + lane] : 0.0;
            // lc[j] = L[j0+j][j0+lane]: column 'lane' of the diagonal block
            double lc[CB];
#pragma unroll
            for (int j = 0; j < CB; ++j) lc[j] = (lane < nb && j < nb) ? A[(int64_t)(j0 + j) * ld + j0 + lane] : 0.0;
#pragma unroll
            for (int jj = 0; jj < CB; ++jj) {
                const int j = CB - 1 - jj;
                if (j < nb) {
                    const double d = readlane_d(lc[j], j);
                    const double xw = readlane_d(zw, j) / d, xb = readlane_d(zb, j) / d;
                    if (lane == j) {
                        zw = xw;
                        zb = xb;
                    } else if (lane < j) {
                        zw -= lc[j] * xw;
                        zb -= lc[j] * xb;
                    }
                }
            }
            if (lane < nb) {
                yw[j0 + lane] = zw;
                yb[j0 + lane] = zb;
            }
        }
        __syncthreads();
        for (int r = t; r < j0; r += 256) {
            double sw = 0.0, sb = 0.0;
            for (int q = 0; q < nb; ++q) {
                const double l = A[(int64_t)(j0 + q) * ld + r];
                sw += l * yw[j0 + q];
                sb += l * yb[j0 + q];
            }
            yw[r] -= sw;
            yb[r] -= sb;
        }
        __syncthreads();
    }
    // Schur complement of the Lagrange row
    __shared__ double red[2][256];
    double pw = 0.0, pb = 0.0;
    for (int r = t; r < n; r += 256) {
        pw += w[r] * yw[r];
        pb += w[r] * yb[r];
    }
    red[0][t] = pw;
    red[1][t] = pb;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            red[0][t] += red[0][t + s];
            red[1][t] += red[1][t + s];
        }
        __syncthreads();
    }
    const double lambda = (red[1][0] - q_tot) / red[0][0];
    for (int r = t; r < n; r += 256) x[r] = yb[r] - lambda * yw[r];
    if (t == 0) x[n] = lambda;
}

__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) b[e] = a[e];
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

__global__ void k_expand(const double *__restrict__ x, const int *__restrict__ src, int N, double *__restrict__ q) {
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a < N) q[a] = src[a] >= 0 ? x[src[a]] : 0.0;
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
    DBuf &d_K = fit->ws[12], &d_ptr = fit->ws[13], &d_src = fit->ws[14], &d_mult = fit->ws[15], &d_w = fit->ws[16],
         &d_diag = fit->ws[17], &d_rhs = fit->ws[18], &d_piv = fit->ws[19];
    DBuf &d_info = fit->ws[24], &d_x = fit->ws[25], &d_exp = fit->ws[26], &d_q = fit->ws[27];
    RP_CHECK(d_K.grow(ctx, sizeof(double) * (size_t)n * ld));
    RP_CHECK(d_ptr.grow(ctx, sizeof(int) * (nr + 1)));
    RP_CHECK(d_src.grow(ctx, sizeof(int) * std::max(nsrc, 1)));
    RP_CHECK(d_mult.grow(ctx, sizeof(double) * std::max(nsrc, 1)));
    RP_CHECK(d_w.grow(ctx, sizeof(double) * nr));
    RP_CHECK(d_diag.grow(ctx, sizeof(double) * nr));
    RP_CHECK(d_rhs.grow(ctx, sizeof(double) * nr));
    RP_CHECK(d_piv.grow(ctx, sizeof(int) * n));
    RP_CHECK(d_info.grow(ctx, sizeof(int)));
    RP_CHECK(d_x.grow(ctx, sizeof(double) * n));
    RP_CHECK(d_exp.grow(ctx, sizeof(int) * N));
    RP_CHECK(d_q.grow(ctx, sizeof(double) * N));
    RP_HIP(ctx, hipMemcpyAsync(d_ptr.p, con->red_ptr, sizeof(int) * (nr + 1), hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipMemcpyAsync(d_src.p, con->red_src, sizeof(int) * nsrc, hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipMemcpyAsync(d_mult.p, con->red_mult, sizeof(double) * nsrc, hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipMemcpyAsync(d_w.p, con->weights, sizeof(double) * nr, hipMemcpyHostToDevice, st));
    if (con->diag_add) RP_HIP(ctx, hipMemcpyAsync(d_diag.p, con->diag_add, sizeof(double) * nr, hipMemcpyHostToDevice, st));
    if (con->rhs_add) RP_HIP(ctx, hipMemcpyAsync(d_rhs.p, con->rhs_add, sizeof(double) * nr, hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipMemcpyAsync(d_exp.p, con->expand_src, sizeof(int) * N, hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipMemsetAsync(d_info.p, 0, sizeof(int), st));

    hipEvent_t e0, e1;
    RP_HIP(ctx, hipEventCreate(&e0));
    RP_HIP(ctx, hipEventCreate(&e1));
    RP_HIP(ctx, hipEventRecord(e0, st));
    KTimer *kts = new KTimer(ctx, KT_SOLVE);
    const int NPg = (int)round_up(N + 2, 128);
    k_assemble<<<ceil_div((int64_t)n * ncols, 256), 256, 0, st>>>(fit->d_gram.as<double>(), NPg, N, d_ptr.as<int>(), d_src.as<int>(),
                                                                  d_mult.as<double>(), d_w.as<double>(),
                                                                  con->diag_add ? d_diag.as<double>() : nullptr,
                                                                  con->rhs_add ? d_rhs.as<double>() : nullptr, con->q_tot, nr,
                                                                  d_K.as<double>(), ld);
    // ---- Cholesky of A (leading nr x nr block) + Schur step
    DBuf &d_K2 = fit->ws[20], &d_yw = fit->ws[21], &d_yb = fit->ws[22], &d_wc = fit->ws[23];
    DBuf &d_cinfo = fit->ws[28], &d_bc = fit->ws[29];
    RP_CHECK(d_K2.grow(ctx, sizeof(double) * (size_t)n * ld));
    RP_CHECK(d_yw.grow(ctx, sizeof(double) * nr));
    RP_CHECK(d_yb.grow(ctx, sizeof(double) * nr));
    RP_CHECK(d_cinfo.grow(ctx, sizeof(int)));
    RP_CHECK(d_wc.grow(ctx, sizeof(double) * nr));
    RP_CHECK(d_bc.grow(ctx, sizeof(double) * nr));
    RP_HIP(ctx, hipMemsetAsync(d_cinfo.p, 0, sizeof(int), st));
    k_copy<<<ceil_div((int64_t)n * ld, 256), 256, 0, st>>>(d_K.as<double>(), d_K2.as<double>(), (int64_t)n * ld);
    // the w and b columns before factorisation
    RP_HIP(ctx, hipMemcpy2DAsync(d_wc.p, sizeof(double), d_K.as<double>() + nr, sizeof(double) * ld, sizeof(double), nr,
                                 hipMemcpyDeviceToDevice, st));
    RP_HIP(ctx, hipMemcpy2DAsync(d_bc.p, sizeof(double), d_K.as<double>() + n, sizeof(double) * ld, sizeof(double), nr,
                                 hipMemcpyDeviceToDevice, st));
    for (int k0 = 0; k0 < nr; k0 += CB) {
        const int nb = std::min(CB, nr - k0);
        k_potrf_diag<<<1, 64, 0, st>>>(d_K.as<double>(), ld, nr, k0, d_K2.as<double>(), d_cinfo.as<int>());
        const int below = nr - (k0 + nb);
        if (below > 0) {
            k_chol_trsm<<<ceil_div(below, 256), 256, 0, st>>>(d_K.as<double>(), ld, nr, k0);
            const int nt = (int)ceil_div(below, 64);
            dim3 g(nt, nt);
            k_chol_update<<<g, 256, 0, st>>>(d_K.as<double>(), ld, nr, k0);
        }
    }
    int cinfo = 0;
    RP_HIP(ctx, hipMemcpyAsync(&cinfo, d_cinfo.p, sizeof(int), hipMemcpyDeviceToHost, st));
    RP_HIP(ctx, hipStreamSynchronize(st));
    fit->lu_fallback = cinfo != 0;
    if (cinfo == 0) {
        k_chol_solve<<<1, 256, 0, st>>>(d_K.as<double>(), ld, nr, d_wc.as<double>(), d_bc.as<double>(), con->q_tot,
                                         d_yw.as<double>(), d_yb.as<double>(), d_x.as<double>());
    } else {
        // indefinite or singular A: LU with partial pivoting on the bordered system
        double *K2 = d_K2.as<double>();
        for (int k0 = 0; k0 < n; k0 += NB) {
            const int nb = std::min(NB, n - k0);
            k_panel<<<1, 1024, 0, st>>>(K2, ld, n, k0, nb, d_piv.as<int>(), d_info.as<int>());
            k_swap_rest<<<ceil_div(ncols, 256), 256, 0, st>>>(K2, ld, ncols, k0, nb, d_piv.as<int>());
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
        int info = 0;
        RP_HIP(ctx, hipMemcpyAsync(&info, d_info.p, sizeof(int), hipMemcpyDeviceToHost, st));
        RP_HIP(ctx, hipStreamSynchronize(st));
        if (info != 0) {
            delete kts;
            hipEventDestroy(e0);
            hipEventDestroy(e1);
            return ctx->fail(RP_ESINGULAR, "KKT matrix is exactly singular (zero pivot at row " + std::to_string(info) + ")");
        }
        k_backsub<<<1, 1024, sizeof(double) * n, st>>>(K2, ld, n, d_x.as<double>());
    }
    if (used_fallback) *used_fallback = 0;
    k_expand<<<ceil_div(N, 256), 256, 0, st>>>(d_x.as<double>(), d_exp.as<int>(), N, d_q.as<double>());
    delete kts;
    RP_HIP(ctx, hipGetLastError());
    RP_HIP(ctx, hipEventRecord(e1, st));
    RP_HIP(ctx, hipMemcpyAsync(q_out, d_q.p, sizeof(double) * N, hipMemcpyDeviceToHost, st));
    if (solution_out) RP_HIP(ctx, hipMemcpyAsync(solution_out, d_x.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    RP_HIP(ctx, hipStreamSynchronize(st));
    ctx->collect();
    float ms = 0.f;
    RP_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    fit->t_solve_ms = ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return RP_OK;
}

}  // namespace rp
