// Constrained least squares on the device (fit_charges steps 4-5, repeat.py:2266-2313).
//
//   A = ATA - n phibar phibar^T,  b = ATb - n phibar Vbar           (repeat.py:2266-2276)
//   symmetry merge of rows/columns and b, group sizes as weights    (repeat.py:2445-2470)
//   RESP / QEq diagonal terms                                       (repeat.py:2511-2535)
//   [[A, w], [w^T, 0]] [q; lambda] = [b; q_tot]                     (repeat.py:2293-2304)
//   solved by LU with partial pivoting (the reference calls LAPACK gesv through
//   scipy.linalg.solve, repeat.py:2307); an exactly zero pivot is reported as
//   RP_ESINGULAR, where the reference falls back to lstsq.
//
// The augmented matrix [K | rhs] (n x (n+1), n = N'+1) is factored in place,
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
    DBuf d_K, d_ptr, d_src, d_mult, d_w, d_diag, d_rhs, d_piv, d_info, d_x, d_exp, d_q;
    RP_CHECK(d_K.alloc(ctx, sizeof(double) * (size_t)n * ld));
    RP_CHECK(d_ptr.alloc(ctx, sizeof(int) * (nr + 1)));
    RP_CHECK(d_src.alloc(ctx, sizeof(int) * std::max(nsrc, 1)));
    RP_CHECK(d_mult.alloc(ctx, sizeof(double) * std::max(nsrc, 1)));
    RP_CHECK(d_w.alloc(ctx, sizeof(double) * nr));
    RP_CHECK(d_diag.alloc(ctx, sizeof(double) * nr));
    RP_CHECK(d_rhs.alloc(ctx, sizeof(double) * nr));
    RP_CHECK(d_piv.alloc(ctx, sizeof(int) * n));
    RP_CHECK(d_info.alloc(ctx, sizeof(int)));
    RP_CHECK(d_x.alloc(ctx, sizeof(double) * n));
    RP_CHECK(d_exp.alloc(ctx, sizeof(int) * N));
    RP_CHECK(d_q.alloc(ctx, sizeof(double) * N));
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
    const int NPg = (int)round_up(N + 2, 128);
    k_assemble<<<ceil_div((int64_t)n * ncols, 256), 256, 0, st>>>(fit->d_gram.as<double>(), NPg, N, d_ptr.as<int>(), d_src.as<int>(),
                                                                  d_mult.as<double>(), d_w.as<double>(),
                                                                  con->diag_add ? d_diag.as<double>() : nullptr,
                                                                  con->rhs_add ? d_rhs.as<double>() : nullptr, con->q_tot, nr,
                                                                  d_K.as<double>(), ld);
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int nb = std::min(NB, n - k0);
        k_panel<<<1, 1024, 0, st>>>(d_K.as<double>(), ld, n, k0, nb, d_piv.as<int>(), d_info.as<int>());
        k_swap_rest<<<ceil_div(ncols, 256), 256, 0, st>>>(d_K.as<double>(), ld, ncols, k0, nb, d_piv.as<int>());
        const int right = ncols - (k0 + nb);
        if (right > 0) {
            k_trsm<<<ceil_div(right, 256), 256, 0, st>>>(d_K.as<double>(), ld, ncols, k0, nb);
            const int down = n - (k0 + nb);
            if (down > 0) {
                dim3 g(ceil_div(right, 64), ceil_div(down, 64));
                k_update<<<g, 256, 0, st>>>(d_K.as<double>(), ld, n, ncols, k0, nb);
            }
        }
    }
    int info = 0;
    RP_HIP(ctx, hipMemcpyAsync(&info, d_info.p, sizeof(int), hipMemcpyDeviceToHost, st));
    RP_HIP(ctx, hipStreamSynchronize(st));
    if (used_fallback) *used_fallback = 0;
    if (info != 0) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return ctx->fail(RP_ESINGULAR, "KKT matrix is exactly singular (zero pivot at row " + std::to_string(info) + ")");
    }
    k_backsub<<<1, 1024, sizeof(double) * n, st>>>(d_K.as<double>(), ld, n, d_x.as<double>());
    k_expand<<<ceil_div(N, 256), 256, 0, st>>>(d_x.as<double>(), d_exp.as<int>(), N, d_q.as<double>());
    RP_HIP(ctx, hipGetLastError());
    RP_HIP(ctx, hipEventRecord(e1, st));
    RP_HIP(ctx, hipMemcpyAsync(q_out, d_q.p, sizeof(double) * N, hipMemcpyDeviceToHost, st));
    if (solution_out) RP_HIP(ctx, hipMemcpyAsync(solution_out, d_x.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    RP_HIP(ctx, hipStreamSynchronize(st));
    float ms = 0.f;
    RP_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    fit->t_solve_ms = ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return RP_OK;
}

}  // namespace rp
