// QEq charge model on the device (ChargeCalculator.compute_qeq_charges, repeat.py:2538-2610, with the
// pair term _ewald_pairwise, repeat.py:2612-2634).
//
// The model needs no ESP: with electronegativities chi and idempotentials J_ii = 2 * hardness, the
// charges solve
//     sum_j q_j = q_tot,     sum_j (J_{i-1,j} - J_{i,j}) q_j = chi_i - chi_{i-1}   (i = 1 .. N-1)
// where, for i != j and fit-type 1,
//     J_ij = sum_{s in shifts} erfc(sqrt(alpha) r)/r [r = |R_i - R_j + s| > 1e-10]      (no cutoff)
//          + sum_k w_k cos(k . (R_i - R_j)),
// over the full shift list and the 754-vector k-set of the fit (fit-type 0: J_ij = 1/|R_i - R_j|).
// cos(k . (R_i - R_j)) = cos(k.R_i) cos(k.R_j) + sin(k.R_i) sin(k.R_j), so the reciprocal part is the
// weighted product C W C^T of the N x 2K table C = [cos(k.R) | sin(k.R)]; it runs as an LDS-tiled fp64
// product, the real part and the diagonal are added in the same tile's epilogue, and the N x N
// system is solved by the device LU with partial pivoting (the reference's linalg.solve; it switches
// to lstsq on an exactly singular matrix, here reported as RP_ESINGULAR).
#include "rp_fit_common.h"

namespace rp {

namespace {

constexpr int QT = 32;   // output tile (QT x QT), 256 threads, 2 x 2 per thread
constexpr int QK = 32;   // depth of one LDS stage

// C[i][k] = cos(k . R_i), C[i][K + k] = sin(k . R_i); w2[k] = w2[K + k] = w_k.
__global__ __launch_bounds__(256) void k_qeq_trig(int N, int K, const double *__restrict__ atoms,
                                                  const double *__restrict__ kvecs, const double *__restrict__ kcoef,
                                                  int ldc, double *__restrict__ C, double *__restrict__ w2) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)N * K) return;
    const int i = (int)(e / K), k = (int)(e % K);
    const double th = (kvecs[3 * k] * atoms[3 * i] + kvecs[3 * k + 1] * atoms[3 * i + 1]) + kvecs[3 * k + 2] * atoms[3 * i + 2];
    double s, c;
    sincos(th, &s, &c);
    C[(int64_t)i * ldc + k] = c;
    C[(int64_t)i * ldc + K + k] = s;
    if (i == 0) {
        w2[k] = kcoef[k];
        w2[K + k] = kcoef[k];
    }
}

struct QeqArgs {
    int N, K2, ldc;
    const double *C, *w2;      // null when fit-type 0
    const double *atoms, *shifts;
    int S;
    double sqrt_alpha;
    int fit_type;
    const double *j_self;
    double *J;                 // N x N
};

__global__ __launch_bounds__(256) void k_qeq_pairs(QeqArgs p) {
    __shared__ double As[QK][QT + 1], Bs[QK][QT + 1];
    const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
    const int i0 = blockIdx.y * QT, j0 = blockIdx.x * QT;
    double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    if (p.C) {
        for (int k0 = 0; k0 < p.K2; k0 += QK) {
            // 32 rows x 32 k of each operand, 4 entries per thread; A carries the weights
            for (int e = t; e < QT * QK; e += 256) {
                const int r = e / QK, kk = e % QK, k = k0 + kk;
                const int ia = i0 + r, jb = j0 + r;
                As[kk][r] = (ia < p.N && k < p.K2) ? p.C[(int64_t)ia * p.ldc + k] * p.w2[k] : 0.0;
                Bs[kk][r] = (jb < p.N && k < p.K2) ? p.C[(int64_t)jb * p.ldc + k] : 0.0;
            }
            __syncthreads();
#pragma unroll 8
            for (int kk = 0; kk < QK; ++kk) {
                const double a0 = As[kk][ty], a1 = As[kk][ty + 16];
                const double b0 = Bs[kk][tx], b1 = Bs[kk][tx + 16];
                acc[0][0] = fma(a0, b0, acc[0][0]);
                acc[0][1] = fma(a0, b1, acc[0][1]);
                acc[1][0] = fma(a1, b0, acc[1][0]);
                acc[1][1] = fma(a1, b1, acc[1][1]);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const int i = i0 + ty + 16 * u, j = j0 + tx + 16 * v;
            if (i >= p.N || j >= p.N) continue;
            double val;
            if (i == j) {
                val = p.j_self[i];
            } else {
                const double dx = p.atoms[3 * i] - p.atoms[3 * j];
                const double dy = p.atoms[3 * i + 1] - p.atoms[3 * j + 1];
                const double dz = p.atoms[3 * i + 2] - p.atoms[3 * j + 2];
                double real = 0.0;
                if (p.fit_type == 0) {
                    const double r = sqrt((dx * dx + dy * dy) + dz * dz);
                    real = r > 1e-10 ? 1.0 / r : 0.0;
                } else {
                    for (int s = 0; s < p.S; ++s) {
                        const double rx = dx + p.shifts[3 * s], ry = dy + p.shifts[3 * s + 1], rz = dz + p.shifts[3 * s + 2];
                        const double r = sqrt((rx * rx + ry * ry) + rz * rz);
                        if (r > 1e-10) real += erfc(p.sqrt_alpha * r) / r;
                    }
                }
                val = real + acc[u][v];
            }
            p.J[(int64_t)i * p.N + j] = val;
        }
}

// [A | b] of the QEq system, row-major with leading dimension ld.
__global__ __launch_bounds__(256) void k_qeq_assemble(int N, const double *__restrict__ J, const double *__restrict__ chi,
                                                      double q_tot, int ld, double *__restrict__ K2) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)N * (N + 1)) return;
    const int i = (int)(e / (N + 1)), c = (int)(e % (N + 1));
    double v;
    if (i == 0) {
        v = c < N ? 1.0 : q_tot;
    } else if (c < N) {
        v = J[(int64_t)(i - 1) * N + c] - J[(int64_t)i * N + c];
    } else {
        v = chi[i] - chi[i - 1];
    }
    K2[(int64_t)i * ld + c] = v;
}

}  // namespace

int run_qeq(rp_fit *fit, const double *chi, const double *j_self, double q_tot, double *q_out, double *j_out) {
    Ctx *ctx = fit->ctx;
    hipStream_t st = ctx->stream;
    const Setup &s = fit->setup;
    const int N = fit->n_atoms;
    if (N <= 0) return ctx->fail(RP_EINVAL, "QEq needs at least one atom");
    const int K = s.fit_type == 1 ? (int)s.kcoefs.size() : 0;
    const int S = (int)(s.shifts.size() / 3);
    const int K2 = 2 * K;
    const int ldc = (int)round_up(std::max(K2, 1), 4);
    const int ld = (int)round_up(N + 1, 4);
    DBuf d_C, d_w2, d_kv, d_kc, d_sh, d_chi, d_js, d_J, d_K, d_piv, d_info, d_x;
    RP_CHECK(d_J.alloc(ctx, sizeof(double) * (size_t)N * N));
    RP_CHECK(d_K.alloc(ctx, sizeof(double) * (size_t)N * ld));
    RP_CHECK(d_chi.alloc(ctx, sizeof(double) * N));
    RP_CHECK(d_js.alloc(ctx, sizeof(double) * N));
    RP_CHECK(d_piv.alloc(ctx, sizeof(int) * N));
    RP_CHECK(d_info.alloc(ctx, sizeof(int)));
    RP_CHECK(d_x.alloc(ctx, sizeof(double) * N));
    RP_CHECK(d_sh.alloc(ctx, sizeof(double) * std::max<size_t>(s.shifts.size(), 3)));
    RP_HIP(ctx, hipMemcpyAsync(d_chi.p, chi, sizeof(double) * N, hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipMemcpyAsync(d_js.p, j_self, sizeof(double) * N, hipMemcpyHostToDevice, st));
    RP_HIP(ctx, hipMemcpyAsync(d_sh.p, s.shifts.data(), sizeof(double) * s.shifts.size(), hipMemcpyHostToDevice, st));
    QeqArgs a{};
    a.N = N;
    a.K2 = K2;
    a.ldc = ldc;
    if (K > 0) {
        RP_CHECK(d_C.alloc(ctx, sizeof(double) * (size_t)N * ldc));
        RP_CHECK(d_w2.alloc(ctx, sizeof(double) * K2));
        RP_CHECK(d_kv.alloc(ctx, sizeof(double) * 3 * K));
        RP_CHECK(d_kc.alloc(ctx, sizeof(double) * K));
        RP_HIP(ctx, hipMemcpyAsync(d_kv.p, s.kvecs.data(), sizeof(double) * 3 * K, hipMemcpyHostToDevice, st));
        RP_HIP(ctx, hipMemcpyAsync(d_kc.p, s.kcoefs.data(), sizeof(double) * K, hipMemcpyHostToDevice, st));
        k_qeq_trig<<<(unsigned)ceil_div((int64_t)N * K, 256), 256, 0, st>>>(N, K, fit->d_atoms.as<double>(), d_kv.as<double>(),
                                                                           d_kc.as<double>(), ldc, d_C.as<double>(),
                                                                           d_w2.as<double>());
        a.C = d_C.as<double>();
        a.w2 = d_w2.as<double>();
    }
    a.atoms = fit->d_atoms.as<double>();
    a.shifts = d_sh.as<double>();
    a.S = S;
    a.sqrt_alpha = s.sqrt_alpha;
    a.fit_type = s.fit_type;
    a.j_self = d_js.as<double>();
    a.J = d_J.as<double>();
    dim3 g((unsigned)ceil_div(N, QT), (unsigned)ceil_div(N, QT));
    k_qeq_pairs<<<g, 256, 0, st>>>(a);
    k_qeq_assemble<<<(unsigned)ceil_div((int64_t)N * (N + 1), 256), 256, 0, st>>>(N, d_J.as<double>(), d_chi.as<double>(),
                                                                                  q_tot, ld, d_K.as<double>());
    RP_HIP(ctx, hipGetLastError());
    if (j_out) RP_HIP(ctx, hipMemcpyAsync(j_out, d_J.p, sizeof(double) * (size_t)N * N, hipMemcpyDeviceToHost, st));
    RP_CHECK(lu_solve_augmented(ctx, d_K.as<double>(), ld, N, d_piv.as<int>(), d_info.as<int>(), d_x.as<double>(), st));
    RP_HIP(ctx, hipMemcpyAsync(q_out, d_x.p, sizeof(double) * N, hipMemcpyDeviceToHost, st));
    RP_HIP(ctx, hipStreamSynchronize(st));
    return RP_OK;
}

}  // namespace rp

extern "C" int rp_fit_qeq(rp_fit *fit, const double *chi, const double *j_self, double q_tot, double *q_out,
                          double *j_out) {
    if (!fit || !chi || !j_self || !q_out) return RP_EINVAL;
    RP_HIP(fit->ctx, hipSetDevice(fit->ctx->device));
    return rp::run_qeq(fit, chi, j_self, q_tot, q_out, j_out);
}
