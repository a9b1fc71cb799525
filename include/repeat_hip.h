/*
 * repeat_hip.h -- C ABI of librepeat_hip.so, the gfx950 (MI355X) HIP
 * implementation of pyREPEAT's charge-fit path:
 *
 *     vdW grid filter -> Ewald basis Phi (real + reciprocal) ->
 *     normal equations -> constrained least squares (KKT)
 *
 * The reference has no FFI: its boundary is a set of Python calls inside
 * repeat.py.  Each entry point below names the reference call it replaces.
 *
 * Conventions
 *   - Every function returns an int status: RP_OK (0) or a negative RP_E*
 *     code.  Nothing calls exit().  rp_last_error() gives the message of the
 *     last failure on a context (or the last global failure when ctx is NULL).
 *   - Host buffers are caller-owned and only read or written during the call.
 *     Device memory is owned by the context and charged against its HBM budget.
 *   - A context binds one device.  Calls on one context must not overlap.
 *   - Units: Bohr, Hartree/e, e.
 */
#ifndef REPEAT_HIP_H
#define REPEAT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define RP_OK 0
#define RP_EINVAL -1     /* bad argument */
#define RP_ENOMEM -2     /* device allocation failed */
#define RP_EHIP -3       /* HIP runtime error */
#define RP_ESINGULAR -4  /* KKT matrix exactly singular (LU zero pivot) */
#define RP_EBUDGET -5    /* request exceeds the context's HBM budget */
#define RP_ESTATE -6     /* call out of order (e.g. solve before accumulate) */
#define RP_EEMPTY -7     /* the vdW filter kept no grid point */

/* ---------------------------------------------------------------- enums */
#define RP_F64 0
#define RP_F32 1

#define RP_ORDER_CUBE 0  /* Gaussian cube text order: x outer, z inner (repeat.py:788-796) */
#define RP_ORDER_FLAT 1  /* flat order f = i + j*nx + k*nx*ny (repeat.py:2086) */

#define RP_PREC_FP64 0   /* plain fp64: fp64 MFMA k-space GEMM, fp64 real space, fp64 Gram */
#define RP_PREC_OZAKI 1  /* split form: reciprocal sum factorized over the lattice axes (fp64),
                            real space in fp64 (erfc as one polynomial), Gram from
                            int8 fixed-point slices on MFMA with exact int32 accumulation,
                            fp64 across chunks and in the KKT solve */
#define RP_PREC_OZAKI_REFINED 2  /* the split form with 3 slices per element, then iterative
                            refinement of the solve: a step forms the residual of the fp64
                            normal equations, Phi^T (V - Phi q) less the centring term, from
                            the fp64 basis kept in HBM (M x (N + 2) doubles), and corrects q
                            and the multiplier with the factors of the split system.  The result
                            converges to the fp64 system's solution.  After each step the error
                            left is bounded by rho / (1 - rho) * max |dq|, with rho the ratio of
                            the last two corrections (RP_REFINE_RHO_PRIOR after the first); steps
                            run while that bound exceeds RP_REFINE_TARGET_E, at most
                            rp_options.refine_steps (rp_fit_solve_report gives the bound). */
#define RP_PREC_FP64_FACT 3  /* fp64 throughout with the split form's factorized reciprocal sum: the fp64
                            basis of RP_PREC_OZAKI, its Gram [Phi | V | 1] on fp64 MFMA, no slices, no
                            refinement; the same normal equations as RP_PREC_FP64 to fp64 rounding,
                            without the K-wide trig tables and GEMM */
#define RP_REFINE_TARGET_E 1e-6
#define RP_REFINE_RHO_PRIOR 1e-2     /* twice the largest contraction measured (5.3e-3, cond(A) 4.9e10) */
#define RP_REFINE_STEPS_DEFAULT 4

/* How rp_fit_solve solved the bordered system (rp_fit_solve's *used_fallback, rp_solve_report.solver). */
#define RP_SOLVER_CHOLESKY 0   /* Cholesky of A and the Schur complement of the Lagrange row */
#define RP_SOLVER_LU 1         /* LU with partial pivoting (a non-positive Cholesky pivot, or a wait that gave up) */
#define RP_SOLVER_LSTSQ 2      /* exactly singular: minimum-norm least squares (repeat.py:2306-2310) */

typedef struct rp_ctx rp_ctx;
typedef struct rp_fit rp_fit;

/* One charge-fit input: the content of a Gaussian cube (read_gauss_cube, repeat.py:770-804). */
typedef struct rp_system {
    int32_t n_grid[3];
    double axis[9];              /* grid step vectors, row i = axis_vector[i] */
    int32_t n_atoms;
    const int32_t *atomic_number;/* Z per atom */
    const double *atom_xyz;      /* n_atoms x 3 */
    const void *esp;             /* nx*ny*nz values */
    int32_t esp_dtype;           /* RP_F64 | RP_F32 */
    int32_t esp_order;           /* RP_ORDER_CUBE | RP_ORDER_FLAT */
} rp_system;

/* Command-line options of repeat.py that reach the fit path (repeat.py:2655-2660, 2680). */
typedef struct rp_options {
    int32_t fit_type;            /* --fit-type: 1 Ewald (default), 0 bare 1/r */
    int32_t precision;           /* RP_PREC_FP64 (--fp64) | RP_PREC_OZAKI | RP_PREC_OZAKI_REFINED | RP_PREC_FP64_FACT */
    double vdw_factor;           /* --vdw-factor, default 1.0 */
    double vdw_max;              /* --vdw-max, default 1000.0 */
    double r_cutoff;             /* --cutoff, default 20.0 */
    int32_t refine_steps;        /* RP_PREC_OZAKI_REFINED: at most this many refinement steps
                                    (<= 0: RP_REFINE_STEPS_DEFAULT); no counterpart in the reference */
    int32_t flags;               /* RP_OPT_* bits, 0 by default; no counterpart in the reference */
} rp_options;

/* RP_PREC_OZAKI_REFINED keeps the fp64 basis of every kept point in HBM (M x (N + 2) doubles) for its
 * refinement when that fits the context's budget and the device's free memory, and otherwise streams:
 * each refinement step recomputes the basis chunk by chunk (one basis pass more per step, one chunk of
 * memory).  Both give the same charges, bitwise.  This flag forces the streamed form. */
#define RP_OPT_STREAM_REFINE 1

/* Sizes of a prepared fit. */
typedef struct rp_fit_info {
    int64_t n_total;             /* nx*ny*nz */
    int64_t n_keep;              /* kept grid points M */
    int64_t n_eff;               /* effective atoms of the filter */
    int32_t n_atoms;             /* N */
    int32_t n_k;                 /* reciprocal vectors K (754 for fit-type 1) */
    int32_t n_shift;             /* real-space shifts (full list) */
    int32_t n_shift_pruned;      /* after AABB pruning */
    double volume;               /* signed cell volume */
    double alpha, sqrt_alpha;
    double v_mean;               /* mean ESP over kept points, subtracted from V */
} rp_fit_info;

/*
 * Constraints of the bordered solve (fit_charges steps 5, repeat.py:2278-2313).
 * The host turns the symmetry file into merge sets: reduced unknown I sums the
 * original atoms red_src[red_ptr[I] .. red_ptr[I+1]) with multiplicities
 * red_mult (the reference's column-then-row merges, repeat.py:2445-2470), and
 * atom a takes the value of reduced unknown expand_src[a] (-1 -> 0.0) after the
 * solve (repeat.py:2472-2489).  Without symmetry: red_* is the identity.
 */
typedef struct rp_constraints {
    double q_tot;                /* --total-charge */
    int32_t n_reduced;           /* N' */
    const int32_t *red_ptr;      /* N'+1 */
    const int32_t *red_src;
    const double *red_mult;
    const double *weights;       /* N': Lagrange row (group sizes) */
    const double *diag_add;      /* N' or NULL: RESP strength / QEq hardness on A[i,i] */
    const double *rhs_add;       /* N' or NULL: RESP strength*target / QEq chi on b[i] */
    const int32_t *expand_src;   /* N */
} rp_constraints;

/* ---------------------------------------------------------------- context */
int rp_ctx_create(int device, uint64_t hbm_budget_bytes, rp_ctx **out);
/* Destroys the context; with fits still alive on it, the last rp_fit_destroy completes it. */
int rp_ctx_destroy(rp_ctx *ctx);
const char *rp_last_error(const rp_ctx *ctx);
int rp_ctx_hbm_in_use(const rp_ctx *ctx, uint64_t *bytes);

/* Kernel timing with HIP events on the context's stream (off by default; enabling resets the
 * counters).  kind: 0 basis (k-space + real space), 1 Gram, 2 filter mask, 3 KKT solve,
 * 4 per-fit tables (grid phases, factorized reciprocal tables, line table), 5 Phi slicing,
 * 6 reduction of the Gram's row-split partials, 7 refinement of the solve (residual passes over
 * the kept basis and the triangular solves of the correction).  total_ms sums the launches of that kind
 * since the last reset. */
int rp_ctx_profile(rp_ctx *ctx, int enable);
/* The same for the kinds whose bit is set (bit k = kind k) only.  Every timed launch adds two event
 * records to the stream, each of which leaves the GPU idle for a few microseconds, so a throughput
 * measurement times only the kinds it reports. */
int rp_ctx_profile_kinds(rp_ctx *ctx, uint32_t kinds);
int rp_ctx_kernel_stats(const rp_ctx *ctx, int kind, double *total_ms, int64_t *launches);

/* ---------------------------------------------------------------- fit path */
/*
 * ChargeCalculator.__init__ (repeat.py:1736-1786): upload the cube, build the
 * 27-image effective atoms and the vdW keep mask on the device
 * (_filter_grid, repeat.py:1922-2113), compact the kept points in ascending
 * flat index with V - mean(V), set up the Ewald sums (EwaldCalculator.__init__,
 * _setup_kspace, repeat.py:1077-1143) and prune the real-space shifts
 * (prepare_pruned_shifts, repeat.py:1708-1731).
 */
int rp_fit_create(rp_ctx *ctx, const rp_system *sys, const rp_options *opt, rp_fit **out);
int rp_fit_destroy(rp_fit *fit);
int rp_fit_info_get(const rp_fit *fit, rp_fit_info *out);

/* Filter results: any pointer may be NULL.  mask: n_total bytes (flat order),
 * flat_idx / v / grid_xyz (M x 3): the kept points as _filter_grid returns them. */
int rp_fit_grid(const rp_fit *fit, uint8_t *mask, int64_t *flat_idx, double *v, double *grid_xyz);

/* Ewald setup: kvecs (K x 3), kcoefs (K), shifts (S x 3), shifts_pruned (S' x 3); NULL skips. */
int rp_fit_ewald(const rp_fit *fit, double *kvecs, double *kcoefs, double *shifts, double *shifts_pruned);

/*
 * fit_charges steps 1-3 (repeat.py:2195-2262): stream Phi tiles over the kept
 * points and accumulate Phi^T Phi, Phi^T V, sum(Phi), sum(V) on the device,
 * in fp64 across tiles, without materialising Phi.
 */
int rp_fit_accumulate(rp_fit *fit);

/* Accumulators of the last rp_fit_accumulate: ata (N x N), atb (N), sum_phi (N), sum_v (1). */
int rp_fit_normal_equations(const rp_fit *fit, double *ata, double *atb, double *sum_phi, double *sum_v);

/*
 * fit_charges steps 4-5 (repeat.py:2266-2313) on the device: centring, symmetry
 * merge, RESP/QEq diagonal terms, bordered [[A, w], [w^T, 0]] solve, expansion to
 * all atoms.  q_out: N charges; solution_out (N'+1, nullable): reduced charges and
 * the Lagrange multiplier.  *used_fallback (nullable) is the RP_SOLVER_* that solved
 * it: the Cholesky; LU with partial pivoting when a Cholesky pivot is not positive (or
 * a triangular-solve wait gave up); when an LU pivot is exactly zero, the minimum-norm
 * least-squares solution, as the reference's np.linalg.lstsq(rcond=None) branch
 * (repeat.py:2306-2310, the same warning on stderr): a Jacobi eigendecomposition of
 * the symmetric bordered matrix, eigenvalues at most 2^-52 (N'+1) max|l| dropped.
 * Under RP_PREC_OZAKI_REFINED the solution of any of the three is refined against
 * the fp64 basis kept by rp_fit_accumulate, each correction solved with the same
 * factors.
 */
int rp_fit_solve(rp_fit *fit, const rp_constraints *con, double *q_out, double *solution_out,
                 int32_t *used_fallback);

/*
 * One whole fit from the device-resident cube of rp_fit_create: vdW filter, Phi streaming and
 * accumulation, constrained solve (the reference's _filter_grid + fit_charges,
 * repeat.py:1922-2113, 2184-2313), with no host-to-device transfer of the grid data.
 */
int rp_fit_refit(rp_fit *fit, const rp_constraints *con, double *q_out);

/* compute_stats (repeat.py:2326-2441): v_coul[m] = (Phi q)[m] for every kept point, streamed: from the
 * basis that RP_PREC_OZAKI_REFINED keeps after rp_fit_accumulate; after an accumulate of the other
 * factorized forms (or a refined one that streamed), the factorized fp64 basis recomputed chunk by
 * chunk; otherwise the fp64 basis of the plain path. */
int rp_fit_phi_matvec(rp_fit *fit, const double *q, double *v_coul);

/*
 * EwaldCalculator.compute_batch_parallel (repeat.py:1145-1197): the basis
 * phi (m x N, row-major, fp64) of the fit's atoms at arbitrary points, with
 * the fit's Ewald setup and pruned shifts.
 */
int rp_fit_phi(rp_fit *fit, int64_t m, const double *grid_xyz, double *phi_out);

/*
 * rp_fit_phi without device allocation (no counterpart in the reference).  For callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  The same basis rows, bitwise, as rp_fit_phi.  Everything on the device goes through a
 * static 16 MiB array per GPU (held until the process ends; its address is looked up when a context is
 * created on that GPU): the fit's k-set, pruned shifts and atoms, then the points in blocks of
 * about 2M / (N + 3) rows.  Limits: N <= RP_PHI_NOALLOC_MAX_ATOMS, K <= RP_PHI_NOALLOC_MAX_K,
 * pruned shifts <= RP_PHI_NOALLOC_MAX_SHIFTS; beyond them RP_EINVAL with the limits in the message.
 * Every call on one GPU uses the same array, so calls that overlap on a GPU (two contexts or threads)
 * race and must be serialised by the caller.
 */
#define RP_PHI_NOALLOC_SCRATCH_BYTES (16u << 20)
#define RP_PHI_NOALLOC_MAX_ATOMS 4096
#define RP_PHI_NOALLOC_MAX_K 16384
#define RP_PHI_NOALLOC_MAX_SHIFTS 4096
int rp_fit_phi_noalloc(rp_fit *fit, int64_t m, const double *grid_xyz, double *phi_out);

/* Device time (ms) of the phases of the last accumulate / solve on this fit:
 * t[0] filter, t[1] accumulate, t[2] solve. */
int rp_fit_timings(const rp_fit *fit, double *t3);

/*
 * Iterative refinement of the last rp_fit_solve (RP_PREC_OZAKI_REFINED; no counterpart in the
 * reference, whose solve is fp64 throughout): the steps that ran and max |dq| (e) of the last
 * correction applied.  0 and 0.0 for the other precisions.
 */
int rp_fit_refine_stats(const rp_fit *fit, int32_t *steps, double *last_correction);

/* How the last rp_fit_solve went (no counterpart in the reference). */
typedef struct rp_solve_report {
    int32_t solver;              /* RP_SOLVER_* */
    int32_t refine_steps;        /* refinement steps run (RP_PREC_OZAKI_REFINED) */
    double last_correction;      /* max |dq| of the last correction (e) */
    double error_bound;          /* refined: estimated max |q - q_converged| left by the steps (e); -1: none */
    int32_t within_target;       /* refined: error_bound <= RP_REFINE_TARGET_E; other forms: 1 */
    int32_t well_conditioned;    /* sensitivity known and <= RP_SENSITIVITY_LIMIT_E */
    double sensitivity;          /* max |dq| (e) when every entry of the solved system moves by 8 ulps
                                    (signs pseudo-random): how far fp64 rounding alone can move these charges,
                                    whatever the implementation; -1 after an LU or least-squares fallback */
} rp_solve_report;
#define RP_SENSITIVITY_LIMIT_E 1e-5
int rp_fit_solve_report(const rp_fit *fit, rp_solve_report *out);

/*
 * Streaming geometry of the basis for n_keep kept points, n_atoms atoms and n_k k-vectors (no
 * counterpart in the reference, whose chunk size follows free RAM, repeat.py:2196-2207): the chunk
 * rows, the padded widths, the Gram's row splits per chunk and the rows of one Gram stage.  Every
 * accumulate allocates and launches from this plan; each row split holds whole stages.
 */
typedef struct rp_chunk_plan_info {
    int64_t rows;                /* kept points per chunk */
    int64_t n_chunks;
    int32_t np;                  /* padded basis width: N + V column + ones column, rounded to 128 */
    int32_t kp;                  /* padded [cos | sin] width of the fp64 k-space GEMM (0: factorized) */
    int32_t splits;              /* Gram row splits per chunk */
    int32_t n_tiles;             /* Gram tiles of the upper triangle */
    int32_t slices;              /* int8 slices per element (0: fp64 Gram) */
    int32_t stage_rows;          /* rows of one Gram stage (k depth) */
} rp_chunk_plan_info;
int rp_chunk_plan(int64_t n_keep, int32_t n_atoms, int32_t n_k, int32_t precision, rp_chunk_plan_info *out);

/*
 * Reload a fit with another system and options (ChargeCalculator.__init__ on a new
 * cube, repeat.py:1736-1786), keeping its device buffers: they only grow, so a
 * fit object can serve a stream of systems without device allocation once it
 * has seen the largest.  Runs the vdW filter like rp_fit_create.
 */
int rp_fit_reset(rp_fit *fit, const rp_system *sys, const rp_options *opt);

/*
 * ChargeCalculator.compute_qeq_charges (repeat.py:2538-2610): the QEq charges of the fit's atoms,
 * from its Ewald setup.  chi (N): electronegativities, j_self (N): idempotentials 2 * hardness (the
 * caller resolves the parameter tables as the reference does).  J_ij for i != j is the Ewald pair
 * term of _ewald_pairwise (repeat.py:2612-2634): erfc(sqrt(alpha) r)/r over the full shift list for
 * every image with r > 1e-10 (no cutoff) plus sum_k w_k cos(k . (R_i - R_j)); fit-type 0 uses
 * 1/|R_i - R_j|.  Rows 1..N-1 of the system are J_{i-1} - J_i = chi_i - chi_{i-1}, row 0 is
 * sum q = q_tot; LU with partial pivoting (RP_ESINGULAR on an exactly singular matrix, where the
 * reference switches to lstsq).  q_out: N charges; j_out (N x N, nullable): the Coulomb matrix.
 */
int rp_fit_qeq(rp_fit *fit, const double *chi, const double *j_self, double q_tot, double *q_out, double *j_out);

/* ---------------------------------------------------------------- serving
 * The reference fits one cube per process (repeat.py:2636-2789).  The serving
 * loop runs many such fits concurrently on one device: submitted requests wait
 * in arrival order and are admitted to a slot as soon as a slot is idle and the
 * HBM budget holds the request's device bound.  Each slot owns a worker thread,
 * a HIP stream and a grow-only fit object, so admitted fits overlap on the
 * device.  A request's charges are bitwise the same as those of rp_fit_create +
 * rp_fit_accumulate + rp_fit_solve on the same inputs.
 */
#define RP_EAGAIN -8      /* rp_poll: the request has not finished yet */

typedef struct rp_server rp_server;

/* One fit request: the system, the options and the constraints of the solve.
 * rp_submit copies everything it points to, so the caller may free it on return. */
typedef struct rp_request {
    rp_system system;
    rp_options options;
    rp_constraints constraints;
} rp_request;

typedef struct rp_result {
    int32_t status;              /* RP_OK or the fit's error code */
    int32_t slot;                /* slot that ran it (-1 if never admitted) */
    int32_t n_atoms;
    int32_t reserved_;
    int64_t n_keep;
    double t_submit, t_admit, t_done;  /* seconds on the server clock (rp_server_now) */
    double device_ms[3];         /* filter, accumulate, solve (HIP events) */
    uint64_t bytes_bound;        /* device bytes the admission reserved for it */
    char error[256];
} rp_result;

typedef struct rp_server_stats {
    int64_t submitted, completed, failed;
    int32_t queued, running, max_running, n_slots;
    uint64_t budget, reserved, peak_reserved;  /* bytes */
} rp_server_stats;

/* n_slots concurrent fits (1..16) on `device` within hbm_budget_bytes of device memory. */
int rp_server_create(int device, int n_slots, uint64_t hbm_budget_bytes, rp_server **out);
/* Finishes every submitted request, then stops the workers and frees the device memory. */
int rp_server_destroy(rp_server *srv);
const char *rp_server_error(const rp_server *srv);
double rp_server_now(const rp_server *srv);
/* Device bytes the admission would reserve for this request (an upper bound of what its fit allocates). */
int rp_request_bytes(const rp_request *req, uint64_t *bytes);
/* Queues a request; RP_EBUDGET if its bound exceeds the whole budget. */
int rp_submit(rp_server *srv, const rp_request *req, uint64_t *ticket);
/* Waits up to wait_ms (negative: no limit) for a request.  On completion fills *res and q_out
 * (n_atoms charges; may be NULL), releases the ticket and returns RP_OK; RP_EAGAIN if it is still
 * queued or running; RP_EINVAL for an unknown ticket. */
int rp_poll(rp_server *srv, uint64_t ticket, int32_t wait_ms, rp_result *res, double *q_out);
int rp_server_stats_get(const rp_server *srv, rp_server_stats *out);

/* ---------------------------------------------------------------- input files
 * Host-side readers (no device work; errors via rp_last_error(NULL)).  Values are parsed
 * multithreaded and bitwise equal to Python's float() on each token.
 */
typedef struct {
    int32_t n_atoms;
    int32_t n_grid[3];
    double origin[3];            /* read and reported; the fit ignores it, as the reference does */
    double axis[9];              /* grid step vectors (rows), Bohr */
    int64_t values_offset;       /* byte offset of the first ESP value */
} rp_cube_header;

/* read_gauss_cube (repeat.py:770-804), header part: atom count, origin, grid and step vectors.
 * A negative atom count is rejected (the reference's reader fails on it too). */
int rp_cube_read_header(const char *path, rp_cube_header *h);

/* read_gauss_cube (repeat.py:770-804), data part: atoms (Z, x y z per line; NULL skips) and the first
 * n_total ESP values in file order (RP_ORDER_CUBE).  Fails on fewer values than the grid needs or on
 * a token float() rejects; values past n_total are checked and ignored, as the reference ignores
 * them.  n_threads <= 0: up to 16 host threads. */
int rp_cube_read(const char *path, const rp_cube_header *h, int32_t *atomic_number, double *atom_xyz,
                 double *esp, int n_threads);

/* The value loop of read_gauss_cube / read_vasp_potential (repeat.py:797, 731): parse every
 * whitespace-separated token of text[0, len), store the first cap into out, total count in *n_found. */
int rp_parse_values(const char *text, int64_t len, double *out, int64_t cap, int n_threads, int64_t *n_found);

#ifdef __cplusplus
}
#endif
#endif /* REPEAT_HIP_H */
