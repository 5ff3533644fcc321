/* luspice.h -- C ABI of the MI355X (gfx950) mixed-precision sparse-LU path.
 *
 * This header adds the batched / resident entry points that the reference
 * has no counterpart for (it factors one matrix per launch and keeps nothing
 * on the device between Newton steps).  The reference-compatible headers sit
 * beside it: lu_gpu.h, lu_gpu_par.h, spmv.h, selector.h, level.h,
 * cpu_ref_lu.h, lu_static.h.
 *
 * Conventions: 0-based COO patterns; return 0 on success, negative LSP_E*
 * codes on failure (lsp_strerror gives the text).  A "batch" holds one or
 * more circuits whose symbolic analysis (COLAMD, pivot sequence, L/U pattern,
 * level sets) is done once on the host at creation; every later call reuses
 * it on the device.  Calls on one batch are not re-entrant; distinct batches
 * may be used from distinct threads with distinct streams.
 */
#ifndef LUSPICE_H
#define LUSPICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LSP_OK = 0,
    LSP_E_ARG = -1,
    LSP_E_ALLOC = -2,
    LSP_E_HIP = -3,
    LSP_E_SINGULAR = -4,      /* structurally or numerically singular */
    LSP_E_STRUCTURE = -5,     /* inconsistent symbolic structure */
    LSP_E_DEVICE = -6,        /* a device wait gave up (spin limit) */
    LSP_E_CAPACITY = -7,      /* a column exceeded the scratch capacity */
    LSP_E_NOT_CONVERGED = -8  /* IR did not reach the tolerance */
};

enum { LSP_PREC_FP32 = 0, LSP_PREC_FP64 = 1, LSP_PREC_FP16 = 2 };

typedef struct {
    int n;
    long nnz;
    const int *row;        /* 0-based, length nnz */
    const int *col;        /* 0-based, length nnz */
    const double *val;     /* values used for the symbolic pivot sequence */
    const int *q;          /* column order (NULL: COLAMD) */
} lsp_circuit_in;

typedef struct {
    double pivot_tol;      /* CPU partial-pivot threshold (splu tol), 1.0 */
    double refactor_ptol;  /* GPU pivot acceptance |piv| >= ptol*max, 1e-3 */
    int ir_outer;          /* restarts (default 7)  */
    int ir_inner;          /* restart length m (default 50) */
    double ir_inner_tol;   /* |g_{j+1}|/||b|| stop (default 1e-10) */
    double ir_outer_tol;   /* true residual stop (default 1e-10) */
    int lds_cap;           /* LDS doubles per wave in the refactor (1024) */
    int blocks_per_cu;     /* persistent grid density (0 = occupancy) */
    int chunk_rows;        /* rows per reduction chunk (0 = auto) */
    int sell_sigma;        /* SELL sort window (0 = whole matrix) */
    int fallback_fp64;     /* 1: stalled IR refactors in fp64 and retries */
    int refactor_blocks;   /* persistent refactor grid (0 = auto) */
    int trisolve_blocks;   /* persistent trisolve grid (0 = auto) */
    int tail_max;          /* largest dense trailing block (0 = 8192, <0 = none) */
    double tail_density;   /* min pattern density of that block (0 = 0.5) */
    int ruiz_sweeps;       /* Ruiz sweeps ahead of FP16 factor storage (0 = 30) */
    int pivot_scaled;      /* 1: choose the pivot sequence on the Ruiz-scaled matrix */
    int poll_backoff;      /* refactor dependency polls: at most this + 1 sleeps apart (0 = 7) */
    double factor_shift;   /* factor-only diagonal shift: the first diagonal entry of each row
                              is factored as a + shift*|a| (0 = none; the residual and SpMV keep
                              the caller's values; test_lu_gpu_f32.cu:354-364) */
    int krylov_spmv;       /* value precision of the Arnoldi products w = A z inside GMRES:
                              0 fp64 (default), 1 fp32, 2 FP16 (values must fit FP16); the
                              residual b - A x is always taken with the fp64 values */
    int repivot;           /* new pivot sequences on the host when a refactor's fixed pivots
                              no longer fit the step's values (the reference's splu picks its
                              pivots anew at every factorization, splu.h:174-185); a circuit
                              given new pivots is refactored again within the same call:
                              0  never: failures are only reported;
                              1  (default) a circuit with a zero pivot gets the pivot sequence
                                 the reference's partial pivoting (tol = pivot_tol) picks on the
                                 step's values; small pivots are left to GMRES-IR and its fp64
                                 fallback (a Newton refactor keeps its pivots, as KLU's does);
                              2  also a circuit with a pivot failing refactor_ptol;
                              3  also whenever that partial pivoting would pick a different
                                 pivot for any column (P equals the reference's at every step;
                                 meant for fp64 factors) */
    int allow_fp16;        /* 1: accept LSP_PREC_FP16 factor storage (Ruiz-scaled).  Default 0:
                              the FP16 factor path once faulted the GPU with no cause shown, so
                              refactor / graph_step refuse FP16 with LSP_E_ARG unless the caller
                              opts in */
    int ir_method;         /* solve: 0 fp64 GMRES-IR (default); 1 fused Richardson steps first
                              (lsp_batch_ir_step, one launch per step for the whole batch, from
                              x = 0, until ||b - A x|| <= ir_outer_tol ||b||), the circuits they
                              leave unconverged after ir_fused_steps are solved by GMRES-IR */
    int ir_fused_steps;    /* most fused steps per solve (0 = 12) */
    int refactor_fp64;     /* with LSP_PREC_FP32: 1 = the refactor runs on fp64 columns (the
                              fp64 factor set) and the fp32 factors the solves read are its
                              rounding, instead of a refactor that reads its earlier columns
                              back in fp32; a better preconditioner for one fp64 refactor */
} lsp_options;

typedef struct {
    int status;            /* LSP_OK or error */
    int bad_pivots;        /* columns failing refactor_ptol (all circuits) */
    int singular_circuits; /* circuits with a zero pivot */
    int first_bad_circuit; /* -1 if none */
    double ms;             /* device time of the refactor launch(es) */
    int repivoted;         /* circuits given a new pivot sequence by this call */
    int pivot_mismatch;    /* columns whose fixed pivot the reference's partial pivoting would
                              not pick on these values (counted when repivot = 3, before any
                              new pivot sequence) */
} lsp_refactor_report;

typedef struct {
    int status;
    int outer;             /* restarts used (max over circuits) */
    int inner_total;       /* inner iterations summed over restarts (max) */
    int converged;         /* circuits that met ir_outer_tol */
    int fallback;          /* circuits that fell back to fp64 factors */
    double max_rel_res;    /* max ||b - A x|| / ||b|| over circuits (INF if one diverged) */
    int singular_fallback; /* fallback circuits whose fp64 refactor met a zero pivot */
} lsp_solve_report;

typedef struct {
    int n;
    long nnz;
    long lnnz, unnz;       /* including the diagonals, as splu counts them */
    long n_updates;        /* scatter count of the left-looking refactor */
    int nlevels, nlv_fwd, nlv_bwd;
    int max_col;
    long sell_slots;
    int tail_T;            /* order of the dense trailing block (0 = none) */
    long tail_updates;     /* left-looking updates replaced by the dense block */
} lsp_circuit_info;

typedef struct {
    double refactor_ms;    /* summed device time of the persistent refactor kernel */
    long refactor_count;
    double fwd_ms, bwd_ms; /* summed device time of forward / backward solves */
    long tri_count;        /* trisolve pairs timed */
    long graph_steps;      /* Newton steps replayed as a HIP graph */
    double dense_ms;       /* summed device time of the dense trailing-block LU */
} lsp_kernel_stats;

typedef struct lsp_batch lsp_batch;

const char *lsp_strerror(int code);
int  lsp_options_default(lsp_options *opt);

int  lsp_batch_create(const lsp_circuit_in *circuits, int ncircuits,
                      const lsp_options *opt, lsp_batch **out);
/* Two-phase creation for callers that admit work by device memory:
 * lsp_batch_prepare runs the host symbolic phase only (no device calls),
 * lsp_batch_estimate_bytes gives the device memory the batch will hold
 * (an upper estimate including fp32 and fp64 factor sets), and
 * lsp_batch_upload allocates and uploads.  lsp_batch_create = both. */
int  lsp_batch_prepare(const lsp_circuit_in *circuits, int ncircuits,
                       const lsp_options *opt, lsp_batch **out);
int  lsp_batch_estimate_bytes(const lsp_batch *b, size_t *bytes);
int  lsp_batch_upload(lsp_batch *b);
/* Device memory currently held by the batch. */
int  lsp_batch_device_bytes(const lsp_batch *b, size_t *bytes);
/* Continuous batching: a prepared batch (host phase done; call
 * lsp_batch_upload) whose circuit i is circuit circuit_of[i] of
 * parts[part_of[i]], moved over without repeating the symbolic phase, so a
 * serving loop can merge newly prepared requests with the ones in flight
 * into one batch (one refactor launch and one GMRES-IR per step for all).
 * Every batch in parts is destroyed by the call, whatever the result (its
 * handle is invalid afterwards); circuits not picked are dropped.  The new
 * batch takes the options of parts[0]. */
int  lsp_batch_adopt(lsp_batch *const *parts, int nparts, const int *part_of,
                     const int *circuit_of, int ncircuits, lsp_batch **out);
/* The circuits lsp_batch_refactor and lsp_batch_solve work on: those with
 * active[c] != 0 (NULL: every circuit).  The others keep their factors and
 * solutions untouched.  graph_step, trisolve and spmv cover every circuit. */
int  lsp_batch_set_active(lsp_batch *b, const int *active);

/* Plain device-memory helpers for callers without their own HIP runtime
 * binding (staging inputs in HBM): allocate, copy (any direction,
 * synchronous), synchronise the device, query free / total memory. */
int  lsp_dev_alloc(void **ptr, size_t bytes);
int  lsp_dev_free(void *ptr);
int  lsp_dev_copy(void *dst, const void *src, size_t bytes);
int  lsp_dev_sync(void);
int  lsp_dev_mem_info(size_t *free_bytes, size_t *total_bytes);
void lsp_batch_destroy(lsp_batch *b);
/* Ruiz equilibration scales of a COO matrix (`sweeps` alternating row and
 * column sweeps, each dividing by the square root of the largest |entry| of
 * R A C in the row / column): R, C of length n.  The scaling of
 * bench/test_lu_gpu_f32.cu:247-269. */
int  lsp_ruiz(int n, long nnz, const int *row, const int *col, const double *val, int sweeps,
              double *R, double *C);
int  lsp_batch_set_stream(lsp_batch *b, void *hip_stream);
int  lsp_batch_info(const lsp_batch *b, int circuit, lsp_circuit_info *info);
int  lsp_batch_sizes(const lsp_batch *b, long *total_n, long *total_nnz);

/* Values for every circuit, concatenated in circuit order, each in its own
 * COO order.  on_device: 1 if vals is a device pointer. */
int  lsp_batch_set_values(lsp_batch *b, const double *vals, int on_device);
/* Numeric refactorization of every circuit (one persistent launch). */
int  lsp_batch_refactor(lsp_batch *b, int prec, lsp_refactor_report *rep);
/* One fused iterative-refinement step for every circuit, in one launch:
 * r = b - A x (fp64 values), x += Q U^-1 L^-1 P r with the current factors
 * (fp64 accumulation); rel_res (optional, one per circuit) gets ||r|| / ||b||
 * of the x passed in.  The reference runs the same step as one kernel for one
 * matrix (src/mp/fusion_bench.cu:119-163) and adds the correction on the
 * host. */
int  lsp_batch_ir_step(lsp_batch *b, const double *rhs, double *x, int on_device, double *rel_res);
/* Smallest and largest |u_kk| of one circuit's current factors: the input of
 * the pivot guard (min |pivot| < 1e-9 max |A| -> fp64 factors), which the
 * reference states in docs/02_precision_budget.md:82-83. */
int  lsp_batch_pivot_range(lsp_batch *b, int circuit, double *min_abs, double *max_abs);
/* Direct solve x = Q U^-1 L^-1 P b with the current factors. */
int  lsp_batch_trisolve(lsp_batch *b, const double *rhs, double *x, int on_device);
/* fp64 GMRES-IR with the current factors as right preconditioner. */
int  lsp_batch_solve(lsp_batch *b, const double *rhs, double *x, int on_device,
                     lsp_solve_report *rep);
/* y = A x with fp64 (prec = LSP_PREC_FP64) or fp32 values, SELL-C-64. */
int  lsp_batch_spmv(lsp_batch *b, const double *x, double *y, int on_device, int prec);
/* Copy the current factors of one circuit to the host in the splu layout
 * (L unit diagonal first, U diagonal last, rows = pivot positions, rows
 * sorted inside each column) as fp64.  Arrays sized from lsp_batch_info. */
int  lsp_batch_export_factors(lsp_batch *b, int circuit, int *pinv,
                              int *Lp, int *Li, double *Lx,
                              int *Up, int *Ui, double *Ux);
/* Device pointers of the vector arenas (length total_n) for zero-copy use. */
int  lsp_batch_device_vectors(lsp_batch *b, double **rhs, double **x);
/* One Newton step on the resident arenas as a replayed HIP graph: refactor
 * in `prec`, one GMRES cycle of up to ir_inner iterations, true residual.
 * Recorded on first use (and when prec changes); no host synchronisation
 * inside.  Circuits not converged after the cycle are finished by the eager
 * solve.  No counterpart in the reference (its graph-free design:
 * spmv.h:5-7); this is the capturable refactor -> trisolve -> IR sequence. */
int  lsp_batch_graph_step(lsp_batch *b, int prec, lsp_refactor_report *rrep,
                          lsp_solve_report *srep);

/* Copies total_n doubles into the resident right-hand-side arena / out of
 * the resident solution arena, ordered on the batch's stream (src / dst may
 * be host or device memory; the call returns when the copy is complete). */
int  lsp_batch_set_rhs(lsp_batch *b, const double *rhs);
int  lsp_batch_get_x(lsp_batch *b, double *x);
/* Device time of the last refactor / trisolve launches (ms, hipEvent). */
int  lsp_batch_last_times(const lsp_batch *b, double *refactor_ms,
                          double *fwd_ms, double *bwd_ms);

/* Text of the last error recorded by the batch ("" if none). */
const char *lsp_batch_last_error(const lsp_batch *b);
/* Running per-kernel device times since creation (or the last reset). */
int  lsp_batch_kernel_stats(lsp_batch *b, lsp_kernel_stats *st, int reset);

/* ---- host-side pieces exposed for tests and tools ---------------------- */
/* COLAMD with all knobs zero on an n x n pattern (0-based COO). */
int  lsp_colamd(int n, long nnz, const int *row, const int *col, int *q);
/* CPU symbolic LU with the reference's pivoting rule; factors in the splu
 * layout.  Returns -1 ok, k singular column, -2 arg, -3 arrays too small. */
int  lsp_symbolic_factor(int n, long nnz, const int *row, const int *col,
                         const double *val, const int *q, double tol,
                         int *pinv, int *Lp, int *Li, double *Lx, long lcap,
                         int *Up, int *Ui, double *Ux, long ucap,
                         long *lnnz, long *unnz);
/* Synthetic MNA circuits.  kind 0: 25x40 grid + 20 sources (n = 1020);
 * kind 1: 316x316 power grid + 50 pads (n = 99906); kind 2: circuit "like
 * kind 0" with n unknowns and long-range fraction lr; kind 3: precision-sweep
 * circuit with n unknowns, ~6 nonzeros per row and conductances spread over
 * up to 15 decades (lr ignored).  Two calls: first
 * with row == NULL to get n and nnz, then with arrays. */
int  lsp_mna_generate(int kind, int n_req, double lr, uint64_t seed, int *n,
                      long *nnz, int *row, int *col, double *val, double *rhs);
/* Newton perturbation: val = val0 * lognormal(0, sigma), rhs ~ U[-1,1]. */
void lsp_mna_perturb_values(uint64_t seed, int step, double sigma, long nnz,
                            const double *val0, double *val, int n, double *rhs);

#ifdef __cplusplus
}
#endif
#endif
