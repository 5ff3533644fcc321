/* lu_static.h -- static-structure sparse LU lifecycle for Newton loops
 * (reference: src/lu_static.h:22-92; same struct fields, functions and return
 * codes; the reference implements it as a CPU stub whose dsh_gmres_ir returns
 * NOT_IMPL).
 *
 *   dsh_symbolic   once per sparsity pattern: COO pattern (1-based row/col,
 *                  as the reference's coo_to_csc reads them) + column order
 *                  (NULL = natural order)
 *   dsh_numeric    every Newton step: new values (1-based COO; a later
 *                  duplicate overwrites an earlier one); pin > 0 is added to
 *                  every diagonal entry present in the pattern.  Factors in
 *                  fp64 on the GPU with the reference's partial pivoting
 *                  (tol = 1.0): the first call picks the pivot sequence on the
 *                  host, later calls refactor over the kept structure and
 *                  re-derive the pivots only when splu would pick another
 *                  pivot on the new values, so P equals splu's every call.
 *                  Fills lnnz, unnz, cap, P, min_diag_U, pivot_growth, and
 *                  factor_cache, which points at the fp64 factors in splu's
 *                  lu_factors layout (splu.h:29-41: pinv, L.Lp/Li/Lx with
 *                  the unit diagonal first, U.Up/Ui/Ux with the diagonal
 *                  last, rows as pivot positions, rows ascending per column)
 *   dsh_trisolve   x = Q U^-1 L^-1 P b with the fp64 factors (the
 *                  reference's lu_solve_perm)
 *   dsh_gmres_ir   fp64 GMRES-IR to rtol with fp32 factors (same pivots) as
 *                  right preconditioner; residual_seq (optional, max_outer+1
 *                  entries) receives ||b - A x||/||b|| after each restart
 *
 * Return codes: 0 OK, 1 SINGULAR, 2 ALLOC, 3 ARG / not converged.
 * dsh_symbolic_count counts dsh_symbolic calls in the process.
 * Calls on one dsh_lu_static are not re-entrant.
 */
#ifndef DSH_LU_STATIC_H
#define DSH_LU_STATIC_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int     n;
    long    cnnz;          /* pattern entries */
    long    lnnz;          /* nnz(L) incl. unit diagonal (after the first numeric) */
    long    unnz;          /* nnz(U) incl. diagonal */
    long    cap;           /* lnnz + unnz */
    int    *Ap;            /* host CSC pointers of the pattern (n+1) */
    int    *Ai;            /* host CSC row ids (0-based, sorted per column) */
    float  *Ax;            /* unused by this implementation (factors stay on the device) */
    int    *P;             /* host copy of pinv after dsh_numeric (n) */
    int    *Q;             /* column order (n) */
    double  min_diag_U;    /* min |U(k,k)| of the last factorization */
    double  pivot_growth;  /* max |U| / max |A| of the last factorization */
    int     dev_allocated;
    void   *factor_cache;  /* fp64 factors in splu's lu_factors layout (host) */
    double  pin;           /* diagonal perturbation added by dsh_numeric */
} dsh_lu_static;

int dsh_symbolic(int n, long cnnz, const int *coo_row, const int *coo_col,
                 const int *colamd_q, dsh_lu_static *F);
int dsh_numeric(dsh_lu_static *F, const int *coo_row, const int *coo_col,
                const double *coo_val, int nnz);
int dsh_trisolve(const dsh_lu_static *F, const double *b, double *x);
int dsh_gmres_ir(dsh_lu_static *F, const double *b, double *x,
                 int max_outer, int max_inner, double rtol,
                 double *residual_seq, int *iters);

int dsh_lu_static_init(dsh_lu_static *F);
int dsh_lu_static_alloc(dsh_lu_static *F);
void dsh_lu_static_free(dsh_lu_static *F);
const char *dsh_lu_status_str(int rc);
long dsh_symbolic_count(void);

#ifdef __cplusplus
}
#endif

#endif
