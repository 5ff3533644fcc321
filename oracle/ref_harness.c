/* oracle/ref_harness.c -- ctypes-friendly entry points around the reference's
 * own splu.h, compiled from /root/reference by oracle/Makefile into
 * oracle/_ref/libref.so.  TEST INFRASTRUCTURE ONLY: the fixtures in
 * tests/golden/ are written through this harness, and bench.py may time it as
 * the CPU baseline.  No reference source is copied: splu.h is included from
 * the reference tree at build time.
 *
 * Compiled twice: without LU_FLOAT (fp64, symbols ref64_*) and with LU_FLOAT
 * (fp32 factor storage and arithmetic, symbols ref32_*), matching
 * src/third_party/splu.h:16-22.
 */
#include <stdlib.h>
#include <string.h>
#include "splu.h"

#ifdef LU_FLOAT
#define REFSYM(x) ref32_##x
#else
#define REFSYM(x) ref64_##x
#endif

/* Factor P*A(:,q) = L*U with the reference's lu_factor and copy the factors
 * out.  Returns -1 ok, k singular column, -3 when the caller's arrays are too
 * small (lnnz/unnz still reported). */
int REFSYM(lu_factor_export)(int n, const int *Ap, const int *Ai,
                             const lu_real *Ax, const int *q, double tol,
                             int *pinv, int *Lp, int *Li, lu_real *Lx, long lcap,
                             int *Up, int *Ui, lu_real *Ux, long ucap,
                             long *lnnz, long *unnz)
{
    lu_factors F;
    memset(&F, 0, sizeof(F));
    long cap0 = (long)Ap[n] + 16;
    F.n = n;
    F.pinv = (int *)malloc((size_t)n * sizeof(int));
    F.L.n = F.U.n = n;
    F.L.Lp = (int *)malloc((size_t)(n + 1) * sizeof(int));
    F.L.Li = (int *)malloc((size_t)cap0 * sizeof(int));
    F.L.Lx = (lu_real *)malloc((size_t)cap0 * sizeof(lu_real));
    F.L.Lcap = cap0;
    F.U.Up = (int *)malloc((size_t)(n + 1) * sizeof(int));
    F.U.Ui = (int *)malloc((size_t)cap0 * sizeof(int));
    F.U.Ux = (lu_real *)malloc((size_t)cap0 * sizeof(lu_real));
    F.U.Ucap = cap0;
    int *xi = (int *)malloc((size_t)(2 * n + 1) * sizeof(int));
    lu_real *x = (lu_real *)malloc((size_t)n * sizeof(lu_real));
    int rc = lu_factor(n, Ap, Ai, Ax, q, &F, (lu_real)tol, xi, x);
    if (rc < 0) {
        if (lnnz) *lnnz = F.L.lnnz;
        if (unnz) *unnz = F.U.unnz;
        if (F.L.lnnz > lcap || F.U.unnz > ucap) {
            rc = -3;
        } else {
            memcpy(pinv, F.pinv, (size_t)n * sizeof(int));
            memcpy(Lp, F.L.Lp, (size_t)(n + 1) * sizeof(int));
            memcpy(Up, F.U.Up, (size_t)(n + 1) * sizeof(int));
            memcpy(Li, F.L.Li, (size_t)F.L.lnnz * sizeof(int));
            memcpy(Lx, F.L.Lx, (size_t)F.L.lnnz * sizeof(lu_real));
            memcpy(Ui, F.U.Ui, (size_t)F.U.unnz * sizeof(int));
            memcpy(Ux, F.U.Ux, (size_t)F.U.unnz * sizeof(lu_real));
        }
    }
    free(F.pinv); free(F.L.Lp); free(F.L.Li); free(F.L.Lx);
    free(F.U.Up); free(F.U.Ui); free(F.U.Ux); free(xi); free(x);
    return rc;
}

/* Full factor + permuted solve (the cpu_lu_ref chain, src/cpu_ref_lu.c:15-58)
 * with a growing capacity instead of the n*n preallocation, so that it runs at
 * n ~ 1e5.  Returns -1 ok, k singular column. */
int REFSYM(factor_solve)(int n, const int *Ap, const int *Ai, const lu_real *Ax,
                         const int *q, double tol, const lu_real *b, lu_real *xout,
                         long *lnnz, long *unnz)
{
    lu_factors F;
    memset(&F, 0, sizeof(F));
    long cap0 = (long)Ap[n] * 4 + 16;
    F.n = n;
    F.pinv = (int *)malloc((size_t)n * sizeof(int));
    F.L.n = F.U.n = n;
    F.L.Lp = (int *)malloc((size_t)(n + 1) * sizeof(int));
    F.L.Li = (int *)malloc((size_t)cap0 * sizeof(int));
    F.L.Lx = (lu_real *)malloc((size_t)cap0 * sizeof(lu_real));
    F.L.Lcap = cap0;
    F.U.Up = (int *)malloc((size_t)(n + 1) * sizeof(int));
    F.U.Ui = (int *)malloc((size_t)cap0 * sizeof(int));
    F.U.Ux = (lu_real *)malloc((size_t)cap0 * sizeof(lu_real));
    F.U.Ucap = cap0;
    int *xi = (int *)malloc((size_t)(2 * n + 1) * sizeof(int));
    lu_real *x = (lu_real *)malloc((size_t)n * sizeof(lu_real));
    int rc = lu_factor(n, Ap, Ai, Ax, q, &F, (lu_real)tol, xi, x);
    if (rc < 0) {
        lu_solve_perm(n, &F, q, b, xout);
        if (lnnz) *lnnz = F.L.lnnz;
        if (unnz) *unnz = F.U.unnz;
    }
    free(F.pinv); free(F.L.Lp); free(F.L.Li); free(F.L.Lx);
    free(F.U.Up); free(F.U.Ui); free(F.U.Ux); free(xi); free(x);
    return rc;
}
