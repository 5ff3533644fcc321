/* cpu_ref_lu.h -- CPU fp64 sparse LU + solve (reference: src/cpu_ref_lu.h:17-19).
 *
 * P * A(:,q) = L * U with the reference's pivoting rule (tol = threshold,
 * diagonal row q[k] preferred), then x = Q U^-1 L^-1 P b.  A is the
 * original-order CSC (0-based); q may be NULL (natural order).  Returns -1
 * on success, k >= 0 for a singular column, -2 for a bad n.  Capacity grows
 * as needed (the reference preallocates n*n + 4096 entries).
 * Implementation: luspice_amd/csrc/host/cpu_ref.cpp (the same symbolic LU the
 * GPU path uses for its pivot sequence).
 */
#ifndef CPU_REF_LU_H
#define CPU_REF_LU_H

#ifdef __cplusplus
extern "C" {
#endif

int cpu_lu_ref(int n, const int *Ap, const int *Ai, const double *Ax,
               const int *q, const double *b, double tol,
               double *x, long *lnnz, long *unnz);

#ifdef __cplusplus
}
#endif
#endif
