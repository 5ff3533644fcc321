/* lu_gpu_par.h -- single-workgroup sparse LU with the reference's launch
 * contract (src/lu_gpu_par.h:14-58).
 *
 *   lu_factor_par_kernel <<<1, 256, LU_PAR_SHARED_BYTES(n)>>>
 *       Same inputs/outputs as lu_factor_kernel (lu_gpu.h).  One workgroup:
 *       the depth-first reach of each column is built by one lane, the
 *       column updates, pivot search (wave64 reductions, first maximum in
 *       reach order) and L/U writes run over the 256 lanes.  Because each
 *       x[i] receives its updates in the same order as the sequential
 *       algorithm, P, L, U and the values equal lu_factor_kernel's bit for
 *       bit.  Workspace: xi[n + 4*LU_PAR_CAP] (only xi[0..n) is used) and the
 *       dynamic LDS (x[n] doubles + n ints).  xg is unused.
 *   trisolve_par_kernel <<<1,256>>>  fp64, result equal to trisolve_kernel.
 *
 * LU_PAR_CAP keeps the reference's value so that callers size xi as before;
 * no frontier cap applies here (singular_out is never -3).
 */
#ifndef DSH_LU_GPU_PAR_H
#define DSH_LU_GPU_PAR_H

#define LU_PAR_CAP 1024
#define LU_PAR_SHARED_BYTES(n) ((size_t)(n) * 12 + (size_t)256)

#ifdef __HIPCC__
#include <hip/hip_runtime.h>

__global__ void lu_factor_par_kernel(
    int n, const int *Ap, const int *Ai, const double *Ax, const int *diagrow,
    int *Lp, int *Li, double *Lx, long Lcap, int *Up, int *Ui, double *Ux, long Ucap,
    int *pinv, int *xi, double *xg, int *lnnz, int *unnz, int *singular_out);

__global__ void trisolve_par_kernel(
    int n, const int *Lp, const int *Li, const double *Lx,
    const int *Up, const int *Ui, const double *Ux,
    const int *pinv, const int *q, const double *b, double *x,
    int *work_i, double *work_d);
#endif

#ifdef __cplusplus
extern "C" {
#endif

int lu_factor_par_launch(int n, const int *Ap, const int *Ai, const double *Ax,
                         const int *diagrow, int *Lp, int *Li, double *Lx, long Lcap, int *Up,
                         int *Ui, double *Ux, long Ucap, int *pinv, int *xi, int *lnnz,
                         int *unnz, int *singular_out, void *stream);
int trisolve_par_launch(int n, const int *Lp, const int *Li, const double *Lx, const int *Up,
                        const int *Ui, const double *Ux, const int *pinv, const int *q,
                        const double *b, double *x, int *work_i, double *work_d, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* DSH_LU_GPU_PAR_H */
