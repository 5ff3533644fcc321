/* lu_gpu.h -- sparse LU kernels with the reference's launch contract
 * (src/lu_gpu.h:23-46 for the fp64 kernels; the fp32 kernels have no header
 * in the reference, their signatures are those of src/lu_gpu_f32.cu:14-172).
 *
 * All take A as q-order CSC with original row ids and diagrow[k] = q[k]
 * (P*A(:,q) = L*U), produce L (unit diagonal first, rows = pivot positions)
 * and U (diagonal last), pinv, lnnz/unnz, and singular_out = -1 ok, k >= 0
 * singular column, -2 capacity overflow.  Kernels never allocate.
 *
 *   lu_factor_kernel        <<<1,1>>>  fp64; the reference's pivoting rule
 *                           (first strict max in depth-first reach order,
 *                           diagrow preferred when |x| >= max) and its
 *                           structural pattern: P, L and U equal the CPU
 *                           oracle's (cpu_lu_ref) bit for bit.
 *                           Workspace: xi[2n+1], x[n].
 *   trisolve_kernel         <<<1,1>>>  x = Q U^-1 L^-1 P b (q = NULL: natural);
 *                           work_i[n], work_d[2n].
 *   lu_factor_f32_kernel    <<<1,1>>>  fp32 storage and arithmetic (the
 *                           reference kernel's precision, equal to splu built
 *                           with LU_FLOAT); x[n] floats.
 *   trisolve_f32_kernel     <<<1,1>>>  fp32 factors, fp64 accumulation.
 *   trisolve_par_f32_kernel <<<1,256>>> same result, scatter spread over the
 *                           block.
 *
 * The extern "C" launchers below enqueue them on a hipStream_t (passed as
 * void*) for C hosts; HIP hosts may also launch the kernels directly.
 * The fast device path is the batched refactorization in luspice.h.
 */
#ifndef DSH_LU_GPU_H
#define DSH_LU_GPU_H

#ifdef __HIPCC__
#include <hip/hip_runtime.h>

__global__ void lu_factor_kernel(
    int n, const int *Ap, const int *Ai, const double *Ax, const int *diagrow,
    int *Lp, int *Li, double *Lx, long Lcap, int *Up, int *Ui, double *Ux, long Ucap,
    int *pinv, int *xi, double *x, int *lnnz, int *unnz, int *singular_out);

__global__ void trisolve_kernel(
    int n, const int *Lp, const int *Li, const double *Lx,
    const int *Up, const int *Ui, const double *Ux,
    const int *pinv, const int *q, const double *b, double *x,
    int *work_i, double *work_d);

__global__ void lu_factor_f32_kernel(
    int n, const int *Ap, const int *Ai, const float *Ax, const int *diagrow,
    int *Lp, int *Li, float *Lx, long Lcap, int *Up, int *Ui, float *Ux, long Ucap,
    int *pinv, int *xi, float *x, int *lnnz, int *unnz, int *singular_out);

__global__ void trisolve_f32_kernel(
    int n, const int *Lp, const int *Li, const float *Lx,
    const int *Up, const int *Ui, const float *Ux,
    const int *pinv, const int *q, const double *b, double *x,
    int *work_i, double *work_d);

__global__ void trisolve_par_f32_kernel(
    int n, const int *Lp, const int *Li, const float *Lx,
    const int *Up, const int *Ui, const float *Ux,
    const int *pinv, const int *q, const double *b, double *x,
    int *work_i, double *work_d);
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Launchers: return 0 when the launch was enqueued. */
int lu_factor_launch(int n, const int *Ap, const int *Ai, const double *Ax, const int *diagrow,
                     int *Lp, int *Li, double *Lx, long Lcap, int *Up, int *Ui, double *Ux,
                     long Ucap, int *pinv, int *xi, double *x, int *lnnz, int *unnz,
                     int *singular_out, void *stream);
int trisolve_launch(int n, const int *Lp, const int *Li, const double *Lx, const int *Up,
                    const int *Ui, const double *Ux, const int *pinv, const int *q,
                    const double *b, double *x, int *work_i, double *work_d, void *stream);
int lu_factor_f32_launch(int n, const int *Ap, const int *Ai, const float *Ax,
                         const int *diagrow, int *Lp, int *Li, float *Lx, long Lcap, int *Up,
                         int *Ui, float *Ux, long Ucap, int *pinv, int *xi, float *x,
                         int *lnnz, int *unnz, int *singular_out, void *stream);
int trisolve_f32_launch(int n, const int *Lp, const int *Li, const float *Lx, const int *Up,
                        const int *Ui, const float *Ux, const int *pinv, const int *q,
                        const double *b, double *x, int *work_i, double *work_d, int parallel,
                        void *stream);

/* trisolve_f32_launch_noalloc -- no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): the same solve as
 * trisolve_f32_launch (same kernels, same result bit for bit) with work_i /
 * work_d taken from a static per-GPU device array of 3 * LU_NOALLOC_MAX_N
 * doubles (6 MiB, held by each GPU the library runs on until the process
 * ends) instead of arguments.  Allocates nothing.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (two streams, or
 * two host threads) race and must be serialised by the caller.  The kernels
 * clear the workspace themselves, so the array's previous contents do not
 * matter.
 * Returns 0 when enqueued, 1 on a launch error, 2 when n < 0 or
 * n > LU_NOALLOC_MAX_N (nothing launched), 3 when the array's device address
 * could not be obtained (nothing launched). */
#define LU_NOALLOC_MAX_N 262144
int trisolve_f32_launch_noalloc(int n, const int *Lp, const int *Li, const float *Lx,
                                const int *Up, const int *Ui, const float *Ux, const int *pinv,
                                const int *q, const double *b, double *x, int parallel,
                                void *stream);

#ifdef __cplusplus
}
#endif

#endif /* DSH_LU_GPU_H */
