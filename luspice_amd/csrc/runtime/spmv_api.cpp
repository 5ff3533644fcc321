// spmv_api.cpp -- device half of the reference's spmv.h contract
// (src/spmv.cu:177-246): csr_alloc_device allocates and copies, spmv_launch
// enqueues y = A x (one thread per row, CSR order) on the given stream,
// spmv_host does a full round trip and times one launch after a warm-up.
#include <hip/hip_runtime.h>

#include "../kernels/launch.hpp"
#include "spmv.h"

extern "C" {

int csr_alloc_device(const CSRMatrix *A, int **d_rowptr, int **d_colidx, double **d_values)
{
    if (!A || !d_rowptr || !d_colidx || !d_values || A->n <= 0) return 1;
    const size_t nb_rp = sizeof(int) * ((size_t)A->n + 1);
    const size_t nb_ci = sizeof(int) * (size_t)(A->nnz > 0 ? A->nnz : 1);
    const size_t nb_vl = sizeof(double) * (size_t)(A->nnz > 0 ? A->nnz : 1);
    if (hipMalloc((void **)d_rowptr, nb_rp) != hipSuccess) return 1;
    if (hipMalloc((void **)d_colidx, nb_ci) != hipSuccess) return 1;
    if (hipMalloc((void **)d_values, nb_vl) != hipSuccess) return 1;
    if (hipMemcpy(*d_rowptr, A->rowptr, nb_rp, hipMemcpyHostToDevice) != hipSuccess) return 1;
    if (A->nnz > 0) {
        if (hipMemcpy(*d_colidx, A->colidx, sizeof(int) * A->nnz, hipMemcpyHostToDevice) != hipSuccess)
            return 1;
        if (hipMemcpy(*d_values, A->values, sizeof(double) * A->nnz, hipMemcpyHostToDevice) != hipSuccess)
            return 1;
    }
    return 0;
}

void spmv_launch(int n, const int *d_rowptr, const int *d_colidx, const double *d_values,
                 const double *d_x, double *d_y, void *stream)
{
    (void)lsp::launch_csr_spmv(n, d_rowptr, d_colidx, d_values, d_x, d_y, (hipStream_t)stream);
}

int spmv_host(const CSRMatrix *A, const double *x, double *y, double *time_ms)
{
    if (!A || !x || !y || A->n <= 0) return 1;
    const int n = A->n;
    int *rp = nullptr, *ci = nullptr;
    double *vl = nullptr, *dx = nullptr, *dy = nullptr;
    int rc = 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (csr_alloc_device(A, &rp, &ci, &vl)) goto done;
    if (hipMalloc((void **)&dx, sizeof(double) * n) != hipSuccess) goto done;
    if (hipMalloc((void **)&dy, sizeof(double) * n) != hipSuccess) goto done;
    if (hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) goto done;
    spmv_launch(n, rp, ci, vl, dx, dy, nullptr);
    if (hipDeviceSynchronize() != hipSuccess) goto done;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
    (void)hipEventRecord(e0, nullptr);
    spmv_launch(n, rp, ci, vl, dx, dy, nullptr);
    (void)hipEventRecord(e1, nullptr);
    if (hipEventSynchronize(e1) != hipSuccess) goto done;
    if (time_ms) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *time_ms = ms;
    }
    if (hipMemcpy(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess) goto done;
    rc = 0;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(rp);
    (void)hipFree(ci);
    (void)hipFree(vl);
    (void)hipFree(dx);
    (void)hipFree(dy);
    return rc;
}

}  // extern "C"
