// spmv.hip -- SpMV kernels (gfx950).
//
// sell_spmv_kernel: SELL-C-sigma with C = 64.  One wave per slice, one lane
//   per row; slot j of row r of slice s sits at off[s] + j*64 + r, so each
//   step of the row loop is one fully coalesced 64-lane load of values and
//   column ids.  Lanes stop at their own row length, so padding is never
//   read.  Each row is summed sequentially in CSR order, hence with
//   contraction off the fp64 result is bitwise equal to a CPU CSR loop
//   (src/spmv.cu:154-171 of the reference is that loop, one thread per row).
//   Reference layout: src/sellc_spmv.cu:170-257 (C = 32, one warp per slice,
//   y returned in slice order and permuted on the host at
//   bench/test_lu_gpu_f32.cu:196-199); here y is written straight to the
//   original row through perm.  Value type V = double, float or __half,
//   accumulation always fp64 (src/mp/fp16_spmv.cu:50-59 pattern).
//   With b != nullptr the kernel writes the residual b - A x instead.
//   Batched: tasks are (matrix, slice) pairs over any number of matrices;
//   vectors of matrix m start at voff[m] in the shared arenas.
//
// csr_spmv_kernel: the reference's spmv_launch contract (one thread per row,
//   256-thread blocks, y = A x in CSR order).
#include "common.hpp"
#include "desc.hpp"

namespace lsp {

constexpr int kSpmvBatch = 8;           // slots of a row loaded per round

template <typename V>
__global__ void __launch_bounds__(256)
sell_spmv_kernel(const SellDesc<V> *mats, const int2 *tasks, int ntask,
                 const int64_t *voff, const double *x, double *y, const double *b,
                 const int *active)
{
    // the task index is wave-uniform: readfirstlane lets the task and its
    // descriptor come through scalar loads
    const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (t >= ntask) return;
    const int2 tk = tasks[t];
    if (active && !active[tk.x]) return;
    const SellDesc<V> M = mats[tk.x];
    const int lane = lane_id();
    const int k = tk.y * kWave + lane;
    if (k >= M.n) return;
    const int64_t vo = voff[tk.x];
    const double *xm = x + vo;
    const int len = M.rlen[k];
    const int64_t base = M.off[tk.y] + lane;
    const int row = M.perm[k];
    const double bi = b ? b[vo + row] : 0.0;
    // kSpmvBatch slots per round: every value and column id of the round is
    // loaded, then every x entry, then the products are added in slot order
    // (the same sum as one slot at a time), so a row of up to kSpmvBatch
    // entries costs two dependent memory round trips instead of two per entry
    double s = 0.0;
    for (int j0 = 0; j0 < len; j0 += kSpmvBatch) {
        V v[kSpmvBatch];
        int c[kSpmvBatch];
#pragma unroll
        for (int u = 0; u < kSpmvBatch; u++) {
            const int64_t e = base + (int64_t)min(j0 + u, len - 1) * kWave;
            v[u] = M.val[e];
            c[u] = M.col[e];
        }
        double xv[kSpmvBatch];
#pragma unroll
        for (int u = 0; u < kSpmvBatch; u++) xv[u] = xm[c[u]];
#pragma unroll
        for (int u = 0; u < kSpmvBatch; u++)
            if (j0 + u < len) s += to_f64(v[u]) * xv[u];
    }
    y[vo + row] = b ? bi - s : s;
}

__global__ void __launch_bounds__(256)
csr_spmv_kernel(int n, const int *__restrict__ rowptr, const int *__restrict__ colidx,
                const double *__restrict__ values, const double *__restrict__ x,
                double *__restrict__ y)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row < n) {
        double s = 0.0;
        for (int e = rowptr[row]; e < rowptr[row + 1]; e++) s += values[e] * x[colidx[e]];
        y[row] = s;
    }
}

// CSR with fp32 values: the same loop reading float values, fp64 x and sum
// (the reference's k_csr32, src/mp/mp_spmv_bench.cu:80-88, keeps x and y in
// fp32; here only the stored values are narrowed).
__global__ void __launch_bounds__(256)
csr_spmv_f32_kernel(int n, const int *__restrict__ rowptr, const int *__restrict__ colidx,
                    const float *__restrict__ values, const double *__restrict__ x,
                    double *__restrict__ y)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row < n) {
        double s = 0.0;
        for (int e = rowptr[row]; e < rowptr[row + 1]; e++) s += (double)values[e] * x[colidx[e]];
        y[row] = s;
    }
}

// out[i] = idx[i] >= 0 ? (Vout) in[idx[i]] : 0   (value repack per refactor)
template <typename Vout>
__global__ void gather_cast_kernel(long count, const int *idx, const double *in, Vout *out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        const int s = idx[i];
        out[i] = from_f64<Vout>(s >= 0 ? in[s] : 0.0);
    }
}

template <typename V>
hipError_t launch_sell_spmv(const SellDesc<V> *mats, const int2 *tasks, int ntask,
                            const int64_t *voff, const double *x, double *y, const double *b,
                            const int *active, hipStream_t st)
{
    if (ntask <= 0) return hipSuccess;
    const int blocks = (ntask + 3) / 4;
    hipLaunchKernelGGL(sell_spmv_kernel<V>, dim3(blocks), dim3(256), 0, st, mats, tasks, ntask,
                       voff, x, y, b, active);
    return hipGetLastError();
}
template hipError_t launch_sell_spmv<double>(const SellDesc<double> *, const int2 *, int,
                                             const int64_t *, const double *, double *,
                                             const double *, const int *, hipStream_t);
template hipError_t launch_sell_spmv<float>(const SellDesc<float> *, const int2 *, int,
                                            const int64_t *, const double *, double *,
                                            const double *, const int *, hipStream_t);
template hipError_t launch_sell_spmv<__half>(const SellDesc<__half> *, const int2 *, int,
                                             const int64_t *, const double *, double *,
                                             const double *, const int *, hipStream_t);

hipError_t launch_csr_spmv(int n, const int *rowptr, const int *colidx, const double *values,
                           const double *x, double *y, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int blocks = (n + 255) / 256;
    hipLaunchKernelGGL(csr_spmv_kernel, dim3(blocks), dim3(256), 0, st, n, rowptr, colidx,
                       values, x, y);
    return hipGetLastError();
}

hipError_t launch_csr_spmv_f32(int n, const int *rowptr, const int *colidx, const float *values,
                               const double *x, double *y, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int blocks = (n + 255) / 256;
    hipLaunchKernelGGL(csr_spmv_f32_kernel, dim3(blocks), dim3(256), 0, st, n, rowptr, colidx,
                       values, x, y);
    return hipGetLastError();
}

template <typename Vout>
hipError_t launch_gather_cast(long count, const int *idx, const double *in, Vout *out,
                              hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    const long blocks = (count + 255) / 256;
    hipLaunchKernelGGL(gather_cast_kernel<Vout>, dim3((unsigned)blocks), dim3(256), 0, st,
                       count, idx, in, out);
    return hipGetLastError();
}
template hipError_t launch_gather_cast<double>(long, const int *, const double *, double *, hipStream_t);
template hipError_t launch_gather_cast<float>(long, const int *, const double *, float *, hipStream_t);
template hipError_t launch_gather_cast<__half>(long, const int *, const double *, __half *, hipStream_t);

}  // namespace lsp
