// legacy.hip -- the reference's one-matrix LU kernels (lu_gpu.h, lu_gpu_par.h)
// on gfx950.
//
// Algorithm: left-looking Gilbert-Peierls LU (J. R. Gilbert, T. Peierls,
// "Sparse partial pivoting in time proportional to arithmetic operations",
// SISC 9(5), 1988) with the reach computed by depth-first search, marks kept
// by negating the column pointer of the visited node (T. A. Davis, "Direct
// Methods for Sparse Linear Systems", SIAM 2006, ch. 3/6); threshold partial
// pivoting with tol = 1 and the diagrow preference, L unit diagonal first,
// U diagonal last, L rows remapped to positions -- the reference's oracle
// semantics (src/third_party/splu.h:151-206, src/cpu_ref_lu.c).  The
// reference's own single-thread kernel scans dense columns instead and drops
// exact zeros from the pattern (src/lu_gpu.cu:34-115); following the oracle
// keeps P, nnz(L), nnz(U) equal to cpu_lu_ref's.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lu_gpu.h"
#include "lu_gpu_par.h"

namespace {

__device__ __forceinline__ int flip(int i) { return -i - 2; }
__device__ __forceinline__ int unflip(int i) { return i < 0 ? flip(i) : i; }

template <typename T> __device__ __forceinline__ T absval(T v) { return v < 0 ? -v : v; }

// Depth-first reach of column `col` of A through the L columns built so far.
// Nodes are original row ids; node j with pinv[j] = J >= 0 leads to the rows
// of L(:,J).  Visited nodes are marked by negating Lp[node]; the marks are
// cleared before returning.  xi[0..) is the DFS stack, xi[top..n) the reach
// in topological order, pstack[] the resume position of each stack level.
__device__ int reach_seq(int n, int *Lp, const int *Li, const int *pinv, const int *Ap,
                         const int *Ai, int col, int *xi, int *pstack)
{
    int top = n;
    for (int p = Ap[col]; p < Ap[col + 1]; p++) {
        const int start = Ai[p];
        if (Lp[start] < 0) continue;
        int head = 0;
        xi[0] = start;
        while (head >= 0) {
            const int j = xi[head];
            const int J = pinv[j];
            if (Lp[j] >= 0) {
                Lp[j] = flip(Lp[j]);
                pstack[head] = J < 0 ? 0 : unflip(Lp[J]);
            }
            const int stop = J < 0 ? 0 : unflip(Lp[J + 1]);
            bool leaf = true;
            for (int q = pstack[head]; q < stop; q++) {
                const int i = Li[q];
                if (Lp[i] < 0) continue;
                pstack[head] = q;
                xi[++head] = i;
                leaf = false;
                break;
            }
            if (leaf) {
                head--;
                xi[--top] = j;
            }
        }
    }
    for (int p = top; p < n; p++) Lp[xi[p]] = unflip(Lp[xi[p]]);
    return top;
}

template <typename T>
__device__ void gp_factor_seq(int n, const int *Ap, const int *Ai, const T *Ax,
                              const int *diagrow, int *Lp, int *Li, T *Lx, long Lcap, int *Up,
                              int *Ui, T *Ux, long Ucap, int *pinv, int *xi, T *x, int *lnnz,
                              int *unnz, int *singular_out)
{
    for (int i = 0; i < n; i++) { x[i] = T(0); pinv[i] = -1; }
    for (int k = 0; k <= n; k++) { Lp[k] = 0; Up[k] = 0; }
    long lnz = 0, unz = 0;
    int *pstack = xi + n;
    for (int k = 0; k < n; k++) {
        Lp[k] = (int)lnz;
        Up[k] = (int)unz;
        const int top = reach_seq(n, Lp, Li, pinv, Ap, Ai, k, xi, pstack);
        for (int p = top; p < n; p++) x[xi[p]] = T(0);
        for (int p = Ap[k]; p < Ap[k + 1]; p++) x[Ai[p]] = Ax[p];
        for (int p = top; p < n; p++) {
            const int j = xi[p];
            const int J = pinv[j];
            if (J < 0) continue;
            x[j] /= Lx[Lp[J]];
            const T xj = x[j];
            for (int q = Lp[J] + 1; q < Lp[J + 1]; q++) x[Li[q]] -= Lx[q] * xj;
        }
        int ipiv = -1;
        T amax = T(-1);
        for (int p = top; p < n; p++) {
            const int i = xi[p];
            if (pinv[i] < 0) {
                const T t = absval(x[i]);
                if (t > amax) { amax = t; ipiv = i; }
            } else {
                if (unz >= Ucap) { *singular_out = -2; return; }
                Ui[unz] = pinv[i];
                Ux[unz++] = x[i];
            }
        }
        if (ipiv == -1 || amax <= T(0)) { *singular_out = k; return; }
        const int dr = diagrow[k];
        if (pinv[dr] < 0 && absval(x[dr]) >= amax) ipiv = dr;
        const T piv = x[ipiv];
        if (unz >= Ucap) { *singular_out = -2; return; }
        Ui[unz] = k;
        Ux[unz++] = piv;
        pinv[ipiv] = k;
        if (lnz >= Lcap) { *singular_out = -2; return; }
        Li[lnz] = ipiv;
        Lx[lnz++] = T(1);
        for (int p = top; p < n; p++) {
            const int i = xi[p];
            if (pinv[i] < 0) {
                if (lnz >= Lcap) { *singular_out = -2; return; }
                Li[lnz] = i;
                Lx[lnz++] = x[i] / piv;
            }
            x[i] = T(0);
        }
    }
    Lp[n] = (int)lnz;
    Up[n] = (int)unz;
    for (long p = 0; p < lnz; p++) Li[p] = pinv[Li[p]];
    *lnnz = (int)lnz;
    *unnz = (int)unz;
    *singular_out = -1;
}

// x = Q U^-1 L^-1 P b; factors of type T, fp64 accumulation.  With nthr > 1
// the scatter of each column is spread over the block; every entry still
// receives its updates in column order, so the result equals nthr == 1.
template <typename T>
__device__ void trisolve_body(int n, const int *Lp, const int *Li, const T *Lx, const int *Up,
                              const int *Ui, const T *Ux, const int *pinv, const int *q,
                              const double *b, double *x, int *rowat, double *work_d, int tid,
                              int nthr)
{
    double *pb = work_d, *y = work_d + n;
    for (int i = tid; i < n; i += nthr) rowat[i] = -1;
    __syncthreads();
    for (int i = tid; i < n; i += nthr) if (pinv[i] >= 0) rowat[pinv[i]] = i;
    __syncthreads();
    for (int k = tid; k < n; k += nthr) pb[k] = rowat[k] >= 0 ? b[rowat[k]] : 0.0;
    __syncthreads();
    for (int k = 0; k < n; k++) {
        const double yk = pb[k] / (double)Lx[Lp[k]];
        if (tid == 0) y[k] = yk;
        for (int p = Lp[k] + 1 + tid; p < Lp[k + 1]; p += nthr) pb[Li[p]] -= (double)Lx[p] * yk;
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; k--) {
        const double yk = y[k] / (double)Ux[Up[k + 1] - 1];
        for (int p = Up[k] + tid; p < Up[k + 1] - 1; p += nthr) y[Ui[p]] -= (double)Ux[p] * yk;
        __syncthreads();
        if (tid == 0) y[k] = yk;
        __syncthreads();
    }
    for (int k = tid; k < n; k += nthr) x[q ? q[k] : k] = y[k];
}

}  // namespace

// ---------------------------------------------------------------------------
__global__ void lu_factor_kernel(int n, const int *Ap, const int *Ai, const double *Ax,
                                 const int *diagrow, int *Lp, int *Li, double *Lx, long Lcap,
                                 int *Up, int *Ui, double *Ux, long Ucap, int *pinv, int *xi,
                                 double *x, int *lnnz, int *unnz, int *singular_out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    gp_factor_seq<double>(n, Ap, Ai, Ax, diagrow, Lp, Li, Lx, Lcap, Up, Ui, Ux, Ucap, pinv, xi,
                          x, lnnz, unnz, singular_out);
}

__global__ void lu_factor_f32_kernel(int n, const int *Ap, const int *Ai, const float *Ax,
                                     const int *diagrow, int *Lp, int *Li, float *Lx, long Lcap,
                                     int *Up, int *Ui, float *Ux, long Ucap, int *pinv, int *xi,
                                     float *x, int *lnnz, int *unnz, int *singular_out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    gp_factor_seq<float>(n, Ap, Ai, Ax, diagrow, Lp, Li, Lx, Lcap, Up, Ui, Ux, Ucap, pinv, xi,
                         x, lnnz, unnz, singular_out);
}

__global__ void trisolve_kernel(int n, const int *Lp, const int *Li, const double *Lx,
                                const int *Up, const int *Ui, const double *Ux, const int *pinv,
                                const int *q, const double *b, double *x, int *work_i,
                                double *work_d)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    trisolve_body<double>(n, Lp, Li, Lx, Up, Ui, Ux, pinv, q, b, x, work_i, work_d, 0, 1);
}

__global__ void trisolve_f32_kernel(int n, const int *Lp, const int *Li, const float *Lx,
                                    const int *Up, const int *Ui, const float *Ux, const int *pinv,
                                    const int *q, const double *b, double *x, int *work_i,
                                    double *work_d)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    trisolve_body<float>(n, Lp, Li, Lx, Up, Ui, Ux, pinv, q, b, x, work_i, work_d, 0, 1);
}

__global__ void __launch_bounds__(256)
trisolve_par_f32_kernel(int n, const int *Lp, const int *Li, const float *Lx, const int *Up,
                        const int *Ui, const float *Ux, const int *pinv, const int *q,
                        const double *b, double *x, int *work_i, double *work_d)
{
    if (blockIdx.x != 0) return;
    trisolve_body<float>(n, Lp, Li, Lx, Up, Ui, Ux, pinv, q, b, x, work_i, work_d, threadIdx.x,
                         blockDim.x);
}

__global__ void __launch_bounds__(256)
trisolve_par_kernel(int n, const int *Lp, const int *Li, const double *Lx, const int *Up,
                    const int *Ui, const double *Ux, const int *pinv, const int *q,
                    const double *b, double *x, int *work_i, double *work_d)
{
    if (blockIdx.x != 0) return;
    trisolve_body<double>(n, Lp, Li, Lx, Up, Ui, Ux, pinv, q, b, x, work_i, work_d, threadIdx.x,
                          blockDim.x);
}

// Single workgroup: lane 0 builds the reach (pstack in LDS), the block applies
// the updates of each reached column, reduces the pivot (first maximum in
// reach order) and lane 0 appends U and L in reach order.
__global__ void __launch_bounds__(256)
lu_factor_par_kernel(int n, const int *Ap, const int *Ai, const double *Ax, const int *diagrow,
                     int *Lp, int *Li, double *Lx, long Lcap, int *Up, int *Ui, double *Ux,
                     long Ucap, int *pinv, int *xi, double *xg, int *lnnz, int *unnz,
                     int *singular_out)
{
    (void)xg;
    if (blockIdx.x != 0) return;
    extern __shared__ double smem[];
    double *x = smem;                          // n doubles
    int *pstack = (int *)(x + n);              // n ints
    __shared__ double red_v[4];
    __shared__ int red_p[4];
    __shared__ int s_top, s_status;
    __shared__ double s_piv;
    const int tid = threadIdx.x, nthr = blockDim.x, lane = tid & 63, wv = tid >> 6;
    for (int i = tid; i < n; i += nthr) { x[i] = 0.0; pinv[i] = -1; }
    for (int k = tid; k <= n; k += nthr) { Lp[k] = 0; Up[k] = 0; }
    if (tid == 0) s_status = -1;
    __syncthreads();
    long lnz = 0, unz = 0;
    for (int k = 0; k < n; k++) {
        if (tid == 0) {
            Lp[k] = (int)lnz;
            Up[k] = (int)unz;
            s_top = reach_seq(n, Lp, Li, pinv, Ap, Ai, k, xi, pstack);
        }
        __syncthreads();
        const int top = s_top;
        for (int p = top + tid; p < n; p += nthr) x[xi[p]] = 0.0;
        __syncthreads();
        if (tid == 0)
            for (int p = Ap[k]; p < Ap[k + 1]; p++) x[Ai[p]] = Ax[p];
        __syncthreads();
        for (int p = top; p < n; p++) {
            const int j = xi[p];
            const int J = pinv[j];
            if (J < 0) continue;
            const double xj = x[j] / Lx[Lp[J]];
            __syncthreads();
            if (tid == 0) x[j] = xj;
            for (int q = Lp[J] + 1 + tid; q < Lp[J + 1]; q += nthr) x[Li[q]] -= Lx[q] * xj;
            __syncthreads();
        }
        // pivot: max |x| over unpivoted reached rows, ties -> earliest in reach order
        double best = -1.0;
        int bestp = 0x7fffffff;
        for (int p = top + tid; p < n; p += nthr) {
            const int i = xi[p];
            if (pinv[i] < 0) {
                const double t = fabs(x[i]);
                if (t > best || (t == best && p < bestp)) { best = t; bestp = p; }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off, 64);
            const int op = __shfl_xor(bestp, off, 64);
            if (ob > best || (ob == best && op < bestp)) { best = ob; bestp = op; }
        }
        if (lane == 0) { red_v[wv] = best; red_p[wv] = bestp; }
        __syncthreads();
        if (tid == 0) {
            for (int w = 1; w < (nthr >> 6); w++)
                if (red_v[w] > best || (red_v[w] == best && red_p[w] < bestp)) {
                    best = red_v[w];
                    bestp = red_p[w];
                }
            int status = -1;
            for (int p = top; p < n && status == -1; p++) {
                const int i = xi[p];
                if (pinv[i] >= 0) {
                    if (unz >= Ucap) { status = -2; break; }
                    Ui[unz] = pinv[i];
                    Ux[unz++] = x[i];
                }
            }
            if (status == -1 && (bestp == 0x7fffffff || best <= 0.0)) status = k;
            if (status == -1) {
                int ipiv = xi[bestp];
                const int dr = diagrow[k];
                if (pinv[dr] < 0 && fabs(x[dr]) >= best) ipiv = dr;
                const double piv = x[ipiv];
                s_piv = piv;
                if (unz >= Ucap || lnz >= Lcap) {
                    status = -2;
                } else {
                    Ui[unz] = k;
                    Ux[unz++] = piv;
                    pinv[ipiv] = k;
                    Li[lnz] = ipiv;
                    Lx[lnz++] = 1.0;
                    for (int p = top; p < n; p++) {
                        const int i = xi[p];
                        if (pinv[i] < 0) {
                            if (lnz >= Lcap) { status = -2; break; }
                            Li[lnz] = i;
                            Lx[lnz++] = x[i] / piv;
                        }
                    }
                }
            }
            s_status = status;
        }
        __syncthreads();
        if (s_status != -1) {
            if (tid == 0) *singular_out = s_status;
            return;
        }
        for (int p = top + tid; p < n; p += nthr) x[xi[p]] = 0.0;
        __syncthreads();
    }
    if (tid == 0) {
        Lp[n] = (int)lnz;
        Up[n] = (int)unz;
        s_top = (int)lnz;
        *lnnz = (int)lnz;
        *unnz = (int)unz;
    }
    __syncthreads();
    for (int p = tid; p < s_top; p += nthr) Li[p] = pinv[Li[p]];
    __syncthreads();
    if (tid == 0) *singular_out = -1;
}

// ---------------------------------------------------------------------------
extern "C" {

int lu_factor_launch(int n, const int *Ap, const int *Ai, const double *Ax, const int *diagrow,
                     int *Lp, int *Li, double *Lx, long Lcap, int *Up, int *Ui, double *Ux,
                     long Ucap, int *pinv, int *xi, double *x, int *lnnz, int *unnz,
                     int *singular_out, void *stream)
{
    hipLaunchKernelGGL(lu_factor_kernel, dim3(1), dim3(1), 0, (hipStream_t)stream, n, Ap, Ai, Ax,
                       diagrow, Lp, Li, Lx, Lcap, Up, Ui, Ux, Ucap, pinv, xi, x, lnnz, unnz,
                       singular_out);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int lu_factor_f32_launch(int n, const int *Ap, const int *Ai, const float *Ax,
                         const int *diagrow, int *Lp, int *Li, float *Lx, long Lcap, int *Up,
                         int *Ui, float *Ux, long Ucap, int *pinv, int *xi, float *x, int *lnnz,
                         int *unnz, int *singular_out, void *stream)
{
    hipLaunchKernelGGL(lu_factor_f32_kernel, dim3(1), dim3(1), 0, (hipStream_t)stream, n, Ap, Ai,
                       Ax, diagrow, Lp, Li, Lx, Lcap, Up, Ui, Ux, Ucap, pinv, xi, x, lnnz, unnz,
                       singular_out);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int trisolve_launch(int n, const int *Lp, const int *Li, const double *Lx, const int *Up,
                    const int *Ui, const double *Ux, const int *pinv, const int *q,
                    const double *b, double *x, int *work_i, double *work_d, void *stream)
{
    hipLaunchKernelGGL(trisolve_kernel, dim3(1), dim3(1), 0, (hipStream_t)stream, n, Lp, Li, Lx,
                       Up, Ui, Ux, pinv, q, b, x, work_i, work_d);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int trisolve_f32_launch(int n, const int *Lp, const int *Li, const float *Lx, const int *Up,
                        const int *Ui, const float *Ux, const int *pinv, const int *q,
                        const double *b, double *x, int *work_i, double *work_d, int parallel,
                        void *stream)
{
    if (parallel)
        hipLaunchKernelGGL(trisolve_par_f32_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, n,
                           Lp, Li, Lx, Up, Ui, Ux, pinv, q, b, x, work_i, work_d);
    else
        hipLaunchKernelGGL(trisolve_f32_kernel, dim3(1), dim3(1), 0, (hipStream_t)stream, n, Lp,
                           Li, Lx, Up, Ui, Ux, pinv, q, b, x, work_i, work_d);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Workspace of trisolve_f32_launch_noalloc: 2n doubles (work_d) then n ints
// (work_i), n <= LU_NOALLOC_MAX_N.  Each GPU holds its own copy for the life
// of the process.  Host code never passes the array's name to a kernel (on
// the host it is only a stand-in): it passes the device address that
// hipGetSymbolAddress returns for the current device, looked up once per
// device and kept.
static __device__ __align__(256) double scratch[3 * (size_t)LU_NOALLOC_MAX_N];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array, or nullptr on failure.
double *scratch_on(int dev)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        double *p = nullptr;
        if (hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int trisolve_f32_launch_noalloc(int n, const int *Lp, const int *Li, const float *Lx, const int *Up,
                                const int *Ui, const float *Ux, const int *pinv, const int *q,
                                const double *b, double *x, int parallel, void *stream)
{
    if (n < 0 || n > LU_NOALLOC_MAX_N) return 2;
    // the scratch copy of the device the launch runs on (the stream's device)
    int cur = 0, dev = 0;
    if (hipGetDevice(&cur) != hipSuccess) return 3;
    dev = cur;
    if (stream && hipStreamGetDevice((hipStream_t)stream, &dev) != hipSuccess) return 3;
    if (dev < 0 || dev >= kMaxDevices) return 3;
    if (dev != cur && hipSetDevice(dev) != hipSuccess) return 3;
    double *ptr = scratch_on(dev);
    if (dev != cur) (void)hipSetDevice(cur);
    if (!ptr) return 3;
    double *work_d = ptr;                          // pb[n], y[n]
    int *work_i = (int *)(ptr + 2 * (size_t)n);    // rowat[n]
    return trisolve_f32_launch(n, Lp, Li, Lx, Up, Ui, Ux, pinv, q, b, x, work_i, work_d, parallel,
                               stream);
}

int lu_factor_par_launch(int n, const int *Ap, const int *Ai, const double *Ax,
                         const int *diagrow, int *Lp, int *Li, double *Lx, long Lcap, int *Up,
                         int *Ui, double *Ux, long Ucap, int *pinv, int *xi, int *lnnz,
                         int *unnz, int *singular_out, void *stream)
{
    const size_t shb = LU_PAR_SHARED_BYTES(n);
    if (shb > 65536 &&
        hipFuncSetAttribute((const void *)lu_factor_par_kernel,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shb) != hipSuccess)
        return 1;
    hipLaunchKernelGGL(lu_factor_par_kernel, dim3(1), dim3(256), shb, (hipStream_t)stream, n, Ap,
                       Ai, Ax, diagrow, Lp, Li, Lx, Lcap, Up, Ui, Ux, Ucap, pinv, xi,
                       (double *)nullptr, lnnz, unnz, singular_out);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int trisolve_par_launch(int n, const int *Lp, const int *Li, const double *Lx, const int *Up,
                        const int *Ui, const double *Ux, const int *pinv, const int *q,
                        const double *b, double *x, int *work_i, double *work_d, void *stream)
{
    hipLaunchKernelGGL(trisolve_par_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, n, Lp, Li,
                       Lx, Up, Ui, Ux, pinv, q, b, x, work_i, work_d);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // extern "C"
