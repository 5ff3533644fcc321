// gmres.hip -- device-resident GMRES(m) pieces for fp64 GMRES-IR (gfx950).
//
// Reference behaviour: bench/test_lu_gpu_f32.cu:540-715 (right-preconditioned
// restarted GMRES, x0 = 0, Arnoldi with two orthogonalisation passes, Givens
// least squares, inner stop |g_{j+1}|/||b|| <= tol, outer stop on the true
// residual).  The reference runs all of it on the host with two PCIe
// crossings per inner iteration (:595-615); here every vector operation and
// the small Hessenberg/Givens work stay on the device.
//
// Orthogonalisation is classical Gram-Schmidt applied twice (CGS2): each pass
// is one block-dot h = V^T w and one update w -= V h, so a pass costs two
// launches instead of 2(j+1) dependent reductions; two passes give the same
// O(eps) loss of orthogonality as the reference's two-pass modified
// Gram-Schmidt.
//
// Batch layout: vectors of matrix m live at voff[m] .. voff[m]+n_m in arenas
// of length ntot; basis vector i is V + i*ntot.  Reductions are two-stage and
// deterministic: row chunks (matrix, r0, r1) write per-chunk partial sums,
// then one thread per (matrix, i) adds them in chunk order.  Matrices whose
// act[m] == 0 are skipped by every kernel.
#include "common.hpp"
#include "launch.hpp"

namespace lsp {


__device__ __forceinline__ double block_sum(double v, double *red)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
    }
    __syncthreads();
    return s;
}

// partial[c][i] = sum over chunk c of V_i * w  (i = 0..nv-1); with V == w and
// nv == 1 this is the squared norm.
__global__ void __launch_bounds__(256) chunk_dots_kernel(KrylovArgs K, const double *V, const double *w, int nv)
{
    __shared__ double red[4];
    const int c = blockIdx.x;
    const int4 ch = K.chunks[c];
    if (!K.act[ch.x]) return;
    const long vo = K.voff[ch.x];
    for (int i = 0; i < nv; i++) {
        const double *vi = V + (long)i * K.ntot + vo;
        double s = 0.0;
        for (int r = ch.y + threadIdx.x; r < ch.z; r += blockDim.x) s += vi[r] * w[vo + r];
        s = block_sum(s, red);
        if (threadIdx.x == 0) K.partial[(long)c * (K.mmax + 1) + i] = s;
    }
}

// out[m*(mmax+1)+i] = sum of partials (mode 0: store, mode 1: add into H
// column j and hcol, mode 2: sqrt into hn)
// One wave per (matrix, i): lanes stride over the matrix's chunks, then a
// wave sum (a single thread walking hundreds of chunks took ~46 us on c1).
__global__ void __launch_bounds__(64) reduce_partials_kernel(KrylovArgs K, int nv, int mode, int j, double *out)
{
    const int t = blockIdx.x;
    const int m = t / nv, i = t % nv;
    if (m >= K.nmat || !K.act[m]) return;
    double s = 0.0;
    for (int c = K.cbeg[m] + (int)threadIdx.x; c < K.cbeg[m + 1]; c += kWave)
        s += K.partial[(long)c * (K.mmax + 1) + i];
    s = wave_sum(s);
    if (threadIdx.x != 0) return;
    if (mode == 0) {
        out[(long)m * (K.mmax + 1) + i] = s;
    } else if (mode == 1) {
        K.hcol[(long)m * (K.mmax + 1) + i] = s;
        K.H[(long)m * (K.mmax + 1) * K.mmax + (long)i * K.mmax + j] += s;
    } else {
        out[m] = sqrt(s);
    }
}

// w -= sum_{i<nv} hcol[i] V_i
__global__ void __launch_bounds__(256) chunk_update_kernel(KrylovArgs K, const double *V, double *w, int nv)
{
    const int c = blockIdx.x;
    const int4 ch = K.chunks[c];
    if (!K.act[ch.x]) return;
    const long vo = K.voff[ch.x];
    const double *h = K.hcol + (long)ch.x * (K.mmax + 1);
    for (int r = ch.y + threadIdx.x; r < ch.z; r += blockDim.x) {
        double acc = w[vo + r];
        for (int i = 0; i < nv; i++) acc -= h[i] * V[(long)i * K.ntot + vo + r];
        w[vo + r] = acc;
    }
}

// dst = src * (1/den[m]) per matrix (dst may alias src); den from beta or hn
__global__ void __launch_bounds__(256) chunk_scale_kernel(KrylovArgs K, const double *src, double *dst,
                                                          const double *den)
{
    const int c = blockIdx.x;
    const int4 ch = K.chunks[c];
    if (!K.act[ch.x]) return;
    const long vo = K.voff[ch.x];
    const double d = den[ch.x];
    for (int r = ch.y + threadIdx.x; r < ch.z; r += blockDim.x) dst[vo + r] = src[vo + r] / d;
}

// Cycle start: outer convergence test on beta, reset H and g.
// One workgroup per matrix: every thread reads the same state and takes the
// same branch; the Hessenberg and g are cleared by all threads together.
__global__ void __launch_bounds__(128) start_cycle_kernel(KrylovArgs K, int *n_active)
{
    const int m = blockIdx.x;
    if (m >= K.nmat) return;
    const bool lead = threadIdx.x == 0;
    const bool conv = K.conv[m] != 0;
    const double rel = K.beta[m] / K.nb[m];
    const bool done = !conv && (rel <= K.outer_tol || !(K.beta[m] > 0.0));
    __syncthreads();                            // every thread has read conv before it changes
    if (lead) {
        K.cycled[m] = 0;
        K.jfin[m] = 0;
        K.act[m] = conv || done ? 0 : 1;
        if (done) K.conv[m] = 1;
    }
    if (conv || done) return;
    double *H = K.H + (long)m * (K.mmax + 1) * K.mmax;
    for (long t = threadIdx.x; t < (long)(K.mmax + 1) * K.mmax; t += blockDim.x) H[t] = 0.0;
    double *g = K.g + (long)m * (K.mmax + 1);
    for (int t = threadIdx.x; t <= K.mmax; t += blockDim.x) g[t] = t == 0 ? K.beta[m] : 0.0;
    if (lead) atomicAdd(n_active, 1);
}

// Givens step for column j; sets act = 0 when the cycle ends for matrix m.
__global__ void givens_kernel(KrylovArgs K, int j, int *n_active)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= K.nmat || !K.act[m]) return;
    const int mm = K.mmax;
    double *H = K.H + (long)m * (mm + 1) * mm;
    double *cs = K.cs + (long)m * mm, *sn = K.sn + (long)m * mm;
    double *g = K.g + (long)m * (mm + 1);
    const double hn = K.hn[m];
    H[(long)(j + 1) * mm + j] = hn;
    for (int i = 0; i < j; i++) {
        const double t0 = H[(long)i * mm + j];
        const double t1 = H[(long)(i + 1) * mm + j];
        H[(long)i * mm + j] = cs[i] * t0 + sn[i] * t1;
        H[(long)(i + 1) * mm + j] = -sn[i] * t0 + cs[i] * t1;
    }
    const double a = H[(long)j * mm + j], c = H[(long)(j + 1) * mm + j];
    const double r = sqrt(a * a + c * c);
    cs[j] = r > 0 ? a / r : 1.0;
    sn[j] = r > 0 ? c / r : 0.0;
    H[(long)j * mm + j] = r;
    H[(long)(j + 1) * mm + j] = 0.0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    K.jfin[m] = j + 1;
    K.cycled[m] = 1;
    const bool done = fabs(g[j + 1]) / K.nb[m] <= K.inner_tol || hn < 1e-14 || j + 1 >= mm;
    if (done) K.act[m] = 0;
    else atomicAdd(n_active, 1);
}

// y = R^-1 g over the first jfin columns; re-activates matrices that cycled
// so the following combine / preconditioner / update launches see them.
__global__ void backsolve_kernel(KrylovArgs K)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= K.nmat) return;
    if (!K.cycled[m]) { K.act[m] = 0; return; }
    const int mm = K.mmax, jf = K.jfin[m];
    const double *H = K.H + (long)m * (mm + 1) * mm;
    const double *g = K.g + (long)m * (mm + 1);
    double *y = K.yv + (long)m * mm;
    for (int i = jf - 1; i >= 0; i--) {
        double s = g[i];
        for (int t = i + 1; t < jf; t++) s -= H[(long)i * mm + t] * y[t];
        y[i] = s / H[(long)i * mm + i];
    }
    K.act[m] = 1;
}

// t = V y (first jfin basis vectors)
__global__ void __launch_bounds__(256) chunk_combine_kernel(KrylovArgs K, const double *V, double *t)
{
    const int c = blockIdx.x;
    const int4 ch = K.chunks[c];
    if (!K.act[ch.x]) return;
    const long vo = K.voff[ch.x];
    const int jf = K.jfin[ch.x];
    const double *y = K.yv + (long)ch.x * K.mmax;
    for (int r = ch.y + threadIdx.x; r < ch.z; r += blockDim.x) {
        double acc = 0.0;
        for (int i = 0; i < jf; i++) acc += y[i] * V[(long)i * K.ntot + vo + r];
        t[vo + r] = acc;
    }
}

// x += z for matrices in the update phase
__global__ void __launch_bounds__(256) chunk_axpy_kernel(KrylovArgs K, double *x, const double *z)
{
    const int c = blockIdx.x;
    const int4 ch = K.chunks[c];
    if (!K.act[ch.x]) return;
    const long vo = K.voff[ch.x];
    for (int r = ch.y + threadIdx.x; r < ch.z; r += blockDim.x) x[vo + r] += z[vo + r];
}

// x = v on active matrices
__global__ void __launch_bounds__(256) chunk_fill_kernel(KrylovArgs K, double *x, double v)
{
    const int c = blockIdx.x;
    const int4 ch = K.chunks[c];
    if (!K.act[ch.x]) return;
    const long vo = K.voff[ch.x];
    for (int r = ch.y + threadIdx.x; r < ch.z; r += blockDim.x) x[vo + r] = v;
}

// Set act for every not-yet-converged matrix (used before the outer residual)
__global__ void activate_kernel(KrylovArgs K)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < K.nmat) K.act[m] = K.conv[m] ? 0 : 1;
}

// act = 1 for all (final residual evaluation, b-norms)
__global__ void set_all_active_kernel(KrylovArgs K, int v)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < K.nmat) K.act[m] = v;
}

__global__ void zero_kernel(double *p, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void zero_int_kernel(int *p, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void copy_kernel(const double *src, double *dst, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
static inline int nb_for(long n, int t) { return (int)((n + t - 1) / t); }

hipError_t k_chunk_dots(const KrylovArgs &K, const double *V, const double *w, int nv, hipStream_t s)
{
    hipLaunchKernelGGL(chunk_dots_kernel, dim3(K.nchunks), dim3(256), 0, s, K, V, w, nv);
    return hipGetLastError();
}
hipError_t k_reduce(const KrylovArgs &K, int nv, int mode, int j, double *out, hipStream_t s)
{
    const int tot = K.nmat * nv;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(tot > 0 ? tot : 1), dim3(kWave), 0, s, K, nv, mode, j, out);
    return hipGetLastError();
}
hipError_t k_chunk_update(const KrylovArgs &K, const double *V, double *w, int nv, hipStream_t s)
{
    hipLaunchKernelGGL(chunk_update_kernel, dim3(K.nchunks), dim3(256), 0, s, K, V, w, nv);
    return hipGetLastError();
}
hipError_t k_chunk_scale(const KrylovArgs &K, const double *src, double *dst, const double *den, hipStream_t s)
{
    hipLaunchKernelGGL(chunk_scale_kernel, dim3(K.nchunks), dim3(256), 0, s, K, src, dst, den);
    return hipGetLastError();
}
hipError_t k_start_cycle(const KrylovArgs &K, int *n_active, hipStream_t s)
{
    hipLaunchKernelGGL(start_cycle_kernel, dim3(K.nmat > 0 ? K.nmat : 1), dim3(128), 0, s, K, n_active);
    return hipGetLastError();
}
hipError_t k_givens(const KrylovArgs &K, int j, int *n_active, hipStream_t s)
{
    hipLaunchKernelGGL(givens_kernel, dim3(nb_for(K.nmat, 128)), dim3(128), 0, s, K, j, n_active);
    return hipGetLastError();
}
hipError_t k_backsolve(const KrylovArgs &K, hipStream_t s)
{
    hipLaunchKernelGGL(backsolve_kernel, dim3(nb_for(K.nmat, 128)), dim3(128), 0, s, K);
    return hipGetLastError();
}
hipError_t k_chunk_combine(const KrylovArgs &K, const double *V, double *t, hipStream_t s)
{
    hipLaunchKernelGGL(chunk_combine_kernel, dim3(K.nchunks), dim3(256), 0, s, K, V, t);
    return hipGetLastError();
}
hipError_t k_chunk_axpy(const KrylovArgs &K, double *x, const double *z, hipStream_t s)
{
    hipLaunchKernelGGL(chunk_axpy_kernel, dim3(K.nchunks), dim3(256), 0, s, K, x, z);
    return hipGetLastError();
}
hipError_t k_chunk_fill(const KrylovArgs &K, double *x, double v, hipStream_t s)
{
    hipLaunchKernelGGL(chunk_fill_kernel, dim3(K.nchunks), dim3(256), 0, s, K, x, v);
    return hipGetLastError();
}
hipError_t k_activate(const KrylovArgs &K, hipStream_t s)
{
    hipLaunchKernelGGL(activate_kernel, dim3(nb_for(K.nmat, 128)), dim3(128), 0, s, K);
    return hipGetLastError();
}
hipError_t k_set_all_active(const KrylovArgs &K, int v, hipStream_t s)
{
    hipLaunchKernelGGL(set_all_active_kernel, dim3(nb_for(K.nmat, 128)), dim3(128), 0, s, K, v);
    return hipGetLastError();
}
hipError_t k_zero(double *p, long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_kernel, dim3(nb_for(n, 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}
hipError_t k_zero_int(int *p, long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_int_kernel, dim3(nb_for(n, 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}
hipError_t k_copy(const double *src, double *dst, long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_kernel, dim3(nb_for(n, 256)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

}  // namespace lsp
