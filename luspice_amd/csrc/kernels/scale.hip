// scale.hip -- Ruiz equilibration on the device (gfx950).
//
// Reference behaviour: test_lu_gpu_f32.cu:247-269 (30 sweeps; each sweep
// scales every row by 1/sqrt(max |a_ij|) over the row, then every column by
// 1/sqrt(max |a_ij|) over the column; b_s = R b and x = C x_s).  Each sweep
// here recomputes |a_ij r_i c_j| from the original values, so A_s = R A C
// holds exactly for the scales that are returned; the reference multiplies
// the values by the accumulated scale inside the loop instead, which departs
// from R A C after the first sweep.
//
// One workgroup per circuit: maxima are taken with 64-bit atomic max on the
// bit patterns of non-negative doubles (ordered like the values), and the
// phases are separated by workgroup barriers, so the whole equilibration is
// one launch.  Used only ahead of FP16 factor storage, where the scaled
// values must fit half precision.
#include "common.hpp"
#include "launch.hpp"

namespace lsp {

__global__ void __launch_bounds__(256) ruiz_kernel(RuizArgs R)
{
    const int m = blockIdx.x;
    if (R.active && !R.active[m]) return;
    const long a0 = R.aoff[m], a1 = R.aoff[m + 1];
    const long v0 = R.voff[m], v1 = R.voff[m + 1];
    const int *row = R.row, *col = R.col;
    double *rs = R.rs, *cs = R.cs;
    unsigned long long *mx = R.mx;
    const double *a = R.ain;
    for (long i = v0 + threadIdx.x; i < v1; i += blockDim.x) { rs[i] = 1.0; cs[i] = 1.0; }
    __syncthreads();
    for (int sweep = 0; sweep < R.sweeps; sweep++) {
        for (int side = 0; side < 2; side++) {
            const int *idx = side == 0 ? row : col;
            double *sc = side == 0 ? rs : cs;
            for (long i = v0 + threadIdx.x; i < v1; i += blockDim.x) mx[i] = 0ull;
            __syncthreads();
            for (long e = a0 + threadIdx.x; e < a1; e += blockDim.x) {
                const double v = fabs(a[e] * rs[v0 + row[e]] * cs[v0 + col[e]]);
                atomicMax(&mx[v0 + idx[e]], (unsigned long long)__double_as_longlong(v));
            }
            __syncthreads();
            for (long i = v0 + threadIdx.x; i < v1; i += blockDim.x) {
                const double t = __longlong_as_double((long long)mx[i]);
                if (t > 0.0) sc[i] *= 1.0 / sqrt(t);
            }
            __syncthreads();
        }
    }
    for (long e = a0 + threadIdx.x; e < a1; e += blockDim.x)
        R.aout[e] = a[e] * rs[v0 + row[e]] * cs[v0 + col[e]];
}

hipError_t launch_ruiz(const RuizArgs &R, int ncirc, hipStream_t st)
{
    hipLaunchKernelGGL(ruiz_kernel, dim3(ncirc), dim3(256), 0, st, R);
    return hipGetLastError();
}

// Factor-only diagonal shift (test_lu_gpu_f32.cu:354-364, test_lu_gpu_ir64.cu:
// 206-212): for each row i the first value entry on the diagonal, in caller
// order, becomes a + shift*|a| in the copy that is factored; the residual keeps
// the unshifted values.  idx lists those entries (global value indices), at
// most one per row; src and dst may be the same buffer.
__global__ void __launch_bounds__(256) shift_kernel(const double *src, double *dst, const long *idx,
                                                    long nidx, double shift)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < nidx; t += (long)gridDim.x * blockDim.x) {
        const long e = idx[t];
        const double a = src[e];
        dst[e] = a + shift * fabs(a);
    }
}

hipError_t launch_shift(const double *src, double *dst, const long *idx, long nidx, double shift,
                        hipStream_t st)
{
    if (nidx <= 0) return hipSuccess;
    const long want = (nidx + 255) / 256, blocks = want < 4096 ? want : 4096;
    hipLaunchKernelGGL(shift_kernel, dim3((unsigned)blocks), dim3(256), 0, st, src, dst, idx, nidx, shift);
    return hipGetLastError();
}

// out[i] = (float) in[i]: the fp32 factor set as the rounding of the fp64 one
__global__ void __launch_bounds__(256) round_f32_kernel(const double *in, float *out, long count)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

hipError_t launch_round_f32(const double *in, float *out, long count, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    const long want = (count + 255) / 256, blocks = want < 8192 ? want : 8192;
    hipLaunchKernelGGL(round_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, in, out, count);
    return hipGetLastError();
}

}  // namespace lsp
