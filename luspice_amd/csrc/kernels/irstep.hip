// irstep.hip -- fused iterative-refinement step (gfx950): one launch does,
// for every circuit of a batch that has not converged,
//   r = b - A x            (fp64 values, SELL-C-64 slices, original row order)
//   stop if ||r|| <= tol ||b||   (the circuit is marked converged, x kept)
//   z = U^-1 L^-1 P r       (stored factors, fp64 accumulation)
//   x = x + Q z
// so a Richardson iteration x_{k+1} = x_k + M^-1 (b - A x_k) costs one
// launch per step for the whole batch.
//
// Reference behaviour: the fused IR step of src/mp/fusion_bench.cu:119-163
// (one block: residual SpMV, then forward and backward substitution level by
// level with a barrier per level; the host adds the correction).  Here one
// workgroup of 256 threads per circuit, all circuits in one grid:
//   * residual: wave w takes slices w, w+4, ...; the squared norms of r and b
//     are reduced in LDS;
//   * forward: the rows below the dense block level by level (one thread per
//     row, one barrier per level); the block's rows gather their sparse part,
//     then L22^-1 in 64-row blocks: wave 0 solves the block's unit triangle
//     in registers (readlane broadcasts), the workgroup subtracts the block
//     from the rows below it;
//   * backward: U22^-1 on the dense block in 64-row blocks from the bottom,
//     then the sparse rows by backward level (the block's rows come first);
//   * the work vector (position order) lives in LDS when the circuit fits,
//     else in the batch's work arena.
#include "common.hpp"
#include "launch.hpp"

namespace lsp {

namespace {

__device__ __forceinline__ double block_sum(double v, double *red)
{
    const int wv = threadIdx.x >> 6;
    v = wave_sum(v);
    if (lane_id() == 0) red[wv] = v;
    __syncthreads();
    const double s = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return s;
}

}  // namespace

template <typename T>
__global__ void __launch_bounds__(256) ir_step_kernel(IrStepArgs A)
{
    extern __shared__ double lds[];
    __shared__ double red[4];
    const int m = blockIdx.x;
    if (A.conv[m]) return;
    const IrDesc<T> D = ((const IrDesc<T> *)A.descs)[m];
    const int tid = threadIdx.x;
    const int wv = tid >> 6;
    const int lane = lane_id();
    const long vo = D.voff;
    const int n = D.n;
    double *y = n <= A.lds_n ? lds : A.yw + vo;
    double *r = A.r + vo;
    double *x = A.x + vo;
    const double *b = A.b + vo;

    // ---- r = b - A x and the two norms ----
    double rr = 0.0, bb = 0.0;
    if (A.first) {                                // x = 0: r = b
        for (int i = tid; i < n; i += 256) {
            const double bi = b[i];
            r[i] = bi;
            x[i] = 0.0;
            bb += bi * bi;
        }
        rr = bb;
    } else
    for (int s = wv; s < D.A.nslices; s += 4) {
        const int k = s * kWave + lane;
        if (k < n) {
            const int len = D.A.rlen[k];
            const int64_t base = D.A.off[s] + lane;
            double sum = 0.0;
            for (int j = 0; j < len; j++) {
                const int64_t e = base + (int64_t)j * kWave;
                sum += D.A.val[e] * x[D.A.col[e]];
            }
            const int row = D.A.perm[k];
            const double bi = b[row];
            const double ri = bi - sum;
            r[row] = ri;
            rr += ri * ri;
            bb += bi * bi;
        }
    }
    rr = block_sum(rr, red);
    bb = block_sum(bb, red);
    if (tid == 0) {
        A.rn2[m] = rr;
        A.bn2[m] = bb;
    }
    if (rr <= A.tol2 * bb) {                       // workgroup-uniform
        if (tid == 0) A.conv[m] = 1;
        return;
    }

    // ---- forward: sparse rows by level ----
    const int k0 = D.tail_k0, TT = D.tail_T;
    for (int l = 0; l < D.nlv_f; l++) {
        const int t1 = D.flvp[l + 1];
        for (int t = D.flvp[l] + tid; t < t1; t += 256) {
            const int i = D.frow[t];
            double acc = r[D.rowat[i]];
            for (int e = D.lrp[i]; e < D.lrp[i + 1]; e++) acc -= to_f64(D.lrow[e]) * y[D.lci[e]];
            y[i] = acc;
        }
        __syncthreads();
    }
    if (TT > 0) {
        // the block's rows: their entries left of the block
        for (int q = tid; q < TT; q += 256) {
            const int i = k0 + q;
            double acc = r[D.rowat[i]];
            for (int e = D.lrp[i]; e < D.tail_lcut[q]; e++) acc -= to_f64(D.lrow[e]) * y[D.lci[e]];
            y[i] = acc;
        }
        __syncthreads();
        // L22^-1 (unit lower, column-major tail_F)
        for (int b0 = 0; b0 < TT; b0 += kWave) {
            const int bw = min(kWave, TT - b0);
            if (wv == 0) {
                const int rr_ = b0 + min(lane, bw - 1);
                double v = lane < bw ? y[k0 + b0 + lane] : 0.0;
                T lv[kWave];
#pragma unroll
                for (int c = 0; c < kWave; c++) lv[c] = D.tail_F[(size_t)(b0 + min(c, bw - 1)) * TT + rr_];
#pragma unroll
                for (int c = 0; c < kWave; c++) {
                    if (c < bw) {
                        const double yc = bcast_f64(v, c);
                        if (lane > c) v -= to_f64(lv[c]) * yc;
                    }
                }
                if (lane < bw) y[k0 + b0 + lane] = v;
            }
            __syncthreads();
            for (int q = b0 + bw + tid; q < TT; q += 256) {
                double acc = 0.0;
                for (int c = 0; c < bw; c++) acc += to_f64(D.tail_F[(size_t)(b0 + c) * TT + q]) * y[k0 + b0 + c];
                y[k0 + q] -= acc;
            }
            __syncthreads();
        }
        // ---- backward: U22^-1 (diagonal included), blocks from the bottom ----
        for (int b0 = ((TT - 1) / kWave) * kWave; b0 >= 0; b0 -= kWave) {
            const int bw = min(kWave, TT - b0);
            if (wv == 0) {
                const int rr_ = b0 + min(lane, bw - 1);
                double v = lane < bw ? y[k0 + b0 + lane] : 0.0;
                const double rd = 1.0 / to_f64(D.tail_F[(size_t)rr_ * TT + rr_]);
                T uv[kWave];
#pragma unroll
                for (int c = 0; c < kWave; c++) uv[c] = D.tail_F[(size_t)(b0 + min(c, bw - 1)) * TT + rr_];
#pragma unroll
                for (int c = kWave - 1; c >= 0; c--) {
                    if (c < bw) {
                        const double zc = bcast_f64(v, c) * bcast_f64(rd, c);
                        if (lane == c) v = zc;
                        if (lane < c) v -= to_f64(uv[c]) * zc;
                    }
                }
                if (lane < bw) y[k0 + b0 + lane] = v;
            }
            __syncthreads();
            for (int q = tid; q < b0; q += 256) {
                double acc = 0.0;
                for (int c = 0; c < bw; c++) acc += to_f64(D.tail_F[(size_t)(b0 + c) * TT + q]) * y[k0 + b0 + c];
                y[k0 + q] -= acc;
            }
            __syncthreads();
        }
    }
    // ---- backward: sparse rows by level ----
    for (int l = 0; l < D.nlv_b; l++) {
        const int t1 = D.blvp[l + 1];
        for (int t = D.blvp[l] + tid; t < t1; t += 256) {
            const int i = D.brow[t];
            double acc = y[i];
            for (int e = D.urp[i]; e < D.urp[i + 1]; e++) acc -= to_f64(D.urow[e]) * y[D.uci[e]];
            y[i] = acc / to_f64(D.diag[i]);
        }
        __syncthreads();
    }
    // ---- x += Q z ----
    for (int i = tid; i < n; i += 256) x[D.q[i]] += y[i];
}

template <typename T>
hipError_t launch_ir_step(const IrStepArgs &A, int nmat, hipStream_t st)
{
    const size_t lds = (size_t)A.lds_n * sizeof(double);
    hipLaunchKernelGGL(ir_step_kernel<T>, dim3(nmat), dim3(256), lds, st, A);
    return hipGetLastError();
}
template hipError_t launch_ir_step<float>(const IrStepArgs &, int, hipStream_t);
template hipError_t launch_ir_step<double>(const IrStepArgs &, int, hipStream_t);

template <typename T>
hipError_t ir_step_set_lds(int lds_n)
{
    return hipFuncSetAttribute((const void *)ir_step_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               lds_n * (int)sizeof(double));
}
template hipError_t ir_step_set_lds<float>(int);
template hipError_t ir_step_set_lds<double>(int);

}  // namespace lsp
