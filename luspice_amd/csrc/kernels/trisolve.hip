// trisolve.hip -- level-ordered, synchronisation-free row-gather triangular
// solves x = Q U^-1 L^-1 P b with fp32/fp64/fp16 factors and fp64
// accumulation (gfx950).
//
// Reference behaviour: lu_solve_perm (src/third_party/splu.h:210-239) and the
// fp32-factor trisolves (src/lu_gpu_f32.cu:255-333: factors read as double,
// fp64 accumulation, pb = P b, x[q[k]] = z[k]).  The reference's
// level-scheduled form (src/mp/fusion_bench.cu:80-116) runs one block with a
// barrier per level; here rows are tasks in a global queue sorted by level,
// each row waits only for the rows it reads (epoch flags), and every CU works
// on the front of the dependency DAG at once.
//
//   forward   y[i] = b[rowat[i]] - sum_{j<i} L(i,j) y[j]
//   backward  z[i] = (y[i] - sum_{j>i} U(i,j) z[j]) / U(i,i),  x[q[i]] = z[i]
//
// One wave per row: lanes stride over the row's entries (coalesced index and
// value loads), wait on their own dependency flags in parallel, and the wave
// sums in fp64 with a DPP reduction.  Solved values are stored agent-coherent
// and flagged once acknowledged; readers load them agent-coherent, so no L2
// write-back or invalidate is needed (see common.hpp).
#include "common.hpp"
#include "launch.hpp"

namespace lsp {


template <typename T, bool kForward>
__global__ void __launch_bounds__(256) trisolve_kernel(TriLaunch L)
{
    const int lane = lane_id();
    const TriDesc<T> *mats = (const TriDesc<T> *)L.mats;
    const int epoch = read_epoch(L.epoch);
    for (;;) {
        const int t = wave_ticket(&L.ctr[0]);
        if (t >= L.ntask) break;
        if (aborted(L.abort_word)) continue;
        const int2 tk = L.task[t];
        if (L.active && !L.active[tk.x]) continue;
        const TriDesc<T> M = mats[tk.x];
        const int i = tk.y;
        const long vo = L.voff[tk.x];
        const int *rp = kForward ? M.lrp : M.urp;
        const int *ci = kForward ? M.lci : M.uci;
        const T *rv = kForward ? M.lrow : M.urow;
        double *y = L.y + vo;
        double *z = L.z + vo;
        const double *src = kForward ? y : z;
        int *flags = kForward ? M.fdone : M.bdone;
        double acc = 0.0;
        bool ok = true;
        const int e0 = rp[i], e1 = rp[i + 1];
        for (int b0 = e0; b0 < e1 && ok; b0 += kWave) {
            const int e = b0 + lane;
            const int j = e < e1 ? ci[e] : 0;
            ok = wave_wait(e < e1 ? &flags[j] : nullptr, epoch, L.abort_word);
            if (!ok) break;
            if (e < e1) acc += to_f64(rv[e]) * ld_coherent_val(&src[j]);
        }
        if (!ok) continue;
        acc = wave_sum(acc);
        if (lane == 0) {
            if (kForward) {
                st_coherent(&y[i], L.b[vo + M.rowat[i]] - acc);
            } else {
                const double zi = (y[i] - acc) / to_f64(M.diag[i]);
                st_coherent(&z[i], zi);
                L.x[vo + M.q[i]] = zi;
            }
            store_complete();
            st_relaxed(&flags[i], epoch);
        }
    }
    if (lane == 0) {
        const int fin = atomicAdd(&L.ctr[1], 1);
        if (fin == L.total_waves - 1) {
            st_relaxed(&L.ctr[0], 0);
            st_relaxed(&L.ctr[1], 0);
            advance_epoch(L.epoch, epoch);
        }
    }
}

template <typename T>
hipError_t launch_trisolve_half(const TriLaunch &L, bool forward, int blocks, hipStream_t st)
{
    if (forward)
        hipLaunchKernelGGL((trisolve_kernel<T, true>), dim3(blocks), dim3(256), 0, st, L);
    else
        hipLaunchKernelGGL((trisolve_kernel<T, false>), dim3(blocks), dim3(256), 0, st, L);
    return hipGetLastError();
}

template hipError_t launch_trisolve_half<float>(const TriLaunch &, bool, int, hipStream_t);
template hipError_t launch_trisolve_half<double>(const TriLaunch &, bool, int, hipStream_t);
template hipError_t launch_trisolve_half<__half>(const TriLaunch &, bool, int, hipStream_t);

}  // namespace lsp
