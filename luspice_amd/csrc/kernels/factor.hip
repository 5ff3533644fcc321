// factor.hip -- persistent left-looking numeric refactorization (gfx950).
//
// Reference behaviour: the numeric phase of src/third_party/splu.h:151-206
// (per column: scatter A(:,k), apply every earlier L column j with U(j,k) != 0
// in ascending pivot order, U(:,k) = x on pivotal rows, pivot = x at the
// pivot row, L(:,k) = x / pivot elsewhere) with the pivot sequence and
// pattern fixed by the CPU symbolic analysis, the way a Newton refactor reuses
// them.  The reference launches one kernel per level
// (bench/test_lu_wave.cu:200-217, src/mp/multiblock_lu.cu:423-431); here one
// persistent launch walks every level of every matrix in the batch.
//
// Mapping to CDNA4
//   * one wave64 per column task; tasks come from a global queue in level
//     order, so a task only waits on tasks handed out earlier, which are held
//     by running waves: the schedule cannot deadlock whatever the occupancy;
//   * the column's working vector x (fp64) lives in LDS when it fits
//     (cap doubles per wave), else in a per-wave global scratch slice;
//   * dependencies are checked 64 at a time (one lane per dependency), then
//     applied in order; each dependency is a gather of fp32/fp64 L values and
//     16-bit local slot indices (precomputed by the plan), so no search and no
//     atomics touch x;
//   * the pivot test (|pivot| >= ptol * max |L(:,k)|, in fp64) uses a wave
//     max-reduction;
//   * dependencies are applied with a one-dependency software pipeline (the
//     next L column's values and slots load while the current one updates x);
//   * the finished column's L part is stored with agent-coherent stores and
//     the column's flag is published once they are acknowledged, so no L2
//     write-back / invalidate is needed on either side.
#include "common.hpp"
#include "launch.hpp"

namespace lsp {


// Applies dependency d of the current chunk: x[rel] -= L(:, j) * x[s_d].
// The first 64 entries of the dependency's L column arrive prefetched in
// (tg, lv); longer columns load the rest here.
template <typename T>
__device__ __forceinline__ void apply_dep(const FactorDesc<T> &M, double *x, int64_t ub, int jl0, int jlen,
                                          double xj, int tg, double lv)
{
    const int lane = lane_id();
    if (xj == 0.0) return;
    if (lane < jlen) x[tg] -= lv * xj;
    for (int e = lane + kWave; e < jlen; e += kWave)
        x[M.rel[ub + e]] -= to_f64(ld_coherent_val(&M.val[jl0 + e])) * xj;
}

template <typename T>
__device__ void factor_column(const FactorDesc<T> &M, int k, double *x,
                              int epoch, int *abort_word, double ptol, int *dbg)
{
    const int lane = lane_id();
    const int64_t c0 = M.cp[k];
    const int m = (int)(M.cp[k + 1] - c0);
    const int dl = M.dp[k] - (int)c0;          // local slot of the diagonal

    for (int i = lane; i < m; i += kWave) x[i] = 0.0;
    wave_fence();
    for (int p = M.aptr[k] + lane; p < M.aptr[k + 1]; p += kWave) {
        int s = M.aslot[p];
        if (s >= 0) x[s] = M.ain[M.asrc[p]];
    }
    wave_fence();

    // A column of the dense trailing block applies only its dependencies
    // outside the block (a prefix of its U slots); the block itself is
    // factored densely afterwards.
    const bool tail = k >= M.tail_k0;
    const int ndep = tail ? M.tail_ds[k - M.tail_k0] : dl;
    int64_t u = M.relp[k];
    for (int s0 = 0; s0 < ndep; s0 += kWave) {
        const int nd = min(kWave, ndep - s0);
        int l0 = 0, len = 0;
        const int *flag = nullptr;
        if (lane < nd) {
            const int j = M.pos[c0 + s0 + lane];
            flag = &M.done[j];
            l0 = M.dp[j] + 1;
            len = (int)(M.cp[j + 1] - l0);
        }
        // offsets of each dependency's update-map entries within the chunk
        int uoff = len;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int v = __shfl_up(uoff, o, kWave);
            if (lane >= o) uoff += v;
        }
        const int chunk_len = __shfl(uoff, kWave - 1, kWave);
        uoff -= len;
        if (!wave_wait(flag, epoch, abort_word, dbg, k)) return;   // aborted; kernel drains
        // software pipeline: the L values and target slots of dependency d+1
        // are loaded while dependency d updates x
        int jl0 = __builtin_amdgcn_readlane(l0, 0), jlen = __builtin_amdgcn_readlane(len, 0);
        int64_t ub = u + __builtin_amdgcn_readlane(uoff, 0);
        int tg = 0;
        double lv = 0.0;
        if (lane < jlen) { tg = M.rel[ub + lane]; lv = to_f64(ld_coherent_val(&M.val[jl0 + lane])); }
        for (int d = 0; d < nd; d++) {
            int njl0 = 0, njlen = 0, ntg = 0;
            int64_t nub = 0;
            double nlv = 0.0;
            if (d + 1 < nd) {
                njl0 = __builtin_amdgcn_readlane(l0, d + 1);
                njlen = __builtin_amdgcn_readlane(len, d + 1);
                nub = u + __builtin_amdgcn_readlane(uoff, d + 1);
                if (lane < njlen) { ntg = M.rel[nub + lane]; nlv = to_f64(ld_coherent_val(&M.val[njl0 + lane])); }
            }
            const double xj = x[s0 + d];
            apply_dep(M, x, ub, jl0, jlen, xj, tg, lv);
            wave_fence();
            jl0 = njl0; jlen = njlen; ub = nub; tg = ntg; lv = nlv;
        }
        u += chunk_len;
    }

    if (tail) {
        // U entries above the block are final; the rest of the column is the
        // Schur-complement column S(:, c) = A22(:, c) - L21 U12(:, c).
        const int c = k - M.tail_k0;
        for (int i = lane; i < ndep; i += kWave) {
            const T v = from_f64<T>(x[i]);
            M.val[c0 + i] = v;
            M.urow[M.dst[c0 + i]] = v;
        }
        double *Sc = M.tail_S + (size_t)c * (size_t)M.tail_T;
        for (int r = lane; r < M.tail_T; r += kWave) Sc[r] = 0.0;
        wave_fence();                           // same-address stores of one wave stay ordered
        for (int s = ndep + lane; s < m; s += kWave) Sc[M.pos[c0 + s] - M.tail_k0] = x[s];
        if (lane == 0) M.tail_colmax[c] = 0ull;
        return;
    }

    const double piv = x[dl];
    double amax = 0.0;
    int nonfinite = 0;
    for (int i = dl + 1 + lane; i < m; i += kWave) {
        const double a = fabs(x[i]);
        amax = fmax(amax, a);
        nonfinite |= !(a <= 1.0e308);
    }
    amax = wave_max(amax);
    nonfinite = __any(nonfinite) || !(fabs(piv) <= 1.0e308);
    if (lane == 0) {
        if (piv == 0.0) {
            atomicMin(&M.mstat[2], k);
        } else if (!(fabs(piv) >= ptol * amax)) {
            atomicAdd(&M.mstat[0], 1);
            atomicMin(&M.mstat[1], k);
        }
        if (nonfinite) atomicAdd(&M.mstat[3], 1);
    }
    // L(:, k) is read by later columns of this launch: coherent stores, then
    // the flag once they are acknowledged.  U, the diagonal and the row-wise
    // copies are read only after the launch.
    for (int i = lane; i < dl; i += kWave) {
        const T v = from_f64<T>(x[i]);
        M.val[c0 + i] = v;
        M.urow[M.dst[c0 + i]] = v;
    }
    if (lane == 0) {
        const T v = from_f64<T>(piv);
        M.val[c0 + dl] = v;
        M.diag[k] = v;
    }
    for (int i = dl + 1 + lane; i < m; i += kWave) {
        const T v = from_f64<T>(x[i] / piv);
        st_coherent(&M.val[c0 + i], v);
        M.lrow[M.dst[c0 + i]] = v;
    }
    store_complete();
    if (lane == 0) st_relaxed(&M.done[k], epoch);
}

template <typename T>
__global__ void __launch_bounds__(256) refactor_kernel(FactorLaunch L)
{
    extern __shared__ double lds[];
    const int wv = threadIdx.x >> 6;
    const int waves_per_block = blockDim.x >> 6;
    double *xl = lds + (long)wv * L.cap;
    double *xg = L.scratch + ((long)blockIdx.x * waves_per_block + wv) * L.scratch_stride;
    const FactorDesc<T> *mats = (const FactorDesc<T> *)L.mats;
    const int epoch = read_epoch(L.epoch);

    for (;;) {
        const int t = wave_ticket(&L.ctr[0]);
        if (t >= L.ntask) break;
        if (aborted(L.abort_word)) continue;                  // drain
        const int2 tk = L.task[t];
        if (L.active && !L.active[tk.x]) continue;
        const FactorDesc<T> M = mats[tk.x];
        const int k = tk.y;
        const int m = (int)(M.cp[k + 1] - M.cp[k]);
        if (m <= L.cap) {
            factor_column<T>(M, k, xl, epoch, L.abort_word, L.ptol, L.dbg);
        } else if (m <= L.scratch_stride) {
            factor_column<T>(M, k, xg, epoch, L.abort_word, L.ptol, L.dbg);
        } else {
            if (lane_id() == 0) st_relaxed(L.abort_word, kErrCapacity);
        }
    }
    // the last wave out re-arms the queue and advances the epoch
    if (lane_id() == 0) {
        const int fin = atomicAdd(&L.ctr[1], 1);
        if (fin == L.total_waves - 1) {
            st_relaxed(&L.ctr[0], 0);
            st_relaxed(&L.ctr[1], 0);
            advance_epoch(L.epoch, epoch);
        }
    }
}

// Per-matrix status words are reset by a kernel (not a memset) so the whole
// refactor sequence stays capturable in a HIP graph.
__global__ void reset_mstat_kernel(int *mstat, int nmat, int *abort_word, int *dbg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (dbg && i < 8) dbg[i] = -1;
    if (i < nmat) {
        mstat[4 * i + 0] = 0;
        mstat[4 * i + 1] = 0x7fffffff;
        mstat[4 * i + 2] = 0x7fffffff;
        mstat[4 * i + 3] = 0;
    }
    if (i == 0 && abort_word) *abort_word = 0;
}

}  // namespace lsp

// ---------------------------------------------------------------------------
// host launchers (C++ linkage, used by the runtime)
// ---------------------------------------------------------------------------
namespace lsp {

template <typename T>
hipError_t launch_refactor(const FactorLaunch &L, int blocks, hipStream_t st)
{
    const size_t lds = (size_t)L.cap * sizeof(double) * 4;
    hipLaunchKernelGGL(refactor_kernel<T>, dim3(blocks), dim3(256), lds, st, L);
    return hipGetLastError();
}

template hipError_t launch_refactor<float>(const FactorLaunch &, int, hipStream_t);
template hipError_t launch_refactor<double>(const FactorLaunch &, int, hipStream_t);
template hipError_t launch_refactor<__half>(const FactorLaunch &, int, hipStream_t);

hipError_t launch_reset_mstat(int *mstat, int nmat, int *abort_word, int *dbg, hipStream_t st)
{
    const int threads = 256;
    const int blocks = (nmat + threads - 1) / threads;
    hipLaunchKernelGGL(reset_mstat_kernel, dim3(blocks > 0 ? blocks : 1), dim3(threads), 0, st,
                       mstat, nmat, abort_word, dbg);
    return hipGetLastError();
}

template <typename T>
hipError_t refactor_set_lds(int cap)
{
    const int bytes = cap * (int)sizeof(double) * 4;
    return hipFuncSetAttribute((const void *)refactor_kernel<T>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}
template hipError_t refactor_set_lds<float>(int);
template hipError_t refactor_set_lds<double>(int);
template hipError_t refactor_set_lds<__half>(int);

}  // namespace lsp
