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
//   * dependencies are fetched 64 at a time (one lane per dependency) and
//     applied in order, by supernode segment, as soon as a leading group of
//     their flags has been seen; rows below a segment are reached through
//     16-bit local slot indices precomputed by the plan, so no search and no
//     atomics touch x;
//   * the finished column's L part is stored with agent-coherent stores and
//     the column's flag is published once they are acknowledged, so no L2
//     write-back / invalidate is needed on either side;
//   * the pivot test (|pivot| >= ptol * max |L(:,k)|, in fp64, wave
//     max-reduction) and the stores read only after the launch (U, diagonal,
//     row-wise copies) follow the flag.
#include <algorithm>

#include "common.hpp"
#include "launch.hpp"

namespace lsp {


constexpr int kGemvBatch = 16;          // L values in flight per GEMV row pass

// Rows below a run's part in a chunk: row i gets sum_t' L(i, j_t') x[j_t']
// over the run's columns t' in [ta, tb), with L(i, j_t') at column offset
// (tb - t' - 1) + i.  Rows i < tri are the run's own later rows (slots
// s0 + tb + i); the rest are L(:, jl) through the map.  G row chunks of 64
// are processed per pass with B columns' values in flight, so a long run
// keeps G * B loads outstanding; every row is updated once.
template <int G, int B, typename T, typename XP>
__device__ __forceinline__ void run_gemv(const T *val, XP x, const LSP_GLOBAL uint16_t *map, int64_t u, double xc,
                                         int lb, int ta, int tb, int tri, int R, int s0)
{
    const int lane = lane_id();
    const int bw = tb - ta;
    for (int i0 = 0; i0 < R; i0 += G * kWave) {
        int i[G], slot[G];
        double acc[G];
#pragma unroll
        for (int g = 0; g < G; g++) {
            i[g] = min(i0 + g * kWave + lane, R - 1);
            acc[g] = 0.0;
        }
        // fp64 values: the target slots first, so their map loads overlap the
        // value loads (measured on c1: 3.92 -> 3.69 ms; with fp32 values the
        // same order measured slower, 3.55 -> 3.70 ms, so they load after)
        if constexpr (sizeof(T) == 8) {
#pragma unroll
            for (int g = 0; g < G; g++) slot[g] = i[g] < tri ? s0 + tb + i[g] : (int)map[u + (i[g] - tri)];
        }
        for (int q0 = 0; q0 < bw; q0 += B) {
            T v[B][G];
#pragma unroll
            for (int h = 0; h < B; h++) {
                const int tp = ta + min(q0 + h, bw - 1);
                const int base = __builtin_amdgcn_readlane(lb, tp) + (tb - tp - 1);
#pragma unroll
                for (int g = 0; g < G; g++) v[h][g] = ld_coherent_val(gp(val) + (uint32_t)base, (uint32_t)i[g]);
            }
#pragma unroll
            for (int h = 0; h < B; h++) {
                if (q0 + h < bw) {
                    const double w = bcast_f64(xc, ta + q0 + h);
#pragma unroll
                    for (int g = 0; g < G; g++) acc[g] = fma(to_f64(v[h][g]), w, acc[g]);
                }
            }
        }
#pragma unroll
        for (int g = 0; g < G; g++) {
            if (i0 + g * kWave + lane < R) {
                if constexpr (sizeof(T) != 8) slot[g] = i[g] < tri ? s0 + tb + i[g] : (int)map[u + (i[g] - tri)];
                x[slot[g]] -= acc[g];
            }
        }
    }
}

// The rows below a single-column part whose column j has not been seen
// published: the values of L(:, j) are loaded until none is unset (the
// values array holds the unset marker before the launch, and every value is
// written once, as a whole), then applied as run_gemv would.  Used for the
// last dependency of a column when it is the column just before it (the
// critical chain of a supernode), which saves the producer's store
// acknowledgement and the flag round trip.  Returns false when abandoned.
template <int G, typename T, typename XP>
__device__ __forceinline__ bool run_gemv_unset(const T *val, XP x, const LSP_GLOBAL uint16_t *map, int64_t u,
                                               double xj, int base, int R, int *abort_word)
{
    const int lane = lane_id();
    for (int i0 = 0; i0 < R; i0 += G * kWave) {
        int i[G], slot[G];
        T v[G];
#pragma unroll
        for (int g = 0; g < G; g++) i[g] = min(i0 + g * kWave + lane, R - 1);
        // the rows' slots do not depend on the predecessor: their loads are
        // in flight while its values are polled, not after they arrive
#pragma unroll
        for (int g = 0; g < G; g++) slot[g] = (int)map[u + i[g]];
        for (int round = 0;; round++) {
#pragma unroll
            for (int g = 0; g < G; g++) v[g] = ld_coherent_val(gp(val) + (uint32_t)base, (uint32_t)i[g]);
            bool miss = false;
#pragma unroll
            for (int g = 0; g < G; g++) miss |= Unset<T>::is(v[g]);
            if (!__any(miss)) break;
            if ((round & 15) == 15 && aborted(abort_word)) return false;
            if (round > kRoundLimit) {
                atomicCAS(abort_word, 0, kErrSpin);
                return false;
            }
            __builtin_amdgcn_s_sleep(2);
        }
#pragma unroll
        for (int g = 0; g < G; g++) {
            if (i0 + g * kWave + lane < R) x[slot[g]] -= to_f64(v[g]) * xj;
        }
    }
    return true;
}

// Marks the factor values of every active matrix unset (before a refactor);
// blockIdx.y is the matrix.
template <typename T>
__global__ void __launch_bounds__(256) fill_unset_kernel(const FactorDesc<T> *mats, const int *active)
{
    const int m = blockIdx.y;
    if (active && !active[m]) return;
    const FactorDesc<T> M = mats[m];
    const long n = (long)M.cp[M.n];
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) M.val[i] = Unset<T>::value();
}

// One column of the left-looking refactorization.  XP is the work vector's
// pointer type: an LSP_LDS pointer when the column fits the wave's LDS slice,
// an LSP_GLOBAL scratch pointer otherwise.
//
// Dependencies are applied by supernode segment.  The U slots of column k
// whose rows j_s .. e-1 are consecutive columns of one supernode D (e =
// min(end of D, k)) form a run; inside the run L(j_t, j_t') is entry
// (t - t' - 1) of column j_t', and below the run every column of D holds
// the same rows (the run's later rows, then L(:, e-1)).  So a run is
//   * a unit-lower triangular solve on the run's x entries, done in registers
//     (lane t holds x[j_t]; step t' broadcasts x[j_t'] with readlane), and
//   * a dense update of the rows below: row i gets sum_t' L(i, j_t') x[j_t'],
//     with the L values read as coalesced runs and one x update per row.
// Columns outside any supernode are runs of length 1.  Slots are taken 64 at
// a time; each lane fetches the metadata and waits on the flag of one slot.
template <typename T, typename XP>
__device__ void factor_column(const FactorDesc<T> &M, int k, XP x,
                              int epoch, int *abort_word, double ptol, double stol, int *dbg, long long *tw,
                              int backoff, int *tctr, TicketReq &treq)
{
    const int lane = lane_id();
    const LSP_GLOBAL int64_t *cp = gp(M.cp);
    const LSP_GLOBAL int *dp = gp(M.dp);
    const LSP_GLOBAL int *pos = gp(M.pos);
    const LSP_GLOBAL int *snend = gp(M.snend);
    const LSP_GLOBAL uint16_t *map = gp(M.rel);
    const int64_t c0 = cp[k];
    const int m = (int)(cp[k + 1] - c0);
    const int dl = dp[k] - (int)c0;            // local slot of the diagonal

    for (int i = lane; i < m; i += kWave) x[i] = 0.0;
    wave_fence();
    {
        const LSP_GLOBAL int *aslot = gp(M.aslot);
        const LSP_GLOBAL int *asrc = gp(M.asrc);
        const LSP_GLOBAL double *ain = gp(M.ain);
        const int a1 = gp(M.aptr)[k + 1];
        for (int p = gp(M.aptr)[k] + lane; p < a1; p += kWave) {
            const int s = aslot[p];
            if (s >= 0) x[s] = ain[asrc[p]];
        }
    }
    wave_fence();

    // A column of the dense trailing block applies only its dependencies
    // outside the block (a prefix of its U slots); the block itself is
    // factored densely afterwards.
    const bool tail = k >= M.tail_k0;
    const int ndep = tail ? gp(M.tail_ds)[k - M.tail_k0] : dl;
    int64_t u = gp(M.relp)[k];                 // map cursor
    for (int s0 = 0; s0 < ndep; s0 += kWave) {
        const int nd = min(kWave, ndep - s0);
        int j = -2, e = -2, lb = 0, nl = 0;
        const int *flag = nullptr;
        if (lane < nd) {
            j = pos[c0 + s0 + lane];
            e = min(snend[j], k);
            lb = dp[j] + 1;                    // first L slot of column j
            nl = (int)(cp[e] - dp[e - 1]) - 1; // |L(:, e-1)|
            flag = &M.done[j];
        }
        const int jp = __shfl_up(j, 1, kWave);
        const int ep = __shfl_up(e, 1, kWave);
        const bool head = lane < nd && (lane == 0 || j != jp + 1 || e != ep);
        const unsigned long long heads = __ballot(head);
        double xc = lane < nd ? (double)x[s0 + lane] : 0.0;
        // Dependencies are applied as soon as a leading group of the chunk's
        // flags has been seen: a column whose last dependency is its
        // predecessor in the same supernode (the critical chain) applies all
        // the others while that one is still being factored, and then only
        // the predecessor's column.  A run that continues past the ready
        // prefix is cut there; its later rows stay in the rows below.
        int ok = flag == nullptr;
        // the last dependency is the column just before this one, alone in
        // its part (the chain step inside a supernode): its values are
        // checked directly instead of its flag
        const bool chain_last = s0 + nd == ndep && __builtin_amdgcn_readlane(j, nd - 1) == k - 1;
        for (int have = 0; have < nd;) {
            if (chain_last && have == nd - 1) {
                const int lbj = __builtin_amdgcn_readlane(lb, nd - 1);
                const int nlj = __builtin_amdgcn_readlane(nl, nd - 1);
                const double xj = bcast_f64(xc, nd - 1);
                if (!run_gemv_unset<8>(M.val, x, map, u, xj, lbj, nlj, abort_word)) return;
                u += nlj;
                wave_fence();
                have = nd;
                break;
            }
            const int p = wave_wait_prefix(flag, epoch, abort_word, ok, have, nd, dbg, k, backoff);
            if (p < 0) return;                                      // aborted; kernel drains
            if (tw) tw[0] = (long long)wall_clock64();
            unsigned long long hs = (heads & ~((1ull << have) - 1)) | (1ull << have);
            if (p < kWave) hs &= (1ull << p) - 1;
            while (hs) {
                const int ta = __builtin_ctzll(hs);
                hs &= hs - 1;
                const int tb = hs ? __builtin_ctzll(hs) : p;
                const int bw = tb - ta;
                const int jl = __builtin_amdgcn_readlane(e, ta) - 1;        // last column of the run
                const int tri = jl - __builtin_amdgcn_readlane(j, tb - 1);  // run rows after this part
                const int nlast = __builtin_amdgcn_readlane(nl, ta);
                if (bw > 1) {
                    if (bw <= 9) {
                        xc = run_triangle<8>(M.val, xc, lb, ta, tb);
                    } else if constexpr (sizeof(T) == 8) {
                        // fp64 values: 16 triangle columns in registers at a
                        // time (the same steps in the same order), which keeps
                        // the kernel at 3 waves per SIMD instead of 2
                        for (int t0 = ta; t0 < tb - 1; t0 += 16) xc = run_triangle<16>(M.val, xc, lb, t0, tb);
                    } else {
                        if (bw <= 25) xc = run_triangle<24>(M.val, xc, lb, ta, tb);
                        else xc = run_triangle<kTri>(M.val, xc, lb, ta, tb);
                    }
                }
                if (lane >= ta && lane < tb) x[s0 + lane] = xc;
                // rows below this part of the run: the run's later rows, then
                // the rows of L(:, jl) through the map
                // (about 32 loads in flight: a narrow part, such as the single
                // column that ends a critical chain step, covers more rows
                // per round trip)
                const int R = tri + nlast;
                // (fp64 values: half the columns in flight per pass, for the
                // register budget of 3 waves per SIMD)
                constexpr int kB8 = sizeof(T) == 8 ? 4 : 8;
                if (bw >= 8) {
                    if (R > 2 * kWave) run_gemv<4, kB8>(M.val, x, map, u, xc, lb, ta, tb, tri, R, s0);
                    else run_gemv<1, kGemvBatch>(M.val, x, map, u, xc, lb, ta, tb, tri, R, s0);
                } else if (R <= kWave) {
                    run_gemv<1, 8>(M.val, x, map, u, xc, lb, ta, tb, tri, R, s0);
                } else if (bw == 1) {
                    run_gemv<8, 1>(M.val, x, map, u, xc, lb, ta, tb, tri, R, s0);
                } else if (bw <= 4) {
                    run_gemv<sizeof(T) == 8 ? 4 : 8, 4>(M.val, x, map, u, xc, lb, ta, tb, tri, R, s0);
                } else {
                    run_gemv<4, kB8>(M.val, x, map, u, xc, lb, ta, tb, tri, R, s0);
                }
                if (tri == 0) u += nlast;
                wave_fence();
                if (lane >= tb && lane < nd) xc = x[s0 + lane];   // later lanes of this chunk
            }
            have = p;
        }
        wave_fence();
    }

    if (tail) {
        // U entries above the block are final; the rest of the column is the
        // Schur-complement column S(:, c) = A22(:, c) - L21 U12(:, c).
        const int c = k - M.tail_k0;
        for (int i = lane; i < ndep; i += kWave) {
            const T v = from_f64<T>(x[i]);
            st_g(&M.val[c0 + i], v);
            st_g(&M.urow[gp(M.dst)[c0 + i]], v);
        }
        LSP_GLOBAL double *Sc = gp(M.tail_S) + (size_t)c * (size_t)M.tail_T;
        for (int r = lane; r < M.tail_T; r += kWave) Sc[r] = 0.0;
        wave_fence();                           // same-address stores of one wave stay ordered
        for (int s = ndep + lane; s < m; s += kWave) Sc[pos[c0 + s] - M.tail_k0] = x[s];
        if (lane == 0) gp(M.tail_colmax)[c] = 0ull;
        return;
    }

    // L(:, k) is read by later columns of this launch: coherent stores, then
    // the flag once they are acknowledged.  The pivot test, U, the diagonal
    // and the row-wise copies are needed only after the launch, so they come
    // after the flag, off the dependency chain.
    const double piv = x[dl];
    const double rpiv = 1.0 / piv;             // one division; L = x * (1 / pivot)
    for (int i = dl + 1 + lane; i < m; i += kWave) st_coherent(&M.val[c0 + i], from_f64<T>(x[i] * rpiv));
    store_complete();
    if (tw) tw[1] = (long long)wall_clock64();  // timeline: flag published
    if (lane == 0) st_relaxed(&M.done[k], epoch);
    // The wave's next ticket is requested now: what follows waits on no
    // other wave, so the ticket is not held behind a dependency, and the
    // atomic's round trip overlaps the bookkeeping.  fp64 storage only: with
    // fp32 / FP16 storage the held request raised the kernel to 193 VGPRs
    // and the fp32 refactor ran 4 % slower on c1 (fp64: 3.30 -> 3.24 ms,
    // 512-circuit Newton batch +8 %).
    if constexpr (sizeof(T) == 8) treq = ticket_issue(tctr);

    // a stored value must be finite in the storage type (65504 for FP16)
    const double vmax = sizeof(T) == 2 ? 65504.0 : sizeof(T) == 4 ? 3.4028234663852886e38 : 1.0e308;
    double amax = 0.0;
    int nonfinite = 0;
    for (int i = dl + 1 + lane; i < m; i += kWave) {
        const double a = fabs(x[i]);
        amax = fmax(amax, a);
        nonfinite |= !(a <= 1.0e308) || !(a <= vmax * fabs(piv));
    }
    for (int i = lane; i < dl; i += kWave) nonfinite |= !(fabs(x[i]) <= vmax);
    amax = wave_max(amax);
    nonfinite = __any(nonfinite) || !(fabs(piv) <= vmax);
    if (lane == 0) {
        if (piv == 0.0) {
            atomicMin(&M.mstat[2], k);
        } else if (!(fabs(piv) >= ptol * amax)) {
            atomicAdd(&M.mstat[0], 1);
            atomicMin(&M.mstat[1], k);
        }
        if (nonfinite) atomicAdd(&M.mstat[3], 1);
    }
    if (stol > 0.0) {
        // would splu (tol = stol) pick this pivot on these values?
        const int dpos = gp(M.dpos)[k];
        double dv = -1.0;
        if (dpos > k) {
            for (int i = dl + 1 + lane; i < m; i += kWave)
                if (pos[c0 + i] == dpos) dv = fabs(x[i]);
            dv = wave_max(dv);
        }
        if (lane == 0) {
            const int v = splu_pivot_verdict(dpos == k, fabs(piv), amax, dv, stol);
            if (v) {
                atomicAdd(&M.mstat[v == 1 ? 4 : 5], 1);
                atomicMin(&M.mstat[6], k);
            }
        }
    }
    const LSP_GLOBAL int *dst = gp(M.dst);
    for (int i = lane; i < dl; i += kWave) {
        const T v = from_f64<T>(x[i]);
        st_g(&M.val[c0 + i], v);
        st_g(&M.urow[dst[c0 + i]], v);
    }
    if (lane == 0) {
        const T v = from_f64<T>(piv);
        st_g(&M.val[c0 + dl], v);
        st_g(&M.diag[k], v);
    }
    for (int i = dl + 1 + lane; i < m; i += kWave) st_g(&M.lrow[dst[c0 + i]], from_f64<T>(x[i] * rpiv));
}

template <typename T>
__global__ void __launch_bounds__(256) refactor_kernel(FactorLaunch L)
{
    extern __shared__ double lds[];
    const int wv = threadIdx.x >> 6;
    const int waves_per_block = blockDim.x >> 6;
    LSP_LDS double *xl = (LSP_LDS double *)lds + (long)wv * L.cap;
    LSP_GLOBAL double *xg = gp(L.scratch) + ((long)blockIdx.x * waves_per_block + wv) * L.scratch_stride;
    const FactorDesc<T> *mats = (const FactorDesc<T> *)L.mats;
    const int epoch = read_epoch(L.epoch);

    const int nq = queue_count(L.nq);
    const int q = queue_index(nq);
    int *const qc = L.ctr + 32 * q;                           // this wave's counter
    TicketReq req{0, -1};
    for (;;) {
        // ticket k of counter q is task k * nq + q; the request comes from
        // the previous column when it reached its bookkeeping
        const int t = (req.leader >= 0 ? ticket_take(req) : wave_ticket(qc)) * nq + q;
        req.leader = -1;
        if (t >= L.ntask) break;
        if (aborted(L.abort_word)) continue;                  // drain
        const int2 tk = L.task[t];
        if (L.active && !L.active[tk.x]) continue;
        const FactorDesc<T> M = mats[tk.x];
        const int k = tk.y;
        const int m = (int)(M.cp[k + 1] - M.cp[k]);
        long long tw[2] = {0, 0};               // end of the last wait, flag published
        long long *wp = L.timeline ? tw : nullptr;
        const long long t0 = L.timeline ? (long long)wall_clock64() : 0;
        if (m <= L.cap) {
            factor_column<T>(M, k, xl, epoch, L.abort_word, L.ptol, L.stol, L.dbg, wp, L.poll_backoff, qc, req);
        } else if (m <= L.scratch_stride) {
            factor_column<T>(M, k, xg, epoch, L.abort_word, L.ptol, L.stol, L.dbg, wp, L.poll_backoff, qc, req);
        } else {
            if (lane_id() == 0) st_relaxed(L.abort_word, kErrCapacity);
        }
        if (L.timeline && lane_id() == 0) {
            long long *r = L.timeline + 4L * t;
            r[0] = t0;
            r[1] = tw[1] ? tw[1] : (long long)wall_clock64();
            r[2] = tw[0];
            r[3] = k;
        }
    }
    // the last wave out re-arms the queue and advances the epoch
    if (leave_queue(L.ctr, nq, L.total_waves)) advance_epoch(L.epoch, epoch);
}

// Per-matrix status words are reset by a kernel (not a memset) so the whole
// refactor sequence stays capturable in a HIP graph.
__global__ void reset_mstat_kernel(int *mstat, int nmat, int *abort_word, int *dbg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (dbg && i < 8) dbg[i] = -1;
    if (i < nmat) {
        int *w = mstat + kMstat * i;
        w[0] = 0;
        w[1] = 0x7fffffff;
        w[2] = 0x7fffffff;
        w[3] = 0;
        w[4] = 0;
        w[5] = 0;
        w[6] = 0x7fffffff;
        w[7] = 0;
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

template <typename T>
hipError_t launch_fill_unset(const void *mats, const int *active, int nmat, long max_slots, hipStream_t st)
{
    const long want = (max_slots + 255) / 256;
    const int bx = (int)std::max(1L, std::min(1024L, want));
    hipLaunchKernelGGL(fill_unset_kernel<T>, dim3(bx, nmat), dim3(256), 0, st, (const FactorDesc<T> *)mats,
                       active);
    return hipGetLastError();
}
template hipError_t launch_fill_unset<float>(const void *, const int *, int, long, hipStream_t);
template hipError_t launch_fill_unset<double>(const void *, const int *, int, long, hipStream_t);
template hipError_t launch_fill_unset<__half>(const void *, const int *, int, long, hipStream_t);

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
