// common.hpp -- device helpers shared by the gfx950 kernels.
//
// Conventions
//   * wave64 everywhere: lane = threadIdx.x & 63, wave = threadIdx.x >> 6.
//   * Cross-workgroup hand-off uses epoch-tagged flags: a producer stores
//     flag[i] = epoch with release semantics at agent scope after its data
//     stores; a consumer spins on an acquire load until flag[i] == epoch.
//     Epochs increase per launch, so flags never need clearing.
//   * Every spin is bounded: after kSpinLimit polls the waiter raises the
//     shared abort word, every other waiter sees it and leaves, and the kernel
//     drains normally.  The host reports the abort as an error.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "desc.hpp"

namespace lsp {

// Address-space qualified pointers.  Descriptor pointers loaded from memory
// are generic to the compiler, which then emits FLAT accesses; a FLAT load is
// counted on both the vector-memory and the LDS counters, so any wait for an
// LDS result also drains every load in flight.  Hot loops therefore access
// device arrays through LSP_GLOBAL pointers and LDS through LSP_LDS pointers.
#define LSP_GLOBAL __attribute__((address_space(1)))
#define LSP_LDS __attribute__((address_space(3)))
template <typename T> __device__ __forceinline__ LSP_GLOBAL T *gp(T *p) { return (LSP_GLOBAL T *)p; }
template <typename T> __device__ __forceinline__ const LSP_GLOBAL T *gp(const T *p)
{
    return (const LSP_GLOBAL T *)p;
}

constexpr int kWave = 64;
constexpr int kSpinLimit = 1 << 22;   // ~ seconds with s_sleep(1) between polls

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ int ld_acquire(const int *p)
{
    return __hip_atomic_load(gp(p), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_relaxed(const int *p)
{
    return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Read-modify-write probes are performed at the device coherence point, so a
// flag written by a wave on another XCD is seen whatever a local cache holds.
__device__ __forceinline__ int ld_coherent(int *p)
{
    return __hip_atomic_fetch_add(p, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void publish(int *p, int v)
{
    __hip_atomic_exchange(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(int *p, int v)
{
    __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-uniform reads for control flow.  A persistent wave's loop must stay
// converged: if its shared state (the task ticket, the abort word) lives in a
// VGPR, the compiler treats every branch on it as divergent and may build a
// loop in which some lanes run on without the others (observed: a ticket
// taken under `if (lane == 0)` and broadcast with a shuffle let lanes 1..63
// loop back to the shuffle with lane 0 parked, re-running one task forever).
// Values read through these helpers sit in SGPRs, so branches on them are
// scalar and the wave moves as one.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Takes the next ticket from a queue counter (the first active lane's atomic).
// ticket_issue sends the atomic and ticket_take reads its result, so a wave
// can request its next ticket while it finishes work that waits on no other
// wave (leader < 0: no request).
struct TicketReq {
    int v;          // per-lane result of the atomic
    int leader;     // the lane that added 1 (wave-uniform)
};
// The leader alone sends the atomic.  The branch holds nothing but the
// atomic and the result is read with readlane, which needs no other lane, so
// no lane can run ahead of the others.  (Files built with the compiler's
// atomic optimizer on turn the atomic into the same single-lane form and
// broadcast its result at once; factor.hip is built with it off, so that
// the result is awaited only where it is read.)
__device__ __forceinline__ TicketReq ticket_issue(int *ctr)
{
    const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
    int v = 0;
    if (lane_id() == leader) v = __hip_atomic_fetch_add(gp(ctr), 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return {v, leader};
}
__device__ __forceinline__ int ticket_take(const TicketReq &r) { return __builtin_amdgcn_readlane(r.v, r.leader); }
__device__ __forceinline__ int wave_ticket(int *ctr) { return ticket_take(ticket_issue(ctr)); }

// Task queues of the persistent kernels.  One counter serving every wave of
// a wide grid bounds the task rate by its atomics (about 85 tasks/us on
// MI355X), so a queue may be split over nq = 1, 2 or 4 counters: task t
// belongs to counter t % nq, wave w of each workgroup takes its tickets from
// counter w % nq, and ticket k of counter q is task k * nq + q.  Every
// counter has waves in every workgroup, and each hands out its tasks in
// order, so the lowest unfinished task is always held by a working wave (its
// counter's waves finish that counter's lower tasks first) and its waits,
// on lower tasks only, end.  Block layout (ints): [0] counter 0, [1] waves
// finished, [32 q] counter q (q >= 1, a cache line each); kQueueInts in all.

__device__ __forceinline__ int queue_count(int nq) { return nq == 2 || nq == 4 ? nq : 1; }

// This wave's counter; wave-uniform: with a per-lane address every lane
// would send its atomic.
__device__ __forceinline__ int queue_index(int nq) { return uniform((int)(threadIdx.x >> 6)) % nq; }

__device__ __forceinline__ int next_task(int *ctr, int nq)
{
    const int q = queue_index(nq);
    return wave_ticket(ctr + 32 * q) * nq + q;
}

// Called once per wave on exit: the last wave out re-arms the counters and
// returns true.
__device__ __forceinline__ bool leave_queue(int *ctr, int nq, int total_waves)
{
    bool last = false;
    if (lane_id() == 0) {
        const int fin = atomicAdd(&ctr[1], 1);
        if (fin == total_waves - 1) {
            for (int q = 0; q < nq; q++) st_relaxed(&ctr[32 * q], 0);
            st_relaxed(&ctr[1], 0);
            last = true;
        }
    }
    return last;
}

// Epochs live in device counters so a launch needs no host-chosen argument and
// a captured graph can be replayed: every wave reads the counter when it
// starts, and the last wave out (after all others have read it) advances it.
__device__ __forceinline__ int read_epoch(const int *counter) { return uniform(ld_relaxed(counter)); }
__device__ __forceinline__ void advance_epoch(int *counter, int epoch)
{
    st_relaxed(counter, epoch < 0x3fffffff ? epoch + 1 : 1);
}

__device__ __forceinline__ bool aborted(const int *abort_word)
{
    return uniform(ld_relaxed(abort_word)) != 0;
}

// Cross-wave hand-off without cache maintenance.  An agent-scope release
// fence writes back the whole L2 (buffer_wbl2) and an acquire fence
// invalidates it (buffer_inv); measured on MI355X they cost ~2.4k / ~1.4k
// clocks alone and up to ~25k / ~17k clocks with a thousand waves fencing.
// Instead, data that another wave reads during the same launch is stored and
// loaded with agent-scope (sc1) accesses, which are coherent across the XCDs'
// L2 caches:
//   producer:  st_coherent(data...); store_complete(); st_relaxed(flag, epoch)
//   consumer:  poll flag (ld_relaxed); ld_coherent_val(data...)
// store_complete waits until this wave's stores are acknowledged, so the
// flag cannot become visible before the data.  The consumer's data loads are
// issued after the flag value has returned (they depend on it through the
// branch), so they cannot read older data.
__device__ __forceinline__ void store_complete() { __builtin_amdgcn_s_waitcnt(0); }

template <typename T> __device__ __forceinline__ void st_coherent(T *p, T v)
{
    __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <> __device__ __forceinline__ void st_coherent<__half>(__half *p, __half v)
{
    __hip_atomic_store(gp((unsigned short *)p), __half_as_ushort(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
// Plain (non-coherent) global store of a factor value.
template <typename T> __device__ __forceinline__ void st_g(T *p, T v) { *gp(p) = v; }
template <> __device__ __forceinline__ void st_g<__half>(__half *p, __half v)
{
    *gp((unsigned short *)p) = __half_as_ushort(v);
}

template <typename T> __device__ __forceinline__ T ld_coherent_val(const T *p)
{
    return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Element i of a wave-uniform base (SGPR base + 32-bit lane offset).
template <typename T> __device__ __forceinline__ T ld_coherent_val(const LSP_GLOBAL T *base, uint32_t i)
{
    return __hip_atomic_load(base + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <> __device__ __forceinline__ __half ld_coherent_val<__half>(const LSP_GLOBAL __half *base, uint32_t i)
{
    return __ushort_as_half(__hip_atomic_load((const LSP_GLOBAL unsigned short *)base + i, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT));
}
template <> __device__ __forceinline__ __half ld_coherent_val<__half>(const __half *p)
{
    return __ushort_as_half(
        __hip_atomic_load(gp((const unsigned short *)p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Element i of a wave-uniform base: agent-coherent (data written during this
// launch) or plain (data written by an earlier launch, which may be cached).
template <typename T> __device__ __forceinline__ T ld_plain(const LSP_GLOBAL T *base, uint32_t i)
{
    return base[i];
}
template <> __device__ __forceinline__ __half ld_plain<__half>(const LSP_GLOBAL __half *base, uint32_t i)
{
    return __ushort_as_half(((const LSP_GLOBAL unsigned short *)base)[i]);
}
template <bool kCoherent, typename T>
__device__ __forceinline__ T ld_val(const LSP_GLOBAL T *base, uint32_t i)
{
    if constexpr (kCoherent) return ld_coherent_val(base, i);
    else return ld_plain(base, i);
}

// "Not yet written" marker of a factor value: a signaling-NaN bit pattern,
// which no arithmetic result has (hardware NaN results are quiet NaNs).
template <typename T> struct Unset;
template <> struct Unset<float> {
    static __device__ __forceinline__ bool is(float v) { return __float_as_uint(v) == 0x7FA5A5A5u; }
    static __device__ __forceinline__ float value() { return __uint_as_float(0x7FA5A5A5u); }
};
template <> struct Unset<double> {
    static __device__ __forceinline__ bool is(double v)
    {
        return (unsigned long long)__double_as_longlong(v) == 0x7FF4A5A5A5A5A5A5ull;
    }
    static __device__ __forceinline__ double value() { return __longlong_as_double(0x7FF4A5A5A5A5A5A5ll); }
};
template <> struct Unset<__half> {
    static __device__ __forceinline__ bool is(__half v) { return __half_as_ushort(v) == 0x7D2Au; }
    static __device__ __forceinline__ __half value() { return __ushort_as_half((unsigned short)0x7D2Au); }
};

// Orders this wave's accesses in the compiler only: the memory operations of
// one wave complete in order, so LDS (and same-address global) traffic between
// lanes of a wave needs no wait.
__device__ __forceinline__ void wave_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Orders this wave's earlier LDS / global accesses before its later ones as
// seen by other lanes of the same wave.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Wave-collective wait: every lane whose flag pointer is non-null waits until
// *flag == epoch; returns true once all of them have seen it, false when the
// wait was abandoned (round limit hit here, or the abort word raised
// elsewhere).  Must be called by all lanes of the wave.
//
// Polling discipline (measured on MI355X): a lane stops reloading once it has
// seen its flag; the wave polls with one relaxed agent-scope load instruction
// per round (an acquire load would also invalidate the cache on every poll);
// rounds back off exponentially up to ~1 us.  Hundreds of waves polling every
// few hundred cycles with all lanes saturate the memory path the producers
// need, and the factorization then crawls.  The caller issues one
// acquire_fence() after a successful wait.
constexpr int kRoundLimit = 1 << 21;

__device__ __forceinline__ bool wave_wait(const int *flag, int epoch, int *abort_word,
                                          int *dbg = nullptr, int who2 = -1, int max_backoff = 7)
{
    int seen = flag ? -1 : epoch;
    int ok = flag == nullptr;
    int backoff = 0;
    for (int round = 0;; round++) {
        if (!ok) {
            seen = ld_relaxed(flag);
            ok = seen == epoch;
        }
        if (__all(ok)) return true;
        if ((round & 15) == 15 && aborted(abort_word)) return false;
        if (round > kRoundLimit) {
            if (!ok && atomicCAS(abort_word, 0, kErrSpin) == 0 && dbg) {
                dbg[1] = seen;
                dbg[2] = epoch;
                dbg[3] = round;
                dbg[4] = who2;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            }
            return false;
        }
        for (int s = 0; s <= backoff; s++) __builtin_amdgcn_s_sleep(4);
        if (backoff < max_backoff) backoff = 2 * backoff + 1;
    }
}

// Wave-collective wait for a growing prefix of lanes: lane l < nd waits on
// its flag (null = nothing to wait for); `ok` holds the lane's seen state
// across calls and starts as (flag == null).  Returns the length p of the
// leading run of lanes whose flags have all been seen, once p > have, or -1
// when the wait was abandoned.  The result is wave-uniform.
__device__ __forceinline__ int wave_wait_prefix(const int *flag, int epoch, int *abort_word, int &ok, int have,
                                                int nd, int *dbg = nullptr, int who2 = -1, int max_backoff = 7)
{
    int seen = epoch;
    int backoff = 0;
    for (int round = 0;; round++) {
        if (!ok) {
            seen = ld_relaxed(flag);
            ok = seen == epoch;
        }
        const unsigned long long miss = ~__ballot(ok);
        const int p = miss ? min(nd, (int)__builtin_ctzll(miss)) : nd;
        if (p > have) return p;
        if ((round & 15) == 15 && aborted(abort_word)) return -1;
        if (round > kRoundLimit) {
            if (!ok && atomicCAS(abort_word, 0, kErrSpin) == 0 && dbg) {
                dbg[1] = seen;
                dbg[2] = epoch;
                dbg[3] = round;
                dbg[4] = who2;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            }
            return -1;
        }
        for (int s = 0; s <= backoff; s++) __builtin_amdgcn_s_sleep(4);
        if (backoff < max_backoff) backoff = 2 * backoff + 1;
    }
}

// The solves poll with one sleep (~0.1 us) between polls: their waits are
// short and their chains long (c1 pair 0.640 -> 0.632 ms against two sleeps,
// 0.662 ms at four; no change on a 512-circuit batch; the factorization's
// cap is ~0.9 us).
constexpr int kSolveBackoff = 0;

// Wave-collective wait on values instead of flags: every lane loads the
// values its non-null pointers name (agent-coherent) until none of them holds
// the unset marker, and returns them in v (0 for null pointers).  Each round
// issues the outstanding loads together, so N chunks of a long dependency
// list cost one round trip instead of N.  A value is
// written once per solve, as a whole, over the marker, so the first value
// that is not the marker is the final one: the producer needs neither its
// store acknowledgement nor a flag, and the consumer has the data with the
// poll that sees it.  Returns false when the wait was abandoned.
template <int N>
__device__ __forceinline__ bool wave_wait_values(const double *const (&p)[N], double (&v)[N], int *abort_word)
{
    unsigned ok = 0;
#pragma unroll
    for (int g = 0; g < N; g++) {
        v[g] = 0.0;
        if (p[g] == nullptr) ok |= 1u << g;
    }
    const unsigned full = (1u << N) - 1;
    int backoff = 0;
    for (int round = 0;; round++) {
#pragma unroll
        for (int g = 0; g < N; g++)
            if (!((ok >> g) & 1)) v[g] = ld_coherent_val(p[g]);
#pragma unroll
        for (int g = 0; g < N; g++)
            if (!((ok >> g) & 1) && !Unset<double>::is(v[g])) ok |= 1u << g;
        if (__all(ok == full)) return true;
        if ((round & 15) == 15 && aborted(abort_word)) return false;
        if (round > kRoundLimit) {
            if (ok != full) atomicCAS(abort_word, 0, kErrSpin);
            return false;
        }
        for (int s = 0; s <= backoff; s++) __builtin_amdgcn_s_sleep(4);
        if (backoff < kSolveBackoff) backoff = 2 * backoff + 1;
    }
}
__device__ __forceinline__ bool wave_wait_value(const double *p, double &v, int *abort_word)
{
    const double *pp[1] = {p};
    double vv[1];
    const bool ok = wave_wait_values<1>(pp, vv, abort_word);
    v = vv[0];
    return ok;
}

// Makes data published before the observed flags visible to this wave.
__device__ __forceinline__ void acquire_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// wave64 reductions (DPP / permute lowered by the compiler from shuffles)
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}
// Would the reference's partial pivoting (splu.h:174-185) pick the fixed
// pivot of this column on the current values?  splu takes the first strict
// maximum of |x| over the rows not yet pivotal (depth-first reach order) and
// then keeps the diagonal row (original row q[k]) instead when its |x| is at
// least tol times that maximum.  Arguments, all magnitudes: pv of the fixed
// pivot, lmax the largest over the other candidate rows, dv of the diagonal
// row when it is one of those rows (negative when it is not a candidate);
// diag_pivot when the fixed pivot is the diagonal row.  Decisions closer than
// a relative 1e-10 to flipping are "uncertain": summation order separates
// these values from splu's, so the host decides those columns.
// Returns 0 agree, 1 disagree, 2 uncertain.
__device__ __forceinline__ int splu_pivot_verdict(bool diag_pivot, double pv, double lmax, double dv,
                                                  double tol)
{
    constexpr double eps = 1e-10;
    if (pv == 0.0) return lmax > 0.0 ? 1 : 0;          // splu: another row, or singular for both
    if (diag_pivot) {                                   // kept iff pv >= tol * max(pv, lmax)
        if (pv >= tol * lmax * (1.0 + eps)) return 0;
        return pv < tol * lmax * (1.0 - eps) ? 1 : 2;
    }
    const double a = fmax(pv, lmax);
    if (dv >= 0.0) {                                    // the diagonal row would win
        if (dv >= tol * a * (1.0 + eps)) return 1;
        if (dv >= tol * a * (1.0 - eps)) return 2;
    }
    if (pv > lmax * (1.0 + eps)) return 0;              // the unique maximum
    return pv < lmax * (1.0 - eps) ? 1 : 2;             // smaller, or a tie broken by reach order
}

__device__ __forceinline__ int wave_min_int(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, kWave));
    return v;
}

template <typename T> __device__ __forceinline__ double to_f64(T v) { return (double)v; }
template <> __device__ __forceinline__ double to_f64<__half>(__half v) { return (double)__half2float(v); }
template <typename T> __device__ __forceinline__ T from_f64(double v) { return (T)v; }
template <> __device__ __forceinline__ __half from_f64<__half>(double v) { return __float2half((float)v); }

// Wave-uniform broadcast of lane `l`'s double (l uniform).
__device__ __forceinline__ double bcast_f64(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

constexpr int kTri = kWave - 1;         // triangle steps held in registers (64-row blocks)

// Unit-lower triangular solve of one run held one entry per lane (lanes ta ..
// tb-1): step q broadcasts x[j_{ta+q}] and subtracts L(j_t, j_{ta+q}) x from
// every later lane t of the run.  The NT triangle columns are all loaded
// before the first step (columns past the run load a valid duplicate and are
// masked), so the chain waits on registers only; a caller may load them
// (load_triangle) before it waits for the right-hand side.
template <int NT, bool kCoherent = true, typename T>
__device__ __forceinline__ void load_triangle(const T *val, int lb, int ta, int tb, T (&tv)[NT])
{
    const int lane = lane_id();
    const int bw = tb - ta;
#pragma unroll
    for (int q = 0; q < NT; q++) {
        const int qq = min(q, bw - 2);
        const int tp = ta + qq;
        const int base = __builtin_amdgcn_readlane(lb, tp);
        const int off = min(max(lane - tp - 1, 0), bw - qq - 2);
        tv[q] = ld_val<kCoherent>(gp(val) + (uint32_t)base, (uint32_t)off);
    }
}
template <int NT, typename T>
__device__ __forceinline__ double apply_triangle(const T (&tv)[NT], double xc, int ta, int tb)
{
    const int lane = lane_id();
#pragma unroll
    for (int q = 0; q < NT; q++) {
        const int tp = min(ta + q, kWave - 1);
        const double ut = bcast_f64(xc, tp);
        const double nx = fma(-to_f64(tv[q]), ut, xc);
        xc = (lane > tp && lane < tb) ? nx : xc;
    }
    return xc;
}
template <int NT, bool kCoherent = true, typename T>
__device__ __forceinline__ double run_triangle(const T *val, double xc, int lb, int ta, int tb)
{
    T tv[NT];
    load_triangle<NT, kCoherent>(val, lb, ta, tb, tv);
    return apply_triangle<NT>(tv, xc, ta, tb);
}
// Upper triangular solve of a block of bw rows held one per lane (row t of
// the block in lane t): column q' of the block's U is U(j_t, j_q') =
// val[dpc_q' - (q' - t)] for t in [q' - ru_q', q'), zero above that run, and
// rd_t = 1 / U(j_t, j_t).  Steps go from the last column to the first: lane
// q' finalises z_q' = num_q' rd_q' and every earlier lane subtracts
// U(j_t, j_q') z_q'.  apply_triangle_upper returns z_t.
template <int NT, bool kCoherent = true, typename T>
__device__ __forceinline__ void load_triangle_upper(const T *val, int dpc, int bw, T (&tv)[NT])
{
    const int lane = lane_id();
#pragma unroll
    for (int q = 0; q < NT; q++) {
        const int qq = min(q, bw - 1);
        const int base = __builtin_amdgcn_readlane(dpc, qq);
        const int off = min(max(qq - lane, 0), qq);
        tv[q] = ld_val<kCoherent>(gp(val) + (uint32_t)(base - off), 0u);
    }
}
template <int NT, typename T>
__device__ __forceinline__ double apply_triangle_upper(const T (&tv)[NT], double num, double rd, int ru, int bw)
{
    const int lane = lane_id();
#pragma unroll
    for (int q = NT - 1; q >= 0; q--) {
        if (q < bw) {
            const double zq = bcast_f64(num * rd, q);
            const int run = __builtin_amdgcn_readlane(ru, q);
            const bool in = lane < q && q - lane <= run;
            const double nn = fma(-to_f64(tv[q]), zq, num);
            num = in ? nn : num;
        }
    }
    return num * rd;
}


}  // namespace lsp
