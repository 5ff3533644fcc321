// trisolve.hip -- level-ordered, synchronisation-free row-gather triangular
// solves x = Q U^-1 L^-1 P b with fp32/fp64/fp16 factors and fp64
// accumulation (gfx950).
//
// Reference behaviour: lu_solve_perm (src/third_party/splu.h:210-239) and the
// fp32-factor trisolves (src/lu_gpu_f32.cu:94-172: factors read as double,
// fp64 accumulation, pb = P b, x[q[k]] = z[k]).  The reference's
// level-scheduled form (src/mp/fusion_bench.cu:80-116) runs one block with a
// barrier per level; here rows are tasks in a global queue sorted by level,
// each row waits only for the values it reads, and every CU works on the
// front of the dependency DAG at once.
//
//   forward   y[i] = b[rowat[i]] - sum_{j<i} L(i,j) y[j]
//   backward  z[i] = (y[i] - sum_{j>i} U(i,j) z[j]) / U(i,i),  x[q[i]] = z[i]
//
// One wave per row: lanes stride over the row's entries (coalesced index and
// value loads), wait on their own dependency values in parallel, and the wave
// sums in fp64 with a DPP reduction.  Rows of a supernode (consecutive
// columns whose L patterns nest, so their diagonal block is dense) first
// gather only their entries outside it, in parallel; one task per supernode
// then solves the dense block in registers, so the solve has one dependency
// level per supernode instead of one per row.
//
// Hand-off by data: before a solve every work value (y, z and the partial
// sums yp, zp) is set to the unset marker; a task stores each value once,
// agent-coherent, and its readers load the values they need until none is
// the marker (wave_wait_values).  No flag, no store acknowledgement on the
// producer's side, and the poll that sees a value already holds it.  No L2
// write-back or invalidate is needed (see common.hpp).
#include "common.hpp"
#include "launch.hpp"

namespace lsp {


constexpr int kGather = 4;              // 64-entry chunks of a row per wait round

// Values of an earlier (later) block's 64 columns loaded per pass of a
// triangle task: all 64 for 4-byte and 2-byte factors, so that the latest
// block's slab is in registers before its values are awaited; two passes of
// 32 for fp64 (register budget).
template <typename T> constexpr int kSolveBatch = sizeof(T) <= 4 ? 64 : 32;

// Forward triangle task: block b (rows s + 64b .. s + 64b + bw - 1) of the
// supernode D = [s, e).  Its rows' partial sums p_i = b_i - sum over L
// entries left of s are in yp; the earlier blocks of D, each solved by a task
// of its own (ticketed before this one), are subtracted first (L(i, j) =
// val[dp[j] + i - j]), 32 columns at a time, their L values loaded before
// their y values are awaited; then the block's unit-lower triangle, loaded
// before any wait (NT steps, NT = 0 for one row), is solved in registers.  A
// supernode of at most 64 rows is one block.
template <int NT, typename T>
__device__ __forceinline__ void tri_forward_nt(const TriDesc<T> &M, int s, int b, int w, double *y,
                                               const double *yp, int *abort_word)
{
    const int lane = lane_id();
    const LSP_GLOBAL int *dp = gp(M.dp);
    const int b0 = kWave * b;
    const int bw = min(kWave, w - b0);
    const int row = s + b0 + lane;
    const int rr = s + b0 + min(lane, bw - 1);
    T tv[NT > 0 ? NT : 1];
    if constexpr (NT > 0) load_triangle<NT, false>(M.val, dp[rr] + 1, 0, bw, tv);
    constexpr int SB = kSolveBatch<T>;
    double acc = 0.0;
    for (int q0 = 0; q0 < b0; q0 += kWave) {            // earlier blocks are full; the last one is the latest
        double yq = 0.0;
        for (int h0 = 0; h0 < kWave; h0 += SB) {
            T v[SB];
#pragma unroll
            for (int h = 0; h < SB; h++) {
                const int jq = s + q0 + h0 + h;
                v[h] = ld_plain(gp(M.val), (uint32_t)(dp[jq] + (rr - jq)));   // written by the refactor
            }
            if (h0 == 0 && !wave_wait_value(&y[s + q0 + lane], yq, abort_word)) return;
#pragma unroll
            for (int h = 0; h < SB; h++) acc = fma(to_f64(v[h]), bcast_f64(yq, h0 + h), acc);
        }
    }
    double p;
    if (!wave_wait_value(lane < bw ? &yp[row] : nullptr, p, abort_word)) return;
    p -= acc;
    if constexpr (NT > 0) p = apply_triangle<NT>(tv, p, 0, bw);
    if (lane < bw) st_coherent(&y[row], p);
}
template <typename T>
__device__ void tri_forward(const TriDesc<T> &M, int s, int b, double *y, const double *yp, int *abort_word)
{
    const int w = M.snend[s] - s;
    const int bw = min(kWave, w - kWave * b);
    if (bw == 1) tri_forward_nt<0>(M, s, b, w, y, yp, abort_word);
    else if (bw <= 9) tri_forward_nt<8>(M, s, b, w, y, yp, abort_word);
    else if (bw <= 25) tri_forward_nt<24>(M, s, b, w, y, yp, abort_word);
    else tri_forward_nt<kTri>(M, s, b, w, y, yp, abort_word);
}

// Backward triangle task: block b from the bottom of D = [s, e).  Its rows'
// partial sums q_i = y_i - sum over U entries right of e are in zp; the
// later blocks of D (solved by earlier tasks) are subtracted first, the
// oldest (the bottom one) first and the one just below this block last
// (U(i, j) = val[dp[j] - (j - i)] within column j's run ru[j], else 0); then
// the block's upper triangle, loaded before any wait, is solved and the
// solution written.
template <int NT, typename T>
__device__ __forceinline__ void tri_backward_nt(const TriDesc<T> &M, int s, int b, int w, double *z,
                                                const double *zp, double *x, const double *cs, int *abort_word)
{
    const int lane = lane_id();
    const LSP_GLOBAL int *dp = gp(M.dp);
    const LSP_GLOBAL int *ru = gp(M.ru);
    const int b1 = w - kWave * b;                        // rows [s, s + b1) not yet solved
    const int bw = min(kWave, b1);
    const int base = s + b1 - bw;
    const int row = base + lane;
    const int rr = base + min(lane, bw - 1);
    const int c0 = base + bw;                            // first column of the later blocks
    const int dpc = dp[rr];
    T tv[NT];
    load_triangle_upper<NT, false>(M.val, dpc, bw, tv);
    const double rd = 1.0 / to_f64(M.diag[rr]);
    const int run = ru[rr];
    constexpr int SB = kSolveBatch<T>;
    double acc = 0.0;
    for (int k = b - 1; k >= 0; k--) {                   // later blocks are full
        const int q0 = kWave * k;
        double zq = 0.0;
        for (int h0 = 0; h0 < kWave; h0 += SB) {
            T v[SB];
            int top[SB];                                 // column jq's run covers rows top .. jq - 1
#pragma unroll
            for (int h = 0; h < SB; h++) {
                const int jq = c0 + q0 + h0 + h;
                const int rj = ru[jq];
                top[h] = jq - rj;
                v[h] = ld_plain(gp(M.val), (uint32_t)(dp[jq] - min(jq - rr, rj)));
            }
            if (h0 == 0 && !wave_wait_value(&z[c0 + q0 + lane], zq, abort_word)) return;
#pragma unroll
            for (int h = 0; h < SB; h++)
                if (rr >= top[h]) acc = fma(to_f64(v[h]), bcast_f64(zq, h0 + h), acc);
        }
    }
    double num;
    if (!wave_wait_value(lane < bw ? &zp[row] : nullptr, num, abort_word)) return;
    num -= acc;
    const double zt = apply_triangle_upper<NT>(tv, num, rd, run, bw);
    if (lane < bw) {
        st_coherent(&z[row], zt);
        const int qo = M.q[row];
        x[qo] = cs ? zt * cs[qo] : zt;
    }
}
template <typename T>
__device__ void tri_backward(const TriDesc<T> &M, int s, int b, double *z, const double *zp, double *x,
                             const double *cs, int *abort_word)
{
    const int w = M.snend[s] - s;
    const int bw = min(kWave, w - kWave * b);
    if (bw <= 8) tri_backward_nt<8>(M, s, b, w, z, zp, x, cs, abort_word);
    else if (bw <= 24) tri_backward_nt<24>(M, s, b, w, z, zp, x, cs, abort_word);
    else tri_backward_nt<kWave>(M, s, b, w, z, zp, x, cs, abort_word);
}

// Packet task (code kTriPacket + o): up to 64 / g row tasks of one level,
// each with at most g kTriPacketLane entries (g = 2, 4, 8, 16 or 32); lanes
// [r g, (r + 1) g) gather row r, lane l taking entries l, l + g, ...  The
// rows' values are polled together, each lane group sums with g-lane
// shuffles, and every row is written as its own task would be.
template <typename T, bool kForward>
__device__ void tri_packet(const TriLaunch &L, const TriDesc<T> &M, int o, long vo)
{
    const int lane = lane_id();
    const LSP_GLOBAL int *pk = gp(kForward ? M.fpk : M.bpk) + o;
    const int g = uniform(pk[0]);
    const int cnt = uniform(pk[1]);
    const int grp = lane >> __builtin_ctz(g);
    const int l = lane & (g - 1);
    const int rc = grp < cnt ? pk[2 + grp] : -1;
    const int i = rc >> 1;
    const bool part = rc & 1;
    int e0 = 0, e1 = 0;
    if (rc >= 0) {
        if (kForward) {
            e0 = M.lrp[i];
            e1 = part ? M.fcut[i] : M.lrp[i + 1];
        } else {
            e0 = part ? M.bcut[i] : M.urp[i];
            e1 = M.urp[i + 1];
        }
    }
    const int *ci = kForward ? M.lci : M.uci;
    const T *rv = kForward ? M.lrow : M.urow;
    double *y = L.y + vo;
    double *z = L.z + vo;
    const double *src = kForward ? y : z;
    const bool lead = rc >= 0 && l == 0;
    // b (forward), y and the diagonal (backward) before the waits
    double pre = 0.0, dia = 1.0;
    if (lead) {
        if (kForward) {
            const int ro = M.rowat[i];
            pre = L.rs ? L.b[vo + ro] * L.rs[vo + ro] : L.b[vo + ro];
        } else {
            pre = y[i];
            if (!part) dia = to_f64(M.diag[i]);
        }
    }
    T vv[kTriPacketLane];
    const double *sp[kTriPacketLane];
#pragma unroll
    for (int h = 0; h < kTriPacketLane; h++) {
        const int e = e0 + l + h * g;
        const bool on = e < e1;
        vv[h] = on ? rv[e] : T(0);
        sp[h] = on ? &src[ci[e]] : nullptr;
    }
    double sv[kTriPacketLane];
    if (!wave_wait_values<kTriPacketLane>(sp, sv, L.abort_word)) return;
    double acc = 0.0;
#pragma unroll
    for (int h = 0; h < kTriPacketLane; h++) acc += to_f64(vv[h]) * sv[h];
    for (int off = g >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kWave);
    const double r = kForward || part ? pre - acc : (pre - acc) / dia;
    if (lead) {
        if (part) st_coherent(kForward ? &L.yp[vo + i] : &L.zp[vo + i], r);
        else st_coherent(kForward ? &y[i] : &z[i], r);
        if (!kForward && !part) {                      // the solution
            const int qo = M.q[i];
            L.x[vo + qo] = L.cs ? r * L.cs[vo + qo] : r;
        }
    }
}

// Packet of small supernode triangles (header -g, g = 4, 8 or 16): up to
// 64 / g supernodes of one level with at most g rows each; lanes
// [r g, (r + 1) g) hold the rows of supernode r.  Every row's partial sum is
// awaited, the dense diagonal block's entries are loaded before the first
// step, and the triangle is solved with g-lane shuffles.
constexpr int kTriPacketRows = 16;      // widest supernode in a triangle packet

template <typename T, bool kForward>
__device__ void tri_tpacket(const TriLaunch &L, const TriDesc<T> &M, int o, long vo)
{
    const int lane = lane_id();
    const LSP_GLOBAL int *pk = gp(kForward ? M.fpk : M.bpk) + o;
    const int g = uniform(-pk[0]);
    const int cnt = uniform(pk[1]);
    const int lg = __builtin_ctz(g);
    const int grp = lane >> lg;
    const int l = lane & (g - 1);
    const int s = grp < cnt ? pk[2 + grp] : 0;
    const int w = grp < cnt ? M.snend[s] - s : 0;
    const bool on = l < w;
    const int row = s + min(l, max(w - 1, 0));
    const int src0 = grp << lg;                         // first lane of this group
    const LSP_GLOBAL int *dp = gp(M.dp);
    if (kForward) {
        // L(s + l, s + q) = val[dp[s + q] + (l - q)] for q < l
        T tv[kTriPacketRows - 1];
#pragma unroll
        for (int q = 0; q < kTriPacketRows - 1; q++) {
            const int j = s + min(q, max(w - 1, 0));
            const int off = on && q < l ? l - q : 0;
            tv[q] = ld_plain(gp(M.val), (uint32_t)(dp[j] + off));
        }
        double p;
        if (!wave_wait_value(on ? &L.yp[vo + row] : nullptr, p, L.abort_word)) return;
#pragma unroll
        for (int q = 0; q < kTriPacketRows - 1; q++) {
            if (q < g - 1) {
                const double yq = __shfl(p, src0 + q, kWave);
                if (on && q < l) p = fma(-to_f64(tv[q]), yq, p);
            }
        }
        if (on) st_coherent(&L.y[vo + row], p);
    } else {
        // U(s + l, s + q) = val[dp[s + q] - (q - l)] for l < q within column
        // s + q's run ru, zero above it
        const LSP_GLOBAL int *ru = gp(M.ru);
        T tv[kTriPacketRows];
        bool in[kTriPacketRows];
#pragma unroll
        for (int q = 0; q < kTriPacketRows; q++) {
            const int j = s + min(q, max(w - 1, 0));
            in[q] = on && q < w && l < q && q - l <= ru[j];
            tv[q] = ld_plain(gp(M.val), (uint32_t)(dp[j] - (in[q] ? q - l : 0)));
        }
        const double rd = on ? 1.0 / to_f64(M.diag[row]) : 1.0;
        double num;
        if (!wave_wait_value(on ? &L.zp[vo + row] : nullptr, num, L.abort_word)) return;
#pragma unroll
        for (int q = kTriPacketRows - 1; q >= 0; q--) {
            if (q < g) {
                const double zq = __shfl(num, src0 + q, kWave) * __shfl(rd, src0 + q, kWave);
                if (in[q]) num = fma(-to_f64(tv[q]), zq, num);
            }
        }
        const double zt = num * rd;
        if (on) {
            st_coherent(&L.z[vo + row], zt);
            const int qo = M.q[row];                   // the solution
            L.x[vo + qo] = L.cs ? zt * L.cs[vo + qo] : zt;
        }
    }
}

// One task of the solve (ticket t).
template <typename T, bool kForward>
__device__ void tri_task(const TriLaunch &L, int t)
{
    const int lane = lane_id();
    const TriDesc<T> *mats = (const TriDesc<T> *)L.mats;
    const int2 tk = L.task[t];
    if (L.active && !L.active[tk.x]) return;
    const TriDesc<T> M = mats[tk.x];
    const long vo = L.voff[tk.x];
    double *y = L.y + vo;
    double *z = L.z + vo;
    const int code = tk.y;
    if (code >= kTriPacket) {
        const int o = code - kTriPacket;
        if (uniform(gp(kForward ? M.fpk : M.bpk)[o]) < 0) tri_tpacket<T, kForward>(L, M, o, vo);
        else tri_packet<T, kForward>(L, M, o, vo);
        return;
    }
    if (code < 0 || code >= kTriWide) {
        int s0 = -code - 1, bk = 0;                      // supernode, block in solve order
        if (code >= 0) {
            const LSP_GLOBAL int *tw = gp(M.twide) + 2 * (code - kTriWide);
            s0 = uniform(tw[0]);
            bk = uniform(tw[1]);
        }
        if (kForward) tri_forward<T>(M, s0, bk, y, L.yp + vo, L.abort_word);
        else tri_backward<T>(M, s0, bk, z, L.zp + vo, L.x + vo, L.cs ? L.cs + vo : nullptr, L.abort_word);
        return;
    }
    const int i = code >> 1;
    const bool part = code & 1;
    const int *ci = kForward ? M.lci : M.uci;
    const T *rv = kForward ? M.lrow : M.urow;
    const double *src = kForward ? y : z;
    int e0, e1;
    if (kForward) {
        e0 = M.lrp[i];
        e1 = part ? M.fcut[i] : M.lrp[i + 1];
    } else {
        e0 = part ? M.bcut[i] : M.urp[i];
        e1 = M.urp[i + 1];
    }
    // what the row's result needs besides its sum -- b (forward), y and the
    // diagonal (backward) -- is loaded before the waits, off the chain
    double pre = 0.0, dia = 1.0;
    if (lane == 0) {
        if (kForward) {
            const int ro = M.rowat[i];
            pre = L.rs ? L.b[vo + ro] * L.rs[vo + ro] : L.b[vo + ro];
        } else {
            pre = y[i];
            if (!part) dia = to_f64(M.diag[i]);
        }
    }
    double acc = 0.0;
    for (int b0 = e0; b0 < e1; b0 += kGather * kWave) {
        const double *sp[kGather];
        T vv[kGather];
#pragma unroll
        for (int g = 0; g < kGather; g++) {
            const int e = b0 + g * kWave + lane;
            vv[g] = e < e1 ? rv[e] : T(0);
            sp[g] = e < e1 ? &src[ci[e]] : nullptr;
        }
        double sv[kGather];
        if (!wave_wait_values<kGather>(sp, sv, L.abort_word)) return;
#pragma unroll
        for (int g = 0; g < kGather; g++) acc += to_f64(vv[g]) * sv[g];
    }
    acc = wave_sum(acc);
    if (lane == 0) {
        const double r = kForward || part ? pre - acc : (pre - acc) / dia;
        if (part) st_coherent(kForward ? &L.yp[vo + i] : &L.zp[vo + i], r);
        else st_coherent(kForward ? &y[i] : &z[i], r);
        if (!kForward && !part) {                      // the solution
            const int qo = M.q[i];
            L.x[vo + qo] = L.cs ? r * L.cs[vo + qo] : r;
        }
    }
}

// Task codes (see the plan): 2i = row i solved whole, 2i+1 = row i's entries
// outside its supernode into a partial sum, -(s+1) = supernode triangle,
// kTriWide + x = one 64-row block of a wide supernode, kTriPacket + o =
// packet of short rows or small triangles.
template <typename T, bool kForward>
__device__ __forceinline__ void trisolve_body(const TriLaunch &L)
{
    // A Krylov iteration after every circuit converged: nothing to solve.
    // Every wave of the grid reads the same count (written by an earlier
    // launch) and returns at once: no ticket and no finish count.
    if (L.n_active && uniform(ld_relaxed(L.n_active)) == 0) return;
    const int nq = queue_count(L.nq);
    for (;;) {
        const int t = next_task(L.ctr, nq);
        if (t >= L.ntask) break;
        if (aborted(L.abort_word)) continue;
        const long long tg = L.timeline ? (long long)wall_clock64() : 0;
        tri_task<T, kForward>(L, t);
        if (L.timeline && lane_id() == 0) {
            // diagnostic: ticket granted, task finished (wall clock)
            L.timeline[2L * t] = tg;
            L.timeline[2L * t + 1] = (long long)wall_clock64();
        }
    }
    leave_queue(L.ctr, nq, L.total_waves);             // the last wave out re-arms the queue
}

template <typename T>
__global__ void __launch_bounds__(256) trisolve_fwd_kernel(TriLaunch L)
{
    trisolve_body<T, true>(L);
}
template <typename T>
__global__ void __launch_bounds__(256) trisolve_bwd_kernel(TriLaunch L)
{
    trisolve_body<T, false>(L);
}

// Sets the solve's work values (y, yp, z, zp) to the unset marker; skipped
// when no circuit is active (the solves then return at once as well).
__global__ void __launch_bounds__(256) solve_unset_kernel(double *y, double *yp, double *z, double *zp, long n,
                                                          const int *n_active)
{
    if (n_active && uniform(ld_relaxed(n_active)) == 0) return;
    const double u = Unset<double>::value();
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
        y[i] = u;
        yp[i] = u;
        z[i] = u;
        zp[i] = u;
    }
}

hipError_t launch_solve_unset(const TriLaunch &L, long n, hipStream_t st)
{
    const long blocks = std::max(1L, std::min(4096L, (n + 255) / 256));
    hipLaunchKernelGGL(solve_unset_kernel, dim3((unsigned)blocks), dim3(256), 0, st, L.y, L.yp, L.z, L.zp, n,
                       L.n_active);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_trisolve_half(const TriLaunch &L, bool forward, int blocks, hipStream_t st)
{
    if (forward)
        hipLaunchKernelGGL(trisolve_fwd_kernel<T>, dim3(blocks), dim3(256), 0, st, L);
    else
        hipLaunchKernelGGL(trisolve_bwd_kernel<T>, dim3(blocks), dim3(256), 0, st, L);
    return hipGetLastError();
}

template hipError_t launch_trisolve_half<float>(const TriLaunch &, bool, int, hipStream_t);
template hipError_t launch_trisolve_half<double>(const TriLaunch &, bool, int, hipStream_t);
template hipError_t launch_trisolve_half<__half>(const TriLaunch &, bool, int, hipStream_t);

}  // namespace lsp
