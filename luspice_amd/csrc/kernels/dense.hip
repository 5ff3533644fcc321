// dense.hip -- the dense trailing block of the factorization and of the
// triangular solves (gfx950).
//
// After a fill-reducing column order, the last columns of a circuit matrix
// (the top separators) are nearly dense: on the 316x316 power grid the last
// ~3000 of 99906 positions hold half of all left-looking updates, and every
// one of their levels has width 1.  A left-looking column there is a chain of
// ~3000 dependent vector steps; factoring the block as a dense matrix turns
// the chain into 64-column panels.
//
// Refactorization (fixed pivot order; no pivot search, as in the sparse part):
//   the persistent refactor kernel applies every column outside the block and
//   leaves S = A22 - L21 U12 (fp64, column-major T x T).  Then, for each
//   64-column panel p:
//     dense_diag_kernel   S[p:p+64, p:p+64] = L11 \ U11       (one wave)
//     dense_trsm_kernel   L21 = A21 U11^-1 (64-row tiles) and
//                         U12 = L11^-1 A12 (64-column tiles)   (one wave each)
//     dense_gemm_kernel   S22 -= L21 U12 (64x64 tiles, fp64 FMA, 4x4 per thread)
//   dense_finish_kernel  applies the pivot test (|pivot| >= ptol max_i |x_i|,
//   i.e. ptol * max |L(:,c)| <= 1), scatters the block into the sparse factor
//   arrays and writes a dense copy in storage precision for the solves.
// Structural zeros stay exactly zero (every product that reaches them has a
// structurally zero factor), so the sparse pattern is reproduced.
//
// Triangular solves: rows of the block are solved after the rows above it
// (forward) and before them (backward):
//   tail_fwd_gather_kernel  y_T = b_T - L21 y  (one wave per row)
//   tail_trsv_kernel<lower> y_T = L22^-1 y_T
//   tail_trsv_kernel<upper> z_T = U22^-1 y_T, x[q] = z, row flags for the
//                           sparse backward solve
// A trsv launch has one workgroup per 64-row block; blocks take tickets in
// solve order and wait (bounded, epoch flags) only on blocks ticketed
// earlier, accumulating each finished block's contribution as it arrives.
#include "common.hpp"
#include "launch.hpp"

namespace lsp {

constexpr int kNb = 64;            // panel width
constexpr int kBsLd = 65;          // padded leading dimension of the GEMM B tile

__device__ __forceinline__ unsigned long long dbits(double v)
{
    return (unsigned long long)__double_as_longlong(v);
}

// Broadcast of lane `src`'s value to the wave (src wave-uniform).
__device__ __forceinline__ double bcast(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffLL), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// ---------------------------------------------------------------------------
// Diagonal block LU, one wave: lane r keeps row r of the 64x64 block in
// registers; row c reaches the other lanes by readlane.  A partial last block
// is padded with identity rows/columns, so every loop is fully unrolled.
__global__ void __launch_bounds__(64) dense_diag_kernel(const DenseJob *jobs, const int *active, int p)
{
    const DenseJob J = jobs[blockIdx.x];
    if (p >= J.T || (active && !active[J.mat])) return;
    const int lane = threadIdx.x;
    const int b = min(kNb, J.T - p);
    const size_t ld = (size_t)J.T;
    double *Sp = J.S + (size_t)p * ld + p;
    const int rr = min(lane, b - 1);
    double a[kNb];
#pragma unroll
    for (int j = 0; j < kNb; j++) {
        const double v = Sp[(size_t)min(j, b - 1) * ld + rr];
        a[j] = (j < b && lane < b) ? v : (j == lane ? 1.0 : 0.0);
    }
#pragma unroll
    for (int c = 0; c < kNb; c++) {
        const double l = a[c] / bcast(a[c], c);
        if (lane > c) a[c] = l;
#pragma unroll
        for (int j = c + 1; j < kNb; j++) {
            const double u = bcast(a[j], c);
            if (lane > c) a[j] -= l * u;
        }
    }
#pragma unroll
    for (int c = 0; c < kNb; c++) {
        const double am = wave_max(lane > c && lane < b ? fabs(a[c]) : 0.0);
        if (lane == 0 && c < b) atomicMax(&J.colmax[p + c], dbits(am));
    }
    if (lane < b) {
#pragma unroll
        for (int j = 0; j < kNb; j++)
            if (j < b) Sp[(size_t)j * ld + lane] = a[j];
    }
}

// ---------------------------------------------------------------------------
// L21 = A21 U11^-1 (64-row tiles) and U12 = L11^-1 A12 (64-column tiles), one
// wave per tile.  The factored diagonal block sits in LDS (read as
// broadcasts); each lane's row (L21) or column (U12) is a padded LDS column,
// so every LDS access is conflict-free, and both tiles are moved between HBM
// and LDS with coalesced accesses (lanes along the rows of a column).
constexpr int kWLd = kNb + 1;                          // padded LDS stride of a tile
constexpr size_t kTrsmLds = sizeof(double) * (kNb * kNb + kNb * kWLd);

__global__ void __launch_bounds__(64) dense_trsm_kernel(const DenseJob *jobs, const int *active, int p)
{
    const DenseJob J = jobs[blockIdx.y];
    if (p >= J.T || (active && !active[J.mat])) return;
    const int b = min(kNb, J.T - p);
    const int rem = J.T - p - b;
    if (rem <= 0) return;
    const int ntile = (rem + kNb - 1) / kNb;
    const int t = blockIdx.x;
    if (t >= 2 * ntile) return;
    extern __shared__ double lds[];
    double *__restrict__ d = lds;                      // d[j*64 + i] = D(i, j), identity-padded
    double *__restrict__ w = lds + kNb * kNb;          // tile, element (row i, col c) at w[c*65 + i]
    const int lane = threadIdx.x;
    const size_t ld = (size_t)J.T;
    const double *Dp = J.S + (size_t)p * ld + p;
    // every load loop is unrolled so its 64 loads are in flight together
    {
        const int rr = min(lane, b - 1);
#pragma unroll
        for (int j = 0; j < kNb; j++) {
            const double v = Dp[(size_t)min(j, b - 1) * ld + rr];
            d[j * kNb + lane] = (j < b && lane < b) ? v : (j == lane ? 1.0 : 0.0);
        }
    }
    if (t < ntile) {
        // rows r0 + lane of L21: x U11 = a; w[c*65 + lane] = x_c of this lane's row
        const int r = p + b + t * kNb + lane;
        const bool ok = r < J.T;
        const int rc = ok ? r : J.T - 1;
#pragma unroll
        for (int c = 0; c < kNb; c++) {
            const double v = J.S[(size_t)(p + min(c, b - 1)) * ld + rc];
            w[c * kWLd + lane] = (ok && c < b) ? v : 0.0;
        }
        wave_fence();
        for (int c = 0; c < b; c++) {
            const double xc = w[c * kWLd + lane] / d[c * kNb + c];
            w[c * kWLd + lane] = xc;
#pragma unroll 8
            for (int j = c + 1; j < b; j++) w[j * kWLd + lane] -= xc * d[j * kNb + c];
        }
        for (int c = 0; c < b; c++) {
            const double v = w[c * kWLd + lane];
            if (ok) J.S[(size_t)(p + c) * ld + r] = v;
            const double am = wave_max(ok ? fabs(v) : 0.0);
            if (lane == 0) atomicMax(&J.colmax[p + c], dbits(am));
        }
    } else {
        // columns c0 + lane of U12: L11 y = a; tile element (row i, column lane)
        // lives at w[i*65 + lane] (transposed on the way in and out)
        const int c0 = p + b + (t - ntile) * kNb;
        const int ncol = min(kNb, J.T - c0);
        const int rl = min(lane, b - 1);
#pragma unroll
        for (int cc = 0; cc < kNb; cc++) {
            const double v = J.S[(size_t)(c0 + min(cc, ncol - 1)) * ld + p + rl];
            w[lane * kWLd + cc] = lane < b ? v : 0.0;
        }
        wave_fence();
        for (int r = 0; r < b; r++) {
            const double yr = w[r * kWLd + lane];
#pragma unroll 8
            for (int i = r + 1; i < b; i++) w[i * kWLd + lane] -= d[r * kNb + i] * yr;
        }
        wave_fence();
        if (lane < b)
            for (int cc = 0; cc < ncol; cc++) J.S[(size_t)(c0 + cc) * ld + p + lane] = w[lane * kWLd + cc];
    }
}

// ---------------------------------------------------------------------------
// S22 -= L21 U12 on 64x64 tiles.  Thread (ty, tx) = (tid % 16, tid / 16) owns
// rows ty + 16 i and columns tx + 16 j (i, j < 4): consecutive lanes touch
// consecutive rows, so the tile loads, the LDS reads and the stores are
// contiguous.
constexpr size_t kGemmLds = sizeof(double) * (kNb * 64 + kNb * kBsLd);

__global__ void __launch_bounds__(256) dense_gemm_kernel(const DenseJob *jobs, const int *active, int p)
{
    const DenseJob J = jobs[blockIdx.y];
    if (p >= J.T || (active && !active[J.mat])) return;
    const int b = min(kNb, J.T - p);
    const int q0 = p + b;
    const int rem = J.T - q0;
    if (rem <= 0) return;
    const int nt = (rem + 63) / 64;
    const int tile = blockIdx.x;
    if (tile >= nt * nt) return;
    const int r0 = q0 + (tile % nt) * 64;
    const int c0 = q0 + (tile / nt) * 64;
    extern __shared__ double lds[];
    double *As = lds;                                  // As[k*64 + i] = L21(r0+i, p+k)
    double *Bs = lds + kNb * 64;                       // Bs[k*65 + j] = U12(p+k, c0+j)
    const int tid = threadIdx.x;
    const size_t ld = (size_t)J.T;
    for (int e = tid; e < b * 64; e += 256) {
        const int k = e >> 6, i = e & 63;
        const int r = r0 + i;
        As[k * 64 + i] = r < J.T ? J.S[(size_t)(p + k) * ld + r] : 0.0;
    }
    for (int e = tid; e < b * 64; e += 256) {
        const int j = e / b, k = e - j * b;
        const int c = c0 + j;
        Bs[k * kBsLd + j] = c < J.T ? J.S[(size_t)c * ld + p + k] : 0.0;
    }
    __syncthreads();
    const int ty = tid & 15, tx = tid >> 4;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
    for (int k = 0; k < b; k++) {
        double a[4], bb[4];
#pragma unroll
        for (int i = 0; i < 4; i++) a[i] = As[k * 64 + ty + 16 * i];
#pragma unroll
        for (int j = 0; j < 4; j++) bb[j] = Bs[k * kBsLd + tx + 16 * j];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = fma(a[i], bb[j], acc[i][j]);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int c = c0 + tx + 16 * j;
        if (c >= J.T) continue;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = r0 + ty + 16 * i;
            if (r < J.T) J.S[(size_t)c * ld + r] -= acc[i][j];
        }
    }
}

// ---------------------------------------------------------------------------
// One wave per block column c: pivot test, scatter of the pattern slots into
// the sparse factors, dense copy in storage precision.
template <typename T>
__global__ void __launch_bounds__(64) dense_finish_kernel(const DenseJob *jobs, const void *mats,
                                                          const int *active, double ptol)
{
    const DenseJob J = jobs[blockIdx.y];
    const int c = blockIdx.x;
    if (c >= J.T || (active && !active[J.mat])) return;
    const FactorDesc<T> M = ((const FactorDesc<T> *)mats)[J.mat];
    const int lane = threadIdx.x;
    const size_t ld = (size_t)J.T;
    const double *Sc = J.S + (size_t)c * ld;
    T *Fc = M.tail_F + (size_t)c * ld;
    for (int r = lane; r < J.T; r += kWave) Fc[r] = from_f64<T>(Sc[r]);
    const int k = M.tail_k0 + c;
    const int64_t c0 = M.cp[k];
    const int m = (int)(M.cp[k + 1] - c0);
    const int dl = M.dp[k] - (int)c0;
    for (int s = M.tail_ds[c] + lane; s < m; s += kWave) {
        const T v = from_f64<T>(Sc[M.pos[c0 + s] - M.tail_k0]);
        M.val[c0 + s] = v;
        if (s < dl) M.urow[M.dst[c0 + s]] = v;
        else if (s == dl) M.diag[k] = v;
        else M.lrow[M.dst[c0 + s]] = v;
    }
    if (lane == 0) {
        const double piv = Sc[c];
        const double lmax = __longlong_as_double((long long)J.colmax[c]);
        if (piv == 0.0) {
            atomicMin(&M.mstat[2], k);
        } else if (!(ptol * lmax <= 1.0)) {
            atomicAdd(&M.mstat[0], 1);
            atomicMin(&M.mstat[1], k);
        }
        if (!(fabs(piv) <= 1.0e308) || !(lmax <= 1.0e308)) atomicAdd(&M.mstat[3], 1);
    }
}

// ---------------------------------------------------------------------------
// Forward solve, rows of the block: y[k0 + r] = b[rowat] - L(k0 + r, :k0) y.
template <typename T>
__global__ void __launch_bounds__(256) tail_fwd_gather_kernel(const void *mats, const int *tails,
                                                              const long *voff, const int *active,
                                                              const double *b, double *y)
{
    const int mi = tails[blockIdx.y];
    if (active && !active[mi]) return;
    const TriDesc<T> M = ((const TriDesc<T> *)mats)[mi];
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= M.tail_T) return;
    const int lane = lane_id();
    const int i = M.tail_k0 + r;
    const long vo = voff[mi];
    const int e1 = M.tail_lcut[r];
    double acc = 0.0;
    for (int e = M.lrp[i] + lane; e < e1; e += kWave) acc += to_f64(M.lrow[e]) * y[vo + M.lci[e]];
    acc = wave_sum(acc);
    if (lane == 0) y[vo + i] = b[vo + M.rowat[i]] - acc;
}

// Dense triangular solve of the block rows (see the file comment).
//   lower: y_T <- L22^-1 y_T (unit diagonal), in place
//   upper: z_T <- U22^-1 y_T, x[q] = z, bdone[row] = epoch of the sparse solve
// Per contributing block the 64 x 64 factor slab (one row per lane, in
// registers) is loaded before its flag is awaited, so the wait overlaps the
// loads; the solved values arrive with one coherent load and are broadcast by
// readlane.  The diagonal block is loaded in bulk before the 64-step solve.
template <typename T, bool kLower>
__global__ void __launch_bounds__(256) tail_trsv_kernel(const void *mats, const int *tails,
                                                        const long *voff, const int *active,
                                                        double *y, double *z, double *x,
                                                        const int *bwd_epoch, int *abort_word)
{
    const int mi = tails[blockIdx.y];
    const TriDesc<T> M = ((const TriDesc<T> *)mats)[mi];
    const int lane = lane_id();
    const int wv = threadIdx.x >> 6;
    __shared__ int s_ticket;
    __shared__ double red[4][kWave];
    const bool live = !(active && !active[mi]);
    const int nblk = (M.tail_T + kNb - 1) / kNb;
    if (live) {
        const int epoch = read_epoch(&M.tail_ctr[2]);            // block flags
        const int bep = kLower ? 0 : read_epoch(bwd_epoch);     // row flags of the sparse solve
        if (threadIdx.x == 0) s_ticket = atomicAdd(&M.tail_ctr[0], 1);
        __syncthreads();
        const int t = uniform(s_ticket);
        if (t < nblk) {
            const int ib = kLower ? t : nblk - 1 - t;
            const int r0 = ib * kNb;
            const int bi = min(kNb, M.tail_T - r0);
            const size_t ld = (size_t)M.tail_T;
            const long vo = voff[mi];
            const int k0 = M.tail_k0;
            const double *src = kLower ? y : z;
            const int rr = r0 + min(lane, bi - 1);               // clamped row of this lane
            double acc = 0.0;
            bool ok = true;
            for (int s = wv; s < t && ok; s += 4) {
                const int jb = kLower ? s : nblk - 1 - s;
                const int c0 = jb * kNb;
                const int bj = min(kNb, M.tail_T - c0);
                T fv[kNb];
#pragma unroll
                for (int c = 0; c < kNb; c++) fv[c] = M.tail_F[(size_t)(c0 + min(c, bj - 1)) * ld + rr];
                ok = wave_wait(lane == 0 ? &M.tail_flag[jb] : nullptr, epoch, abort_word);
                if (!ok) break;
                const double sv = lane < bj ? ld_coherent_val(&src[vo + k0 + c0 + lane]) : 0.0;
#pragma unroll
                for (int c = 0; c < kNb; c++) {
                    const double yc = bcast(sv, c);              // zero beyond bj
                    acc += to_f64(fv[c]) * yc;
                }
            }
            red[wv][lane] = lane < bi ? acc : 0.0;
            __syncthreads();
            const bool all_ok = __syncthreads_and(ok);
            if (all_ok && wv == 0) {
                // row lane of the diagonal block: dg[c] = F(r0 + lane, r0 + c)
                T dg[kNb];
#pragma unroll
                for (int c = 0; c < kNb; c++) dg[c] = M.tail_F[(size_t)(r0 + min(c, bi - 1)) * ld + rr];
                double tr = 0.0;
                if (lane < bi)
                    tr = y[vo + k0 + r0 + lane] - (red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane]);
                if (kLower) {
#pragma unroll
                    for (int c = 0; c < kNb; c++) {
                        const double yc = bcast(tr, c);
                        if (lane > c && c < bi) tr -= to_f64(dg[c]) * yc;
                    }
                    if (lane < bi) st_coherent(&y[vo + k0 + r0 + lane], tr);
                } else {
#pragma unroll
                    for (int c = kNb - 1; c >= 0; c--) {
                        if (c < bi) {
                            const double zc = bcast(tr, c) / bcast(to_f64(dg[c]), c);
                            if (lane == c) tr = zc;
                            if (lane < c) tr -= to_f64(dg[c]) * zc;
                        }
                    }
                    if (lane < bi) {
                        const int row = k0 + r0 + lane;
                        st_coherent(&z[vo + row], tr);
                        x[vo + M.q[row]] = tr;
                    }
                }
                store_complete();
                if (!kLower && lane < bi) st_relaxed(&M.bdone[k0 + r0 + lane], bep);
                if (lane == 0) st_relaxed(&M.tail_flag[ib], epoch);
            }
        }
        // the last workgroup of this matrix re-arms its ticket counter
        if (threadIdx.x == 0) {
            const int fin = atomicAdd(&M.tail_ctr[1], 1);
            if (fin == (int)gridDim.x - 1) {
                st_relaxed(&M.tail_ctr[0], 0);
                st_relaxed(&M.tail_ctr[1], 0);
                advance_epoch(&M.tail_ctr[2], epoch);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
hipError_t dense_set_lds()
{
    hipError_t e = hipFuncSetAttribute((const void *)dense_gemm_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGemmLds);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void *)dense_trsm_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kTrsmLds);
}

hipError_t launch_dense_lu(const DenseJob *jobs, int njobs, int maxT, const int *active, hipStream_t st)
{
    for (int p = 0; p < maxT; p += kNb) {
        const int rem = maxT - p - min(kNb, maxT - p);
        hipLaunchKernelGGL(dense_diag_kernel, dim3(njobs), dim3(64), 0, st, jobs, active, p);
        if (rem > 0) {
            const int nt = (rem + kNb - 1) / kNb;
            hipLaunchKernelGGL(dense_trsm_kernel, dim3(2 * nt, njobs), dim3(64), kTrsmLds, st, jobs, active, p);
            hipLaunchKernelGGL(dense_gemm_kernel, dim3(nt * nt, njobs), dim3(256), kGemmLds, st, jobs,
                               active, p);
        }
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_dense_finish(const DenseJob *jobs, int njobs, int maxT, const void *mats,
                               const int *active, double ptol, hipStream_t st)
{
    hipLaunchKernelGGL(dense_finish_kernel<T>, dim3(maxT, njobs), dim3(64), 0, st, jobs, mats, active, ptol);
    return hipGetLastError();
}
template hipError_t launch_dense_finish<float>(const DenseJob *, int, int, const void *, const int *, double, hipStream_t);
template hipError_t launch_dense_finish<double>(const DenseJob *, int, int, const void *, const int *, double, hipStream_t);
template hipError_t launch_dense_finish<__half>(const DenseJob *, int, int, const void *, const int *, double, hipStream_t);

template <typename T>
hipError_t launch_tail_solve(const void *mats, const int *tails, int ntails, int maxT, const long *voff,
                             const int *active, const double *b, double *y, double *z, double *x,
                             bool forward, const int *bwd_epoch, int *abort_word, hipStream_t st)
{
    const int nblk = (maxT + kNb - 1) / kNb;
    if (forward) {
        hipLaunchKernelGGL(tail_fwd_gather_kernel<T>, dim3((maxT + 3) / 4, ntails), dim3(256), 0, st, mats,
                           tails, voff, active, b, y);
        hipLaunchKernelGGL((tail_trsv_kernel<T, true>), dim3(nblk, ntails), dim3(256), 0, st, mats, tails,
                           voff, active, y, z, x, bwd_epoch, abort_word);
    } else {
        hipLaunchKernelGGL((tail_trsv_kernel<T, false>), dim3(nblk, ntails), dim3(256), 0, st, mats, tails,
                           voff, active, y, z, x, bwd_epoch, abort_word);
    }
    return hipGetLastError();
}
template hipError_t launch_tail_solve<float>(const void *, const int *, int, int, const long *, const int *,
                                             const double *, double *, double *, double *, bool, const int *, int *,
                                             hipStream_t);
template hipError_t launch_tail_solve<double>(const void *, const int *, int, int, const long *, const int *,
                                              const double *, double *, double *, double *, bool, const int *, int *,
                                              hipStream_t);
template hipError_t launch_tail_solve<__half>(const void *, const int *, int, int, const long *, const int *,
                                              const double *, double *, double *, double *, bool, const int *, int *,
                                              hipStream_t);

}  // namespace lsp
