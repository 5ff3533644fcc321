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
//   leaves S = A22 - L21 U12 (fp64, column-major T x T).  Then, for the
//   64-column panels p:
//     dense_diag_kernel   (panel 0) S[p:p+64, p:p+64] = L11 \ U11, and
//                         L11^-1, U11^-1
//     dense_trsm_kernel   (panel 0) L21 = A21 U11^-1 (64-row tiles) and
//                         U12 = L11^-1 A12 (64-column tiles), as tile products
//     dense_gemm_kernel   S22 -= L21 U12 (64x64 tiles, fp64 matrix cores); its
//                         first tile then factors the next panel's diagonal
//                         block, and the next panel's L21 / U12 tiles wait
//                         for it and multiply by its inverses, so each later
//                         panel costs one launch
//   dense_finish_kernel  applies the pivot test (|pivot| >= ptol max_i |x_i|,
//   i.e. ptol * max |L(:,c)| <= 1), scatters the block into the sparse factor
//   arrays and writes a dense copy in storage precision for the solves.
// Structural zeros stay exactly zero (every product that reaches them has a
// structurally zero factor), so the sparse pattern is reproduced.
//
// Triangular solves: rows of the block are solved after the rows above it
// (forward) and before them (backward):
//   tail_fwd_gather_kernel  yp_T = b_T - L21 y  (one wave per row)
//   tail_trsv_kernel<lower> y_T = L22^-1 yp_T
//   tail_trsv_kernel<upper> z_T = U22^-1 y_T, x[q] = z
// A trsv launch has one workgroup per two 64-row blocks; workgroups take
// tickets in solve order and wait (bounded) only on blocks ticketed earlier,
// accumulating each finished block's contribution as it arrives.  Blocks
// hand over their values as the sparse solves do: y and z are unset before
// the solve and a waiting block polls the values themselves.
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
// Diagonal block: LU of the 64x64 block and the inverses of both factors,
// blocked by 16-column sub-panels (256 threads).  Thread (r, s) = (tid % 16,
// tid / 16) keeps elements (r + 16a, s + 16b), a, b < 4, of A, of X = L^-1
// and of Y = U^-1 in registers; a partial last block is padded with identity
// rows and columns.  For each sub-panel k (columns k0 = 16k .. k0 + 15):
//   1. the owners publish, in LDS, the panel columns (rows >= k0), the block
//      row [R | I | A12] (R: rows k0.. of X left of k0; A12: rows k0.. of A
//      right of the panel) and delta - Yacc over the panel columns;
//   2. wave 0 factors the 64 x 16 panel, one row per lane: step c broadcasts
//      the pivot row by readlane, every lane below it forms l = A(i, c) / p
//      (times the reciprocal) and updates its row -- the same operations, in
//      the same order, as the unblocked right-looking elimination, so the
//      factor is the same bit for bit;
//   3. wave 1 applies L11^-1 to the block row, one column per lane: D R, D =
//      L11^-1 and U12 = L11^-1 A12 in one forward substitution; wave 2 solves
//      V U11 = delta - Yacc, one row per lane, for the panel columns of Y;
//   4. every thread takes its final panel values and applies the rank-16
//      updates: A22 -= L21 U12, the rows of X below the panel -= L21 [D R | D]
//      (X <- M_k^-1 X for the elementary block factor M_k of L), and
//      Yacc(:, right of the panel) += V U12.
// Four barriers per sub-panel instead of one per column.  Outputs: L\U over
// the block in S, L^-1 (unit lower) and U^-1 (upper) into the job's
// workspace W (column-major, leading dimension 64), and the column maxima of
// |L| for the pivot test.
constexpr int kDgLs = 80;                  // row stride of the panel and delta - Yacc buffers
constexpr int kDgRs = 68;                  // row stride of the block-row buffer
constexpr int kDgLp = 0;                   // [16][80]: A(i, k0 + q) at q * 80 + i
constexpr int kDgRow = 16 * kDgLs;         // [16][68]: block row at q * 68 + j
constexpr int kDgW = kDgRow + 16 * kDgRs;  // [16][80]: delta - Yacc(i, k0 + q), then V, at q * 80 + i
constexpr int kDgRp = kDgW + 16 * kDgLs;   // [16]: pivot reciprocals of the sub-panel
constexpr size_t kDiagLds = sizeof(double) * (kDgRp + 16);

__device__ __forceinline__ double max16(double v)
{
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}

// The block body, run by all 256 threads of a workgroup: panel p of job J,
// its input read from src (column-major, leading dimension lds: the block in
// S itself, or a copy in LDS that the caller has made visible with a
// barrier); sm is kDiagLds bytes of LDS, which may overlap src.
__device__ __forceinline__ void diag_block(const DenseJob &J, int p, const double *src, size_t lds, double *sm)
{
    const int tid = threadIdx.x;
    const int r = tid & 15, s = tid >> 4;
    const int lane = tid & 63, wid = tid >> 6;
    const int b = min(kNb, J.T - p);
    const size_t ld = (size_t)J.T;
    double *Sp = J.S + (size_t)p * ld + p;
    double A[4][4], X[4][4], Y[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const int i = r + 16 * a;
#pragma unroll
        for (int bb = 0; bb < 4; bb++) {
            const int j = s + 16 * bb;
            const double v = src[(size_t)min(j, b - 1) * lds + min(i, b - 1)];
            A[a][bb] = (i < b && j < b) ? v : (i == j ? 1.0 : 0.0);
            X[a][bb] = 0.0;
            Y[a][bb] = 0.0;
        }
    }
    __syncthreads();                                   // src read by every thread: sm may overlap it
    double *Lp = sm + kDgLp, *Rw = sm + kDgRow, *Wv = sm + kDgW, *Rp = sm + kDgRp;
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
        const int k0 = 16 * kb;
        // 1. publish
#pragma unroll
        for (int a = kb; a < 4; a++) Lp[s * kDgLs + r + 16 * a] = A[a][kb];
#pragma unroll
        for (int bb = 0; bb < 4; bb++)
            Rw[r * kDgRs + s + 16 * bb] = bb < kb ? X[kb][bb] : bb == kb ? (r == s ? 1.0 : 0.0) : A[kb][bb];
#pragma unroll
        for (int a = 0; a <= kb; a++) Wv[s * kDgLs + r + 16 * a] = (a == kb && r == s ? 1.0 : 0.0) - Y[a][kb];
        __syncthreads();
        // 2. panel LU on wave 0 (rows above k0 are read and left alone)
        if (wid == 0) {
            double v[16];
#pragma unroll
            for (int q = 0; q < 16; q++) v[q] = Lp[q * kDgLs + lane];
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int c = k0 + q;
                const double rp = 1.0 / bcast(v[q], c);
                const double l = lane > c ? v[q] * rp : 0.0;
                if (lane > c) v[q] = l;
#pragma unroll
                for (int j = q + 1; j < 16; j++) v[j] = fma(-l, bcast(v[j], c), v[j]);
                if (lane == 0) Rp[q] = rp;
            }
            if (lane >= k0) {
#pragma unroll
                for (int q = 0; q < 16; q++) Lp[q * kDgLs + lane] = v[q];
            }
        }
        __syncthreads();
        // 3. triangular solves with the panel's diagonal block; the next
        //    column of L11 / row of U11 is loaded while the current one is used
        if (wid == 1) {
            double w[16], lc[2][16];
#pragma unroll
            for (int q = 0; q < 16; q++) w[q] = Rw[q * kDgRs + lane];
#pragma unroll
            for (int q = 1; q < 16; q++) lc[0][q] = Lp[k0 + q];
#pragma unroll
            for (int qp = 0; qp < 15; qp++) {
                if (qp + 1 < 15) {
#pragma unroll
                    for (int q = qp + 2; q < 16; q++) lc[(qp + 1) & 1][q] = Lp[(qp + 1) * kDgLs + k0 + q];
                }
#pragma unroll
                for (int q = qp + 1; q < 16; q++) w[q] = fma(-lc[qp & 1][q], w[qp], w[q]);
            }
#pragma unroll
            for (int q = 0; q < 16; q++) Rw[q * kDgRs + lane] = w[q];
        } else if (wid == 2) {
            if (lane < k0 + 16) {                      // rows below hold no V
                double w[16], uc[2][16], rq[2];
#pragma unroll
                for (int q = 0; q < 16; q++) w[q] = Wv[q * kDgLs + lane];
                rq[0] = Rp[0];
#pragma unroll
                for (int j = 1; j < 16; j++) uc[0][j] = Lp[j * kDgLs + k0];
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    if (q + 1 < 16) {
                        rq[(q + 1) & 1] = Rp[q + 1];
#pragma unroll
                        for (int j = q + 2; j < 16; j++) uc[(q + 1) & 1][j] = Lp[j * kDgLs + k0 + q + 1];
                    }
                    w[q] = w[q] * rq[q & 1];
#pragma unroll
                    for (int j = q + 1; j < 16; j++) w[j] = fma(-w[q], uc[q & 1][j], w[j]);
                }
#pragma unroll
                for (int q = 0; q < 16; q++) Wv[q * kDgLs + lane] = w[q];
            }
        }
        __syncthreads();
        // 4. final panel values, then the rank-16 updates (operands of step
        //    q + 1 loaded while step q is applied)
#pragma unroll
        for (int a = kb; a < 4; a++) A[a][kb] = Lp[s * kDgLs + r + 16 * a];
#pragma unroll
        for (int bb = 0; bb < 4; bb++) {
            const double v = Rw[r * kDgRs + s + 16 * bb];
            if (bb <= kb) X[kb][bb] = v;
            else A[kb][bb] = v;
        }
#pragma unroll
        for (int a = 0; a <= kb; a++) Y[a][kb] = Wv[s * kDgLs + r + 16 * a];
        if (kb < 3) {
            double lq[2][4], rq[2][4], vq[2][4];
#pragma unroll
            for (int a = kb + 1; a < 4; a++) lq[0][a] = Lp[r + 16 * a];
#pragma unroll
            for (int bb = 0; bb < 4; bb++) rq[0][bb] = Rw[s + 16 * bb];
#pragma unroll
            for (int a = 0; a <= kb; a++) vq[0][a] = Wv[r + 16 * a];
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int cb = q & 1, nb = cb ^ 1;
                if (q + 1 < 16) {
#pragma unroll
                    for (int a = kb + 1; a < 4; a++) lq[nb][a] = Lp[(q + 1) * kDgLs + r + 16 * a];
#pragma unroll
                    for (int bb = 0; bb < 4; bb++) rq[nb][bb] = Rw[(q + 1) * kDgRs + s + 16 * bb];
#pragma unroll
                    for (int a = 0; a <= kb; a++) vq[nb][a] = Wv[(q + 1) * kDgLs + r + 16 * a];
                }
#pragma unroll
                for (int a = kb + 1; a < 4; a++)
#pragma unroll
                    for (int bb = 0; bb < 4; bb++) {
                        if (bb > kb) A[a][bb] = fma(-lq[cb][a], rq[cb][bb], A[a][bb]);
                        else X[a][bb] = fma(-lq[cb][a], rq[cb][bb], X[a][bb]);
                    }
#pragma unroll
                for (int a = 0; a <= kb; a++)
#pragma unroll
                    for (int bb = kb + 1; bb < 4; bb++) Y[a][bb] = fma(vq[cb][a], rq[cb][bb], Y[a][bb]);
            }
        }
        __syncthreads();                               // LDS reads done before the next sub-panel's writes
    }
    double *W = J.W;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const int i = r + 16 * a;
#pragma unroll
        for (int bb = 0; bb < 4; bb++) {
            const int j = s + 16 * bb;
            if (i < b && j < b) Sp[(size_t)j * ld + i] = A[a][bb];
            // W is read by other workgroups of the same update launch
            st_coherent(&W[j * kNb + i], i > j ? X[a][bb] : (i == j ? 1.0 : 0.0));
            st_coherent(&W[kNb * kNb + j * kNb + i], i <= j ? Y[a][bb] : 0.0);
        }
    }
    // column maxima of |L| inside the block: the 16 threads of one s are
    // consecutive lanes of one wave
#pragma unroll
    for (int bb = 0; bb < 4; bb++) {
        const int j = s + 16 * bb;
        double am = 0.0;
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const int i = r + 16 * a;
            if (i > j && i < b) am = fmax(am, fabs(A[a][bb]));
        }
        am = max16(am);
        if (r == 0 && j < b) atomicMax(&J.colmax[p + j], dbits(am));
    }
}

__global__ void __launch_bounds__(256) dense_diag_kernel(const DenseJob *jobs, const int *active, int p)
{
    extern __shared__ double lds[];
    const DenseJob J = jobs[blockIdx.x];
    if (p >= J.T || (active && !active[J.mat])) return;
    if (threadIdx.x == 0) st_relaxed(J.flag, 0);      // no later panel final yet (read by the next launches)
    diag_block(J, p, J.S + (size_t)p * J.T + p, (size_t)J.T, lds);
}

// ---------------------------------------------------------------------------
// 64x64x64 tile product on 256 threads: thread (ty, tx) = (tid % 16, tid / 16)
// accumulates rows ty + 16 i and columns tx + 16 j (i, j < 4) of As * Bs,
// As[k*64 + i] = left(i, k), Bs[k*65 + j] = right(k, j).  Consecutive lanes
// read consecutive LDS words.
constexpr size_t kGemmLds = sizeof(double) * (kNb * 64 + kNb * kBsLd);
static_assert(kDiagLds <= kGemmLds, "tile 0 of the update factors the next diagonal block in its own LDS");

__device__ __forceinline__ void tile_mm(const double *As, const double *Bs, int kmax, double (&acc)[4][4])
{
    const int tid = threadIdx.x;
    const int ty = tid & 15, tx = tid >> 4;
#pragma unroll 4
    for (int k = 0; k < kmax; k++) {
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
}

// L21 = A21 U11^-1 (64-row tiles) and U12 = L11^-1 A12 (64-column tiles) as
// tile products with the inverses from the diagonal kernel.
__global__ void __launch_bounds__(256) dense_trsm_kernel(const DenseJob *jobs, const int *active, int p)
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
    double *As = lds;
    double *Bs = lds + kNb * 64;
    const int tid = threadIdx.x;
    const size_t ld = (size_t)J.T;
    const double *W = J.W;
    const bool lower = t < ntile;
    const int r0 = p + b + t * kNb;                    // L21 tile rows
    const int c0 = p + b + (t - ntile) * kNb;          // U12 tile columns
    {
        double va[16], vb[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int e = tid + 256 * u;
            const int k = e >> 6, i = e & 63;
            if (lower) {
                const int rr = r0 + i;
                va[u] = (rr < J.T && k < b) ? J.S[(size_t)(p + k) * ld + rr] : 0.0;
                vb[u] = W[kNb * kNb + i * kNb + k];                     // U^-1(k, i)
            } else {
                va[u] = W[k * kNb + i];                                 // L^-1(i, k)
                const int cc = c0 + k;                                  // column k of the tile, row i
                vb[u] = (cc < J.T && i < b) ? J.S[(size_t)cc * ld + p + i] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int e = tid + 256 * u;
            const int k = e >> 6, i = e & 63;
            As[k * 64 + i] = va[u];
            if (lower) Bs[k * kBsLd + i] = vb[u];
            else Bs[i * kBsLd + k] = vb[u];
        }
    }
    __syncthreads();
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
    tile_mm(As, Bs, kNb, acc);
    const int ty = tid & 15, tx = tid >> 4;
    if (lower) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int c = tx + 16 * j;
            double am = 0.0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int rr = r0 + ty + 16 * i;
                if (rr < J.T && c < b) {
                    J.S[(size_t)(p + c) * ld + rr] = acc[i][j];
                    am = fmax(am, fabs(acc[i][j]));
                }
            }
            am = max16(am);
            if (ty == 0 && c < b) atomicMax(&J.colmax[p + c], dbits(am));
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int cc = c0 + tx + 16 * j;
            if (cc >= J.T) continue;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int rr = ty + 16 * i;
                if (rr < b) J.S[(size_t)cc * ld + p + rr] = acc[i][j];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// S22 -= L21 U12 on 64x64 tiles with fp64 matrix cores.  Wave w computes the
// 32x32 quadrant (rows 32 (w & 1), columns 32 (w >> 1)) as 2x2 products of
// v_mfma_f64_16x16x4f64, whose operands and result sit in the lanes as
//   A(i, k): lane i + 16k      B(k, j): lane j + 16k
//   D(i, j): lane j + 16 (i % 4), register i / 4.
// K (the panel width) is zero-padded to a multiple of 4 in LDS.  The result
// goes through LDS so that the read-modify-write of S runs down columns.
typedef double v4d __attribute__((ext_vector_type(4)));

// acc = As * Bs over k < kmax (a multiple of 4), As[k*64 + i], Bs[k*65 + j]
__device__ __forceinline__ void mfma_tile(const double *As, const double *Bs, int kmax, v4d (&acc)[2][2])
{
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int rb = (w & 1) * 32, cb = (w >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) acc[x][y] = (v4d){0.0, 0.0, 0.0, 0.0};
    for (int k4 = 0; k4 < kmax; k4 += 4) {
        const int k = k4 + lk;
        double av[2], bv[2];
#pragma unroll
        for (int x = 0; x < 2; x++) av[x] = As[k * 64 + rb + 16 * x + li];
#pragma unroll
        for (int y = 0; y < 2; y++) bv[y] = Bs[k * kBsLd + cb + 16 * y + li];
#pragma unroll
        for (int x = 0; x < 2; x++)
#pragma unroll
            for (int y = 0; y < 2; y++) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], acc[x][y], 0, 0, 0);
    }
}

// The product to Cs[j*65 + i] (the caller has finished reading As / Bs if
// they overlap Cs).
__device__ __forceinline__ void acc_to_lds(double *Cs, const v4d (&acc)[2][2])
{
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    const int rb = (w & 1) * 32, cb = (w >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = rb + 16 * x + 4 * r + lk;
                const int j = cb + 16 * y + li;
                Cs[j * kBsLd + i] = acc[x][y][r];
            }
}

// Update launch for panel p (columns p .. q0 - 1, q0 = p + 64): tile (ti, tj)
// of the trailing matrix from (q0, q0).  Tiles go in the order
//   (0, 0), the next panel's diagonal block: updated, then factored here
//        (look-ahead, while the other tiles run), W written agent-coherent,
//        and J.flag set to q0 / 64;
//   the next panel's edge tiles, column (ti, 0) and row (0, tj): updated,
//        then, once J.flag says the diagonal block is factored, turned into
//        L21 = A21 U11^-1 and U12 = L11^-1 A12 with the inverses in W;
//   the interior tiles (ti, tj >= 1): updated.
// An edge tile only waits on tile 0 of its job, which comes first in the
// grid (dispatch order), so it is resident or done.  The wait is bounded
// and raises the abort word if it gives up.
__global__ void __launch_bounds__(256, 2) dense_gemm_kernel(const DenseJob *jobs, const int *active, int p,
                                                            int *abort_word)
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
    const int ni = nt - 1;
    int ti = 0, tj = 0;
    if (tile > 2 * ni) {
        ti = 1 + (tile - 1 - 2 * ni) % ni;
        tj = 1 + (tile - 1 - 2 * ni) / ni;
    } else if (tile > ni) {
        tj = tile - ni;
    } else {
        ti = tile;
    }
    const int edge = tile == 0 ? 0 : tj == 0 ? 1 : ti == 0 ? 2 : 0;   // 1: L21 tile, 2: U12 tile
    const int r0 = q0 + ti * 64;
    const int c0 = q0 + tj * 64;
    extern __shared__ double lds[];
    double *As = lds;                                  // As[k*64 + i] = L21(r0+i, p+k)
    double *Bs = lds + kNb * 64;                       // Bs[k*65 + j] = U12(p+k, c0+j)
    const int tid = threadIdx.x;
    const size_t ld = (size_t)J.T;
    const int b4 = (b + 3) & ~3;
    // 16 elements per thread of each operand; all loads are issued before
    // the first LDS store so they are in flight together
    {
        double va[16], vb[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int e = tid + 256 * u;
            const int k = e >> 6, i = e & 63;            // A: column k, row i
            const int rr = r0 + i;
            va[u] = (rr < J.T && k < b) ? J.S[(size_t)(p + k) * ld + rr] : 0.0;
            const int c = c0 + k;                         // B: column k of the tile, row i of the panel
            vb[u] = (c < J.T && i < b) ? J.S[(size_t)c * ld + p + i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int e = tid + 256 * u;
            const int k = e >> 6, i = e & 63;
            As[k * 64 + i] = va[u];
            Bs[i * kBsLd + k] = vb[u];
        }
    }
    __syncthreads();
    v4d acc[2][2];
    mfma_tile(As, Bs, b4, acc);
    __syncthreads();                                   // As/Bs reads done: reuse LDS for the product
    double *Cs = lds;                                  // Cs[j*65 + i]
    acc_to_lds(Cs, acc);
    __syncthreads();
    double cv[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int e = tid + 256 * u;
        const int j = e >> 6, i = e & 63;
        const int rr = r0 + i, c = c0 + j;
        cv[u] = (rr < J.T && c < J.T) ? J.S[(size_t)c * ld + rr] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int e = tid + 256 * u;
        const int j = e >> 6, i = e & 63;
        const int rr = r0 + i, c = c0 + j;
        const double v = cv[u] - Cs[j * kBsLd + i];
        cv[u] = v;
        if (!edge && rr < J.T && c < J.T) J.S[(size_t)c * ld + rr] = v;
        if (tile == 0) Cs[j * kBsLd + i] = v;           // element read by this thread alone above
    }
    if (tile == 0) {
        // the next panel's diagonal block is final (every earlier panel's
        // update has reached it): factor it while the other tiles run
        __syncthreads();
        diag_block(J, q0, Cs, kBsLd, lds);
        store_complete();                              // this wave's W stores acknowledged
        __syncthreads();
        if (tid == 0) st_relaxed(J.flag, q0 / kNb);
        return;
    }
    if (!edge) return;
    // Edge tile of the next panel: the updated tile in LDS as the operand it
    // will be (L21: left, As layout; U12: right, Bs layout), then the wait
    // for the diagonal block's inverses.
    __syncthreads();                                   // Cs reads done
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int e = tid + 256 * u;
        const int j = e >> 6, i = e & 63;
        if (edge == 1) As[j * 64 + i] = cv[u];         // left(i, k = j)
        else Bs[i * kBsLd + j] = cv[u];                // right(k = i, j)
    }
    int ok = 1;
    if (tid < kWave)
        ok = wave_wait(tid == 0 ? J.flag : nullptr, q0 / kNb, abort_word, nullptr, -1, kSolveBackoff);
    if (!__syncthreads_and(ok)) return;
    const double *W = J.W;
    {
        double wv[16];
#pragma unroll
        for (int u = 0; u < 16; u++) wv[u] = ld_coherent_val(&W[(edge == 1 ? kNb * kNb : 0) + tid + 256 * u]);
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int e = tid + 256 * u;
            const int j = e >> 6, k = e & 63;
            if (edge == 1) Bs[k * kBsLd + j] = wv[u];  // U11^-1(k, j)
            else As[e] = wv[u];                        // L11^-1(i = k, j): As[j*64 + i]
        }
    }
    __syncthreads();
    mfma_tile(As, Bs, kNb, acc);
    __syncthreads();
    acc_to_lds(Cs, acc);
    __syncthreads();
    // the panel's columns (L21) or rows (U12) of this tile, and |L| maxima
    double am[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
        const int e = tid + 256 * u;
        const int j = e >> 6, i = e & 63;
        const int rr = r0 + i, c = c0 + j;
        const double v = Cs[j * kBsLd + i];
        const bool in = rr < J.T && c < J.T;
        if (in) J.S[(size_t)c * ld + rr] = v;
        am[u] = in ? fabs(v) : 0.0;
    }
    if (edge == 1) {
        // column j = e >> 6 = tid / 64 + 4u: a wave holds 64 rows of one column
#pragma unroll
        for (int u = 0; u < 16; u++) {
            double m = am[u];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, kWave));
            const int j = (tid >> 6) + 4 * u;
            if ((tid & 63) == 0) atomicMax(&J.colmax[q0 + j], dbits(m));
        }
    }
}

// ---------------------------------------------------------------------------
// One wave per block column c: pivot test, scatter of the pattern slots into
// the sparse factors, dense copy in storage precision.
template <typename T>
__global__ void __launch_bounds__(64) dense_finish_kernel(const DenseJob *jobs, const void *mats,
                                                          const int *active, double ptol,
                                                          double stol)
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
        const double vmax = sizeof(T) == 2 ? 65504.0 : sizeof(T) == 4 ? 3.4028234663852886e38 : 1.0e308;
        if (!(fabs(piv) <= vmax) || !(lmax <= vmax)) atomicAdd(&M.mstat[3], 1);
        if (stol > 0.0) {
            // splu's choice on these values, in units of |pivot|: the other
            // candidates are the multipliers L(r, c), r > c, of the block
            const int dpos = M.dpos[k];
            const double md = dpos > k ? fabs(Sc[dpos - M.tail_k0]) : -1.0;
            const int v = piv == 0.0 ? (lmax > 0.0 ? 1 : 0)
                                     : splu_pivot_verdict(dpos == k, 1.0, lmax, md, stol);
            if (v) {
                atomicAdd(&M.mstat[v == 1 ? 4 : 5], 1);
                atomicMin(&M.mstat[6], k);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Forward solve, rows of the block: yp[k0 + r] = b[rowat] - L(k0 + r, :k0) y.
template <typename T>
__global__ void __launch_bounds__(256) tail_fwd_gather_kernel(const void *mats, const int *tails,
                                                              const long *voff, const int *active,
                                                              const double *b, const double *y, double *yp,
                                                              const double *rs)
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
    if (lane == 0) {
        const int ro = M.rowat[i];
        yp[vo + i] = (rs ? b[vo + ro] * rs[vo + ro] : b[vo + ro]) - acc;
    }
}

// Dense triangular solve of the block rows (see the file comment), two
// 64-row blocks (128 rows, in solve order) per workgroup:
//   lower: y_T <- L22^-1 yp_T (unit diagonal)
//   upper: z_T <- U22^-1 y_T, x[q] = z (the sparse backward solve, a later
//          launch, reads these z values)
// Waves 2..7 accumulate the contributions of the blocks solved before this
// pair (waves 2, 4, 6 for the first block's rows, 3, 5, 7 for the second's,
// three earlier blocks each in turn): per block the 64 x 64 factor slab (one
// row per lane, in registers) is loaded before the block's values are
// awaited, so the wait overlaps the loads; the polls return the solved
// values, broadcast by readlane.  Waves 0 and 1 hold the two diagonal blocks,
// right-hand sides and diagonal reciprocals.  After the first block is
// solved, the six accumulator waves apply it to the second block's rows (a
// slice of its columns each, from LDS), and wave 1 solves the second block:
// one hand-off through memory per 128 rows instead of per 64.
constexpr int kTailAcc = 6;                 // accumulator waves per workgroup
constexpr int kTailSlice = (kNb + kTailAcc - 1) / kTailAcc;   // coupling columns per accumulator wave

template <typename T, bool kLower>
__global__ void __launch_bounds__(512) tail_trsv_kernel(const void *mats, const int *tails,
                                                        const long *voff, const int *active,
                                                        double *y, const double *yp, double *z, double *x,
                                                        int *abort_word, const double *cs)
{
    const int mi = tails[blockIdx.y];
    const TriDesc<T> M = ((const TriDesc<T> *)mats)[mi];
    const int lane = lane_id();
    const int wv = threadIdx.x >> 6;
    __shared__ int s_ticket;
    __shared__ double red[kTailAcc][kWave];            // accumulators' sums over the earlier blocks
    __shared__ double red2[kTailAcc][kWave];           // the first block's contribution to the second's rows
    __shared__ double ys[kWave];                       // the first block's solution
    const bool live = !(active && !active[mi]);
    const int nblk = (M.tail_T + kNb - 1) / kNb;
    const int npair = (nblk + 1) / 2;
    if (live) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(&M.tail_ctr[0], 1);
        __syncthreads();
        const int t = uniform(s_ticket);
        if (t < npair) {
            const size_t ld = (size_t)M.tail_T;
            const long vo = voff[mi];
            const int k0 = M.tail_k0;
            const double *src = kLower ? y : z;
            const int nsub = min(2, nblk - 2 * t);      // blocks in this pair
            // block in solve order so -> its first row
            auto row0 = [&](int so) { return (kLower ? so : nblk - 1 - so) * kNb; };
            // the rows of the pair's block u, clamped per lane
            const int u = wv < 2 ? wv : (wv - 2) & 1;  // the block whose rows this wave handles
            const bool has = u < nsub;
            const int r0 = has ? row0(2 * t + u) : 0;
            const int bi = has ? min(kNb, M.tail_T - r0) : 1;
            const int rr = r0 + min(lane, bi - 1);
            double acc = 0.0;
            bool ok = true;
            T dg[kNb];                                  // waves 0, 1: row lane of the diagonal block
            T fc[kTailSlice];                           // waves 2..7: row of block 1 against a slice of block 0
            double rd = 1.0;                            // waves 0, 1, upper: 1 / U(rr, rr)
            double yb = 0.0;                            // waves 0, 1: right-hand side of row rr
            int qo = 0;                                 // waves 0, 1, upper: output index of row rr
            if (wv < 2) {
                if (has) {
#pragma unroll
                    for (int c = 0; c < kNb; c++) dg[c] = M.tail_F[(size_t)(r0 + min(c, bi - 1)) * ld + rr];
                    if (!kLower) {
                        // zero past the block's end: those lanes then pass zeros
                        rd = lane < bi ? 1.0 / to_f64(M.tail_F[(size_t)rr * ld + rr]) : 0.0;
                        qo = M.q[k0 + rr];
                    }
                    yb = kLower ? yp[vo + k0 + rr] : y[vo + k0 + rr];
                }
            } else {
                if (nsub == 2) {
                    // slice of the coupling block: rows of block 1, columns of block 0
                    // (block 0 is partial only for the upper solve's first pair;
                    // its missing columns meet zeros in ys)
                    const int ra = row0(2 * t), rb = row0(2 * t + 1);
                    const int ba = min(kNb, M.tail_T - ra);
                    const int bb = min(kNb, M.tail_T - rb);
                    const int rr1 = rb + min(lane, bb - 1);
                    const int c0 = (wv - 2) * kTailSlice;
#pragma unroll
                    for (int c = 0; c < kTailSlice; c++)
                        fc[c] = M.tail_F[(size_t)(ra + min(c0 + c, ba - 1)) * ld + rr1];
                }
                if (has) {
                    for (int s = (wv - 2) >> 1; s < 2 * t && ok; s += 3) {
                        const int c0 = row0(s);
                        const int bj = min(kNb, M.tail_T - c0);
                        T fv[kNb];
#pragma unroll
                        for (int c = 0; c < kNb; c++) fv[c] = M.tail_F[(size_t)(c0 + min(c, bj - 1)) * ld + rr];
                        double sv;                                   // zero beyond bj
                        ok = wave_wait_value(lane < bj ? &src[vo + k0 + c0 + lane] : nullptr, sv, abort_word);
                        if (!ok) break;
#pragma unroll
                        for (int c = 0; c < kNb; c++) {
                            const double yc = bcast(sv, c);          // zero beyond bj
                            acc += to_f64(fv[c]) * yc;
                        }
                    }
                }
                red[wv - 2][lane] = has && lane < bi ? acc : 0.0;
            }
            const bool all_ok = __syncthreads_and(ok);
            // triangle of one block on its solver wave: tr = right-hand side
            // minus the contributions gathered, solved in place
            auto solve = [&](double tr) {
                if (kLower) {
#pragma unroll
                    for (int c = 0; c < kNb; c++) {
                        const double yc = bcast(tr, c);
                        if (lane > c && c < bi) tr -= to_f64(dg[c]) * yc;
                    }
                } else {
                    // z_c = tr_c / U(c, c), one broadcast per step; rows past
                    // bi hold tr = 0 and rd = 0 and broadcast zeros.  Row c's
                    // tr is final from step c on, so every row's z is the
                    // same product tr * rd, formed once at the end.
#pragma unroll
                    for (int c = kNb - 1; c >= 0; c--) {
                        const double zc = bcast(tr * rd, c);
                        if (lane < c) tr -= to_f64(dg[c]) * zc;
                    }
                    tr = tr * rd;
                }
                if (lane < bi) {
                    if (kLower) {
                        st_coherent(&y[vo + k0 + rr], tr);
                    } else {
                        st_coherent(&z[vo + k0 + rr], tr);
                        x[vo + qo] = cs ? tr * cs[vo + qo] : tr;
                    }
                }
                return tr;
            };
            if (all_ok && wv == 0) {
                double tr = 0.0;
                if (lane < bi) tr = yb - (red[0][lane] + red[2][lane] + red[4][lane]);
                tr = solve(tr);
                ys[lane] = lane < bi ? tr : 0.0;
            }
            if (nsub == 2) {
                __syncthreads();                        // ys written
                if (all_ok && wv >= 2) {
                    const int c0 = (wv - 2) * kTailSlice;
                    double a2 = 0.0;
#pragma unroll
                    for (int c = 0; c < kTailSlice; c++)
                        if (c0 + c < kNb) a2 += to_f64(fc[c]) * ys[c0 + c];
                    red2[wv - 2][lane] = a2;
                }
                __syncthreads();                        // red2 written
                if (all_ok && wv == 1) {
                    double tr = 0.0;
                    if (lane < bi) {
                        tr = yb - (red[1][lane] + red[3][lane] + red[5][lane]);
#pragma unroll
                        for (int w = 0; w < kTailAcc; w++) tr -= red2[w][lane];
                    }
                    solve(tr);
                }
            }
        }
        // the last workgroup of this matrix re-arms its ticket counter
        if (threadIdx.x == 0) {
            const int fin = atomicAdd(&M.tail_ctr[1], 1);
            if (fin == (int)gridDim.x - 1) {
                st_relaxed(&M.tail_ctr[0], 0);
                st_relaxed(&M.tail_ctr[1], 0);
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
                               (int)kGemmLds);
}

// Panel 0's diagonal block has a launch of its own; every later diagonal
// block is factored by tile 0 of the preceding trailing update.  For a job
// whose T is below maxT the update kernel returns before its tiles once
// p + 64 >= T, so each of its panels is factored exactly once.
hipError_t launch_dense_lu(const DenseJob *jobs, int njobs, int maxT, const int *active, int *abort_word,
                           hipStream_t st)
{
    if (maxT <= 0) return hipSuccess;
    hipLaunchKernelGGL(dense_diag_kernel, dim3(njobs), dim3(256), kDiagLds, st, jobs, active, 0);
    for (int p = 0; p < maxT; p += kNb) {
        const int rem = maxT - p - min(kNb, maxT - p);
        if (rem > 0) {
            const int nt = (rem + kNb - 1) / kNb;
            if (p == 0)
                hipLaunchKernelGGL(dense_trsm_kernel, dim3(2 * nt, njobs), dim3(256), kGemmLds, st, jobs, active, p);
            hipLaunchKernelGGL(dense_gemm_kernel, dim3(nt * nt, njobs), dim3(256), kGemmLds, st, jobs,
                               active, p, abort_word);
        }
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_dense_finish(const DenseJob *jobs, int njobs, int maxT, const void *mats,
                               const int *active, double ptol, double stol, hipStream_t st)
{
    hipLaunchKernelGGL(dense_finish_kernel<T>, dim3(maxT, njobs), dim3(64), 0, st, jobs, mats, active, ptol,
                       stol);
    return hipGetLastError();
}
template hipError_t launch_dense_finish<float>(const DenseJob *, int, int, const void *, const int *, double, double, hipStream_t);
template hipError_t launch_dense_finish<double>(const DenseJob *, int, int, const void *, const int *, double, double, hipStream_t);
template hipError_t launch_dense_finish<__half>(const DenseJob *, int, int, const void *, const int *, double, double, hipStream_t);

template <typename T>
hipError_t launch_tail_solve(const void *mats, const int *tails, int ntails, int maxT, const long *voff,
                             const int *active, const double *b, double *y, double *yp, double *z,
                             double *x, bool forward, int *abort_word, const double *rs, const double *cs,
                             hipStream_t st)
{
    const int nblk = (maxT + kNb - 1) / kNb;
    const int npair = (nblk + 1) / 2;                   // two 64-row blocks per workgroup
    if (forward) {
        hipLaunchKernelGGL(tail_fwd_gather_kernel<T>, dim3((maxT + 3) / 4, ntails), dim3(256), 0, st, mats,
                           tails, voff, active, b, y, yp, rs);
        hipLaunchKernelGGL((tail_trsv_kernel<T, true>), dim3(npair, ntails), dim3(512), 0, st, mats, tails,
                           voff, active, y, yp, z, x, abort_word, cs);
    } else {
        hipLaunchKernelGGL((tail_trsv_kernel<T, false>), dim3(npair, ntails), dim3(512), 0, st, mats, tails,
                           voff, active, y, yp, z, x, abort_word, cs);
    }
    return hipGetLastError();
}
template hipError_t launch_tail_solve<float>(const void *, const int *, int, int, const long *, const int *,
                                              const double *, double *, double *, double *, double *, bool,
                                              int *, const double *, const double *, hipStream_t);
template hipError_t launch_tail_solve<double>(const void *, const int *, int, int, const long *, const int *,
                                              const double *, double *, double *, double *, double *, bool,
                                              int *, const double *, const double *, hipStream_t);
template hipError_t launch_tail_solve<__half>(const void *, const int *, int, int, const long *, const int *,
                                              const double *, double *, double *, double *, double *, bool,
                                              int *, const double *, const double *, hipStream_t);

}  // namespace lsp
