// batch.cpp -- host runtime of the device-resident batch.
//
// create(): per circuit, on host threads: COLAMD (unless q is given), the CPU
//   symbolic LU with the reference's pivoting rule, the static plan (column
//   patterns, A scatter map, update map, levels, row-wise copies) and the
//   SELL-C-64 layout.  Then every structure array of every circuit is
//   concatenated into one device arena and the per-circuit descriptors are
//   written.  Task lists of all circuits are merged by level so that one
//   persistent launch advances every circuit's dependency front at once.
// refactor(): value repack for SpMV + one persistent refactor launch.
// solve(): fp64 GMRES-IR, every vector on the device; the host reads one
//   word per inner step (the number of circuits still iterating).
#include "batch.hpp"

#include <type_traits>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

#include "../host/csc.h"

namespace lsp {

static_assert(kTriPacket == ::kTriPacketBase && kTriPacketLane == ::kTriPacketPerLane,
              "packet layout of the plan and the solve kernel");


Batch::~Batch()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (auto &m : mem_) if (m.p) (void)hipFree(m.p);
    if (h_pin_) (void)hipHostFree(h_pin_);
    for (auto &e : ev_) if (e) (void)hipEventDestroy(e);
    if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

int Batch::check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return LSP_OK;
    last_error = std::string(what) + ": " + hipGetErrorString(e);
    return LSP_E_HIP;
}

void *Batch::dalloc(size_t bytes)
{
    if (bytes == 0) bytes = 16;
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    mem_.push_back({p, bytes});
    return p;
}

void Batch::dfree(void *p)
{
    if (!p) return;
    for (auto it = mem_.begin(); it != mem_.end(); ++it) {
        if (it->p == p) {
            (void)hipFree(p);
            mem_.erase(it);
            return;
        }
    }
}

// Krylov storage for restart length mm (basis, Hessenberg, rotations).
int Batch::alloc_krylov(int mm)
{
    const int nc = (int)circ_.size();
    for (void *p : {(void *)d_V_, (void *)d_Z_, (void *)d_partial_, (void *)d_H_, (void *)d_cs_, (void *)d_sn_,
                    (void *)d_g_, (void *)d_yv_, (void *)d_hcol_})
        dfree(p);
    d_V_ = (double *)dalloc((size_t)(mm + 1) * ntot_ * 8);
    d_Z_ = (double *)dalloc((size_t)mm * ntot_ * 8);
    d_partial_ = (double *)dalloc((size_t)nchunks_ * (mm + 1) * 8);
    d_H_ = (double *)dalloc((size_t)nc * (mm + 1) * mm * 8);
    d_cs_ = (double *)dalloc((size_t)nc * mm * 8);
    d_sn_ = (double *)dalloc((size_t)nc * mm * 8);
    d_g_ = (double *)dalloc((size_t)nc * (mm + 1) * 8);
    d_yv_ = (double *)dalloc((size_t)nc * mm * 8);
    d_hcol_ = (double *)dalloc((size_t)nc * (mm + 1) * 8);
    if (!d_V_ || !d_Z_ || !d_partial_ || !d_H_ || !d_cs_ || !d_sn_ || !d_g_ || !d_yv_ || !d_hcol_)
        return LSP_E_ALLOC;
    mm_alloc_ = mm;
    return LSP_OK;
}

int Batch::set_ir_params(int outer, int inner, double inner_tol, double outer_tol)
{
    if (outer < 1 || inner < 1 || !(inner_tol > 0) || !(outer_tol > 0)) return LSP_E_ARG;
    // a recorded step graph holds the Krylov buffers and tolerances of its
    // recording: record again on the next graph_step
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    graph_prec_ = -1;
    if (inner > mm_alloc_) {
        int rc = alloc_krylov(inner);
        if (rc) return rc;
    }
    opt_.ir_outer = outer;
    opt_.ir_inner = inner;
    opt_.ir_inner_tol = inner_tol;
    opt_.ir_outer_tol = outer_tol;
    return LSP_OK;
}

int Batch::diag_stats(int c, double *min_abs, double *max_abs)
{
    if (c < 0 || c >= (int)circ_.size() || cur_prec_ < 0) return LSP_E_ARG;
    const int n = circ_[c].n;
    std::vector<double> d(n);
    int rc;
    if (cur_prec_ == LSP_PREC_FP64) {
        rc = check(hipMemcpy(d.data(), f64_.diag + circ_[c].o_n, n * 8, hipMemcpyDeviceToHost), "diag");
    } else if (cur_prec_ == LSP_PREC_FP32) {
        std::vector<float> f(n);
        rc = check(hipMemcpy(f.data(), f32_.diag + circ_[c].o_n, n * 4, hipMemcpyDeviceToHost), "diag");
        for (int i = 0; i < n; i++) d[i] = f[i];
    } else {
        std::vector<__half> f(n);
        rc = check(hipMemcpy(f.data(), f16_.diag + circ_[c].o_n, n * 2, hipMemcpyDeviceToHost), "diag");
        for (int i = 0; i < n; i++) d[i] = (double)__half2float(f[i]);
    }
    if (rc) return rc;
    double lo = INFINITY, hi = 0;
    for (double v : d) {
        lo = std::min(lo, std::fabs(v));
        hi = std::max(hi, std::fabs(v));
    }
    if (min_abs) *min_abs = lo;
    if (max_abs) *max_abs = hi;
    return LSP_OK;
}

template <typename X>
X *Batch::upload(const std::vector<X> &v)
{
    X *d = (X *)dalloc(v.size() * sizeof(X));
    if (!d) return nullptr;
    if (!v.empty() && hipMemcpy(d, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice) != hipSuccess)
        return nullptr;
    return d;
}

int Batch::sync_read_int(const int *d, int *h, int count)
{
    int rc = check(hipMemcpyAsync(h, d, count * sizeof(int), hipMemcpyDeviceToHost, stream_),
                   "read flags");
    if (rc) return rc;
    rc = check(hipStreamSynchronize(stream_), "sync");
    if (!rc) account_trisolve();
    return rc;
}

// Adds the device time of the last forward/backward trisolve pair to the
// running statistics; called after a stream synchronisation, when its events
// are complete.
void Batch::account_trisolve()
{
    if (!tri_pending_) return;
    float a = 0, b = 0;
    if (hipEventElapsedTime(&a, ev_[2], ev_[4]) == hipSuccess &&
        hipEventElapsedTime(&b, ev_[4], ev_[3]) == hipSuccess) {
        t_fwd_ = a;
        t_bwd_ = b;
        stats_.fwd_ms += a;
        stats_.bwd_ms += b;
        stats_.tri_count++;
    }
    tri_pending_ = false;
}

// ---------------------------------------------------------------------------
// Rows below the dense block (positions < tail_k0) grouped by forward and by
// backward solve level: a counting sort of the plan's row levels.  The level
// count is taken from the levels themselves (the plan's nlv_f / nlv_b count
// the supernodal solve's levels, which can be fewer).  Row lists are padded
// to n entries so that they share the n-sized arena offsets.
static int by_level(const std::vector<int> &lv, int k0, int n, std::vector<int> &row, std::vector<int> &ptr)
{
    if (k0 < 0 || k0 > n || (int)lv.size() < k0) return LSP_E_STRUCTURE;
    int nlv = 0;
    for (int i = 0; i < k0; i++) {
        if (lv[i] < 0) return LSP_E_STRUCTURE;
        nlv = std::max(nlv, lv[i] + 1);
    }
    ptr.assign(nlv + 1, 0);
    for (int i = 0; i < k0; i++) ptr[lv[i] + 1]++;
    for (int l = 0; l < nlv; l++) ptr[l + 1] += ptr[l];
    row.assign(n, 0);
    std::vector<int> next(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < k0; i++) row[next[lv[i]]++] = i;
    return LSP_OK;
}

static int ir_level_lists(CircuitHost &C)
{
    const LspPlan &P = C.plan;
    int rc = by_level(P.lvf, P.tail_k0, C.n, C.frow, C.flvp);
    return rc ? rc : by_level(P.lvb, P.tail_k0, C.n, C.brow, C.blvp);
}

static int build_circuit(CircuitHost &C, const lsp_circuit_in &in, const lsp_options &opt)
{
    C.n = in.n;
    C.nnz = in.nnz;
    C.row.assign(in.row, in.row + in.nnz);
    C.col.assign(in.col, in.col + in.nnz);
    C.val0.assign(in.val, in.val + in.nnz);
    for (long e = 0; e < in.nnz; e++)
        if (C.row[e] < 0 || C.row[e] >= C.n || C.col[e] < 0 || C.col[e] >= C.n) return LSP_E_ARG;
    C.q.resize(C.n);
    if (in.q) {
        std::vector<char> seen(C.n, 0);
        for (int k = 0; k < C.n; k++) {
            if (in.q[k] < 0 || in.q[k] >= C.n || seen[in.q[k]]) return LSP_E_ARG;
            seen[in.q[k]] = 1;
            C.q[k] = in.q[k];
        }
    } else if (lsp_colamd_coo(C.n, C.nnz, C.row.data(), C.col.data(), C.q.data()) != 0) {
        return LSP_E_STRUCTURE;
    }
    std::vector<int> Ap, Ai, src;
    lsp_coo_to_csc(C.n, C.nnz, C.row.data(), C.col.data(), C.q.data(), Ap, Ai, src);
    std::vector<double> Ax(C.nnz);
    for (long p = 0; p < C.nnz; p++) Ax[p] = C.val0[src[p]];
    if (opt.pivot_scaled) {
        // pivot on the equilibrated matrix R A C (the order the reference
        // uses when Ruiz is on: scale, then order and factor)
        std::vector<double> R(C.n), Cs(C.n);
        lsp_ruiz_scales(C.n, C.nnz, C.row.data(), C.col.data(), C.val0.data(),
                        opt.ruiz_sweeps > 0 ? opt.ruiz_sweeps : 30, R.data(), Cs.data());
        for (int k = 0; k < C.n; k++)
            for (int p = Ap[k]; p < Ap[k + 1]; p++) Ax[p] *= R[Ai[p]] * Cs[C.q[k]];
    }
    int rc = lsp_gp_factor(C.n, Ap.data(), Ai.data(), Ax.data(), C.q.data(), opt.pivot_tol, C.sym);
    if (rc != -1) return LSP_E_SINGULAR;
    if (lsp_build_plan(C.n, Ap.data(), Ai.data(), src.data(), C.sym, C.q.data(), C.plan) != 0)
        return LSP_E_STRUCTURE;
    if (C.plan.max_col > 65535) return LSP_E_CAPACITY;
    {
        const int tmax = opt.tail_max == 0 ? 8192 : opt.tail_max;
        const double tden = opt.tail_density > 0 ? opt.tail_density : 0.5;
        if (tmax > 0 && lsp_plan_set_tail(C.plan, 64, tmax, tden) < 0) return LSP_E_STRUCTURE;
    }
    std::vector<int> rp, ci, rsrc;
    lsp_coo_to_csr(C.n, C.nnz, C.row.data(), C.col.data(), rp, ci, rsrc);
    if (lsp_build_sell(C.n, rp.data(), ci.data(), rsrc.data(), 64, opt.sell_sigma, C.sell) != 0)
        return LSP_E_STRUCTURE;
    return ir_level_lists(C);
}

template <typename T>
int Batch::ensure_factor_set(FactorSet<T> &fs)
{
    if (fs.ready) return LSP_OK;
    fs.val = (T *)dalloc(nslot_tot_ * sizeof(T));
    fs.lrow = (T *)dalloc(nl_tot_ * sizeof(T));
    fs.urow = (T *)dalloc(nu_tot_ * sizeof(T));
    fs.diag = (T *)dalloc(ntot_ * sizeof(T));
    fs.tailF = (T *)dalloc(nt2_tot_ * sizeof(T));
    if (!fs.val || !fs.lrow || !fs.urow || !fs.diag || !fs.tailF) return LSP_E_ALLOC;
    const int nc = (int)circ_.size();
    std::vector<FactorDesc<T>> fd(nc);
    std::vector<TriDesc<T>> td(nc);
    std::vector<IrDesc<T>> id(nc);
    for (int m = 0; m < nc; m++) {
        const CircuitHost &C = circ_[m];
        FactorDesc<T> &F = fd[m];
        F.cp = d_cp_ + C.o_np1;
        F.dp = d_dp_ + C.o_n;
        F.pos = d_pos_ + C.o_slot;
        F.aptr = d_aptr_ + C.o_np1;
        F.asrc = d_asrc_ + C.o_a;
        F.aslot = d_aslot_ + C.o_a;
        // FP16 reads the Ruiz-scaled copy; a factor-only shift also factors that copy
        const bool own_copy = std::is_same<T, __half>::value || opt_.factor_shift > 0.0;
        F.ain = (own_copy ? d_ains_ : d_ain_) + C.aoff;
        F.relp = d_relp_ + C.o_np1;
        F.rel = d_rel_ + C.o_rel;
        F.snend = d_snend_ + C.o_n;
        F.dst = d_dst_ + C.o_slot;
        F.dpos = d_dpos_ + C.o_n;
        F.val = fs.val + C.o_slot;
        F.lrow = fs.lrow + C.o_l;
        F.urow = fs.urow + C.o_u;
        F.diag = fs.diag + C.o_n;
        F.done = d_done_ + C.o_n;
        F.mstat = d_mstat_ + kMstat * m;
        F.n = C.n;
        F.tail_ds = d_tail_ds_ + C.o_t;
        F.tail_S = d_tailS_ + C.o_t2;
        F.tail_colmax = d_colmax_ + C.o_t;
        F.tail_F = fs.tailF + C.o_t2;
        F.tail_k0 = C.plan.tail_k0;
        F.tail_T = C.plan.tail_T;
        TriDesc<T> &D = td[m];
        D.lrp = d_lrp_ + C.o_np1;
        D.lci = d_lci_ + C.o_l;
        D.urp = d_urp_ + C.o_np1;
        D.uci = d_uci_ + C.o_u;
        D.lrow = F.lrow;
        D.urow = F.urow;
        D.diag = F.diag;
        D.rowat = d_rowat_ + C.o_n;
        D.q = d_q_ + C.o_n;
        D.twide = d_twide_ + C.o_tw;
        D.n = C.n;
        D.tail_F = fs.tailF + C.o_t2;
        D.val = F.val;
        D.dp = F.dp;
        D.snend = d_snend_ + C.o_n;
        D.ru = d_ru_ + C.o_n;
        D.fcut = d_fcut_ + C.o_n;
        D.bcut = d_bcut_ + C.o_n;
        D.fpk = d_fpk_ + C.o_fpk;
        D.bpk = d_bpk_ + C.o_bpk;
        D.tail_lcut = d_tail_lcut_ + C.o_t;
        D.tail_ctr = d_tail_ctr_ + 4 * m;
        D.tail_k0 = C.plan.tail_k0;
        D.tail_T = C.plan.tail_T;
        IrDesc<T> &I = id[m];
        I.A.off = d_soff_ + C.o_slp1;
        I.A.rlen = d_srlen_ + C.o_n;
        I.A.col = d_scol_ + C.o_ss;
        I.A.val = d_sval64_ + C.o_ss;
        I.A.perm = d_sperm_ + C.o_n;
        I.A.nslices = C.sell.nslices;
        I.A.n = C.n;
        I.lrp = D.lrp; I.lci = D.lci; I.urp = D.urp; I.uci = D.uci;
        I.lrow = D.lrow; I.urow = D.urow; I.diag = D.diag; I.tail_F = D.tail_F;
        I.rowat = D.rowat; I.q = D.q; I.tail_lcut = D.tail_lcut;
        I.frow = d_frow_ + C.o_n;
        I.flvp = d_flvp_ + C.o_lf;
        I.brow = d_brow_ + C.o_n;
        I.blvp = d_blvp_ + C.o_lb;
        I.voff = C.voff;
        I.n = C.n;
        I.tail_k0 = C.plan.tail_k0;
        I.tail_T = C.plan.tail_T;
        I.nlv_f = (int)C.flvp.size() - 1;
        I.nlv_b = (int)C.blvp.size() - 1;
    }
    fs.fdesc = upload(fd);
    fs.tdesc = upload(td);
    fs.idesc = upload(id);
    if (!fs.fdesc || !fs.tdesc || !fs.idesc) return LSP_E_ALLOC;
    fs.ready = true;
    return LSP_OK;
}

int Batch::create(const lsp_circuit_in *cin, int nc, const lsp_options *opt)
{
    int rc = prepare(cin, nc, opt);
    return rc ? rc : upload_all();
}

// Host phase: symbolic analysis of every circuit (no device calls).
int Batch::prepare(const lsp_circuit_in *cin, int nc, const lsp_options *opt)
{
    if (!cin || nc <= 0 || !circ_.empty()) return LSP_E_ARG;
    if (opt) opt_ = *opt; else lsp_options_default(&opt_);
    if (opt_.repivot < 0 || opt_.repivot > 3) return LSP_E_ARG;
    if (opt_.ir_method < 0 || opt_.ir_method > 1 ||
        opt_.ir_fused_steps < 0 || opt_.refactor_fp64 < 0 || opt_.refactor_fp64 > 1)
        return LSP_E_ARG;
    refactor_tl_path_ = getenv("LSP_REFACTOR_TIMELINE");     // diagnostic dumps
    trisolve_tl_path_ = getenv("LSP_TRISOLVE_TIMELINE");
    circ_.resize(nc);
    std::vector<int> rcs(nc, 0);
    {
        std::atomic<int> next{0};
        int nthreads = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 16u);
        nthreads = std::min(nthreads, nc);
        std::vector<std::thread> pool;
        for (int t = 0; t < nthreads; t++)
            pool.emplace_back([&]() {
                for (int m = next++; m < nc; m = next++) rcs[m] = build_circuit(circ_[m], cin[m], opt_);
            });
        for (auto &th : pool) th.join();
    }
    for (int m = 0; m < nc; m++) {
        if (rcs[m] != LSP_OK) {
            last_error = "circuit " + std::to_string(m) + ": symbolic phase failed";
            circ_.clear();
            return rcs[m];
        }
    }
    return LSP_OK;
}

// Moves circuit picks[i].second of batch picks[i].first into circuit i of this
// (empty) batch.  Only host structures move (pattern, pivot sequence, plan,
// SELL layout, values); the device phase is upload_all().  Every source batch
// is emptied: its circuits not picked are dropped.  The options are those of
// the first source.
int Batch::adopt(const std::vector<std::pair<Batch *, int>> &picks)
{
    if (!circ_.empty() || uploaded_ || picks.empty()) return LSP_E_ARG;
    std::vector<std::pair<Batch *, int>> sorted(picks);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) return LSP_E_ARG;   // picked twice
    for (const auto &pk : picks) {
        if (!pk.first || pk.first == this || pk.second < 0 || pk.second >= (int)pk.first->circ_.size())
            return LSP_E_ARG;
        if (pk.first->circ_[pk.second].n <= 0) return LSP_E_ARG;
    }
    opt_ = picks[0].first->opt_;
    refactor_tl_path_ = picks[0].first->refactor_tl_path_;
    trisolve_tl_path_ = picks[0].first->trisolve_tl_path_;
    circ_.reserve(picks.size());
    for (const auto &pk : picks) circ_.push_back(std::move(pk.first->circ_[pk.second]));
    for (const auto &pk : picks) pk.first->circ_.clear();
    return LSP_OK;
}

int Batch::set_active(const int *mask)
{
    if (!mask) {
        active_.clear();
        return LSP_OK;
    }
    active_.assign(mask, mask + circ_.size());
    for (int &a : active_) a = a ? 1 : 0;
    return LSP_OK;
}

// Device bytes upload_all() and the first fp32 + fp64 refactors will hold
// (an upper estimate used for admission; device_bytes() reports the actual).
size_t Batch::estimate_device_bytes() const
{
    double n = 0, nnz = 0, slots = 0, rel = 0, nlu = 0, sell = 0, slices = 0, T = 0, T2 = 0;
    double maxcol = 0;
    const double nc = (double)circ_.size();
    for (const auto &C : circ_) {
        const LspPlan &P = C.plan;
        n += C.n; nnz += C.nnz; slots += (double)P.cp[C.n]; rel += (double)P.relp[C.n];
        nlu += (double)(P.n_l + P.n_u); sell += (double)C.sell.slots; slices += C.sell.nslices;
        T += P.tail_T; T2 += (double)P.tail_T * P.tail_T;
        maxcol = std::max(maxcol, (double)P.max_col);
    }
    const double m = std::max(1, opt_.ir_inner);
    double b = 0;
    b += 16 * (n + nc) + 4 * n + 8 * slots + 4 * (n + nc) + 8 * nnz;     // cp relp dp pos dst aptr asrc aslot
    b += 8 * (n + nc) + 4 * nlu + 24 * n + 2 * rel;                       // lrp urp lci uci rowat q snend ru fcut bcut rel
    b += 8 * (slices + nc) + 8 * n + 8 * sell;                            // SELL structure
    b += 12 * n + 8 * nnz + 12 * sell;                                    // flags, values
    if (opt_.krylov_spmv == 2) b += 2 * sell;                              // FP16 SELL values
    b += 16 * nnz + 24 * n + 16 * (nc + 1);                               // Ruiz: row/col, scaled values, scales
    if (opt_.factor_shift > 0.0) b += 8 * n;                               // shifted diagonal entries
    b += 8 * T2 + 16 * T + 65536.0 * nc;                                   // dense block work
    b += 24 * n + 8 * slices + 16 * n;                                    // task lists, row packets
    b += 64 * n + 8 * (2 * m + 1) * n;                                    // vectors, Krylov bases V and Z
    b += (4 + 8) * (slots + nlu + n + T2);                                 // fp32 + fp64 factor sets
    if (maxcol > std::max(1, opt_.lds_cap)) b += 256.0 * 8 * 4 * maxcol * 8;   // global column scratch
    return (size_t)(1.05 * b) + (1u << 20);
}

size_t Batch::device_bytes() const
{
    size_t s = 0;
    for (const auto &mm : mem_) s += mm.bytes;
    return s;
}

// Device phase: concatenates the circuits' structures and uploads them.
// Also the second half of a rebuild (release_device first): the stream,
// events, pinned words and the b / x arenas are kept.
int Batch::upload_all()
{
    if (circ_.empty() || uploaded_) return LSP_E_ARG;
    const int nc = (int)circ_.size();
    int dev = 0;
    hipDeviceProp_t prop;
    int rc = check(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    rc = check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    if (rc) return rc;
    cu_count_ = prop.multiProcessorCount;
    if (!stream_) {
        rc = check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "stream");
        if (rc) return rc;
        own_stream_ = true;
    }

    // ---- concatenate structures ----
    long o_np1 = 0, o_n = 0, o_slot = 0, o_a = 0, o_rel = 0, o_l = 0, o_u = 0, o_sl = 0,
         o_slp1 = 0, o_ss = 0, o_t = 0, o_t2 = 0, o_fpk = 0, o_bpk = 0, o_tw = 0;
    for (auto &C : circ_) {
        C.o_fpk = o_fpk; C.o_bpk = o_bpk; C.o_tw = o_tw;
        o_fpk += (long)C.plan.fpk.size(); o_bpk += (long)C.plan.bpk.size(); o_tw += (long)C.plan.twide.size();
        C.o_np1 = o_np1; C.o_n = o_n; C.o_slot = o_slot; C.o_a = o_a; C.o_rel = o_rel;
        C.o_l = o_l; C.o_u = o_u; C.o_sl = o_sl; C.o_slp1 = o_slp1; C.o_ss = o_ss;
        C.o_t = o_t; C.o_t2 = o_t2;
        const long T = C.plan.tail_T;
        o_t += T; o_t2 += T * T;
        C.voff = o_n;
        C.aoff = o_a;
        o_np1 += C.n + 1; o_n += C.n; o_slot += C.plan.cp[C.n]; o_a += C.nnz;
        o_rel += C.plan.relp[C.n]; o_l += C.plan.n_l; o_u += C.plan.n_u;
        o_sl += C.sell.nslices; o_slp1 += C.sell.nslices + 1; o_ss += C.sell.slots;
    }
    ntot_ = o_n; nnz_tot_ = o_a; nslot_tot_ = o_slot; nl_tot_ = o_l; nu_tot_ = o_u;
    nsell_tot_ = o_ss;
    nt_tot_ = o_t; nt2_tot_ = o_t2;
    {
        std::vector<int64_t> cp, relp, soff;
        std::vector<int> dp, pos, aptr, asrc, aslot, dst, lrp, lci, urp, uci, rowat, q, snend, ru, fcut, bcut, fpk, bpk,
            twide;
        std::vector<int> dpos;
        std::vector<int> srlen, scol, sperm, ssrc;
        std::vector<uint16_t> rel;
        cp.reserve(o_np1); relp.reserve(o_np1); rel.reserve(o_rel);
        for (auto &C : circ_) {
            const LspPlan &P = C.plan;
            cp.insert(cp.end(), P.cp.begin(), P.cp.end());
            relp.insert(relp.end(), P.relp.begin(), P.relp.end());
            dp.insert(dp.end(), P.dp.begin(), P.dp.end());
            pos.insert(pos.end(), P.pos.begin(), P.pos.end());
            aptr.insert(aptr.end(), P.aptr.begin(), P.aptr.end());
            asrc.insert(asrc.end(), P.asrc.begin(), P.asrc.end());
            aslot.insert(aslot.end(), P.aslot.begin(), P.aslot.end());
            dst.insert(dst.end(), P.dst.begin(), P.dst.end());
            lrp.insert(lrp.end(), P.lrp.begin(), P.lrp.end());
            lci.insert(lci.end(), P.lci.begin(), P.lci.end());
            urp.insert(urp.end(), P.urp.begin(), P.urp.end());
            uci.insert(uci.end(), P.uci.begin(), P.uci.end());
            rowat.insert(rowat.end(), P.rowat.begin(), P.rowat.end());
            q.insert(q.end(), P.q.begin(), P.q.end());
            snend.insert(snend.end(), P.snend.begin(), P.snend.end());
            ru.insert(ru.end(), P.ru.begin(), P.ru.end());
            fcut.insert(fcut.end(), P.fcut.begin(), P.fcut.end());
            bcut.insert(bcut.end(), P.bcut.begin(), P.bcut.end());
            fpk.insert(fpk.end(), P.fpk.begin(), P.fpk.end());
            bpk.insert(bpk.end(), P.bpk.begin(), P.bpk.end());
            twide.insert(twide.end(), P.twide.begin(), P.twide.end());
            rel.insert(rel.end(), P.rel.begin(), P.rel.end());
            for (int k = 0; k < C.n; k++) dpos.push_back(P.pinv[P.q[k]]);
            const LspSell &S = C.sell;
            soff.insert(soff.end(), S.off.begin(), S.off.end());
            std::vector<int> rl(C.n);
            for (int k = 0; k < C.n; k++) rl[k] = 0;
            // stored row lengths in slice order
            {
                std::vector<int> rp, ci, rsrc;
                lsp_coo_to_csr(C.n, C.nnz, C.row.data(), C.col.data(), rp, ci, rsrc);
                for (int k = 0; k < C.n; k++) rl[k] = rp[S.perm[k] + 1] - rp[S.perm[k]];
            }
            srlen.insert(srlen.end(), rl.begin(), rl.end());
            scol.insert(scol.end(), S.col.begin(), S.col.end());
            sperm.insert(sperm.end(), S.perm.begin(), S.perm.end());
            for (int s : S.src) ssrc.push_back(s >= 0 ? (int)(s + C.aoff) : -1);
        }
        d_cp_ = upload(cp); d_relp_ = upload(relp); d_soff_ = upload(soff);
        d_dp_ = upload(dp); d_pos_ = upload(pos); d_aptr_ = upload(aptr);
        d_asrc_ = upload(asrc); d_aslot_ = upload(aslot); d_dst_ = upload(dst);
        d_lrp_ = upload(lrp); d_lci_ = upload(lci); d_urp_ = upload(urp); d_uci_ = upload(uci);
        d_rowat_ = upload(rowat); d_q_ = upload(q); d_rel_ = upload(rel); d_snend_ = upload(snend);
        d_ru_ = upload(ru); d_fcut_ = upload(fcut); d_bcut_ = upload(bcut);
        d_fpk_ = upload(fpk); d_bpk_ = upload(bpk); d_twide_ = upload(twide);
        d_srlen_ = upload(srlen); d_scol_ = upload(scol); d_sperm_ = upload(sperm);
        d_ssrc_ = upload(ssrc);
        d_dpos_ = upload(dpos);
        if (!d_dpos_) return LSP_E_ALLOC;
        // fused IR step level lists (built by the host phase)
        {
            std::vector<int> frow, flvp, brow, blvp;
            long lf = 0, lb = 0;
            for (auto &C : circ_) {
                C.o_lf = lf;
                C.o_lb = lb;
                frow.insert(frow.end(), C.frow.begin(), C.frow.end());
                brow.insert(brow.end(), C.brow.begin(), C.brow.end());
                flvp.insert(flvp.end(), C.flvp.begin(), C.flvp.end());
                blvp.insert(blvp.end(), C.blvp.begin(), C.blvp.end());
                lf += (long)C.flvp.size();
                lb += (long)C.blvp.size();
            }
            d_frow_ = upload(frow); d_flvp_ = upload(flvp); d_brow_ = upload(brow); d_blvp_ = upload(blvp);
            std::vector<int> zc(circ_.size(), 0);
            std::vector<double> zd(circ_.size(), 0.0);
            d_irconv_ = upload(zc); d_rn2_ = upload(zd); d_bn2_ = upload(zd);
            if (!d_frow_ || !d_flvp_ || !d_brow_ || !d_blvp_ || !d_irconv_ || !d_rn2_ || !d_bn2_)
                return LSP_E_ALLOC;
        }
        if (!d_cp_ || !d_relp_ || !d_soff_ || !d_dp_ || !d_pos_ || !d_aptr_ || !d_asrc_ ||
            !d_aslot_ || !d_dst_ || !d_lrp_ || !d_lci_ || !d_urp_ || !d_uci_ || !d_rowat_ ||
            !d_q_ || !d_rel_ || !d_snend_ || !d_ru_ || !d_fcut_ || !d_bcut_ || !d_fpk_ || !d_bpk_ || !d_twide_ || !d_srlen_ || !d_scol_ || !d_sperm_ || !d_ssrc_)
            return LSP_E_ALLOC;
    }
    // flags start at 0; epochs start at 1
    {
        std::vector<int> zeros(ntot_, 0);
        d_done_ = upload(zeros);
        std::vector<int> z6(8, 0);
        std::vector<int> zq(3 * kQueueInts, 0);    // refactor, forward, backward queue blocks
        d_ctr_ = upload(zq); d_abort_ = upload(z6);
        std::vector<int> e1(4, 1);                 // epochs start at 1, flags at 0
        d_epoch_ = upload(e1);
        std::vector<int> ms(kMstat * nc, 0);
        d_mstat_ = upload(ms);
        d_dbg_ = upload(z6);
        if (!d_done_ || !d_ctr_ || !d_abort_ || !d_mstat_ || !d_epoch_)
            return LSP_E_ALLOC;
    }
    // dense trailing blocks
    {
        std::vector<int> tds, tlc, tails;
        std::vector<DenseJob> jobs;
        d_tailS_ = (double *)dalloc(nt2_tot_ * sizeof(double));
        d_colmax_ = (unsigned long long *)dalloc(nt_tot_ * sizeof(unsigned long long));
        int ntail = 0;
        for (const auto &C : circ_) ntail += C.plan.tail_T > 0;
        d_tailW_ = (double *)dalloc((size_t)std::max(ntail, 1) * 2 * 64 * 64 * sizeof(double));
        d_dflag_ = upload(std::vector<int>(std::max(ntail, 1), 0));
        if (!d_tailS_ || !d_colmax_ || !d_tailW_ || !d_dflag_) return LSP_E_ALLOC;
        for (int m = 0; m < nc; m++) {
            const CircuitHost &C = circ_[m];
            const LspPlan &P = C.plan;
            tds.insert(tds.end(), P.tail_ds.begin(), P.tail_ds.end());
            tlc.insert(tlc.end(), P.tail_lcut.begin(), P.tail_lcut.end());
            if (P.tail_T > 0) {
                DenseJob J{};
                J.S = d_tailS_ + C.o_t2;
                J.W = d_tailW_ + jobs.size() * 2 * 64 * 64;
                J.colmax = d_colmax_ + C.o_t;
                J.flag = d_dflag_ + jobs.size();
                J.T = P.tail_T;
                J.mat = m;
                jobs.push_back(J);
                tails.push_back(m);
                max_tail_T_ = std::max(max_tail_T_, P.tail_T);
            }
        }
        njobs_ = (int)jobs.size();
        d_tail_ds_ = upload(tds);
        d_tail_lcut_ = upload(tlc);
        d_djobs_ = upload(jobs);
        d_tails_ = upload(tails);
        d_tail_ctr_ = upload(std::vector<int>(4 * nc, 0));
        if (!d_tail_ds_ || !d_tail_lcut_ || !d_djobs_ || !d_tails_ || !d_tail_ctr_)
            return LSP_E_ALLOC;
        if (njobs_) {
            rc = check(dense_set_lds(), "dense lds attr");
            if (rc) return rc;
        }
    }
    d_ain_ = (double *)dalloc(nnz_tot_ * sizeof(double));
    // Ruiz equilibration for FP16 factor storage: per-entry circuit-local
    // row / column, scaled values, row / column scales, scratch
    {
        std::vector<int> ar, ac;
        std::vector<long> ao(nc + 1, 0), vo(nc + 1, 0);
        ar.reserve(nnz_tot_); ac.reserve(nnz_tot_);
        for (int m = 0; m < nc; m++) {
            const CircuitHost &C = circ_[m];
            ar.insert(ar.end(), C.row.begin(), C.row.end());
            ac.insert(ac.end(), C.col.begin(), C.col.end());
            ao[m + 1] = ao[m] + C.nnz;
            vo[m + 1] = vo[m] + C.n;
        }
        d_arow_ = upload(ar); d_acol_ = upload(ac);
        d_raoff_ = upload(ao); d_rvoff_ = upload(vo);
        d_ains_ = (double *)dalloc(nnz_tot_ * sizeof(double));
        d_rscale_ = (double *)dalloc(ntot_ * sizeof(double));
        d_cscale_ = (double *)dalloc(ntot_ * sizeof(double));
        d_rmx_ = (unsigned long long *)dalloc(ntot_ * sizeof(unsigned long long));
        if (!d_arow_ || !d_acol_ || !d_raoff_ || !d_rvoff_ || !d_ains_ || !d_rscale_ || !d_cscale_ || !d_rmx_)
            return LSP_E_ALLOC;
        if (!(opt_.factor_shift >= 0.0)) return LSP_E_ARG;   // negative or NaN
        if (opt_.factor_shift > 0.0) {
            // the first value entry of each diagonal, in caller order
            // (test_lu_gpu_f32.cu:356-361 stops at the first match per row)
            std::vector<long> sidx;
            for (int m = 0; m < nc; m++) {
                const CircuitHost &C = circ_[m];
                std::vector<char> seen(C.n, 0);
                for (long e = 0; e < C.nnz; e++)
                    if (C.row[e] == C.col[e] && !seen[C.row[e]]) {
                        seen[C.row[e]] = 1;
                        sidx.push_back(C.aoff + e);
                    }
            }
            nsidx_ = (long)sidx.size();
            if (nsidx_ && !(d_sidx_ = upload(sidx))) return LSP_E_ALLOC;
        }
    }
    d_sval64_ = (double *)dalloc(nsell_tot_ * sizeof(double));
    d_sval32_ = (float *)dalloc(nsell_tot_ * sizeof(float));
    if (!d_ain_ || !d_sval64_ || !d_sval32_) return LSP_E_ALLOC;
    if (opt_.krylov_spmv < 0 || opt_.krylov_spmv > 2) return LSP_E_ARG;
    if (opt_.krylov_spmv == 2 && !(d_sval16_ = (__half *)dalloc(nsell_tot_ * sizeof(__half))))
        return LSP_E_ALLOC;
    {
        std::vector<double> v0(nnz_tot_);
        for (auto &C : circ_) std::copy(C.val0.begin(), C.val0.end(), v0.begin() + C.aoff);
        rc = check(hipMemcpy(d_ain_, v0.data(), v0.size() * sizeof(double), hipMemcpyHostToDevice),
                   "upload values");
        if (rc) return rc;
    }
    // SELL descriptors and slice tasks
    {
        std::vector<SellDesc<double>> s64(nc);
        std::vector<SellDesc<float>> s32(nc);
        std::vector<int2> st;
        for (int m = 0; m < nc; m++) {
            const CircuitHost &C = circ_[m];
            s64[m].off = d_soff_ + C.o_slp1;
            s64[m].rlen = d_srlen_ + C.o_n;
            s64[m].col = d_scol_ + C.o_ss;
            s64[m].val = d_sval64_ + C.o_ss;
            s64[m].perm = d_sperm_ + C.o_n;
            s64[m].nslices = C.sell.nslices;
            s64[m].n = C.n;
            s32[m].off = s64[m].off; s32[m].rlen = s64[m].rlen; s32[m].col = s64[m].col;
            s32[m].val = d_sval32_ + C.o_ss; s32[m].perm = s64[m].perm;
            s32[m].nslices = C.sell.nslices; s32[m].n = C.n;
            for (int s = 0; s < C.sell.nslices; s++) st.push_back(make_int2(m, s));
        }
        d_sdesc64_ = upload(s64);
        d_sdesc32_ = upload(s32);
        if (d_sval16_) {
            std::vector<SellDesc<__half>> s16(nc);
            for (int m = 0; m < nc; m++) {
                s16[m].off = s64[m].off; s16[m].rlen = s64[m].rlen; s16[m].col = s64[m].col;
                s16[m].val = d_sval16_ + circ_[m].o_ss; s16[m].perm = s64[m].perm;
                s16[m].nslices = s64[m].nslices; s16[m].n = s64[m].n;
            }
            if (!(d_sdesc16_ = upload(s16))) return LSP_E_ALLOC;
        }
        d_stask_ = upload(st);
        nstask_ = (int)st.size();
        if (!d_sdesc64_ || !d_sdesc32_ || !d_stask_) return LSP_E_ALLOC;
    }
    // merged task lists: (level, circuit, index)
    {
        auto merge = [&](auto key_of, auto order_of, long &width) {
            std::vector<std::pair<long, int2>> all;
            all.reserve(ntot_);
            for (int m = 0; m < nc; m++) {
                const LspPlan &P = circ_[m].plan;
                const std::vector<int> &ord = order_of(P);
                for (size_t t = 0; t < ord.size(); t++) all.push_back({(long)key_of(P, t), make_int2(m, ord[t])});
            }
            std::stable_sort(all.begin(), all.end(),
                             [](const auto &a, const auto &b) { return a.first < b.first; });
            std::vector<int2> out(all.size());
            width = 0;
            for (size_t i = 0, run = 0; i < all.size(); i++) {
                out[i] = all[i].second;
                run = (i > 0 && all[i].first == all[i - 1].first) ? run + 1 : 1;
                width = std::max(width, (long)run);
            }
            return out;
        };
        auto tf = merge([](const LspPlan &P, size_t t) { return P.fkey[P.ftask[t]]; },
                        [](const LspPlan &P) -> const std::vector<int> & { return P.ftask; },
                        width_f_);
        auto tw = merge([](const LspPlan &P, size_t t) { return P.fwd_key[t]; },
                        [](const LspPlan &P) -> const std::vector<int> & { return P.fwd_task; },
                        width_tf_);
        auto tb = merge([](const LspPlan &P, size_t t) { return P.bwd_key[t]; },
                        [](const LspPlan &P) -> const std::vector<int> & { return P.bwd_task; },
                        width_tb_);
        nfw_tot_ = (long)tw.size();
        nbw_tot_ = (long)tb.size();
        d_ftask_ = upload(tf); d_fwtask_ = upload(tw); d_bwtask_ = upload(tb);
        if (!d_ftask_ || !d_fwtask_ || !d_bwtask_) return LSP_E_ALLOC;
    }
    // persistent grids
    {
        // LDS doubles per refactor wave: the requested cap, raised towards the
        // longest column up to 2048 (two 4-wave blocks per CU's 160 KB; on
        // c1 that beats one block holding every column: 4.65 vs 4.91 ms).  A
        // negative request is taken as is.  Longer columns use global scratch.
        int maxcol_all = 0;
        for (auto &C : circ_) maxcol_all = std::max(maxcol_all, C.plan.max_col);
        int cap = opt_.lds_cap > 0 ? opt_.lds_cap : opt_.lds_cap < 0 ? -opt_.lds_cap : 1024;
        if (opt_.lds_cap >= 0 && maxcol_all > cap) cap = std::min(2048, (maxcol_all + 255) / 256 * 256);
        opt_.lds_cap = cap;
        const size_t lds = (size_t)cap * 8 * 4;
        if (lds > 65536) {
            rc = check(refactor_set_lds<float>(cap), "lds attr");
            if (!rc) rc = check(refactor_set_lds<double>(cap), "lds attr");
            if (!rc) rc = check(refactor_set_lds<__half>(cap), "lds attr");
            if (rc) return rc;
        }
        int per_cu = opt_.blocks_per_cu;
        if (per_cu <= 0) {
            const size_t lds_cu = 160 * 1024;
            per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, lds_cu / lds));
        }
        // Waves beyond the widest dependency level only poll; size each
        // persistent grid to about twice that width (4 waves per block).
        auto blocks_for = [&](long width, long cap_blocks) {
            long want = (2 * width + 3) / 4;
            return (int)std::max(1L, std::min(cap_blocks, want));
        };
        fblocks_ = blocks_for(width_f_, (long)per_cu * cu_count_);
        // The solves' waves mostly hold a task that waits for its values:
        // one 4-wave block per CU keeps the queue ahead of the ready front
        // (c1 pair 0.950 -> 0.919 ms against the width rule alone, 512
        // circuits 27.4k -> 29.0k circuit steps/s; two blocks per CU and
        // more were slower on both).
        tblocks_f_ = blocks_for(width_tf_, cu_count_);
        tblocks_b_ = blocks_for(width_tb_, cu_count_);
        if (opt_.refactor_blocks > 0) fblocks_ = opt_.refactor_blocks;
        if (opt_.trisolve_blocks > 0) tblocks_f_ = tblocks_b_ = opt_.trisolve_blocks;
        // Wide grids split each task queue over 4 counters: one counter's
        // atomics held the task rate to about 85 tasks/us (c1 refactor 3.54
        // -> 3.11 ms, solve pair 0.856 -> 0.728 ms; 512 circuits 29.2k ->
        // 46.6k fp32 and 39.1k -> 56.8k fp64 circuit steps/s; 2 counters
        // gave about two thirds of that).  Small grids keep one counter,
        // shared by all their waves.
        nq_f_ = fblocks_ >= 16 ? 4 : 1;
        nq_t_ = std::min(tblocks_f_, tblocks_b_) >= 16 ? 4 : 1;
        int maxcol = 0;
        for (auto &C : circ_) maxcol = std::max(maxcol, C.plan.max_col);
        scratch_stride_ = maxcol > cap ? maxcol : 1;
        d_scratch_ = (double *)dalloc((size_t)fblocks_ * 4 * scratch_stride_ * sizeof(double));
        if (!d_scratch_) return LSP_E_ALLOC;
    }
    // vectors + Krylov state
    {
        const int mm = std::max(1, opt_.ir_inner);
        if (!d_b_) d_b_ = (double *)dalloc(ntot_ * 8);     // kept by a rebuild (with their contents)
        if (!d_x_) d_x_ = (double *)dalloc(ntot_ * 8);
        d_r_ = (double *)dalloc(ntot_ * 8); d_w_ = (double *)dalloc(ntot_ * 8);
        d_z_ = (double *)dalloc(ntot_ * 8); d_t_ = (double *)dalloc(ntot_ * 8);
        d_ty_ = (double *)dalloc(ntot_ * 8); d_tz_ = (double *)dalloc(ntot_ * 8);
        d_typ_ = (double *)dalloc(ntot_ * 8); d_tzp_ = (double *)dalloc(ntot_ * 8);
        std::vector<long> voff(nc);
        std::vector<int4> chunks;
        std::vector<int> cbeg(nc + 1, 0);
        int cr = opt_.chunk_rows > 0 ? opt_.chunk_rows : 512;
        for (int m = 0; m < nc; m++) {
            voff[m] = circ_[m].voff;
            cbeg[m] = (int)chunks.size();
            for (int r0 = 0; r0 < circ_[m].n; r0 += cr)
                chunks.push_back(make_int4(m, r0, std::min(circ_[m].n, r0 + cr), 0));
        }
        cbeg[nc] = (int)chunks.size();
        nchunks_ = (int)chunks.size();
        d_voff_ = upload(voff);
        d_chunks_ = upload(chunks);
        d_cbeg_ = upload(cbeg);
        rc = alloc_krylov(mm);
        if (rc) return rc;
        d_beta_ = (double *)dalloc(nc * 8); d_nb_ = (double *)dalloc(nc * 8);
        d_hn_ = (double *)dalloc(nc * 8);
        d_act_ = (int *)dalloc(nc * 4); d_jfin_ = (int *)dalloc(nc * 4);
        d_conv_ = (int *)dalloc(nc * 4); d_cycled_ = (int *)dalloc(nc * 4);
        d_nact_ = (int *)dalloc(16); d_subset_ = (int *)dalloc(nc * 4);
        if (!d_b_ || !d_x_ || !d_r_ || !d_w_ || !d_z_ || !d_t_ || !d_ty_ || !d_tz_ || !d_typ_ || !d_tzp_ ||
            !d_voff_ || !d_chunks_ || !d_cbeg_ || !d_act_)
            return LSP_E_ALLOC;
        if (!h_pin_ && hipHostMalloc((void **)&h_pin_, 4096, hipHostMallocDefault) != hipSuccess)
            return LSP_E_ALLOC;
    }
    for (auto &e : ev_) {
        if (e) continue;
        rc = check(hipEventCreate(&e), "event");
        if (rc) return rc;
    }
    // initial SELL values
    rc = check(launch_gather_cast<double>(nsell_tot_, d_ssrc_, d_ain_, d_sval64_, stream_), "gather");
    if (!rc) rc = check(launch_gather_cast<float>(nsell_tot_, d_ssrc_, d_ain_, d_sval32_, stream_), "gather");
    if (!rc && d_sval16_)
        rc = check(launch_gather_cast<__half>(nsell_tot_, d_ssrc_, d_ain_, d_sval16_, stream_), "gather");
    if (!rc) rc = check(hipStreamSynchronize(stream_), "sync");
    uploaded_ = rc == LSP_OK;
    return rc;
}

// Frees every device array of the batch except the b / x arenas (callers may
// hold their addresses, and b carries the caller's right-hand sides) and
// resets the derived sizes; upload_all() then rebuilds from circ_.  The
// stream, events and pinned host words stay.
void Batch::release_device()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    graph_prec_ = -1;
    double *keep_b = d_b_, *keep_x = d_x_;
    std::vector<DevMem> kept;
    for (auto &m : mem_) {
        if (m.p == keep_b || m.p == keep_x) kept.push_back(m);
        else if (m.p) (void)hipFree(m.p);
    }
    mem_.swap(kept);
    static_cast<DeviceArrays &>(*this) = DeviceArrays{};
    d_b_ = keep_b;
    d_x_ = keep_x;
    cur_prec_ = -1;
    uploaded_ = false;
    tri_pending_ = false;
}

int Batch::set_stream(hipStream_t s)
{
    if (!s) return LSP_E_ARG;
    if (own_stream_ && stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
    own_stream_ = false;
    stream_ = s;
    return LSP_OK;
}

int Batch::set_values(const double *vals, bool on_device)
{
    if (!vals) return LSP_E_ARG;
    int rc = check(hipMemcpyAsync(d_ain_, vals, nnz_tot_ * sizeof(double),
                                  on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                  stream_), "set values");
    if (rc) return rc;
    rc = check(launch_gather_cast<double>(nsell_tot_, d_ssrc_, d_ain_, d_sval64_, stream_), "gather");
    if (!rc) rc = check(launch_gather_cast<float>(nsell_tot_, d_ssrc_, d_ain_, d_sval32_, stream_), "gather");
    if (!rc && d_sval16_)
        rc = check(launch_gather_cast<__half>(nsell_tot_, d_ssrc_, d_ain_, d_sval16_, stream_), "gather");
    return rc;
}

KrylovArgs Batch::kargs() const
{
    KrylovArgs K{};
    K.nmat = (int)circ_.size();
    K.mmax = std::max(1, opt_.ir_inner);
    K.ntot = ntot_;
    K.voff = d_voff_;
    K.chunks = d_chunks_;
    K.nchunks = nchunks_;
    K.cbeg = d_cbeg_;
    K.partial = d_partial_;
    K.H = d_H_; K.cs = d_cs_; K.sn = d_sn_; K.g = d_g_; K.yv = d_yv_;
    K.beta = d_beta_; K.nb = d_nb_; K.hn = d_hn_; K.hcol = d_hcol_;
    K.act = d_act_; K.jfin = d_jfin_; K.conv = d_conv_; K.cycled = d_cycled_;
    K.inner_tol = opt_.ir_inner_tol;
    K.outer_tol = opt_.ir_outer_tol;
    return K;
}

// ---------------------------------------------------------------------------
template <typename T>
int Batch::refactor_t(FactorSet<T> &fs, const int *d_active)
{
    int rc = ensure_factor_set(fs);
    if (rc) return rc;
    FactorLaunch L{};
    L.mats = fs.fdesc;
    L.task = d_ftask_;
    L.active = d_active;
    L.ntask = (int)ntot_;
    L.epoch = d_epoch_;
    L.timeline = nullptr;
    if (refactor_tl_path_) {                                  // diagnostic dump
        if (!d_timeline_) d_timeline_ = (long long *)dalloc((size_t)ntot_ * 4 * sizeof(long long));
        L.timeline = d_timeline_;
    }
    L.ctr = d_ctr_;
    L.abort_word = d_abort_;
    L.scratch = d_scratch_;
    L.scratch_stride = scratch_stride_;
    L.cap = opt_.lds_cap;
    L.total_waves = fblocks_ * 4;
    L.nq = nq_f_;
    L.poll_backoff = opt_.poll_backoff > 0 ? opt_.poll_backoff : 7;
    L.ptol = opt_.refactor_ptol;
    L.stol = opt_.repivot == 3 ? opt_.pivot_tol : 0.0;
    L.dbg = d_dbg_;
    rc = check(launch_reset_mstat(d_mstat_, (int)circ_.size(), d_abort_, d_dbg_, stream_), "reset");
    if (rc) return rc;
    long max_slots = 0;
    for (const auto &C : circ_) max_slots = std::max(max_slots, (long)C.plan.cp[C.n]);
    rc = check(launch_fill_unset<T>(fs.fdesc, d_active, (int)circ_.size(), max_slots, stream_), "unset");
    if (rc) return rc;
    if (!capturing_) (void)hipEventRecord(ev_[0], stream_);   // the persistent kernel alone, from here
    rc = check(launch_refactor<T>(L, fblocks_, stream_), "refactor launch");
    if (!capturing_) (void)hipEventRecord(ev_[1], stream_);      // the persistent kernel alone
    if (!rc && njobs_) {
        rc = check(launch_dense_lu(d_djobs_, njobs_, max_tail_T_, d_active, d_abort_, stream_), "dense lu");
        if (!rc)
            rc = check(launch_dense_finish<T>(d_djobs_, njobs_, max_tail_T_, fs.fdesc, d_active,
                                              opt_.refactor_ptol, L.stol, stream_), "dense finish");
    }
    if (!capturing_) (void)hipEventRecord(ev_[5], stream_);      // + dense trailing block
    return rc;
}

// The refactor launch sequence of one precision (no host synchronisation).
int Batch::launch_refactor_prec(int prec, const int *d_active)
{
    const bool shift = opt_.factor_shift > 0.0;
    if (shift && (prec == LSP_PREC_FP32 || prec == LSP_PREC_FP64)) {
        // the factored copy: caller values with the diagonal shifted
        int rc = check(k_copy(d_ain_, d_ains_, nnz_tot_, stream_), "shift copy");
        if (!rc) rc = check(launch_shift(d_ain_, d_ains_, d_sidx_, nsidx_, opt_.factor_shift, stream_), "shift");
        if (rc) return rc;
    }
    if (prec == LSP_PREC_FP32 && opt_.refactor_fp64) {
        // refactor on fp64 columns, then the fp32 set the solves read is its
        // rounding (every array, so exports and pivot ranges see fp32 storage)
        int rc = ensure_factor_set(f32_);
        if (!rc) rc = refactor_t(f64_, d_active);
        if (!rc) rc = check(launch_round_f32(f64_.val, f32_.val, nslot_tot_, stream_), "round val");
        if (!rc) rc = check(launch_round_f32(f64_.lrow, f32_.lrow, nl_tot_, stream_), "round lrow");
        if (!rc) rc = check(launch_round_f32(f64_.urow, f32_.urow, nu_tot_, stream_), "round urow");
        if (!rc) rc = check(launch_round_f32(f64_.diag, f32_.diag, ntot_, stream_), "round diag");
        if (!rc) rc = check(launch_round_f32(f64_.tailF, f32_.tailF, nt2_tot_, stream_), "round tail");
        return rc;
    }
    if (prec == LSP_PREC_FP32) return refactor_t(f32_, d_active);
    if (prec == LSP_PREC_FP64) return refactor_t(f64_, d_active);
    if (prec == LSP_PREC_FP16) {
        // FP16 storage only after Ruiz equilibration: R A C is factored, and
        // the solves apply R to their input and C to their output
        RuizArgs R{};
        R.ain = d_ain_; R.aout = d_ains_; R.row = d_arow_; R.col = d_acol_;
        R.aoff = d_raoff_; R.voff = d_rvoff_; R.rs = d_rscale_; R.cs = d_cscale_; R.mx = d_rmx_;
        R.active = d_active; R.sweeps = opt_.ruiz_sweeps > 0 ? opt_.ruiz_sweeps : 30;
        int rc = check(launch_ruiz(R, (int)circ_.size(), stream_), "ruiz");
        if (!rc && shift)     // the shift is relative, so it commutes with the diagonal scaling
            rc = check(launch_shift(d_ains_, d_ains_, d_sidx_, nsidx_, opt_.factor_shift, stream_), "shift");
        return rc ? rc : refactor_t(f16_, d_active);
    }
    return LSP_E_ARG;
}

// One refactor launch sequence over the circuits of `subset` (all when null)
// and a host read of their status words (kMstat per circuit into ms).
int Batch::refactor_pass(int prec, const std::vector<int> *subset, std::vector<int> &ms)
{
    const int *d_active = nullptr;
    int rc;
    if (subset) {
        rc = check(hipMemcpyAsync(d_subset_, subset->data(), circ_.size() * sizeof(int),
                                  hipMemcpyHostToDevice, stream_), "subset");
        if (rc) return rc;
        rc = check(hipStreamSynchronize(stream_), "sync");
        if (rc) return rc;
        d_active = d_subset_;
    }
    rc = launch_refactor_prec(prec, d_active);
    if (rc) return rc;
    cur_prec_ = prec;
    ms.assign((size_t)kMstat * circ_.size(), 0);
    rc = check(hipMemcpyAsync(ms.data(), d_mstat_, ms.size() * 4, hipMemcpyDeviceToHost, stream_), "mstat");
    if (!rc) rc = check(hipMemcpyAsync(h_pin_, d_abort_, 4, hipMemcpyDeviceToHost, stream_), "abort");
    if (!rc) rc = check(hipStreamSynchronize(stream_), "refactor sync");
    if (rc) return rc;
    float t_kernel = 0.0f, t_all = 0.0f;
    (void)hipEventElapsedTime(&t_kernel, ev_[0], ev_[1]);
    (void)hipEventElapsedTime(&t_all, ev_[0], ev_[5]);
    t_refactor_ = t_all;
    stats_.refactor_ms += t_kernel;
    stats_.dense_ms += t_all - t_kernel;
    stats_.refactor_count++;
    if (refactor_tl_path_ && d_timeline_) {
        std::vector<long long> tl((size_t)ntot_ * 4);
        (void)hipMemcpy(tl.data(), d_timeline_, tl.size() * sizeof(long long), hipMemcpyDeviceToHost);
        std::vector<int2> tk(ntot_);
        (void)hipMemcpy(tk.data(), d_ftask_, ntot_ * sizeof(int2), hipMemcpyDeviceToHost);
        if (FILE *f = fopen(refactor_tl_path_, "wb")) {
            fwrite(tl.data(), sizeof(long long), tl.size(), f);
            fwrite(tk.data(), sizeof(int2), tk.size(), f);
            fclose(f);
        }
    }
    if (h_pin_[0] == kErrSpin) {
        int dbg[8] = {0};
        (void)hipMemcpy(dbg, d_dbg_, sizeof(dbg), hipMemcpyDeviceToHost);
        int flag_now = -99;
        if (dbg[0] >= 0 && dbg[0] < ntot_)
            (void)hipMemcpy(&flag_now, d_done_ + dbg[0], 4, hipMemcpyDeviceToHost);
        last_error = "refactor wait gave up: column " + std::to_string(dbg[4]) + " waiting on " +
                     std::to_string(dbg[0]) + " saw " + std::to_string(dbg[1]) + " epoch " +
                     std::to_string(dbg[2]) + " spins " + std::to_string(dbg[3]) +
                     "; flag in memory after the launch: " + std::to_string(flag_now);
    }
    return LSP_OK;
}

// New pivot sequences for the candidate circuits, chosen on the host by the
// reference's partial pivoting (lsp_gp_factor, tol = pivot_tol) on each
// circuit's current values.  A circuit whose new sequence equals its current
// one keeps its structure; if any changes, the device state is rebuilt with
// every circuit's current values (release_device + upload_all).
// Whether a circuit's status words (kMstat) call for a new pivot sequence
// under opt_.repivot: 1 a zero pivot, 2 also a pivot failing refactor_ptol,
// 3 also any column where splu (tol = pivot_tol) would pick another pivot.
bool Batch::repivot_wanted(const int *w) const
{
    const bool zero = w[2] != 0x7fffffff;
    const bool unstable = w[0] > 0;
    const bool differs = w[4] > 0 || w[5] > 0;
    return (opt_.repivot >= 1 && zero) || (opt_.repivot >= 2 && unstable) || (opt_.repivot >= 3 && differs);
}

int Batch::repivot(const std::vector<int> &cands, int *changed)
{
    *changed = 0;
    if (cands.empty()) return LSP_OK;
    std::vector<double> vals(nnz_tot_);
    int rc = check(hipMemcpyAsync(vals.data(), d_ain_, nnz_tot_ * sizeof(double), hipMemcpyDeviceToHost,
                                  stream_), "values");
    if (!rc) rc = check(hipStreamSynchronize(stream_), "sync");
    if (rc) return rc;
    std::vector<CircuitHost> fresh(cands.size());
    std::vector<int> rcs(cands.size(), LSP_OK);
    {
        std::atomic<int> next{0};
        const int nt = (int)std::min<size_t>(cands.size(),
                                             std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; t++)
            pool.emplace_back([&]() {
                for (int i = next++; i < (int)cands.size(); i = next++) {
                    const CircuitHost &C = circ_[cands[i]];
                    lsp_circuit_in in{};
                    in.n = C.n;
                    in.nnz = C.nnz;
                    in.row = C.row.data();
                    in.col = C.col.data();
                    in.val = vals.data() + C.aoff;
                    in.q = C.q.data();
                    rcs[i] = build_circuit(fresh[i], in, opt_);
                }
            });
        for (auto &th : pool) th.join();
    }
    for (size_t i = 0; i < cands.size(); i++) {
        if (rcs[i] == LSP_E_SINGULAR) continue;             // singular for the reference too
        if (rcs[i] != LSP_OK) return rcs[i];
        CircuitHost &C = circ_[cands[i]];
        if (fresh[i].plan.pinv == C.plan.pinv) continue;    // the host confirms the current pivots
        C.sym = std::move(fresh[i].sym);
        C.plan = std::move(fresh[i].plan);
        C.sell = std::move(fresh[i].sell);
        C.frow = std::move(fresh[i].frow);
        C.flvp = std::move(fresh[i].flvp);
        C.brow = std::move(fresh[i].brow);
        C.blvp = std::move(fresh[i].blvp);
        (*changed)++;
    }
    if (*changed == 0) return LSP_OK;
    // upload_all() takes every circuit's values from val0
    for (auto &C : circ_) std::copy(vals.begin() + C.aoff, vals.begin() + C.aoff + C.nnz, C.val0.begin());
    release_device();
    rc = upload_all();
    if (rc && last_error.empty()) last_error = "rebuild after new pivot sequences failed";
    return rc;
}

// FP16 factor storage is refused unless the options opt in (allow_fp16): the
// path once faulted the GPU and the fault's cause was not shown.
int Batch::check_prec(int prec)
{
    if (prec == LSP_PREC_FP32 || prec == LSP_PREC_FP64) return LSP_OK;
    if (prec == LSP_PREC_FP16 && opt_.allow_fp16) return LSP_OK;
    last_error = prec == LSP_PREC_FP16 ? "FP16 factor storage is off (lsp_options.allow_fp16 = 0)"
                                       : "unknown factor precision";
    return LSP_E_ARG;
}

int Batch::refactor(int prec, lsp_refactor_report *rep, const std::vector<int> *subset)
{
    const int nc = (int)circ_.size();
    std::vector<int> ms;
    int rc = check_prec(prec);
    if (rc) return rc;
    if (!subset && !active_.empty()) subset = &active_;
    rc = refactor_pass(prec, subset, ms);
    if (rc) return rc;
    auto in_subset = [&](int m) { return !subset || (*subset)[m]; };
    lsp_refactor_report r{};
    for (int m = 0; m < nc; m++)
        if (in_subset(m)) r.pivot_mismatch += ms[kMstat * m + 4];
    if (opt_.repivot > 0 && h_pin_[0] == kErrNone) {
        std::vector<int> cands;
        for (int m = 0; m < nc; m++) {
            if (!in_subset(m)) continue;
            const int *w = &ms[kMstat * m];
            if (repivot_wanted(w)) cands.push_back(m);
        }
        int changed = 0;
        rc = repivot(cands, &changed);
        if (rc) return rc;
        if (changed) {
            // the rebuild dropped every factor set: refactor every circuit
            r.repivoted = changed;
            rc = refactor_pass(prec, nullptr, ms);
            if (rc) return rc;
        }
    }
    const int abort_code = h_pin_[0];
    r.first_bad_circuit = -1;
    r.ms = t_refactor_;
    for (int m = 0; m < nc; m++) {
        if (!in_subset(m)) continue;
        const int *w = &ms[kMstat * m];
        r.bad_pivots += w[0];
        if (w[2] != 0x7fffffff) r.singular_circuits++;
        if ((w[0] || w[2] != 0x7fffffff || w[3]) && r.first_bad_circuit < 0) r.first_bad_circuit = m;
    }
    r.status = abort_code == kErrSpin ? LSP_E_DEVICE
             : abort_code == kErrCapacity ? LSP_E_CAPACITY
             : r.singular_circuits ? LSP_E_SINGULAR : LSP_OK;
    if (rep) *rep = r;
    return r.status;
}

// ---------------------------------------------------------------------------
template <typename T>
int Batch::trisolve_t(FactorSet<T> &fs, const double *d_b, double *d_x, const int *d_active, const int *d_nact)
{
    if (!fs.ready) return LSP_E_ARG;
    TriLaunch Lf{};
    Lf.mats = fs.tdesc;
    Lf.task = d_fwtask_;
    Lf.voff = d_voff_;
    Lf.active = d_active;
    Lf.b = d_b;
    Lf.x = d_x;
    Lf.y = d_ty_;
    Lf.z = d_tz_;
    Lf.yp = d_typ_;
    Lf.zp = d_tzp_;
    Lf.ntask = (int)nfw_tot_;
    Lf.ctr = d_ctr_ + kQueueInts;
    Lf.abort_word = d_abort_ + 1;
    Lf.total_waves = tblocks_f_ * 4;
    Lf.nq = nq_t_;
    const bool scaled = std::is_same<T, __half>::value;
    Lf.rs = scaled ? d_rscale_ : nullptr;
    Lf.cs = scaled ? d_cscale_ : nullptr;
    Lf.n_active = d_nact;
    Lf.timeline = nullptr;
    if (trisolve_tl_path_) {                                 // diagnostic dump (Batch::trisolve)
        if (!d_tri_tl_) d_tri_tl_ = (long long *)dalloc((size_t)(nfw_tot_ + nbw_tot_) * 2 * sizeof(long long));
        Lf.timeline = d_tri_tl_;
    }
    TriLaunch Lb = Lf;
    Lb.timeline = Lf.timeline ? Lf.timeline + 2 * nfw_tot_ : nullptr;
    Lb.task = d_bwtask_;
    Lb.ntask = (int)nbw_tot_;
    Lb.ctr = d_ctr_ + 2 * kQueueInts;
    Lb.total_waves = tblocks_b_ * 4;
    Lb.nq = nq_t_;
    if (!capturing_) (void)hipEventRecord(ev_[2], stream_);
    int rc = check(launch_solve_unset(Lf, ntot_, stream_), "solve unset");
    if (!rc) rc = check(launch_trisolve_half<T>(Lf, true, tblocks_f_, stream_), "trisolve fwd");
    if (!rc && njobs_)
        rc = check(launch_tail_solve<T>(fs.tdesc, d_tails_, njobs_, max_tail_T_, d_voff_, d_active, d_b,
                                        d_ty_, d_typ_, d_tz_, d_x, true, d_abort_ + 1, Lf.rs, Lf.cs,
                                        stream_),
                   "tail fwd");
    if (!capturing_) (void)hipEventRecord(ev_[4], stream_);
    if (!rc && njobs_)
        rc = check(launch_tail_solve<T>(fs.tdesc, d_tails_, njobs_, max_tail_T_, d_voff_, d_active, d_b,
                                        d_ty_, d_typ_, d_tz_, d_x, false, d_abort_ + 1, Lf.rs, Lf.cs,
                                        stream_),
                   "tail bwd");
    if (!rc) rc = check(launch_trisolve_half<T>(Lb, false, tblocks_b_, stream_), "trisolve bwd");
    if (!capturing_) (void)hipEventRecord(ev_[3], stream_);
    tri_pending_ = true;
    return rc;
}

int Batch::launch_trisolve_prec(int prec, const double *d_b, double *d_x, const int *active, const int *d_nact)
{
    if (prec == LSP_PREC_FP32) return trisolve_t(f32_, d_b, d_x, active, d_nact);
    if (prec == LSP_PREC_FP64) return trisolve_t(f64_, d_b, d_x, active, d_nact);
    if (prec == LSP_PREC_FP16) return trisolve_t(f16_, d_b, d_x, active, d_nact);
    return LSP_E_ARG;
}

// Arnoldi product w = A z with the value precision chosen by krylov_spmv
// (fp64 accumulation in every case; the residual always uses launch_spmv64)
int Batch::launch_spmv_krylov(const double *d_xv, double *d_y, const int *active)
{
    const int64_t *vo = (const int64_t *)d_voff_;
    if (opt_.krylov_spmv == 1)
        return check(launch_sell_spmv<float>(d_sdesc32_, d_stask_, nstask_, vo, d_xv, d_y, nullptr, active,
                                             stream_), "spmv32");
    if (opt_.krylov_spmv == 2)
        return check(launch_sell_spmv<__half>(d_sdesc16_, d_stask_, nstask_, vo, d_xv, d_y, nullptr, active,
                                              stream_), "spmv16");
    return launch_spmv64(d_xv, d_y, nullptr, active);
}

int Batch::launch_spmv64(const double *d_xv, double *d_y, const double *d_bv, const int *active)
{
    return check(launch_sell_spmv<double>(d_sdesc64_, d_stask_, nstask_, (const int64_t *)d_voff_,
                                          d_xv, d_y, d_bv, active, stream_), "spmv");
}

int Batch::trisolve(const double *b, double *x, bool on_device)
{
    if (cur_prec_ < 0) return LSP_E_ARG;
    const double *db = b;
    double *dx = x;
    int rc;
    if (!on_device) {
        rc = check(hipMemcpyAsync(d_b_, b, ntot_ * 8, hipMemcpyHostToDevice, stream_), "b");
        if (rc) return rc;
        db = d_b_;
        dx = d_x_;
    }
    rc = check(k_zero_int(d_abort_ + 1, 1, stream_), "abort reset");
    if (rc) return rc;
    rc = launch_trisolve_prec(cur_prec_, db, dx, nullptr);
    if (rc) return rc;
    if (!on_device) {
        rc = check(hipMemcpyAsync(x, d_x_, ntot_ * 8, hipMemcpyDeviceToHost, stream_), "x");
        if (rc) return rc;
    }
    rc = check(hipMemcpyAsync(h_pin_, d_abort_ + 1, 4, hipMemcpyDeviceToHost, stream_), "abort");
    if (!rc) rc = check(hipStreamSynchronize(stream_), "sync");
    if (rc) return rc;
    account_trisolve();
    const char *tl_path = trisolve_tl_path_;
    if (tl_path && d_tri_tl_) {
        // [nfw, nbw] (int64), forward then backward task lists (int32 x 2 per
        // task), then per task: ticket granted, task finished (int64 ticks)
        const size_t nt = (size_t)(nfw_tot_ + nbw_tot_);
        std::vector<long long> tl(2 * nt);
        std::vector<int2> tf(nfw_tot_), tb(nbw_tot_);
        (void)hipMemcpy(tl.data(), d_tri_tl_, tl.size() * sizeof(long long), hipMemcpyDeviceToHost);
        (void)hipMemcpy(tf.data(), d_fwtask_, tf.size() * sizeof(int2), hipMemcpyDeviceToHost);
        (void)hipMemcpy(tb.data(), d_bwtask_, tb.size() * sizeof(int2), hipMemcpyDeviceToHost);
        if (FILE *f = fopen(tl_path, "wb")) {
            const long long hdr[2] = {nfw_tot_, nbw_tot_};
            fwrite(hdr, sizeof(long long), 2, f);
            fwrite(tf.data(), sizeof(int2), tf.size(), f);
            fwrite(tb.data(), sizeof(int2), tb.size(), f);
            fwrite(tl.data(), sizeof(long long), tl.size(), f);
            fclose(f);
        }
    }
    return h_pin_[0] ? LSP_E_DEVICE : LSP_OK;
}

int Batch::spmv(const double *x, double *y, bool on_device, int prec)
{
    const double *dx = x;
    double *dy = y;
    int rc;
    if (!on_device) {
        rc = check(hipMemcpyAsync(d_t_, x, ntot_ * 8, hipMemcpyHostToDevice, stream_), "x");
        if (rc) return rc;
        dx = d_t_;
        dy = d_w_;
    }
    if (prec == LSP_PREC_FP32)
        rc = check(launch_sell_spmv<float>(d_sdesc32_, d_stask_, nstask_, (const int64_t *)d_voff_,
                                           dx, dy, nullptr, nullptr, stream_), "spmv32");
    else
        rc = launch_spmv64(dx, dy, nullptr, nullptr);
    if (rc) return rc;
    if (!on_device) {
        rc = check(hipMemcpyAsync(y, d_w_, ntot_ * 8, hipMemcpyDeviceToHost, stream_), "y");
        if (rc) return rc;
    }
    return check(hipStreamSynchronize(stream_), "sync");
}

// ---------------------------------------------------------------------------
int Batch::gmres_ir(const double *d_bv, double *d_xv, int prec, const std::vector<char> *subset,
                    lsp_solve_report *rep, bool keep_x)
{
    const KrylovArgs K = kargs();
    const int nc = (int)circ_.size();
    const int mm = K.mmax;
    hipStream_t s = stream_;
    int rc;
#define LSP_TRY(expr) do { rc = (expr); if (rc) return rc; } while (0)
#define LSP_HIP(expr, what) LSP_TRY(check((expr), what))
    // conv = 1 for circuits outside the subset, 0 inside; x = 0 inside
    // unless keep_x (x then holds a first iterate, e.g. after fused steps)
    {
        std::vector<int> cv(nc, 0);
        if (subset) for (int m = 0; m < nc; m++) cv[m] = (*subset)[m] ? 0 : 1;
        LSP_HIP(hipMemcpyAsync(d_conv_, cv.data(), nc * 4, hipMemcpyHostToDevice, s), "conv");
        LSP_HIP(hipStreamSynchronize(s), "sync");
    }
    LSP_HIP(k_activate(K, s), "activate");
    LSP_HIP(k_chunk_dots(K, d_bv, d_bv, 1, s), "dots");
    LSP_HIP(k_reduce(K, 1, 2, 0, d_nb_, s), "reduce");
    if (!keep_x) LSP_HIP(k_chunk_fill(K, d_xv, 0.0, s), "zero x");
    int outer_used = 0, inner_total = 0;
    res_hist0_.clear();
    for (int outer = 0; outer < opt_.ir_outer; outer++) {
        LSP_HIP(k_activate(K, s), "activate");
        LSP_TRY(launch_spmv64(d_xv, d_r_, d_bv, d_act_));
        LSP_HIP(k_chunk_dots(K, d_r_, d_r_, 1, s), "dots");
        LSP_HIP(k_reduce(K, 1, 2, 0, d_beta_, s), "reduce");
        LSP_HIP(k_zero_int(d_nact_, 1, s), "zero");
        LSP_HIP(k_start_cycle(K, d_nact_, s), "start");
        {
            double bn[2];
            LSP_HIP(hipMemcpyAsync(bn, d_beta_, 8, hipMemcpyDeviceToHost, s), "beta0");
            LSP_HIP(hipMemcpyAsync(bn + 1, d_nb_, 8, hipMemcpyDeviceToHost, s), "nb0");
            LSP_TRY(sync_read_int(d_nact_, h_pin_, 1));
            res_hist0_.push_back(bn[1] > 0 ? bn[0] / bn[1] : bn[0]);
        }
        if (h_pin_[0] == 0) break;
        outer_used++;
        LSP_HIP(k_chunk_scale(K, d_r_, d_V_, d_beta_, s), "scale");
        // Z_j = M^-1 V_j is kept, so the update x += M^-1 (V y) = Z y needs
        // no further solve (the flexible-GMRES form of right preconditioning)
        for (int j = 0; j < mm; j++) {
            double *Vj = d_V_ + (long)j * ntot_;
            double *Zj = d_Z_ + (long)j * ntot_;
            LSP_TRY(launch_trisolve_prec(prec, Vj, Zj, d_act_));
            LSP_TRY(launch_spmv_krylov(Zj, d_w_, d_act_));
            for (int pass = 0; pass < 2; pass++) {
                LSP_HIP(k_chunk_dots(K, d_V_, d_w_, j + 1, s), "dots");
                LSP_HIP(k_reduce(K, j + 1, 1, j, nullptr, s), "reduce");
                LSP_HIP(k_chunk_update(K, d_V_, d_w_, j + 1, s), "update");
            }
            LSP_HIP(k_chunk_dots(K, d_w_, d_w_, 1, s), "dots");
            LSP_HIP(k_reduce(K, 1, 2, 0, d_hn_, s), "reduce");
            LSP_HIP(k_chunk_scale(K, d_w_, d_V_ + (long)(j + 1) * ntot_, d_hn_, s), "scale");
            LSP_HIP(k_zero_int(d_nact_, 1, s), "zero");
            LSP_HIP(k_givens(K, j, d_nact_, s), "givens");
            inner_total++;
            LSP_TRY(sync_read_int(d_nact_, h_pin_, 1));
            if (h_pin_[0] == 0) break;
        }
        LSP_HIP(k_backsolve(K, s), "backsolve");
        LSP_HIP(k_chunk_combine(K, d_Z_, d_t_, s), "combine");
        LSP_HIP(k_chunk_axpy(K, d_xv, d_t_, s), "axpy");
    }
    // final true residual of every circuit in the subset
    {
        std::vector<int> am(nc, 1);
        if (subset) for (int m = 0; m < nc; m++) am[m] = (*subset)[m] ? 1 : 0;
        LSP_HIP(hipMemcpyAsync(d_act_, am.data(), nc * 4, hipMemcpyHostToDevice, s), "act");
        LSP_TRY(launch_spmv64(d_xv, d_r_, d_bv, d_act_));
        LSP_HIP(k_chunk_dots(K, d_r_, d_r_, 1, s), "dots");
        LSP_HIP(k_reduce(K, 1, 2, 0, d_beta_, s), "reduce");
        std::vector<double> beta(nc), nb(nc);
        LSP_HIP(hipMemcpyAsync(beta.data(), d_beta_, nc * 8, hipMemcpyDeviceToHost, s), "beta");
        LSP_HIP(hipMemcpyAsync(nb.data(), d_nb_, nc * 8, hipMemcpyDeviceToHost, s), "nb");
        LSP_HIP(hipMemcpyAsync(h_pin_, d_abort_ + 1, 4, hipMemcpyDeviceToHost, s), "abort");
        LSP_HIP(hipStreamSynchronize(s), "sync");
        account_trisolve();
        if (h_pin_[0]) return LSP_E_DEVICE;
        lsp_solve_report r{};
        r.outer = outer_used;
        r.inner_total = inner_total;
        for (int m = 0; m < nc; m++) {
            if (subset && !(*subset)[m]) continue;
            double rel = nb[m] > 0 ? beta[m] / nb[m] : beta[m];
            if (!(rel == rel)) rel = INFINITY;
            r.max_rel_res = std::max(r.max_rel_res, rel);
            if (rel <= opt_.ir_outer_tol * 10 || rel <= 1e-9) r.converged++;
        }
        r.status = LSP_OK;
        if (rep) *rep = r;
    }
#undef LSP_TRY
#undef LSP_HIP
    return LSP_OK;
}

// ---------------------------------------------------------------------------
// Fused IR steps (irstep.hip)
template <typename T>
int Batch::ir_launch(FactorSet<T> &fs, const double *d_b, double *d_x, double tol2, bool first)
{
    if (!fs.idesc) return LSP_E_ARG;
    int rc = ir_prepare();
    if (rc) return rc;
    IrStepArgs A{};
    A.descs = fs.idesc;
    A.b = d_b;
    A.x = d_x;
    A.r = d_r_;
    A.yw = d_ty_;
    A.conv = d_irconv_;
    A.rn2 = d_rn2_;
    A.bn2 = d_bn2_;
    A.tol2 = tol2;
    A.lds_n = ir_lds_n_;
    A.first = first ? 1 : 0;
    return check(launch_ir_step<T>(A, (int)circ_.size(), stream_), "ir step");
}

// The fused step's LDS work vector: up to 8192 doubles (64 KB, two workgroups
// per CU); larger circuits use the work arena.  Set once, before any capture.
int Batch::ir_prepare()
{
    if (ir_lds_n_) return LSP_OK;
    int maxn = 0;
    for (const auto &C : circ_) maxn = std::max(maxn, C.n);
    const int lds_n = std::min(8192, std::max(64, maxn));
    int rc = check(ir_step_set_lds<float>(lds_n), "ir lds attr");
    if (!rc) rc = check(ir_step_set_lds<double>(lds_n), "ir lds attr");
    if (!rc) ir_lds_n_ = lds_n;
    return rc;
}

bool Batch::use_fused(int prec) const
{
    return opt_.ir_method == 1 && prec != LSP_PREC_FP16;
}

int Batch::fused_steps() const
{
    return opt_.ir_fused_steps > 0 ? opt_.ir_fused_steps : 12;
}

int Batch::ir_launch_prec(int prec, const double *d_b, double *d_x, double tol2, bool first)
{
    if (prec == LSP_PREC_FP32) return ir_launch(f32_, d_b, d_x, tol2, first);
    if (prec == LSP_PREC_FP64) return ir_launch(f64_, d_b, d_x, tol2, first);
    return LSP_E_ARG;                        // FP16 factors are solved with their Ruiz scales by GMRES-IR
}

int Batch::fused_ir(const double *d_bv, double *d_xv, int prec, const std::vector<char> *subset, int max_steps,
                    std::vector<int> &conv, lsp_solve_report *rep)
{
    const int nc = (int)circ_.size();
    conv.assign(nc, 0);
    std::vector<int> skip(nc, 0);
    for (int m = 0; m < nc; m++) skip[m] = subset && !(*subset)[m] ? 1 : 0;
    int rc = check(hipMemcpyAsync(d_irconv_, skip.data(), nc * 4, hipMemcpyHostToDevice, stream_), "conv");
    if (rc) return rc;
    const double tol2 = opt_.ir_outer_tol * opt_.ir_outer_tol;
    std::vector<int> cv(nc);
    int steps = 0, done = 0;
    const int want = nc - (int)std::count(skip.begin(), skip.end(), 1);
    while (steps < max_steps) {
        rc = ir_launch_prec(prec, d_bv, d_xv, tol2, steps == 0);      // the first step starts from x = 0
        if (rc) return rc;
        steps++;
        rc = check(hipMemcpyAsync(cv.data(), d_irconv_, nc * 4, hipMemcpyDeviceToHost, stream_), "conv");
        if (!rc) rc = check(hipStreamSynchronize(stream_), "sync");
        if (rc) return rc;
        done = 0;
        for (int m = 0; m < nc; m++) done += !skip[m] && cv[m];
        if (done == want) break;
    }
    std::vector<double> rn(nc), bn(nc);
    rc = check(hipMemcpyAsync(rn.data(), d_rn2_, nc * 8, hipMemcpyDeviceToHost, stream_), "rn");
    if (!rc) rc = check(hipMemcpyAsync(bn.data(), d_bn2_, nc * 8, hipMemcpyDeviceToHost, stream_), "bn");
    if (!rc) rc = check(hipStreamSynchronize(stream_), "sync");
    if (rc) return rc;
    lsp_solve_report r{};
    r.outer = 1;
    r.inner_total = steps;
    for (int m = 0; m < nc; m++) {
        if (skip[m]) continue;
        conv[m] = cv[m];
        r.converged += cv[m];
        double rel = bn[m] > 0 ? std::sqrt(rn[m] / bn[m]) : std::sqrt(rn[m]);
        if (!(rel == rel)) rel = INFINITY;
        r.max_rel_res = std::max(r.max_rel_res, rel);
    }
    r.status = LSP_OK;
    if (rep) *rep = r;
    return LSP_OK;
}

int Batch::ir_step(const double *b, double *x, bool on_device, double *rel_res)
{
    if (cur_prec_ < 0 || circ_.empty()) return LSP_E_ARG;
    const int nc = (int)circ_.size();
    const double *db = b;
    double *dx = x;
    int rc;
    if (!on_device) {
        rc = check(hipMemcpyAsync(d_b_, b, ntot_ * 8, hipMemcpyHostToDevice, stream_), "b");
        if (!rc) rc = check(hipMemcpyAsync(d_x_, x, ntot_ * 8, hipMemcpyHostToDevice, stream_), "x");
        if (rc) return rc;
        db = d_b_;
        dx = d_x_;
    }
    std::vector<int> z(nc, 0);
    rc = check(hipMemcpyAsync(d_irconv_, z.data(), nc * 4, hipMemcpyHostToDevice, stream_), "conv");
    if (!rc) rc = ir_launch_prec(cur_prec_, db, dx, -1.0, false);   // no tolerance: every circuit steps
    if (rc) return rc;
    std::vector<double> rn(nc), bn(nc);
    rc = check(hipMemcpyAsync(rn.data(), d_rn2_, nc * 8, hipMemcpyDeviceToHost, stream_), "rn");
    if (!rc) rc = check(hipMemcpyAsync(bn.data(), d_bn2_, nc * 8, hipMemcpyDeviceToHost, stream_), "bn");
    if (!rc && !on_device) rc = check(hipMemcpyAsync(x, d_x_, ntot_ * 8, hipMemcpyDeviceToHost, stream_), "x");
    if (!rc) rc = check(hipStreamSynchronize(stream_), "sync");
    if (rc) return rc;
    if (rel_res)
        for (int m = 0; m < nc; m++) rel_res[m] = bn[m] > 0 ? std::sqrt(rn[m] / bn[m]) : std::sqrt(rn[m]);
    return LSP_OK;
}

int Batch::solve(const double *b, double *x, bool on_device, lsp_solve_report *rep)
{
    if (cur_prec_ < 0) return LSP_E_ARG;
    const double *db = b;
    double *dx = x;
    int rc;
    if (!on_device) {
        rc = check(hipMemcpyAsync(d_b_, b, ntot_ * 8, hipMemcpyHostToDevice, stream_), "b");
        if (rc) return rc;
        db = d_b_;
        dx = d_x_;
    }
    rc = check(k_zero_int(d_abort_ + 1, 1, stream_), "abort reset");
    if (rc) return rc;
    // the circuits solved: the set_active mask, else all
    const int nc = (int)circ_.size();
    std::vector<char> live;
    int nlive = nc;
    if (!active_.empty()) {
        live.assign(active_.begin(), active_.end());
        nlive = (int)std::count(active_.begin(), active_.end(), 1);
    }
    lsp_solve_report r{};
    if (use_fused(cur_prec_)) {
        // fused Richardson steps first; the circuits they leave are solved by GMRES-IR
        std::vector<int> conv;
        rc = fused_ir(db, dx, cur_prec_, live.empty() ? nullptr : &live, fused_steps(), conv, &r);
        if (rc) return rc;
        if (r.converged < nlive) {
            std::vector<char> rest(nc, 0);
            for (int m = 0; m < nc; m++) rest[m] = (live.empty() || live[m]) && !conv[m];
            lsp_solve_report rg{};
            rc = gmres_ir(db, dx, cur_prec_, &rest, &rg, true);
            if (rc) return rc;
            r.converged += rg.converged;
            r.outer += rg.outer;
            r.inner_total += rg.inner_total;
            // the residual norms of every circuit, as gmres_ir's final check left them
            std::vector<double> beta(nc), nb(nc);
            (void)hipMemcpy(beta.data(), d_beta_, nc * 8, hipMemcpyDeviceToHost);
            (void)hipMemcpy(nb.data(), d_nb_, nc * 8, hipMemcpyDeviceToHost);
            std::vector<double> rn(nc), bn(nc);
            (void)hipMemcpy(rn.data(), d_rn2_, nc * 8, hipMemcpyDeviceToHost);
            (void)hipMemcpy(bn.data(), d_bn2_, nc * 8, hipMemcpyDeviceToHost);
            r.max_rel_res = 0.0;
            for (int m = 0; m < nc; m++) {
                if (!(live.empty() || live[m])) continue;
                double rel = rest[m] ? (nb[m] > 0 ? beta[m] / nb[m] : beta[m])
                                     : (bn[m] > 0 ? std::sqrt(rn[m] / bn[m]) : std::sqrt(rn[m]));
                if (!(rel == rel)) rel = INFINITY;
                r.max_rel_res = std::max(r.max_rel_res, rel);
            }
            // the fp64 fallback below reads the final residuals from d_beta_ /
            // d_nb_: give the fused circuits theirs
            for (int m = 0; m < nc; m++)
                if (!rest[m]) {
                    beta[m] = std::sqrt(rn[m]);
                    nb[m] = std::sqrt(bn[m]);
                }
            (void)hipMemcpy(d_beta_, beta.data(), nc * 8, hipMemcpyHostToDevice);
            (void)hipMemcpy(d_nb_, nb.data(), nc * 8, hipMemcpyHostToDevice);
        }
    } else {
        rc = gmres_ir(db, dx, cur_prec_, live.empty() ? nullptr : &live, &r);
        if (rc) return rc;
    }
    // per-circuit fallback to fp64 factors when IR stalls
    if (opt_.fallback_fp64 && cur_prec_ != LSP_PREC_FP64 && r.converged < nlive) {
        // identify stalled circuits from the final residuals
        std::vector<double> beta(nc), nb(nc);
        (void)hipMemcpy(beta.data(), d_beta_, nc * 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(nb.data(), d_nb_, nc * 8, hipMemcpyDeviceToHost);
        std::vector<int> sub(nc, 0);
        std::vector<char> subc(nc, 0);
        int cnt = 0;
        for (int m = 0; m < nc; m++) {
            if (!live.empty() && !live[m]) continue;
            double rel = nb[m] > 0 ? beta[m] / nb[m] : beta[m];
            if (!(rel <= opt_.ir_outer_tol * 10 || rel <= 1e-9)) { sub[m] = 1; subc[m] = 1; cnt++; }
        }
        if (cnt) {
            const int keep = cur_prec_;
            lsp_refactor_report rr{};
            rc = refactor(LSP_PREC_FP64, &rr, &sub);
            // a rebuild (new pivot sequences) refactored every circuit in fp64
            // and dropped the other factor sets
            if (!rr.repivoted) cur_prec_ = keep;
            if (rc && rc != LSP_E_SINGULAR) return rc;
            r.singular_fallback = rr.singular_circuits;
            rc = check(k_zero_int(d_abort_ + 1, 1, stream_), "abort reset");
            if (rc) return rc;
            lsp_solve_report r2{};
            rc = gmres_ir(db, dx, LSP_PREC_FP64, &subc, &r2);
            if (rc) return rc;
            r.fallback = cnt;
            r.converged = r.converged + r2.converged;
            r.outer = std::max(r.outer, r2.outer);
            r.inner_total += r2.inner_total;
            // recompute the max residual over all circuits
            std::vector<double> b2(nc);
            (void)hipMemcpy(b2.data(), d_beta_, nc * 8, hipMemcpyDeviceToHost);
            r.max_rel_res = 0;
            for (int m = 0; m < nc; m++) {
                if (!live.empty() && !live[m]) continue;
                double rel = sub[m] ? (nb[m] > 0 ? b2[m] / nb[m] : b2[m])
                                    : (nb[m] > 0 ? beta[m] / nb[m] : beta[m]);
                if (!(rel == rel)) rel = INFINITY;             // a diverged circuit stays visible
                r.max_rel_res = std::max(r.max_rel_res, rel);
            }
        }
    }
    if (!on_device) {
        rc = check(hipMemcpyAsync(x, d_x_, ntot_ * 8, hipMemcpyDeviceToHost, stream_), "x");
        if (!rc) rc = check(hipStreamSynchronize(stream_), "sync");
        if (rc) return rc;
    }
    r.status = r.converged == nlive ? LSP_OK : LSP_E_NOT_CONVERGED;
    if (rep) *rep = r;
    return r.status;
}

// ---------------------------------------------------------------------------
int Batch::info(int c, lsp_circuit_info *inf) const
{
    if (c < 0 || c >= (int)circ_.size() || !inf) return LSP_E_ARG;
    const CircuitHost &C = circ_[c];
    inf->n = C.n;
    inf->nnz = C.nnz;
    inf->lnnz = C.sym.lnnz;
    inf->unnz = C.sym.unnz;
    inf->n_updates = C.plan.n_updates;
    inf->nlevels = C.plan.nlevels;
    inf->nlv_fwd = C.plan.nlv_f;
    inf->nlv_bwd = C.plan.nlv_b;
    inf->max_col = C.plan.max_col;
    inf->sell_slots = C.sell.slots;
    inf->tail_T = C.plan.tail_T;
    inf->tail_updates = C.plan.tail_updates;
    return LSP_OK;
}

int Batch::export_factors(int c, int *pinv, int *Lp, int *Li, double *Lx, int *Up, int *Ui,
                          double *Ux)
{
    if (c < 0 || c >= (int)circ_.size() || cur_prec_ < 0) return LSP_E_ARG;
    const CircuitHost &C = circ_[c];
    const LspPlan &P = C.plan;
    const long ns = P.cp[C.n];
    std::vector<double> v(ns);
    int rc;
    if (cur_prec_ == LSP_PREC_FP64) {
        rc = check(hipMemcpy(v.data(), f64_.val + C.o_slot, ns * 8, hipMemcpyDeviceToHost), "export");
    } else if (cur_prec_ == LSP_PREC_FP32) {
        std::vector<float> f(ns);
        rc = check(hipMemcpy(f.data(), f32_.val + C.o_slot, ns * 4, hipMemcpyDeviceToHost), "export");
        for (long i = 0; i < ns; i++) v[i] = f[i];
    } else {
        std::vector<__half> f(ns);
        rc = check(hipMemcpy(f.data(), f16_.val + C.o_slot, ns * 2, hipMemcpyDeviceToHost), "export");
        for (long i = 0; i < ns; i++) v[i] = (double)__half2float(f[i]);
    }
    if (rc) return rc;
    if (pinv) std::copy(P.pinv.begin(), P.pinv.end(), pinv);
    long l = 0, u = 0;
    for (int k = 0; k < C.n; k++) {
        if (Lp) Lp[k] = (int)l;
        if (Up) Up[k] = (int)u;
        const long c0 = P.cp[k], d = P.dp[k], c1 = P.cp[k + 1];
        for (long s = c0; s < d; s++) {
            if (Ui) Ui[u] = P.pos[s];
            if (Ux) Ux[u] = v[s];
            u++;
        }
        if (Ui) Ui[u] = k;
        if (Ux) Ux[u] = v[d];
        u++;
        if (Li) Li[l] = k;
        if (Lx) Lx[l] = 1.0;
        l++;
        for (long s = d + 1; s < c1; s++) {
            if (Li) Li[l] = P.pos[s];
            if (Lx) Lx[l] = v[s];
            l++;
        }
    }
    if (Lp) Lp[C.n] = (int)l;
    if (Up) Up[C.n] = (int)u;
    return LSP_OK;
}

// ---------------------------------------------------------------------------
// One Newton step as a replayed HIP graph: refactor (prec) + one GMRES cycle
// of up to m inner iterations + the true residual, on the resident b / x.
// Nothing inside synchronises with the host: circuits that converge stop
// through the device activity mask, so a fixed launch sequence serves every
// step.  The graph is recorded on first use (and again when prec changes);
// epochs and queue counters live on the device, so replays need no
// arguments.  After the replay the host reads the reports once; a circuit
// still above the tolerance is finished by the eager solve (restart from 0,
// fp64 fallback as in solve()).
constexpr int kGraphInner = 8;          // GMRES inner iterations recorded in the step graph

int Batch::record_step(int prec)
{
    const KrylovArgs K = kargs();
    const int nc = (int)circ_.size();
    const int mm = K.mmax;
    hipStream_t s = stream_;
    const double *d_bv = d_b_;
    double *d_xv = d_x_;
    int rc;
#define LSP_TRY(expr) do { rc = (expr); if (rc) return rc; } while (0)
#define LSP_HIP(expr, what) LSP_TRY(check((expr), what))
    LSP_TRY(launch_refactor_prec(prec, nullptr));
    LSP_HIP(k_zero_int(d_abort_ + 1, 1, s), "abort reset");
    if (use_fused(prec)) {
        // fused Richardson steps: a converged circuit's later launches return
        // at once, so the recorded count only bounds the work
        LSP_HIP(k_zero_int(d_irconv_, nc, s), "conv");
        const double tol2 = opt_.ir_outer_tol * opt_.ir_outer_tol;
        for (int j = 0; j < fused_steps(); j++) LSP_TRY(ir_launch_prec(prec, d_bv, d_xv, tol2, j == 0));
        return LSP_OK;
    }
    LSP_HIP(k_zero_int(d_conv_, nc, s), "conv");
    LSP_HIP(k_activate(K, s), "activate");
    LSP_HIP(k_chunk_dots(K, d_bv, d_bv, 1, s), "dots");
    LSP_HIP(k_reduce(K, 1, 2, 0, d_nb_, s), "reduce");
    LSP_HIP(k_chunk_fill(K, d_xv, 0.0, s), "zero x");
    LSP_HIP(k_activate(K, s), "activate");
    LSP_TRY(launch_spmv64(d_xv, d_r_, d_bv, d_act_));
    LSP_HIP(k_chunk_dots(K, d_r_, d_r_, 1, s), "dots");
    LSP_HIP(k_reduce(K, 1, 2, 0, d_beta_, s), "reduce");
    LSP_HIP(k_zero_int(d_nact_, 1, s), "zero");
    LSP_HIP(k_start_cycle(K, d_nact_, s), "start");
    LSP_HIP(k_chunk_scale(K, d_r_, d_V_, d_beta_, s), "scale");
    // a replay runs every recorded iteration: record the few a step usually
    // needs; an idle iteration's solves exit at once (n_active == 0)
    const int mg = std::min(mm, kGraphInner);
    for (int j = 0; j < mg; j++) {
        double *Vj = d_V_ + (long)j * ntot_;
        double *Zj = d_Z_ + (long)j * ntot_;
        LSP_TRY(launch_trisolve_prec(prec, Vj, Zj, d_act_, d_nact_));
        LSP_TRY(launch_spmv_krylov(Zj, d_w_, d_act_));
        for (int pass = 0; pass < 2; pass++) {
            LSP_HIP(k_chunk_dots(K, d_V_, d_w_, j + 1, s), "dots");
            LSP_HIP(k_reduce(K, j + 1, 1, j, nullptr, s), "reduce");
            LSP_HIP(k_chunk_update(K, d_V_, d_w_, j + 1, s), "update");
        }
        LSP_HIP(k_chunk_dots(K, d_w_, d_w_, 1, s), "dots");
        LSP_HIP(k_reduce(K, 1, 2, 0, d_hn_, s), "reduce");
        LSP_HIP(k_chunk_scale(K, d_w_, d_V_ + (long)(j + 1) * ntot_, d_hn_, s), "scale");
        LSP_HIP(k_zero_int(d_nact_, 1, s), "zero");
        LSP_HIP(k_givens(K, j, d_nact_, s), "givens");
    }
    LSP_HIP(k_backsolve(K, s), "backsolve");
    LSP_HIP(k_chunk_combine(K, d_Z_, d_t_, s), "combine");
    LSP_HIP(k_chunk_axpy(K, d_xv, d_t_, s), "axpy");
    LSP_HIP(k_set_all_active(K, 1, s), "act");
    LSP_TRY(launch_spmv64(d_xv, d_r_, d_bv, d_act_));
    LSP_HIP(k_chunk_dots(K, d_r_, d_r_, 1, s), "dots");
    LSP_HIP(k_reduce(K, 1, 2, 0, d_beta_, s), "reduce");
#undef LSP_TRY
#undef LSP_HIP
    return LSP_OK;
}

int Batch::graph_step(int prec, lsp_refactor_report *rrep, lsp_solve_report *srep)
{
    if (!stream_ || circ_.empty()) return LSP_E_ARG;
    int rc = check_prec(prec);
    if (rc) return rc;
    if (!graph_exec_ || graph_prec_ != prec) {
        // every allocation happens before the capture
        if (prec == LSP_PREC_FP32) rc = ensure_factor_set(f32_);
        else if (prec == LSP_PREC_FP64) rc = ensure_factor_set(f64_);
        else if (prec == LSP_PREC_FP16) rc = ensure_factor_set(f16_);
        else return LSP_E_ARG;
        // the fp64-column refactor of the mixed path writes both sets
        if (!rc && prec == LSP_PREC_FP32 && opt_.refactor_fp64) rc = ensure_factor_set(f64_);
        if (rc) return rc;
        if (use_fused(prec)) {
            rc = ir_prepare();
            if (rc) return rc;
        }
        if (graph_exec_) { (void)hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr; }
        if (graph_) { (void)hipGraphDestroy(graph_); graph_ = nullptr; }
        rc = check(hipStreamSynchronize(stream_), "sync");
        if (rc) return rc;
        rc = check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
        if (rc) return rc;
        capturing_ = true;
        const int rc_rec = record_step(prec);
        capturing_ = false;
        hipGraph_t g = nullptr;
        rc = check(hipStreamEndCapture(stream_, &g), "end capture");
        if (rc_rec || rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc_rec ? rc_rec : rc;
        }
        graph_ = g;
        rc = check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "graph instantiate");
        if (rc) return rc;
        graph_prec_ = prec;
    }
    rc = check(hipGraphLaunch(graph_exec_, stream_), "graph launch");
    if (rc) return rc;
    cur_prec_ = prec;
    const int nc = (int)circ_.size();
    const bool fused = use_fused(prec);
    std::vector<int> ms(kMstat * nc), jf(nc), cv(nc, 0);
    std::vector<double> beta(nc), nb(nc);
    int ab[2] = {0, 0};
    rc = check(hipMemcpyAsync(ms.data(), d_mstat_, ms.size() * 4, hipMemcpyDeviceToHost, stream_), "mstat");
    if (fused) {
        // squared residual norms of the last step each circuit ran
        if (!rc) rc = check(hipMemcpyAsync(beta.data(), d_rn2_, nc * 8, hipMemcpyDeviceToHost, stream_), "rn2");
        if (!rc) rc = check(hipMemcpyAsync(nb.data(), d_bn2_, nc * 8, hipMemcpyDeviceToHost, stream_), "bn2");
        if (!rc) rc = check(hipMemcpyAsync(cv.data(), d_irconv_, nc * 4, hipMemcpyDeviceToHost, stream_), "conv");
    } else {
        if (!rc) rc = check(hipMemcpyAsync(beta.data(), d_beta_, nc * 8, hipMemcpyDeviceToHost, stream_), "beta");
        if (!rc) rc = check(hipMemcpyAsync(nb.data(), d_nb_, nc * 8, hipMemcpyDeviceToHost, stream_), "nb");
        if (!rc) rc = check(hipMemcpyAsync(jf.data(), d_jfin_, nc * 4, hipMemcpyDeviceToHost, stream_), "jfin");
    }
    if (!rc) rc = check(hipMemcpyAsync(h_pin_, d_abort_, 8, hipMemcpyDeviceToHost, stream_), "abort");
    if (!rc) rc = check(hipStreamSynchronize(stream_), "graph sync");
    if (rc) return rc;
    ab[0] = h_pin_[0];
    ab[1] = h_pin_[1];
    stats_.graph_steps++;
    lsp_refactor_report r{};
    r.first_bad_circuit = -1;
    for (int m = 0; m < nc; m++) {
        const int *w = &ms[kMstat * m];
        r.bad_pivots += w[0];
        r.pivot_mismatch += w[4];
        if (w[2] != 0x7fffffff) r.singular_circuits++;
        if ((w[0] || w[2] != 0x7fffffff || w[3]) && r.first_bad_circuit < 0) r.first_bad_circuit = m;
    }
    r.status = ab[0] == kErrSpin ? LSP_E_DEVICE : ab[0] == kErrCapacity ? LSP_E_CAPACITY
             : r.singular_circuits ? LSP_E_SINGULAR : LSP_OK;
    if (opt_.repivot > 0 && ab[0] == kErrNone) {
        std::vector<int> cands;
        for (int m = 0; m < nc; m++) {
            const int *w = &ms[kMstat * m];
            if (repivot_wanted(w)) cands.push_back(m);
        }
        int changed = 0;
        rc = repivot(cands, &changed);
        if (rc) return rc;
        if (changed) {
            // new pivot sequences dropped the graph and the factors: this
            // step runs eagerly (the next graph_step records a new graph)
            lsp_refactor_report r2{};
            rc = refactor(prec, &r2);
            r2.repivoted += changed;
            r2.pivot_mismatch = r.pivot_mismatch;
            if (rrep) *rrep = r2;
            if (rc) return rc;
            lsp_solve_report s2{};
            rc = solve(d_b_, d_x_, true, &s2);
            if (srep) *srep = s2;
            return rc == LSP_E_NOT_CONVERGED ? LSP_OK : rc;
        }
    }
    if (rrep) *rrep = r;
    if (r.status == LSP_E_DEVICE || r.status == LSP_E_CAPACITY) return r.status;
    if (ab[1]) return LSP_E_DEVICE;
    lsp_solve_report sr{};
    sr.outer = 1;
    int conv = 0, jmax = 0;
    for (int m = 0; m < nc; m++) {
        if (fused) {
            beta[m] = std::sqrt(beta[m]);
            nb[m] = std::sqrt(nb[m]);
        }
        double rel = nb[m] > 0 ? beta[m] / nb[m] : beta[m];
        if (!(rel == rel)) rel = INFINITY;
        sr.max_rel_res = std::max(sr.max_rel_res, rel);
        if (fused ? cv[m] != 0 : (rel <= opt_.ir_outer_tol * 10 || rel <= 1e-9)) conv++;
        jmax = std::max(jmax, jf[m]);
    }
    sr.inner_total = fused ? fused_steps() : jmax;
    sr.converged = conv;
    sr.status = LSP_OK;
    if (conv < nc || r.status != LSP_OK) {
        // finish eagerly (rare): full IR from x = 0 with the fp64 fallback
        lsp_solve_report s2{};
        rc = solve(d_b_, d_x_, true, &s2);
        if (rc) return rc;
        s2.inner_total += sr.inner_total;
        sr = s2;
    }
    if (srep) *srep = sr;
    return LSP_OK;
}

int Batch::device_vectors(double **b, double **x)
{
    if (b) *b = d_b_;
    if (x) *x = d_x_;
    return LSP_OK;
}

int Batch::copy_vector(const double *p, bool to_rhs)
{
    int rc = to_rhs ? check(hipMemcpyAsync(d_b_, p, ntot_ * 8, hipMemcpyDefault, stream_), "rhs")
                    : check(hipMemcpyAsync((void *)p, d_x_, ntot_ * 8, hipMemcpyDefault, stream_), "x");
    if (!rc) rc = check(hipStreamSynchronize(stream_), "sync");
    return rc;
}

int Batch::kernel_stats(lsp_kernel_stats *st, int reset)
{
    if (st) *st = stats_;
    if (reset) stats_ = lsp_kernel_stats{};
    return LSP_OK;
}

int Batch::last_times(double *rf, double *fw, double *bw) const
{
    if (rf) *rf = t_refactor_;
    if (fw) *fw = t_fwd_;
    if (bw) *bw = t_bwd_;
    return LSP_OK;
}

template int Batch::ensure_factor_set<float>(FactorSet<float> &);
template int Batch::ensure_factor_set<double>(FactorSet<double> &);
template int Batch::ensure_factor_set<__half>(FactorSet<__half> &);

}  // namespace lsp
