// engine.hip -- networks, work spaces and the GPU layer loops behind the
// C ABI (include/dnnf90.h, include/dnnf90_amd.h).
//
// The reference runs one point per call through the layer loop of
// lib/train_module.f90:241-450.  Here a call runs a batch of points: every
// layer is one linear-map GEMM over the na*Np column axis plus one
// Faa di Bruno launch, forward; one adjoint launch plus a gradient GEMM and a
// transposed linear-map GEMM, backward.  The per-point entries of the
// reference ABI are batches of one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/dnnf90_amd.h"
#include "codegen.hpp"
#include "kernels.hpp"
#include "tables.hpp"

#include <map>

namespace dnnf90 {

// generated Faa di Bruno kernels: on unless DNNF90_CODEGEN=0 or switched off
// through dnnf90_set_codegen; code objects are cached by source text
static int codegen_env()
{
    const char *e = std::getenv("DNNF90_CODEGEN");
    return (e && e[0] == '0') ? 0 : 1;
}
static std::atomic<int> g_codegen{codegen_env()};
static std::mutex g_gen_mu;
static std::map<std::string, std::vector<char>> g_gen_cache;

static bool codegen_on() { return g_codegen.load(std::memory_order_relaxed) != 0; }

// with the generated kernels: each hidden layer keeps S and the adjoint
// recomputes h from it (on unless switched off by dnnf90_set_store_s), or
// the layer keeps h as the table-driven kernels need
static int g_store_s = 1;

// one point through small networks in one launch each way (the generated
// one-point kernels); on unless switched off by dnnf90_set_point_kernels
static int g_point = 1;

// edge variants of the generated kernels (the first hidden layer's S formed
// from the input plane, the last hidden layer's Tbar from the output
// layer's Sbar): on unless switched off by dnnf90_set_edge_kernels
static int g_edge = 1;

// the fused layer kernels for the collocation step of FP32 work spaces
// (codegen.cpp, fused section): on unless switched off by
// dnnf90_set_fused_layers
static int g_fused_layers = 1;

static thread_local std::string g_err;

static int fail(int rc, const std::string &m)
{
    g_err = m;
    return rc;
}

#define HIPCHK(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            g_err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return e_ == hipErrorOutOfMemory ? -4 : -3;                              \
        }                                                                            \
    } while (0)

template <typename T>
static int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) return 0;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        g_err = std::string("hipMalloc: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? -4 : -3;
    }
    return 0;
}

// a failed launch or stream call: -4 when the device is out of memory, else -3
static int hip_fail(hipError_t e, const char *where)
{
    g_err = std::string(where) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? -4 : -3;
}

template <typename T>
static void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

template <typename T>
static int upload(T **dst, const std::vector<T> &src)
{
    int rc = dalloc(dst, src.empty() ? 1 : src.size());
    if (rc) return rc;
    if (!src.empty()) HIPCHK(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

// ---------------------------------------------------------------------------
// device copy of one table family's schedules
struct DevTables {
    int na = 0, k = 0, d0 = 0, iact = 0, nentry = 0;
    int *ind_e1 = nullptr, *slot_group = nullptr, *grp_term = nullptr, *grp_store = nullptr;
    TermRec64 *t64 = nullptr, *p64 = nullptr;
    TermRec32 *t32 = nullptr, *p32 = nullptr;
    double *tpc = nullptr;
    int *pstart = nullptr, *pnum = nullptr;
    bool ready = false;

    int build(const TableSet &t, const FwdSchedule &s)
    {
        na = t.na; k = t.k; d0 = t.d0; iact = t.iact; nentry = s.nentry;
        std::vector<TermRec64> h64(s.term_s.size()), q64(t.pair_ia.size());
        std::vector<TermRec32> h32(s.term_s.size()), q32(t.pair_ia.size());
        for (size_t i = 0; i < h64.size(); ++i) {
            h64[i] = TermRec64{s.term_s[i], s.term_b[i], s.term_c[i]};
            h32[i] = TermRec32{s.term_s[i], s.term_b[i], (float)s.term_c[i], 0};
        }
        for (size_t i = 0; i < q64.size(); ++i) {
            q64[i] = TermRec64{t.pair_ia[i], t.pair_id[i], t.pair_binom[i]};
            q32[i] = TermRec32{t.pair_ia[i], t.pair_id[i], (float)t.pair_binom[i], 0};
        }
        int rc = 0;
        if ((rc = upload(&ind_e1, t.ind_e1)) || (rc = upload(&slot_group, s.slot_group)) ||
            (rc = upload(&grp_term, s.grp_term)) || (rc = upload(&grp_store, s.grp_store)) ||
            (rc = upload(&t64, h64)) || (rc = upload(&t32, h32)) || (rc = upload(&pstart, t.pair_start)) ||
            (rc = upload(&pnum, t.pair_num)) || (rc = upload(&p64, q64)) || (rc = upload(&p32, q32)) ||
            (rc = upload(&tpc, t.tpc)))
            return rc;
        ready = true;
        return 0;
    }
    void release()
    {
        dfree(ind_e1); dfree(slot_group); dfree(grp_term); dfree(grp_store); dfree(t64); dfree(t32);
        dfree(pstart); dfree(pnum); dfree(p64); dfree(p32); dfree(tpc);
        ready = false;
    }
    DevFwdSched fwd() const
    {
        DevFwdSched f;
        f.slot_group = slot_group; f.grp_term = grp_term; f.grp_store = grp_store;
        f.t64 = t64; f.t32 = t32;
        f.na = na; f.k = k; f.nentry = nentry;
        return f;
    }
    DevPairs pairs() const
    {
        DevPairs p;
        p.start = pstart; p.num = pnum; p.p64 = p64; p.p32 = p32;
        p.na = na;
        return p;
    }
    ActParams act() const
    {
        ActParams a;
        a.iact = iact; a.nq = k + 1; a.tpc = tpc; a.ldt = k + 3;
        return a;
    }
};

struct Net {
    int nlayer = 0, nw = 0;
    std::vector<int> dims;
    std::vector<long> woff;  // per layer l >= 1
    TableSet tab;
    FwdSchedule sched;
    std::vector<double> w;   // host master copy
    long host_ver = 1, dev_ver = 0, dev32_ver = 0;
    DevTables dt;
    double *d_w64 = nullptr;
    float *d_w32 = nullptr;
    GenKernels gen[2];       // [0] FP32, [1] FP64
    bool gen_tried[2] = {false, false};
    FusedKernels fz;         // the fused layer kernels (FP32), compiled on first use
    bool fz_tried = false;
    std::mutex mu;
    // weight uploads and device-side weight steps run on the network's own
    // stream after the last launch of every stream that read the weights:
    // one event per such stream, recorded after its launches (used_on)
    hipStream_t ns = nullptr;
    mutable std::mutex use_mu;
    mutable std::vector<std::pair<hipStream_t, hipEvent_t>> uses;

    // kernels just enqueued on s read (or, for the filter, wrote) the device weights
    void used_on(hipStream_t s) const
    {
        std::lock_guard<std::mutex> lk(use_mu);
        for (auto &u : uses)
            if (u.first == s) { (void)hipEventRecord(u.second, s); return; }
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return;
        (void)hipEventRecord(e, s);
        uses.emplace_back(s, e);
    }
    // a work space's own stream is going away: drop its entry
    void forget(hipStream_t s) const
    {
        std::lock_guard<std::mutex> lk(use_mu);
        for (size_t i = 0; i < uses.size(); ++i)
            if (uses[i].first == s) {
                (void)hipEventSynchronize(uses[i].second);
                (void)hipEventDestroy(uses[i].second);
                uses.erase(uses.begin() + (long)i);
                return;
            }
    }
    // the network's stream, waiting for every recorded reader
    int own_stream_after_readers()
    {
        if (!ns) HIPCHK(hipStreamCreateWithFlags(&ns, hipStreamNonBlocking));
        std::lock_guard<std::mutex> lk(use_mu);
        for (auto &u : uses) HIPCHK(hipStreamWaitEvent(ns, u.second, 0));
        return 0;
    }

    // the generated kernels of one precision (compiled on first use)
    const GenKernels *generated(bool fp64)
    {
        if (!codegen_on()) return nullptr;
        const int i = fp64 ? 1 : 0;
        if (!gen_tried[i]) {
            gen_tried[i] = true;
            if (gen_bell_usable(tab, sched)) {
                const std::string src = gen_bell_source(tab, sched, fp64);
                std::lock_guard<std::mutex> lk(g_gen_mu);
                auto it = g_gen_cache.find(src);
                if (it != g_gen_cache.end()) {
                    (void)gen_bell_load(it->second, fp64, &gen[i]);
                } else {
                    std::string log;
                    if (gen_bell_compile(tab, sched, fp64, &gen[i], &log) == 0) g_gen_cache[src] = gen[i].code;
                }
            }
        }
        return gen[i].ok ? &gen[i] : nullptr;
    }

    int ensure_device(bool need32)
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!dt.ready) {
            int rc = dt.build(tab, sched);
            if (rc) return rc;
        }
        (void)generated(true);
        if (need32) (void)generated(false);
        if (!d_w64) {
            int rc = dalloc(&d_w64, (size_t)nw);
            if (rc) return rc;
        }
        if (dev_ver != host_ver || (need32 && dev32_ver != host_ver)) {
            // kernels of other streams may still read the previous weights:
            // the copy waits for this network's readers only
            int rc = own_stream_after_readers();
            if (rc) return rc;
            if (dev_ver != host_ver) {
                HIPCHK(hipMemcpyAsync(d_w64, w.data(), sizeof(double) * nw, hipMemcpyHostToDevice, ns));
                dev_ver = host_ver;
            }
            if (need32 && dev32_ver != dev_ver) {
                if (!d_w32 && (rc = dalloc(&d_w32, (size_t)nw))) return rc;
                launch_convert_f64_f32(d_w64, d_w32, nw, ns);
                dev32_ver = dev_ver;
            }
            HIPCHK(hipStreamSynchronize(ns));
        }
        return 0;
    }
    void release()
    {
        if (ns) (void)hipStreamSynchronize(ns);
        for (auto &u : uses) {
            (void)hipEventSynchronize(u.second);
            (void)hipEventDestroy(u.second);
        }
        uses.clear();
        if (ns) (void)hipStreamDestroy(ns);
        ns = nullptr;
        dt.release();
        dfree(d_w64);
        dfree(d_w32);
        gen[0].release();
        gen[1].release();
        fz.release();
    }

    // the fused layer kernels of the family (compiled on first use; code
    // objects cached by source text)
    const FusedKernels *fused()
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!codegen_on()) return nullptr;
        if (!fz_tried) {
            fz_tried = true;
            if (gen_fused_usable(tab, sched)) {
                const std::string src = gen_fused_source(tab, sched);
                std::lock_guard<std::mutex> lg(g_gen_mu);
                auto it = g_gen_cache.find(src);
                if (it != g_gen_cache.end()) {
                    (void)gen_fused_load(it->second, &fz);
                } else {
                    std::string log;
                    if (gen_fused_compile(tab, sched, &fz, &log) == 0) g_gen_cache[src] = fz.code;
                }
            }
        }
        return fz.ok ? &fz : nullptr;
    }
};

// gradient accumulator: kernels of any work space's stream add into d_g
// (one thread at a time per accumulator, the reference's contract); the
// event marks the last of them, and the accumulator's own stream does the
// clears and read-outs after it
struct Grad {
    int nw = 0;
    double *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
    int init()
    {
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        return 0;
    }
    // before kernels on st touch d_g: after the accumulator's last writer
    int before(hipStream_t st) const
    {
        HIPCHK(hipStreamWaitEvent(st, ev, 0));
        return 0;
    }
    // after kernels on st wrote d_g
    int after(hipStream_t st) const
    {
        HIPCHK(hipEventRecord(ev, st));
        return 0;
    }
    void release()
    {
        if (ev) (void)hipEventSynchronize(ev);
        dfree(d_g); dfree(d_m); dfree(d_v);
        if (ev) (void)hipEventDestroy(ev);
        if (s) (void)hipStreamDestroy(s);
        ev = nullptr;
        s = nullptr;
    }
};

// kernel categories timed by HIP events when profiling is on
enum {
    CAT_LINEAR = 0, CAT_BELL_FWD, CAT_BELL_BWD, CAT_GRAD, CAT_LINEAR_T, CAT_OTHER, CAT_GRAD_REDUCE,
    CAT_KF_PASS, CAT_KF_GAIN, NCAT
};

struct KProf {
    bool on = false;
    int only = -1;               // bracket this category alone (-1: every one)
    std::vector<hipEvent_t> ev;  // begin/end pairs
    std::vector<int> cat;        // one per pair
    size_t used = 0;             // events recorded since the last resolve
    double ms[NCAT] = {};
    long cnt[NCAT] = {};
};

// planes of one batched work space: [rows][na*Np] in T
struct BatchWork {
    int prec = PREC_FP64, maxpt = 0, npt = 0;
    long np_cap = 0, np = 0;
    size_t esz = 8;
    hipStream_t own = nullptr, stream = nullptr;
    void *t0 = nullptr, *tl = nullptr, *p0 = nullptr, *p1 = nullptr, *scratch = nullptr;
    std::vector<void *> tpl, hpl;  // per hidden layer (hpl unused when recomputing)
    // hpl[l] holds S of layer l when the last forward pass kept S (s_kept:
    // generated kernels, the adjoint recomputes h in registers), else h.
    // recompute: that plane is not kept from the forward pass; the reverse
    // pass recomputes it from the stored T of the layer below (K1, and the
    // Faa di Bruno forward when it holds h) into the plane hs
    bool recompute = false, s_kept = false;
    bool pt_used = false;   // the last forward went through the one-point kernel
    bool edge_in = false;   // the last forward formed the first hidden layer's S from T0 (no S plane kept)
    // the caller drains the stream before returning (the filter's
    // observation loop), so the forwards and adjoints record no reader event
    bool defer_use = false;
    bool rows_next = false; // the adjoint of this forward writes per-point rows (batched kernels)
    // debug mode (dnnf90_set_poison): the planes, scratch and partials below
    // are filled with NaN at creation and before every batched forward
    bool poison = false;
    std::vector<std::pair<void *, size_t>> poisoned;
    void *hs = nullptr;
    double bytes = 0.0, bytes_stored = 0.0;   // device bytes held / needed with every h kept
    void *hplane(int l) const { return recompute ? hs : hpl[l]; }
    double *partial = nullptr;
    long partial_cap = 0;
    double *d_x = nullptr, *d_seed = nullptr, *d_tout = nullptr, *d_rsq = nullptr;
    double *d_rsq_part = nullptr;
    double *d_coeff = nullptr;
    int *d_slot = nullptr, *d_nonlin = nullptr, nterm_cap = 0;
    // NGD metric and solve (allocated on first use)
    double *ngd_G = nullptr, *ngd_part = nullptr, *ngd_u = nullptr, *ngd_trace = nullptr, *ngd_work = nullptr;
    long ngd_G_cap = 0, ngd_part_cap = 0, ngd_u_cap = 0, ngd_work_cap = 0;
    double *ngd_J = nullptr;         // dual route: rows row-contiguous, FP64
    double *ngd_v = nullptr;         // dual route: b, J b, R y
    long ngd_J_cap = 0, ngd_v_cap = 0;
    double *comm_buf = nullptr;      // committee: running mean and M2
    long comm_cap = 0;
    int *ngd_info = nullptr;
    std::vector<double> h_coeff;
    std::vector<int> h_slot, h_nonlin;
    bool fwd_valid = false;
    // where the last collocation step left its output row: 0 nowhere (or
    // used since), 1 the output plane tl, 2 d_tout ([point][slot], the
    // fused kernels)
    int step_out = 0;
    long fwd_ver = -1;
    KProf prof;

    void mark(int c)
    {
        if (prof.used == prof.ev.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            prof.ev.push_back(e);
        }
        if (c >= 0) prof.cat.push_back(c);
        (void)hipEventRecord(prof.ev[prof.used++], stream);
    }
    int resolve()
    {
        if (prof.used == 0) return 0;
        HIPCHK(hipStreamSynchronize(stream));
        for (size_t i = 0; i + 1 < prof.used; i += 2) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, prof.ev[i], prof.ev[i + 1]));
            prof.ms[prof.cat[i / 2]] += ms;
            prof.cnt[prof.cat[i / 2]]++;
        }
        prof.used = 0;
        prof.cat.clear();
        return 0;
    }

    void release()
    {
        for (auto ev : prof.ev) (void)hipEventDestroy(ev);
        prof.ev.clear();
        auto f = [](void *&p) { if (p) (void)hipFree(p); p = nullptr; };
        f(t0); f(tl); f(p0); f(p1); f(scratch); f(hs);
        for (auto &p : tpl) f(p);
        for (auto &p : hpl) f(p);
        tpl.clear(); hpl.clear();
        dfree(partial); dfree(d_x); dfree(d_seed); dfree(d_tout); dfree(d_rsq); dfree(d_rsq_part);
        dfree(d_coeff); dfree(d_slot); dfree(d_nonlin);
        dfree(ngd_G); dfree(ngd_part); dfree(ngd_u); dfree(ngd_trace); dfree(ngd_info); dfree(ngd_work);
        dfree(ngd_J); dfree(ngd_v); dfree(comm_buf);
        if (own) (void)hipStreamDestroy(own);
        own = stream = nullptr;
    }
};

// brackets the launches of one kernel category with events when profiling
struct PScope {
    BatchWork &b;
    bool act;
    PScope(BatchWork &bb, int c) : b(bb), act(bb.prof.on && (bb.prof.only < 0 || bb.prof.only == c))
    {
        if (act) {
            if (b.prof.used >= 8192) (void)b.resolve();
            b.mark(c);
        }
    }
    ~PScope()
    {
        if (act) b.mark(-1);
    }
};

static long round_up(long v, long m) { return (v + m - 1) / m * m; }

static int g_poison = 0;

// every byte 0xFF: a NaN in FP32 and in FP64
static hipError_t poison_planes(BatchWork &b)
{
    for (auto &pb : b.poisoned)
        if (pb.first && pb.second)
            if (hipError_t e = hipMemsetAsync(pb.first, 0xFF, pb.second, b.stream)) return e;
    return hipSuccess;
}

static int valloc(void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return 0;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        g_err = std::string("hipMalloc: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? -4 : -3;
    }
    return 0;
}

// gradient GEMM split over the column axis (chunks of the K5 kernel): the
// one-block split K5 takes one chunk per workgroup it keeps resident, the
// tiled kernels theirs (grad_chunks).  any: the larger of the two, whatever
// the split-kernel switch (a work space's partials are sized once)
static void grad_split(long ncol, int M, int Kd, int esz, int prec, int *nchunk, long *chunk, bool any = false,
                       bool fused = false)
{
    grad_chunks(ncol, M, Kd, esz, nchunk, chunk);
    if (prec != PREC_BF16X6) return;
    // fused: the layer's K5 runs with its K6 (k_grad_lint_x6), which keeps
    // three workgroups a CU; any: room for whichever runs
    int wgs = fused ? grad_lint_x6_wgs(M, Kd) : grad_x6s_wgs(M, Kd, any);
    if (any) wgs = std::max(wgs, grad_lint_x6_wgs(M, Kd, true));
    if (wgs <= 0) return;
    const long nc = std::min((long)wgs, std::max(1L, ncol / 128));
    const long ch = (((ncol + nc - 1) / nc) + 127) / 128 * 128;
    if (ch <= 0) return;
    const int n = (int)((ncol + ch - 1) / ch);
    if (any && n <= *nchunk) return;
    *nchunk = n;
    *chunk = ch;
}

static int bwork_init(BatchWork &b, const Net &nt, int maxpt, int prec, double budget)
{
    b.prec = prec;
    b.maxpt = maxpt;
    b.esz = prec == PREC_FP64 ? 8 : 4;
    b.np_cap = round_up(std::max(maxpt, 1), 64);
    const int na = nt.tab.na, L = nt.nlayer;
    const size_t col = (size_t)na * b.np_cap;
    int nmax = 0, hsum = 0;
    for (int l = 0; l < L; ++l) nmax = std::max(nmax, nt.dims[l]);
    for (int l = 1; l < L - 1; ++l) hsum += nt.dims[l];
    int rc;
    // device bytes: a plane of n rows is n * col * esz
    auto plane = [&](int n) { return (double)n * col * b.esz; };
    long pc = 0;
    for (int l = 1; l < L; ++l) {
        int nc;
        long ch;
        grad_split((long)col, nt.dims[l], nt.dims[l - 1] + 1, (int)b.esz, prec, &nc, &ch, true);
        pc = std::max(pc, (long)(nc + grad_reduce_splits(nc)) * nt.dims[l] * (nt.dims[l - 1] + 1));
    }
    if (L >= 3) {
        // the first hidden layer's gradient from its adjoint kernel: one
        // partial a 64-point block
        const int nc = (int)(b.np_cap / 64);
        pc = std::max(pc, (long)(nc + grad_reduce_splits(nc)) * nt.dims[1] * (nt.dims[0] + 1));
    }
    const int nfwd = bell_fwd_threads(nt.sched.nentry, (int)b.esz);
    const double fixed = plane(nt.dims[0]) + plane(nt.dims[L - 1]) + 2 * plane(nmax) + 8.0 * pc +
                         (nfwd == 0 ? (double)nt.sched.nentry * nmax * b.np_cap * b.esz : 0.0) +
                         8.0 * ((double)b.np_cap * (nt.dims[0] + 2.0 * nt.dims[L - 1] * na) + b.np_cap / 256 + 2);
    b.bytes_stored = fixed + 2 * plane(hsum);
    b.recompute = budget > 0.0 && b.bytes_stored > budget && L > 2;
    b.bytes = b.recompute ? fixed + plane(hsum) + plane(nmax) : b.bytes_stored;
    HIPCHK(hipStreamCreateWithFlags(&b.own, hipStreamNonBlocking));
    b.stream = b.own;
    if ((rc = valloc(&b.t0, (size_t)nt.dims[0] * col * b.esz))) return rc;
    if ((rc = valloc(&b.tl, (size_t)nt.dims[L - 1] * col * b.esz))) return rc;
    if ((rc = valloc(&b.p0, (size_t)nmax * col * b.esz))) return rc;
    if ((rc = valloc(&b.p1, (size_t)nmax * col * b.esz))) return rc;
    b.tpl.assign(L, nullptr);
    b.hpl.assign(L, nullptr);
    for (int l = 1; l < L - 1; ++l) {
        if ((rc = valloc(&b.tpl[l], (size_t)nt.dims[l] * col * b.esz))) return rc;
        if (!b.recompute && (rc = valloc(&b.hpl[l], (size_t)nt.dims[l] * col * b.esz))) return rc;
    }
    if (b.recompute && (rc = valloc(&b.hs, (size_t)nmax * col * b.esz))) return rc;
    // LDS-resident operand store unless the set is too large
    if (nfwd == 0) {
        if ((rc = valloc(&b.scratch, (size_t)nt.sched.nentry * nmax * b.np_cap * b.esz))) return rc;
    }
    b.partial_cap = pc;
    if ((rc = dalloc(&b.partial, (size_t)pc))) return rc;
    if ((rc = dalloc(&b.d_x, (size_t)b.np_cap * nt.dims[0]))) return rc;
    if ((rc = dalloc(&b.d_seed, (size_t)b.np_cap * nt.dims[L - 1] * na))) return rc;
    if ((rc = dalloc(&b.d_tout, (size_t)b.np_cap * nt.dims[L - 1] * na))) return rc;
    if ((rc = dalloc(&b.d_rsq_part, (size_t)(b.np_cap / 256 + 1)))) return rc;
    if ((rc = dalloc(&b.d_rsq, 1))) return rc;
    b.poison = g_poison != 0;
    if (b.poison) {
        const size_t pl = col * b.esz;
        b.poisoned = {{b.t0, nt.dims[0] * pl}, {b.tl, nt.dims[L - 1] * pl}, {b.p0, nmax * pl}, {b.p1, nmax * pl},
                      {b.partial, sizeof(double) * (size_t)pc}};
        for (int l = 1; l < L - 1; ++l) {
            b.poisoned.push_back({b.tpl[l], nt.dims[l] * pl});
            if (b.hpl[l]) b.poisoned.push_back({b.hpl[l], nt.dims[l] * pl});
        }
        if (b.hs) b.poisoned.push_back({b.hs, nmax * pl});
        if (b.scratch) b.poisoned.push_back({b.scratch, (size_t)nt.sched.nentry * nmax * b.np_cap * b.esz});
        HIPCHK(poison_planes(b));
        HIPCHK(hipStreamSynchronize(b.stream));
    }
    return 0;
}

// ---------------------------------------------------------------------------
// the layer loops

template <typename T>
static const GenKernels *gen_for(const Net &nt)
{
    const GenKernels &g = nt.gen[sizeof(T) == 8 ? 1 : 0];
    return (codegen_on() && g.ok) ? &g : nullptr;
}

// the edge kernels a batched step of many points takes (a step of one point,
// or one writing per-point rows, drops the two gradient-forming ones):
// in -- the first hidden layer forms S from the input plane (forward and
// adjoint); in_grad -- its adjoint also forms that layer's weight gradient;
// out -- the last hidden layer forms Tbar from the output layer's Sbar;
// out_grad -- that adjoint also forms the output layer's weight gradient
struct EdgePlan {
    bool in = false, in_grad = false, out = false, out_grad = false;
};

static EdgePlan edge_plan(const Net &nt, bool fp64, bool s_kept)
{
    EdgePlan p;
    const GenKernels &g = nt.gen[fp64 ? 1 : 0];
    const int L = nt.nlayer;
    if (!s_kept || !g_edge || L < 3) return p;
    p.in = g.fwdt_in && g.bwds_i;
    const bool one_hidden = L == 3 && p.in;
    p.out = one_hidden ? g.bwds_io : g.bwds_o;
    p.out_grad = p.out && nt.dims[L - 1] == 1 && g.bwds_og && !one_hidden;
    p.in_grad = p.in && ((one_hidden && p.out) ? g.bwds_iog : g.bwds_ig);
    return p;
}

template <typename T> static const T *weights(const Net &nt);
template <> const double *weights<double>(const Net &nt) { return nt.d_w64; }
template <> const float *weights<float>(const Net &nt) { return nt.d_w32; }

// K1 of layer l from the stored T of the layer below, into the layer's
// kept plane (s_kept) or p0
template <typename T>
static void layer_linear(const Net &nt, BatchWork &b, int l, long ncol, long Np)
{
    const int L = nt.nlayer, nd = nt.dims[l], ndm = nt.dims[l - 1];
    const T *X = (const T *)(l == 1 ? b.t0 : b.tpl[l - 1]);
    T *out = (T *)(l == L - 1 ? b.tl : b.s_kept ? b.hplane(l) : b.p0);
    PScope p(b, CAT_LINEAR);
    launch_gemm_wx<T>(b.prec, false, weights<T>(nt) + nt.woff[l], ndm + 1, X, out, nd, ndm + 1, ncol, Np, b.stream);
}

// K1 of layer l, then (hidden layers) the Faa di Bruno step writing T of
// layer l and, unless the layer keeps S, h
template <typename T>
static hipError_t layer_forward(const Net &nt, BatchWork &b, int l, long ncol, long Np)
{
    const int L = nt.nlayer, nd = nt.dims[l];
    hipStream_t s = b.stream;
    layer_linear<T>(nt, b, l, ncol, Np);
    if (l < L - 1) {
        PScope p(b, CAT_BELL_FWD);
        const GenKernels *gk = gen_for<T>(nt);
        if (b.s_kept)
            return gen_launch_fwdt(nt.gen[sizeof(T) == 8 ? 1 : 0], b.hplane(l), b.tpl[l], nd, Np, nt.dt.act(), s);
        else if (gk)
            return gen_launch_fwd(*gk, b.p0, b.tpl[l], b.hplane(l), nd, Np, nt.dt.act(), s);
        else
            launch_bell_fwd<T>((const T *)b.p0, (T *)b.tpl[l], (T *)b.hplane(l), nd, Np, nt.dt.fwd(), nt.dt.act(),
                               (T *)b.scratch, s);
    }
    return hipSuccess;
}

// the one-point kernels take the work space when it holds one FP64 point,
// the layers keep S (the adjoint recomputes h), nothing is recomputed, and
// the network fits: at most PT_MAXL layers, every width at most PT_MAXW
// lanes and width x slots at most PT_MAXWA (a plane's LDS rows; wider
// layers go faster through the batched kernels anyway, which spread a layer
// over many workgroups)
template <typename T>
static const GenKernels *point_kernels(const Net &nt, const BatchWork &b)
{
    if (sizeof(T) != 8 || !g_point || b.npt != 1 || !b.s_kept || b.recompute || b.rows_next || nt.nlayer > PT_MAXL)
        return nullptr;
    const GenKernels &g = nt.gen[1];
    if (!g.ok || !g.pt_fwd || !g.pt_bwd || g.pt_scratch != 0) return nullptr;
    for (int l = 0; l < nt.nlayer; ++l)
        if (nt.dims[l] > PT_MAXW || (long)nt.dims[l] * nt.tab.na > PT_MAXWA) return nullptr;
    return &g;
}

static PtNet point_net(const Net &nt, const BatchWork &b)
{
    PtNet p{};
    const int L = nt.nlayer;
    p.L = L;
    p.Np = b.np;
    p.ldr = (long)nt.tab.na * b.np;
    for (int l = 0; l < L; ++l) {
        p.dims[l] = nt.dims[l];
        p.woff[l] = nt.woff[l];
        p.T[l] = (double *)(l == 0 ? b.t0 : (l == L - 1 ? b.tl : b.tpl[l]));
        p.S[l] = (double *)(l > 0 && l < L - 1 ? b.hplane(l) : nullptr);
    }
    return p;
}

// the forward's launches on the work space's stream; the first failed
// launch of a generated kernel is returned (the others surface through
// hipGetLastError)
template <typename T>
static hipError_t forward_t(const Net &nt, BatchWork &b, const double *d_x, int npt)
{
    b.step_out = 0;
    const int na = nt.tab.na, L = nt.nlayer;
    const long Np = round_up(std::max(npt, 1), 64);
    const long ncol = (long)na * Np;
    b.np = Np;
    b.npt = npt;
    const GenKernels *gk = gen_for<T>(nt);
    b.s_kept = g_store_s && gk && gk->s_ok();
    hipStream_t s = b.stream;
    if (const GenKernels *pk = point_kernels<T>(nt, b)) {
        b.pt_used = true;
        b.edge_in = false;
        hipError_t e;
        {
            PScope p(b, CAT_BELL_FWD);
            e = gen_launch_pt_fwd(*pk, (const double *)weights<T>(nt), d_x, nt.dt.ind_e1, point_net(nt, b),
                                  nt.dt.act(), s);
        }
        if (!b.defer_use) nt.used_on(s);
        return e;
    }
    b.pt_used = false;
    if (b.poison)
        if (hipError_t e = poison_planes(b)) return e;
    {
        PScope p(b, CAT_OTHER);
        launch_input<T>(d_x, npt, nt.dims[0], na, nt.dt.ind_e1, (T *)b.t0, Np, s);
    }
    hipError_t e = hipSuccess;
    const GenKernels &gg = nt.gen[sizeof(T) == 8 ? 1 : 0];
    b.edge_in = edge_plan(nt, sizeof(T) == 8, b.s_kept).in;
    int l0 = 1;
    if (b.edge_in) {
        // the first hidden layer straight from T0 (no K1, no S plane)
        PScope p(b, CAT_BELL_FWD);
        e = gen_launch_fwdt_in(gg, b.t0, weights<T>(nt) + nt.woff[1], b.tpl[1], nt.dims[1], Np, nt.dt.act(), s);
        l0 = 2;
    }
    for (int l = l0; l < L && e == hipSuccess; ++l) e = layer_forward<T>(nt, b, l, ncol, Np);
    if (!b.defer_use) nt.used_on(s);
    return e;
}

// whether layer l's K5 and K6 run as one kernel in a batched BF16X6 step
// (layers 2 .. L-1 whose K6 is not inside the last hidden layer's adjoint)
static bool lint_fused(const BatchWork &b, int l, int L, bool out_fused, int nd, int ndm)
{
    return b.prec == PREC_BF16X6 && l > 1 && !(l == L - 1 && out_fused) && grad_lint_x6_ok(nd, ndm + 1);
}

// adjoint from the output seed already in p0 (rows < nout).  rows_t == null:
// gradient added to grad; else the per-point rows are written to rows_t.
template <typename T>
static hipError_t backward_t(const Net &nt, BatchWork &b, double *grad, float *rows_t, long ldr)
{
    const int na = nt.tab.na, L = nt.nlayer;
    const long Np = b.np, ncol = (long)na * Np;
    hipStream_t s = b.stream;
    const T *W = weights<T>(nt);
    const DevPairs pr = nt.dt.pairs();
    if (!rows_t && b.pt_used) {
        // the forward went through the one-point kernel: so does the adjoint
        hipError_t e;
        {
            PScope p(b, CAT_BELL_BWD);
            e = gen_launch_pt_bwd(nt.gen[1], (const double *)W, grad, (const double *)b.p0, point_net(nt, b),
                                  nt.dt.act(), s);
        }
        if (!b.defer_use) nt.used_on(s);
        return e;
    }
    hipError_t e = hipSuccess;
    // sbp holds Sbar of the layer being reversed, tbp Tbar of the layer
    // below; they trade places after the last hidden layer when its Tbar is
    // formed from the output layer's Sbar (no K6 of the output layer)
    void *sbp = b.p0, *tbp = b.p1;
    const GenKernels &gg = nt.gen[sizeof(T) == 8 ? 1 : 0];
    const int nout = nt.dims[L - 1];
    const EdgePlan plan = edge_plan(nt, sizeof(T) == 8, b.s_kept);
    const bool out_fused = plan.out;
    // one output row: the last hidden layer's adjoint also forms the output
    // layer's weight gradient (T of the layer recomputed beside h)
    const bool out_grad = plan.out_grad && !rows_t && b.npt > 1;
    for (int l = L - 1; l >= 1 && e == hipSuccess; --l) {
        const int nd = nt.dims[l], ndm = nt.dims[l - 1];
        if (l < L - 1) {
            const bool in_s = l == 1 && b.edge_in;
            // p0 is free here (its S-bar was consumed by layer l+1); the
            // recomputed T of layer l equals the stored one bit for bit.
            // A layer that keeps S needs K1 alone; the first hidden layer
            // of an edge forward keeps none.
            if (b.recompute && !in_s) {
                if (b.s_kept)
                    layer_linear<T>(nt, b, l, ncol, Np);
                else if ((e = layer_forward<T>(nt, b, l, ncol, Np)) != hipSuccess)
                    break;
            }
            const GenKernels *gk = gen_for<T>(nt);
            const bool out_t = l == L - 2 && out_fused;
            if (in_s && !rows_t && b.npt > 1 && plan.in_grad) {
                // the first hidden layer's adjoint forms the layer's weight
                // gradient itself: no Sbar plane, no K5 of this layer
                const int nc = (int)(Np / 64);
                {
                    PScope p(b, CAT_BELL_BWD);
                    e = gen_launch_bwds_grad1(gg, out_t, tbp, nd, Np, nt.dt.act(), b.t0, W + nt.woff[1], sbp,
                                              W + nt.woff[L - 1], nout, b.partial, s);
                }
                if (e == hipSuccess) {
                    PScope p(b, CAT_GRAD_REDUCE);
                    launch_grad_reduce(b.partial, nc, (long)nd * (ndm + 1), grad + nt.woff[l], s);
                }
                break;   // the first hidden layer is the last one reversed
            }
            if (out_t && out_grad) {
                {
                    PScope p(b, CAT_BELL_BWD);
                    e = gen_launch_bwds_gradout(gg, b.hplane(l), tbp, nd, Np, nt.dt.act(), sbp, W + nt.woff[L - 1],
                                                b.partial, s);
                }
                if (e != hipSuccess) break;
                {
                    PScope p(b, CAT_GRAD_REDUCE);
                    launch_grad_reduce(b.partial, (int)(Np / 64), (long)nd + 1, grad + nt.woff[L - 1], s);
                }
                std::swap(sbp, tbp);
            } else {
                PScope p(b, CAT_BELL_BWD);
                if (b.s_kept) {
                    const T *W1 = W + nt.woff[1];
                    if (out_t) {
                        e = gen_launch_bwds_edge(gg, in_s, true, nullptr, b.hplane(l), tbp, nd, Np, nt.dt.act(), b.t0,
                                                 W1, sbp, W + nt.woff[L - 1], nout, s);
                        std::swap(sbp, tbp);
                    } else {
                        e = gen_launch_bwds_edge(gg, in_s, false, tbp, b.hplane(l), sbp, nd, Np, nt.dt.act(), b.t0,
                                                 W1, nullptr, nullptr, 0, s);
                    }
                } else if (gk) {
                    e = gen_launch_bwd(*gk, tbp, b.hplane(l), sbp, nd, Np, s);
                } else {
                    launch_bell_bwd<T>((const T *)tbp, (const T *)b.hplane(l), (T *)sbp, nd, Np, pr, s);
                }
            }
            if (e != hipSuccess) break;
        }
        const T *X = (const T *)(l == 1 ? b.t0 : b.tpl[l - 1]);
        bool k6_done = false;   // this layer's K6 ran inside its K5
        if (l == L - 1 && out_grad) {
            // the output layer's gradient comes with the next layer's adjoint
        } else if (rows_t) {
            PScope p(b, CAT_GRAD);
            launch_point_rows<T>((const T *)sbp, X, nd, ndm + 1, na, Np, rows_t + nt.woff[l] * ldr, ldr, s);
        } else if (b.npt == 1) {
            // one point (the per-point ABI, the Kalman filter): one thread a
            // weight instead of the batched chunked product and its reduction
            PScope p(b, CAT_GRAD);
            launch_grad_point1<T>((const T *)sbp, X, nd, ndm + 1, na, Np, grad + nt.woff[l], s);
        } else {
            int nc;
            long ch;
            const bool fuse = lint_fused(b, l, L, out_fused, nd, ndm);
            grad_split(ncol, nd, ndm + 1, (int)b.esz, b.prec, &nc, &ch, false, fuse);
            {
                PScope p(b, CAT_GRAD);
                // K5 and this layer's K6 in one pass over S-bar where that kernel applies
                if constexpr (sizeof(T) == 4)
                    if (fuse)
                        k6_done = launch_grad_lint_x6((const float *)sbp, (const float *)X,
                                                      (const float *)(W + nt.woff[l]), (float *)tbp, nd, ndm + 1,
                                                      ncol, Np, b.partial, nc, ch, s);
                if (!k6_done)
                    launch_gemm_grad<T>(b.prec, (const T *)sbp, X, nd, ndm + 1, ncol, Np, b.partial, nc, ch, s);
            }
            PScope p(b, CAT_GRAD_REDUCE);
            launch_grad_reduce(b.partial, nc, (long)nd * (ndm + 1), grad + nt.woff[l], s);
        }
        if (l > 1 && !(l == L - 1 && out_fused) && !k6_done) {
            PScope p(b, CAT_LINEAR_T);
            launch_gemm_wx<T>(b.prec, true, W + nt.woff[l], ndm + 1, (const T *)sbp, (T *)tbp, ndm, nd,
                              ncol, Np, s);
        }
    }
    if (!b.defer_use) nt.used_on(s);
    return e;
}

// ---------------------------------------------------------------------------
// registries (lib/c_api_module.f90:38-62)

static constexpr int MAXH = 64;

template <typename Obj>
struct Registry {
    Obj *obj[MAXH + 1] = {};
    int gen[MAXH + 1] = {};      // generation of the slot (networks)
    int owner[MAXH + 1] = {};    // network a dependent belongs to
    int ogen[MAXH + 1] = {};     // generation of that network when built
    int slot()
    {
        for (int i = 1; i <= MAXH; ++i)
            if (!obj[i]) return i;
        return -1;
    }
    Obj *get(int h) const { return (h >= 1 && h <= MAXH) ? obj[h] : nullptr; }
};

static Registry<Net> g_net;
static Registry<TableSet> g_tab;
static Registry<BatchWork> g_work;   // evaluation work spaces (wid)
static Registry<BatchWork> g_twork;  // training work spaces (tid)
static Registry<BatchWork> g_bwork;  // batched work spaces (bid)
static Registry<Grad> g_grad;
static TableSet g_current;
static bool g_current_ready = false;

static bool dependent_ok(const Registry<BatchWork> &r, int h, int nid)
{
    return r.get(h) && r.owner[h] == nid && r.ogen[h] == g_net.gen[nid];
}
static bool grad_ok(int gid, int nid)
{
    return g_grad.get(gid) && g_grad.owner[gid] == nid && g_grad.ogen[gid] == g_net.gen[nid];
}

static int make_net(const TableSet &ts, int nlayer, const int *dims)
{
    if (nlayer < 2 || !dims) return fail(-1, "nlayer must be >= 2");
    for (int l = 0; l < nlayer; ++l)
        if (dims[l] < 1) return fail(-1, "every width must be >= 1");
    if (dims[0] != ts.d0) return fail(-1, "input width differs from the table set's d0");
    int id = g_net.slot();
    if (id < 0) return fail(-1, "no free network handle");
    Net *nt = new Net();
    nt->nlayer = nlayer;
    nt->dims.assign(dims, dims + nlayer);
    nt->tab = ts;
    nt->sched.build(ts);
    nt->woff.assign(nlayer, 0);
    long off = 0;
    for (int l = 1; l < nlayer; ++l) {
        nt->woff[l] = off;
        off += (long)dims[l] * (dims[l - 1] + 1);
    }
    nt->nw = (int)off;
    nt->w.assign(off, 0.0);
    g_net.obj[id] = nt;
    return id;
}

static std::string c_path(const char *f) { return f ? std::string(f) : std::string(); }

// nn_weight.dat (lib/net_module.f90:146-215): epoch; "func" code;
// "Activation_out" code; nlayer; widths; per layer a header record then one
// list-directed record per neuron (bias, then the incoming weights)
static int load_weights(const std::string &path, const TableSet &ts, int *nid_out)
{
    std::ifstream in(path);
    if (!in) return fail(-2, "no such file: " + path);
    std::string line;
    auto next = [&](std::string &l) { return (bool)std::getline(in, l); };
    if (!next(line)) return fail(-1, "empty weight file");
    std::string key;
    int ifunc = -1, iout = -1, nlayer = 0;
    if (!next(line)) return fail(-1, "truncated weight file");
    { std::istringstream is(line); is >> key >> ifunc; }
    if (!next(line)) return fail(-1, "truncated weight file");
    { std::istringstream is(line); is >> key >> iout; }
    if (ifunc < 0 || ifunc > 5) return fail(-2, "weight file names an unknown activation");
    if (iout != 0) return fail(-2, "weight file has a nonlinear output layer");
    if (ifunc != ts.iact) return fail(-2, "weight file activation differs from the table set's");
    if (!next(line)) return fail(-1, "truncated weight file");
    nlayer = std::atoi(line.c_str());
    if (nlayer < 2) return fail(-1, "bad layer count");
    std::vector<int> dims(nlayer);
    for (int l = 0; l < nlayer; ++l) {
        if (!next(line)) return fail(-1, "truncated weight file");
        dims[l] = std::atoi(line.c_str());
    }
    int id = make_net(ts, nlayer, dims.data());
    if (id < 0) return id;
    Net *nt = g_net.obj[id];
    for (int l = 1; l < nlayer; ++l) {
        if (!next(line)) { delete nt; g_net.obj[id] = nullptr; return fail(-1, "truncated weight file"); }
        for (int j = 0; j < dims[l]; ++j) {
            int need = dims[l - 1] + 1, got = 0;
            double *row = &nt->w[nt->woff[l] + (long)j * need];
            while (got < need) {
                if (!next(line)) { delete nt; g_net.obj[id] = nullptr; return fail(-1, "truncated weight file"); }
                for (char &c : line)
                    if (c == 'D' || c == 'd') c = 'E';
                std::istringstream is(line);
                double v;
                while (got < need && (is >> v)) row[got++] = v;
            }
        }
    }
    nt->host_ver++;
    *nid_out = id;
    return 0;
}

static const Net *net_or_null(int nid) { return g_net.get(nid); }

template <typename T>
static int run_forward(Net &nt, BatchWork &b, int npt, const double *x, double *tout, bool dev)
{
    if (npt < 0 || npt > b.maxpt) return fail(-1, "npt outside [0, maxpt]");
    int rc = nt.ensure_device(b.prec != PREC_FP64);
    if (rc) return rc;
    const double *dx = x;
    if (!dev) {
        if (npt > 0)
            HIPCHK(hipMemcpyAsync(b.d_x, x, sizeof(double) * npt * nt.dims[0], hipMemcpyHostToDevice, b.stream));
        dx = b.d_x;
    }
    if (hipError_t e = forward_t<T>(nt, b, dx, npt)) return hip_fail(e, "forward launch");
    if (tout) {
        const int nout = nt.dims[nt.nlayer - 1], na = nt.tab.na;
        double *dt = dev ? tout : b.d_tout;
        launch_output<T>((const T *)b.tl, nout, na, npt, b.np, dt, b.stream);
        if (!dev && npt > 0)
            HIPCHK(hipMemcpyAsync(tout, b.d_tout, sizeof(double) * npt * nout * na, hipMemcpyDeviceToHost,
                                  b.stream));
    }
    HIPCHK(hipGetLastError());
    b.fwd_valid = true;
    b.fwd_ver = nt.dev_ver;
    return 0;
}

template <typename T>
static int run_backward_seed(Net &nt, BatchWork &b, const double *seed, Grad &g, bool dev)
{
    if (!b.fwd_valid || b.fwd_ver != nt.dev_ver || nt.host_ver != nt.dev_ver)
        return fail(-1, "backward needs a forward on the same work space and weights");
    const int nout = nt.dims[nt.nlayer - 1], na = nt.tab.na;
    const double *ds = seed;
    if (!dev) {
        if (b.npt > 0)
            HIPCHK(hipMemcpyAsync(b.d_seed, seed, sizeof(double) * b.npt * nout * na, hipMemcpyHostToDevice,
                                  b.stream));
        ds = b.d_seed;
    }
    {
        PScope p(b, CAT_OTHER);
        launch_seed_user<T>(ds, b.npt, nout, na, (T *)b.p0, b.np, b.stream);
    }
    int rc = g.before(b.stream);
    if (rc) return rc;
    hipError_t e = backward_t<T>(nt, b, g.d_g, nullptr, 0);
    if ((rc = g.after(b.stream))) return rc;
    if (e) return hip_fail(e, "adjoint launch");
    HIPCHK(hipGetLastError());
    return 0;
}

// value-only batched step (lib/train_module.f90:525-631, net_grad_batch):
// the work space's planes are used with one slot (row stride Np); Zd of
// hidden layer l sits after Z in the same plane.  Gradient added to g.
template <typename T>
static int run_grad_value(Net &nt, BatchWork &b, int npt, const double *x, const double *y, double coef, Grad &g,
                          bool dev)
{
    const int L = nt.nlayer;
    if (npt < 1 || npt > b.maxpt) return fail(-1, "npt outside [1, maxpt]");
    if (nt.dims[L - 1] != 1) return fail(-1, "the value path is written for a single output");
    int rc = nt.ensure_device(b.prec != PREC_FP64);
    if (rc) return rc;
    const long Np = round_up(npt, 64);
    hipStream_t s = b.stream;
    const double *dx = x, *dy = y;
    if (!dev) {
        HIPCHK(hipMemcpyAsync(b.d_x, x, sizeof(double) * npt * nt.dims[0], hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(b.d_seed, y, sizeof(double) * npt, hipMemcpyHostToDevice, s));
        dx = b.d_x;
        dy = b.d_seed;
    }
    const T *W = weights<T>(nt);
    const ActParams act = nt.dt.act();
    auto Z = [&](int l) { return (T *)(l == 0 ? b.t0 : (l == L - 1 ? b.tl : b.tpl[l])); };
    auto Zd = [&](int l) { return (T *)b.tpl[l] + (long)nt.dims[l] * Np; };
    {
        PScope p(b, CAT_OTHER);
        launch_input<T>(dx, npt, nt.dims[0], 1, nt.dt.ind_e1, Z(0), Np, s);
    }
    for (int l = 1; l < L; ++l) {
        const int nd = nt.dims[l], ndm = nt.dims[l - 1];
        T *out = l < L - 1 ? (T *)b.p0 : Z(l);
        {
            PScope p(b, CAT_LINEAR);
            launch_gemm_wx<T>(b.prec, false, W + nt.woff[l], ndm + 1, Z(l - 1), out, nd, ndm + 1, Np, Np, s);
        }
        if (l < L - 1) {
            PScope p(b, CAT_BELL_FWD);
            launch_act_value<T>(out, Z(l), Zd(l), (long)nd * Np, act, s);
        }
    }
    {
        PScope p(b, CAT_OTHER);
        launch_seed_value<T>(Z(L - 1), dy, coef, npt, Np, (T *)b.p0, s);
    }
    if ((rc = g.before(s))) return rc;
    for (int l = L - 1; l >= 1; --l) {
        const int nd = nt.dims[l], ndm = nt.dims[l - 1];
        if (l < L - 1) {
            PScope p(b, CAT_BELL_BWD);
            launch_mul_value<T>((const T *)b.p1, Zd(l), (T *)b.p0, (long)nd * Np, s);
        }
        int nc;
        long ch;
        grad_split(Np, nd, ndm + 1, (int)b.esz, b.prec, &nc, &ch);
        {
            PScope p(b, CAT_GRAD);
            launch_gemm_grad<T>(b.prec, (const T *)b.p0, Z(l - 1), nd, ndm + 1, Np, Np, b.partial, nc, ch, s);
        }
        {
            PScope p(b, CAT_GRAD_REDUCE);
            launch_grad_reduce(b.partial, nc, (long)nd * (ndm + 1), g.d_g + nt.woff[l], s);
        }
        if (l > 1) {
            PScope p(b, CAT_LINEAR_T);
            launch_gemm_wx<T>(b.prec, true, W + nt.woff[l], ndm + 1, (const T *)b.p0, (T *)b.p1, ndm, nd, Np, Np, s);
        }
    }
    nt.used_on(s);
    if ((rc = g.after(s))) return rc;
    b.fwd_valid = false;   // the planes now hold the one-slot layout
    HIPCHK(hipGetLastError());
    return 0;
}

// residual terms are host arrays (they describe the operator); the device
// copy is refreshed only when they change, so repeated steps never block
static int set_residual(BatchWork &b, int nterm, const double *coeff, const int *slot,
                        const int *nonlin, int na, DevResidual *r)
{
    if (nterm < 0) return fail(-1, "nterm < 0");
    std::vector<double> c(nterm);
    std::vector<int> s(nterm), nl(nterm);
    for (int i = 0; i < nterm; ++i) {
        c[i] = coeff[i];
        if (slot[i] < 1 || slot[i] > na) return fail(-1, "residual slot outside [1, nderiv]");
        s[i] = slot[i] - 1;
        nl[i] = (nonlin && nonlin[i]) ? 1 : 0;
    }
    if (c != b.h_coeff || s != b.h_slot || nl != b.h_nonlin) {
        if (nterm > b.nterm_cap) {
            dfree(b.d_coeff); dfree(b.d_slot); dfree(b.d_nonlin);
            int rc;
            if ((rc = dalloc(&b.d_coeff, nterm)) || (rc = dalloc(&b.d_slot, nterm)) ||
                (rc = dalloc(&b.d_nonlin, nterm)))
                return rc;
            b.nterm_cap = nterm;
        }
        if (nterm > 0) {
            HIPCHK(hipStreamSynchronize(b.stream));
            HIPCHK(hipMemcpy(b.d_coeff, c.data(), sizeof(double) * nterm, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(b.d_slot, s.data(), sizeof(int) * nterm, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(b.d_nonlin, nl.data(), sizeof(int) * nterm, hipMemcpyHostToDevice));
        }
        b.h_coeff = c;
        b.h_slot = s;
        b.h_nonlin = nl;
    }
    r->nterm = nterm;
    r->coeff = b.d_coeff;
    r->slot = b.d_slot;
    r->nonlin = b.d_nonlin;
    std::vector<int> sorted(s);
    std::sort(sorted.begin(), sorted.end());
    r->distinct = (std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end() &&
                   (sorted.empty() || sorted[0] > 0)) ? 1 : 0;
    return 0;
}

// the fused layer kernels take a collocation step (not per-point rows) of an
// FP32 work space that keeps its planes (no per-layer recompute: h of each
// hidden layer is kept) when the network has at least two hidden layers (at
// most RS_MAX layers in all: one reduction segment a layer), every one FZ_W
// wide, one output row, at most FZ_MAXD0 inputs, and the family's fused
// module compiled (codegen.cpp)
static const FusedKernels *fused_plan(Net &nt, const BatchWork &b)
{
    const int L = nt.nlayer;
    if (!g_fused_layers || !codegen_on() || b.prec != PREC_FP32 || b.recompute || L < 4 || L > RS_MAX)
        return nullptr;
    if (nt.dims[L - 1] != 1 || nt.dims[0] > FZ_MAXD0) return nullptr;
    for (int l = 1; l < L - 1; ++l)
        if (nt.dims[l] != FZ_W) return nullptr;
    return nt.fused();
}

// grow the gradient partials to n doubles (the work space's stream drained first)
static int grow_partial(BatchWork &b, long n)
{
    if (n <= b.partial_cap) return 0;
    HIPCHK(hipStreamSynchronize(b.stream));
    for (auto &pb : b.poisoned)
        if (pb.first == b.partial) pb = {nullptr, 0};
    dfree(b.partial);
    b.partial_cap = 0;
    int rc = dalloc(&b.partial, (size_t)n);
    if (rc) return rc;
    b.partial_cap = n;
    if (b.poison) {
        b.poisoned.push_back({b.partial, sizeof(double) * (size_t)n});
        HIPCHK(hipMemsetAsync(b.partial, 0xFF, sizeof(double) * (size_t)n, b.stream));
    }
    return 0;
}

// The collocation step on the fused layer kernels.  The work space's planes
// hold the point-blocked layout ([block][slot][4 points][neuron]): T of the
// hidden layers but the last (tpl), h of the hidden layers but the first
// (hpl), the seed and the T-bar planes the adjoint passes down (p0, p1).
// Forward: the first hidden layer from the inputs, then one kernel a layer
// (K1 + Faa di Bruno, T and h); the last one also forms the output row, the
// residual, the seed, the output layer's gradient partials and the sums of
// R^2.  Then per hidden layer from the last: one kernel (h read; pair sums,
// K5, T-bar of the layer below), the first hidden layer's forming S and h
// from the inputs and its gradient.  The planes are left in that layout: no
// backward_batch may follow.
static int run_residual_fused(Net &nt, BatchWork &b, const FusedKernels &f, int npt, const double *x,
                              const DevResidual &r, double fac, Grad &g, double *rsq, bool dev)
{
    if (npt < 0 || npt > b.maxpt) return fail(-1, "npt outside [0, maxpt]");
    int rc = nt.ensure_device(true);
    if (rc) return rc;
    const int L = nt.nlayer, na = nt.tab.na, d0 = nt.dims[0];
    hipStream_t s = b.stream;
    const double *dx = x;
    if (!dev) {
        if (npt > 0) HIPCHK(hipMemcpyAsync(b.d_x, x, sizeof(double) * npt * d0, hipMemcpyHostToDevice, s));
        dx = b.d_x;
    }
    const long Np = round_up(std::max(npt, 1), 64);
    const int nblk = (int)(Np / FZ_P);
    const int grid = std::min(nblk, f.grid);
    auto H = [&](int l) { return (float *)b.hpl[l]; };
    b.np = Np;
    b.npt = npt;
    b.pt_used = false;
    b.fwd_valid = false;
    // partials, each layer's in its own range, all reduced by one launch
    // after the adjoint: the output layer's per block (nblk x 65), the
    // blocks' sums of R^2 (nblk), then the adjoint kernels' of layers L-2 ..
    // 2 (grid x 64 x 65 each) and of layer 1 (grid x 64 x (d0 + 1))
    const long hid = (long)grid * FZ_W * (FZ_W + 1);
    if ((rc = grow_partial(b, (long)nblk * (FZ_W + 2) + (L - 3) * hid + (long)grid * FZ_W * (d0 + 1)))) return rc;
    if (b.poison)
        if (hipError_t e = poison_planes(b)) return hip_fail(e, "poison");
    const float *W = nt.d_w32;
    const ActParams act = nt.dt.act();
    if ((rc = g.before(s))) return rc;
    hipError_t e;
    {
        PScope p(b, CAT_BELL_FWD);
        e = fz_launch_fwd_in(f, dx, npt, W + nt.woff[1], (float *)b.tpl[1], nblk, act, s);
    }
    for (int l = 2; l < L - 2 && e == hipSuccess; ++l) {
        PScope p(b, CAT_BELL_FWD);
        e = fz_launch_fwd(f, (const float *)b.tpl[l - 1], W + nt.woff[l], (float *)b.tpl[l], H(l), nblk, act, s);
    }
    double *gpart = b.partial, *rpart = b.partial + (long)nblk * (FZ_W + 1);
    if (e == hipSuccess) {
        PScope p(b, CAT_BELL_FWD);
        e = fz_launch_fwd_last(f, (const float *)b.tpl[L - 3], W + nt.woff[L - 2], H(L - 2), nblk, act,
                               W + nt.woff[L - 1], npt, r, fac, (float *)b.p0, gpart, rpart, b.d_tout, s);
    }
    ReduceSegs red;
    red.add(gpart, nblk, FZ_W + 1, g.d_g + nt.woff[L - 1]);
    if (rsq) red.add(rpart, nblk, 1, b.d_rsq, true);
    double *part = rpart + nblk;
    void *tb = b.p0, *tbo = b.p1;
    for (int l = L - 2; l >= 2 && e == hipSuccess; --l, part += hid) {
        const bool last = l == L - 2;
        PScope p(b, CAT_BELL_BWD);
        e = fz_launch_bwd(f, last, (const float *)b.tpl[l - 1], (const float *)tb, H(l), W + nt.woff[l],
                          W + nt.woff[L - 1], (float *)tbo, part, nblk, act, grid, s);
        red.add(part, grid, (long)FZ_W * (FZ_W + 1), g.d_g + nt.woff[l]);
        std::swap(tb, tbo);
    }
    if (e == hipSuccess) {
        {
            PScope p(b, CAT_BELL_BWD);
            e = fz_launch_bwd_in(f, dx, npt, (const float *)tb, W + nt.woff[1], part, nblk, act, grid, s);
        }
        red.add(part, grid, (long)FZ_W * (d0 + 1), g.d_g + nt.woff[1]);
        if (e == hipSuccess) {
            PScope p(b, CAT_GRAD_REDUCE);
            launch_reduce_segs(red, s);
        }
    }
    if ((rc = g.after(s))) return rc;
    if (e) return hip_fail(e, "fused step launch");
    b.step_out = 2;
    if (!b.defer_use) nt.used_on(s);
    if (rsq) {
        HIPCHK(hipMemcpyAsync(rsq, b.d_rsq, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    HIPCHK(hipGetLastError());
    return 0;
}

template <typename T>
static int run_residual(Net &nt, BatchWork &b, int npt, const double *x, int nterm,
                        const double *coeff, const int *slot, const int *nonlin, double fac,
                        Grad *g, double *rsq, float *rows_t, long ldr, bool dev)
{
    const int na = nt.tab.na, nout = nt.dims[nt.nlayer - 1];
    DevResidual r;
    int rc = set_residual(b, nterm, coeff, slot, nonlin, na, &r);
    if (rc) return rc;
    if constexpr (sizeof(T) == 4)
        if (!rows_t && g && r.nterm <= FZ_MAXT)
            if (const FusedKernels *f = fused_plan(nt, b)) return run_residual_fused(nt, b, *f, npt, x, r, fac, *g, rsq, dev);
    b.rows_next = rows_t != nullptr;
    rc = run_forward<T>(nt, b, npt, x, nullptr, dev);
    b.rows_next = false;
    if (rc) return rc;
    const int nblk = (int)((b.np + 255) / 256);
    // rows: the seed is dR/dT, set_pinn_seed(t, R=1, fac=1) (sgd_batch_module.f90:742)
    {
        PScope p(b, CAT_OTHER);
        launch_seed_residual<T>((const T *)b.tl, nout, na, npt, b.np, r, rows_t ? 1.0 : fac, rows_t != nullptr,
                                (T *)b.p0, b.d_rsq_part, b.stream);
    }
    hipError_t e;
    if (rows_t) {
        e = backward_t<T>(nt, b, nullptr, rows_t, ldr);
    } else {
        if ((rc = g->before(b.stream))) return rc;
        e = backward_t<T>(nt, b, g->d_g, nullptr, 0);
        if ((rc = g->after(b.stream))) return rc;
    }
    if (e) return hip_fail(e, "adjoint launch");
    b.step_out = b.pt_used ? 0 : 1;
    if (rsq) {
        launch_reduce_sum(b.d_rsq_part, nblk, b.d_rsq, b.stream);
        HIPCHK(hipMemcpyAsync(rsq, b.d_rsq, sizeof(double), hipMemcpyDeviceToHost, b.stream));
        HIPCHK(hipStreamSynchronize(b.stream));
    }
    HIPCHK(hipGetLastError());
    return 0;
}


// committee buffers: running mean and M2 of len values each
static int grow_comm(BatchWork &b, long len)
{
    if (2 * len <= b.comm_cap) return 0;
    HIPCHK(hipStreamSynchronize(b.stream));
    dfree(b.comm_buf);
    b.comm_cap = 0;
    int rc = dalloc(&b.comm_buf, (size_t)(2 * len));
    if (!rc) b.comm_cap = 2 * len;
    return rc;
}

// ---------------------------------------------------------------------------
// extended Kalman filter (lib/kalman_module.f90).  The filter's row is the
// gradient vector in this library's weight order, which is the reference's
// flatten order l = 2..nlayer, neuron, bias then fan-in (:378-386).

struct KFilter {
    int nw = 0, decoupled = 0;
    int nt = 0;                       // dense: tiles per side
    int nb = 0, maxb = 0;             // decoupled: blocks, largest block
    long plen = 0;                    // doubles held for P
    double *P = nullptr;
    std::vector<int> bsz, boff;
    std::vector<long> poff;
    long *d_poff = nullptr;
    int *d_bsz = nullptr, *d_boff = nullptr;
    double *pj = nullptr, *kv = nullptr, *part = nullptr, *g = nullptr, *wprior = nullptr, *r = nullptr;
    double *d_x = nullptr, *d_seed = nullptr, *d_r0 = nullptr;
    long x_cap = 0, seed_cap = 0, r0_cap = 0;
    KfScalars *st = nullptr;
    bool prior = false;
    hipError_t lerr = hipSuccess;     // first failed launch of the current call
    void release()
    {
        dfree(P); dfree(d_poff); dfree(d_bsz); dfree(d_boff); dfree(pj); dfree(kv); dfree(part); dfree(g);
        dfree(wprior); dfree(r); dfree(d_x); dfree(d_seed); dfree(d_r0); dfree(st);
    }
};

static Registry<KFilter> g_kf;

static bool kf_ok(int kid, int nid)
{
    return g_kf.get(kid) && g_kf.owner[kid] == nid && g_kf.ogen[kid] == g_net.gen[nid];
}

// grow a device buffer to n doubles
static int kf_grow(double **p, long *cap, long n)
{
    if (n <= *cap) return 0;
    dfree(*p);
    *cap = 0;
    int rc = dalloc(p, (size_t)std::max(1L, n));
    if (!rc) *cap = n;
    return rc;
}

// kf_init (kalman_module.f90:107-167, kf_block_map :171-185)
static int kf_build(const Net &nt, KFilter &kf, double p0, double lam_ini, double lam0, int decoupled)
{
    kf.nw = nt.nw;
    kf.decoupled = decoupled != 0;
    int rc;
    const long padw = kf.decoupled ? kf.nw : (long)round_up(std::max(kf.nw, 1), 64);
    if (!kf.decoupled) {
        kf.nt = (int)(padw / 64);
        kf.plen = (long)kf.nt * (kf.nt + 1) / 2 * 64 * 64;
        if ((rc = dalloc(&kf.P, (size_t)kf.plen)) || (rc = dalloc(&kf.part, (size_t)kf.nt * padw))) return rc;
        launch_kf_init_tiles(kf.P, kf.nt, kf.nw, p0, 0);
    } else {
        long m = 0, po = 0;
        for (int l = 1; l < nt.nlayer; ++l)
            for (int j = 0; j < nt.dims[l]; ++j) {
                const int n = nt.dims[l - 1] + 1;
                kf.bsz.push_back(n);
                kf.boff.push_back((int)m);
                kf.poff.push_back(po);
                m += n;
                po += (long)n * n;
                kf.maxb = std::max(kf.maxb, n);
            }
        kf.nb = (int)kf.bsz.size();
        kf.plen = po;
        if (kf.maxb > 8192) return fail(-1, "decoupled filter blocks are limited to fan-in 8191");
        if ((rc = dalloc(&kf.P, (size_t)kf.plen)) || (rc = upload(&kf.d_poff, kf.poff)) ||
            (rc = upload(&kf.d_bsz, kf.bsz)) || (rc = upload(&kf.d_boff, kf.boff)))
            return rc;
        launch_kf_init_blocks(kf.P, kf.d_poff, kf.d_bsz, kf.nb, p0, 0);
    }
    if ((rc = dalloc(&kf.pj, (size_t)padw)) || (rc = dalloc(&kf.kv, (size_t)padw)) ||
        (rc = dalloc(&kf.g, (size_t)std::max(1, kf.nw))) || (rc = dalloc(&kf.r, 1)) || (rc = dalloc(&kf.st, 1)))
        return rc;
    launch_fill(kf.pj, padw, 0.0, 0);
    launch_fill(kf.kv, padw, 0.0, 0);
    KfScalars h{};
    h.lam = lam_ini;
    h.lam0 = lam0;
    h.gs = 1.0;
    HIPCHK(hipMemcpy(kf.st, &h, sizeof(h), hipMemcpyHostToDevice));
    HIPCHK(hipStreamSynchronize(nullptr));   // the fills above, on the null stream
    return 0;
}

// the pass over P: the pending downdate, then (g) pj = P (scale g)
static void kf_pass(KFilter &kf, const double *g, double scale, hipStream_t s)
{
    if (!kf.decoupled) {
        launch_kf_tiles(kf.P, kf.nt, kf.nw, g, scale, kf.kv, kf.pj, kf.st, g != nullptr, kf.part, s);
        if (g) launch_kf_sum_part(kf.part, kf.nt, kf.pj, s);
    } else {
        launch_kf_blocks(kf.P, kf.d_poff, kf.d_bsz, kf.d_boff, kf.nb, kf.maxb, g, scale, kf.kv, kf.pj, kf.st,
                         g != nullptr, s);
    }
}

// one update from the row g (kf_update_grad :353-493, kf_iekf_iter :532-648);
// b (optional) times the kernels when its profiling is on
static void kf_step(Net &nt, KFilter &kf, const double *g, double scale, int xmode, double v, const double *seed,
                    const double *T, int nsl, bool iekf, bool final, hipStream_t s, BatchWork *b = nullptr)
{
    if (b && b->prof.on) {
        // the pass category brackets the one covariance kernel
        {
            PScope p(*b, CAT_KF_PASS);
            if (!kf.decoupled)
                launch_kf_tiles(kf.P, kf.nt, kf.nw, g, scale, kf.kv, kf.pj, kf.st, true, kf.part, s);
            else
                launch_kf_blocks(kf.P, kf.d_poff, kf.d_bsz, kf.d_boff, kf.nb, kf.maxb, g, scale, kf.kv, kf.pj,
                                 kf.st, true, s);
        }
        PScope p(*b, CAT_KF_GAIN);
        if (!kf.decoupled) launch_kf_sum_part(kf.part, kf.nt, kf.pj, s);
        launch_kf_gain(kf.st, g, scale, kf.pj, kf.nw, xmode, v, seed, T, nsl, kf.r, kf.wprior, nt.d_w64, final, s);
        launch_kf_apply(kf.st, kf.pj, kf.kv, nt.d_w64, kf.wprior, kf.nw, iekf, s);
        return;
    }
    kf_pass(kf, g, scale, s);
    launch_kf_gain(kf.st, g, scale, kf.pj, kf.nw, xmode, v, seed, T, nsl, kf.r, kf.wprior, nt.d_w64, final, s);
    launch_kf_apply(kf.st, kf.pj, kf.kv, nt.d_w64, kf.wprior, kf.nw, iekf, s);
}

// one observation at the device point dx: the row dy/dw from the seed (or
// the collocation residual's dR/dT, res), then the update
static void kf_observe(Net &nt, BatchWork &b, KFilter &kf, const double *dx, const double *dseed,
                       const DevResidual *res, const double *r0, int xmode, double v, double scale, bool iekf,
                       bool final)
{
    hipStream_t s = b.stream;
    const int nout = nt.dims[nt.nlayer - 1], na = nt.tab.na;
    {
        PScope p(b, CAT_OTHER);
        launch_fill(kf.g, kf.nw, 0.0, s);
    }
    // kf_finish drains the stream after the observations: no reader events
    b.defer_use = true;
    hipError_t e = forward_t<double>(nt, b, dx, 1);
    if (e && !kf.lerr) kf.lerr = e;
    {
        PScope p(b, CAT_OTHER);
        if (xmode == KF_XI_LINEAR) launch_output<double>((const double *)b.tl, nout, na, 1, b.np, b.d_tout, s);
        if (res)
            launch_seed_residual<double>((const double *)b.tl, nout, na, 1, b.np, *res, 1.0, true, (double *)b.p0,
                                         b.d_rsq_part, s, r0, kf.r);
        else
            launch_seed_user<double>(dseed, 1, nout, na, (double *)b.p0, b.np, s);
    }
    e = backward_t<double>(nt, b, kf.g, nullptr, 0);
    b.defer_use = false;
    if (e && !kf.lerr) kf.lerr = e;
    kf_step(nt, kf, kf.g, scale, xmode, v, dseed, b.d_tout, nout * na, iekf, final, s, &b);
}

// after the updates: weights back to the host copy, the stream drained and
// the filter's state checked
static int kf_finish(Net &nt, KFilter &kf, hipStream_t s)
{
    KfScalars h;
    HIPCHK(hipMemcpyAsync(&h, kf.st, sizeof(h), hipMemcpyDeviceToHost, s));
    {
        std::lock_guard<std::mutex> lk(nt.mu);
        HIPCHK(hipMemcpyAsync(nt.w.data(), nt.d_w64, sizeof(double) * nt.nw, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        nt.host_ver++;
        nt.dev_ver = nt.host_ver;   // the device copy is the updated one
    }
    HIPCHK(hipGetLastError());
    if (kf.lerr) {
        const hipError_t e = kf.lerr;
        kf.lerr = hipSuccess;
        return hip_fail(e, "filter launch");
    }
    if (h.bad)
        return fail(-5, "nonpositive innovation denominator: the covariance has lost positive definiteness");
    return 0;
}

// what an observation sequence needs: an FP64 work space, the filter's net,
// a sound filter, the points (and seeds) on the device
static int kf_begin(Net &nt, BatchWork &b, KFilter &kf, int nobs, const double *x, const double *seed,
                    bool dev, const double **dx, const double **dseed)
{
    if (b.prec != PREC_FP64) return fail(-1, "the filter runs on an FP64 work space");
    if (nobs < 0 || (nobs > 0 && !x)) return fail(-1, "bad observation arguments");
    int rc = nt.ensure_device(false);
    if (rc) return rc;
    KfScalars h;
    HIPCHK(hipMemcpy(&h, kf.st, sizeof(h), hipMemcpyDeviceToHost));
    if (h.bad) return fail(-5, "the filter met a nonpositive denominator earlier; create it again");
    const int d0 = nt.dims[0], nsl = nt.dims[nt.nlayer - 1] * nt.tab.na;
    *dx = x;
    if (dseed) *dseed = seed;
    if (!dev && nobs > 0) {
        if ((rc = kf_grow(&kf.d_x, &kf.x_cap, (long)nobs * d0))) return rc;
        HIPCHK(hipMemcpyAsync(kf.d_x, x, sizeof(double) * nobs * d0, hipMemcpyHostToDevice, b.stream));
        *dx = kf.d_x;
        if (seed && dseed) {
            if ((rc = kf_grow(&kf.d_seed, &kf.seed_cap, (long)nobs * nsl))) return rc;
            HIPCHK(hipMemcpyAsync(kf.d_seed, seed, sizeof(double) * nobs * nsl, hipMemcpyHostToDevice, b.stream));
            *dseed = kf.d_seed;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// one-point gradient without allocation (dnnf90_grad_point_noalloc): the
// planes of the generated one-point kernels, packed with slot stride 1
// (every hidden layer's T and S, the input and output T, the seed plane),
// live in a fixed device array.  Each GPU holds its own copy until the
// process ends; its address is taken once per device with
// hipGetSymbolAddress and used for every access from the host.  Every call
// on one GPU shares it, so calls that overlap on one GPU race.
static constexpr long NOALLOC_DOUBLES = 49152;   // 384 KiB per device
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
static constexpr int NOALLOC_MAXDEV = 64;
static double *g_scratch_ptr[NOALLOC_MAXDEV] = {};
static std::mutex g_scratch_mu;

// the current device's copy of scratch (its address taken on the first call)
static int scratch_of_device(double **out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAXDEV) return fail(-1, "device index above the no-allocation table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// seed plane [nout][na]: output 1 takes the seed, the others zero (as
// dnnf90_grad_point seeds output 1 only)
__global__ void k_noalloc_seed(const double *__restrict__ seed, int nout, int na, double *__restrict__ sb)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < nout * na) sb[u] = u < na ? seed[u] : 0.0;
}

}  // namespace dnnf90

// ===========================================================================
// C ABI
using namespace dnnf90;

extern "C" {

const char *dnnf90_last_error(void) { return g_err.c_str(); }

static bool table_args_ok(int d0, int k, int nseed, const int *seeds)
{
    if (d0 < 1 || k < 1 || k > 15 || nseed < 0) return false;
    if (nseed > 0 && !seeds) return false;
    for (int i = 0; i < d0 * nseed; ++i)
        if (seeds[i] < 0 || seeds[i] > k) return false;
    return true;
}

static int build_current(int d0, int k, int nseed, const int *seeds)
{
    if (!table_args_ok(d0, k, nseed, seeds)) return fail(-1, "bad table arguments");
    TableSet t;
    std::string err;
    if (!t.build(d0, k, nseed, seeds, &err)) return fail(-1, err);
    g_current = std::move(t);
    g_current_ready = true;
    return 0;
}

int dnnf90_tables_init_dense(int d0, int k) { return build_current(d0, k, 0, nullptr); }

int dnnf90_tables_init_closure(int d0, int k, int nseed, const int *seeds)
{
    return build_current(d0, k, nseed, seeds);
}

int dnnf90_nderiv(void) { return g_current_ready ? g_current.na : 0; }

int dnnf90_alpha(int ia, int *a)
{
    if (!g_current_ready || ia < 1 || ia > g_current.na || !a) return fail(-1, "slot outside [1, nderiv]");
    for (int v = 0; v < g_current.d0; ++v) a[v] = g_current.alpha[(size_t)(ia - 1) * g_current.d0 + v];
    return 0;
}

static int new_tabset(int d0, int k, int nseed, const int *seeds)
{
    if (!table_args_ok(d0, k, nseed, seeds)) return fail(-1, "bad table arguments");
    int id = g_tab.slot();
    if (id < 0) return fail(-1, "no free table-set handle");
    int rc = build_current(d0, k, nseed, seeds);
    if (rc) return rc;
    g_tab.obj[id] = new TableSet(g_current);
    return id;
}

int dnnf90_tabset_dense(int d0, int k) { return new_tabset(d0, k, 0, nullptr); }

int dnnf90_tabset_closure(int d0, int k, int nseed, const int *seeds)
{
    return new_tabset(d0, k, nseed, seeds);
}

int dnnf90_tabset_nderiv(int tid)
{
    TableSet *t = g_tab.get(tid);
    return t ? t->na : fail(-1, "bad table-set handle");
}

int dnnf90_tabset_activation(int tid, int iact)
{
    TableSet *t = g_tab.get(tid);
    if (!t) return fail(-1, "bad table-set handle");
    if (iact < 0 || iact > 5) return fail(-2, "activation code outside 0..5");
    t->iact = iact;
    return 0;
}

int dnnf90_net_create(int nlayer, const int *dims)
{
    if (!g_current_ready) return fail(-1, "build the tables first");
    TableSet ts = g_current;
    ts.iact = ACT_TANH;
    return make_net(ts, nlayer, dims);
}

int dnnf90_net_create_ts(int tid, int nlayer, const int *dims)
{
    TableSet *t = g_tab.get(tid);
    if (!t) return fail(-1, "bad table-set handle");
    return make_net(*t, nlayer, dims);
}

int dnnf90_net_load(const char *filename)
{
    std::string p = c_path(filename);
    {
        std::ifstream f(p);
        if (!f) return fail(-2, "no such file: " + p);
    }
    if (!g_current_ready) return fail(-1, "build the tables first");
    TableSet ts = g_current;
    ts.iact = ACT_TANH;
    int id;
    int rc = load_weights(p, ts, &id);
    return rc ? rc : id;
}

int dnnf90_net_load_ts(int tid, const char *filename)
{
    TableSet *t = g_tab.get(tid);
    if (!t) return fail(-1, "bad table-set handle");
    std::string p = c_path(filename);
    {
        std::ifstream f(p);
        if (!f) return fail(-2, "no such file: " + p);
    }
    int id;
    int rc = load_weights(p, *t, &id);
    return rc ? rc : id;
}

int dnnf90_net_set_layer(int nid, int l, const double *wrow)
{
    Net *nt = g_net.get(nid);
    if (!nt || l < 2 || l > nt->nlayer || !wrow) return fail(-1, "bad network handle or layer");
    std::lock_guard<std::mutex> lk(nt->mu);
    const long n = (long)nt->dims[l - 1] * (nt->dims[l - 2] + 1);
    std::memcpy(&nt->w[nt->woff[l - 1]], wrow, sizeof(double) * n);
    nt->host_ver++;
    return 0;
}

int dnnf90_net_free(int nid)
{
    Net *nt = g_net.get(nid);
    if (!nt) return fail(-1, "bad network handle");
    nt->release();
    delete nt;
    g_net.obj[nid] = nullptr;
    g_net.gen[nid]++;
    return 0;
}

static int make_work(Registry<BatchWork> &reg, int nid, int maxpt, int prec, double budget = 0.0)
{
    Net *nt = g_net.get(nid);
    if (!nt) return fail(-1, "bad network handle");
    if (maxpt < 1 || prec < 0 || prec > 3) return fail(-1, "maxpt >= 1 and prec in 0..3");
    int id = reg.slot();
    if (id < 0) return fail(-1, "no free work-space handle");
    int rc = nt->ensure_device(prec != PREC_FP64);
    if (rc) return rc;
    BatchWork *b = new BatchWork();
    rc = bwork_init(*b, *nt, maxpt, prec, budget);
    if (rc) {
        b->release();
        delete b;
        return rc;
    }
    reg.obj[id] = b;
    reg.owner[id] = nid;
    reg.ogen[id] = g_net.gen[nid];
    return id;
}

static int free_work(Registry<BatchWork> &reg, int id)
{
    BatchWork *b = reg.get(id);
    if (!b) return fail(-1, "bad work-space handle");
    (void)hipStreamSynchronize(b->stream);
    if (const Net *nt = g_net.get(reg.owner[id]))
        if (reg.ogen[id] == g_net.gen[reg.owner[id]]) nt->forget(b->own);
    b->release();
    delete b;
    reg.obj[id] = nullptr;
    return 0;
}

int dnnf90_work_create(int nid) { return make_work(g_work, nid, 1, PREC_FP64); }
int dnnf90_work_free(int wid) { return free_work(g_work, wid); }

int dnnf90_eval(int nid, int wid, const double *x, double *t)
{
    Net *nt = g_net.get(nid);
    if (!nt || !dependent_ok(g_work, wid, nid) || !x || !t) return fail(-1, "bad handle or argument");
    BatchWork &b = *g_work.obj[wid];
    const int na = nt->tab.na, nout = nt->dims[nt->nlayer - 1];
    std::vector<double> tout((size_t)nout * na);
    b.defer_use = true;   // the stream is drained below: no reader event
    int rc = run_forward<double>(*nt, b, 1, x, tout.data(), false);
    b.defer_use = false;
    if (rc) {
        nt->used_on(b.stream);   // not drained on this path
        return rc;
    }
    HIPCHK(hipStreamSynchronize(b.stream));
    std::memcpy(t, tout.data(), sizeof(double) * na);  // output 1
    return 0;
}

int dnnf90_twork_create(int nid) { return make_work(g_twork, nid, 1, PREC_FP64); }
int dnnf90_twork_free(int tid) { return free_work(g_twork, tid); }

int dnnf90_grad_create(int nid)
{
    Net *nt = g_net.get(nid);
    if (!nt) return fail(-1, "bad network handle");
    int id = g_grad.slot();
    if (id < 0) return fail(-1, "no free accumulator handle");
    Grad *g = new Grad();
    g->nw = nt->nw;
    int rc = g->init();
    if (!rc) rc = dalloc(&g->d_g, (size_t)std::max(1, nt->nw));
    if (rc) { g->release(); delete g; return rc; }
    launch_fill(g->d_g, nt->nw, 0.0, g->s);
    if (hipStreamSynchronize(g->s) != hipSuccess) { g->release(); delete g; return fail(-3, "device failure"); }
    g_grad.obj[id] = g;
    g_grad.owner[id] = nid;
    g_grad.ogen[id] = g_net.gen[nid];
    return id;
}

int dnnf90_grad_free(int gid)
{
    Grad *g = g_grad.get(gid);
    if (!g) return fail(-1, "bad accumulator handle");
    g->release();   // after the accumulator's last writer
    delete g;
    g_grad.obj[gid] = nullptr;
    return 0;
}

int dnnf90_grad_zero(int gid)
{
    Grad *g = g_grad.get(gid);
    if (!g) return fail(-1, "bad accumulator handle");
    int rc = g->before(g->s);
    if (rc) return rc;
    launch_fill(g->d_g, g->nw, 0.0, g->s);
    if ((rc = g->after(g->s))) return rc;
    HIPCHK(hipStreamSynchronize(g->s));
    return 0;
}

int dnnf90_grad_point(int nid, int tid, const double *x, const double *seed, int gid)
{
    Net *nt = g_net.get(nid);
    if (!nt || !dependent_ok(g_twork, tid, nid) || !grad_ok(gid, nid) || !x || !seed)
        return fail(-1, "bad handle or argument");
    BatchWork &b = *g_twork.obj[tid];
    b.defer_use = true;   // the stream is drained below: no reader events
    int rc = run_forward<double>(*nt, b, 1, x, nullptr, false);
    if (rc) {
        b.defer_use = false;
        nt->used_on(b.stream);   // not drained on this path
        return rc;
    }
    const int na = nt->tab.na, nout = nt->dims[nt->nlayer - 1];
    std::vector<double> sm((size_t)nout * na, 0.0);
    std::memcpy(sm.data(), seed, sizeof(double) * na);  // output 1 only
    rc = run_backward_seed<double>(*nt, b, sm.data(), *g_grad.obj[gid], false);
    b.defer_use = false;
    if (rc) {
        nt->used_on(b.stream);   // not drained on this path
        return rc;
    }
    HIPCHK(hipStreamSynchronize(b.stream));
    return 0;
}

int dnnf90_grad_point_noalloc(int nid, const double *x, const double *seed, double *grad, void *stream)
{
    Net *nt = g_net.get(nid);
    if (!nt || !x || !seed || !grad) return fail(-1, "bad handle or argument");
    {
        std::lock_guard<std::mutex> lk(nt->mu);
        if (!nt->dt.ready || !nt->gen_tried[1] || !nt->d_w64)
            return fail(-1, "network not prepared on the device: create a work space of it first "
                            "(dnnf90_twork_create builds its tables, kernels and weights there)");
    }
    const GenKernels &g = nt->gen[1];
    const int L = nt->nlayer, na = nt->tab.na, nout = nt->dims[L - 1];
    bool fits = codegen_on() && g.ok && g.pt_fwd && g.pt_bwd && g.pt_scratch == 0 && L <= PT_MAXL;
    for (int l = 0; l < L && fits; ++l) fits = nt->dims[l] <= PT_MAXW && (long)nt->dims[l] * na <= PT_MAXWA;
    if (!fits)
        return fail(-1, "the network is outside the one-point kernels (at most 16 layers, widths <= 256, "
                        "width x slots <= 1536, a family of at most 48 slots and 96 operand entries)");
    long need = (long)nout * na;                                   // seed plane
    for (int l = 0; l < L; ++l) need += (long)nt->dims[l] * na;    // T of every layer
    for (int l = 1; l < L - 1; ++l) need += (long)nt->dims[l] * na;  // S of the hidden layers
    if (need > NOALLOC_DOUBLES)
        return fail(-1, "the one-point planes need " + std::to_string(need) + " doubles of scratch; the maximum is " +
                            std::to_string(NOALLOC_DOUBLES));
    double *sp = nullptr;
    int rc = scratch_of_device(&sp);
    if (rc) return rc;
    if ((rc = nt->ensure_device(false))) return rc;   // prepared: at most a weight copy, no allocation
    hipStream_t s = (hipStream_t)stream;
    PtNet p{};
    p.L = L;
    p.Np = 1;
    p.ldr = na;
    double *q = sp;
    double *sb = q;
    q += (long)nout * na;
    for (int l = 0; l < L; ++l) {
        p.dims[l] = nt->dims[l];
        p.woff[l] = nt->woff[l];
        p.T[l] = q;
        q += (long)nt->dims[l] * na;
    }
    for (int l = 1; l < L - 1; ++l) {
        p.S[l] = q;
        q += (long)nt->dims[l] * na;
    }
    hipLaunchKernelGGL(k_noalloc_seed, dim3((nout * na + 255) / 256), dim3(256), 0, s, seed, nout, na, sb);
    hipError_t e = gen_launch_pt_fwd(g, nt->d_w64, x, nt->dt.ind_e1, p, nt->dt.act(), s);
    if (e == hipSuccess) e = gen_launch_pt_bwd(g, nt->d_w64, grad, sb, p, nt->dt.act(), s);
    nt->used_on(s);
    if (e != hipSuccess) return hip_fail(e, "one-point launch");
    HIPCHK(hipGetLastError());
    return 0;
}

int dnnf90_adam_step(int nid, int gid, double eta, double beta1, double beta2, double eps,
                     int nbatch, int istep)
{
    Net *nt = g_net.get(nid);
    if (!nt || !grad_ok(gid, nid)) return fail(-1, "bad handle");
    Grad &g = *g_grad.obj[gid];
    int rc = nt->ensure_device(false);
    if (rc) return rc;
    // the step writes the weights: on the network's stream, after the
    // weights' readers and the accumulator's last writer
    if ((rc = nt->own_stream_after_readers()) || (rc = g.before(nt->ns))) return rc;
    hipStream_t s = nt->ns;
    if (!g.d_m) {
        if ((rc = dalloc(&g.d_m, (size_t)std::max(1, g.nw))) || (rc = dalloc(&g.d_v, (size_t)std::max(1, g.nw))))
            return rc;
        launch_fill(g.d_m, g.nw, 0.0, s);
        launch_fill(g.d_v, g.nw, 0.0, s);
    }
    // beta**istep as Fortran evaluates a real to an integer power: binary
    // exponentiation, squaring the base and multiplying in the set bits
    auto powi = [](double a, int b) {
        double r = 1.0;
        while (true) {
            if (b & 1) r *= a;
            b /= 2;
            if (b == 0) break;
            a *= a;
        }
        return r;
    };
    const int is = std::max(istep, 1);
    const double bc1 = 1.0 - powi(beta1, is), bc2 = 1.0 - powi(beta2, is);
    launch_adam(nt->d_w64, g.d_m, g.d_v, g.d_g, g.nw, eta, beta1, beta2, eps, (double)nbatch, bc1, bc2, s);
    if ((rc = g.after(s))) return rc;
    std::lock_guard<std::mutex> lk(nt->mu);
    HIPCHK(hipMemcpyAsync(nt->w.data(), nt->d_w64, sizeof(double) * nt->nw, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    nt->host_ver++;
    nt->dev_ver = nt->host_ver;  // the device copy is the updated one
    return 0;
}

// ---------------------------------------------------------------------------
// extensions (include/dnnf90_amd.h)

int dnnf90_tabset_counts(int tid, int *counts)
{
    TableSet *t = g_tab.get(tid);
    if (!t || !counts) return fail(-1, "bad table-set handle");
    counts[0] = t->na;
    counts[1] = t->npair();
    counts[2] = t->nfq();
    counts[3] = t->k;
    return 0;
}

int dnnf90_tabset_free(int tid)
{
    TableSet *t = g_tab.get(tid);
    if (!t) return fail(-1, "bad table-set handle");
    // networks hold their own copy of the tables, so none depends on the slot
    delete t;
    g_tab.obj[tid] = nullptr;
    return 0;
}

int dnnf90_tabset_write_order(int tid, const char *filename)
{
    TableSet *t = g_tab.get(tid);
    if (!t || !filename) return fail(-1, "bad table-set handle or file name");
    FILE *f = std::fopen(filename, "w");
    if (!f) return fail(-2, std::string("cannot write ") + filename);
    // lib/multi_index_bell_module.f90:451-471: two comment lines, then
    // "ia  |alpha|  alpha(1:D0)" one line per slot, 1-based ia
    std::fprintf(f, "# column order of high-order derivative targets\n");
    std::fprintf(f, "# ia  |alpha|  alpha(1:D0)\n");
    for (int a = 0; a < t->na; ++a) {
        std::fprintf(f, "%d  %d  ", a + 1, t->deg[a]);
        for (int i = 0; i < t->d0; ++i) std::fprintf(f, " %d", t->alpha[(size_t)a * t->d0 + i]);
        std::fprintf(f, "\n");
    }
    const bool ok = std::fclose(f) == 0;
    return ok ? 0 : fail(-2, std::string("cannot write ") + filename);
}

int dnnf90_tabset_lists(int tid, int *pair_start, int *pair_num, int *pair_ia, int *pair_id,
                        double *pair_binom, int *fq_start, int *fq_num, int *fq_ib, int *fq_id,
                        double *fq_c)
{
    TableSet *t = g_tab.get(tid);
    if (!t) return fail(-1, "bad table-set handle");
    auto cp = [](auto *dst, const auto &src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), sizeof(src[0]) * src.size());
    };
    cp(pair_start, t->pair_start); cp(pair_num, t->pair_num); cp(pair_ia, t->pair_ia);
    cp(pair_id, t->pair_id); cp(pair_binom, t->pair_binom); cp(fq_start, t->fq_start);
    cp(fq_num, t->fq_num); cp(fq_ib, t->fq_ib); cp(fq_id, t->fq_id); cp(fq_c, t->fq_c);
    return 0;
}

int dnnf90_net_nweights(int nid)
{
    Net *nt = g_net.get(nid);
    return nt ? nt->nw : fail(-1, "bad network handle");
}

int dnnf90_net_nlayer(int nid)
{
    Net *nt = g_net.get(nid);
    return nt ? nt->nlayer : fail(-1, "bad network handle");
}

int dnnf90_net_dims(int nid, int *dims)
{
    Net *nt = g_net.get(nid);
    if (!nt || !dims) return fail(-1, "bad network handle");
    for (int l = 0; l < nt->nlayer; ++l) dims[l] = nt->dims[l];
    return 0;
}

int dnnf90_net_nderiv(int nid)
{
    Net *nt = g_net.get(nid);
    return nt ? nt->tab.na : fail(-1, "bad network handle");
}

int dnnf90_net_get_layer(int nid, int l, double *wrow)
{
    Net *nt = g_net.get(nid);
    if (!nt || l < 2 || l > nt->nlayer || !wrow) return fail(-1, "bad network handle or layer");
    std::lock_guard<std::mutex> lk(nt->mu);
    const long n = (long)nt->dims[l - 1] * (nt->dims[l - 2] + 1);
    std::memcpy(wrow, &nt->w[nt->woff[l - 1]], sizeof(double) * n);
    return 0;
}

int dnnf90_net_get_weights(int nid, double *w)
{
    Net *nt = g_net.get(nid);
    if (!nt || !w) return fail(-1, "bad network handle");
    std::lock_guard<std::mutex> lk(nt->mu);
    std::memcpy(w, nt->w.data(), sizeof(double) * nt->nw);
    return 0;
}

int dnnf90_net_set_weights(int nid, const double *w)
{
    Net *nt = g_net.get(nid);
    if (!nt || !w) return fail(-1, "bad network handle");
    std::lock_guard<std::mutex> lk(nt->mu);
    std::memcpy(nt->w.data(), w, sizeof(double) * nt->nw);
    nt->host_ver++;
    return 0;
}

int dnnf90_net_save(int nid, const char *filename)
{
    Net *nt = g_net.get(nid);
    if (!nt || !filename) return fail(-1, "bad network handle");
    FILE *f = std::fopen(filename, "w");
    if (!f) return fail(-2, "cannot open " + std::string(filename));
    std::fprintf(f, "0\nfunc %d\nActivation_out 0\n%d\n", nt->tab.iact, nt->nlayer);
    for (int l = 0; l < nt->nlayer; ++l) std::fprintf(f, "%d\n", nt->dims[l]);
    for (int l = 1; l < nt->nlayer; ++l) {
        std::fprintf(f, " #l=           %d\n", l + 1);
        const int nrow = nt->dims[l], ncol = nt->dims[l - 1] + 1;
        for (int j = 0; j < nrow; ++j) {
            for (int i = 0; i < ncol; ++i) {
                std::fprintf(f, "%25.16E", nt->w[nt->woff[l] + (long)j * ncol + i]);
                if ((i + 1) % 100 == 0 && i + 1 < ncol) std::fprintf(f, "\n");
            }
            std::fprintf(f, "\n");
        }
    }
    std::fclose(f);
    return 0;
}

int dnnf90_grad_get(int gid, double *g)
{
    Grad *gr = g_grad.get(gid);
    if (!gr || !g) return fail(-1, "bad accumulator handle");
    int rc = gr->before(gr->s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(g, gr->d_g, sizeof(double) * gr->nw, hipMemcpyDeviceToHost, gr->s));
    HIPCHK(hipStreamSynchronize(gr->s));
    return 0;
}

int dnnf90_grad_set(int gid, const double *g)
{
    Grad *gr = g_grad.get(gid);
    if (!gr || !g) return fail(-1, "bad accumulator handle");
    int rc = gr->before(gr->s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(gr->d_g, g, sizeof(double) * gr->nw, hipMemcpyHostToDevice, gr->s));
    if ((rc = gr->after(gr->s))) return rc;
    HIPCHK(hipStreamSynchronize(gr->s));
    return 0;
}

int dnnf90_bwork_create(int nid, int maxpt, int prec) { return make_work(g_bwork, nid, maxpt, prec); }

int dnnf90_bwork_create_budget(int nid, int maxpt, int prec, double budget_bytes)
{
    return make_work(g_bwork, nid, maxpt, prec, budget_bytes);
}

int dnnf90_bwork_info(int bid, double *info)
{
    BatchWork *b = g_bwork.get(bid);
    if (!b || !info) return fail(-1, "bad work-space handle or argument");
    info[0] = b->recompute ? 1.0 : 0.0;
    info[1] = b->bytes;
    info[2] = b->bytes_stored;
    info[3] = (double)b->maxpt;
    info[4] = (double)b->prec;
    return 0;
}
int dnnf90_bwork_free(int bid) { return free_work(g_bwork, bid); }

int dnnf90_bwork_step_output(int bid, double *tout, int flags)
{
    BatchWork *b = g_bwork.get(bid);
    if (!b || !tout) return fail(-1, "bad work-space handle or argument");
    Net *nt = g_net.get(g_bwork.owner[bid]);
    if (!nt || !b->step_out) return fail(-1, "no collocation step output held by this work space");
    const bool dev = flags & DNNF90_DEVICE_PTRS;
    const int nout = nt->dims[nt->nlayer - 1], na = nt->tab.na, npt = b->npt;
    const size_t bytes = sizeof(double) * (size_t)npt * nout * na;
    double *dt = dev ? tout : b->d_tout;
    if (b->step_out == 1) {
        if (b->prec == PREC_FP64) launch_output<double>((const double *)b->tl, nout, na, npt, b->np, dt, b->stream);
        else launch_output<float>((const float *)b->tl, nout, na, npt, b->np, dt, b->stream);
    } else if (dev && npt > 0) {
        HIPCHK(hipMemcpyAsync(tout, b->d_tout, bytes, hipMemcpyDeviceToDevice, b->stream));
    }
    if (!dev && npt > 0) HIPCHK(hipMemcpyAsync(tout, b->d_tout, bytes, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipGetLastError());
    return 0;
}

int dnnf90_bwork_set_stream(int bid, void *stream)
{
    BatchWork *b = g_bwork.get(bid);
    if (!b) return fail(-1, "bad work-space handle");
    b->stream = stream ? (hipStream_t)stream : b->own;
    return 0;
}

// NGD step on the GPU (app/optimizer_module.f90:221-469, primal route):
// G = R R^T over the first nrow columns of the transposed FP32 rows, then
// u = (G/nb + cI)^{-1} (b/nb) by an FP64 Cholesky
extern "C++" {
// grow a device buffer to n elements
template <typename T>
static int grow(T **p, long *cap, long n)
{
    if (*cap >= n) return 0;
    dfree(*p);
    int rc = dalloc(p, (size_t)n);
    if (rc) return rc;
    *cap = n;
    return 0;
}

// Gram product chunking over ncol columns of an n x n result
static void gram_split(int n, long ncol, int *nc, long *ch)
{
    const int tiles = ((n + 127) / 128) * ((n + 143) / 144);
    int c = (int)std::min<long>(std::max(1, 1024 / tiles), std::max(1L, ncol / 512));
    long h = round_up((ncol + c - 1) / c, 16);
    *nc = (int)((ncol + h - 1) / h);
    *ch = h;
}
}  // extern "C++"

int dnnf90_ngd_solve(int nid, int bid, int nrow, const float *rows_t, long ld, const double *b, int nb,
                     double mu, int mode, double *u, int flags)
{
    Net *nt = g_net.get(nid);
    if (!nt || !dependent_ok(g_bwork, bid, nid) || !rows_t || !b || !u || nrow < 1 || nb < 1 || mu < 0.0 ||
        (mode != 0 && mode != 1))
        return fail(-1, "bad handle or argument");
    const long ncol = round_up(nrow, 16);
    if (ld < ncol || ld % 16 != 0) return fail(-1, "row stride must be a multiple of 16 and >= nrow");
    const bool dual = flags & DNNF90_NGD_DUAL;
    if (dual && mu <= 0.0) return fail(-1, "the dual route needs a positive damping");
    BatchWork &bw = *g_bwork.obj[bid];
    const int nw = nt->nw;
    const bool dev = flags & DNNF90_DEVICE_PTRS;
    // the factored system: nw x nw (primal) or nrow x nrow (dual)
    const int n = dual ? nrow : nw;
    const long ldj = round_up(nw, 16);          // dual: J row stride
    const long gcol = dual ? ldj : ncol;        // Gram reduction length
    const long len = (long)n * n;
    int nc;
    long ch;
    gram_split(n, gcol, &nc, &ch);
    int rc;
    if ((rc = grow(&bw.ngd_G, &bw.ngd_G_cap, len))) return rc;
    if ((rc = grow(&bw.ngd_part, &bw.ngd_part_cap, nc > 1 ? (long)(nc + grad_reduce_splits(nc)) * len : 0)))
        return rc;
    if ((rc = grow(&bw.ngd_u, &bw.ngd_u_cap, (long)nw))) return rc;
    if ((rc = grow(&bw.ngd_work, &bw.ngd_work_cap, ngd_work_doubles(n)))) return rc;
    if (!bw.ngd_trace && ((rc = dalloc(&bw.ngd_trace, 1)) || (rc = dalloc(&bw.ngd_info, 1)))) return rc;
    if (dual) {
        if ((rc = grow(&bw.ngd_J, &bw.ngd_J_cap, (long)nrow * ldj))) return rc;
        if ((rc = grow(&bw.ngd_v, &bw.ngd_v_cap, 2L * nw + nrow))) return rc;
    }
    hipStream_t s = bw.stream;
    if (nc > 1 && !dual) launch_fill(bw.ngd_G, len, 0.0, s);
    if (!dual) {
        {
            PScope p(bw, CAT_GRAD);
            launch_gram(rows_t, nw, ld, bw.ngd_part, nc, ch, bw.ngd_G, s);
        }
        HIPCHK(hipMemcpyAsync(bw.ngd_u, b, sizeof(double) * nw,
                              dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
        PScope p(bw, CAT_OTHER);
        if (ngd_factor_solve(bw.ngd_G, nw, nw, (double)nb, mu, mode, bw.ngd_u, bw.ngd_trace, bw.ngd_info,
                             bw.ngd_work, s))
            return fail(-3, "NGD factorization launch failed");
    } else {
        double *db = bw.ngd_v, *jv = bw.ngd_v + nw, *ry = bw.ngd_v + nw + nrow;
        HIPCHK(hipMemcpyAsync(db, b, sizeof(double) * nw, dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
        {
            PScope p(bw, CAT_GRAD);
            launch_rows_transpose(rows_t, ld, nw, nrow, bw.ngd_J, ldj, s);
            launch_gram64(bw.ngd_J, nrow, ldj, bw.ngd_G, s);
        }
        PScope p(bw, CAT_OTHER);
        launch_rows_matvec(bw.ngd_J, ldj, nrow, nw, db, jv, s);          // J b
        if (ngd_factor_solve(bw.ngd_G, nrow, nrow, (double)nb, mu, mode, jv, bw.ngd_trace, bw.ngd_info,
                             bw.ngd_work, s))                // y = K^-1 J b / nb
            return fail(-3, "NGD factorization launch failed");
        launch_rows_matvec(rows_t, ld, nw, nrow, jv, ry, s);            // R y = J^T y
        launch_dual_combine(nw, db, ry, (double)nb, mu, mode, bw.ngd_trace, bw.ngd_u, s);
    }
    int info = 0;
    HIPCHK(hipMemcpyAsync(&info, bw.ngd_info, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(u, bw.ngd_u, sizeof(double) * nw, dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    if (info > 0) return fail(-5, "damped metric not positive definite at column " + std::to_string(info));
    return 0;
}

int dnnf90_set_codegen(int on)
{
    const int prev = codegen_on() ? 1 : 0;
    g_codegen = on ? 1 : 0;
    return prev;
}

int dnnf90_set_point_kernels(int on)
{
    const int prev = g_point;
    g_point = on ? 1 : 0;
    return prev;
}

int dnnf90_set_split_kernels(int on) { return set_lin_x6c(on ? 1 : 0); }

int dnnf90_set_edge_kernels(int on)
{
    const int prev = g_edge;
    g_edge = on ? 1 : 0;
    return prev;
}

int dnnf90_set_fused_grad(int on) { return set_grad_lint_x6(on ? 1 : 0); }

int dnnf90_set_fused_layers(int on)
{
    const int prev = g_fused_layers;
    g_fused_layers = on ? 1 : 0;
    return prev;
}

int dnnf90_set_poison(int on)
{
    const int prev = g_poison;
    g_poison = on ? 1 : 0;
    return prev;
}

int dnnf90_set_store_s(int on)
{
    const int prev = g_store_s;
    g_store_s = on ? 1 : 0;
    return prev;
}

int dnnf90_net_kernel_info(int nid, int prec, int *info)
{
    Net *nt = g_net.get(nid);
    if (!nt || !info || prec < 0 || prec > 3) return fail(-1, "bad network handle or argument");
    int rc = nt->ensure_device(prec != PREC_FP64);
    if (rc) return rc;
    const GenKernels &g = nt->gen[prec == PREC_FP64 ? 1 : 0];
    const bool on = codegen_on() && g.ok, ks = on && g_store_s && g.s_ok();
    info[0] = on ? (ks ? 2 : 1) : 0;
    info[1] = on ? (ks ? g.fwdt_regs : g.fwd_regs) : 0;
    info[2] = on ? (ks ? g.bwds_regs : g.bwd_regs) : 0;
    info[3] = on ? (ks ? g.fwdt_scratch : g.fwd_scratch) : 0;
    info[4] = on ? (ks ? g.bwds_scratch : g.bwd_scratch) : 0;
    info[5] = nt->sched.nentry;
    return 0;
}

int dnnf90_net_fused_info(int nid, int *info)
{
    Net *nt = g_net.get(nid);
    if (!nt || !info) return fail(-1, "bad network handle or argument");
    int rc = nt->ensure_device(true);
    if (rc) return rc;
    BatchWork probe;
    probe.prec = PREC_FP32;
    const FusedKernels *f = fused_plan(*nt, probe);
    constexpr int NK = FusedKernels::NK;
    for (int i = 0; i < 2 + 2 * NK; ++i) info[i] = 0;
    if (!f) {
        // the network's shape or the switch rules them out (0), or the
        // module did not come up (FusedKernels::status)
        if (nt->fz_tried) info[0] = nt->fz.status == 0 ? -1 : nt->fz.status;
        if (nt->fz_tried)
            for (int i = 0; i < NK; ++i) {
                info[1 + i] = nt->fz.regs[i];
                info[1 + NK + i] = nt->fz.scratch[i];
            }
        return 0;
    }
    info[0] = 1;
    for (int i = 0; i < NK; ++i) {
        info[1 + i] = f->regs[i];
        info[1 + NK + i] = f->scratch[i];
    }
    info[1 + 2 * NK] = f->grid;
    return 0;
}

int dnnf90_net_step_info(int nid, int prec, int *info)
{
    Net *nt = g_net.get(nid);
    if (!nt || !info || prec < 0 || prec > 3) return fail(-1, "bad network handle or argument");
    int rc = nt->ensure_device(prec != PREC_FP64);
    if (rc) return rc;
    const GenKernels &g = nt->gen[prec == PREC_FP64 ? 1 : 0];
    const bool s_kept = codegen_on() && g.ok && g_store_s && g.s_ok();
    const EdgePlan p = edge_plan(*nt, prec == PREC_FP64, s_kept);
    info[0] = p.in;
    info[1] = p.in_grad;
    info[2] = p.out;
    info[3] = p.out_grad;
    // layers whose K5 and K6 are one kernel (bit l)
    info[4] = 0;
    const int L = nt->nlayer;
    for (int l = 2; l < L; ++l)
        if (prec == PREC_BF16X6 && !(l == L - 1 && p.out) && grad_lint_x6_ok(nt->dims[l], nt->dims[l - 1] + 1))
            info[4] |= 1 << l;
    return 0;
}

int dnnf90_bwork_profile(int bid, int on)
{
    BatchWork *b = g_bwork.get(bid);
    if (!b) return fail(-1, "bad work-space handle");
    int rc = b->resolve();
    if (rc) return rc;
    if (on < 0 || on >= 2 + NCAT) return fail(-1, "profile: 0, 1 or 2 + category");
    b->prof.on = on != 0;
    b->prof.only = on >= 2 ? on - 2 : -1;
    return 0;
}

int dnnf90_bwork_kernel_ms(int bid, double *ms, long *count)
{
    BatchWork *b = g_bwork.get(bid);
    if (!b || !ms || !count) return fail(-1, "bad work-space handle");
    int rc = b->resolve();
    if (rc) return rc;
    for (int c = 0; c < NCAT; ++c) {
        ms[c] = b->prof.ms[c];
        count[c] = b->prof.cnt[c];
        b->prof.ms[c] = 0.0;
        b->prof.cnt[c] = 0;
    }
    return 0;
}

int dnnf90_bwork_sync(int bid)
{
    BatchWork *b = g_bwork.get(bid);
    if (!b) return fail(-1, "bad work-space handle");
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int dnnf90_forward_batch(int nid, int bid, int npt, const double *x, double *tout, int flags)
{
    Net *nt = g_net.get(nid);
    if (!nt || !dependent_ok(g_bwork, bid, nid) || (!x && npt > 0)) return fail(-1, "bad handle or argument");
    BatchWork &b = *g_bwork.obj[bid];
    const bool dev = flags & DNNF90_DEVICE_PTRS;
    int rc = b.prec == PREC_FP64 ? run_forward<double>(*nt, b, npt, x, tout, dev)
                                 : run_forward<float>(*nt, b, npt, x, tout, dev);
    if (rc) return rc;
    if (!dev && tout) HIPCHK(hipStreamSynchronize(b.stream));
    return 0;
}

int dnnf90_backward_batch(int nid, int bid, const double *seed, int gid, int flags)
{
    Net *nt = g_net.get(nid);
    if (!nt || !dependent_ok(g_bwork, bid, nid) || !grad_ok(gid, nid) || !seed)
        return fail(-1, "bad handle or argument");
    BatchWork &b = *g_bwork.obj[bid];
    const bool dev = flags & DNNF90_DEVICE_PTRS;
    int rc = b.prec == PREC_FP64 ? run_backward_seed<double>(*nt, b, seed, *g_grad.obj[gid], dev)
                                 : run_backward_seed<float>(*nt, b, seed, *g_grad.obj[gid], dev);
    if (rc) return rc;
    if (!dev) HIPCHK(hipStreamSynchronize(b.stream));
    return 0;
}

int dnnf90_grad_batch(int nid, int bid, int npt, const double *x, const double *y, double coef, int gid, int flags)
{
    Net *nt = g_net.get(nid);
    if (!nt || !dependent_ok(g_bwork, bid, nid) || !grad_ok(gid, nid) || !x || !y)
        return fail(-1, "bad handle or argument");
    BatchWork &b = *g_bwork.obj[bid];
    const bool dev = flags & DNNF90_DEVICE_PTRS;
    int rc = b.prec == PREC_FP64 ? run_grad_value<double>(*nt, b, npt, x, y, coef, *g_grad.obj[gid], dev)
                                 : run_grad_value<float>(*nt, b, npt, x, y, coef, *g_grad.obj[gid], dev);
    if (rc) return rc;
    if (!dev) HIPCHK(hipStreamSynchronize(b.stream));   // the host x and y are read asynchronously
    return 0;
}

int dnnf90_residual_step_batch(int nid, int bid, int npt, const double *x, int nterm,
                               const double *coeff, const int *slot, const int *nonlin, double fac,
                               int gid, double *rsq, int flags)
{
    Net *nt = g_net.get(nid);
    if (!nt || !dependent_ok(g_bwork, bid, nid) || !grad_ok(gid, nid) || (!x && npt > 0) ||
        (nterm > 0 && (!coeff || !slot)))
        return fail(-1, "bad handle or argument");
    BatchWork &b = *g_bwork.obj[bid];
    const bool dev = flags & DNNF90_DEVICE_PTRS;
    int rc = b.prec == PREC_FP64
                 ? run_residual<double>(*nt, b, npt, x, nterm, coeff, slot, nonlin, fac, g_grad.obj[gid], rsq,
                                        nullptr, 0, dev)
                 : run_residual<float>(*nt, b, npt, x, nterm, coeff, slot, nonlin, fac, g_grad.obj[gid], rsq,
                                       nullptr, 0, dev);
    if (rc) return rc;
    if (!dev) HIPCHK(hipStreamSynchronize(b.stream));
    return 0;
}

int dnnf90_residual_rows_batch(int nid, int bid, int npt, const double *x, int nterm,
                               const double *coeff, const int *slot, const int *nonlin, float *rows_t,
                               long ld, int flags)
{
    Net *nt = g_net.get(nid);
    if (!nt || !dependent_ok(g_bwork, bid, nid) || (!x && npt > 0) || !rows_t ||
        (nterm > 0 && (!coeff || !slot)))
        return fail(-1, "bad handle or argument");
    BatchWork &b = *g_bwork.obj[bid];
    if (ld < round_up(std::max(npt, 1), 64)) return fail(-1, "row stride below the padded point count");
    const bool dev = flags & DNNF90_DEVICE_PTRS;
    int rc = b.prec == PREC_FP64
                 ? run_residual<double>(*nt, b, npt, x, nterm, coeff, slot, nonlin, 1.0, nullptr, nullptr,
                                        rows_t, ld, dev)
                 : run_residual<float>(*nt, b, npt, x, nterm, coeff, slot, nonlin, 1.0, nullptr, nullptr,
                                       rows_t, ld, dev);
    if (rc) return rc;
    if (!dev) HIPCHK(hipStreamSynchronize(b.stream));
    return 0;
}

// ---------------------------------------------------------------------------
// extended Kalman filter and committee (include/dnnf90_amd.h)

int dnnf90_kf_create(int nid, double p0, double lam_ini, double lam0, int decoupled)
{
    Net *nt = g_net.get(nid);
    if (!nt) return fail(-1, "bad network handle");
    if (!(p0 > 0.0)) return fail(-1, "p0 must be positive");
    if (!(lam_ini > 0.0 && lam_ini <= 1.0 && lam0 > 0.0 && lam0 <= 1.0))
        return fail(-1, "forgetting parameters must lie in (0, 1]");
    int id = g_kf.slot();
    if (id < 0) return fail(-1, "no free filter handle");
    KFilter *kf = new KFilter();
    int rc = kf_build(*nt, *kf, p0, lam_ini, lam0, decoupled);
    if (rc) {
        kf->release();
        delete kf;
        return rc;
    }
    g_kf.obj[id] = kf;
    g_kf.owner[id] = nid;
    g_kf.ogen[id] = g_net.gen[nid];
    return id;
}

int dnnf90_kf_set(int kid, double gate, double q)
{
    KFilter *kf = g_kf.get(kid);
    if (!kf) return fail(-1, "bad filter handle");
    if (gate < 0.0 || q < 0.0) return fail(-1, "gate and q must be >= 0");
    // every update call drains its stream before it returns: the state is settled
    KfScalars h;
    HIPCHK(hipMemcpy(&h, kf->st, sizeof(h), hipMemcpyDeviceToHost));
    h.gate = gate;
    h.q = q;
    HIPCHK(hipMemcpy(kf->st, &h, sizeof(h), hipMemcpyHostToDevice));
    return 0;
}

int dnnf90_kf_free(int kid)
{
    KFilter *kf = g_kf.get(kid);
    if (!kf) return fail(-1, "bad filter handle");
    kf->release();   // every update call drained its stream before it returned
    delete kf;
    g_kf.obj[kid] = nullptr;
    return 0;
}

int dnnf90_kf_update_grad(int nid, int kid, int gid, double xi, double scale)
{
    Net *nt = g_net.get(nid);
    if (!nt || !kf_ok(kid, nid) || !grad_ok(gid, nid)) return fail(-1, "bad handle");
    KFilter &kf = *g_kf.obj[kid];
    int rc = nt->ensure_device(false);
    if (rc) return rc;
    // on the network's stream (it writes the weights), after the weights'
    // readers and the accumulator's last writer
    if ((rc = nt->own_stream_after_readers())) return rc;
    const Grad &g = *g_grad.obj[gid];
    if ((rc = g.before(nt->ns))) return rc;
    kf_step(*nt, kf, g.d_g, scale, KF_XI_GIVEN, xi, nullptr, nullptr, 0, false, true, nt->ns);
    return kf_finish(*nt, kf, nt->ns);
}

int dnnf90_kf_update_obs(int nid, int kid, int bid, int nobs, const double *x, const double *seed,
                         const double *target, const double *pred, int flags)
{
    Net *nt = g_net.get(nid);
    if (!nt || !kf_ok(kid, nid) || !dependent_ok(g_bwork, bid, nid) || (nobs > 0 && (!seed || !target)))
        return fail(-1, "bad handle or argument");
    KFilter &kf = *g_kf.obj[kid];
    BatchWork &b = *g_bwork.obj[bid];
    const double *dx, *ds;
    int rc = kf_begin(*nt, b, kf, nobs, x, seed, flags & DNNF90_DEVICE_PTRS, &dx, &ds);
    if (rc) return rc;
    const int d0 = nt->dims[0], nsl = nt->dims[nt->nlayer - 1] * nt->tab.na;
    for (int n = 0; n < nobs; ++n) {
        // xi = target - pred, pred either given or the linear functional seed.T
        const int xm = pred ? KF_XI_GIVEN : KF_XI_LINEAR;
        const double v = pred ? target[n] - pred[n] : target[n];
        kf_observe(*nt, b, kf, dx + (long)n * d0, ds + (long)n * nsl, nullptr, nullptr, xm, v, 1.0, false, true);
    }
    b.fwd_valid = false;
    return kf_finish(*nt, kf, b.stream);
}

int dnnf90_kf_update_resid(int nid, int kid, int bid, int nobs, const double *x, const double *seed,
                           const double *resid, const double *rnoise, int flags)
{
    Net *nt = g_net.get(nid);
    if (!nt || !kf_ok(kid, nid) || !dependent_ok(g_bwork, bid, nid) || (nobs > 0 && (!seed || !resid)))
        return fail(-1, "bad handle or argument");
    for (int n = 0; rnoise && n < nobs; ++n)
        if (!(rnoise[n] > 0.0)) return fail(-1, "observation noise must be > 0");
    KFilter &kf = *g_kf.obj[kid];
    BatchWork &b = *g_bwork.obj[bid];
    const double *dx, *ds;
    int rc = kf_begin(*nt, b, kf, nobs, x, seed, flags & DNNF90_DEVICE_PTRS, &dx, &ds);
    if (rc) return rc;
    const int d0 = nt->dims[0], nsl = nt->dims[nt->nlayer - 1] * nt->tab.na;
    for (int n = 0; n < nobs; ++n) {
        const double sc = rnoise ? 1.0 / std::sqrt(rnoise[n]) : 1.0;
        kf_observe(*nt, b, kf, dx + (long)n * d0, ds + (long)n * nsl, nullptr, nullptr, KF_XI_GIVEN, -resid[n], sc,
                   false, true);
    }
    b.fwd_valid = false;
    return kf_finish(*nt, kf, b.stream);
}

int dnnf90_kf_update_residual(int nid, int kid, int bid, int nobs, const double *x, int nterm,
                              const double *coeff, const int *slot, const int *nonlin, const double *rsrc,
                              int flags)
{
    Net *nt = g_net.get(nid);
    if (!nt || !kf_ok(kid, nid) || !dependent_ok(g_bwork, bid, nid) || (nterm > 0 && (!coeff || !slot)))
        return fail(-1, "bad handle or argument");
    KFilter &kf = *g_kf.obj[kid];
    BatchWork &b = *g_bwork.obj[bid];
    if (b.prec != PREC_FP64) return fail(-1, "the filter runs on an FP64 work space");
    DevResidual res;
    int rc = set_residual(b, nterm, coeff, slot, nonlin, nt->tab.na, &res);
    if (rc) return rc;
    const double *dx;
    rc = kf_begin(*nt, b, kf, nobs, x, nullptr, flags & DNNF90_DEVICE_PTRS, &dx, nullptr);
    if (rc) return rc;
    const double *dr0 = rsrc;
    if (rsrc && !(flags & DNNF90_DEVICE_PTRS) && nobs > 0) {
        if ((rc = kf_grow(&kf.d_r0, &kf.r0_cap, nobs))) return rc;
        HIPCHK(hipMemcpyAsync(kf.d_r0, rsrc, sizeof(double) * nobs, hipMemcpyHostToDevice, b.stream));
        dr0 = kf.d_r0;
    }
    const int d0 = nt->dims[0];
    for (int n = 0; n < nobs; ++n)
        kf_observe(*nt, b, kf, dx + (long)n * d0, nullptr, &res, dr0 ? dr0 + n : nullptr, KF_XI_RESID, 0.0, 1.0,
                   false, true);
    b.fwd_valid = false;
    return kf_finish(*nt, kf, b.stream);
}

int dnnf90_kf_iekf_begin(int nid, int kid)
{
    Net *nt = g_net.get(nid);
    if (!nt || !kf_ok(kid, nid)) return fail(-1, "bad handle");
    KFilter &kf = *g_kf.obj[kid];
    int rc = nt->ensure_device(false);
    if (rc) return rc;
    if (!kf.wprior && (rc = dalloc(&kf.wprior, (size_t)std::max(1, kf.nw)))) return rc;
    HIPCHK(hipMemcpy(kf.wprior, nt->d_w64, sizeof(double) * kf.nw, hipMemcpyDeviceToDevice));
    kf.prior = true;
    return 0;
}

int dnnf90_kf_iekf_iter(int nid, int kid, int bid, const double *x, const double *seed, double resid,
                        double rnoise, int final, int flags)
{
    Net *nt = g_net.get(nid);
    if (!nt || !kf_ok(kid, nid) || !dependent_ok(g_bwork, bid, nid) || !seed)
        return fail(-1, "bad handle or argument");
    KFilter &kf = *g_kf.obj[kid];
    if (!kf.prior) return fail(-1, "dnnf90_kf_iekf_begin first");
    if (!(rnoise > 0.0)) return fail(-1, "observation noise must be > 0");
    BatchWork &b = *g_bwork.obj[bid];
    const double *dx, *ds;
    int rc = kf_begin(*nt, b, kf, 1, x, seed, flags & DNNF90_DEVICE_PTRS, &dx, &ds);
    if (rc) return rc;
    kf_observe(*nt, b, kf, dx, ds, nullptr, nullptr, KF_XI_IEKF, resid, 1.0 / std::sqrt(rnoise), true,
               final != 0);
    b.fwd_valid = false;
    return kf_finish(*nt, kf, b.stream);
}

int dnnf90_kf_get(int kid, double *P, double *info)
{
    KFilter *kfp = g_kf.get(kid);
    if (!kfp) return fail(-1, "bad filter handle");
    KFilter &kf = *kfp;
    KfScalars h;
    HIPCHK(hipMemcpy(&h, kf.st, sizeof(h), hipMemcpyDeviceToHost));
    if (P) {
        if (h.pending && !h.bad) {   // settle the downdate the last update left owing
            kf_pass(kf, nullptr, 1.0, 0);
            launch_kf_clear(kf.st, 0);
            HIPCHK(hipStreamSynchronize(nullptr));
        }
        std::vector<double> hp((size_t)kf.plen);
        HIPCHK(hipMemcpy(hp.data(), kf.P, sizeof(double) * kf.plen, hipMemcpyDeviceToHost));
        if (!kf.decoupled) {
            // P(nw, nw) column-major (symmetric), from the lower tiles
            const long n = kf.nw;
            for (long c = 0; c < n; ++c)
                for (long r = 0; r < n; ++r) {
                    const long gr = std::max(r, c), gc = std::min(r, c);
                    const long I = gr / 64, J = gc / 64;
                    P[c * n + r] = hp[(size_t)((I * (I + 1) / 2 + J) * 4096 + (gr % 64) * 64 + gc % 64)];
                }
        } else {
            // Pb(maxb, maxb, nb) column-major, zero outside each block
            const long mb = kf.maxb;
            std::fill(P, P + mb * mb * kf.nb, 0.0);
            for (int b = 0; b < kf.nb; ++b)
                for (long c = 0; c < kf.bsz[b]; ++c)
                    for (long r = 0; r < kf.bsz[b]; ++r)
                        P[b * mb * mb + c * mb + r] = hp[(size_t)(kf.poff[b] + r * kf.bsz[b] + c)];
        }
    }
    if (info) {
        info[0] = h.lam;
        info[1] = h.lam0;
        info[2] = h.gate;
        info[3] = h.q;
        info[4] = h.ngated;
        info[5] = kf.nw;
        info[6] = kf.decoupled ? kf.nb : 0;
        info[7] = kf.decoupled ? kf.maxb : 0;
        info[8] = h.bad;
        info[9] = (double)kf.plen * sizeof(double);
    }
    return 0;
}

int dnnf90_comm_eval_batch(int nmem, const int *nids, int bid, int npt, const double *x, double *tmean,
                           double *tstd, int flags)
{
    if (nmem < 1 || !nids || (npt > 0 && (!x || !tmean || !tstd))) return fail(-1, "bad argument");
    Net *n0 = g_net.get(nids[0]);
    if (!n0 || !dependent_ok(g_bwork, bid, nids[0])) return fail(-1, "bad handle");
    BatchWork &b = *g_bwork.obj[bid];
    if (npt < 0 || npt > b.maxpt) return fail(-1, "npt outside [0, maxpt]");
    std::vector<Net *> mem(nmem);
    for (int k = 0; k < nmem; ++k) {
        Net *nk = g_net.get(nids[k]);
        if (!nk) return fail(-1, "bad member handle");
        // members share the shape and the table family (committee_module.f90:44-50)
        if (nk->dims != n0->dims || nk->tab.alpha != n0->tab.alpha || nk->tab.iact != n0->tab.iact)
            return fail(-1, "committee members differ in shape or tables");
        mem[k] = nk;
    }
    const bool dev = flags & DNNF90_DEVICE_PTRS;
    const int nout = n0->dims[n0->nlayer - 1], na = n0->tab.na;
    const long len = (long)npt * nout * na;
    hipStream_t s = b.stream;
    int rc;
    for (Net *nk : mem)
        if ((rc = nk->ensure_device(b.prec != PREC_FP64))) return rc;
    if (len > 0 && (rc = grow_comm(b, len))) return rc;
    double *mean = dev ? tmean : b.comm_buf, *m2 = dev ? tstd : b.comm_buf + len;
    const double *dx = x;
    if (!dev && npt > 0) {
        HIPCHK(hipMemcpyAsync(b.d_x, x, sizeof(double) * npt * n0->dims[0], hipMemcpyHostToDevice, s));
        dx = b.d_x;
    }
    launch_fill(mean, len, 0.0, s);
    launch_fill(m2, len, 0.0, s);
    for (int k = 0; k < nmem; ++k) {
        if (b.prec == PREC_FP64) {
            if (hipError_t e = forward_t<double>(*mem[k], b, dx, npt)) return hip_fail(e, "committee forward");
            launch_output<double>((const double *)b.tl, nout, na, npt, b.np, b.d_tout, s);
        } else {
            if (hipError_t e = forward_t<float>(*mem[k], b, dx, npt)) return hip_fail(e, "committee forward");
            launch_output<float>((const float *)b.tl, nout, na, npt, b.np, b.d_tout, s);
        }
        launch_comm_welford(b.d_tout, mean, m2, len, k + 1, s);
    }
    launch_comm_finish(m2, len, nmem, s);
    if (!dev && len > 0) {
        HIPCHK(hipMemcpyAsync(tmean, mean, sizeof(double) * len, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(tstd, m2, sizeof(double) * len, hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    b.fwd_valid = false;
    return 0;
}

}  // extern "C"
