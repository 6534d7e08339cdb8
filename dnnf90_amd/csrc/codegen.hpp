// codegen.hpp -- run-time generated Faa di Bruno kernels (see codegen.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "kernels.hpp"
#include "tables.hpp"

namespace dnnf90 {

// fwd/bwd: the forward writes T and h, the adjoint reads Tbar and h.
// fwdt/bwds: the forward writes T only, the layer keeps S, and the adjoint
// reads Tbar and S and recomputes h in registers (one plane less per hidden
// layer); used when neither spills.
// one-point kernels (FP64 families with at most PT_MAXA slots and PT_MAXE
// operand entries; networks of at most PT_MAXL layers whose widths fit one
// workgroup, width x slots at most PT_MAXWA, the LDS rows of a plane): the
// whole forward, and the whole adjoint with the weight gradient, each one
// launch
constexpr int PT_MAXL = 16, PT_MAXA = 48, PT_MAXE = 96, PT_MAXW = 256, PT_MAXWA = 1536;
struct PtNet {                    // mirrors the generated kernels' argument
    int L, pad;
    long Np, ldr;
    int dims[PT_MAXL];
    long woff[PT_MAXL];
    double *T[PT_MAXL];
    double *S[PT_MAXL];
};

struct GenKernels {
    hipModule_t mod = nullptr;
    hipFunction_t fwd = nullptr, bwd = nullptr, fwdt = nullptr, bwds = nullptr;
    hipFunction_t pt_fwd = nullptr, pt_bwd = nullptr;   // FP64 small families only
    // edge variants: fwdt_in -- the first hidden layer, S formed from the
    // input plane T0 and the layer's weights; bwds_i likewise in the adjoint;
    // bwds_o -- the last hidden layer, Tbar formed from the output layer's
    // Sbar and weights; bwds_io both (one hidden layer).  Null where they
    // would need scratch.
    hipFunction_t fwdt_in = nullptr, bwds_i = nullptr, bwds_o = nullptr, bwds_io = nullptr;
    // the first hidden layer's adjoint forming that layer's weight gradient
    // (per point block) instead of its Sbar plane; iog with bwds_o's Tbar
    hipFunction_t bwds_ig = nullptr, bwds_iog = nullptr;
    // the last hidden layer's adjoint (one output row) with the output
    // layer's weight gradient
    hipFunction_t bwds_og = nullptr;
    std::vector<char> code;
    int fwd_regs = 0, bwd_regs = 0, fwd_scratch = 0, bwd_scratch = 0;
    int fwdt_regs = 0, bwds_regs = 0, fwdt_scratch = 0, bwds_scratch = 0;
    int pt_scratch = 0;
    bool ok = false;
    bool s_ok() const { return ok && fwdt_scratch == 0 && bwds_scratch == 0; }
    void release();
};

// HIP source of the straight-line forward and adjoint kernels of one family
std::string gen_bell_source(const TableSet &t, const FwdSchedule &s, bool fp64);
// whether the family is small enough to unroll
bool gen_bell_usable(const TableSet &t, const FwdSchedule &s);
// whether the FP64 module carries the one-point kernels for this family
bool gen_point_family(const TableSet &t, const FwdSchedule &s);
// compile with hipRTC and load on the current device; 0 on success
int gen_bell_compile(const TableSet &t, const FwdSchedule &s, bool fp64, GenKernels *g, std::string *log);
// load compiled code (from gen_bell_compile or a cache of it) into g;
// 0, or -3 with g released and the HIP error state cleared
int gen_bell_load(const std::vector<char> &code, bool fp64, GenKernels *g);

hipError_t gen_launch_fwd(const GenKernels &g, const void *S, void *Tout, void *H, int nrow, long Np,
                          const ActParams &ap, hipStream_t s);
hipError_t gen_launch_bwd(const GenKernels &g, const void *Tb, const void *H, void *Sb, int nrow, long Np,
                          hipStream_t s);
hipError_t gen_launch_fwdt(const GenKernels &g, const void *S, void *Tout, int nrow, long Np, const ActParams &ap,
                           hipStream_t s);
hipError_t gen_launch_bwds(const GenKernels &g, const void *Tb, const void *S, void *Sb, int nrow, long Np,
                           const ActParams &ap, hipStream_t s);
hipError_t gen_launch_fwdt_in(const GenKernels &g, const void *T0, const void *W1, void *Tout, int nrow, long Np,
                              const ActParams &ap, hipStream_t s);
// the adjoint from Tb (or, out_t, from the output layer's SbL and WL over
// nout rows) and S (or, in_s, S formed from T0 and W1), into Sb
hipError_t gen_launch_bwds_edge(const GenKernels &g, bool in_s, bool out_t, const void *Tb, const void *S, void *Sb,
                                int nrow, long Np, const ActParams &ap, const void *T0, const void *W1,
                                const void *SbL, const void *WL, int nout, hipStream_t s);
// the first hidden layer's adjoint with its weight gradient: part gets
// [Np/64][nrow][d0+1] point-block sums (bias first), no Sbar plane
hipError_t gen_launch_bwds_grad1(const GenKernels &g, bool out_t, const void *Tb, int nrow, long Np,
                                 const ActParams &ap, const void *T0, const void *W1, const void *SbL, const void *WL,
                                 int nout, double *part, hipStream_t s);
// the last hidden layer's adjoint from the single output row SbL (weights
// WL), writing Sb, with the output layer's gradient: part gets
// [Np/64][nrow + 1] point-block sums (bias first)
hipError_t gen_launch_bwds_gradout(const GenKernels &g, const void *S, void *Sb, int nrow, long Np, const ActParams &ap,
                                   const void *SbL, const void *WL, double *part, hipStream_t s);
hipError_t gen_launch_pt_fwd(const GenKernels &g, const double *W, const double *x, const int *ind_e1, const PtNet &p,
                             const ActParams &ap, hipStream_t s);
hipError_t gen_launch_pt_bwd(const GenKernels &g, const double *W, double *grad, const double *Sb, const PtNet &p,
                             const ActParams &ap, hipStream_t s);

// Fused layer kernels (FP32 storage; codegen.cpp): hidden layers FZ_W wide,
// point blocks of FZ_P points, slots at most FZ_MAXA, inputs at most
// FZ_MAXD0.  The adjoint's staged block in LDS: rows of FZ_PJ floats (one
// point's 64 neurons and 4 of padding), slots of FZ_XS (four rows), which
// keeps the linear maps' operand reads at most 2-way bank-conflicted.
constexpr int FZ_W = 64, FZ_P = 4, FZ_MAXA = 48, FZ_MAXD0 = 8, FZ_PJ = 68, FZ_XS = 272;
// terms of the collocation residual the fused last forward holds in LDS
constexpr int FZ_MAXT = 64;
// scratch (bytes a lane) a fused kernel may hold: loop-invariant values the
// register allocator parks there and reloads a few times a block
constexpr int FZ_MAX_SCRATCH = 256;
struct FusedKernels {
    enum Role { FWD_IN, FWD, FWD_LAST, BWD_LAST, BWD, BWD_IN, NK };
    int grid = 0;   // workgroups of the adjoint kernels: one resident round
    // 1 loaded; else why not: -2 hipRTC failed, -3 module or kernel missing,
    // -4 a kernel spills past FZ_MAX_SCRATCH, -5 no occupancy reported
    int status = 0;
    hipModule_t mod = nullptr;
    hipFunction_t fn[NK] = {};
    int regs[NK] = {}, scratch[NK] = {};
    std::vector<char> code;
    bool ok = false;
    void release();
};
bool gen_fused_usable(const TableSet &t, const FwdSchedule &s);
std::string gen_fused_source(const TableSet &t, const FwdSchedule &s);
int gen_fused_compile(const TableSet &t, const FwdSchedule &s, FusedKernels *f, std::string *log);
// 0, or -3 (a kernel missing or spilling) with f released
int gen_fused_load(const std::vector<char> &code, FusedKernels *f);
hipError_t fz_launch_fwd_in(const FusedKernels &f, const double *x, int npt, const float *W1, float *Tout, int nblk,
                            const ActParams &ap, hipStream_t s);
// forward of a hidden layer l >= 2 but the last, one block a workgroup: T
// and h of the layer
hipError_t fz_launch_fwd(const FusedKernels &f, const float *Tin, const float *W, float *Tout, float *H, int nblk,
                         const ActParams &ap, hipStream_t s);
// forward of the last hidden layer: h, the output row (into tout,
// [point][slot]), the collocation seed into Sbo ([block][slot][4 points]),
// the output layer's gradient partials gpart[block][65] (bias first) and
// rpart[block] = sum of R^2
hipError_t fz_launch_fwd_last(const FusedKernels &f, const float *Tin, const float *W, float *H, int nblk,
                              const ActParams &ap, const float *WL, int npt, const DevResidual &r, double fac,
                              float *Sbo, double *gpart, double *rpart, double *tout, hipStream_t s);
// adjoint of a hidden layer l >= 2 from its h plane (last: T-bar from the
// output row's seed in Tb)
hipError_t fz_launch_bwd(const FusedKernels &f, bool last, const float *Tin, const float *Tb, const float *H,
                         const float *W, const float *WL, float *Tbout, double *part, int nblk, const ActParams &ap,
                         int grid, hipStream_t s);
hipError_t fz_launch_bwd_in(const FusedKernels &f, const double *x, int npt, const float *Tb, const float *W1,
                            double *part, int nblk, const ActParams &ap, int grid, hipStream_t s);

}  // namespace dnnf90
