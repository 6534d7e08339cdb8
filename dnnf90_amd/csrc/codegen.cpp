// codegen.cpp -- per-family Faa di Bruno kernels generated at run time.
//
// The table-driven kernels (bell.hip) keep every lane's operands in LDS
// because register arrays cannot be indexed by run-time tables.  For the
// families a network actually carries, the forward lists
// (lib/multi_index_bell_module.f90:299-332) and pair lists (:267-297) are
// known once the tables are built, so this file prints the whole step as
// straight-line HIP -- one named register per operand, coefficients as exact
// hexadecimal literals, terms in the reference's order -- and compiles it
// with hipRTC for the device.  No LDS, no table loads; occupancy is set by
// registers alone.  The activation derivatives come from act.inc, the same
// text bell.hip compiles.  Families too large to unroll keep the
// table-driven kernels.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cstdio>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

#include "codegen.hpp"

namespace dnnf90 {

#define DNNF90_ACT_SRC(...) #__VA_ARGS__
static const char *kActSrc =
#include "act.inc"
    ;
#undef DNNF90_ACT_SRC

static std::string lit(double v, bool fp64)
{
    char buf[64];
    if (!fp64) v = (double)(float)v;
    std::snprintf(buf, sizeof(buf), "(real)(%a)", v);
    return buf;
}

static int kq_of(int k)
{
    return k <= 3 ? 5 : k <= 4 ? 6 : k <= 7 ? 9 : 17;
}

// The recurrence over slots 1..na-1 of one lane: for every slot the partial
// Bell sums v of each q (terms in the reference's order, stored entries kept
// as registers), T_a = d1 S_a + sum_q d_q v and h_a = d2 S_a + sum_q d_(q+1) v.
// REC_T stores T_a, REC_H stores h_a, REC_HREG assigns h_a to register h<a>.
// REC_TDOT adds T_a * Sbar_L(0, a) (the single output row, through slz)
// into gdot.  REC_TOUT hands T_a to the kernel's FZ_OUT(a, t) (the fused
// last hidden layer's output row).  st: the slot stride of the T store and
// of slz.
enum { REC_T = 1, REC_H = 2, REC_HREG = 4, REC_TDOT = 8, REC_TOUT = 16 };

static void emit_recurrence(std::ostringstream &o, const TableSet &t, const FwdSchedule &s, bool fp64, int m,
                            const char *st = "Np", const std::function<void(int)> &pre_slot = nullptr)
{
    const bool wt = (m & (REC_T | REC_TDOT | REC_TOUT)) != 0, wh = (m & (REC_H | REC_HREG)) != 0;
    for (int e = t.na; e < s.nentry; ++e) o << "  real e" << e << ";\n";
    for (int a = 1; a < t.na; ++a) {
        const int g0 = s.slot_group[a], g1 = s.slot_group[a + 1];
        if (pre_slot) pre_slot(a);
        o << "  {\n";
        if (wt) o << "    real t = d1 * e" << a << ";\n";
        if (wh) o << "    real h = d2 * e" << a << ";\n";
        for (int g = g0; g < g1; ++g) {
            const int q = s.grp_q[g];
            o << "    {\n      real v = (real)0;\n";
            for (int it = s.grp_term[g]; it < s.grp_term[g + 1]; ++it)
                o << "      v += " << lit(s.term_c[it], fp64) << " * e" << s.term_s[it] << " * e" << s.term_b[it]
                  << ";\n";
            if (s.grp_store[g] >= 0) o << "      e" << s.grp_store[g] << " = v;\n";
            if (wt) o << "      t += d" << q << " * v;\n";
            if (wh) o << "      h += d" << q + 1 << " * v;\n";
            o << "    }\n";
        }
        if (m & REC_T) o << "    Tj[" << a << "L * " << st << "] = t;\n";
        if (m & REC_H) o << "    Hj[" << a << "L * Np] = h;\n";
        if (m & REC_HREG) o << "    h" << a << " = h;\n";
        if (m & REC_TDOT) o << "    gdot += t * slz[" << a << "L * " << st << "];\n";
        if (m & REC_TOUT) o << "    FZ_OUT(" << a << ", t);\n";
        o << "  }\n";
    }
}

// Sbar_0 = sum_a Tbar_a h_a and the pair sums Sbar_b = sum binom Tbar_ia h_id
// (lib/multi_index_bell_module.f90:267-297) from registers t<a>, h<a>; into
// sb[b * Np], or (reg) into registers sb<b> the caller declares
static void emit_pair_sums(std::ostringstream &o, const TableSet &t, bool fp64, bool reg = false,
                           const std::function<void(int)> &post_slot = nullptr)
{
    auto put = [&](int b) {
        if (reg) o << "    sb" << b << " = v;\n";
        else o << "    sb[" << b << "L * Np] = v;\n";
    };
    o << "  {\n    real v = t0 * h0;\n";
    for (int a = 1; a < t.na; ++a) o << "    v += t" << a << " * h" << a << ";\n";
    put(0);
    o << "  }\n";
    if (post_slot) post_slot(0);
    for (int b = 1; b < t.na; ++b) {
        o << "  {\n    real v = (real)0;\n";
        for (int ip = t.pair_start[b]; ip < t.pair_start[b] + t.pair_num[b]; ++ip)
            o << "    v += " << lit(t.pair_binom[ip], fp64) << " * t" << t.pair_ia[ip] << " * h" << t.pair_id[ip]
              << ";\n";
        put(b);
        o << "  }\n";
        if (post_slot) post_slot(b);
    }
}

bool gen_point_family(const TableSet &t, const FwdSchedule &s)
{
    return t.na <= PT_MAXA && s.nentry <= PT_MAXE;
}

// One point through the whole network in one workgroup of 256 threads (FP64;
// the per-point ABI and the Kalman filter).  Every plane a phase reads is
// staged in LDS as [row][slot]; the work space's planes (column 0 of each
// slot) are written where later kernels read them (the forward's S and T)
// and read where the forward left them (the adjoint's S and T).  Forward,
// per layer: every (neuron, slot) dot product of K1 by one thread (bias
// first, then the inputs in order), a barrier, the Faa di Bruno forward one
// lane a neuron (the generated T-only body), a barrier.  Adjoint, from the
// seed in Sb, per layer: the Faa di Bruno adjoint one lane a neuron (h
// recomputed from S, then the pair sums), the weight gradient one thread a
// weight (the slots in order, as the one-point gradient kernel sums them)
// added into grad, the transposed map one thread a (neuron, slot) of the
// layer below, barriers between phases.  Networks whose width x slots stay
// within PT_MAXWA; the batched kernels take the rest.
template <typename LoadS>
void emit_point_kernels(std::ostringstream &o, const TableSet &t, const FwdSchedule &s, LoadS load_s)
{
    const int na = t.na;
    o << "#define PT_MAXL " << PT_MAXL << "\n"
      << "#define PT_LDS " << PT_MAXWA << "\n"
      << "struct PtNet { int L, pad; long Np, ldr; int dims[PT_MAXL]; long woff[PT_MAXL]; real *T[PT_MAXL]; "
         "real *S[PT_MAXL]; };\n";
    o << "extern \"C\" __global__ __launch_bounds__(256) void dnnf90_pt_fwd_f64(const real *__restrict__ W, "
         "const double *__restrict__ x, const int *__restrict__ ind_e1, PtNet p, ActParams ap)\n{\n"
      << "  __shared__ real xs[PT_LDS];   // T of the layer below\n"
      << "  __shared__ real ss[PT_LDS];   // S of this layer\n"
      << "  const int tid = threadIdx.x;\n"
      << "  const long GNp = p.Np, ldr = p.ldr;\n"
      << "  for (int u = tid; u < p.dims[0] * " << na << "; u += 256) {\n"
      << "    const int i = u / " << na << ", a = u % " << na << ";\n"
      << "    const real v = a == 0 ? (real)x[i] : (a == ind_e1[i] ? (real)1 : (real)0);\n"
      << "    xs[u] = v;\n"
      << "    p.T[0][i * ldr + a * GNp] = v;\n"
      << "  }\n"
      << "  __syncthreads();\n"
      << "  for (int l = 1; l < p.L; ++l) {\n"
      << "    const int nn = p.dims[l], m = p.dims[l - 1];\n"
      << "    const real *Wl = W + p.woff[l];\n"
      << "    real *outg = l < p.L - 1 ? p.S[l] : p.T[l];\n"
      << "    for (int u = tid; u < nn * " << na << "; u += 256) {\n"
      << "      const int jj = u / " << na << ", a = u % " << na << ";\n"
      << "      const real *wr = Wl + (long)jj * (m + 1);\n"
      << "      real v = a == 0 ? wr[0] : (real)0;\n"
      << "      for (int i = 0; i < m; ++i) v += wr[1 + i] * xs[i * " << na << " + a];\n"
      << "      ss[u] = v;\n"
      << "      outg[jj * ldr + a * GNp] = v;\n"
      << "    }\n"
      << "    __syncthreads();\n"
      << "    if (l < p.L - 1) {\n"
      << "      if (tid < nn) {\n"
      << "        const long Np = 1;   // LDS rows: slot stride 1\n"
      << "        const real *Sj = ss + tid * " << na << ";\n"
      << "        real *Tj = xs + tid * " << na << ";   // the next layer's input; K1 is done with xs\n";
    load_s();
    o << "        Tj[0] = d0;\n";
    emit_recurrence(o, t, s, true, REC_T);
    o << "        for (int a = 0; a < " << na << "; ++a) p.T[l][tid * ldr + a * GNp] = Tj[a];\n"
      << "      }\n"
      << "      __syncthreads();\n"
      << "    }\n"
      << "  }\n"
      << "}\n";

    o << "extern \"C\" __global__ __launch_bounds__(256) void dnnf90_pt_bwd_f64(const real *__restrict__ W, "
         "double *__restrict__ grad, const real *__restrict__ Sb, PtNet p, ActParams ap)\n{\n"
      << "  __shared__ real ss[PT_LDS];    // S of this layer\n"
      << "  __shared__ real xs[PT_LDS];    // T of the layer below\n"
      << "  __shared__ real sbs[PT_LDS];   // Sbar of this layer\n"
      << "  __shared__ real tbs[PT_LDS];   // Tbar of this layer\n"
      << "  const int tid = threadIdx.x;\n"
      << "  const long GNp = p.Np, ldr = p.ldr;\n"
      << "  for (int u = tid; u < p.dims[p.L - 1] * " << na << "; u += 256)\n"
      << "    sbs[u] = Sb[(u / " << na << ") * ldr + (u % " << na << ") * GNp];\n"
      << "  for (int l = p.L - 1; l >= 1; --l) {\n"
      << "    const int nn = p.dims[l], m = p.dims[l - 1];\n"
      << "    const real *Wl = W + p.woff[l];\n"
      << "    for (int u = tid; u < m * " << na << "; u += 256)\n"
      << "      xs[u] = p.T[l - 1][(u / " << na << ") * ldr + (u % " << na << ") * GNp];\n"
      << "    if (l < p.L - 1)\n"
      << "      for (int u = tid; u < nn * " << na << "; u += 256)\n"
      << "        ss[u] = p.S[l][(u / " << na << ") * ldr + (u % " << na << ") * GNp];\n"
      << "    __syncthreads();\n"
      << "    if (l < p.L - 1) {\n"
      << "      if (tid < nn) {\n"
      << "        const long Np = 1;   // LDS rows: slot stride 1\n"
      << "        const real *Sj = ss + tid * " << na << ";\n"
      << "        const real *tb = tbs + tid * " << na << ";\n"
      << "        real *sb = sbs + tid * " << na << ";\n";
    load_s();
    o << "        const real h0 = d1;\n";
    for (int a = 1; a < na; ++a) o << "        real h" << a << ";\n";
    emit_recurrence(o, t, s, true, REC_HREG);
    for (int a = 0; a < na; ++a) o << "        const real t" << a << " = tb[" << a << "L * Np];\n";
    emit_pair_sums(o, t, true);
    o << "      }\n"
      << "      __syncthreads();\n"
      << "    }\n"
      << "    for (int u = tid; u < nn * (m + 1); u += 256) {\n"
      << "      const int jj = u / (m + 1), kk = u % (m + 1);\n"
      << "      const real *sj = sbs + jj * " << na << ";\n"
      << "      double g;\n"
      << "      if (kk == 0) {\n"
      << "        g = (double)sj[0];\n"
      << "      } else {\n"
      << "        const real *xk = xs + (kk - 1) * " << na << ";\n"
      << "        g = 0.0;\n"
      << "        for (int a = 0; a < " << na << "; ++a) g += (double)sj[a] * (double)xk[a];\n"
      << "      }\n"
      << "      grad[p.woff[l] + u] += g;\n"
      << "    }\n"
      << "    if (l > 1) {\n"
      << "      for (int u = tid; u < m * " << na << "; u += 256) {\n"
      << "        const int i = u / " << na << ", a = u % " << na << ";\n"
      << "        real v = (real)0;\n"
      << "        for (int jj = 0; jj < nn; ++jj) v += Wl[(long)jj * (m + 1) + 1 + i] * sbs[jj * " << na << " + a];\n"
      << "        tbs[u] = v;\n"
      << "      }\n"
      << "    }\n"
      << "    __syncthreads();\n"
      << "  }\n"
      << "}\n";
}

std::string gen_bell_source(const TableSet &t, const FwdSchedule &s, bool fp64)
{
    std::ostringstream o;
    const int na = t.na, kq = kq_of(t.k);
    const bool bes = t.iact == ACT_BESSEL_J0 || t.iact == ACT_BESSEL_J1;
    const char *sfx = fp64 ? "_f64" : "_f32";
    // hipRTC supplies the HIP device environment itself (no runtime header)
    o << "struct ActParams { int iact, nq; const double *tpc; int ldt; };\n"
      << kActSrc << "\n"
      << "typedef " << (fp64 ? "double" : "float") << " real;\n"
      // the sum of v over the wave's 64 lanes by a fixed butterfly (lane 0's
      // result is used: one fixed order)
      << "__device__ __forceinline__ double wsum64(double v)\n{\n"
      << "  for (int o = 32; o > 0; o >>= 1) {\n"
      << "    const int src = (int)(((threadIdx.x & 63) ^ o) << 2);\n"
      << "    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);\n"
      << "    const int lo = __builtin_amdgcn_ds_bpermute(src, (int)(unsigned)b);\n"
      << "    const int hi = __builtin_amdgcn_ds_bpermute(src, (int)(unsigned)(b >> 32));\n"
      << "    v += __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);\n"
      << "  }\n  return v;\n}\n";
    // the edge kernels' row blocks of one point block share operands (the
    // input values, the output layer's Sbar): the block index is remapped so
    // that each XCD (workgroups go to the 8 XCDs in turn) runs a contiguous
    // run of point blocks, all row blocks of each together, and the shared
    // operands stay in that XCD's L2.  pb / rb: point and row block.
    auto block_ids = [&](bool xcd) {
        if (!xcd) {
            o << "  const unsigned pb = blockIdx.x, rb = blockIdx.y;\n";
            return;
        }
        o << "  unsigned lin = blockIdx.x + gridDim.x * blockIdx.y;\n"
          << "  {\n    const unsigned G = gridDim.x * gridDim.y;\n"
          << "    if ((G & 7u) == 0u) lin = (lin & 7u) * (G >> 3) + (lin >> 3);\n  }\n"
          << "  const unsigned pb = lin / gridDim.y, rb = lin % gridDim.y;\n";
    };
    auto head_x = [&](const char *name, const char *args, bool xcd) {
        o << "extern \"C\" __global__ __launch_bounds__(256) void " << name << sfx << "(" << args << ")\n{\n"
          << "  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;\n";
        block_ids(xcd);
        o << "  const int j = rb * (blockDim.x >> 6) + wave;\n"
          << "  if (j >= nrow) return;\n"
          << "  const long n = (long)pb * 64 + lane;\n"
          << "  const long ldr = " << na << "L * Np;\n";
    };
    auto head = [&](const char *name, const char *args) { head_x(name, args, false); };
    // S of one lane and the activation derivatives sigma^(q)(S_0), q < kq, in FP64
    auto load_s = [&]() {
        for (int a = 0; a < na; ++a) o << "  const real e" << a << " = Sj[" << a << "L * Np];\n";
        o << "  double d[" << kq << "];\n"
          << "  act_derivs<" << kq << ", " << (bes ? "true" : "false") << ">(ap, (double)e0, d);\n";
        for (int q = 0; q < kq; ++q) o << "  const real d" << q << " = (real)d[" << q << "];\n";
    };
    // ---- forward: T_a and h_a of every slot from the S_a of one lane
    head("dnnf90_bell_fwd", "const real *__restrict__ S, real *__restrict__ Tout, real *__restrict__ H, int nrow, "
                            "long Np, ActParams ap");
    o << "  const real *Sj = S + (long)j * ldr + n;\n"
      << "  real *Tj = Tout + (long)j * ldr + n;\n"
      << "  real *Hj = H + (long)j * ldr + n;\n";
    load_s();
    o << "  Tj[0] = d0;\n  Hj[0] = d1;\n";
    emit_recurrence(o, t, s, fp64, REC_T | REC_H);
    o << "}\n";

    // ---- adjoint: Sbar from Tbar and h of one lane
    head("dnnf90_bell_bwd", "const real *__restrict__ Tb, const real *__restrict__ H, real *__restrict__ Sb, int nrow, "
                            "long Np");
    o << "  const real *tb = Tb + (long)j * ldr + n;\n"
      << "  const real *hh = H + (long)j * ldr + n;\n"
      << "  real *sb = Sb + (long)j * ldr + n;\n";
    for (int a = 0; a < na; ++a)
        o << "  const real t" << a << " = tb[" << a << "L * Np], h" << a << " = hh[" << a << "L * Np];\n";
    emit_pair_sums(o, t, fp64);
    o << "}\n";

    // ---- forward, T only: the layer keeps S and the adjoint recomputes h
    head("dnnf90_bell_fwdt", "const real *__restrict__ S, real *__restrict__ Tout, int nrow, long Np, ActParams ap");
    o << "  const real *Sj = S + (long)j * ldr + n;\n"
      << "  real *Tj = Tout + (long)j * ldr + n;\n";
    load_s();
    o << "  Tj[0] = d0;\n";
    emit_recurrence(o, t, s, fp64, REC_T);
    o << "}\n";

    // ---- adjoint from Tbar and S: h of every slot recomputed in registers
    // (the same expressions as the forward above, so bit-identical), then
    // the pair sums
    head("dnnf90_bell_bwds", "const real *__restrict__ Tb, const real *__restrict__ S, real *__restrict__ Sb, int nrow, "
                             "long Np, ActParams ap");
    o << "  const real *Sj = S + (long)j * ldr + n;\n"
      << "  const real *tb = Tb + (long)j * ldr + n;\n"
      << "  real *sb = Sb + (long)j * ldr + n;\n";
    load_s();
    o << "  const real h0 = d1;\n";
    for (int a = 1; a < na; ++a) o << "  real h" << a << ";\n";
    emit_recurrence(o, t, s, fp64, REC_HREG);
    for (int a = 0; a < na; ++a) o << "  const real t" << a << " = tb[" << a << "L * Np];\n";
    emit_pair_sums(o, t, fp64);
    o << "}\n";

    // ---- the layers next to the input and the output.  The first hidden
    // layer's S needs no plane: T0 holds x at slot 0 and 1 at slot e_i, so
    // S_0 = b + sum_i w_i x_i (summed in FP64, then rounded), S at e_i = w_i
    // and every other slot is zero (the zero operands fold away in the
    // recurrence).  The last hidden layer's Tbar needs no plane either:
    // Tbar_a = sum_k W_L(k, j) Sbar_L(k, a) over the nout output rows.
    auto load_s_in = [&]() {
        const int d0 = t.d0;
        o << "  const real *wj = W1 + (long)j * " << d0 + 1 << ";\n"
          << "  double s0 = (double)wj[0];\n";
        for (int i = 0; i < d0; ++i) o << "  s0 += (double)wj[" << i + 1 << "] * (double)T0[" << i << "L * ldr + n];\n";
        o << "  const real e0 = (real)s0;\n";
        for (int a = 1; a < na; ++a) {
            int src = -1;
            for (int i = 0; i < d0; ++i)
                if (t.ind_e1[i] == a) src = i;
            if (src >= 0) o << "  const real e" << a << " = wj[" << src + 1 << "];\n";
            else o << "  const real e" << a << " = (real)0;\n";
        }
        o << "  double d[" << kq << "];\n"
          << "  act_derivs<" << kq << ", " << (bes ? "true" : "false") << ">(ap, (double)e0, d);\n";
        for (int q = 0; q < kq; ++q) o << "  const real d" << q << " = (real)d[" << q << "];\n";
    };
    auto load_tb_out = [&]() {
        for (int a = 0; a < na; ++a) o << "  real t" << a << " = (real)0;\n";
        o << "  for (int k = 0; k < nout; ++k) {\n"
          << "    const real wk = WL[(long)k * (nrow + 1) + 1 + j];\n"
          << "    const real *sl = SbL + (long)k * ldr + n;\n";
        for (int a = 0; a < na; ++a) o << "    t" << a << " += wk * sl[" << a << "L * Np];\n";
        o << "  }\n";
    };
    head_x("dnnf90_bell_fwdt_in", "const real *__restrict__ T0, const real *__restrict__ W1, real *__restrict__ Tout, "
                                  "int nrow, long Np, ActParams ap", true);
    o << "  real *Tj = Tout + (long)j * ldr + n;\n";
    load_s_in();
    o << "  Tj[0] = d0;\n";
    emit_recurrence(o, t, s, fp64, REC_T);
    o << "}\n";
    const char *bargs = "const real *__restrict__ Tb, const real *__restrict__ S, real *__restrict__ Sb, int nrow, "
                        "long Np, ActParams ap, const real *__restrict__ T0, const real *__restrict__ W1, "
                        "const real *__restrict__ SbL, const real *__restrict__ WL, int nout";
    for (int v = 1; v < 4; ++v) {
        const bool in_s = v & 1, out_t = v & 2;
        const std::string name = std::string("dnnf90_bell_bwds_") + (in_s ? "i" : "") + (out_t ? "o" : "");
        head_x(name.c_str(), bargs, true);
        o << "  real *sb = Sb + (long)j * ldr + n;\n";
        if (in_s) {
            load_s_in();
        } else {
            o << "  const real *Sj = S + (long)j * ldr + n;\n";
            load_s();
        }
        o << "  const real h0 = d1;\n";
        for (int a = 1; a < na; ++a) o << "  real h" << a << ";\n";
        emit_recurrence(o, t, s, fp64, REC_HREG);
        if (out_t) {
            load_tb_out();
        } else {
            o << "  const real *tb = Tb + (long)j * ldr + n;\n";
            for (int a = 0; a < na; ++a) o << "  const real t" << a << " = tb[" << a << "L * Np];\n";
        }
        emit_pair_sums(o, t, fp64);
        o << "}\n";
    }
    // ---- the first hidden layer's adjoint with its weight gradient.  T0 is
    // x at slot 0 and 1 at slot e_i, so the layer's gradient needs Sbar at
    // slot 0 and at the slots e_i alone: g(j, bias) = sum_n Sbar_0 and
    // g(j, i) = sum_n (Sbar_0 x_i + Sbar_(e_i)).  The lanes of a workgroup
    // (64 points) are summed in FP64 by a fixed butterfly over the wave
    // (wsum64); part[pb][j][c] takes the sums and a fixed-order reduction
    // adds the point blocks.  No Sbar plane is written.
    for (int v = 1; v < 4; v += 2) {
        const bool out_t = v & 2;
        const int d0 = t.d0;
        const std::string name = std::string("dnnf90_bell_bwds_i") + (out_t ? "o" : "") + "g";
        o << "extern \"C\" __global__ __launch_bounds__(256) void " << name << sfx << "(" << bargs
          << ", double *__restrict__ part)\n{\n"
          << "  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;\n";
        block_ids(true);
        o << "  const int j0 = rb * (blockDim.x >> 6) + wave;\n"
          << "  const bool jok = j0 < nrow;\n"
          << "  const int j = jok ? j0 : nrow - 1;   // every lane takes part in the butterflies\n"
          << "  const long n = (long)pb * 64 + lane;\n"
          << "  const long ldr = " << na << "L * Np;\n";
        load_s_in();
        o << "  const real h0 = d1;\n";
        for (int a = 1; a < na; ++a) o << "  real h" << a << ";\n";
        emit_recurrence(o, t, s, fp64, REC_HREG);
        if (out_t) {
            load_tb_out();
        } else {
            o << "  const real *tb = Tb + (long)j * ldr + n;\n";
            for (int a = 0; a < na; ++a) o << "  const real t" << a << " = tb[" << a << "L * Np];\n";
        }
        // Sbar at slot 0 and at the slots e_i, as emit_pair_sums forms them
        o << "  real sbar0;\n  {\n    real v = t0 * h0;\n";
        for (int a = 1; a < na; ++a) o << "    v += t" << a << " * h" << a << ";\n";
        o << "    sbar0 = v;\n  }\n";
        o << "  double *pj = part + ((long)pb * nrow + j) * " << d0 + 1 << ";\n"
          << "  {\n    const double g = wsum64(jok ? (double)sbar0 : 0.0);\n"
          << "    if (jok && lane == 0) pj[0] = g;\n  }\n";
        for (int i = 0; i < d0; ++i) {
            const int b = t.ind_e1[i];
            o << "  {\n    double g = (double)sbar0 * (double)T0[" << i << "L * ldr + n];\n";
            if (b >= 0) {
                o << "    real v = (real)0;\n";
                for (int ip = t.pair_start[b]; ip < t.pair_start[b] + t.pair_num[b]; ++ip)
                    o << "    v += " << lit(t.pair_binom[ip], fp64) << " * t" << t.pair_ia[ip] << " * h"
                      << t.pair_id[ip] << ";\n";
                o << "    g += (double)v;\n";
            }
            o << "    g = wsum64(jok ? g : 0.0);\n"
              << "    if (jok && lane == 0) pj[" << i + 1 << "] = g;\n  }\n";
        }
        o << "}\n";
    }
    // ---- the last hidden layer's adjoint (Tbar from the single output
    // row) with the output layer's weight gradient: T of the layer is
    // recomputed beside h from the same sums, g(0, 1 + j) = sum_n sum_a
    // Sbar_L(0, a) T_a(j) and the bias sum_n Sbar_L(0, 0), per 64-point block
    // in part[pb][c] (c = 0 bias, 1 + j weights), summed in FP64 by the
    // wave butterfly; the Sbar plane of the layer is written as bwds_o does
    {
        o << "extern \"C\" __global__ __launch_bounds__(256) void dnnf90_bell_bwds_og" << sfx << "(" << bargs
          << ", double *__restrict__ part)\n{\n"
          << "  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;\n";
        block_ids(true);
        o << "  const int j0 = rb * (blockDim.x >> 6) + wave;\n"
          << "  const bool jok = j0 < nrow;\n"
          << "  const int j = jok ? j0 : nrow - 1;   // every lane takes part in the butterflies\n"
          << "  const long n = (long)pb * 64 + lane;\n"
          << "  const long ldr = " << na << "L * Np;\n"
          << "  real *sb = Sb + (long)j * ldr + n;\n"
          << "  const real *slz = SbL + n;   // the output row (nout is 1)\n"
          << "  const real *Sj = S + (long)j * ldr + n;\n";
        load_s();
        o << "  const real sl00 = slz[0];\n"
          << "  real gdot = d0 * sl00;\n"
          << "  const real h0 = d1;\n";
        for (int a = 1; a < na; ++a) o << "  real h" << a << ";\n";
        emit_recurrence(o, t, s, fp64, REC_HREG | REC_TDOT);
        load_tb_out();
        o << "  if (jok) {\n";
        emit_pair_sums(o, t, fp64);
        o << "  }\n"
          << "  {\n    const double g = wsum64(jok ? (double)gdot : 0.0);\n"
          << "    if (lane == 0 && jok) part[(long)pb * (nrow + 1) + 1 + j] = g;\n  }\n"
          << "  if (wave == 0 && rb == 0) {\n"
          << "    const double g = wsum64((double)sl00);\n"
          << "    if (lane == 0) part[(long)pb * (nrow + 1)] = g;\n"
          << "  }\n}\n";
    }
    if (fp64 && gen_point_family(t, s)) emit_point_kernels(o, t, s, load_s);
    return o.str();
}

bool gen_bell_usable(const TableSet &t, const FwdSchedule &s)
{
    // straight-line size and live registers stay modest below these
    return t.nfq() <= 4096 && t.npair() <= 4096 && s.nentry <= 320 && t.na <= 160;
}

// hipRTC for gfx950 (packed FP32 instructions off, as for every kernel of
// the library); 0 and the code object, or -3 with the log
static int rtc_compile(const std::string &src, const char *name, std::vector<char> *code, std::string *log)
{
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), name, 0, nullptr, nullptr) != HIPRTC_SUCCESS) return -3;
    const char *opts[] = {"--gpu-architecture=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                          "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (rc != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string l(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &l[0]);
        if (log) *log = l;
        hiprtcDestroyProgram(&prog);
        return -3;
    }
    size_t sz = 0;
    hiprtcGetCodeSize(prog, &sz);
    code->assign(sz, 0);
    hiprtcGetCode(prog, code->data());
    hiprtcDestroyProgram(&prog);
    return 0;
}

int gen_bell_compile(const TableSet &t, const FwdSchedule &s, bool fp64, GenKernels *g, std::string *log)
{
    *g = GenKernels();
    std::vector<char> code;
    if (rtc_compile(gen_bell_source(t, s, fp64), "dnnf90_bell_gen.hip", &code, log)) return -3;
    return gen_bell_load(code, fp64, g);
}

int gen_bell_load(const std::vector<char> &code, bool fp64, GenKernels *g)
{
    *g = GenKernels();
    g->code = code;
    const std::string sfx = fp64 ? "_f64" : "_f32";
    auto get = [&](hipFunction_t *f, const char *name) {
        return hipModuleGetFunction(f, g->mod, (name + sfx).c_str()) == hipSuccess;
    };
    if (hipModuleLoadData(&g->mod, g->code.data()) != hipSuccess || !get(&g->fwd, "dnnf90_bell_fwd") ||
        !get(&g->bwd, "dnnf90_bell_bwd") || !get(&g->fwdt, "dnnf90_bell_fwdt") || !get(&g->bwds, "dnnf90_bell_bwds")) {
        // the table-driven kernels take over; leave no error behind for
        // the next launch check
        (void)hipGetLastError();
        g->release();
        return -3;
    }
    (void)hipFuncGetAttribute(&g->fwd_scratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, g->fwd);
    (void)hipFuncGetAttribute(&g->bwd_scratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, g->bwd);
    (void)hipFuncGetAttribute(&g->fwd_regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, g->fwd);
    (void)hipFuncGetAttribute(&g->bwd_regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, g->bwd);
    (void)hipFuncGetAttribute(&g->fwdt_scratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, g->fwdt);
    (void)hipFuncGetAttribute(&g->bwds_scratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, g->bwds);
    (void)hipFuncGetAttribute(&g->fwdt_regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, g->fwdt);
    (void)hipFuncGetAttribute(&g->bwds_regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, g->bwds);
    // the edge variants: used where they compile without scratch
    {
        hipFunction_t *f[7] = {&g->fwdt_in, &g->bwds_i,  &g->bwds_o, &g->bwds_io,
                               &g->bwds_ig,  &g->bwds_iog, &g->bwds_og};
        const char *nm[7] = {"dnnf90_bell_fwdt_in", "dnnf90_bell_bwds_i",   "dnnf90_bell_bwds_o", "dnnf90_bell_bwds_io",
                             "dnnf90_bell_bwds_ig", "dnnf90_bell_bwds_iog", "dnnf90_bell_bwds_og"};
        for (int i = 0; i < 7; ++i) {
            int sc = 0;
            if (!get(f[i], nm[i]) ||
                hipFuncGetAttribute(&sc, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, *f[i]) != hipSuccess || sc != 0)
                *f[i] = nullptr;
        }
        // the output-gradient variant holds T beside h: where that leaves one
        // wave a SIMD (the registers of large families spill into the
        // accumulation registers) it streams slower than the separate
        // gradient kernel, so it is kept only at two resident workgroups a CU
        int occ = 0;
        if (g->bwds_og && (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, g->bwds_og, 256, 0) != hipSuccess ||
                           occ < 2))
            g->bwds_og = nullptr;
        (void)hipGetLastError();
    }
    // the one-point kernels are in the FP64 modules of small families only
    if (fp64 && hipModuleGetFunction(&g->pt_fwd, g->mod, "dnnf90_pt_fwd_f64") == hipSuccess &&
        hipModuleGetFunction(&g->pt_bwd, g->mod, "dnnf90_pt_bwd_f64") == hipSuccess) {
        int a = 0, b = 0;
        (void)hipFuncGetAttribute(&a, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, g->pt_fwd);
        (void)hipFuncGetAttribute(&b, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, g->pt_bwd);
        g->pt_scratch = std::max(a, b);
    } else {
        g->pt_fwd = g->pt_bwd = nullptr;
        (void)hipGetLastError();
    }
    g->ok = true;
    return 0;
}

void GenKernels::release()
{
    if (mod) (void)hipModuleUnload(mod);
    mod = nullptr;
    fwd = bwd = fwdt = bwds = pt_fwd = pt_bwd = nullptr;
    fwdt_in = bwds_i = bwds_o = bwds_io = bwds_ig = bwds_iog = bwds_og = nullptr;
    ok = false;
}

hipError_t gen_launch_fwd(const GenKernels &g, const void *S, void *Tout, void *H, int nrow, long Np,
                          const ActParams &ap, hipStream_t s)
{
    ActParams a = ap;
    void *params[] = {&S, &Tout, &H, &nrow, &Np, &a};
    return hipModuleLaunchKernel(g.fwd, (unsigned)(Np / 64), (unsigned)((nrow + 3) / 4), 1, 256, 1, 1, 0, s, params,
                                 nullptr);
}

hipError_t gen_launch_bwd(const GenKernels &g, const void *Tb, const void *H, void *Sb, int nrow, long Np,
                          hipStream_t s)
{
    void *params[] = {&Tb, &H, &Sb, &nrow, &Np};
    return hipModuleLaunchKernel(g.bwd, (unsigned)(Np / 64), (unsigned)((nrow + 3) / 4), 1, 256, 1, 1, 0, s, params,
                                 nullptr);
}

hipError_t gen_launch_fwdt(const GenKernels &g, const void *S, void *Tout, int nrow, long Np, const ActParams &ap,
                           hipStream_t s)
{
    ActParams a = ap;
    void *params[] = {&S, &Tout, &nrow, &Np, &a};
    return hipModuleLaunchKernel(g.fwdt, (unsigned)(Np / 64), (unsigned)((nrow + 3) / 4), 1, 256, 1, 1, 0, s, params,
                                 nullptr);
}

hipError_t gen_launch_bwds(const GenKernels &g, const void *Tb, const void *S, void *Sb, int nrow, long Np,
                           const ActParams &ap, hipStream_t s)
{
    ActParams a = ap;
    void *params[] = {&Tb, &S, &Sb, &nrow, &Np, &a};
    return hipModuleLaunchKernel(g.bwds, (unsigned)(Np / 64), (unsigned)((nrow + 3) / 4), 1, 256, 1, 1, 0, s, params,
                                 nullptr);
}

hipError_t gen_launch_fwdt_in(const GenKernels &g, const void *T0, const void *W1, void *Tout, int nrow, long Np,
                              const ActParams &ap, hipStream_t s)
{
    ActParams a = ap;
    void *params[] = {&T0, &W1, &Tout, &nrow, &Np, &a};
    return hipModuleLaunchKernel(g.fwdt_in, (unsigned)(Np / 64), (unsigned)((nrow + 3) / 4), 1, 256, 1, 1, 0, s,
                                 params, nullptr);
}

hipError_t gen_launch_bwds_edge(const GenKernels &g, bool in_s, bool out_t, const void *Tb, const void *S, void *Sb,
                                int nrow, long Np, const ActParams &ap, const void *T0, const void *W1,
                                const void *SbL, const void *WL, int nout, hipStream_t s)
{
    hipFunction_t f = in_s ? (out_t ? g.bwds_io : g.bwds_i) : (out_t ? g.bwds_o : g.bwds);
    if (!f) return hipErrorInvalidDeviceFunction;
    ActParams a = ap;
    void *params[] = {&Tb, &S, &Sb, &nrow, &Np, &a, &T0, &W1, &SbL, &WL, &nout};
    if (!in_s && !out_t) return gen_launch_bwds(g, Tb, S, Sb, nrow, Np, ap, s);
    return hipModuleLaunchKernel(f, (unsigned)(Np / 64), (unsigned)((nrow + 3) / 4), 1, 256, 1, 1, 0, s, params,
                                 nullptr);
}

hipError_t gen_launch_bwds_grad1(const GenKernels &g, bool out_t, const void *Tb, int nrow, long Np,
                                 const ActParams &ap, const void *T0, const void *W1, const void *SbL, const void *WL,
                                 int nout, double *part, hipStream_t s)
{
    hipFunction_t f = out_t ? g.bwds_iog : g.bwds_ig;
    if (!f) return hipErrorInvalidDeviceFunction;
    ActParams a = ap;
    const void *S = nullptr;
    void *Sb = nullptr;
    void *params[] = {&Tb, &S, &Sb, &nrow, &Np, &a, &T0, &W1, &SbL, &WL, &nout, &part};
    return hipModuleLaunchKernel(f, (unsigned)(Np / 64), (unsigned)((nrow + 3) / 4), 1, 256, 1, 1, 0, s, params,
                                 nullptr);
}

hipError_t gen_launch_bwds_gradout(const GenKernels &g, const void *S, void *Sb, int nrow, long Np, const ActParams &ap,
                                   const void *SbL, const void *WL, double *part, hipStream_t s)
{
    if (!g.bwds_og) return hipErrorInvalidDeviceFunction;
    ActParams a = ap;
    const void *Tb = nullptr, *T0 = nullptr, *W1 = nullptr;
    int nout = 1;
    void *params[] = {&Tb, &S, &Sb, &nrow, &Np, &a, &T0, &W1, &SbL, &WL, &nout, &part};
    return hipModuleLaunchKernel(g.bwds_og, (unsigned)(Np / 64), (unsigned)((nrow + 3) / 4), 1, 256, 1, 1, 0, s, params,
                                 nullptr);
}

hipError_t gen_launch_pt_fwd(const GenKernels &g, const double *W, const double *x, const int *ind_e1, const PtNet &p,
                             const ActParams &ap, hipStream_t s)
{
    ActParams a = ap;
    PtNet q = p;
    void *params[] = {&W, &x, &ind_e1, &q, &a};
    return hipModuleLaunchKernel(g.pt_fwd, 1, 1, 1, 256, 1, 1, 0, s, params, nullptr);
}

hipError_t gen_launch_pt_bwd(const GenKernels &g, const double *W, double *grad, const double *Sb, const PtNet &p,
                             const ActParams &ap, hipStream_t s)
{
    ActParams a = ap;
    PtNet q = p;
    void *params[] = {&W, &grad, &Sb, &q, &a};
    return hipModuleLaunchKernel(g.pt_bwd, 1, 1, 1, 256, 1, 1, 0, s, params, nullptr);
}

// ---------------------------------------------------------------------------
// Fused layer kernels (FP32 storage) for networks whose hidden layers are all
// FZ_W = 64 wide, with at least two hidden layers and one output row.
//
// Their planes are point-blocked: a block of FZ_P = 4 points holds, slot by
// slot and point by point, the 64 neurons' values ([block][slot][point]
// [neuron], na KB a block, contiguous; a wave's 64 lanes then touch four
// 64-byte runs in lane order, which streams at 6.7 TB/s where the neuron-
// outer order took 4.3), and one workgroup of four waves covers one
// block, every neuron and every slot.  Wave w, lane l owns neuron
// j = 16 w + (l mod 16) at point p = l / 16.
//   K1 (lib/train_module.f90:262-278), in the forward: the block of T below
//   is split on staging into three bf16 pieces (x = h + m + l, each the
//   round-to-nearest bf16 of what the previous ones leave) held as
//   X2[slot][piece][point][64 neurons], 128-byte rows whose 16-byte chunks
//   sit swizzled by the row (chunk c at c ^ (row mod 8): the eight lanes
//   of an A-operand read meet no bank twice); the weights are split the same
//   way into registers; v_mfma_f32_16x16x32_bf16 forms the six significant
//   products of a 32-term k-step, moved into FP64 after each k-step (bias at
//   slot 0): rows (point, slot), columns the 16 neurons of the wave, k 32
//   neurons of the block, leaving lane l with slots 4t .. 4t+3 of its
//   (neuron, point l / 16) -- the recurrence's operands, in registers, so S
//   never goes to memory.
//   K5 (:408-422), in the adjoint: rows the wave's neurons, columns 16
//   inputs, k the four points of the block, slot by slot on
//   v_mfma_f32_16x16x4f32 against the FP32 block of T below staged as
//   X[slot * FZ_XS + point * FZ_PJ + neuron]; the lane's own S-bar is the A
//   operand; FP32 over a block, FP64 across blocks (LDS accumulators),
//   per-workgroup FP64 partials summed in a fixed order by the gradient
//   reduction.
//   K6 (:430-444): S-bar transposed through X, rows (point, slot), columns
//   the wave's 16 inputs, k the 64 neurons, FP32 MFMA, two tiles at a time.
// The maps are emitted interleaved with the per-slot arithmetic: slot a of
// the recurrence needs S of its sub-indices only (earlier slots in the graded
// order), so the MFMAs of tile q + 1 go out before the recurrence of tile q's
// slots, and each slot's K5 right after its pair sum -- the matrix pipe runs
// while the wave issues the vector work.
// Forward: the first hidden layer forms S from the inputs (fwd_in, T
// written); every other hidden layer reads the block of T below, maps it and
// writes its own T and h (fwd); the last writes h only and forms the output
// row from T held in LDS, summed over the neurons in FP64, with the
// residual, the seed, the output layer's gradient partials and the sums of
// R^2 (fwd_last).  Adjoint: the last hidden layer takes T-bar = the output
// weight times the seed (bwd_last), the others read it (bwd); both read h,
// form the pair sums, the layer's K5 and T-bar of the layer below (K6); the
// first hidden layer (bwd_in) forms S and h again from the inputs (no linear
// map: cheaper than a plane of h) and its gradient from S-bar at slot 0 and
// the slots e_i, writing no plane.
static const char *kFusedHelpers = R"FZ(
// lane offsets pass through an empty asm at the top of every block: nothing
// derived from them is hoisted out of the block loop into live registers
#define FZ_OPAQUE(x) asm volatile("" : "+v"(x))
#define FZ_NU ((NA * 64 + 255) / 256)
// one FP32 block [slot][point][neuron] (global) into X[slot * XS + point *
// PJ + neuron], in two halves: the global loads into registers (issued a
// block ahead) and the LDS writes
__device__ __forceinline__ void fz_stage_load(const float *__restrict__ src, flt4 (&v)[FZ_NU])
{
  const int tid = threadIdx.x;
  int so = tid;
  FZ_OPAQUE(so);
#pragma unroll
  for (int k = 0; k < FZ_NU; ++k)
    if (tid + 256 * k < NA * 64) v[k] = reinterpret_cast<const flt4 *>(src)[so + 256 * k];
}
__device__ __forceinline__ void fz_stage_store(const flt4 (&v)[FZ_NU], float *__restrict__ X)
{
  const int tid = threadIdx.x;
  int xo = (tid >> 6) * XS + ((tid >> 4) & 3) * PJ + (tid & 15) * 4;
  FZ_OPAQUE(xo);
#pragma unroll
  for (int k = 0; k < FZ_NU; ++k)
    if (tid + 256 * k < NA * 64) *reinterpret_cast<flt4 *>(X + xo + k * 4 * XS) = v[k];
}
__device__ __forceinline__ void fz_zero_pad(float *__restrict__ X)
{
  for (int u = threadIdx.x; u < (4 * NT - NA) * XS; u += 256) X[NA * XS + u] = 0.f;
}
// bf16 pieces: a pair converts in one instruction and widens back by a
// shift and a mask
__device__ __forceinline__ flt2 fz_widen2(bf16x2 v)
{
  const unsigned b = __builtin_bit_cast(unsigned, v);
  return flt2{__builtin_bit_cast(float, b << 16), __builtin_bit_cast(float, b & 0xffff0000u)};
}
__device__ __forceinline__ void fz_split2(flt2 x, bf16x2 &h, bf16x2 &m, bf16x2 &l)
{
  h = __builtin_convertvector(x, bf16x2);
  const flt2 r = x - fz_widen2(h);
  m = __builtin_convertvector(r, bf16x2);
  l = __builtin_convertvector(r - fz_widen2(m), bf16x2);
}
// element (slot, piece, point, neuron col) of the split block: rows
// (slot * 3 + piece) * 4 + point of 64 bf16, the 16-byte chunk col / 8 at
// (col / 8) ^ (2 (slot mod 4) + point mod 2): the eight lanes of one
// A-operand read (four slots, two points) meet eight different chunks
__device__ __forceinline__ int fz_x2(int slot, int c, int pt, int col)
{
  return ((slot * 3 + c) * 4 + pt) * 64 + ((((col >> 3) ^ (((slot & 3) << 1) | (pt & 1))) << 3) | (col & 7));
}
// the FP32 block of T below into X2 as three bf16 pieces (row (slot * 3 +
// piece) * 4 + point); each thread four neurons of a (slot, point) at a time
__device__ __forceinline__ void fz_stage_x6(const float *__restrict__ src, __bf16 *__restrict__ X2)
{
  const int tid = threadIdx.x;
  int so = tid;
  FZ_OPAQUE(so);
  flt4 v[FZ_NU];
#pragma unroll
  for (int k = 0; k < FZ_NU; ++k)
    if (tid + 256 * k < NA * 64) v[k] = reinterpret_cast<const flt4 *>(src)[so + 256 * k];
  const int pt = (tid >> 4) & 3, col = (tid & 15) * 4;
#pragma unroll
  for (int k = 0; k < FZ_NU; ++k)
    if (tid + 256 * k < NA * 64) {
      const int slot = (tid >> 6) + 4 * k;
      bf16x2 h0, m0, l0, h1, m1, l1;
      fz_split2(flt2{v[k][0], v[k][1]}, h0, m0, l0);
      fz_split2(flt2{v[k][2], v[k][3]}, h1, m1, l1);
      *reinterpret_cast<bf16x4 *>(X2 + fz_x2(slot, 0, pt, col)) = bf16x4{h0[0], h0[1], h1[0], h1[1]};
      *reinterpret_cast<bf16x4 *>(X2 + fz_x2(slot, 1, pt, col)) = bf16x4{m0[0], m0[1], m1[0], m1[1]};
      *reinterpret_cast<bf16x4 *>(X2 + fz_x2(slot, 2, pt, col)) = bf16x4{l0[0], l0[1], l1[0], l1[1]};
    }
}
__device__ __forceinline__ void fz_zero_pad_x6(__bf16 *__restrict__ X2)
{
  for (int u = threadIdx.x; u < (4 * NT - NA) * XS2; u += 256) X2[NA * XS2 + u] = (__bf16)0.f;
}
// the lane's A-operand offset into X2 for tile 0, piece 0, k-step 0: row
// (point, slot) lane mod 16, k = 8 (lane / 16) .. +8.  Tile t adds 4 t slots
// (48 t rows; the same swizzle); the next k-step is chunk ^ 4 (offset ^ 32);
// piece c is 4 c rows on (+256 c; the same swizzle).
__device__ __forceinline__ int fz_x6_offset()
{
  const int lane = threadIdx.x & 63;
  return fz_x2(lane & 3, 0, (lane & 15) >> 2, 8 * (lane >> 4));
}
// the weight pieces of the lane's column (neuron j), k as above
__device__ __forceinline__ void fz_w_split(const float *__restrict__ wj, bf16x8 (&wh)[2], bf16x8 (&wm)[2],
                                           bf16x8 (&wl)[2])
{
  const int kq = 8 * ((threadIdx.x & 63) >> 4);
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
      bf16x2 h, m, l;
      fz_split2(flt2{wj[32 * ks + kq + e], wj[32 * ks + kq + e + 1]}, h, m, l);
      wh[ks][e] = h[0]; wh[ks][e + 1] = h[1];
      wm[ks][e] = m[0]; wm[ks][e + 1] = m[1];
      wl[ks][e] = l[0]; wl[ks][e + 1] = l[1];
    }
}
// K1 tile t: the six products of each 32-term k-step (small terms first)
// accumulated in FP32, then moved into FP64; slot 0 takes the bias
__device__ __forceinline__ flt4 fz_tile_x6(const __bf16 *__restrict__ X2, int xb0, int xb1, int t,
                                           const bf16x8 (&wh)[2], const bf16x8 (&wm)[2], const bf16x8 (&wl)[2],
                                           float bias)
{
  double d[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    const int x0 = ks ? xb1 : xb0;
    const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(X2 + x0 + t * 48 * 64);
    const bf16x8 am = *reinterpret_cast<const bf16x8 *>(X2 + x0 + 256 + t * 48 * 64);
    const bf16x8 al = *reinterpret_cast<const bf16x8 *>(X2 + x0 + 512 + t * 48 * 64);
    flt4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, wh[ks], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, wl[ks], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, wm[ks], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, wh[ks], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, wm[ks], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, wh[ks], c, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) d[r] += (double)c[r];
  }
  flt4 out;
#pragma unroll
  for (int r = 0; r < 4; ++r) out[r] = (float)(d[r] + ((t == 0 && r == 0) ? (double)bias : 0.0));
  return out;
}
// K6 tile t: rows (point (lane mod 16) / 4, slot 4 t + lane mod 4), k the
// input lane / 16 + 4 s at step s; xa = the lane's offset for t = 0
__device__ __forceinline__ flt4 fz_tile(const float *__restrict__ X, int xa, int t, const float (&b)[16])
{
  flt4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 16; ++s) c = __builtin_amdgcn_mfma_f32_16x16x4f32(X[xa + t * 4 * XS + 4 * s], b[s], c, 0, 0, 0);
  return c;
}
// two tiles at once: two independent accumulator chains for the MFMA pipe
__device__ __forceinline__ void fz_tile2(const float *__restrict__ X, int xa, int t0, const float (&b)[16], flt4 &c0,
                                         flt4 &c1)
{
  c0 = flt4{0.f, 0.f, 0.f, 0.f};
  c1 = c0;
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(X[xa + t0 * 4 * XS + 4 * s], b[s], c0, 0, 0, 0);
    c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(X[xa + (t0 + 1) * 4 * XS + 4 * s], b[s], c1, 0, 0, 0);
  }
}
// sum over the lanes l, l ^ 16, l ^ 32, l ^ 48 (the four points of a neuron)
__device__ __forceinline__ double fz_sum4(double v)
{
  for (int o = 16; o <= 32; o <<= 1) {
    const int src = (int)(((threadIdx.x & 63) ^ o) << 2);
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_ds_bpermute(src, (int)(unsigned)b);
    const int hi = __builtin_amdgcn_ds_bpermute(src, (int)(unsigned)(b >> 32));
    v += __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
  }
  return v;
}
// K5 of one slot: g<ib> += Sbar_a (row: the lane's neuron, k: its point) x
// T_a of the block below (k: point, column: input 16 ib + lane mod 16);
// xg = (lane / 16) PJ + lane mod 16
#define FZ_G(a, sbv)                                                                               \
  {                                                                                                \
    g0 = __builtin_amdgcn_mfma_f32_16x16x4f32((sbv), X[xg + (a) * XS], g0, 0, 0, 0);               \
    g1 = __builtin_amdgcn_mfma_f32_16x16x4f32((sbv), X[xg + (a) * XS + 16], g1, 0, 0, 0);          \
    g2 = __builtin_amdgcn_mfma_f32_16x16x4f32((sbv), X[xg + (a) * XS + 32], g2, 0, 0, 0);          \
    g3 = __builtin_amdgcn_mfma_f32_16x16x4f32((sbv), X[xg + (a) * XS + 48], g3, 0, 0, 0);          \
  }
)FZ";

bool gen_fused_usable(const TableSet &t, const FwdSchedule &s)
{
    return gen_bell_usable(t, s) && t.na <= FZ_MAXA && t.d0 <= FZ_MAXD0;
}

std::string gen_fused_source(const TableSet &t, const FwdSchedule &s)
{
    std::ostringstream o;
    const int na = t.na, kq = kq_of(t.k), d0 = t.d0, nt = (na + 3) / 4;
    const bool bes = t.iact == ACT_BESSEL_J0 || t.iact == ACT_BESSEL_J1;
    // the activation, its derivative count and the polynomial row length
    // are the family's: compile-time here (the argument layout is the
    // host's ActParams), so one activation's code is compiled and nothing of
    // it is hoisted out of the block loops.  Multiply-adds are fused inside
    // an expression only (each the same way wherever it is emitted).
    o << "#pragma clang fp contract(on)\n"
      << "struct ActParams { int iact_rt, nq_rt; const double *tpc; int ldt_rt; static constexpr int iact = "
      << t.iact << ", nq = " << t.k + 1 << ", ldt = " << t.k + 3 << "; };\n"
      << kActSrc << "\n"
      << "typedef float real;\n"
      << "typedef float flt4 __attribute__((ext_vector_type(4)));\n"
      << "typedef float flt2 __attribute__((ext_vector_type(2)));\n"
      << "typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));\n"
      << "typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));\n"
      << "typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));\n"
      << "#define NA " << na << "\n#define NT " << nt << "\n#define D0 " << d0 << "\n#define XS " << FZ_XS
      << "\n#define PJ " << FZ_PJ << "\n#define BLK (NA * 256)\n"
      // the split block: 3 pieces x 4 points x 64 bf16 a slot; the last
      // hidden layer's T (FP32, rows of PJ) takes slot a's place once read
      << "#define XS2 (3 * 4 * 64)\n#define XT (XS2 / 2)\n"
      << kFusedHelpers;
    auto derivs = [&]() {
        // the table pointer passes through an empty asm in every block: its
        // loads are not hoisted out of the loop into live registers
        o << "  double d[" << kq << "];\n"
          << "  ActParams apl = ap;\n  asm volatile(\"\" : \"+s\"(apl.tpc));\n"
          << "  apl.tpc = (const double *)(((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)apl.tpc >> 32)) << 32) | (unsigned)__builtin_amdgcn_readfirstlane((unsigned)(unsigned long long)apl.tpc));\n"
          << "  act_derivs<" << kq << ", " << (bes ? "true" : "false") << ">(apl, (double)e0, d);\n";
        for (int q = 0; q < kq; ++q) o << "  const real d" << q << " = (real)d[" << q << "];\n";
    };
    // S of the lane from the staged block of T below (K1): tiles 0 and 1 and
    // the activation derivatives here, tile q + 1 before the recurrence
    // reaches tile q's slots (tile_hook)
    auto tile = [&](int q) {
        o << "  {\n    const flt4 sv = fz_tile_x6(X2, xb0, xb1, " << q << ", wh, wm, wl, bias);\n";
        for (int r = 0; r < 4 && 4 * q + r < na; ++r) o << "    e" << 4 * q + r << " = sv[" << r << "];\n";
        o << "  }\n";
    };
    auto s_from_x = [&]() {
        o << "  bf16x8 wh[2], wm[2], wl[2];\n  fz_w_split(W + j * 65 + 1, wh, wm, wl);\n"
          << "  const int xb0 = fz_x6_offset(), xb1 = xb0 ^ 32;\n";
        for (int a = 0; a < na; ++a) o << "  real e" << a << ";\n";
        tile(0);
        if (nt > 1) tile(1);
        derivs();
    };
    const std::function<void(int)> tile_hook = [&](int a) {
        if (a % 4 == 0 && a / 4 + 1 < nt) tile(a / 4 + 1);
    };
    // the same with a barrier after each tile: tile q reads the staged
    // slots 4q .. 4q+3 only, so once every wave has passed the barrier after
    // tile q + 1 the slots up to 4q + 7 may take the layer's own T (the last
    // hidden layer's output row reads it from there)
    const std::function<void(int)> tile_hook_sync = [&](int a) {
        if (a % 4 == 0 && a / 4 + 1 < nt) {
            tile(a / 4 + 1);
            o << "  __syncthreads();\n";
        }
    };
    // S of the lane from the inputs (the FP32 input values, S_0 summed in FP64)
    auto s_from_in = [&]() {
        o << "  const long n = (long)pb * 4 + p;\n"
          << "  const float *wj = W1 + j * " << d0 + 1 << ";\n"
          << "  double s0 = (double)wj[0];\n";
        for (int i = 0; i < d0; ++i)
            o << "  const float xv" << i << " = n < npt ? (float)xin[n * " << d0 << " + " << i << "] : 0.f;\n"
              << "  s0 += (double)wj[" << i + 1 << "] * (double)xv" << i << ";\n";
        o << "  const real e0 = (real)s0;\n";
        for (int a = 1; a < na; ++a) {
            int src = -1;
            for (int i = 0; i < d0; ++i)
                if (t.ind_e1[i] == a) src = i;
            if (src >= 0) o << "  const real e" << a << " = wj[" << src + 1 << "];\n";
            else o << "  const real e" << a << " = (real)0;\n";
        }
        derivs();
    };
    auto head = [&](const char *name, const char *args, int waves_per_eu) {
        o << "extern \"C\" __global__ __launch_bounds__(256, " << waves_per_eu << ") void " << name << "(" << args
          << ")\n{\n"
          << "  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;\n"
          << "  const int j = 16 * w + (lane & 15), p = lane >> 4;\n"
          << "  const long Np = 256;   // slot stride inside a block\n"
          // lane offsets: K6 A-operand row (xa), K5 B operand (xg), the
          // transposed S-bar row and the last layer's T in LDS (xs), the
          // (neuron, point) in a block (lo), the weight rows of K6's B
          // operand (wv)
          << "  int xa = (lane & 3) * XS + ((lane & 15) >> 2) * PJ + (lane >> 4);\n"
          << "  int xg = (lane >> 4) * PJ + (lane & 15), xs = p * PJ + j, lo = p * 64 + j;\n"
          << "  int wv = (lane >> 4) * 65 + 1 + j;\n";
    };
    auto opaque = [&]() {
        o << "  FZ_OPAQUE(xa);\n  FZ_OPAQUE(xg);\n  FZ_OPAQUE(xs);\n  FZ_OPAQUE(lo);\n  FZ_OPAQUE(wv);\n";
    };
    auto xbuf = [&]() {
        o << "  __shared__ __attribute__((aligned(16))) float X[4 * NT * XS];\n"
          << "  fz_zero_pad(X);\n";
    };
    // K5 of the block into the workgroup's FP64 accumulators in LDS (each
    // lane owns 16 of them: rows 16 w + 4 (lane / 16) + r, columns
    // 16 ib + lane mod 16), FP32 within the block; its MFMAs go out slot by
    // slot right after each pair sum (k5_hook)
    auto k5_begin = [&]() { o << "  flt4 g0 = {0.f, 0.f, 0.f, 0.f}, g1 = g0, g2 = g0, g3 = g0;\n"; };
    const std::function<void(int)> k5_hook = [&](int b) { o << "  FZ_G(" << b << ", sb" << b << ")\n"; };
    auto k5_end = [&]() {
        o << "  {\n    double *gl = Gs + (16 * w + 4 * (lane >> 4)) * 64 + (lane & 15);\n"
          << "#pragma unroll\n    for (int r = 0; r < 4; ++r) {\n"
          << "      gl[r * 64] += (double)g0[r];\n      gl[r * 64 + 16] += (double)g1[r];\n"
          << "      gl[r * 64 + 32] += (double)g2[r];\n      gl[r * 64 + 48] += (double)g3[r];\n    }\n"
          << "    gb += (double)sb0;\n  }\n";
    };
    auto k5_out = [&]() {
        o << "  gb = fz_sum4(gb);\n"
          << "  double *pw = part + (long)blockIdx.x * (64 * 65);\n"
          << "  if (lane < 16) pw[j * 65] = gb;\n"
          << "  {\n    const double *gl = Gs + (16 * w + 4 * (lane >> 4)) * 64 + (lane & 15);\n"
          << "#pragma unroll\n    for (int ib = 0; ib < 4; ++ib)\n#pragma unroll\n      for (int r = 0; r < 4; ++r)\n"
          << "        pw[(16 * w + 4 * (lane >> 4) + r) * 65 + 1 + 16 * ib + (lane & 15)] = gl[r * 64 + 16 * ib];\n  }\n";
    };
    auto decl_h = [&]() {
        o << "  const real h0 = d1;\n";
        for (int a = 1; a < na; ++a) o << "  real h" << a << ";\n";
    };
    auto decl_sb = [&]() {
        for (int a = 0; a < na; ++a) o << "  real sb" << a << ";\n";
    };

    // ---- forward, first hidden layer
    head("dnnf90_fz_fwd_in",
         "const double *__restrict__ xin, int npt, const float *__restrict__ W1, float *__restrict__ Tout, int nblk, "
         "ActParams ap",
         1);
    o << "  const int pb = blockIdx.x;\n  if (pb >= nblk) return;\n";
    s_from_in();
    o << "  real *Tj = Tout + (long)pb * BLK + lo;\n  Tj[0] = d0;\n";
    emit_recurrence(o, t, s, false, REC_T);
    o << "}\n";

    // ---- forward, other hidden layers; one block a workgroup.  The last
    // hidden layer also forms the output row (summed over the neurons in
    // FP64 from T held in the staged block's place in LDS), the residual of
    // each point and the seed S-bar of the output row (k_seed_residual's
    // arithmetic, misc.hip; app/pinn_module.f90:32-65), the output layer's
    // weight-gradient partials of the block (bias first) and its sum of R^2:
    // no output plane, no seed kernel, no T plane of that layer.
    o << "struct DevResidual { int nterm; const double *coeff; const int *slot; const int *nonlin; int distinct; };\n";
    for (int last = 0; last < 2; ++last) {
        if (last)
            head("dnnf90_fz_fwd_last",
                 "const float *__restrict__ Tin, const float *__restrict__ W, int nblk, ActParams ap, "
                 "const float *__restrict__ WL, int npt, DevResidual rs, double fac, float *__restrict__ Sbo, "
                 "double *__restrict__ gpart, double *__restrict__ rpart, double *__restrict__ tout, "
                 "float *__restrict__ Hout",
                 2);
        else
            head("dnnf90_fz_fwd",
                 "const float *__restrict__ Tin, const float *__restrict__ W, float *__restrict__ Tout, int nblk, "
                 "ActParams ap, float *__restrict__ Hout",
                 2);
        o << "  __shared__ __attribute__((aligned(16))) __bf16 X2[4 * NT * XS2];\n"
          << "  fz_zero_pad_x6(X2);\n";
        o << "  const int pb = blockIdx.x;\n  if (pb >= nblk) return;\n";
        if (last)
            // the residual's terms and the output weights in LDS, loaded
            // beside the staged block (the output section reads them many
            // times a block)
            o << "  float *Xf = reinterpret_cast<float *>(X2);\n"
              << "#define FZ_OUT(a, tv) Xf[xs + (a) * XT] = (tv);\n"
              << "  __shared__ double Rc[" << FZ_MAXT << "];\n  __shared__ int Rsl[" << FZ_MAXT << "], Rnl["
              << FZ_MAXT << "];\n"
              << "  if (tid < rs.nterm) {\n    Rc[tid] = rs.coeff[tid];\n    Rsl[tid] = rs.slot[tid];\n"
              << "    Rnl[tid] = rs.nonlin[tid];\n  }\n"
              << "  __shared__ float WLs[64];\n  if (tid < 64) WLs[tid] = WL[1 + tid];\n";
        o << "  const float bias = W[j * 65];\n"
          << "  fz_stage_x6(Tin + (long)pb * BLK, X2);\n  __syncthreads();\n";
        s_from_x();
        // the last hidden layer's T goes to the output row only, through the
        // staged block's place in LDS, each slot once every wave has read it
        // (the tiles keep their place in the recurrence, with a barrier each)
        if (last) {
            o << "  __syncthreads();\n  FZ_OUT(0, d0);\n";
        } else {
            o << "  real *Tj = Tout + (long)pb * BLK + lo;\n  Tj[0] = d0;\n";
        }
        o << "  real *Hj = Hout + (long)pb * BLK + lo;\n  Hj[0] = d1;\n";
        emit_recurrence(o, t, s, false, (last ? REC_TOUT : REC_T) | REC_H, "Np", last ? tile_hook_sync : tile_hook);
        if (last)
            o << "#undef FZ_OUT\n"
              << "  __shared__ double Ot[NA * 4];   // the output row, FP32-rounded, (slot, point)\n"
              << "  __shared__ double Rq[4];        // R of the block's points\n"
              << "  __shared__ float Sz[NA * 4];    // the seed\n"
              << "  __shared__ double Op[4 * NA * 4 > 256 ? 4 * NA * 4 : 256];   // per-wave partial sums\n"
              << "  __syncthreads();\n"
              // wave w sums its 16 neurons of each (slot, point), the four
              // partials meet in wave order
              << "#pragma unroll\n"
              << "  for (int k = 0; k < 3; ++k) {\n"
              << "    const int u = lane + 64 * k;\n"
              << "    if (u < NA * 4) {\n"
              << "      const float *xu = Xf + (u >> 2) * XT + (u & 3) * PJ + 16 * w;\n"
              << "      double v = 0.0;\n"
              << "#pragma unroll\n"
              << "      for (int m = 0; m < 16; ++m) v += (double)WLs[16 * w + m] * (double)xu[m];\n"
              << "      Op[w * (NA * 4) + u] = v;\n"
              << "    }\n"
              << "  }\n"
              << "  __syncthreads();\n"
              << "  if (tid < NA * 4) {\n"
              << "    const double v = ((Op[tid] + Op[NA * 4 + tid]) + Op[2 * NA * 4 + tid]) + Op[3 * NA * 4 + tid];\n"
              << "    const double ov = (double)(float)(v + (tid < 4 ? (double)WL[0] : 0.0));\n"
              << "    Ot[tid] = ov;\n"
              << "    Sz[tid] = 0.f;\n"
              << "    const long n = (long)pb * 4 + (tid & 3);\n"
              << "    if (n < npt) tout[n * NA + (tid >> 2)] = ov;   // [point][slot], as the forward returns it\n"
              << "  }\n"
              << "  __syncthreads();\n"
              // one lane a point: R, then the seed of each term's slot in
              // the terms' order (k_seed_residual's arithmetic), slot 0's
              // nonlinear part last
              << "  if (tid < 4) {\n"
              << "    double R = 0.0;\n"
              << "    if ((long)pb * 4 + tid < npt) {\n"
              << "      const double t0 = Ot[tid];\n"
              << "      for (int k = 0; k < rs.nterm; ++k) {\n"
              << "        const double tv = Ot[Rsl[k] * 4 + tid];\n"
              << "        R += Rnl[k] ? Rc[k] * t0 * tv : Rc[k] * tv;\n"
              << "      }\n"
              << "      double s0 = 0.0;\n"
              << "      for (int k = 0; k < rs.nterm; ++k) {\n"
              << "        const double add = Rnl[k] ? fac * R * Rc[k] * t0 : fac * R * Rc[k];\n"
              << "        if (Rnl[k]) s0 += fac * R * Rc[k] * Ot[Rsl[k] * 4 + tid];\n"
              << "        float &sv = Sz[Rsl[k] * 4 + tid];\n"
              << "        sv = rs.distinct ? (float)(0.0 + add) : (float)((double)sv + add);\n"
              << "      }\n"
              << "      Sz[tid] = rs.distinct ? (float)(0.0 + s0) : (float)((double)Sz[tid] + s0);\n"
              << "    }\n"
              << "    Rq[tid] = R;\n"
              << "  }\n"
              << "  __syncthreads();\n"
              << "  if (tid < NA * 4) Sbo[(long)pb * (NA * 4) + tid] = Sz[tid];\n"
              // wave w sums slots [w NA, (w + 1) NA) of the (slot, point)
              // pairs for every neuron, the partials meet in wave order
              << "  {\n"
              << "    double g = 0.0;\n"
              << "#pragma unroll\n"
              << "    for (int i = 0; i < NA; ++i) {\n"
              << "      const int u = w * NA + i;\n"
              << "      g += (double)Sz[u] * (double)Xf[(u >> 2) * XT + (u & 3) * PJ + lane];\n"
              << "    }\n"
              << "    Op[w * 64 + lane] = g;\n"
              << "  }\n"
              << "  __syncthreads();\n"
              << "  if (tid < 64) {\n"
              << "    gpart[(long)pb * 65 + 1 + tid] = ((Op[tid] + Op[64 + tid]) + Op[128 + tid]) + Op[192 + tid];\n"
              << "  } else if (tid == 64) {\n"
              << "    gpart[(long)pb * 65] = (double)Sz[0] + (double)Sz[1] + (double)Sz[2] + (double)Sz[3];\n"
              << "  } else if (tid == 65) {\n"
              << "    rpart[pb] = Rq[0] * Rq[0] + Rq[1] * Rq[1] + Rq[2] * Rq[2] + Rq[3] * Rq[3];\n"
              << "  }\n";
        o << "}\n";
    }

    // ---- adjoint of a hidden layer l >= 2 (last: l = L-1, T-bar = the
    // output weight times the output row's seed; else T-bar of l read from
    // the plane the layer above wrote).  Per block: h read, the pair sums
    // into S-bar (registers), the layer's K5 against the staged block of T
    // below, then S-bar transposed into LDS and T-bar of l-1 = W' S-bar (K6)
    // written out for the layer below.  Each block's staged block below, h
    // and the seed are loaded into registers during the block before it (h
    // and the seed before its pair sums, the block below before its K6);
    // T-bar of the layer above is loaded at the top of the block (both
    // planes a block ahead would not fit the registers).
    for (int last = 0; last < 2; ++last) {
        head(last ? "dnnf90_fz_bwd_last" : "dnnf90_fz_bwd",
             "const float *__restrict__ Tin, const float *__restrict__ Tb, const float *__restrict__ W, "
             "const float *__restrict__ WL, float *__restrict__ Tbout, double *__restrict__ part, int nblk, "
             "ActParams ap, const float *__restrict__ H",
             2);
        xbuf();
        o << "  __shared__ double Gs[64 * 64];\n"
          << "  for (int u = tid; u < 64 * 64; u += 256) Gs[u] = 0.0;\n"
          << "  double gb = 0.0;\n";
        if (last) o << "  const real wl = WL[1 + j];\n  __shared__ float Sz[NA * 4];   // the block's output-row seed\n";
        auto next_ops = [&](const char *pbv) {
            o << "    const real *hp = H + (long)" << pbv << " * BLK + lo;\n";
            for (int a = 0; a < na; ++a) o << "    hn" << a << " = hp[" << a * 256 << "];\n";
            if (last) o << "    szn = tid < NA * 4 ? Tb[(long)" << pbv << " * (NA * 4) + tid] : 0.f;\n";
        };
        o << "  flt4 xn[FZ_NU];\n";
        for (int a = 0; a < na; ++a) o << "  real hn" << a << ";\n";
        if (last) o << "  float szn;\n";
        o << "  if ((int)blockIdx.x < nblk) {\n    fz_stage_load(Tin + (long)blockIdx.x * BLK, xn);\n";
        next_ops("blockIdx.x");
        o << "  }\n";
        o << "  for (int pb = blockIdx.x; pb < nblk; pb += gridDim.x) {\n";
        opaque();
        o << "  const int pn = pb + gridDim.x;\n"
          << "  __syncthreads();\n  fz_stage_store(xn, X);\n";
        if (last) o << "  if (tid < NA * 4) Sz[tid] = szn;\n";
        if (!last) {
            o << "  const real *tb = Tb + (long)pb * BLK + lo;\n";
            for (int a = 0; a < na; ++a) o << "  const real t" << a << " = tb[" << a * 256 << "];\n";
        }
        for (int a = 0; a < na; ++a) o << "  const real h" << a << " = hn" << a << ";\n";
        o << "  if (pn < nblk) {\n";
        next_ops("pn");
        o << "  }\n  __syncthreads();\n";
        o << "  __builtin_amdgcn_sched_barrier(0);\n";
        if (last)
            for (int a = 0; a < na; ++a) o << "  const real t" << a << " = wl * Sz[" << a * 4 << " + p];\n";
        decl_sb();
        k5_begin();
        emit_pair_sums(o, t, false, true, k5_hook);
        k5_end();
        o << "  if (pn < nblk) fz_stage_load(Tin + (long)pn * BLK, xn);\n";
        // xa once more through the asm: K6 forms its own LDS addresses
        o << "  __syncthreads();\n  {\n";
        for (int a = 0; a < na; ++a) o << "    X[xs + " << a << " * XS] = sb" << a << ";\n";
        o << "  }\n  __syncthreads();\n  FZ_OPAQUE(xa);\n  {\n    float v[16];\n"
          << "#pragma unroll\n    for (int s = 0; s < 16; ++s) v[s] = W[wv + 4 * s * 65];\n"
          << "    real *tbo = Tbout + (long)pb * BLK + lo;\n";
        for (int q = 0; q < nt; ++q) {
            if (q + 1 < nt) {
                o << "    {\n      flt4 c, c1;\n      fz_tile2(X, xa, " << q << ", v, c, c1);\n";
                for (int r = 0; r < 4 && 4 * q + r < na; ++r)
                    o << "      tbo[" << (4 * q + r) * 256 << "] = c[" << r << "];\n";
                for (int r = 0; r < 4 && 4 * q + 4 + r < na; ++r)
                    o << "      tbo[" << (4 * q + 4 + r) * 256 << "] = c1[" << r << "];\n";
                o << "    }\n    __builtin_amdgcn_sched_barrier(0);\n";
                ++q;
                continue;
            }
            o << "    {\n      const flt4 c = fz_tile(X, xa, " << q << ", v);\n";
            for (int r = 0; r < 4 && 4 * q + r < na; ++r) o << "      tbo[" << (4 * q + r) * 256 << "] = c[" << r << "];\n";
            o << "    }\n    __builtin_amdgcn_sched_barrier(0);\n";
        }
        o << "  }\n  }\n";
        k5_out();
        o << "}\n";
    }

    // ---- adjoint, first hidden layer: S from the inputs, T-bar from the
    // plane layer 2 wrote, the layer's gradient from S-bar at slot 0 and the
    // slots e_i (no plane written)
    head("dnnf90_fz_bwd_in",
         "const double *__restrict__ xin, int npt, const float *__restrict__ Tb, const float *__restrict__ W1, "
         "double *__restrict__ part, int nblk, ActParams ap",
         2);
    o << "  double g[" << d0 + 1 << "];\n";
    for (int c = 0; c <= d0; ++c) o << "  g[" << c << "] = 0.0;\n";
    o << "  for (int pb = blockIdx.x; pb < nblk; pb += gridDim.x) {\n  FZ_OPAQUE(lo);\n";
    s_from_in();
    decl_h();
    emit_recurrence(o, t, s, false, REC_HREG);
    o << "  const real *tb = Tb + (long)pb * BLK + lo;\n";
    for (int a = 0; a < na; ++a) o << "  const real t" << a << " = tb[" << a * 256 << "];\n";
    o << "  real sbar0;\n  {\n    real v = t0 * h0;\n";
    for (int a = 1; a < na; ++a) o << "    v += t" << a << " * h" << a << ";\n";
    o << "    sbar0 = v;\n  }\n"
      << "  g[0] += (double)sbar0;\n";
    for (int i = 0; i < d0; ++i) {
        const int bb = t.ind_e1[i];
        o << "  {\n    double gi = (double)sbar0 * (double)xv" << i << ";\n";
        if (bb >= 0) {
            o << "    real v = (real)0;\n";
            for (int ip = t.pair_start[bb]; ip < t.pair_start[bb] + t.pair_num[bb]; ++ip)
                o << "    v += " << lit(t.pair_binom[ip], false) << " * t" << t.pair_ia[ip] << " * h"
                  << t.pair_id[ip] << ";\n";
            o << "    gi += (double)v;\n";
        }
        o << "    g[" << i + 1 << "] += gi;\n  }\n";
    }
    o << "  }\n";
    for (int c = 0; c <= d0; ++c) o << "  g[" << c << "] = fz_sum4(g[" << c << "]);\n";
    o << "  if (lane < 16) {\n    double *pj = part + ((long)blockIdx.x * 64 + j) * " << d0 + 1 << ";\n";
    for (int c = 0; c <= d0; ++c) o << "    pj[" << c << "] = g[" << c << "];\n";
    o << "  }\n}\n";
    return o.str();
}

int gen_fused_compile(const TableSet &t, const FwdSchedule &s, FusedKernels *f, std::string *log)
{
    *f = FusedKernels();
    std::vector<char> code;
    if (rtc_compile(gen_fused_source(t, s), "dnnf90_fz_gen.hip", &code, log)) {
        f->status = -2;
        return -3;
    }
    return gen_fused_load(code, f);
}

int gen_fused_load(const std::vector<char> &code, FusedKernels *f)
{
    *f = FusedKernels();
    f->code = code;
    static const char *const nm[FusedKernels::NK] = {"dnnf90_fz_fwd_in",   "dnnf90_fz_fwd", "dnnf90_fz_fwd_last",
                                                      "dnnf90_fz_bwd_last", "dnnf90_fz_bwd", "dnnf90_fz_bwd_in"};
    int st = hipModuleLoadData(&f->mod, f->code.data()) == hipSuccess ? 1 : -3;
    for (int i = 0; st == 1 && i < FusedKernels::NK; ++i) {
        if (hipModuleGetFunction(&f->fn[i], f->mod, nm[i]) != hipSuccess ||
            hipFuncGetAttribute(&f->regs[i], HIP_FUNC_ATTRIBUTE_NUM_REGS, f->fn[i]) != hipSuccess ||
            hipFuncGetAttribute(&f->scratch[i], HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, f->fn[i]) != hipSuccess)
            st = -3;
        // a few loop-invariant values may live in scratch (reloaded a few
        // times a block); a kernel spilling more is not used (the unfused
        // kernels take the step)
        else if (f->scratch[i] > FZ_MAX_SCRATCH)
            st = -4;
    }
    // the adjoint kernels stream their blocks in one resident round
    int occ = 0, dev = 0, ncu = 0;
    if (st == 1 &&
        !(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, f->fn[FusedKernels::BWD], 256, 0) == hipSuccess &&
          hipGetDevice(&dev) == hipSuccess &&
          hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && occ > 0))
        st = -5;
    if (st != 1) {
        (void)hipGetLastError();
        f->release();   // the registers and scratch read stay for the report
        f->status = st;
        return -3;
    }
    f->grid = occ * ncu;
    f->status = 1;
    f->ok = true;
    return 0;
}

void FusedKernels::release()
{
    if (mod) (void)hipModuleUnload(mod);
    mod = nullptr;
    for (auto &k : fn) k = nullptr;
    ok = false;
}

static hipError_t fz_launch(const FusedKernels &f, int role, unsigned grid, void **params, hipStream_t s)
{
    return hipModuleLaunchKernel(f.fn[role], grid, 1, 1, 256, 1, 1, 0, s, params, nullptr);
}

hipError_t fz_launch_fwd_in(const FusedKernels &f, const double *x, int npt, const float *W1, float *Tout, int nblk,
                            const ActParams &ap, hipStream_t s)
{
    ActParams a = ap;
    void *params[] = {&x, &npt, &W1, &Tout, &nblk, &a};
    return fz_launch(f, FusedKernels::FWD_IN, (unsigned)nblk, params, s);
}

hipError_t fz_launch_fwd(const FusedKernels &f, const float *Tin, const float *W, float *Tout, float *H, int nblk,
                         const ActParams &ap, hipStream_t s)
{
    ActParams a = ap;
    void *params[] = {&Tin, &W, &Tout, &nblk, &a, &H};
    return fz_launch(f, FusedKernels::FWD, (unsigned)nblk, params, s);
}

hipError_t fz_launch_fwd_last(const FusedKernels &f, const float *Tin, const float *W, float *H, int nblk,
                              const ActParams &ap, const float *WL, int npt, const DevResidual &r, double fac,
                              float *Sbo, double *gpart, double *rpart, double *tout, hipStream_t s)
{
    ActParams a = ap;
    DevResidual rr = r;
    void *params[] = {&Tin, &W, &nblk, &a, &WL, &npt, &rr, &fac, &Sbo, &gpart, &rpart, &tout, &H};
    return fz_launch(f, FusedKernels::FWD_LAST, (unsigned)nblk, params, s);
}

hipError_t fz_launch_bwd(const FusedKernels &f, bool last, const float *Tin, const float *Tb, const float *H,
                         const float *W, const float *WL, float *Tbout, double *part, int nblk, const ActParams &ap,
                         int grid, hipStream_t s)
{
    ActParams a = ap;
    void *params[] = {&Tin, &Tb, &W, &WL, &Tbout, &part, &nblk, &a, &H};
    return fz_launch(f, last ? FusedKernels::BWD_LAST : FusedKernels::BWD, (unsigned)grid, params, s);
}

hipError_t fz_launch_bwd_in(const FusedKernels &f, const double *x, int npt, const float *Tb, const float *W1,
                            double *part, int nblk, const ActParams &ap, int grid, hipStream_t s)
{
    ActParams a = ap;
    void *params[] = {&x, &npt, &Tb, &W1, &part, &nblk, &a};
    return fz_launch(f, FusedKernels::BWD_IN, (unsigned)grid, params, s);
}

}  // namespace dnnf90
