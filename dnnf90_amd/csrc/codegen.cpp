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
// into gdot.
enum { REC_T = 1, REC_H = 2, REC_HREG = 4, REC_TDOT = 8 };

static void emit_recurrence(std::ostringstream &o, const TableSet &t, const FwdSchedule &s, bool fp64, int m)
{
    const bool wt = (m & (REC_T | REC_TDOT)) != 0, wh = (m & (REC_H | REC_HREG)) != 0;
    for (int e = t.na; e < s.nentry; ++e) o << "  real e" << e << ";\n";
    for (int a = 1; a < t.na; ++a) {
        const int g0 = s.slot_group[a], g1 = s.slot_group[a + 1];
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
        if (m & REC_T) o << "    Tj[" << a << "L * Np] = t;\n";
        if (m & REC_H) o << "    Hj[" << a << "L * Np] = h;\n";
        if (m & REC_HREG) o << "    h" << a << " = h;\n";
        if (m & REC_TDOT) o << "    gdot += t * slz[" << a << "L * Np];\n";
        o << "  }\n";
    }
}

// Sbar_0 = sum_a Tbar_a h_a and the pair sums Sbar_b = sum binom Tbar_ia h_id
// (lib/multi_index_bell_module.f90:267-297) from registers t<a>, h<a>
static void emit_pair_sums(std::ostringstream &o, const TableSet &t, bool fp64)
{
    o << "  {\n    real v = t0 * h0;\n";
    for (int a = 1; a < t.na; ++a) o << "    v += t" << a << " * h" << a << ";\n";
    o << "    sb[0] = v;\n  }\n";
    for (int b = 1; b < t.na; ++b) {
        o << "  {\n    real v = (real)0;\n";
        for (int ip = t.pair_start[b]; ip < t.pair_start[b] + t.pair_num[b]; ++ip)
            o << "    v += " << lit(t.pair_binom[ip], fp64) << " * t" << t.pair_ia[ip] << " * h" << t.pair_id[ip]
              << ";\n";
        o << "    sb[" << b << "L * Np] = v;\n  }\n";
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

int gen_bell_compile(const TableSet &t, const FwdSchedule &s, bool fp64, GenKernels *g, std::string *log)
{
    *g = GenKernels();
    const std::string src = gen_bell_source(t, s, fp64);
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "dnnf90_bell_gen.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        return -3;
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
    std::vector<char> code(sz);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
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

}  // namespace dnnf90
