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

std::string gen_bell_source(const TableSet &t, const FwdSchedule &s, bool fp64)
{
    std::ostringstream o;
    const int na = t.na, kq = kq_of(t.k);
    const bool bes = t.iact == ACT_BESSEL_J0 || t.iact == ACT_BESSEL_J1;
    // hipRTC supplies the HIP device environment itself (no runtime header)
    o << "struct ActParams { int iact, nq; const double *tpc; int ldt; };\n"
      << kActSrc << "\n"
      << "typedef " << (fp64 ? "double" : "float") << " real;\n";

    // ---- forward: T_a and h_a of every slot from the S_a of one lane
    const char *sfx = fp64 ? "_f64" : "_f32";
    o << "extern \"C\" __global__ __launch_bounds__(256) void dnnf90_bell_fwd" << sfx << "(const real *__restrict__ S, "
         "real *__restrict__ Tout, real *__restrict__ H, int nrow, long Np, ActParams ap)\n{\n"
      << "  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;\n"
      << "  const int j = blockIdx.y * (blockDim.x >> 6) + wave;\n"
      << "  if (j >= nrow) return;\n"
      << "  const long n = (long)blockIdx.x * 64 + lane;\n"
      << "  const long ldr = " << na << "L * Np;\n"
      << "  const real *Sj = S + (long)j * ldr + n;\n"
      << "  real *Tj = Tout + (long)j * ldr + n;\n"
      << "  real *Hj = H + (long)j * ldr + n;\n";
    for (int a = 0; a < na; ++a) o << "  const real e" << a << " = Sj[" << a << "L * Np];\n";
    o << "  double d[" << kq << "];\n"
      << "  act_derivs<" << kq << ", " << (bes ? "true" : "false") << ">(ap, (double)e0, d);\n";
    for (int q = 0; q < kq; ++q) o << "  const real d" << q << " = (real)d[" << q << "];\n";
    o << "  Tj[0] = d0;\n  Hj[0] = d1;\n";
    // stored entries are function-scope registers
    for (int e = na; e < s.nentry; ++e) o << "  real e" << e << ";\n";
    for (int a = 1; a < na; ++a) {
        const int g0 = s.slot_group[a], g1 = s.slot_group[a + 1];
        o << "  {\n    real t = d1 * e" << a << ", h = d2 * e" << a << ";\n";
        for (int g = g0; g < g1; ++g) {
            const int q = s.grp_q[g];
            o << "    {\n      real v = (real)0;\n";
            for (int it = s.grp_term[g]; it < s.grp_term[g + 1]; ++it)
                o << "      v += " << lit(s.term_c[it], fp64) << " * e" << s.term_s[it] << " * e" << s.term_b[it]
                  << ";\n";
            if (s.grp_store[g] >= 0) o << "      e" << s.grp_store[g] << " = v;\n";
            o << "      t += d" << q << " * v;\n      h += d" << q + 1 << " * v;\n    }\n";
        }
        o << "    Tj[" << a << "L * Np] = t;\n    Hj[" << a << "L * Np] = h;\n  }\n";
    }
    o << "}\n";

    // ---- adjoint: Sbar from Tbar and h of one lane
    o << "extern \"C\" __global__ __launch_bounds__(256) void dnnf90_bell_bwd" << sfx << "(const real *__restrict__ Tb, "
         "const real *__restrict__ H, real *__restrict__ Sb, int nrow, long Np)\n{\n"
      << "  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;\n"
      << "  const int j = blockIdx.y * (blockDim.x >> 6) + wave;\n"
      << "  if (j >= nrow) return;\n"
      << "  const long n = (long)blockIdx.x * 64 + lane;\n"
      << "  const long ldr = " << na << "L * Np;\n"
      << "  const real *tb = Tb + (long)j * ldr + n;\n"
      << "  const real *hh = H + (long)j * ldr + n;\n"
      << "  real *sb = Sb + (long)j * ldr + n;\n";
    for (int a = 0; a < na; ++a)
        o << "  const real t" << a << " = tb[" << a << "L * Np], h" << a << " = hh[" << a << "L * Np];\n";
    o << "  {\n    real v = t0 * h0;\n";
    for (int a = 1; a < na; ++a) o << "    v += t" << a << " * h" << a << ";\n";
    o << "    sb[0] = v;\n  }\n";
    for (int b = 1; b < na; ++b) {
        o << "  {\n    real v = (real)0;\n";
        for (int ip = t.pair_start[b]; ip < t.pair_start[b] + t.pair_num[b]; ++ip)
            o << "    v += " << lit(t.pair_binom[ip], fp64) << " * t" << t.pair_ia[ip] << " * h" << t.pair_id[ip]
              << ";\n";
        o << "    sb[" << b << "L * Np] = v;\n  }\n";
    }
    o << "}\n";
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
    if (hipModuleLoadData(&g->mod, g->code.data()) != hipSuccess ||
        hipModuleGetFunction(&g->fwd, g->mod, fp64 ? "dnnf90_bell_fwd_f64" : "dnnf90_bell_fwd_f32") != hipSuccess ||
        hipModuleGetFunction(&g->bwd, g->mod, fp64 ? "dnnf90_bell_bwd_f64" : "dnnf90_bell_bwd_f32") != hipSuccess) {
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
    g->ok = true;
    return 0;
}

void GenKernels::release()
{
    if (mod) (void)hipModuleUnload(mod);
    mod = nullptr;
    fwd = bwd = nullptr;
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

}  // namespace dnnf90
