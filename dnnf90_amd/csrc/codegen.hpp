// codegen.hpp -- run-time generated Faa di Bruno kernels (see codegen.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "kernels.hpp"
#include "tables.hpp"

namespace dnnf90 {

struct GenKernels {
    hipModule_t mod = nullptr;
    hipFunction_t fwd = nullptr, bwd = nullptr;
    std::vector<char> code;
    int fwd_regs = 0, bwd_regs = 0, fwd_scratch = 0, bwd_scratch = 0;
    bool ok = false;
    void release();
};

// HIP source of the straight-line forward and adjoint kernels of one family
std::string gen_bell_source(const TableSet &t, const FwdSchedule &s, bool fp64);
// whether the family is small enough to unroll
bool gen_bell_usable(const TableSet &t, const FwdSchedule &s);
// compile with hipRTC and load on the current device; 0 on success
int gen_bell_compile(const TableSet &t, const FwdSchedule &s, bool fp64, GenKernels *g, std::string *log);
// load compiled code (from gen_bell_compile or a cache of it) into g;
// 0, or -3 with g released and the HIP error state cleared
int gen_bell_load(const std::vector<char> &code, bool fp64, GenKernels *g);

hipError_t gen_launch_fwd(const GenKernels &g, const void *S, void *Tout, void *H, int nrow, long Np,
                          const ActParams &ap, hipStream_t s);
hipError_t gen_launch_bwd(const GenKernels &g, const void *Tb, const void *H, void *Sb, int nrow, long Np,
                          hipStream_t s);

}  // namespace dnnf90
