// mfma.hpp -- MFMA operand types and wrappers shared by the GEMM kernels.
//
// Layouts of the 16x16 family on gfx950 (wave of 64): lane l feeds row
// l mod 16 of A and column l mod 16 of B; the k index a lane feeds is l/16
// (x4 forms) or 8*(l/16)+e (bf16 x32 form).  Result: column l mod 16, rows
// 4*(l/16)+r for the FP32-accumulating forms and l/16+4r for
// v_mfma_f64_16x16x4f64 (measured on gfx950).
#ifndef DNNF90_MFMA_HPP
#define DNNF90_MFMA_HPP
#include <hip/hip_runtime.h>

namespace dnnf90 {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef dbl4 V;
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
    static __device__ __forceinline__ V mma(double a, double b, V c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<float> {
    typedef flt4 V;
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
    static __device__ __forceinline__ V mma(float a, float b, V c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// one 16-byte vector of T
template <typename T> struct Vec2x16;
template <> struct Vec2x16<float> { typedef float V __attribute__((ext_vector_type(4))); };
template <> struct Vec2x16<double> { typedef double V __attribute__((ext_vector_type(2))); };

// x = h + m + l (bf16 pieces, each the round-to-nearest bf16 of what the
// previous ones leave; the residuals are exact in FP32), so a product of two
// split operands keeps the six terms of relative size 2^-16 or larger.
// Elements go in pairs: one packed conversion per pair and level, and a
// bf16 pair widens back to FP32 by a shift and a mask.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float flt2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ flt2 widen2(bf16x2 v)
{
    const unsigned b = __builtin_bit_cast(unsigned, v);
    return flt2{__builtin_bit_cast(float, b << 16), __builtin_bit_cast(float, b & 0xffff0000u)};
}

__device__ __forceinline__ void split3(const float (&v)[8], bf16x8 &h, bf16x8 &m, bf16x8 &l)
{
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const flt2 x = {v[2 * p], v[2 * p + 1]};
        const bf16x2 a = __builtin_convertvector(x, bf16x2);
        const flt2 r = x - widen2(a);
        const bf16x2 b = __builtin_convertvector(r, bf16x2);
        const bf16x2 c = __builtin_convertvector(r - widen2(b), bf16x2);
        h[2 * p] = a[0];
        h[2 * p + 1] = a[1];
        m[2 * p] = b[0];
        m[2 * p + 1] = b[1];
        l[2 * p] = c[0];
        l[2 * p + 1] = c[1];
    }
}

// c += a * b over the six significant terms of the split product, small
// terms first
__device__ __forceinline__ flt4 mma_x6(const bf16x8 &ah, const bf16x8 &am, const bf16x8 &al, const bf16x8 &bh,
                                       const bf16x8 &bm, const bf16x8 &bl, flt4 c)
{
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
    return c;
}

}  // namespace dnnf90
#endif
