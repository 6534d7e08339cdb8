// gemm_x6.hip -- K1, K6 and K5 on BF16 three-way split products.
//
// An FP32 operand x is carried as three bf16 pieces x = h + m + l (mfma.hpp
// split3), and a product a*b as the six terms whose size is 2^-16 |a b| or
// more; every term is one v_mfma_f32_16x16x32_bf16, exact products summed in
// FP32.  Six bf16 MFMAs cover 32 k at 96 cycles where FP32 MFMA needs eight
// v_mfma_f32_16x16x4f32 at 256 cycles, with product error near FP32's.  The
// forward map K1 (whose accumulation error feeds every carried derivative)
// moves each 32-term MFMA sum into an FP64 accumulator, which keeps the
// forward error of an FP32-storage pass below 1e-6; K6 and K5 accumulate in
// FP32 (K5's chunk partials are then summed in FP64 like every K5).
//
//   K1  S   = W  [1; T]   (lib/train_module.f90:262-278)
//   K6  Tb  = W' Sbar     (lib/train_module.f90:430-444)
//   K5  dW += Sbar [1; T]^T (lib/train_module.f90:408-422)
#include <hip/hip_bf16.h>

#include <algorithm>

#include "kernels.hpp"
#include "mfma.hpp"

namespace dnnf90 {

// register prefetch depth (chunks of 32 k) of the split linear maps: the
// forward map K1 and the transposed map K6
#ifndef LIN_X6_KP
#define LIN_X6_KP 3
#endif
#ifndef LIN_X6_KP_T
#define LIN_X6_KP_T 3
#endif
// largest row tile (in 16-row blocks) of K1 and of K6.  K1 at 64 rows: two
// workgroups share a column range (its rows come from L2 the second time)
// and two fit a CU, against one at 128 rows (11 % faster on a 128-wide
// layer); K6 at 128 rows with four column blocks a step
#ifndef LIN_X6_RBT_MAX
#define LIN_X6_RBT_MAX 4
#endif
#ifndef LIN_X6_RBT_MAX_T
#define LIN_X6_RBT_MAX_T 8
#endif
// 16-column blocks a wave streams per step.  Where the split weights' LDS
// leaves room for one workgroup per CU, occupancy is one wave per SIMD
// whatever the registers, and the _DEEP settings (more column blocks, a
// longer prefetch ring) keep more loads in flight; elsewhere those cost
// occupancy.  Measured: four blocks make K6 14 % faster on a 128-wide layer
// (128-row tile) and 20 % on a 256-wide one (64-row tile), K1 17 % faster on
// a 256-wide layer; a ring of 4 or 6 chunks changes < 2 % (6 spills K6 at
// 128 rows).
#ifndef LIN_X6_NCB
#define LIN_X6_NCB 2
#endif
#ifndef LIN_X6_NCB_DEEP
#define LIN_X6_NCB_DEEP 4
#endif
#ifndef LIN_X6_NCB_DEEP_T
#define LIN_X6_NCB_DEEP_T 4
#endif
#ifndef LIN_X6_KP_DEEP
#define LIN_X6_KP_DEEP 3
#endif
// column order of a wave step: 0 -- block j is columns 16 j .. 16 j + 15
// (dword loads and stores, sixteen lanes cover 64 bytes of a row); 1 --
// lane n owns the NCB consecutive columns NCB n .. NCB n + NCB - 1, one per
// block (NCB-wide vector loads and stores, sixteen lanes cover 64 NCB bytes)
#ifndef LIN_X6_VEC
#define LIN_X6_VEC 1
#endif

// ---------------------------------------------------------------------------
// K1 / K6.  The workgroup's 16*RBT rows of A (K1: the weight columns after
// the bias; K6: the transposed weights) are split once into LDS in operand
// order, Ap[rb][kc][piece][lane] = pieces of A[16 rb + lane%16][32 kc +
// 8 (lane/16) .. +7], zero past Kd.  A wave streams units of 16 NCB
// columns of B as NCB MFMA column blocks: lane (n = lane%16, q = lane/16)
// reads rows 32 kc + 8 q .. +7 of its NCB columns (LIN_X6_VEC: the
// consecutive columns NCB n .., one NCB-wide vector load a row, so sixteen
// lanes read 64 NCB contiguous bytes; block j takes column NCB n + j), splits
// them and multiplies by every row block; the next chunk's rows are loaded
// before the current chunk's MFMAs.  The results go out the same way, one
// NCB-wide vector store per row.  K1's bias is added in the epilogue to
// the first Np columns (slot 0; NCB*16 columns never straddle Np).
template <bool kTrans, bool kAcc64, int RBT, int NCB, int KP>
__global__ __launch_bounds__(256) void k_lin_x6(const float *__restrict__ W, int ldw, const float *__restrict__ B,
                                                float *__restrict__ C, int M, int Kd, long ncol, long Np, int nrt,
                                                int ncw, long cupw)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char x6_lds[];
    bf16x8 *Ap = reinterpret_cast<bf16x8 *>(x6_lds);
    const int G = gridDim.x, bid = blockIdx.x;
    const int w = (bid & 7) * (G >> 3) + (bid >> 3);
    const int rt = w % nrt, cw = w / nrt;
    if (cw >= ncw) return;
    const long ncu = ncol / (16 * NCB);
    const long cu0 = (long)cw * cupw, cu1 = min(ncu, cu0 + cupw);
    if (cu0 >= cu1) return;
    const int nkc = (Kd + 31) >> 5;
    const int tid = threadIdx.x, lane = tid & 63, q = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int m0 = rt * 16 * RBT;
    for (int u = tid; u < RBT * nkc * 64; u += 256) {
        const int rb = u / (nkc * 64), rem = u - rb * nkc * 64, kc = rem >> 6, ln = rem & 63;
        const int m = m0 + 16 * rb + (ln & 15), k = 32 * kc + 8 * (ln >> 4);
        float v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int kk = k + e;
            v[e] = (m < M && kk < Kd) ? (kTrans ? W[(long)kk * ldw + 1 + m] : W[(long)m * ldw + 1 + kk]) : 0.f;
        }
        bf16x8 h, mm, l;
        split3(v, h, mm, l);
        bf16x8 *dst = Ap + (long)(rb * nkc + kc) * 3 * 64 + ln;
        dst[0] = h;
        dst[64] = mm;
        dst[128] = l;
    }
    float *Bias = reinterpret_cast<float *>(Ap + (long)RBT * nkc * 3 * 64);
    if (!kTrans && tid < 16 * RBT) Bias[tid] = m0 + tid < M ? W[(long)(m0 + tid) * ldw] : 0.f;
    __syncthreads();
    if (cu0 + wave >= cu1) return;
    // the wave's iterations: it -> (unit cu0 + wave + 4 (it / nkc), chunk
    // it % nkc); the rows of iteration it + KP - 1 are loaded while
    // iteration it is multiplied (a ring of KP register buffers, indexed
    // statically by unrolling the loop KP times)
    const long nu = (cu1 - cu0 - wave + 3) / 4;
    const long nit = nu * nkc;
    // column of block j held by this lane, relative to the unit's first
    const auto col_of = [&](int j) { return LIN_X6_VEC ? NCB * (lane & 15) + j : 16 * j + (lane & 15); };
    const float *bl = B + col_of(0);
    float rb_[KP][NCB][8];
    auto load = [&](float (&dst)[NCB][8], long it) {
        const long cul = cu0 + wave + 4 * (it / nkc);
        const int kc = (int)(it % nkc);
        if constexpr (LIN_X6_VEC && NCB > 1) {
            typedef float fv __attribute__((ext_vector_type(NCB)));
            const float *src = bl + cul * 16 * NCB;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const fv v = *reinterpret_cast<const fv *>(src + (long)min(32 * kc + 8 * q + e, Kd - 1) * ncol);
#pragma unroll
                for (int j = 0; j < NCB; ++j) dst[j][e] = v[j];
            }
        } else {
#pragma unroll
            for (int j = 0; j < NCB; ++j) {
                const float *src = bl + cul * 16 * NCB + 16 * j;
#pragma unroll
                for (int e = 0; e < 8; ++e) dst[j][e] = src[(long)min(32 * kc + 8 * q + e, Kd - 1) * ncol];
            }
        }
    };
    flt4 cacc[RBT][NCB];
    double acc64[kAcc64 ? RBT : 1][kAcc64 ? NCB : 1][4];
    auto step = [&](const float (&rc)[NCB][8], long it) {
        const long cu = cu0 + wave + 4 * (it / nkc);
        const int kc = (int)(it % nkc);
        if (kc == 0 || kAcc64) {
#pragma unroll
            for (int rb = 0; rb < RBT; ++rb)
#pragma unroll
                for (int j = 0; j < NCB; ++j) cacc[rb][j] = flt4{0, 0, 0, 0};
        }
        if (kAcc64 && kc == 0) {
#pragma unroll
            for (int rb = 0; rb < (kAcc64 ? RBT : 1); ++rb)
#pragma unroll
                for (int j = 0; j < (kAcc64 ? NCB : 1); ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc64[rb][j][r] = 0.0;
        }
        bf16x8 bh[NCB], bm[NCB], blo[NCB];
#pragma unroll
        for (int j = 0; j < NCB; ++j) split3(rc[j], bh[j], bm[j], blo[j]);
        const bf16x8 *Ak = Ap + (long)kc * 3 * 64 + lane;
#pragma unroll
        for (int rb = 0; rb < RBT; ++rb) {
            const bf16x8 *a = Ak + (long)rb * nkc * 3 * 64;
            const bf16x8 ah = a[0], am = a[64], al = a[128];
#pragma unroll
            for (int j = 0; j < NCB; ++j) cacc[rb][j] = mma_x6(ah, am, al, bh[j], bm[j], blo[j], cacc[rb][j]);
        }
        if (kAcc64) {
#pragma unroll
            for (int rb = 0; rb < (kAcc64 ? RBT : 1); ++rb)
#pragma unroll
                for (int j = 0; j < (kAcc64 ? NCB : 1); ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc64[rb][j][r] += (double)cacc[rb][j][r];
        }
        if (kc == nkc - 1) {
            const bool bias = !kTrans && cu * 16 * NCB < Np;
#pragma unroll
            for (int rb = 0; rb < RBT; ++rb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int mr = 16 * rb + 4 * q + r, m = m0 + mr;
                    const float add = bias ? Bias[mr] : 0.f;
                    float val[NCB];
#pragma unroll
                    for (int j = 0; j < NCB; ++j) {
                        if (kAcc64)
                            val[j] = (float)(acc64[kAcc64 ? rb : 0][kAcc64 ? j : 0][r] + (double)add);
                        else
                            val[j] = cacc[rb][j][r] + add;
                    }
                    if (m >= M) continue;
                    float *dst = C + (long)m * ncol + cu * 16 * NCB;
                    if constexpr (LIN_X6_VEC && NCB > 1) {
                        typedef float fv __attribute__((ext_vector_type(NCB)));
                        fv v;
#pragma unroll
                        for (int j = 0; j < NCB; ++j) v[j] = val[j];
                        *reinterpret_cast<fv *>(dst + col_of(0)) = v;
                    } else {
#pragma unroll
                        for (int j = 0; j < NCB; ++j) dst[col_of(j)] = val[j];
                    }
                }
        }
    };
#pragma unroll
    for (int j = 0; j < KP - 1; ++j)
        if (j < nit) load(rb_[j], j);
    for (long it = 0; it < nit; it += KP) {
#pragma unroll
        for (int j = 0; j < KP; ++j) {
            if (it + j + KP - 1 < nit) load(rb_[(j + KP - 1) % KP], it + j + KP - 1);
            if (it + j < nit) step(rb_[j], it + j);
        }
    }
}

// K1 / K6 where the reduction is whole 32-row chunks (Kd = 32 NKC, NKC
// fixed at compile time).  Same operand layout, products, summation order
// and results (bit for bit) as k_lin_x6, with less VALU work a step: the
// chunk loop is unrolled, a load or store address is a wave-uniform row base
// plus one 32-bit lane offset (no per-lane 64-bit row arithmetic), and the
// NKC chunks of the wave's next unit are loaded into the registers of the
// chunk just split, so a whole unit of loads is in flight during the
// products (PD chunks ahead: the chunk g of the wave's sequence is held in
// register buffer g mod PD, refilled with chunk g + PD right after it is
// split; PD divides NKC).  The lane offsets need 24 ncol + 16 NCB < 2^30
// (host-checked).  WPE: the waves per SIMD the registers are held to.
template <bool kTrans, bool kAcc64, int RBT, int NCB, int NKC, int PD, int WPE>
__global__ __launch_bounds__(256, WPE) void k_lin_x6c(const float *__restrict__ W, int ldw, const float *__restrict__ B,
                                                 float *__restrict__ C, int M, long ncol, long Np, int nrt, int ncw,
                                                 long cupw)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char x6_lds[];
    bf16x8 *Ap = reinterpret_cast<bf16x8 *>(x6_lds);
    const int G = gridDim.x, bid = blockIdx.x;
    const int w = (bid & 7) * (G >> 3) + (bid >> 3);
    const int rt = w % nrt, cw = w / nrt;
    if (cw >= ncw) return;
    const long ncu = ncol / (16 * NCB);
    const long cu0 = (long)cw * cupw, cu1 = min(ncu, cu0 + cupw);
    if (cu0 >= cu1) return;
    const int tid = threadIdx.x, lane = tid & 63, q = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int m0 = rt * 16 * RBT;
    for (int u = tid; u < RBT * NKC * 64; u += 256) {
        const int rb = u / (NKC * 64), rem = u - rb * NKC * 64, kc = rem >> 6, ln = rem & 63;
        const int m = m0 + 16 * rb + (ln & 15), k = 32 * kc + 8 * (ln >> 4);
        float v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int kk = k + e;
            v[e] = m < M ? (kTrans ? W[(long)kk * ldw + 1 + m] : W[(long)m * ldw + 1 + kk]) : 0.f;
        }
        bf16x8 h, mm, l;
        split3(v, h, mm, l);
        bf16x8 *dst = Ap + (long)(rb * NKC + kc) * 3 * 64 + ln;
        dst[0] = h;
        dst[64] = mm;
        dst[128] = l;
    }
    float *Bias = reinterpret_cast<float *>(Ap + (long)RBT * NKC * 3 * 64);
    if (!kTrans && tid < 16 * RBT) Bias[tid] = m0 + tid < M ? W[(long)(m0 + tid) * ldw] : 0.f;
    __syncthreads();
    if (cu0 + wave >= cu1) return;
    const int nu = (int)((cu1 - cu0 - wave + 3) / 4);
    const int col0 = NCB * (lane & 15);   // lane n: columns NCB n .. NCB n + NCB - 1
    // byte offsets of this lane: loads, rows 8 q + e of a chunk; stores,
    // rows 4 q + r of a row block.  Loads and stores are raw buffer accesses:
    // a scalar resource at the chunk's (row block's) first row and unit, the
    // row within it as a scalar offset, the lane's offset in one VGPR
    const unsigned ld_off = ((unsigned)(8 * q) * (unsigned)ncol + (unsigned)col0) * 4u;
    const unsigned st_off = ((unsigned)(4 * q) * (unsigned)ncol + (unsigned)col0) * 4u;
    const long rowb = ncol * 4;   // bytes per row
    const int rowb32 = (int)rowb; // e rowb for e < 8 stays below 2^31 (host-checked)
    typedef float fv __attribute__((ext_vector_type(NCB)));
    const auto rsrc = [](const void *p) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, -1, 0x00020000);
    };
    const auto ldv = [](__amdgpu_buffer_rsrc_t r, unsigned vo, int so) -> fv {
        if constexpr (NCB == 4) return __builtin_bit_cast(fv, __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, 0));
        else if constexpr (NCB == 2) return __builtin_bit_cast(fv, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
        else return __builtin_bit_cast(fv, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
    };
    const auto stv = [](fv v, __amdgpu_buffer_rsrc_t r, unsigned vo, int so) {
        if constexpr (NCB == 4) {
            typedef unsigned u4 __attribute__((ext_vector_type(4)));
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, vo, so, 0);
        } else if constexpr (NCB == 2) {
            typedef unsigned u2 __attribute__((ext_vector_type(2)));
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, vo, so, 0);
        } else {
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vo, so, 0);
        }
    };
    static_assert(NKC % PD == 0, "PD divides NKC");
    float buf[PD][NCB][8];
    auto load = [&](float (&dst)[NCB][8], int kc, long cu) {
        const __amdgpu_buffer_rsrc_t r =
            rsrc(reinterpret_cast<const char *>(B) + (long)(32 * kc) * rowb + cu * (16 * NCB * 4));
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const fv v = ldv(r, ld_off, e * rowb32);
#pragma unroll
            for (int j = 0; j < NCB; ++j) dst[j][e] = v[j];
        }
    };
#pragma unroll
    for (int g = 0; g < PD; ++g)
        if (g / NKC < nu) load(buf[g], g % NKC, cu0 + wave + 4 * (g / NKC));
    for (int i = 0; i < nu; ++i) {
        const long cu = cu0 + wave + 4L * i;
        // the weight pieces are read from LDS again for every unit: hidden
        // from the compiler, which would otherwise hold all of them in
        // registers across the loop (RBT NKC 12 VGPRs)
        int la = lane;
        asm volatile("" : "+v"(la));
        flt4 cacc[RBT][NCB];
        double acc64[kAcc64 ? RBT : 1][kAcc64 ? NCB : 1][4];
#pragma unroll
        for (int kc = 0; kc < NKC; ++kc) {
            bf16x8 bh[NCB], bm[NCB], blo[NCB];
#pragma unroll
            for (int j = 0; j < NCB; ++j) split3(buf[kc % PD][j], bh[j], bm[j], blo[j]);
            const int di = (kc + PD) / NKC;
            if (i + di < nu) load(buf[kc % PD], (kc + PD) % NKC, cu + 4 * di);
            const bf16x8 *Ak = Ap + kc * 3 * 64 + la;
#pragma unroll
            for (int rb = 0; rb < RBT; ++rb) {
                const bf16x8 *a = Ak + rb * NKC * 3 * 64;
                const bf16x8 ah = a[0], am = a[64], al = a[128];
#pragma unroll
                for (int j = 0; j < NCB; ++j)
                    cacc[rb][j] = mma_x6(ah, am, al, bh[j], bm[j], blo[j],
                                         (kc == 0 || kAcc64) ? flt4{0, 0, 0, 0} : cacc[rb][j]);
            }
            if constexpr (kAcc64) {
#pragma unroll
                for (int rb = 0; rb < RBT; ++rb)
#pragma unroll
                    for (int j = 0; j < NCB; ++j)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            acc64[rb][j][r] = kc == 0 ? 0.0 + (double)cacc[rb][j][r]
                                                      : acc64[rb][j][r] + (double)cacc[rb][j][r];
            }
        }
        const bool bias = !kTrans && cu * 16 * NCB < Np;
        char *cb = reinterpret_cast<char *>(C) + cu * (16 * NCB * 4);
#pragma unroll
        for (int rb = 0; rb < RBT; ++rb) {
            const __amdgpu_buffer_rsrc_t r = rsrc(cb + (long)(m0 + 16 * rb) * rowb);
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                const int mr = 16 * rb + 4 * q + r4;
                const float add = bias ? Bias[mr] : 0.f;
                fv v;
#pragma unroll
                for (int j = 0; j < NCB; ++j) {
                    if constexpr (kAcc64)
                        v[j] = (float)(acc64[rb][j][r4] + (double)add);
                    else
                        v[j] = cacc[rb][j][r4] + add;
                }
                if (m0 + mr < M) stv(v, r, st_off, r4 * rowb32);
            }
        }
    }
}

// prefetch distance (chunks) of the fixed-chunk maps where a CU holds
// several workgroups, and where it holds one
#ifndef LIN_X6C_PD
#define LIN_X6C_PD 2
#endif
// 1: the general kernel's grid is one round at the resident workgroups per
// CU that the runtime reports (registers and LDS), not at the LDS bound
// alone -- the 64-row K1 of a 128-wide layer holds 240 registers, so two
// workgroups a CU, not three (its grid was 1.5 rounds; one round measured
// 5 % faster)
// 1: the 128 x 128 forward map K1 takes k_lin_x6b (one workgroup a CU, B
// split once; 6 % faster than k_lin_x6 on dense4_k3)
#ifndef LIN_X6B
#define LIN_X6B 1
#endif
#ifndef LIN_X6_OCC
#define LIN_X6_OCC 1
#endif
#ifndef LIN_X6C_PD_DEEP
#define LIN_X6C_PD_DEEP 4
#endif

// 0: every split linear map takes k_lin_x6 (the fixed-chunk kernels off)
static int g_lin_x6c = 1;
int set_lin_x6c(int on)
{
    const int prev = g_lin_x6c;
    if (on >= 0) g_lin_x6c = on ? 1 : 0;
    return prev;
}

// K1 (and, by the same code, K6) of a 128 x 128 map (M = 128 rows, Kd = 128
// = four 32-row chunks) in one workgroup per CU: the split weights (96 KB) and one 64-column
// step of split B (48 KB) in LDS.  A step: wave w loads chunk w of B --
// lane (n, q) the rows 32 w + 8 q .. +7 of the four consecutive columns
// 4 n .. 4 n + 3, one 16-byte load a row -- splits it and writes the pieces
// to LDS where that same lane's operand sits (block j = column 4 n + j), so
// B is split once (k_lin_x6 at 64-row tiles splits it twice); after a
// barrier wave w multiplies row blocks 2 w, 2 w + 1 by all four column
// blocks over the four chunks, flushes each chunk's sum to FP64 (K1) as
// k_lin_x6 does, and stores its rows as 16-byte vectors; the next step's
// loads are in flight meanwhile.  Same products in the same order as
// k_lin_x6: the same bits.
template <bool kTrans, bool kAcc64>
__global__ __launch_bounds__(256, 1) void k_lin_x6b(const float *__restrict__ W, int ldw, const float *__restrict__ B,
                                                    float *__restrict__ C, long ncol, long Np, long cupw)
{
    constexpr int RBT = 8, NKC = 4, NCB = 4;
    extern __shared__ __attribute__((aligned(16))) unsigned char x6_lds[];
    bf16x8 *Ap = reinterpret_cast<bf16x8 *>(x6_lds);
    bf16x8 *Bq = Ap + RBT * NKC * 3 * 64;                                   // [kc][j][piece][lane]
    float *Bias = reinterpret_cast<float *>(Bq + NKC * NCB * 3 * 64);
    const long ncu = ncol / 64;
    const long cu0 = (long)blockIdx.x * cupw, cu1 = min(ncu, cu0 + cupw);
    if (cu0 >= cu1) return;
    const int tid = threadIdx.x, lane = tid & 63, q = lane >> 4, n = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int u = tid; u < RBT * NKC * 64; u += 256) {
        const int rb = u / (NKC * 64), rem = u - rb * NKC * 64, kc = rem >> 6, ln = rem & 63;
        const int m = 16 * rb + (ln & 15), k = 32 * kc + 8 * (ln >> 4);
        float v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e)
            v[e] = kTrans ? W[(long)(k + e) * ldw + 1 + m] : W[(long)m * ldw + 1 + k + e];
        bf16x8 h, mm, l;
        split3(v, h, mm, l);
        bf16x8 *dst = Ap + (long)(rb * NKC + kc) * 3 * 64 + ln;
        dst[0] = h;
        dst[64] = mm;
        dst[128] = l;
    }
    if (!kTrans && tid < 128) Bias[tid] = W[(long)tid * ldw];
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    const long rowb = ncol * 4;
    const int rowb32 = (int)rowb;   // 7 rowb < 2^31 (host-checked)
    const auto rsrc = [](const void *p) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, -1, 0x00020000);
    };
    const unsigned ld_off = ((unsigned)(8 * q) * (unsigned)ncol + 4u * n) * 4u;
    const unsigned st_off = ((unsigned)(4 * q) * (unsigned)ncol + 4u * n) * 4u;
    float raw[8][4];   // rows 32 wave + 8 q + e, columns 4 n .. 4 n + 3
    auto load = [&](long cu) {
        const __amdgpu_buffer_rsrc_t r =
            rsrc(reinterpret_cast<const char *>(B) + (long)(32 * wave) * rowb + cu * 256);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const f4 v = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, ld_off, e * rowb32, 0));
#pragma unroll
            for (int j = 0; j < 4; ++j) raw[e][j] = v[j];
        }
    };
    load(cu0);
    for (long cu = cu0; cu < cu1; ++cu) {
        // this step's chunk: split into LDS (the previous step's reads ended
        // at its closing barrier)
#pragma unroll
        for (int j = 0; j < NCB; ++j) {
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = raw[e][j];
            bf16x8 h, mm, l;
            split3(v, h, mm, l);
            bf16x8 *dst = Bq + (wave * NCB + j) * 3 * 64 + lane;
            dst[0] = h;
            dst[64] = mm;
            dst[128] = l;
        }
        if (cu + 1 < cu1) load(cu + 1);
        __syncthreads();
        flt4 cacc[kAcc64 ? 1 : 2][NCB];
        double acc64[kAcc64 ? 2 : 1][kAcc64 ? NCB : 1][4];
#pragma unroll
        for (int kc = 0; kc < NKC; ++kc) {
            bf16x8 ah[2], am[2], al[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const bf16x8 *a = Ap + ((2 * wave + i) * NKC + kc) * 3 * 64 + lane;
                ah[i] = a[0];
                am[i] = a[64];
                al[i] = a[128];
            }
            // one column block's pieces live at a time (each accumulator's
            // own sequence over kc is unchanged)
#pragma unroll
            for (int j = 0; j < NCB; ++j) {
                const bf16x8 *sb = Bq + (kc * NCB + j) * 3 * 64 + lane;
                const bf16x8 bh = sb[0], bm = sb[64], bl = sb[128];
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    if constexpr (kAcc64) {
                        // the chunk's sum goes to FP64 at once (no FP32
                        // accumulator held across column blocks)
                        const flt4 c = mma_x6(ah[i], am[i], al[i], bh, bm, bl, flt4{0, 0, 0, 0});
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            acc64[i][j][r] = kc == 0 ? 0.0 + (double)c[r] : acc64[i][j][r] + (double)c[r];
                    } else {
                        cacc[i][j] = mma_x6(ah[i], am[i], al[i], bh, bm, bl, kc == 0 ? flt4{0, 0, 0, 0} : cacc[i][j]);
                    }
                }
            }
        }
        const bool bias = !kTrans && cu * 64 < Np;
        char *cb = reinterpret_cast<char *>(C) + cu * 256;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int rb = 2 * wave + i;
            const __amdgpu_buffer_rsrc_t r = rsrc(cb + (long)(16 * rb) * rowb);
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                const float add = bias ? Bias[16 * rb + 4 * q + r4] : 0.f;
                f4 v;
#pragma unroll
                for (int j = 0; j < NCB; ++j) {
                    if constexpr (kAcc64)
                        v[j] = (float)(acc64[i][j][r4] + (double)add);
                    else
                        v[j] = cacc[i][j][r4] + add;
                }
                __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, st_off, r4 * rowb32, 0);
            }
        }
        __syncthreads();
    }
}

template <bool kTrans, bool kAcc64>
static bool lin_x6b_try(const float *W, int ldw, const float *B, float *C, int M, int Kd, long ncol, long Np,
                        hipStream_t s)
{
    // K1 only: K6 at its 128-row tile already splits B once (measured equal)
    if (kTrans || !LIN_X6B || !g_lin_x6c || M != 128 || Kd != 128 || ncol % 64 != 0 || 24 * ncol + 64 >= (1L << 28))
        return false;
    constexpr size_t lds = (size_t)(8 * 4 + 4 * 4) * 3 * 64 * 16 + 128 * sizeof(float);
    lds_allow(reinterpret_cast<const void *>(&k_lin_x6b<kTrans, kAcc64>), lds);
    const long ncu = ncol / 64;
    const long cupw = (ncu + 255) / 256;   // one workgroup a CU (its LDS)
    const long g = (ncu + cupw - 1) / cupw;
    hipLaunchKernelGGL((k_lin_x6b<kTrans, kAcc64>), dim3((unsigned)g), dim3(256), lds, s, W, ldw, B, C, ncol, Np,
                       cupw);
    return true;
}

template <bool kTrans, bool kAcc64, int RBT, int NCB, int NKC, int PD, int WPE>
static void lin_x6c_launch(const float *W, int ldw, const float *B, float *C, int M, long ncol, long Np, size_t lds,
                           hipStream_t s)
{
    const void *fn = reinterpret_cast<const void *>(&k_lin_x6c<kTrans, kAcc64, RBT, NCB, NKC, PD, WPE>);
    // resident workgroups per CU, from the registers and the LDS (queried
    // once per instantiation and LDS size), so the grid is one full round
    const int per_cu = std::min(occupancy_cached(fn, lds), 4);
    const int nrt = (M + 16 * RBT - 1) / (16 * RBT);
    const long ncu = ncol / (16 * NCB);
    const long want = std::max(1L, (long)256 * per_cu / nrt);
    long cupw = (ncu + want - 1) / want;
    cupw = (cupw + 3) / 4 * 4;
    const long ncw = (ncu + cupw - 1) / cupw;
    long g = (long)nrt * ncw;
    g = (g + 7) / 8 * 8;
    hipLaunchKernelGGL((k_lin_x6c<kTrans, kAcc64, RBT, NCB, NKC, PD, WPE>), dim3((unsigned)g), dim3(256), lds, s, W, ldw, B, C,
                       M, ncol, Np, nrt, (int)ncw, cupw);
}

// PDMAX: the prefetch distance in chunks, at most (a whole unit where one
// workgroup fills a CU anyway; two chunks elsewhere, which keeps two waves a
// SIMD, held there by the register bound WPE)
template <bool kTrans, bool kAcc64, int RBT, int NCB, int PDMAX, bool DEEP>
static bool lin_x6c_try(const float *W, int ldw, const float *B, float *C, int M, int Kd, long ncol, long Np,
                        size_t lds, hipStream_t s)
{
    if (!g_lin_x6c || !LIN_X6_VEC || Kd % 32 != 0 || ncol % (16 * NCB) != 0 || 24 * ncol + 16 * NCB >= (1L << 28))
        return false;
    constexpr int WPE = DEEP ? 1 : 2;
    switch (Kd / 32) {
    case 1: lin_x6c_launch<kTrans, kAcc64, RBT, NCB, 1, 1, WPE>(W, ldw, B, C, M, ncol, Np, lds, s); return true;
    case 2:
        lin_x6c_launch<kTrans, kAcc64, RBT, NCB, 2, std::min(2, PDMAX), WPE>(W, ldw, B, C, M, ncol, Np, lds, s);
        return true;
    default: return false;   // four and eight chunks: k_lin_x6 measured faster (128- and 256-wide layers)
    }
}

template <bool kTrans, bool kAcc64, int RBT, int NCB, int KP>
static void lin_x6_launch(const float *W, int ldw, const float *B, float *C, int M, int Kd, long ncol, long Np,
                          size_t lds, int per_cu, hipStream_t s)
{
    const void *fn = reinterpret_cast<const void *>(&k_lin_x6<kTrans, kAcc64, RBT, NCB, KP>);
#if LIN_X6_OCC
    per_cu = std::min(per_cu, occupancy_cached(fn, lds));
#else
    lds_allow(fn, lds);
#endif
    const int nrt = (M + 16 * RBT - 1) / (16 * RBT);
    const long ncu = ncol / (16 * NCB);
    const long want = std::max(1L, (long)256 * per_cu / nrt);
    long cupw = (ncu + want - 1) / want;
    cupw = (cupw + 3) / 4 * 4;
    const long ncw = (ncu + cupw - 1) / cupw;
    long g = (long)nrt * ncw;
    g = (g + 7) / 8 * 8;
    hipLaunchKernelGGL((k_lin_x6<kTrans, kAcc64, RBT, NCB, KP>), dim3((unsigned)g), dim3(256), lds, s, W, ldw, B, C, M,
                       Kd, ncol, Np, nrt, (int)ncw, cupw);
}

template <bool kTrans, bool kAcc64, int RBT>
static void lin_x6_go(const float *W, int ldw, const float *B, float *C, int M, int Kd, long ncol, long Np,
                      hipStream_t s)
{
    const int nkc = (Kd + 31) / 32;
    const size_t lds = (size_t)RBT * nkc * 3 * 64 * 16 + 16 * RBT * sizeof(float);
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, (160 * 1024) / lds));
    if (per_cu == 1) {
        if (lin_x6c_try<kTrans, kAcc64, RBT, kTrans ? LIN_X6_NCB_DEEP_T : LIN_X6_NCB_DEEP, LIN_X6C_PD_DEEP, true>(
                W, ldw, B, C, M, Kd, ncol, Np, lds, s))
            return;
    } else if (lin_x6c_try<kTrans, kAcc64, RBT, LIN_X6_NCB, LIN_X6C_PD, false>(W, ldw, B, C, M, Kd, ncol, Np, lds,
                                                                                s)) {
        return;
    }
    if (per_cu == 1)
        lin_x6_launch<kTrans, kAcc64, RBT, kTrans ? LIN_X6_NCB_DEEP_T : LIN_X6_NCB_DEEP, LIN_X6_KP_DEEP>(
            W, ldw, B, C, M, Kd, ncol, Np, lds, per_cu, s);
    else
        lin_x6_launch<kTrans, kAcc64, RBT, LIN_X6_NCB, kTrans ? LIN_X6_KP_T : LIN_X6_KP>(W, ldw, B, C, M, Kd, ncol,
                                                                                          Np, lds, per_cu, s);
}

template <bool kTrans, bool kAcc64>
static void lin_x6_pick(const float *W, int ldw, const float *B, float *C, int M, int Kd, long ncol, long Np,
                        hipStream_t s)
{
    if (lin_x6b_try<kTrans, kAcc64>(W, ldw, B, C, M, Kd, ncol, Np, s)) return;
    const size_t per_rb = (size_t)((Kd + 31) / 32) * 3 * 64 * 16;
    int rbt = kTrans ? LIN_X6_RBT_MAX_T : LIN_X6_RBT_MAX;
    while (rbt > 1 && 16 * (rbt / 2) >= M) rbt /= 2;
    while (rbt > 1 && rbt * per_rb > (size_t)144 * 1024) rbt /= 2;
    switch (rbt) {
    case 8:
        // only K6 reaches 128-row tiles (K1's cap is LIN_X6_RBT_MAX); K1 at
        // 128 rows would spill its FP64 accumulators, so it is not built
        if constexpr (kTrans || LIN_X6_RBT_MAX >= 8) lin_x6_go<kTrans, kAcc64, 8>(W, ldw, B, C, M, Kd, ncol, Np, s);
        break;
    case 4: lin_x6_go<kTrans, kAcc64, 4>(W, ldw, B, C, M, Kd, ncol, Np, s); break;
    case 2: lin_x6_go<kTrans, kAcc64, 2>(W, ldw, B, C, M, Kd, ncol, Np, s); break;
    default: lin_x6_go<kTrans, kAcc64, 1>(W, ldw, B, C, M, Kd, ncol, Np, s); break;
    }
}

void launch_lin_x6(bool trans_a, const float *W, int ldw, const float *B, float *C, int M, int Kd, long ncol,
                   long Np, hipStream_t s)
{
    if (trans_a) lin_x6_pick<true, false>(W, ldw, B, C, M, Kd, ncol, Np, s);
    else lin_x6_pick<false, true>(W, ldw, B, C, M, Kd - 1, ncol, Np, s);
}

// ---------------------------------------------------------------------------
// K5.  As k_grad_nt (gemm.hip): lane (p = lane%16, q = lane/16) reads the 8
// consecutive columns c0 + 8 q .. +7 of row p of each A and B block with two
// 16-byte loads -- exactly the 32-k slice of one bf16 MFMA -- splits them and
// forms the six-term products; a wave owns the (16 RB) x (16 CB) tile over
// every fourth 32-column step of the chunk, the next step's rows loaded
// before the current step's MFMAs.  A rows are split once per step, B rows
// one block at a time.  The bias column (sum of A over the first Np columns)
// is added up from the loaded A rows by the first k tile.
template <int RB, int CB, int ST>
__global__ __launch_bounds__(256) void k_grad_x6(const float *__restrict__ A, const float *__restrict__ B, int M,
                                                 int Kd, long ncol, long Np, long chunk, int ntm, int ntk,
                                                 int nchunk, double *__restrict__ partial)
{
    constexpr int SC = 32;              // columns per wave step
    constexpr int LD = 16 * CB + 1;     // reduction row: the tile, then the bias column
    typedef float f4 __attribute__((ext_vector_type(4)));
    __shared__ double red[16 * RB * LD];
    const int G = gridDim.x, bid = blockIdx.x;
    const int w = (bid & 7) * (G >> 3) + (bid >> 3);
    const int ntile = ntm * ntk;
    const int tile = w % ntile, chk = w / ntile;
    if (chk >= nchunk) return;
    const int Kg = Kd - 1;   // GEMM column k <-> weight column k + 1
    const int m0 = (tile % ntm) * 16 * RB, k0 = (tile / ntm) * 16 * CB;
    const bool do_bias = k0 == 0;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int p = lane & 15, q = lane >> 4;
    const long cb = (long)chk * chunk, ce = min(ncol, cb + chunk);

    const float *pa[RB];
    const float *pb[CB];
#pragma unroll
    for (int i = 0; i < RB; ++i) pa[i] = A + (long)min(m0 + 16 * i + p, M - 1) * ncol + 8 * q;
#pragma unroll
    for (int j = 0; j < CB; ++j) pb[j] = B + (long)min(k0 + 16 * j + p, Kg - 1) * ncol + 8 * q;
    flt4 acc[RB][CB];
    float bacc[RB];
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        bacc[i] = 0.f;
#pragma unroll
        for (int j = 0; j < CB; ++j) acc[i][j] = flt4{0, 0, 0, 0};
    }
    // one register stage of raw rows: a step splits them into bf16 pieces,
    // then reuses the registers for the next step's loads, which are in
    // flight during the step's MFMAs
    // ST stages: the rows of the wave's step k + ST are loaded into the
    // stage step k has just split (the loop unrolled ST times so that every
    // stage is indexed statically)
    float ar[ST][RB][8], br[ST][CB][8];
    auto ld8 = [](float (&d)[8], const float *src) {
        const f4 x = *reinterpret_cast<const f4 *>(src), y = *reinterpret_cast<const f4 *>(src + 4);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            d[e] = x[e];
            d[4 + e] = y[e];
        }
    };
    auto load = [&](float (&a)[RB][8], float (&b)[CB][8], long c) {
#pragma unroll
        for (int i = 0; i < RB; ++i) ld8(a[i], pa[i] + c);
#pragma unroll
        for (int j = 0; j < CB; ++j) ld8(b[j], pb[j] + c);
    };
    auto step = [&](float (&a)[RB][8], float (&b)[CB][8], long c) {
        if (do_bias && c < Np) {   // a step never straddles Np (multiples of 32)
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int e = 0; e < 8; ++e) bacc[i] += a[i][e];
        }
        bf16x8 ah[RB], am[RB], al[RB], bh[CB], bm[CB], bl[CB];
#pragma unroll
        for (int i = 0; i < RB; ++i) split3(a[i], ah[i], am[i], al[i]);
#pragma unroll
        for (int j = 0; j < CB; ++j) split3(b[j], bh[j], bm[j], bl[j]);
        if (c + ST * 4 * SC < ce) load(a, b, c + ST * 4 * SC);
#pragma unroll
        for (int j = 0; j < CB; ++j)
#pragma unroll
            for (int i = 0; i < RB; ++i) acc[i][j] = mma_x6(ah[i], am[i], al[i], bh[j], bm[j], bl[j], acc[i][j]);
    };
    const long c0 = cb + (long)wave * SC;
#pragma unroll
    for (int st = 0; st < ST; ++st)
        if (c0 + st * 4 * SC < ce) load(ar[st], br[st], c0 + st * 4 * SC);
    for (long c = c0; c < ce; c += ST * 4 * SC) {
#pragma unroll
        for (int st = 0; st < ST; ++st)
            if (c + st * 4 * SC < ce) step(ar[st], br[st], c + st * 4 * SC);
    }
    for (int wv = 0; wv < 4; ++wv) {
        if (wave == wv) {
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int idx = (16 * i + 4 * q + r) * LD + 16 * j + p;
                        red[idx] = (wv == 0 ? 0.0 : red[idx]) + (double)acc[i][j][r];
                    }
            if (do_bias) {
                for (int qq = 0; qq < 4; ++qq) {
                    if (q == qq) {
#pragma unroll
                        for (int i = 0; i < RB; ++i) {
                            const int idx = (16 * i + p) * LD + 16 * CB;
                            red[idx] = (wv == 0 && qq == 0 ? 0.0 : red[idx]) + (double)bacc[i];
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                }
            }
        }
        __syncthreads();
    }
    double *P = partial + (long)chk * M * Kd;
    for (int idx = tid; idx < 16 * RB * LD; idx += 256) {
        const int r = idx / LD, cidx = idx % LD;
        const int m = m0 + r;
        if (m >= M) continue;
        if (cidx < 16 * CB) {
            const int k = k0 + cidx;
            if (k < Kg) P[(long)m * Kd + 1 + k] = red[idx];
        } else if (do_bias) {
            P[(long)m * Kd] = red[idx];
        }
    }
}

// K5 with the whole weight block (at most 128 x 128) in one workgroup, for
// one column chunk.  A step of 32 columns: every A (Sbar) and B (T) row
// fragment of the step -- row p of a 16-row block, columns 8 q .. 8 q + 7,
// exactly one lane's operand of a bf16 MFMA -- is loaded by one thread,
// split once and written to LDS in operand order; after a barrier, wave
// (wr, wk) multiplies the pieces of its RW row blocks by its CW column blocks
// (the four waves own disjoint quarters of the block, so no cross-wave sum),
// and a second barrier frees the LDS for the next step, whose loads were
// issued before the products.  Against k_grad_x6 each element is split once
// (there the two 64-column tiles of a 128-wide block split Sbar twice), and
// two workgroups share a CU, one splitting while the other multiplies.  The
// bias column (sum of Sbar over the first Np columns) is summed in FP64 by
// the threads holding the A fragments and met in LDS in a fixed order.  The
// chunk partials and their reduction are those of k_grad_x6.
// Blocks larger than 128 x 128 are cut into ntm x ntk tiles of 128 x 128,
// one workgroup each; the tiles of one chunk are adjacent work items (one
// XCD, so the second read of an operand comes from L2).
template <int MB, int KB, int WR>
__global__ __launch_bounds__(256, 2) void k_grad_x6s(const float *__restrict__ A, const float *__restrict__ B, int M,
                                                      int Kd, long ncol, long Np, long chunk, int nchunk, int ntm,
                                                      int ntk, double *__restrict__ partial)
{
    constexpr int WK = 4 / WR, RW = MB / WR, CW = KB / WK;
    static_assert(RW * WR == MB && CW * WK == KB, "wave grid divides the block grid");
    constexpr int NF = (MB + KB) * 64;   // fragments a step
    constexpr int NJ = (NF + 255) / 256; // fragments a thread
    __shared__ bf16x8 pc[(MB + KB) * 3 * 64];
    const int G = gridDim.x, bid = blockIdx.x;
    const int w = (bid & 7) * (G >> 3) + (bid >> 3);
    const int ntile = ntm * ntk, tile = w % ntile, chk = w / ntile;
    if (chk >= nchunk) return;
    const int m0 = (tile % ntm) * 16 * MB, k0 = (tile / ntm) * 16 * KB;
    const bool do_bias = k0 == 0;
    const int Kg = Kd - 1;
    const long cb = (long)chk * chunk, ce = min(ncol, cb + chunk);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wave % WR, wk = wave / WR;
    typedef float f4 __attribute__((ext_vector_type(4)));
    const float *src[NJ];
    bool live[NJ], isA[NJ];
    double bsum[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int f = tid + 256 * j, blk = f >> 6, ln = f & 63, r16 = ln & 15, qq = ln >> 4;
        bsum[j] = 0.0;
        isA[j] = f < NF && blk < MB;
        if (f >= NF) {
            live[j] = false;
            src[j] = A;
        } else if (blk < MB) {
            const int m = m0 + 16 * blk + r16;
            live[j] = m < M;
            src[j] = A + (long)min(m, M - 1) * ncol + cb + 8 * qq;
        } else {
            const int k = k0 + 16 * (blk - MB) + r16;
            live[j] = k < Kg;
            src[j] = B + (long)min(k, Kg - 1) * ncol + cb + 8 * qq;
        }
    }
    float raw[NJ][8];
    auto load = [&](long off) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            if (!live[j]) continue;
            const f4 x = *reinterpret_cast<const f4 *>(src[j] + off), y = *reinterpret_cast<const f4 *>(src[j] + off + 4);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                raw[j][e] = x[e];
                raw[j][4 + e] = y[e];
            }
        }
    };
    flt4 acc[RW][CW];
#pragma unroll
    for (int i = 0; i < RW; ++i)
#pragma unroll
        for (int j = 0; j < CW; ++j) acc[i][j] = flt4{0, 0, 0, 0};
    if (cb < ce) load(0);
    for (long c = cb; c < ce; c += 32) {
        const bool bias = c < Np;   // a 32-column step never straddles Np
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int f = tid + 256 * j;
            if (f >= NF) continue;
            bf16x8 h, m, l;
            if (live[j]) {
                if (isA[j] && bias && do_bias) {
                    double t = 0.0;
#pragma unroll
                    for (int e = 0; e < 8; ++e) t += (double)raw[j][e];
                    bsum[j] += t;
                }
                split3(raw[j], h, m, l);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) h[e] = m[e] = l[e] = (__bf16)0.f;
            }
            bf16x8 *d = pc + (f >> 6) * 3 * 64 + (f & 63);
            d[0] = h;
            d[64] = m;
            d[128] = l;
        }
        if (c + 32 < ce) load(c + 32 - cb);
        __syncthreads();
        bf16x8 bh[CW], bm[CW], bl[CW];
#pragma unroll
        for (int j = 0; j < CW; ++j) {
            const bf16x8 *s = pc + (MB + wk * CW + j) * 3 * 64 + lane;
            bh[j] = s[0];
            bm[j] = s[64];
            bl[j] = s[128];
        }
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const bf16x8 *s = pc + (wr * RW + i) * 3 * 64 + lane;
            const bf16x8 ah = s[0], am = s[64], al = s[128];
#pragma unroll
            for (int j = 0; j < CW; ++j) acc[i][j] = mma_x6(ah, am, al, bh[j], bm[j], bl[j], acc[i][j]);
        }
        __syncthreads();
    }
    double *P = partial + (long)chk * M * Kd;
    const int q = lane >> 4;
#pragma unroll
    for (int i = 0; i < RW; ++i)
#pragma unroll
        for (int j = 0; j < CW; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + 16 * (wr * RW + i) + 4 * q + r, k = k0 + 16 * (wk * CW + j) + (lane & 15);
                if (m < M && k < Kg) P[(long)m * Kd + 1 + k] = (double)acc[i][j][r];
            }
    // bias column: the four q partials of row m met in LDS, summed in order
    if (!do_bias) return;
    double *red = reinterpret_cast<double *>(pc);   // 16 MB x 4 doubles, free after the last barrier
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int f = tid + 256 * j;
        if (isA[j]) red[((f >> 6) * 16 + (f & 15)) * 4 + ((f & 63) >> 4)] = bsum[j];
    }
    __syncthreads();
    for (int m = tid; m < 16 * MB && m0 + m < M; m += 256) {
        const double *rr = red + m * 4;
        P[(long)(m0 + m) * Kd] = ((rr[0] + rr[1]) + rr[2]) + rr[3];
    }
}

// resident workgroups a CU the one-block K5 is given at most: each is a
// column chunk with its own FP64 partial, so more of them cost reduction
// traffic (four measured slower than two: dense4_k3 reduction 0.078 ->
// 0.100 ms, K5 no faster)
#ifndef GRAD_X6S_MAXWG
#define GRAD_X6S_MAXWG 2
#endif

// the wave grid of k_grad_x6s for an MB x KB block grid (0: not handled)
static int grad_x6s_wr(int MB, int KB)
{
    if (MB >= 2 && KB >= 2) return 2;
    if (MB >= 4) return 4;
    if (KB >= 4) return 1;
    return 0;
}

// 16-row blocks of a tile (a power of two, at most 8) and tiles along n
static int pow2_blocks(int n)
{
    const int b = (n + 15) / 16;
    return b <= 1 ? 1 : b <= 2 ? 2 : b <= 4 ? 4 : 8;
}
static int tiles_of(int n) { return (n + 127) / 128; }

template <int MB, int KB>
static const void *grad_x6s_fn()
{
    constexpr int WR = (MB >= 2 && KB >= 2) ? 2 : (MB >= 4 ? 4 : 1);
    return reinterpret_cast<const void *>(&k_grad_x6s<MB, KB, WR>);
}

template <int MB, int KB>
static void grad_x6s_go(const float *A, const float *B, int M, int Kd, long ncol, long Np, double *partial,
                        int nchunk, long chunk, hipStream_t s)
{
    constexpr int WR = (MB >= 2 && KB >= 2) ? 2 : (MB >= 4 ? 4 : 1);
    const int ntm = tiles_of(M), ntk = tiles_of(Kd - 1);
    long g = (long)ntm * ntk * nchunk;
    g = (g + 7) / 8 * 8;
    hipLaunchKernelGGL((k_grad_x6s<MB, KB, WR>), dim3((unsigned)g), dim3(256), 0, s, A, B, M, Kd, ncol, Np, chunk,
                       nchunk, ntm, ntk, partial);
}

// dispatch over the power-of-two block counts: fn(MB, KB) instantiated
#define GRAD_X6S_CASES(X)                                                                                            \
    X(8, 8) X(8, 4) X(8, 2) X(8, 1) X(4, 8) X(4, 4) X(4, 2) X(4, 1) X(2, 8) X(2, 4) X(2, 2) X(1, 8) X(1, 4)

int grad_x6s_wgs(int M, int Kd, bool any)
{
    if (!any && !set_lin_x6c(-1)) return 0;
    const int MB = pow2_blocks(M), KB = pow2_blocks(Kd - 1);
    if (!MB || !KB || !grad_x6s_wr(MB, KB)) return 0;
    const void *fn = nullptr;
#define X(a, b)                                                                                                      \
    if (MB == a && KB == b) fn = grad_x6s_fn<a, b>();
    GRAD_X6S_CASES(X)
#undef X
    if (!fn) return 0;
    // resident workgroups a CU, queried once per instantiation
    const int n = occupancy_cached(fn, 0);
    return std::max(1, 256 * std::min(n, GRAD_X6S_MAXWG) / (tiles_of(M) * tiles_of(Kd - 1)));
}

static bool launch_grad_x6s(const float *A, const float *B, int M, int Kd, long ncol, long Np, double *partial,
                            int nchunk, long chunk, hipStream_t s)
{
    if (!set_lin_x6c(-1)) return false;
    const int MB = pow2_blocks(M), KB = pow2_blocks(Kd - 1);
    if (!MB || !KB || !grad_x6s_wr(MB, KB) || chunk % 32 != 0) return false;
#define X(a, b)                                                                                                      \
    if (MB == a && KB == b) {                                                                                        \
        grad_x6s_go<a, b>(A, B, M, Kd, ncol, Np, partial, nchunk, chunk, s);                                        \
        return true;                                                                                                 \
    }
    GRAD_X6S_CASES(X)
#undef X
    return false;
}

// K5 and K6 of one layer in one pass over S-bar (BF16x6): the one-block K5
// above (the whole M x Kd weight block in one workgroup, one column chunk
// a workgroup), and for the same 32-column steps T-bar = W' S-bar, so that
// S-bar is read from HBM once for both.  K6 needs S-bar as its B operand,
// eight consecutive rows of one column a lane, where K5 holds eight
// consecutive columns of one row: the threads that load S-bar's fragments
// also write their raw values to LDS transposed (St[column][row]), and K6
// reads its fragments from there (two 16-byte reads a lane) and splits them.
// Each wave holds its row block of W' split in registers, in the operand
// order of k_lin_x6c's LDS, and each 16 x 16 block of T-bar takes the six products
// of every 32-row chunk of S-bar in chunk order with an FP32 accumulator, as
// k_lin_x6c does: the same products in the same order, so T-bar has the same
// bits as the separate K6.  The K5 partials are k_grad_x6s's over this
// kernel's column chunks (three workgroups a CU: W' in registers, 42 KB of
// LDS).  For layers of 32 or 64 rows over at most 64 inputs (MB in {2, 4},
// KB <= 4).
template <int MB, int KB, int WR>
__global__ __launch_bounds__(256, 3) void k_grad_lint_x6(const float *__restrict__ A, const float *__restrict__ B,
                                                         const float *__restrict__ W, float *__restrict__ Tb, int M,
                                                         int Kd, long ncol, long Np, long chunk, int nchunk,
                                                         double *__restrict__ partial)
{
    constexpr int WK = 4 / WR, RW = MB / WR, CW = KB / WK;
    static_assert(RW * WR == MB && CW * WK == KB, "wave grid divides the block grid");
    constexpr int NF = (MB + KB) * 64;   // fragments a step
    constexpr int NJ = (NF + 255) / 256; // fragments a thread
    constexpr int NKC = MB / 2;          // 32-row chunks of S-bar (K6's reduction)
    constexpr int STL = 16 * MB + 4;     // St row length (floats)
    static_assert(MB == 2 || MB == 4, "K6 reduction of whole 32-row chunks");
    static_assert(KB <= 4, "one K6 row block a wave");
    __shared__ bf16x8 pc[(MB + KB) * 3 * 64];
    __shared__ __attribute__((aligned(16))) float St[64 * STL];
    const int G = gridDim.x, bid = blockIdx.x;
    const int chk = (bid & 7) * (G >> 3) + (bid >> 3);
    if (chk >= nchunk) return;
    const int Kg = Kd - 1;
    const long cb = (long)chk * chunk, ce = min(ncol, cb + chunk);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wave % WR, wk = wave / WR;
    typedef float f4 __attribute__((ext_vector_type(4)));
    // W' pieces of the wave's K6 row block rb = wave (rows of T-bar: the Kg
    // outputs; reduction: the M rows of S-bar), in registers, as k_lin_x6c's
    // LDS holds them: lane (m = lane % 16, q = lane / 16) has rows
    // 32 kc + 8 q .. +7 of output row 16 rb + m
    bf16x8 wh[NKC], wm[NKC], wl[NKC];
    {
        const int m = 16 * wave + (lane & 15);
#pragma unroll
        for (int kc = 0; kc < NKC; ++kc) {
            const int k = 32 * kc + 8 * (lane >> 4);
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = (wave < KB && m < Kg && k + e < M) ? W[(long)(k + e) * Kd + 1 + m] : 0.f;
            split3(v, wh[kc], wm[kc], wl[kc]);
        }
    }
    const float *src[NJ];
    bool live[NJ], isA[NJ];
    double bsum[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int f = tid + 256 * j, blk = f >> 6, ln = f & 63, r16 = ln & 15, qq = ln >> 4;
        bsum[j] = 0.0;
        isA[j] = f < NF && blk < MB;
        if (f >= NF) {
            live[j] = false;
            src[j] = A;
        } else if (blk < MB) {
            const int m = 16 * blk + r16;
            live[j] = m < M;
            src[j] = A + (long)min(m, M - 1) * ncol + cb + 8 * qq;
        } else {
            const int k = 16 * (blk - MB) + r16;
            live[j] = k < Kg;
            src[j] = B + (long)min(k, Kg - 1) * ncol + cb + 8 * qq;
        }
    }
    float raw[NJ][8];
    auto load = [&](long off) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            if (!live[j]) continue;
            const f4 x = *reinterpret_cast<const f4 *>(src[j] + off), y = *reinterpret_cast<const f4 *>(src[j] + off + 4);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                raw[j][e] = x[e];
                raw[j][4 + e] = y[e];
            }
        }
    };
    flt4 acc[RW][CW];
#pragma unroll
    for (int i = 0; i < RW; ++i)
#pragma unroll
        for (int j = 0; j < CW; ++j) acc[i][j] = flt4{0, 0, 0, 0};
    if (cb < ce) load(0);
    const int q = lane >> 4, n16 = lane & 15;
    for (long c = cb; c < ce; c += 32) {
        const bool bias = c < Np;   // a 32-column step never straddles Np
        // steps pair up: St holds the S-bar columns of two steps, and K6 runs
        // on the second over all 64 (a chunk is a whole number of 64-column pairs)
        const int half = (int)((c - cb) >> 5) & 1;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int f = tid + 256 * j;
            if (f >= NF) continue;
            bf16x8 h, m, l;
            if (live[j]) {
                if (isA[j] && bias) {
                    double t = 0.0;
#pragma unroll
                    for (int e = 0; e < 8; ++e) t += (double)raw[j][e];
                    bsum[j] += t;
                }
                split3(raw[j], h, m, l);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) h[e] = m[e] = l[e] = (__bf16)0.f;
            }
            bf16x8 *d = pc + (f >> 6) * 3 * 64 + (f & 63);
            d[0] = h;
            d[64] = m;
            d[128] = l;
            if (isA[j]) {
                // row 16 blk + r16 of S-bar, columns 8 qq .. 8 qq + 7, transposed
                const int row = 16 * (f >> 6) + (f & 15), c8 = 32 * half + 8 * ((f & 63) >> 4);
#pragma unroll
                for (int e = 0; e < 8; ++e) St[(c8 + e) * STL + row] = live[j] ? raw[j][e] : 0.f;
            }
        }
        if (c + 32 < ce) load(c + 32 - cb);
        __syncthreads();
        bf16x8 bh[CW], bm[CW], bl[CW];
#pragma unroll
        for (int j = 0; j < CW; ++j) {
            const bf16x8 *s = pc + (MB + wk * CW + j) * 3 * 64 + lane;
            bh[j] = s[0];
            bm[j] = s[64];
            bl[j] = s[128];
        }
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const bf16x8 *s = pc + (wr * RW + i) * 3 * 64 + lane;
            const bf16x8 ah = s[0], am = s[64], al = s[128];
#pragma unroll
            for (int j = 0; j < CW; ++j) acc[i][j] = mma_x6(ah, am, al, bh[j], bm[j], bl[j], acc[i][j]);
        }
        // K6 on the second step of a pair: row block rb = wave of T-bar over
        // the pair's 64 columns as four MFMA column blocks, block j taking
        // column 4 n + j (n = lane % 16), so a lane stores four adjacent
        // columns of a row (sixteen lanes: 256 bytes)
        if (half && wave < KB) {
            {
                const int rb = wave;
                flt4 t[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    t[j] = flt4{0, 0, 0, 0};
#pragma unroll
                    for (int kc = 0; kc < NKC; ++kc) {
                        const float *sp = St + (4 * n16 + j) * STL + 32 * kc + 8 * q;
                        const f4 x = *reinterpret_cast<const f4 *>(sp), y = *reinterpret_cast<const f4 *>(sp + 4);
                        float v[8];
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            v[e] = x[e];
                            v[4 + e] = y[e];
                        }
                        bf16x8 sh, sm, sl;
                        split3(v, sh, sm, sl);
                        t[j] = mma_x6(wh[kc], wm[kc], wl[kc], sh, sm, sl, kc == 0 ? flt4{0, 0, 0, 0} : t[j]);
                    }
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * rb + 4 * q + r;
                    if (row < Kg)
                        *reinterpret_cast<f4 *>(Tb + (long)row * ncol + (c - 32) + 4 * n16) =
                            f4{t[0][r], t[1][r], t[2][r], t[3][r]};
                }
            }
        }
        __syncthreads();
    }
    double *P = partial + (long)chk * M * Kd;
#pragma unroll
    for (int i = 0; i < RW; ++i)
#pragma unroll
        for (int j = 0; j < CW; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = 16 * (wr * RW + i) + 4 * q + r, k = 16 * (wk * CW + j) + n16;
                if (m < M && k < Kg) P[(long)m * Kd + 1 + k] = (double)acc[i][j][r];
            }
    double *red = reinterpret_cast<double *>(pc);   // 16 MB x 4 doubles, free after the last barrier
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int f = tid + 256 * j;
        if (isA[j]) red[((f >> 6) * 16 + (f & 15)) * 4 + ((f & 63) >> 4)] = bsum[j];
    }
    __syncthreads();
    for (int m = tid; m < 16 * MB && m < M; m += 256) {
        const double *rr = red + m * 4;
        P[(long)m * Kd] = ((rr[0] + rr[1]) + rr[2]) + rr[3];
    }
}

// 0: K5 and K6 of a layer as separate kernels even where the fused one applies
static int g_grad_lint = 1;
int set_grad_lint_x6(int on)
{
    const int prev = g_grad_lint;
    if (on >= 0) g_grad_lint = on ? 1 : 0;
    return prev;
}

// whether the fused K5 + K6 takes an M x Kd weight block (M rows: the
// layer; Kd - 1: the layer below); chunks as grad_x6s_wgs gives them
static bool grad_lint_x6_fits(int M, int Kd)
{
    const int kb = pow2_blocks(Kd - 1);
    return (M == 32 || M == 64) && Kd - 1 >= 1 && Kd - 1 <= 64 && !(M == 32 && kb == 1);
}

bool grad_lint_x6_ok(int M, int Kd) { return g_grad_lint && set_lin_x6c(-1) && grad_lint_x6_fits(M, Kd); }

// column chunks the fused kernel takes: one resident round at three
// workgroups a CU (its LDS allows three), 0 where it does not apply
int grad_lint_x6_wgs(int M, int Kd, bool any)
{
    return (any ? grad_lint_x6_fits(M, Kd) : grad_lint_x6_ok(M, Kd)) ? 256 * 3 : 0;
}

bool launch_grad_lint_x6(const float *A, const float *B, const float *W, float *Tb, int M, int Kd, long ncol, long Np,
                         double *partial, int nchunk, long chunk, hipStream_t s)
{
    if (!grad_lint_x6_ok(M, Kd) || chunk % 64 != 0 || ncol % 64 != 0) return false;
    const int MB = M / 16, KB = pow2_blocks(Kd - 1);
    long g = ((long)nchunk + 7) / 8 * 8;
#define GL(a, b)                                                                                                     \
    if (MB == a && KB == b) {                                                                                        \
        constexpr int WR = (a >= 2 && b >= 2) ? 2 : (a >= 4 ? 4 : 1);                                               \
        hipLaunchKernelGGL((k_grad_lint_x6<a, b, WR>), dim3((unsigned)g), dim3(256), 0, s, A, B, W, Tb, M, Kd, ncol,  \
                           Np, chunk, nchunk, partial);                                                              \
        return true;                                                                                                 \
    }
    GL(4, 4) GL(4, 2) GL(4, 1) GL(2, 4) GL(2, 2)
#undef GL
    return false;
}

// register stages of raw rows of the split K5 at 128 and at 64 rows (two
// stages spill at 128 rows, and gain nothing measurable at 64)
#ifndef GRAD_X6_ST8
#define GRAD_X6_ST8 1
#endif
#ifndef GRAD_X6_ST4
#define GRAD_X6_ST4 1
#endif

// tile of the split K5: (16 RB) x 64 of the weight block, RB = 8 where the
// block has more than 64 rows (each split element then feeds more MFMAs)
// weight blocks taller than this take the 128-row tile (64-row tiles with
// two register stages measured 6-13 % slower on 128- and 256-wide layers)
#ifndef GRAD_X6_RB8_ABOVE
#define GRAD_X6_RB8_ABOVE 64
#endif
int grad_x6_rb(int M) { return M > GRAD_X6_RB8_ABOVE ? 8 : 4; }

void launch_grad_x6(const float *A, const float *B, int M, int Kd, long ncol, long Np, double *partial,
                    int nchunk, long chunk, hipStream_t s)
{
    if (launch_grad_x6s(A, B, M, Kd, ncol, Np, partial, nchunk, chunk, s)) return;
    const int rb = grad_x6_rb(M);
    const int ntm = (M + 16 * rb - 1) / (16 * rb), ntk = (Kd - 1 + 63) / 64;
    long g = (long)ntm * ntk * nchunk;
    g = (g + 7) / 8 * 8;
    if (rb == 8)
        hipLaunchKernelGGL((k_grad_x6<8, 4, GRAD_X6_ST8>), dim3((unsigned)g), dim3(256), 0, s, A, B, M, Kd, ncol, Np,
                           chunk, ntm, ntk, nchunk, partial);
    else
        hipLaunchKernelGGL((k_grad_x6<4, 4, GRAD_X6_ST4>), dim3((unsigned)g), dim3(256), 0, s, A, B, M, Kd, ncol, Np,
                           chunk, ntm, ntk, nchunk, partial);
}

}  // namespace dnnf90
