// gemm.hip -- the linear maps of the derivative-carrying pass on MFMA.
//
//   K1  S   = W  [1; T]   (lib/train_module.f90:262-278)
//   K6  Tb  = W' Sbar     (lib/train_module.f90:430-444, bias column skipped)
//   K5  dW += Sbar [1; T]^T summed over the column axis
//                         (lib/train_module.f90:408-422)
//
// The column axis is slot * Np + point for every layer, so each map is one
// GEMM over na*Np columns.  Row 0 of [1; T] is the bias channel: 1 on the
// first Np columns (slot 0), 0 elsewhere; it is generated, never stored.
// The MFMA operand layouts are described in mfma.hpp; the BF16 three-way
// split forms of K1, K5 and K6 are in gemm_x6.hip.
#include <hip/hip_bf16.h>

#include <algorithm>

#include "kernels.hpp"
#include "mfma.hpp"

namespace dnnf90 {

// value of row gk of [1; B] (bias row first) at column gc
template <typename T, bool kBiasRow>
__device__ __forceinline__ T bext(const T *__restrict__ B, long ncol, long Np, int gk, int Kd, long gc)
{
    if (gk >= Kd) return T(0);
    if (kBiasRow && gk == 0) return gc < Np ? T(1) : T(0);
    return B[(long)(kBiasRow ? gk - 1 : gk) * ncol + gc];
}

// ---------------------------------------------------------------------------
// K1 / K6, FP64 and FP32.  A workgroup owns 16*RBT rows of C.  Their block of
// A (K1: the weight columns after the bias; K6: the transposed weights) is
// staged once into LDS in MFMA operand order,
//     As[rb][s][lane] = A[16 rb + lane%16][4 s + lane/16],
// zero-padded to a whole number of KC-step chunks, so the A operand of every
// MFMA is one conflict-free 256-byte LDS row.  The k extent of A is the whole
// reduction, so B is never staged: each wave streams 16-column blocks of B
// from HBM straight into registers (lane: column lane%16, rows 4 s + lane/16;
// sixteen lanes read 64 contiguous bytes of a row), KC k-steps at a time with
// the next chunk's loads issued before the current chunk's MFMAs, and writes
// its 16-column strip of C.  K1's bias enters in the epilogue: it feeds slot
// 0 only, i.e. the first Np columns (Np is a multiple of 64, so a 16-column
// block lies wholly inside or outside).  The four waves of a workgroup take
// adjacent column blocks.
// kAcc64 (FP32 K1 of the mixed mode): every MFMA starts from zero and its
// four-term sums are accumulated in FP64, which removes the FP32
// accumulation error of the long k sums (the dominant error of an
// all-FP32 forward pass).
template <typename T, bool kAcc64> struct LinAcc {
    typename Mfma<T>::V v;
    __device__ __forceinline__ void zero() { v = typename Mfma<T>::V{0, 0, 0, 0}; }
    __device__ __forceinline__ void mma(T a, T b) { v = Mfma<T>::mma(a, b, v); }
    __device__ __forceinline__ T get(int r, T add) const { return v[r] + add; }
};
template <> struct LinAcc<float, true> {
    double v[4];
    __device__ __forceinline__ void zero() { v[0] = v[1] = v[2] = v[3] = 0.0; }
    __device__ __forceinline__ void mma(float a, float b)
    {
        const flt4 d = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, flt4{0, 0, 0, 0}, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] += (double)d[r];
    }
    __device__ __forceinline__ float get(int r, float add) const { return (float)(v[r] + (double)add); }
};

template <typename T, bool kTrans, bool kAcc64, int RBT, int KC>
__global__ __launch_bounds__(256) void k_lin(const T *__restrict__ W, int ldw, const T *__restrict__ B,
                                             T *__restrict__ C, int M, int Kd, long ncol, long Np, int nrt,
                                             int ncw, long cbpw)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lin_lds[];
    T *As = reinterpret_cast<T *>(lin_lds);
    // work item: consecutive items on one XCD (workgroup ids are dealt to the
    // eight XCDs in turn), row tiles of one column range adjacent, so the
    // second row tile's reads of B hit that XCD's L2
    const int G = gridDim.x, bid = blockIdx.x;
    const int w = (bid & 7) * (G >> 3) + (bid >> 3);
    const int rt = w % nrt, cw = w / nrt;
    if (cw >= ncw) return;
    const long ncb = ncol >> 4;
    const long cb0 = (long)cw * cbpw, cb1 = min(ncb, cb0 + cbpw);
    if (cb0 >= cb1) return;
    const int ks = (Kd + 3) >> 2, nch = (ks + KC - 1) / KC, ksp = nch * KC;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int m0 = rt * 16 * RBT;
    for (int e = tid; e < RBT * ksp * 64; e += 256) {
        const int rb = e / (ksp * 64), rem = e - rb * ksp * 64, s = rem >> 6, ln = rem & 63;
        const int m = m0 + 16 * rb + (ln & 15), k = 4 * s + (ln >> 4);
        T v = T(0);
        if (m < M && k < Kd) v = kTrans ? W[(long)k * ldw + 1 + m] : W[(long)m * ldw + 1 + k];
        As[e] = v;
    }
    T *Bias = As + RBT * ksp * 64;   // K1: the bias column of the tile's rows
    if (!kTrans && tid < 16 * RBT) Bias[tid] = m0 + tid < M ? W[(long)(m0 + tid) * ldw] : T(0);
    __syncthreads();
    long cb = cb0 + wave;
    if (cb >= cb1) return;
    // rows of B this lane reads, clamped into B (A is zero past Kd)
    const int kq = lane >> 4;
    const T *bl = B + (lane & 15);
    T bc[KC], bn[KC];
    auto load = [&](T (&dst)[KC], long cbl, int ch) {
        const T *src = bl + cbl * 16;
#pragma unroll
        for (int e = 0; e < KC; ++e) dst[e] = src[(long)min(4 * (ch * KC + e) + kq, Kd - 1) * ncol];
    };
    LinAcc<T, kAcc64> acc[RBT];
    load(bc, cb, 0);
    int ch = 0;
    for (;;) {
        long cbn = cb;
        int chn = ch + 1;
        if (chn == nch) {
            chn = 0;
            cbn = cb + 4;
        }
        const bool more = cbn < cb1;
        if (more) load(bn, cbn, chn);
        if (ch == 0) {
#pragma unroll
            for (int rb = 0; rb < RBT; ++rb) acc[rb].zero();
        }
        const T *Ac = As + (long)ch * KC * 64 + lane;
#pragma unroll
        for (int e = 0; e < KC; ++e)
#pragma unroll
            for (int rb = 0; rb < RBT; ++rb) acc[rb].mma(Ac[(rb * ksp + e) * 64], bc[e]);
        if (ch == nch - 1) {
            const long c = cb * 16 + (lane & 15);
            const bool bias = !kTrans && cb * 16 < Np;
#pragma unroll
            for (int rb = 0; rb < RBT; ++rb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m0 + 16 * rb + Mfma<T>::row(lane, r);
                    const T add = bias ? Bias[16 * rb + Mfma<T>::row(lane, r)] : T(0);
                    if (m < M) C[(long)m * ncol + c] = acc[rb].get(r, add);
                }
        }
        if (!more) break;
#pragma unroll
        for (int e = 0; e < KC; ++e) bc[e] = bn[e];
        cb = cbn;
        ch = chn;
    }
}

// ---------------------------------------------------------------------------
// K1 / K6 with BF16 hi+lo split operands: x = hi + lo, hi = bf16(x),
// lo = bf16(x - hi); a product keeps lo*hi + hi*lo + hi*hi (lo*lo is below
// 2^-32 relative) with FP32 accumulation on v_mfma_f32_16x16x32_bf16.  The
// operands are split while staging into LDS as [row][k] so a lane's eight
// consecutive k are one 16-byte read.
static constexpr int XBK = 32, XPAD = 8, BN = 64;

__device__ __forceinline__ void split8(const float (&v)[8], bf16x8 &hi, bf16x8 &lo)
{
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 h = (__bf16)v[e];
        hi[e] = h;
        lo[e] = (__bf16)(v[e] - (float)h);
    }
}

template <bool kTrans, int BM>
__global__ __launch_bounds__(256) void k_gemm_wx_bf3(const float *__restrict__ W, int ldw,
                                                     const float *__restrict__ B, float *__restrict__ C,
                                                     int M, int Kd, long ncol, long Np)
{
    constexpr int RB = BM / 32;
    __shared__ __attribute__((aligned(16))) __bf16 Ah[BM][XBK + XPAD];
    __shared__ __attribute__((aligned(16))) __bf16 Al[BM][XBK + XPAD];
    __shared__ __attribute__((aligned(16))) __bf16 Bh[BN][XBK + XPAD];
    __shared__ __attribute__((aligned(16))) __bf16 Bl[BN][XBK + XPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * BM;
    const long n0 = (long)blockIdx.x * BN;
    flt4 acc[RB][2];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) acc[i][jj] = flt4{0, 0, 0, 0};

    for (int k0 = 0; k0 < Kd; k0 += XBK) {
        // A: BM rows x 32 k; each thread splits runs of 8 consecutive k of one row
        for (int u = tid; u < BM * (XBK / 8); u += 256) {
            const int m = u / (XBK / 8), kq = (u % (XBK / 8)) * 8, gm = m0 + m;
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int gk = k0 + kq + e;
                v[e] = (gm < M && gk < Kd) ? (kTrans ? W[(long)gk * ldw + 1 + gm] : W[(long)gm * ldw + gk]) : 0.f;
            }
            bf16x8 hi, lo;
            split8(v, hi, lo);
            *reinterpret_cast<bf16x8 *>(&Ah[m][kq]) = hi;
            *reinterpret_cast<bf16x8 *>(&Al[m][kq]) = lo;
        }
        // B: 32 k x 64 columns; a thread takes 8 consecutive k of one column
        // (each of its 8 loads is one coalesced 64-column row segment per wave)
        {
            const int c = tid & 63, kq = (tid >> 6) * 8;
            const long gc = n0 + c;
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = bext<float, !kTrans>(B, ncol, Np, k0 + kq + e, Kd, gc);
            bf16x8 hi, lo;
            split8(v, hi, lo);
            *reinterpret_cast<bf16x8 *>(&Bh[c][kq]) = hi;
            *reinterpret_cast<bf16x8 *>(&Bl[c][kq]) = lo;
        }
        __syncthreads();
        bf16x8 ah[RB], al[RB], bh[2], bl[2];
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const int r = wm * (BM / 2) + i * 16 + (lane & 15);
            ah[i] = *reinterpret_cast<const bf16x8 *>(&Ah[r][8 * (lane >> 4)]);
            al[i] = *reinterpret_cast<const bf16x8 *>(&Al[r][8 * (lane >> 4)]);
        }
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
            const int r = wn * 32 + jj * 16 + (lane & 15);
            bh[jj] = *reinterpret_cast<const bf16x8 *>(&Bh[r][8 * (lane >> 4)]);
            bl[jj] = *reinterpret_cast<const bf16x8 *>(&Bl[r][8 * (lane >> 4)]);
        }
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) {
                acc[i][jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[jj], acc[i][jj], 0, 0, 0);
                acc[i][jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[jj], acc[i][jj], 0, 0, 0);
                acc[i][jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[jj], acc[i][jj], 0, 0, 0);
            }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm * (BM / 2) + i * 16 + 4 * (lane >> 4) + r;
                const long gc = n0 + wn * 32 + jj * 16 + (lane & 15);
                if (gm < M) C[(long)gm * ncol + gc] = acc[i][jj][r];
            }
}

template <typename T, bool kTrans, bool kAcc64, int RBT, int KC>
static void lin_go(const T *W, int ldw, const T *B, T *C, int M, int Kd, long ncol, long Np, hipStream_t s)
{
    const int ks = (Kd + 3) / 4, ksp = (ks + KC - 1) / KC * KC;
    const size_t lds = ((size_t)RBT * ksp * 64 + 16 * RBT) * sizeof(T);
    lds_allow(reinterpret_cast<const void *>(&k_lin<T, kTrans, kAcc64, RBT, KC>), lds);
    const int nrt = (M + 16 * RBT - 1) / (16 * RBT);
    const long ncb = ncol / 16;
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, (160 * 1024) / std::max<size_t>(lds, 1)));
    const long want = std::max(1L, (long)256 * per_cu / nrt);
    long cbpw = (ncb + want - 1) / want;
    cbpw = (cbpw + 3) / 4 * 4;
    long ncw = (ncb + cbpw - 1) / cbpw;
    long g = (long)nrt * ncw;
    g = (g + 7) / 8 * 8;
    hipLaunchKernelGGL((k_lin<T, kTrans, kAcc64, RBT, KC>), dim3((unsigned)g), dim3(256), lds, s, W, ldw, B, C, M,
                       Kd, ncol, Np, nrt, (int)ncw, cbpw);
}

template <typename T, bool kTrans, bool kAcc64, int RBT>
static void lin_kc(const T *W, int ldw, const T *B, T *C, int M, int Kd, long ncol, long Np, hipStream_t s)
{
    // chunks of 8 k-steps; a reduction of under 32 terms is one exact chunk
    const int ks = (Kd + 3) / 4;
    if (ks >= 8) lin_go<T, kTrans, kAcc64, RBT, 8>(W, ldw, B, C, M, Kd, ncol, Np, s);
    else if (ks >= 4) lin_go<T, kTrans, kAcc64, RBT, 4>(W, ldw, B, C, M, Kd, ncol, Np, s);
    else lin_go<T, kTrans, kAcc64, RBT, 1>(W, ldw, B, C, M, Kd, ncol, Np, s);
}

// Kd: length of the reduction (K1: the ndm weight columns after the bias;
// K6: nd)
template <typename T, bool kTrans, bool kAcc64>
static void lin_pick(const T *W, int ldw, const T *B, T *C, int M, int Kd, long ncol, long Np, hipStream_t s)
{
    // rows per workgroup: the whole of M up to 128, fewer where the staged
    // block would pass 144 KiB of LDS
    const size_t per_rb = (size_t)(((Kd + 3) / 4 + 7) / 8 * 8) * 64 * sizeof(T);
    int rbt = 8;
    while (rbt > 1 && 16 * (rbt / 2) >= M) rbt /= 2;
    while (rbt > 1 && rbt * per_rb > (size_t)144 * 1024) rbt /= 2;
    switch (rbt) {
    case 8: lin_kc<T, kTrans, kAcc64, 8>(W, ldw, B, C, M, Kd, ncol, Np, s); break;
    case 4: lin_kc<T, kTrans, kAcc64, 4>(W, ldw, B, C, M, Kd, ncol, Np, s); break;
    case 2: lin_kc<T, kTrans, kAcc64, 2>(W, ldw, B, C, M, Kd, ncol, Np, s); break;
    default: lin_kc<T, kTrans, kAcc64, 1>(W, ldw, B, C, M, Kd, ncol, Np, s); break;
    }
}

template <typename T>
void launch_gemm_wx(int prec, bool trans_a, const T *W, int ldw, const T *B, T *C, int M, int Kd, long ncol,
                    long Np, hipStream_t s)
{
    if constexpr (sizeof(T) == 4) {
        if (prec == PREC_BF16X3) {
            const bool big = M > 64;
            const int bm = big ? 128 : 64;
            dim3 grid((unsigned)(ncol / BN), (unsigned)((M + bm - 1) / bm));
#define LAUNCH(K, ...) hipLaunchKernelGGL((K), grid, dim3(256), 0, s, __VA_ARGS__)
            if (trans_a) {
                if (big) LAUNCH((k_gemm_wx_bf3<true, 128>), W, ldw, B, C, M, Kd, ncol, Np);
                else LAUNCH((k_gemm_wx_bf3<true, 64>), W, ldw, B, C, M, Kd, ncol, Np);
            } else {
                if (big) LAUNCH((k_gemm_wx_bf3<false, 128>), W, ldw, B, C, M, Kd, ncol, Np);
                else LAUNCH((k_gemm_wx_bf3<false, 64>), W, ldw, B, C, M, Kd, ncol, Np);
            }
#undef LAUNCH
            return;
        }
        if (prec == PREC_BF16X6) {
            launch_lin_x6(trans_a, W, ldw, B, C, M, Kd, ncol, Np, s);
            return;
        }
        // FP32 mode: the forward map accumulates in FP64
        if (trans_a) lin_pick<T, true, false>(W, ldw, B, C, M, Kd, ncol, Np, s);
        else lin_pick<T, false, true>(W, ldw, B, C, M, Kd - 1, ncol, Np, s);
        return;
    }
    if (trans_a) lin_pick<T, true, false>(W, ldw, B, C, M, Kd, ncol, Np, s);
    else lin_pick<T, false, false>(W, ldw, B, C, M, Kd - 1, ncol, Np, s);
}

// ---------------------------------------------------------------------------
// K5: partial[chunk][m][k] = sum over the chunk's columns of A[m][c] Bext[k][c].
// The reduction runs along the contiguous column axis, so no operand is
// staged in LDS: the sum over columns may visit them in any order as long as
// A and B visit them alike, so lane (p = lane%16, q = lane/16) of a 16x16x4
// MFMA takes, for row p of its A block and row p of its B block, the V
// consecutive columns c0 + q V .. c0 + q V + V-1, and
// feeds them to V successive MFMAs as its k = q element (V: one or two
// 16-byte loads).  A wave owns the
// whole (16 RB) x (16 CB) output tile and every fourth step of the
// workgroup's column chunk; the next step's loads are issued before the
// current step's MFMAs.  The four waves' tiles are summed in LDS in a fixed
// order and written as one FP64 partial per chunk.
// With a bias row (K5), B holds the Kd-1 rows after it and the bias column
// of the gradient, the sum of A over the first Np columns (slot 0), is added
// up by the lanes of the first k tile from the A values they already hold.

template <typename T, int RB, int CB, int NH>
__global__ __launch_bounds__(256) void k_grad_nt(const T *__restrict__ A, const T *__restrict__ B, int M, int Kd,
                                                 long ncol, long Np, long chunk, int bias_row, int ntm, int ntk,
                                                 int nchunk, double *__restrict__ partial)
{
    constexpr int H = 16 / sizeof(T);   // elements per 16-byte load
    constexpr int V = NH * H;           // columns per lane per step
    constexpr int SC = 4 * V;           // columns per wave step
    constexpr int LD = 16 * CB + 1;     // reduction row: the tile, then the bias column
    typedef typename Vec2x16<T>::V VT;
    __shared__ double red[16 * RB * LD];
    const int G = gridDim.x, bid = blockIdx.x;
    const int w = (bid & 7) * (G >> 3) + (bid >> 3);
    const int ntile = ntm * ntk;
    const int tile = w % ntile, chk = w / ntile;
    if (chk >= nchunk) return;
    const int koff = bias_row ? 1 : 0, Kg = Kd - koff;   // GEMM columns k <-> weight column k + koff
    const int m0 = (tile % ntm) * 16 * RB, k0 = (tile / ntm) * 16 * CB;
    const bool do_bias = bias_row && k0 == 0;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int p = lane & 15, q = lane >> 4;
    const long cb = (long)chk * chunk, ce = min(ncol, cb + chunk);

    // per-lane row pointers, clamped into the operands: a row past M or Kg
    // only feeds output entries that are never written
    const T *pa[RB];
    const T *pb[CB];
#pragma unroll
    for (int i = 0; i < RB; ++i) pa[i] = A + (long)min(m0 + 16 * i + p, M - 1) * ncol + q * V;
#pragma unroll
    for (int j = 0; j < CB; ++j) pb[j] = B + (long)min(k0 + 16 * j + p, Kg - 1) * ncol + q * V;
    typename Mfma<T>::V acc[RB][CB];
    T bacc[RB];
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        bacc[i] = T(0);
#pragma unroll
        for (int j = 0; j < CB; ++j) acc[i][j] = typename Mfma<T>::V{0, 0, 0, 0};
    }

    VT a0[RB][NH], b0[CB][NH], a1[RB][NH], b1[CB][NH];
    auto load = [&](VT (&ar)[RB][NH], VT (&br)[CB][NH], long c) {
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int h = 0; h < NH; ++h) ar[i][h] = *reinterpret_cast<const VT *>(pa[i] + c + h * H);
#pragma unroll
        for (int j = 0; j < CB; ++j)
#pragma unroll
            for (int h = 0; h < NH; ++h) br[j][h] = *reinterpret_cast<const VT *>(pb[j] + c + h * H);
    };
    auto mul = [&](const VT (&ar)[RB][NH], const VT (&br)[CB][NH], long c) {
        if (do_bias && c < Np) {   // a step never straddles Np (multiples of 16)
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int h = 0; h < NH; ++h)
#pragma unroll
                    for (int e = 0; e < H; ++e) bacc[i] += ar[i][h][e];
        }
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int e = 0; e < H; ++e)
#pragma unroll
                for (int i = 0; i < RB; ++i)
#pragma unroll
                    for (int j = 0; j < CB; ++j) acc[i][j] = Mfma<T>::mma(ar[i][h][e], br[j][h][e], acc[i][j]);
    };
    long c = cb + (long)wave * SC;
    if (c < ce) load(a0, b0, c);
    while (c < ce) {
        const long c1 = c + 4 * SC;
        if (c1 < ce) load(a1, b1, c1);
        mul(a0, b0, c);
        if (c1 >= ce) break;
        const long c2 = c1 + 4 * SC;
        if (c2 < ce) load(a0, b0, c2);
        mul(a1, b1, c1);
        c = c2;
    }
    // fixed-order sum of the four waves' tiles (bias: also of the four lanes
    // q that share a row)
    for (int wv = 0; wv < 4; ++wv) {
        if (wave == wv) {
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int idx = (16 * i + Mfma<T>::row(lane, r)) * LD + 16 * j + p;
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
            if (k < Kg) P[(long)m * Kd + koff + k] = red[idx];
        } else if (do_bias) {
            P[(long)m * Kd] = red[idx];
        }
    }
}

// ---------------------------------------------------------------------------
// The Gauss-Newton Gram matrix (G = R R^T over the point axis).
// Block tile 16*RB*4 rows (four waves stacked along m) x 144 columns of the
// weight block (nine 16-wide MFMA blocks), so a weight block of up to
// 128 x 144 is one tile and each operand streams from HBM once.
static constexpr int GN = 144, GC = 16;

template <typename T, int RB>
__global__ __launch_bounds__(256) void k_gemm_grad(const T *__restrict__ A, const T *__restrict__ B, int M,
                                                   int Kd, long ncol, long Np, long chunk, int bias_row,
                                                   int lower, double *__restrict__ partial)
{
    constexpr int BMg = 64 * RB;
    __shared__ T As[GC][BMg + 1];
    __shared__ T Bs[GC][GN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BMg, k0 = blockIdx.z * GN;
    if (lower && k0 > m0 + BMg - 1) return;   // tile wholly above the diagonal
    const long cb = (long)blockIdx.x * chunk;
    const long ce = min(ncol, cb + chunk);
    typename Mfma<T>::V acc[RB][9];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int jj = 0; jj < 9; ++jj) acc[i][jj] = typename Mfma<T>::V{0, 0, 0, 0};

    for (long c0 = cb; c0 < ce; c0 += GC) {
        // A rows m0.., 16 columns: threads take 8 consecutive columns of a row
        for (int u = tid; u < BMg * 2; u += 256) {
            const int r = u >> 1, cq = (u & 1) * 8, gm = m0 + r;
            T v[8];
            if (gm < M) {
                const T *src = A + (long)gm * ncol + c0 + cq;
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = src[e];
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = T(0);
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) As[cq + e][r] = v[e];
        }
        for (int u = tid; u < GN * 2; u += 256) {
            const int r = u >> 1, cq = (u & 1) * 8, gk = k0 + r;
            T v[8];
            if (gk < Kd && !(bias_row && gk == 0)) {
                const T *src = B + (long)(bias_row ? gk - 1 : gk) * ncol + c0 + cq;
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = src[e];
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = (gk < Kd && c0 + cq + e < Np) ? T(1) : T(0);
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) Bs[cq + e][r] = v[e];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GC; kk += 4) {
            T a[RB], b[9];
#pragma unroll
            for (int i = 0; i < RB; ++i) a[i] = As[kk + (lane >> 4)][wave * 16 * RB + i * 16 + (lane & 15)];
#pragma unroll
            for (int jj = 0; jj < 9; ++jj) b[jj] = Bs[kk + (lane >> 4)][jj * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int jj = 0; jj < 9; ++jj) acc[i][jj] = Mfma<T>::mma(a[i], b[jj], acc[i][jj]);
        }
        __syncthreads();
    }
    double *P = partial + (long)blockIdx.x * M * Kd;
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int jj = 0; jj < 9; ++jj)
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int gm = m0 + wave * 16 * RB + i * 16 + Mfma<T>::row(lane, rr);
                const int gk = k0 + jj * 16 + (lane & 15);
                if (gm < M && gk < Kd) P[(long)gm * Kd + gk] = (double)acc[i][jj][rr];
            }
}

// fixed-order two-pass reduction: part2[s][i] = sum of chunk range s, then
// grad[i] += sum_s part2[s][i]
__global__ void k_reduce_pass1(const double *__restrict__ partial, int nchunk, int nsplit, long len,
                               double *__restrict__ part2)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int sidx = blockIdx.y;
    if (i >= len) return;
    const int per = (nchunk + nsplit - 1) / nsplit;
    const int c0 = sidx * per, c1 = min(nchunk, c0 + per);
    double s = 0.0;
    for (int c = c0; c < c1; ++c) s += partial[(long)c * len + i];
    part2[(long)sidx * len + i] = s;
}

__global__ void k_reduce_pass2(const double *__restrict__ part2, int nsplit, long len, double *__restrict__ grad)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    double s = 0.0;
    for (int sidx = 0; sidx < nsplit; ++sidx) s += part2[(long)sidx * len + i];
    grad[i] += s;
}

// the same two-level sum in one launch: thread (slice, weight) sums its
// chunk range in order (k_reduce_pass1's loop), the slice sums meet in LDS
// and are added in slice order (k_reduce_pass2's loop) -- the same
// operations in the same order, so the same bits.  256 / nsplit weights a
// workgroup.
__global__ __launch_bounds__(256) void k_reduce_fused(const double *__restrict__ partial, int nchunk, int nsplit,
                                                      long len, double *__restrict__ grad)
{
    __shared__ double red[256];
    const int wpb = 256 / nsplit, t = threadIdx.x;
    const int wi = t % wpb, sidx = t / wpb;
    const long i = (long)blockIdx.x * wpb + wi;
    if (i < len) {
        const int per = (nchunk + nsplit - 1) / nsplit;
        const int c0 = sidx * per, c1 = min(nchunk, c0 + per);
        double s = 0.0;
        // loads unrolled so several are in flight; the sum stays in chunk order
#pragma unroll 8
        for (int c = c0; c < c1; ++c) s += partial[(long)c * len + i];
        red[t] = s;
    }
    __syncthreads();
    if (sidx == 0 && i < len) {
        double s = 0.0;
        for (int k = 0; k < nsplit; ++k) s += red[k * wpb + wi];
        grad[i] += s;
    }
}

// several of k_reduce_fused's reductions in one launch (blockIdx.y picks
// the segment), each with its own split count: the same bits as one launch
// a segment, but for segments of 1024 chunks or more, which take 64 slices
// (short rows of many point blocks: the output layer's partials and the
// sums of R^2); set: the segment's output receives the sum instead of
// adding it
__host__ __device__ static inline int seg_splits(int nchunk) { return nchunk >= 1024 ? 64 : grad_reduce_splits(nchunk); }

__global__ __launch_bounds__(256) void k_reduce_segs(ReduceSegs r)
{
    __shared__ double red[256];
    const int g = blockIdx.y;
    const long len = r.len[g];
    const int nchunk = r.nchunk[g], nsplit = seg_splits(nchunk);
    const int wpb = 256 / nsplit, t = threadIdx.x;
    if ((long)blockIdx.x * wpb >= len) return;   // the whole workgroup: this segment is shorter
    const int wi = t % wpb, sidx = t / wpb;
    const long i = (long)blockIdx.x * wpb + wi;
    const double *partial = r.partial[g];
    if (i < len) {
        const int per = (nchunk + nsplit - 1) / nsplit;
        const int c0 = sidx * per, c1 = min(nchunk, c0 + per);
        double s = 0.0;
#pragma unroll 8
        for (int c = c0; c < c1; ++c) s += partial[(long)c * len + i];
        red[t] = s;
    }
    __syncthreads();
    if (sidx == 0 && i < len) {
        double s = 0.0;
        for (int k = 0; k < nsplit; ++k) s += red[k * wpb + wi];
        double *out = r.out[g];
        out[i] = r.set[g] ? s : out[i] + s;
    }
}

void launch_reduce_segs(const ReduceSegs &r, hipStream_t s)
{
    long nb = 0;
    for (int g = 0; g < r.n; ++g) {
        const long wpb = 256 / seg_splits(r.nchunk[g]);
        nb = std::max(nb, (r.len[g] + wpb - 1) / wpb);
    }
    if (r.n > 0 && nb > 0)
        hipLaunchKernelGGL(k_reduce_segs, dim3((unsigned)nb, (unsigned)r.n), dim3(256), 0, s, r);
}

#ifndef GRAD_WGS
#define GRAD_WGS 256
#endif

// K5 tile columns: 64, or 128 for FP32 blocks wider than 64
static int grad_cb(int esz, int Kg) { return (esz == 4 && Kg > 64) ? 8 : 4; }

void grad_chunks(long ncol, int M, int Kd, int esz, int *nchunk, long *chunk)
{
    // one resident round of workgroups (one per CU: the gradient kernels
    // hold a whole output tile per wave) over tiles x chunks, so every
    // workgroup streams one long column range; chunks are whole multiples of
    // 128 columns (every wave step stays inside its chunk), and a short
    // column range (a few points) is spread down to one 128-column step per
    // workgroup rather than walked by a handful of them
    const int cb = grad_cb(esz, Kd - 1);
    const int tiles = ((M + 63) / 64) * ((Kd - 1 + 16 * cb - 1) / (16 * cb));
    long nc = std::max(1, GRAD_WGS / tiles);
    nc = std::min(nc, std::max(1L, ncol / 128));
    const long ch = (((ncol + nc - 1) / nc) + 127) / 128 * 128;
    *nchunk = (int)((ncol + ch - 1) / ch);
    *chunk = ch;
}

template <typename T>
void launch_gemm_grad(int prec, const T *A, const T *B, int M, int Kd, long ncol, long Np, double *partial,
                      int nchunk, long chunk_cols, hipStream_t s)
{
    const int cb = grad_cb((int)sizeof(T), Kd - 1);
    const int ntm = (M + 63) / 64, ntk = (Kd - 1 + 16 * cb - 1) / (16 * cb);
    long g = (long)ntm * ntk * nchunk;
    g = (g + 7) / 8 * 8;
    bool done = false;
    if constexpr (sizeof(T) == 4) {
        if (prec == PREC_BF16X6) {
            launch_grad_x6(A, B, M, Kd, ncol, Np, partial, nchunk, chunk_cols, s);
            done = true;
        } else if (cb == 8) {
            hipLaunchKernelGGL((k_grad_nt<T, 4, 8, 2>), dim3((unsigned)g), dim3(256), 0, s, A, B, M, Kd, ncol, Np,
                               chunk_cols, 1, ntm, ntk, nchunk, partial);
            done = true;
        }
    }
    if (!done)
        hipLaunchKernelGGL((k_grad_nt<T, 4, 4, sizeof(T) == 4 ? 2 : 1>), dim3((unsigned)g), dim3(256), 0, s, A, B, M,
                           Kd, ncol, Np, chunk_cols, 1, ntm, ntk, nchunk, partial);
}

void launch_grad_reduce(double *partial, int nchunk, long len, double *grad, hipStream_t s)
{
    const int ns = grad_reduce_splits(nchunk);
    const long wpb = 256 / ns;
    hipLaunchKernelGGL(k_reduce_fused, dim3((unsigned)((len + wpb - 1) / wpb)), dim3(256), 0, s, partial, nchunk, ns,
                       len, grad);
}

// the Gram matrix of Gauss-Newton rows: G[i][j] = sum_n R[i][n] R[j][n]
// (rows stored transposed, R[w][n]), 128 x 144 tiles; only tiles holding
// lower-triangle entries are formed (the factorization reads no other)
void launch_gram(const float *R, int nw, long ld, double *partial, int nchunk, long chunk_cols, double *G,
                 hipStream_t s)
{
    dim3 grid((unsigned)nchunk, (unsigned)((nw + 127) / 128), (unsigned)((nw + GN - 1) / GN));
    if (nchunk == 1) {  // one chunk: the tiles write G directly
        hipLaunchKernelGGL((k_gemm_grad<float, 2>), grid, dim3(256), 0, s, R, R, nw, nw, ld, 0L, chunk_cols, 0, 1, G);
        return;
    }
    hipLaunchKernelGGL((k_gemm_grad<float, 2>), grid, dim3(256), 0, s, R, R, nw, nw, ld, 0L, chunk_cols, 0,
                       1, partial);
    const long len = (long)nw * nw;
    const int ns = grad_reduce_splits(nchunk);
    double *part2 = partial + (long)nchunk * len;
    const unsigned gx = (unsigned)((len + 255) / 256);
    hipLaunchKernelGGL(k_reduce_pass1, dim3(gx, (unsigned)ns), dim3(256), 0, s, partial, nchunk, ns, len, part2);
    hipLaunchKernelGGL(k_reduce_pass2, dim3(gx), dim3(256), 0, s, part2, ns, len, G);
}

void launch_gram64(const double *R, int n, long ld, double *G, hipStream_t s)
{
    dim3 grid(1u, (unsigned)((n + 127) / 128), (unsigned)((n + GN - 1) / GN));
    hipLaunchKernelGGL((k_gemm_grad<double, 2>), grid, dim3(256), 0, s, R, R, n, n, ld, 0L, ld, 0, 1, G);
}

// per-point gradient rows (no reduction over points): the Jacobian rows of
// the NGD metric, stored transposed, rowsT[w][n]
template <typename T>
__global__ void k_point_rows(const T *__restrict__ A, const T *__restrict__ B, int M, int Kd, int na, long Np,
                             float *__restrict__ rowsT, long ldr)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y, j = blockIdx.z;
    if (n >= Np) return;
    const long ldc = (long)na * Np;
    double s;
    if (k == 0) {
        s = (double)A[(long)j * ldc + n];
    } else {
        s = 0.0;
        for (int a = 0; a < na; ++a)
            s += (double)A[(long)j * ldc + a * Np + n] * (double)B[(long)(k - 1) * ldc + a * Np + n];
    }
    rowsT[((long)j * Kd + k) * ldr + n] = (float)s;
}

template <typename T>
void launch_point_rows(const T *A, const T *B, int M, int Kd, int na, long Np, float *rowsT, long ldr,
                       hipStream_t s)
{
    dim3 grid((unsigned)((Np + 255) / 256), (unsigned)Kd, (unsigned)M);
    hipLaunchKernelGGL((k_point_rows<T>), grid, dim3(256), 0, s, A, B, M, Kd, na, Np, rowsT, ldr);
}

// weight gradient of one point (column 0 of every slot), added to grad[M][Kd]:
// the bias column takes slot 0 of Sbar, weight column k the sum over slots
// in slot order, in FP64 -- one thread per weight, no partials, no reduction
template <typename T>
__global__ __launch_bounds__(256) void k_grad_point1(const T *__restrict__ A, const T *__restrict__ B, int M, int Kd,
                                                     int na, long Np, double *__restrict__ grad)
{
    const int u = blockIdx.x * 256 + threadIdx.x;
    if (u >= M * Kd) return;
    const int j = u / Kd, k = u - j * Kd;
    const long ldc = (long)na * Np;
    double s;
    if (k == 0) {
        s = (double)A[(long)j * ldc];
    } else {
        // eight slots' loads in flight at a time, then their terms in order
        const T *pa = A + (long)j * ldc, *pb = B + (long)(k - 1) * ldc;
        s = 0.0;
        int a = 0;
        for (; a + 8 <= na; a += 8) {
            double x[8], y[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                x[e] = (double)pa[(long)(a + e) * Np];
                y[e] = (double)pb[(long)(a + e) * Np];
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) s += x[e] * y[e];
        }
        for (; a < na; ++a) s += (double)pa[(long)a * Np] * (double)pb[(long)a * Np];
    }
    grad[u] += s;
}

template <typename T>
void launch_grad_point1(const T *A, const T *B, int M, int Kd, int na, long Np, double *grad, hipStream_t s)
{
    hipLaunchKernelGGL((k_grad_point1<T>), dim3((unsigned)((M * Kd + 255) / 256)), dim3(256), 0, s, A, B, M, Kd, na,
                       Np, grad);
}

#define INST(T)                                                                                             \
    template void launch_grad_point1<T>(const T *, const T *, int, int, int, long, double *, hipStream_t);   \
    template void launch_gemm_wx<T>(int, bool, const T *, int, const T *, T *, int, int, long, long,         \
                                    hipStream_t);                                                           \
    template void launch_gemm_grad<T>(int, const T *, const T *, int, int, long, long, double *, int, long,  \
                                      hipStream_t);                                                         \
    template void launch_point_rows<T>(const T *, const T *, int, int, int, long, float *, long, hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace dnnf90
