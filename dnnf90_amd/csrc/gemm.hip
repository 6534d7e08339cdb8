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
//
// MFMA layouts (16x16 family, wave of 64): lane l feeds row l mod 16 of A
// and column l mod 16 of B; the k index a lane feeds is l/16 (x4 forms) or
// 8*(l/16)+e (bf16 x32 form).  Result: column l mod 16, rows 4*(l/16)+r for
// the FP32-accumulating forms and l/16+4r for v_mfma_f64_16x16x4f64
// (measured on gfx950).
#include <hip/hip_bf16.h>

#include "kernels.hpp"

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

// value of row gk of [1; B] (bias row first) at column gc
template <typename T, bool kBiasRow>
__device__ __forceinline__ T bext(const T *__restrict__ B, long ncol, long Np, int gk, int Kd, long gc)
{
    if (gk >= Kd) return T(0);
    if (kBiasRow && gk == 0) return gc < Np ? T(1) : T(0);
    return B[(long)(kBiasRow ? gk - 1 : gk) * ncol + gc];
}

// ---------------------------------------------------------------------------
// K1 / K6, FP64 and FP32.  Block tile BM x 64, k-tile 16, four waves in 2x2;
// a wave owns (BM/2) x 32 = (BM/32) x 2 MFMA blocks.  The next k-tile is
// loaded into registers while the current one is multiplied.
static constexpr int BN = 64, BK = 16;

template <typename T, bool kTrans, int BM>
__global__ __launch_bounds__(256) void k_gemm_wx(const T *__restrict__ W, int ldw, const T *__restrict__ B,
                                                 T *__restrict__ C, int M, int Kd, long ncol, long Np)
{
    constexpr int RB = BM / 32;            // row blocks per wave
    constexpr int AE = BM * BK / 256;      // A elements per thread per k-tile
    __shared__ T As[BK][BM];
    __shared__ T Bs[BK][BN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * BM;
    const long n0 = (long)blockIdx.x * BN;
    typename Mfma<T>::V acc[RB][2];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) acc[i][jj] = typename Mfma<T>::V{0, 0, 0, 0};

    // staging assignment: A: kTrans ? (k = tid / (BM/AE), m run of AE) : (m = tid / (BK/AE)... )
    T ra[AE], rb[4];
    auto load = [&](int k0) {
        if (!kTrans) {
            // A[m][k] = W[m][k], k contiguous: thread reads AE consecutive k of one row
            constexpr int TPR = BK / AE;   // threads per row
            const int m = tid / TPR, kq = (tid % TPR) * AE, gm = m0 + m;
#pragma unroll
            for (int e = 0; e < AE; ++e) {
                const int gk = k0 + kq + e;
                ra[e] = (gm < M && gk < Kd) ? W[(long)gm * ldw + gk] : T(0);
            }
        } else {
            // A[m][k] = W[k][m+1], m contiguous
            constexpr int TPR = BM / AE;
            const int k = tid / TPR, mq = (tid % TPR) * AE, gk = k0 + k;
#pragma unroll
            for (int e = 0; e < AE; ++e) {
                const int gm = m0 + mq + e;
                ra[e] = (gm < M && gk < Kd) ? W[(long)gk * ldw + 1 + gm] : T(0);
            }
        }
        const int k = tid >> 4, cq = (tid & 15) * 4, gk = k0 + k;
        const long gc = n0 + cq;
        if (gk < Kd && !(!kTrans && gk == 0)) {
            const T *src = B + (long)(kTrans ? gk : gk - 1) * ncol + gc;
#pragma unroll
            for (int e = 0; e < 4; ++e) rb[e] = src[e];
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) rb[e] = bext<T, !kTrans>(B, ncol, Np, gk, Kd, gc + e);
        }
    };
    auto store = [&]() {
        if (!kTrans) {
            constexpr int TPR = BK / AE;
            const int m = tid / TPR, kq = (tid % TPR) * AE;
#pragma unroll
            for (int e = 0; e < AE; ++e) As[kq + e][m] = ra[e];
        } else {
            constexpr int TPR = BM / AE;
            const int k = tid / TPR, mq = (tid % TPR) * AE;
#pragma unroll
            for (int e = 0; e < AE; ++e) As[k][mq + e] = ra[e];
        }
        const int k = tid >> 4, cq = (tid & 15) * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e) Bs[k][cq + e] = rb[e];
    };

    load(0);
    for (int k0 = 0; k0 < Kd; k0 += BK) {
        store();
        __syncthreads();
        if (k0 + BK < Kd) load(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T a[RB], b[2];
#pragma unroll
            for (int i = 0; i < RB; ++i) a[i] = As[kk + (lane >> 4)][wm * (BM / 2) + i * 16 + (lane & 15)];
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) b[jj] = Bs[kk + (lane >> 4)][wn * 32 + jj * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) acc[i][jj] = Mfma<T>::mma(a[i], b[jj], acc[i][jj]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm * (BM / 2) + i * 16 + Mfma<T>::row(lane, r);
                const long gc = n0 + wn * 32 + jj * 16 + (lane & 15);
                if (gm < M) C[(long)gm * ncol + gc] = acc[i][jj][r];
            }
}

// ---------------------------------------------------------------------------
// K1 / K6 with BF16 hi+lo split operands: x = hi + lo, hi = bf16(x),
// lo = bf16(x - hi); a product keeps lo*hi + hi*lo + hi*hi (lo*lo is below
// 2^-32 relative) with FP32 accumulation on v_mfma_f32_16x16x32_bf16.  The
// operands are split while staging into LDS as [row][k] so a lane's eight
// consecutive k are one 16-byte read.
static constexpr int XBK = 32, XPAD = 8;

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

template <typename T>
void launch_gemm_wx(int prec, bool trans_a, const T *W, int ldw, const T *B, T *C, int M, int Kd, long ncol,
                    long Np, hipStream_t s)
{
    const bool big = M > 64;
    const int bm = big ? 128 : 64;
    dim3 grid((unsigned)(ncol / BN), (unsigned)((M + bm - 1) / bm));
#define LAUNCH(K, ...) hipLaunchKernelGGL((K), grid, dim3(256), 0, s, __VA_ARGS__)
    if constexpr (sizeof(T) == 4) {
        if (prec == PREC_BF16X3) {
            if (trans_a) {
                if (big) LAUNCH((k_gemm_wx_bf3<true, 128>), W, ldw, B, C, M, Kd, ncol, Np);
                else LAUNCH((k_gemm_wx_bf3<true, 64>), W, ldw, B, C, M, Kd, ncol, Np);
            } else {
                if (big) LAUNCH((k_gemm_wx_bf3<false, 128>), W, ldw, B, C, M, Kd, ncol, Np);
                else LAUNCH((k_gemm_wx_bf3<false, 64>), W, ldw, B, C, M, Kd, ncol, Np);
            }
            return;
        }
    }
    if (trans_a) {
        if (big) LAUNCH((k_gemm_wx<T, true, 128>), W, ldw, B, C, M, Kd, ncol, Np);
        else LAUNCH((k_gemm_wx<T, true, 64>), W, ldw, B, C, M, Kd, ncol, Np);
    } else {
        if (big) LAUNCH((k_gemm_wx<T, false, 128>), W, ldw, B, C, M, Kd, ncol, Np);
        else LAUNCH((k_gemm_wx<T, false, 64>), W, ldw, B, C, M, Kd, ncol, Np);
    }
#undef LAUNCH
}

// ---------------------------------------------------------------------------
// K5: partial[chunk][m][k] = sum over the chunk's columns of A[m][c] Bext[k][c].
// Block tile 16*RB*4 rows (four waves stacked along m) x 144 columns of the
// weight block (nine 16-wide MFMA blocks), so a weight block of up to
// 128 x 144 is one tile and each operand streams from HBM once.
static constexpr int GN = 144, GC = 16;

template <typename T, int RB>
__global__ __launch_bounds__(256) void k_gemm_grad(const T *__restrict__ A, const T *__restrict__ B, int M,
                                                   int Kd, long ncol, long Np, long chunk, int bias_row,
                                                   double *__restrict__ partial)
{
    constexpr int BMg = 64 * RB;
    __shared__ T As[GC][BMg + 1];
    __shared__ T Bs[GC][GN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BMg, k0 = blockIdx.z * GN;
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

int grad_reduce_splits(int nchunk) { return nchunk >= 64 ? 16 : (nchunk >= 16 ? 4 : 1); }

template <typename T>
void launch_gemm_grad(int prec, const T *A, const T *B, int M, int Kd, long ncol, long Np, double *partial,
                      int nchunk, long chunk_cols, double *grad, hipStream_t s)
{
    (void)prec;
    const int rb = M > 64 ? 2 : 1;
    dim3 grid((unsigned)nchunk, (unsigned)((M + 64 * rb - 1) / (64 * rb)), (unsigned)((Kd + GN - 1) / GN));
    if (rb == 2)
        hipLaunchKernelGGL((k_gemm_grad<T, 2>), grid, dim3(256), 0, s, A, B, M, Kd, ncol, Np, chunk_cols, 1,
                           partial);
    else
        hipLaunchKernelGGL((k_gemm_grad<T, 1>), grid, dim3(256), 0, s, A, B, M, Kd, ncol, Np, chunk_cols, 1,
                           partial);
    const long len = (long)M * Kd;
    const int ns = grad_reduce_splits(nchunk);
    double *part2 = partial + (long)nchunk * len;
    const unsigned gx = (unsigned)((len + 255) / 256);
    hipLaunchKernelGGL(k_reduce_pass1, dim3(gx, (unsigned)ns), dim3(256), 0, s, partial, nchunk, ns, len, part2);
    hipLaunchKernelGGL(k_reduce_pass2, dim3(gx), dim3(256), 0, s, part2, ns, len, grad);
}

// the Gram matrix of Gauss-Newton rows: G[i][j] = sum_n R[i][n] R[j][n]
// (rows stored transposed, R[w][n]); same tile as K5 without a bias row
void launch_gram(const float *R, int nw, long ld, double *partial, int nchunk, long chunk_cols, double *G,
                 hipStream_t s)
{
    dim3 grid((unsigned)nchunk, (unsigned)((nw + 127) / 128), (unsigned)((nw + GN - 1) / GN));
    if (nchunk == 1) {  // one chunk: the tiles write G directly
        hipLaunchKernelGGL((k_gemm_grad<float, 2>), grid, dim3(256), 0, s, R, R, nw, nw, ld, 0L, chunk_cols, 0, G);
        return;
    }
    hipLaunchKernelGGL((k_gemm_grad<float, 2>), grid, dim3(256), 0, s, R, R, nw, nw, ld, 0L, chunk_cols, 0,
                       partial);
    const long len = (long)nw * nw;
    const int ns = grad_reduce_splits(nchunk);
    double *part2 = partial + (long)nchunk * len;
    const unsigned gx = (unsigned)((len + 255) / 256);
    hipLaunchKernelGGL(k_reduce_pass1, dim3(gx, (unsigned)ns), dim3(256), 0, s, partial, nchunk, ns, len, part2);
    hipLaunchKernelGGL(k_reduce_pass2, dim3(gx), dim3(256), 0, s, part2, ns, len, G);
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

#define INST(T)                                                                                             \
    template void launch_gemm_wx<T>(int, bool, const T *, int, const T *, T *, int, int, long, long,         \
                                    hipStream_t);                                                           \
    template void launch_gemm_grad<T>(int, const T *, const T *, int, int, long, long, double *, int, long,  \
                                      double *, hipStream_t);                                               \
    template void launch_point_rows<T>(const T *, const T *, int, int, int, long, float *, long, hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace dnnf90
