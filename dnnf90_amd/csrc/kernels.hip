// kernels.hip -- gfx950 kernels of the derivative-carrying forward pass and
// its exact adjoint.
//
//   K1/K6  linear maps W*T and W^T*Sbar: one MFMA GEMM over the na*Np
//          column axis (FP64: v_mfma_f64_16x16x4; FP32: v_mfma_f32_16x16x4;
//          BF16x3: v_mfma_f32_16x16x32_bf16 on hi/lo split operands)
//   K2/K3  Faa di Bruno step with sigma^(q), one lane per (point, neuron)
//          (lib/train_module.f90:283-325 and multi_index_bell_module.f90:723-867)
//   K4     adjoint pair sums (lib/train_module.f90:381-403)
//   K5     weight gradient: split-K MFMA over the column axis, FP64 partials
//          reduced in a fixed order (lib/train_module.f90:408-422)
//   K7     output-layer seeds (app/pinn_module.f90:32-65)
#include "kernels.hpp"

#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>

namespace dnnf90 {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

static constexpr int KQMAX = 17;  // sigma^(0..16): k <= 15 needs q <= k+1

// ---------------------------------------------------------------------------
// sigma^(q)(a), q = 0..ap.nq, always evaluated in FP64
// (multi_index_bell_module.f90:723-867; GELU: x Phi^(q) + q Phi^(q-1))
__device__ __forceinline__ void act_derivs(const ActParams &ap, double a, double (&d)[KQMAX])
{
    const int nq = ap.nq;
#pragma unroll
    for (int q = 0; q < KQMAX; ++q) d[q] = 0.0;
    if (ap.iact == 1) {  // sin(a + q pi/2)
        const double pio2 = 1.57079632679489661923;
#pragma unroll
        for (int q = 0; q < KQMAX; ++q)
            if (q <= nq) d[q] = sin(a + (double)q * pio2);
    } else if (ap.iact == 2) {  // erf: (2/sqrt(pi)) (-1)^(q-1) H_(q-1)(a) e^(-a^2)
        const double c2spi = 1.12837916709551257390;
        double pm = exp(-a * a), pc = 2.0 * a * pm;  // psi_0, psi_1
        d[0] = erf(a);
        if (nq >= 1) d[1] = c2spi * pm;
        if (nq >= 2) d[2] = -c2spi * pc;
#pragma unroll
        for (int n = 1; n < KQMAX - 2; ++n) {
            if (n + 2 <= nq) {
                double pn = 2.0 * a * pc - 2.0 * (double)n * pm;  // psi_(n+1)
                d[n + 2] = ((n + 1) & 1 ? -c2spi : c2spi) * pn;
                pm = pc;
                pc = pn;
            }
        }
    } else if (ap.iact == 5) {  // erf-GELU
        const double rs2pi = 0.39894228040143267794, rs2 = 0.70710678118654752440;
        double ps[KQMAX];
        ps[0] = rs2pi * exp(-0.5 * a * a);
        ps[1] = a * ps[0];
#pragma unroll
        for (int n = 1; n < KQMAX - 1; ++n) ps[n + 1] = a * ps[n] - (double)n * ps[n - 1];
        double phm = 0.5 * erfc(-a * rs2);  // Phi^(q-1)
        d[0] = a * phm;
#pragma unroll
        for (int q = 1; q < KQMAX; ++q) {
            if (q <= nq) {
                double ph = ((q - 1) & 1 ? -1.0 : 1.0) * ps[q - 1];  // Phi^(q)
                d[q] = a * ph + (double)q * phm;
                phm = ph;
            }
        }
    } else if (ap.iact == 3 || ap.iact == 4) {
        // J0^(q) = (-1)^q 2^-q sum_k (-1)^k C(q,k) J_(q-2k);  J1 shifted by one
        const int sh = ap.iact == 4 ? 1 : 0;
#pragma unroll
        for (int q = 0; q < KQMAX; ++q) {
            if (q <= nq) {
                double term = 0.0, c = 1.0;
                for (int kk = 0; kk <= q; ++kk) {
                    int m = sh + q - 2 * kk;
                    double jm = jn(m < 0 ? -m : m, a);
                    if (m < 0 && ((-m) & 1)) jm = -jm;
                    term += ((kk & 1) ? -1.0 : 1.0) * c * jm;
                    c = c * (double)(q - kk) / (double)(kk + 1);
                }
                d[q] = ((q & 1) ? -term : term) / ldexp(1.0, q);
            }
        }
    } else {  // tanh: Horner on P_q(t)
        const double t = tanh(a);
#pragma unroll
        for (int q = 0; q < KQMAX; ++q) {
            if (q <= nq) {
                const double *c = ap.tpc + q * ap.ldt;
                double v = c[q + 1];
                for (int i = q; i >= 0; --i) v = v * t + c[i];
                d[q] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// K2/K3: forward Faa di Bruno step.  Lane = point, wave = neuron.  Per lane
// the operand store E holds S_a (entries < na) and the B(a,q) that later
// terms read; it lives in LDS ([entry][lane] so that every access is
// conflict free) or, for sets too large for LDS, in a global scratch plane.
template <typename T> __device__ __forceinline__ T coef(const DevFwdSched &s, int it);
template <> __device__ __forceinline__ double coef<double>(const DevFwdSched &s, int it) { return s.term_c64[it]; }
template <> __device__ __forceinline__ float coef<float>(const DevFwdSched &s, int it) { return s.term_c32[it]; }

template <typename T, bool kLDS>
__global__ __launch_bounds__(256) void k_bell_fwd(const T *__restrict__ S, T *__restrict__ Tout,
                                                  T *__restrict__ H, int nrow, long Np,
                                                  DevFwdSched sch, ActParams ap,
                                                  T *__restrict__ scratch)
{
    extern __shared__ unsigned char smem_raw[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const int j = blockIdx.y * nwave + wave;
    if (j >= nrow) return;
    const long n = (long)blockIdx.x * 64 + lane;
    const int na = sch.na;
    const long ldr = (long)na * Np;
    const T *Sj = S + (long)j * ldr + n;
    T *E;
    long es;
    if (kLDS) {
        E = reinterpret_cast<T *>(smem_raw) + threadIdx.x;
        es = blockDim.x;
    } else {
        E = scratch + (long)j * Np + n;
        es = (long)nrow * Np;
    }
    for (int a = 0; a < na; ++a) E[a * es] = Sj[a * Np];

    double d[KQMAX];
    act_derivs(ap, (double)E[0], d);
    T dq[KQMAX];
#pragma unroll
    for (int q = 0; q < KQMAX; ++q) dq[q] = (T)d[q];

    T *Tj = Tout + (long)j * ldr + n;
    T *Hj = H + (long)j * ldr + n;
    Tj[0] = dq[0];
    Hj[0] = dq[1];
    for (int a = 1; a < na; ++a) {
        const int g0 = sch.slot_group[a];
        const int p = sch.slot_group[a + 1] - g0 + 1;  // |a|
        const T s = E[a * es];
        T ttv = dq[1] * s, htv = dq[2] * s;
#pragma unroll
        for (int q = 2; q < KQMAX - 1; ++q) {
            if (q <= p) {
                const int g = g0 + q - 2;
                T bsv = T(0);
                const int t1 = sch.grp_term[g + 1];
                for (int it = sch.grp_term[g]; it < t1; ++it) {
                    const int2 sb = sch.term_sb[it];
                    bsv += coef<T>(sch, it) * E[sb.x * es] * E[sb.y * es];
                }
                const int st = sch.grp_store[g];
                if (st >= 0) E[st * es] = bsv;
                ttv += dq[q] * bsv;
                htv += dq[q + 1] * bsv;
            }
        }
        Tj[a * Np] = ttv;
        Hj[a * Np] = htv;
    }
}

static int pick_threads(long bytes_per_thread)
{
    // 64 KiB per workgroup keeps two workgroups resident per CU
    for (int nt = 256; nt >= 64; nt >>= 1)
        if (bytes_per_thread * nt <= 65536) return nt;
    return 0;
}

int bell_fwd_threads(int nentry, int elem_bytes) { return pick_threads((long)nentry * elem_bytes); }
int bell_bwd_threads(int na, int elem_bytes) { return pick_threads(2L * na * elem_bytes); }

template <typename T>
void launch_bell_fwd(const T *S, T *Tout, T *H, int nrow, long Np, const DevFwdSched &sch,
                     const ActParams &act, T *scratch, hipStream_t s)
{
    const int nt = bell_fwd_threads(sch.nentry, sizeof(T));
    if (nt > 0) {
        dim3 grid((unsigned)(Np / 64), (unsigned)((nrow + nt / 64 - 1) / (nt / 64)));
        size_t lds = (size_t)sch.nentry * nt * sizeof(T);
        hipLaunchKernelGGL((k_bell_fwd<T, true>), grid, dim3(nt), lds, s, S, Tout, H, nrow, Np, sch,
                           act, (T *)nullptr);
    } else {
        dim3 grid((unsigned)(Np / 64), (unsigned)((nrow + 3) / 4));
        hipLaunchKernelGGL((k_bell_fwd<T, false>), grid, dim3(256), 0, s, S, Tout, H, nrow, Np, sch,
                           act, scratch);
    }
}

// ---------------------------------------------------------------------------
// K4: adjoint pair sums.  Sbar_0 = sum_a Tbar_a h_a;
// Sbar_b = sum_{pairs(b)} C(a,b) Tbar_a h_(a-b)
template <typename T> __device__ __forceinline__ T pbinom(const DevPairs &p, int ip);
template <> __device__ __forceinline__ double pbinom<double>(const DevPairs &p, int ip) { return p.binom64[ip]; }
template <> __device__ __forceinline__ float pbinom<float>(const DevPairs &p, int ip) { return p.binom32[ip]; }

template <typename T, bool kLDS>
__global__ __launch_bounds__(256) void k_bell_bwd(const T *__restrict__ Tb, const T *__restrict__ H,
                                                  T *__restrict__ Sb, int nrow, long Np,
                                                  DevPairs pr)
{
    extern __shared__ unsigned char smem_raw[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const int j = blockIdx.y * nwave + wave;
    if (j >= nrow) return;
    const long n = (long)blockIdx.x * 64 + lane;
    const int na = pr.na;
    const long ldr = (long)na * Np;
    const T *tb = Tb + (long)j * ldr + n;
    const T *hh = H + (long)j * ldr + n;
    const T *et, *eh;
    long es;
    if (kLDS) {
        T *E = reinterpret_cast<T *>(smem_raw) + threadIdx.x;
        es = blockDim.x;
        for (int a = 0; a < na; ++a) {
            E[a * es] = tb[a * Np];
            E[(na + a) * es] = hh[a * Np];
        }
        et = E;
        eh = E + (long)na * es;
    } else {
        et = tb;
        eh = hh;
        es = Np;
    }
    T *sb = Sb + (long)j * ldr + n;
    T acc = et[0] * eh[0];
    for (int a = 1; a < na; ++a) acc += et[a * es] * eh[a * es];
    sb[0] = acc;
    for (int b = 1; b < na; ++b) {
        T v = T(0);
        const int p0 = pr.start[b], p1 = p0 + pr.num[b];
        for (int ip = p0; ip < p1; ++ip) {
            const int2 q = pr.ia_id[ip];
            v += pbinom<T>(pr, ip) * et[q.x * es] * eh[q.y * es];
        }
        sb[b * Np] = v;
    }
}

template <typename T>
void launch_bell_bwd(const T *Tb, const T *H, T *Sb, int nrow, long Np, const DevPairs &pr,
                     hipStream_t s)
{
    const int nt = bell_bwd_threads(pr.na, sizeof(T));
    if (nt > 0) {
        dim3 grid((unsigned)(Np / 64), (unsigned)((nrow + nt / 64 - 1) / (nt / 64)));
        hipLaunchKernelGGL((k_bell_bwd<T, true>), grid, dim3(nt), (size_t)2 * pr.na * nt * sizeof(T), s,
                           Tb, H, Sb, nrow, Np, pr);
    } else {
        dim3 grid((unsigned)(Np / 64), (unsigned)((nrow + 3) / 4));
        hipLaunchKernelGGL((k_bell_bwd<T, false>), grid, dim3(256), 0, s, Tb, H, Sb, nrow, Np, pr);
    }
}

// ---------------------------------------------------------------------------
// input layer: T(i, 0) = x_i, T(i, e_i) = 1, all else 0
// (lib/train_module.f90:253-257)
template <typename T>
__global__ void k_input(const double *__restrict__ x, int npt, int d0, int na,
                        const int *__restrict__ ind_e1, T *__restrict__ T0, long Np)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (c >= (long)na * Np) return;
    const int a = (int)(c / Np);
    const long n = c - (long)a * Np;
    T v = T(0);
    if (a == 0) v = n < npt ? (T)x[n * d0 + i] : T(0);
    else if (a == ind_e1[i]) v = T(1);
    T0[(long)i * na * Np + c] = v;
}

template <typename T>
void launch_input(const double *x, int npt, int d0, int na, const int *ind_e1, T *T0, long Np,
                  hipStream_t s)
{
    dim3 grid((unsigned)(((long)na * Np + 255) / 256), (unsigned)d0);
    hipLaunchKernelGGL((k_input<T>), grid, dim3(256), 0, s, x, npt, d0, na, ind_e1, T0, Np);
}

// ---------------------------------------------------------------------------
// MFMA GEMMs.  Block tile 64x64, four waves in a 2x2 arrangement, each wave
// a 32x32 tile of four 16x16 MFMA blocks.  In the 16x16 family lane l feeds
// row (l mod 16) of A and column (l mod 16) of B and holds four results of
// column (l mod 16): rows 4*(l/16)+r for the FP32-accumulating forms, rows
// l/16+4r for v_mfma_f64_16x16x4f64 (measured on gfx950).
static constexpr int GBM = 64, GBN = 64, GBK = 16, XBK = 32;

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

// B-extended row gk of the column tile [gc, gc+4): row 0 is the bias channel
template <typename T, bool kBiasRow>
__device__ __forceinline__ void load_brow4(const T *__restrict__ B, long ncol, long Np, int gk,
                                           int Kd, long gc, T (&v)[4])
{
    if (gk >= Kd) {
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = T(0);
    } else if (kBiasRow && gk == 0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = (gc + e < Np) ? T(1) : T(0);
    } else {
        const T *src = B + (long)(kBiasRow ? gk - 1 : gk) * ncol + gc;
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = src[e];
    }
}

template <typename T, bool kTrans>
__global__ __launch_bounds__(256) void k_gemm_wx(const T *__restrict__ W, int ldw,
                                                 const T *__restrict__ B, T *__restrict__ C,
                                                 int M, int Kd, long ncol, long Np)
{
    __shared__ T As[GBK][GBM];
    __shared__ T Bs[GBK][GBN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * GBM;
    const long n0 = (long)blockIdx.x * GBN;
    typename Mfma<T>::V acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) acc[i][jj] = typename Mfma<T>::V{0, 0, 0, 0};

    for (int k0 = 0; k0 < Kd; k0 += GBK) {
        if (!kTrans) {
            const int m = tid >> 2, kq = (tid & 3) * 4, gm = m0 + m;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int gk = k0 + kq + e;
                As[kq + e][m] = (gm < M && gk < Kd) ? W[(long)gm * ldw + gk] : T(0);
            }
        } else {
            const int k = tid >> 4, mq = (tid & 15) * 4, gk = k0 + k;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int gm = m0 + mq + e;
                As[k][mq + e] = (gm < M && gk < Kd) ? W[(long)gk * ldw + 1 + gm] : T(0);
            }
        }
        {
            const int k = tid >> 4, cq = (tid & 15) * 4;
            T v[4];
            load_brow4<T, !kTrans>(B, ncol, Np, k0 + k, Kd, n0 + cq, v);
#pragma unroll
            for (int e = 0; e < 4; ++e) Bs[k][cq + e] = v[e];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 4) {
            T a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[kk + (lane >> 4)][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) b[jj] = Bs[kk + (lane >> 4)][wn * 32 + jj * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) acc[i][jj] = Mfma<T>::mma(a[i], b[jj], acc[i][jj]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm * 32 + i * 16 + Mfma<T>::row(lane, r);
                const long gc = n0 + wn * 32 + jj * 16 + (lane & 15);
                if (gm < M) C[(long)gm * ncol + gc] = acc[i][jj][r];
            }
}

// BF16 hi+lo split: x = hi + lo with hi = bf16(x), lo = bf16(x - hi); the
// product keeps hi*hi + hi*lo + lo*hi (the lo*lo term is below FP32 rounding)
__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo)
{
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

template <bool kTrans>
__global__ __launch_bounds__(256) void k_gemm_wx_bf3(const float *__restrict__ W, int ldw,
                                                     const float *__restrict__ B,
                                                     float *__restrict__ C, int M, int Kd,
                                                     long ncol, long Np)
{
    __shared__ __attribute__((aligned(16))) __bf16 Ah[GBM][XBK + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Al[GBM][XBK + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Bh[GBN][XBK + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Bl[GBN][XBK + 8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * GBM;
    const long n0 = (long)blockIdx.x * GBN;
    flt4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) acc[i][jj] = flt4{0, 0, 0, 0};

    for (int k0 = 0; k0 < Kd; k0 += XBK) {
        if (!kTrans) {
            const int m = tid >> 2, kq = (tid & 3) * 8, gm = m0 + m;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int gk = k0 + kq + e;
                const float v = (gm < M && gk < Kd) ? W[(long)gm * ldw + gk] : 0.f;
                split_bf16(v, Ah[m][kq + e], Al[m][kq + e]);
            }
        } else {
            const int k = tid >> 3, mq = (tid & 7) * 8, gk = k0 + k;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int gm = m0 + mq + e;
                const float v = (gm < M && gk < Kd) ? W[(long)gk * ldw + 1 + gm] : 0.f;
                split_bf16(v, Ah[mq + e][k], Al[mq + e][k]);
            }
        }
        {
            const int k = tid >> 3, cq = (tid & 7) * 8;
            float v[8];
            load_brow4<float, !kTrans>(B, ncol, Np, k0 + k, Kd, n0 + cq, *reinterpret_cast<float(*)[4]>(v));
            load_brow4<float, !kTrans>(B, ncol, Np, k0 + k, Kd, n0 + cq + 4,
                                       *reinterpret_cast<float(*)[4]>(v + 4));
#pragma unroll
            for (int e = 0; e < 8; ++e) split_bf16(v[e], Bh[cq + e][k], Bl[cq + e][k]);
        }
        __syncthreads();
        bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int r = wm * 32 + i * 16 + (lane & 15);
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
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) {
                acc[i][jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[jj], acc[i][jj], 0, 0, 0);
                acc[i][jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[jj], acc[i][jj], 0, 0, 0);
                acc[i][jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[jj], acc[i][jj], 0, 0, 0);
            }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
                const long gc = n0 + wn * 32 + jj * 16 + (lane & 15);
                if (gm < M) C[(long)gm * ncol + gc] = acc[i][jj][r];
            }
}

template <typename T>
void launch_gemm_wx(int prec, bool trans_a, const T *W, int ldw, const T *B, T *C, int M, int Kd,
                    long ncol, long Np, hipStream_t s)
{
    dim3 grid((unsigned)(ncol / GBN), (unsigned)((M + GBM - 1) / GBM));
    if constexpr (sizeof(T) == 4) {
        if (prec == PREC_BF16X3) {
            if (trans_a)
                hipLaunchKernelGGL((k_gemm_wx_bf3<true>), grid, dim3(256), 0, s, W, ldw, B, C, M, Kd, ncol, Np);
            else
                hipLaunchKernelGGL((k_gemm_wx_bf3<false>), grid, dim3(256), 0, s, W, ldw, B, C, M, Kd, ncol, Np);
            return;
        }
    }
    if (trans_a)
        hipLaunchKernelGGL((k_gemm_wx<T, true>), grid, dim3(256), 0, s, W, ldw, B, C, M, Kd, ncol, Np);
    else
        hipLaunchKernelGGL((k_gemm_wx<T, false>), grid, dim3(256), 0, s, W, ldw, B, C, M, Kd, ncol, Np);
}

// K5: partial[chunk][m][k] = sum_{c in chunk} A[m][c] Bext[k][c]
template <typename T>
__global__ __launch_bounds__(256) void k_gemm_grad(const T *__restrict__ A, const T *__restrict__ B,
                                                   int M, int Kd, long ncol, long Np, long chunk,
                                                   double *__restrict__ partial)
{
    __shared__ T As[GBK][GBM + 1];
    __shared__ T Bs[GBK][GBN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * GBM, k0 = blockIdx.z * GBN;
    const long cb = (long)blockIdx.x * chunk;
    const long ce = min(ncol, cb + chunk);
    typename Mfma<T>::V acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) acc[i][jj] = typename Mfma<T>::V{0, 0, 0, 0};

    const int r = tid >> 2, cq = (tid & 3) * 4;
    for (long c0 = cb; c0 < ce; c0 += GBK) {
        {
            const int gm = m0 + r;
            T v[4];
            if (gm < M) {
                const T *src = A + (long)gm * ncol + c0 + cq;
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = src[e];
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = T(0);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) As[cq + e][r] = v[e];
            T w[4];
            load_brow4<T, true>(B, ncol, Np, k0 + r, Kd, c0 + cq, w);
#pragma unroll
            for (int e = 0; e < 4; ++e) Bs[cq + e][r] = w[e];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 4) {
            T a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[kk + (lane >> 4)][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) b[jj] = Bs[kk + (lane >> 4)][wn * 32 + jj * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int jj = 0; jj < 2; ++jj) acc[i][jj] = Mfma<T>::mma(a[i], b[jj], acc[i][jj]);
        }
        __syncthreads();
    }
    double *P = partial + (long)blockIdx.x * M * Kd;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int gm = m0 + wm * 32 + i * 16 + Mfma<T>::row(lane, rr);
                const int gk = k0 + wn * 32 + jj * 16 + (lane & 15);
                if (gm < M && gk < Kd) P[(long)gm * Kd + gk] = (double)acc[i][jj][rr];
            }
}

// grad[i] += sum_chunk partial[chunk][i], chunks in ascending order
__global__ void k_reduce_partials(const double *__restrict__ partial, int nchunk, long len,
                                  double *__restrict__ grad)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += partial[(long)c * len + i];
    grad[i] += s;
}

template <typename T>
void launch_gemm_grad(int prec, const T *A, const T *B, int M, int Kd, long ncol, long Np,
                      double *partial, int nchunk, long chunk_cols, double *grad, hipStream_t s)
{
    (void)prec;
    dim3 grid((unsigned)nchunk, (unsigned)((M + GBM - 1) / GBM), (unsigned)((Kd + GBN - 1) / GBN));
    hipLaunchKernelGGL((k_gemm_grad<T>), grid, dim3(256), 0, s, A, B, M, Kd, ncol, Np, chunk_cols,
                       partial);
    const long len = (long)M * Kd;
    hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s,
                       partial, nchunk, len, grad);
}

// per-point gradient rows (no reduction over points): the Jacobian rows of
// the NGD metric, stored transposed, rowsT[w][n]
template <typename T>
__global__ void k_point_rows(const T *__restrict__ A, const T *__restrict__ B, int M, int Kd,
                             int na, long Np, float *__restrict__ rowsT, long ldr)
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
void launch_point_rows(const T *A, const T *B, int M, int Kd, int na, long Np, float *rowsT,
                       long ldr, hipStream_t s)
{
    dim3 grid((unsigned)((Np + 255) / 256), (unsigned)Kd, (unsigned)M);
    hipLaunchKernelGGL((k_point_rows<T>), grid, dim3(256), 0, s, A, B, M, Kd, na, Np, rowsT, ldr);
}

// ---------------------------------------------------------------------------
// K7: output-layer seeds
template <typename T>
__global__ void k_seed_user(const double *__restrict__ seed, int npt, int nout, int na,
                            T *__restrict__ Tb, long Np)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (c >= (long)na * Np) return;
    const int a = (int)(c / Np);
    const long n = c - (long)a * Np;
    Tb[(long)j * na * Np + c] = n < npt ? (T)seed[(n * nout + j) * na + a] : T(0);
}

template <typename T>
void launch_seed_user(const double *seed, int npt, int nout, int na, T *Tb, long Np, hipStream_t s)
{
    dim3 grid((unsigned)(((long)na * Np + 255) / 256), (unsigned)nout);
    hipLaunchKernelGGL((k_seed_user<T>), grid, dim3(256), 0, s, seed, npt, nout, na, Tb, Np);
}

template <typename T>
__global__ __launch_bounds__(256) void k_seed_residual(const T *__restrict__ TL, int nout, int na,
                                                       int npt, long Np, DevResidual rs,
                                                       double fac, int unit_r,
                                                       T *__restrict__ Tb,
                                                       double *__restrict__ rsq_part)
{
    __shared__ double red[256];
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double r2 = 0.0;
    if (n < Np) {
        for (int j = 1; j < nout; ++j)
            for (int a = 0; a < na; ++a) Tb[((long)j * na + a) * Np + n] = T(0);
        for (int a = 0; a < na; ++a) Tb[(long)a * Np + n] = T(0);
        if (n < npt) {
            const double t0 = (double)TL[n];
            double R = 0.0;
            for (int k = 0; k < rs.nterm; ++k) {
                const double tv = (double)TL[(long)rs.slot[k] * Np + n];
                R += rs.nonlin[k] ? rs.coeff[k] * t0 * tv : rs.coeff[k] * tv;
            }
            r2 = R * R;
            if (unit_r) R = 1.0;  // Gauss-Newton row seed dR/dT
            double s0 = 0.0;
            for (int k = 0; k < rs.nterm; ++k) {
                const long o = (long)rs.slot[k] * Np + n;
                double add;
                if (rs.nonlin[k]) {
                    add = fac * R * rs.coeff[k] * t0;
                    s0 += fac * R * rs.coeff[k] * (double)TL[o];
                } else {
                    add = fac * R * rs.coeff[k];
                }
                Tb[o] = (T)((double)Tb[o] + add);
            }
            Tb[n] = (T)((double)Tb[n] + s0);
        }
    }
    red[threadIdx.x] = r2;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) rsq_part[blockIdx.x] = red[0];
}

template <typename T>
void launch_seed_residual(const T *TL, int nout, int na, int npt, long Np, const DevResidual &r,
                          double fac, bool unit_r, T *Tb, double *rsq_part, hipStream_t s)
{
    hipLaunchKernelGGL((k_seed_residual<T>), dim3((unsigned)((Np + 255) / 256)), dim3(256), 0, s, TL,
                       nout, na, npt, Np, r, fac, unit_r ? 1 : 0, Tb, rsq_part);
}

template <typename T>
__global__ void k_output(const T *__restrict__ TL, int nout, int na, int npt, long Np,
                         double *__restrict__ tout)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tot = (long)npt * nout * na;
    if (i >= tot) return;
    const int a = (int)(i % na);
    const long r = i / na;
    const int j = (int)(r % nout);
    const long n = r / nout;
    tout[i] = (double)TL[((long)j * na + a) * Np + n];
}

template <typename T>
void launch_output(const T *TL, int nout, int na, int npt, long Np, double *tout, hipStream_t s)
{
    const long tot = (long)npt * nout * na;
    if (tot == 0) return;
    hipLaunchKernelGGL((k_output<T>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, TL, nout,
                       na, npt, Np, tout);
}

// ---------------------------------------------------------------------------
__global__ void k_convert(const double *__restrict__ s, float *__restrict__ d, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}

void launch_convert_f64_f32(const double *src, float *dst, long n, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_convert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
}

__global__ void k_reduce_sum(const double *__restrict__ p, int n, double *__restrict__ out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

void launch_reduce_sum(const double *part, int n, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(256), 0, s, part, n, out);
}

// Adam, elementwise in the order of lib/train_module.f90:672-682
__global__ void k_adam(double *__restrict__ w, double *__restrict__ m, double *__restrict__ v,
                       const double *__restrict__ g, long n, double eta, double b1, double b2,
                       double eps, double nbatch, double bc1, double bc2)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double gv = g[i] / nbatch;
    const double mi = b1 * m[i] + (1.0 - b1) * gv;
    const double vi = b2 * v[i] + (1.0 - b2) * (gv * gv);
    m[i] = mi;
    v[i] = vi;
    w[i] = w[i] - eta * (mi / bc1) / (sqrt(vi / bc2) + eps);
}

void launch_adam(double *w, double *m, double *v, const double *g, long n, double eta,
                 double beta1, double beta2, double eps, double nbatch, double bc1, double bc2,
                 hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, m, v, g, n,
                           eta, beta1, beta2, eps, nbatch, bc1, bc2);
}

__global__ void k_fill(double *__restrict__ p, long n, double v)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill(double *p, long n, double v, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

__global__ void k_axpy(double *__restrict__ y, const double *__restrict__ x, long n, double a)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

void launch_axpy(double *y, const double *x, long n, double a, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, y, x, n, a);
}

// explicit instantiations
#define INST(T)                                                                                     \
    template void launch_input<T>(const double *, int, int, int, const int *, T *, long, hipStream_t); \
    template void launch_bell_fwd<T>(const T *, T *, T *, int, long, const DevFwdSched &,             \
                                     const ActParams &, T *, hipStream_t);                           \
    template void launch_bell_bwd<T>(const T *, const T *, T *, int, long, const DevPairs &,          \
                                     hipStream_t);                                                   \
    template void launch_gemm_wx<T>(int, bool, const T *, int, const T *, T *, int, int, long, long,  \
                                    hipStream_t);                                                    \
    template void launch_gemm_grad<T>(int, const T *, const T *, int, int, long, long, double *, int, \
                                      long, double *, hipStream_t);                                  \
    template void launch_point_rows<T>(const T *, const T *, int, int, int, long, float *, long,      \
                                       hipStream_t);                                                 \
    template void launch_seed_user<T>(const double *, int, int, int, T *, long, hipStream_t);         \
    template void launch_seed_residual<T>(const T *, int, int, int, long, const DevResidual &, double, \
                                          bool, T *, double *, hipStream_t);                         \
    template void launch_output<T>(const T *, int, int, int, long, double *, hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace dnnf90
