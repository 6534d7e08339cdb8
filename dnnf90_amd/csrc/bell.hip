// bell.hip -- the Faa di Bruno step (K2 + K3) and its adjoint (K4) on gfx950.
//
//   forward  (lib/train_module.f90:283-325):
//     B(a,1) = S_a,  B(a,q) = sum_{fq(a,q)} C(a,b)/q * S_b * B(a-b, q-1),
//     T_a = sum_q sigma^(q)(S_0) B(a,q),  h_a = sum_q sigma^(q+1)(S_0) B(a,q)
//   adjoint  (lib/train_module.f90:381-403):
//     Sbar_0 = sum_a Tbar_a h_a,  Sbar_b = sum_{pair(b)} C(a,b) Tbar_a h_(a-b)
//
// One lane per (point, neuron): a wave covers 64 consecutive points of one
// neuron, so every slot read or write is one contiguous 64-element row
// segment.  The per-lane operand store (S_a and the B(a, q >= 2) a later
// term reads) lives in LDS as [entry][lane]; sets too large for LDS keep it
// in a global scratch plane of the same shape.  The term tables are
// wave-uniform and read through the scalar cache.
#include "kernels.hpp"

namespace dnnf90 {

#define DNNF90_ACT_SRC(...) __VA_ARGS__
#include "act.inc"
#undef DNNF90_ACT_SRC

template <typename T> struct Recs;
template <> struct Recs<double> {
    typedef TermRec64 R;
    static __device__ __forceinline__ const R *fwd(const DevFwdSched &s) { return s.t64; }
    static __device__ __forceinline__ const R *pairs(const DevPairs &p) { return p.p64; }
};
template <> struct Recs<float> {
    typedef TermRec32 R;
    static __device__ __forceinline__ const R *fwd(const DevFwdSched &s) { return s.t32; }
    static __device__ __forceinline__ const R *pairs(const DevPairs &p) { return p.p32; }
};

// acc += sum over records [t0, t1) of c * E[s] * E[b], in record order; four
// records per step so their scalar loads and LDS reads are in flight together
template <typename T, typename R>
__device__ __forceinline__ T term_sum(const R *__restrict__ rec, int t0, int t1, const T *E, long es, T acc)
{
    int it = t0;
    for (; it + 4 <= t1; it += 4) {
        const R r0 = rec[it], r1 = rec[it + 1], r2 = rec[it + 2], r3 = rec[it + 3];
        const T s0 = E[r0.s * es], b0 = E[r0.b * es], s1 = E[r1.s * es], b1 = E[r1.b * es];
        const T s2 = E[r2.s * es], b2 = E[r2.b * es], s3 = E[r3.s * es], b3 = E[r3.b * es];
        acc += (T)r0.c * s0 * b0;
        acc += (T)r1.c * s1 * b1;
        acc += (T)r2.c * s2 * b2;
        acc += (T)r3.c * s3 * b3;
    }
    for (; it < t1; ++it) {
        const R r = rec[it];
        acc += (T)r.c * E[r.s * es] * E[r.b * es];
    }
    return acc;
}

// the adjoint pair sum reads its two operands from separate stores
template <typename T, typename R>
__device__ __forceinline__ T pair_sum(const R *__restrict__ rec, int t0, int t1, const T *Et, const T *Eh, long es,
                                      T acc)
{
    int it = t0;
    for (; it + 4 <= t1; it += 4) {
        const R r0 = rec[it], r1 = rec[it + 1], r2 = rec[it + 2], r3 = rec[it + 3];
        const T s0 = Et[r0.s * es], b0 = Eh[r0.b * es], s1 = Et[r1.s * es], b1 = Eh[r1.b * es];
        const T s2 = Et[r2.s * es], b2 = Eh[r2.b * es], s3 = Et[r3.s * es], b3 = Eh[r3.b * es];
        acc += (T)r0.c * s0 * b0;
        acc += (T)r1.c * s1 * b1;
        acc += (T)r2.c * s2 * b2;
        acc += (T)r3.c * s3 * b3;
    }
    for (; it < t1; ++it) {
        const R r = rec[it];
        acc += (T)r.c * Et[r.s * es] * Eh[r.b * es];
    }
    return acc;
}

// copy n slot rows of one lane (stride src_s) into its operand store
// (stride dst_s); sixteen loads are in flight before the first store, so a
// lane's whole slot column is mostly fetched in one or two round trips
template <typename T>
__device__ __forceinline__ void gather_rows(const T *__restrict__ src, long src_s, T *dst, long dst_s, int n)
{
    constexpr int G = 16;
    int a = 0;
    for (; a + G <= n; a += G) {
        T v[G];
#pragma unroll
        for (int e = 0; e < G; ++e) v[e] = src[(long)(a + e) * src_s];
#pragma unroll
        for (int e = 0; e < G; ++e) dst[(long)(a + e) * dst_s] = v[e];
    }
    if (a < n) {
        T v[G];
#pragma unroll
        for (int e = 0; e < G; ++e)
            if (a + e < n) v[e] = src[(long)(a + e) * src_s];
#pragma unroll
        for (int e = 0; e < G; ++e)
            if (a + e < n) dst[(long)(a + e) * dst_s] = v[e];
    }
}

template <typename T, bool kLDS, int KQ, bool kBes>
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
    T *E;
    long es;
    if (kLDS) {
        E = reinterpret_cast<T *>(smem_raw) + threadIdx.x;
        es = blockDim.x;
    } else {
        E = scratch + (long)j * Np + n;
        es = (long)nrow * Np;
    }
    gather_rows<T>(S + (long)j * ldr + n, Np, E, es, na);

    double d[KQ];
    act_derivs<KQ, kBes>(ap, (double)E[0], d);
    T dq[KQ];
#pragma unroll
    for (int q = 0; q < KQ; ++q) dq[q] = (T)d[q];

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
        for (int q = 2; q + 1 < KQ; ++q) {
            if (q <= p) {
                const int g = g0 + q - 2;
                const T bsv = term_sum<T>(Recs<T>::fwd(sch), sch.grp_term[g], sch.grp_term[g + 1], E, es, T(0));
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
    // at most 64 KiB per workgroup, so that two or more stay resident per CU
    for (int nt = 256; nt >= 64; nt >>= 1)
        if (bytes_per_thread * nt <= 65536) return nt;
    return 0;
}

int bell_fwd_threads(int nentry, int elem_bytes) { return pick_threads((long)nentry * elem_bytes); }
int bell_bwd_threads(int na, int elem_bytes) { return pick_threads(2L * na * elem_bytes); }

template <typename T, int KQ, bool kBes>
static void launch_fwd_kq2(const T *S, T *Tout, T *H, int nrow, long Np, const DevFwdSched &sch,
                           const ActParams &act, T *scratch, hipStream_t s)
{
    const int nt = bell_fwd_threads(sch.nentry, sizeof(T));
    if (nt > 0) {
        dim3 grid((unsigned)(Np / 64), (unsigned)((nrow + nt / 64 - 1) / (nt / 64)));
        size_t lds = (size_t)sch.nentry * nt * sizeof(T);
        hipLaunchKernelGGL((k_bell_fwd<T, true, KQ, kBes>), grid, dim3(nt), lds, s, S, Tout, H, nrow, Np, sch,
                           act, (T *)nullptr);
    } else {
        dim3 grid((unsigned)(Np / 64), (unsigned)((nrow + 3) / 4));
        hipLaunchKernelGGL((k_bell_fwd<T, false, KQ, kBes>), grid, dim3(256), 0, s, S, Tout, H, nrow, Np, sch,
                           act, scratch);
    }
}

template <typename T, int KQ>
static void launch_fwd_kq(const T *S, T *Tout, T *H, int nrow, long Np, const DevFwdSched &sch,
                          const ActParams &act, T *scratch, hipStream_t s)
{
    if (act.iact == 3 || act.iact == 4)
        launch_fwd_kq2<T, KQ, true>(S, Tout, H, nrow, Np, sch, act, scratch, s);
    else
        launch_fwd_kq2<T, KQ, false>(S, Tout, H, nrow, Np, sch, act, scratch, s);
}

// sigma^(0..k+1) needs k+2 values; instantiations cover k <= 3, 4, 7, 15
template <typename T>
void launch_bell_fwd(const T *S, T *Tout, T *H, int nrow, long Np, const DevFwdSched &sch,
                     const ActParams &act, T *scratch, hipStream_t s)
{
    if (sch.k <= 3)
        launch_fwd_kq<T, 5>(S, Tout, H, nrow, Np, sch, act, scratch, s);
    else if (sch.k <= 4)
        launch_fwd_kq<T, 6>(S, Tout, H, nrow, Np, sch, act, scratch, s);
    else if (sch.k <= 7)
        launch_fwd_kq<T, 9>(S, Tout, H, nrow, Np, sch, act, scratch, s);
    else
        launch_fwd_kq<T, 17>(S, Tout, H, nrow, Np, sch, act, scratch, s);
}

template <typename T, bool kLDS>
__global__ __launch_bounds__(256) void k_bell_bwd(const T *__restrict__ Tb, const T *__restrict__ H,
                                                  T *__restrict__ Sb, int nrow, long Np, DevPairs pr)
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
        gather_rows<T>(tb, Np, E, es, na);
        gather_rows<T>(hh, Np, E + (long)na * es, es, na);
        et = E;
        eh = E + (long)na * es;
    } else {
        et = tb;
        eh = hh;
        es = Np;
    }
    T *sb = Sb + (long)j * ldr + n;
    T acc = et[0] * eh[0];
#pragma unroll 4
    for (int a = 1; a < na; ++a) acc += et[a * es] * eh[a * es];
    sb[0] = acc;
    const typename Recs<T>::R *rec = Recs<T>::pairs(pr);
    for (int b = 1; b < na; ++b) {
        const int p0 = pr.start[b];
        sb[b * Np] = pair_sum<T>(rec, p0, p0 + pr.num[b], et, eh, es, T(0));
    }
}

template <typename T>
void launch_bell_bwd(const T *Tb, const T *H, T *Sb, int nrow, long Np, const DevPairs &pr, hipStream_t s)
{
    const int nt = bell_bwd_threads(pr.na, sizeof(T));
    if (nt > 0) {
        dim3 grid((unsigned)(Np / 64), (unsigned)((nrow + nt / 64 - 1) / (nt / 64)));
        hipLaunchKernelGGL((k_bell_bwd<T, true>), grid, dim3(nt), (size_t)2 * pr.na * nt * sizeof(T), s, Tb, H,
                           Sb, nrow, Np, pr);
    } else {
        dim3 grid((unsigned)(Np / 64), (unsigned)((nrow + 3) / 4));
        hipLaunchKernelGGL((k_bell_bwd<T, false>), grid, dim3(256), 0, s, Tb, H, Sb, nrow, Np, pr);
    }
}

// ---------------------------------------------------------------------------
// value-only path (lib/train_module.f90:525-631): Z = sigma(S), Zd =
// sigma'(S) elementwise over n values; the adjoint is Sbar = Tbar * Zd
template <typename T, bool kBes>
__global__ void k_act_value(const T *__restrict__ S, T *__restrict__ Z, T *__restrict__ Zd, long n, ActParams ap)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d[2];
    act_derivs<2, kBes>(ap, (double)S[i], d);
    Z[i] = (T)d[0];
    Zd[i] = (T)d[1];
}

template <typename T>
__global__ void k_mul_value(const T *__restrict__ Tb, const T *__restrict__ Zd, T *__restrict__ Sb, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Sb[i] = Tb[i] * Zd[i];
}

template <typename T>
void launch_act_value(const T *S, T *Z, T *Zd, long n, const ActParams &ap, hipStream_t s)
{
    const unsigned g = (unsigned)((n + 255) / 256);
    const bool bes = ap.iact == 3 || ap.iact == 4;   // J0, J1: the series variant of act_derivs
    if (bes)
        hipLaunchKernelGGL((k_act_value<T, true>), dim3(g), dim3(256), 0, s, S, Z, Zd, n, ap);
    else
        hipLaunchKernelGGL((k_act_value<T, false>), dim3(g), dim3(256), 0, s, S, Z, Zd, n, ap);
}

template <typename T>
void launch_mul_value(const T *Tb, const T *Zd, T *Sb, long n, hipStream_t s)
{
    hipLaunchKernelGGL((k_mul_value<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Tb, Zd, Sb, n);
}

#define INST(T)                                                                                         \
    template void launch_act_value<T>(const T *, T *, T *, long, const ActParams &, hipStream_t);         \
    template void launch_mul_value<T>(const T *, const T *, T *, long, hipStream_t);                     \
    template void launch_bell_fwd<T>(const T *, T *, T *, int, long, const DevFwdSched &, const ActParams &, \
                                     T *, hipStream_t);                                                  \
    template void launch_bell_bwd<T>(const T *, const T *, T *, int, long, const DevPairs &, hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace dnnf90
