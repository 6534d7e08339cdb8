// kalman.hip -- the extended Kalman filter over the weights
// (lib/kalman_module.f90:107-659) with its covariance resident in HBM, and
// the committee's running mean and spread (lib/committee_module.f90:79-111).
//
// Dense covariance.  P is symmetric at every step (p0 I at the start, and
// every update symmetrises it, :473-480), so only its lower triangle is
// kept: 64 x 64 tiles (I, J), I >= J, tile (I, J) at index I (I + 1) / 2 + J,
// row-major inside the tile; diagonal tiles are kept whole.  Rows and
// columns past nw are zero and stay zero (kv and pj are zero there, and q
// is added to real diagonal entries only).
//
// One observation reads and writes every tile once.  The pass first applies
// the downdate the previous observation left pending,
//     P <- sym((P - kv pj^T) / lam) + q I                 (:466-487)
// and then, on the updated tile, forms its share of P j for the new row j:
// the row sums of tile (I, J) against j over the columns of J go to the rows
// of I, and (I != J) its column sums against j over the rows of I go to the
// rows of J.  part[J][rows of I] and part[I][rows of J] receive them, and a
// second kernel sums part over the column-tile index in a fixed order, so
// the result does not depend on scheduling.  The reference sweeps the full
// nw x nw matrix three times per observation (product, downdate,
// symmetrisation); here it is one read and one write of half of it.
//
// Decoupled covariance (:71-83): one bsz x bsz block per neuron, bsz = its
// fan-in + 1, kept whole and row-major, packed one after the other.  One
// workgroup owns a block, applies the pending downdate and forms the block's
// P j in the same sweep, the sum over rows running in the reference's order.
//
// The element-wise arithmetic (downdate, symmetrisation, q, gate, gain,
// weight step, lambda schedule) is the reference's operation for operation
// with floating-point contraction off, so for equal inputs it agrees bit for
// bit; the products P j and j.P j are sums in another order (dense filter).
#include "kernels.hpp"

namespace dnnf90 {

static constexpr int KT = 64;     // covariance tile edge

__device__ inline void kf_tile_of(long b, int &I, int &J)
{
    long i = (long)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= b) ++i;
    while (i * (i + 1) / 2 > b) --i;
    I = (int)i;
    J = (int)(b - i * (i + 1) / 2);
}

// P = p0 I over the lower tiles
__global__ __launch_bounds__(256) void k_kf_init_tiles(double *__restrict__ P, int nw, double p0)
{
    int I, J;
    kf_tile_of(blockIdx.x, I, J);
    double *tp = P + (long)blockIdx.x * KT * KT;
    for (int e = threadIdx.x; e < KT * KT; e += 256) {
        const int r = e >> 6, c = e & 63;
        const long gr = (long)I * KT + r, gc = (long)J * KT + c;
        tp[e] = (gr == gc && gr < nw) ? p0 : 0.0;
    }
}

__global__ __launch_bounds__(256) void k_kf_init_blocks(double *__restrict__ P, const long *__restrict__ poff,
                                                        const int *__restrict__ bsz, double p0)
{
    const int n = bsz[blockIdx.x];
    double *pb = P + poff[blockIdx.x];
    for (long e = threadIdx.x; e < (long)n * n; e += 256) pb[e] = (e / n == e % n) ? p0 : 0.0;
}

// pending downdate of tile (I, J), then (want) its share of P j.  Thread t
// owns column c = t % 64 of rows t / 64 + 4 i, i < 16: its sixteen loads are
// issued together, kv and pj of the tile's rows come from LDS and those of
// its column from registers.
__global__ __launch_bounds__(256) void k_kf_tiles(double *__restrict__ P, int nt, int nw,
                                                  const double *__restrict__ g, double scale,
                                                  const double *__restrict__ kv, const double *__restrict__ pj,
                                                  const KfScalars *__restrict__ st, int want,
                                                  double *__restrict__ part)
{
#pragma clang fp contract(off)
    __shared__ double S[KT][KT + 1];
    __shared__ double jr[KT], jc[KT], kvr[KT], pjr[KT];
    __shared__ double red[4][KT];
    int I, J;
    kf_tile_of(blockIdx.x, I, J);
    const int t = threadIdx.x, c = t & 63, w = t >> 6;
    double *tp = P + (long)blockIdx.x * KT * KT;
    double v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = tp[t + 256 * i];
    const bool pend = st->pending && !st->bad;
    if (t < KT) {
        const long gr = (long)I * KT + t, gc = (long)J * KT + t;
        if (pend) {
            kvr[t] = kv[gr];
            pjr[t] = pj[gr];
        }
        if (want) {
            jr[t] = gr < nw ? scale * g[gr] : 0.0;   // j over the rows of I
            jc[t] = gc < nw ? scale * g[gc] : 0.0;   // j over the rows of J
        }
    }
    __syncthreads();
    if (pend) {
        const double lam = st->lam_dd, q = st->q;
        const long gc = (long)J * KT + c;
        const double kvc = kv[gc], pjc = pj[gc];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int r = w + 4 * i;
            const long gr = (long)I * KT + r;
            // P'(gr,gc) and P'(gc,gr) from the one stored value (P symmetric)
            const double a = (v[i] - kvr[r] * pjc) / lam;
            if (gr != gc) {
                const double b = (v[i] - kvc * pjr[r]) / lam;
                v[i] = 0.5 * (a + b);
            } else {
                v[i] = a;
                if (q > 0.0 && gr < nw) v[i] = v[i] + q;
            }
            tp[t + 256 * i] = v[i];
        }
    }
    if (!want) return;
    const long ldp = (long)nt * KT;
#pragma unroll
    for (int i = 0; i < 16; ++i) S[w + 4 * i][c] = v[i];
    if (I != J) {
        // rows of J: column sums against j over the rows of I, four waves of
        // sixteen rows each, combined in a fixed order
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) s += v[i] * jr[w + 4 * i];
        red[w][c] = s;
    }
    __syncthreads();
    {
        // rows of I: four lanes per row, sixteen columns each
        const int r = t >> 2, qd = t & 3;
        double s = 0.0;
        for (int cc = qd * 16; cc < qd * 16 + 16; ++cc) s += S[r][cc] * jc[cc];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        if (qd == 0) part[(long)J * ldp + (long)I * KT + r] = s;
    }
    if (I != J && t < KT) part[(long)I * ldp + (long)J * KT + t] = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
}

// pj[row] = sum over column tiles k of part[k][row]: 32 rows a workgroup,
// eight k-groups (k = g, g + 8, ...) per row, combined in a fixed order
__global__ __launch_bounds__(256) void k_kf_sum_part(const double *__restrict__ part, int nt, double *__restrict__ pj)
{
    __shared__ double red[8][32];
    const long ldp = (long)nt * KT;
    const int rl = threadIdx.x & 31, kg = threadIdx.x >> 5;
    const long row = (long)blockIdx.x * 32 + rl;   // ldp is a multiple of 32
    double s = 0.0;
    for (int k = kg; k < nt; k += 8) s += part[(long)k * ldp + row];
    red[kg][rl] = s;
    __syncthreads();
    if (threadIdx.x < 32)
        pj[row] = ((red[0][rl] + red[1][rl]) + (red[2][rl] + red[3][rl])) +
                  ((red[4][rl] + red[5][rl]) + (red[6][rl] + red[7][rl]));
}

// decoupled: one workgroup per neuron block (kalman_module.f90:389-396,
// :447-464); thread c owns column c, rows in ascending order
__global__ __launch_bounds__(256) void k_kf_blocks(double *__restrict__ P, const long *__restrict__ poff,
                                                   const int *__restrict__ bsz, const int *__restrict__ boff,
                                                   const double *__restrict__ g, double scale,
                                                   const double *__restrict__ kv, double *__restrict__ pj,
                                                   const KfScalars *__restrict__ st, int want)
{
#pragma clang fp contract(off)
    extern __shared__ double acc[];
    const int n = bsz[blockIdx.x], o = boff[blockIdx.x];
    double *pb = P + poff[blockIdx.x];
    const bool pend = st->pending && !st->bad;
    const double lam = st->lam_dd, q = st->q;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const double kvc = kv[o + c], pjc = pj[o + c];
        double s = 0.0;
        for (int r = 0; r < n; ++r) {
            double v = pb[(long)r * n + c];
            if (pend) {
                const double a = (v - kv[o + r] * pjc) / lam;
                if (r != c) {
                    const double b = (v - kvc * pj[o + r]) / lam;
                    v = 0.5 * (a + b);
                } else {
                    v = a;
                    if (q > 0.0) v = v + q;
                }
                pb[(long)r * n + c] = v;
            }
            if (want) s += v * (scale * g[o + r]);
        }
        acc[c] = s;
    }
    if (!want) return;
    __syncthreads();   // every thread has read the old pj of the block
    for (int c = threadIdx.x; c < n; c += blockDim.x) pj[o + c] = acc[c];
}

// denominator, innovation, gate and lambda schedule (one workgroup)
__global__ __launch_bounds__(256) void k_kf_gain(KfScalars *__restrict__ st, const double *__restrict__ g,
                                                 double scale, const double *__restrict__ pj, int nw, int xmode,
                                                 double v, const double *__restrict__ seed,
                                                 const double *__restrict__ T, int nsl,
                                                 const double *__restrict__ rv,
                                                 const double *__restrict__ wprior, const double *__restrict__ w,
                                                 int final)
{
#pragma clang fp contract(off)
    __shared__ double red[2][256];
    if (st->bad) return;
    const int t = threadIdx.x;
    double s = 0.0, c = 0.0;
    for (int m = t; m < nw; m += 256) {
        const double jm = scale * g[m];
        s += jm * pj[m];
        if (xmode == KF_XI_IEKF) c += jm * (wprior[m] - w[m]);
    }
    red[0][t] = s;
    red[1][t] = c;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            red[0][t] += red[0][t + h];
            red[1][t] += red[1][t + h];
        }
        __syncthreads();
    }
    if (t) return;
    const double lam = st->lam;
    double denom = lam + red[0][0];
    if (!(denom > 0.0)) {   // also catches a NaN
        st->bad = 1;
        return;
    }
    double xi;
    if (xmode == KF_XI_LINEAR) {
        double pred = 0.0;
        for (int i = 0; i < nsl; ++i) pred = pred + seed[i] * T[i];
        xi = scale * (v - pred);
    } else if (xmode == KF_XI_RESID) {
        xi = scale * (v - rv[0]);
    } else if (xmode == KF_XI_IEKF) {
        xi = -scale * v - red[1][0];
    } else {
        xi = scale * v;
    }
    double gs = 1.0;
    int gated = 0;
    if (xmode != KF_XI_IEKF && st->gate > 0.0) {
        const double gate = st->gate, jpj = denom - lam;
        if (xi * xi > gate * gate * denom) {
            const double s2 = gate * gate * lam / (xi * xi - gate * gate * jpj);
            gs = sqrt(s2);
            xi = gs * xi;
            denom = lam + s2 * jpj;
            st->ngated += 1;
            gated = 1;
        }
    }
    st->xi = xi;
    st->denom = denom;
    st->gs = gs;
    st->gated = gated;
    if (final) {
        st->lam_dd = lam;
        st->lam = st->lam0 * lam + 1.0 - st->lam0;
        st->pending = 1;
    } else {
        st->pending = 0;
    }
}

// gain kv = pj / denom and the weight step (:421-431, :593-608)
__global__ __launch_bounds__(256) void k_kf_apply(const KfScalars *__restrict__ st, double *__restrict__ pj,
                                                  double *__restrict__ kv, double *__restrict__ w,
                                                  const double *__restrict__ wprior, int nw, int iekf)
{
#pragma clang fp contract(off)
    const int m = blockIdx.x * 256 + threadIdx.x;
    if (m >= nw || st->bad) return;
    double p = pj[m];
    if (st->gated) {
        p = st->gs * p;
        pj[m] = p;
    }
    const double k = p / st->denom;
    kv[m] = k;
    w[m] = (iekf ? wprior[m] : w[m]) + k * st->xi;
}

__global__ void k_kf_clear(KfScalars *st) { st->pending = 0; }

void launch_kf_init_tiles(double *P, int nt, int nw, double p0, hipStream_t s)
{
    const long nt2 = (long)nt * (nt + 1) / 2;
    hipLaunchKernelGGL(k_kf_init_tiles, dim3((unsigned)nt2), dim3(256), 0, s, P, nw, p0);
}

void launch_kf_init_blocks(double *P, const long *poff, const int *bsz, int nb, double p0, hipStream_t s)
{
    hipLaunchKernelGGL(k_kf_init_blocks, dim3((unsigned)nb), dim3(256), 0, s, P, poff, bsz, p0);
}

void launch_kf_tiles(double *P, int nt, int nw, const double *g, double scale, const double *kv,
                     const double *pj, const KfScalars *st, bool want, double *part, hipStream_t s)
{
    const long nt2 = (long)nt * (nt + 1) / 2;
    hipLaunchKernelGGL(k_kf_tiles, dim3((unsigned)nt2), dim3(256), 0, s, P, nt, nw, g, scale, kv, pj, st,
                       want ? 1 : 0, part);
}

void launch_kf_sum_part(const double *part, int nt, double *pj, hipStream_t s)
{
    const long rows = (long)nt * KT;
    hipLaunchKernelGGL(k_kf_sum_part, dim3((unsigned)(rows / 32)), dim3(256), 0, s, part, nt, pj);
}

void launch_kf_blocks(double *P, const long *poff, const int *bsz, const int *boff, int nb, int maxb,
                      const double *g, double scale, const double *kv, double *pj, const KfScalars *st, bool want,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_kf_blocks, dim3((unsigned)nb), dim3(256), sizeof(double) * maxb, s, P, poff, bsz, boff, g,
                       scale, kv, pj, st, want ? 1 : 0);
}

void launch_kf_gain(KfScalars *st, const double *g, double scale, const double *pj, int nw, int xmode, double v,
                    const double *seed, const double *T, int nsl, const double *rv, const double *wprior,
                    const double *w, bool final, hipStream_t s)
{
    hipLaunchKernelGGL(k_kf_gain, dim3(1), dim3(256), 0, s, st, g, scale, pj, nw, xmode, v, seed, T, nsl, rv,
                       wprior, w, final ? 1 : 0);
}

void launch_kf_apply(const KfScalars *st, double *pj, double *kv, double *w, const double *wprior, int nw,
                     bool iekf, hipStream_t s)
{
    hipLaunchKernelGGL(k_kf_apply, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, st, pj, kv, w, wprior, nw,
                       iekf ? 1 : 0);
}

void launch_kf_clear(KfScalars *st, hipStream_t s) { hipLaunchKernelGGL(k_kf_clear, dim3(1), dim3(1), 0, s, st); }

// ---------------------------------------------------------------------------
// committee: Welford over the members, in the reference's operation order
__global__ __launch_bounds__(256) void k_comm_welford(const double *__restrict__ t, double *__restrict__ mean,
                                                      double *__restrict__ m2, long n, int k)
{
#pragma clang fp contract(off)
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double d = t[i] - mean[i];
    m2[i] = m2[i] + d * d * (double)(k - 1) / (double)k;
    mean[i] = mean[i] + d / (double)k;
}

__global__ __launch_bounds__(256) void k_comm_finish(double *__restrict__ m2, long n, int nmem)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    m2[i] = nmem > 1 ? sqrt(m2[i] / (double)(nmem - 1)) : 0.0;
}

void launch_comm_welford(const double *t, double *mean, double *m2, long n, int k, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_comm_welford, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t, mean, m2, n, k);
}

void launch_comm_finish(double *m2, long n, int nmem, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_comm_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m2, n, nmem);
}

}  // namespace dnnf90
