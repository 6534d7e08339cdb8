// misc.hip -- input layer, output-layer seeds, output gather and the small
// elementwise kernels (Adam, fills) of the derivative-carrying pass.
#include "kernels.hpp"

#include <map>
#include <mutex>
#include <utility>

namespace dnnf90 {

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
                                                       double *__restrict__ rsq_part,
                                                       const double *__restrict__ r0,
                                                       double *__restrict__ r_out)
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
            double R = r0 ? r0[n] : 0.0;
#pragma unroll 4
            for (int k = 0; k < rs.nterm; ++k) {
                const double tv = (double)TL[(long)rs.slot[k] * Np + n];
                R += rs.nonlin[k] ? rs.coeff[k] * t0 * tv : rs.coeff[k] * tv;
            }
            r2 = R * R;
            if (r_out) r_out[n] = R;
            if (unit_r) R = 1.0;  // Gauss-Newton row seed dR/dT
            double s0 = 0.0;
            if (rs.distinct) {
                // each seed slot is written once, from the zero it starts at
                // (no read-back of the plane just cleared; the same values)
#pragma unroll 4
                for (int k = 0; k < rs.nterm; ++k) {
                    const long o = (long)rs.slot[k] * Np + n;
                    double add;
                    if (rs.nonlin[k]) {
                        add = fac * R * rs.coeff[k] * t0;
                        s0 += fac * R * rs.coeff[k] * (double)TL[o];
                    } else {
                        add = fac * R * rs.coeff[k];
                    }
                    Tb[o] = (T)(0.0 + add);
                }
                Tb[n] = (T)(0.0 + s0);
            } else {
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
                          double fac, bool unit_r, T *Tb, double *rsq_part, hipStream_t s,
                          const double *r0, double *r_out)
{
    hipLaunchKernelGGL((k_seed_residual<T>), dim3((unsigned)((Np + 255) / 256)), dim3(256), 0, s, TL,
                       nout, na, npt, Np, r, fac, unit_r ? 1 : 0, Tb, rsq_part, r0, r_out);
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
    // the reference's arithmetic, operation for operation, unfused (its
    // build multiplies and adds separately): the same bits
#pragma clang fp contract(off)
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

// value-only output seed (lib/train_module.f90:585-586): D = coef (Z - Y)
// on the npt points of the single output, 0 on the padding
template <typename T>
__global__ void k_seed_value(const T *__restrict__ Z, const double *__restrict__ y, double coef, int npt, long Np,
                             T *__restrict__ D)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= Np) return;
    D[n] = n < npt ? (T)(coef * ((double)Z[n] - y[n])) : T(0);
}

template <typename T>
void launch_seed_value(const T *Z, const double *y, double coef, int npt, long Np, T *D, hipStream_t s)
{
    hipLaunchKernelGGL((k_seed_value<T>), dim3((unsigned)((Np + 255) / 256)), dim3(256), 0, s, Z, y, coef, npt, Np, D);
}

#define INST(T)                                                                                     \
    template void launch_input<T>(const double *, int, int, int, const int *, T *, long, hipStream_t); \
    template void launch_seed_user<T>(const double *, int, int, int, T *, long, hipStream_t);         \
    template void launch_seed_residual<T>(const T *, int, int, int, long, const DevResidual &, double, \
                                          bool, T *, double *, hipStream_t, const double *, double *); \
    template void launch_output<T>(const T *, int, int, int, long, double *, hipStream_t);              \
    template void launch_seed_value<T>(const T *, const double *, double, int, long, T *, hipStream_t);
INST(double)
INST(float)
#undef INST

static std::mutex g_attr_mu;
static std::map<const void *, size_t> g_lds_limit;
static std::map<std::pair<const void *, size_t>, int> g_occ;

static void lds_allow_locked(const void *fn, size_t lds)
{
    if (lds <= 64 * 1024) return;
    size_t &cur = g_lds_limit[fn];
    if (lds > cur) {
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        cur = lds;
    }
}

void lds_allow(const void *fn, size_t lds)
{
    std::lock_guard<std::mutex> lk(g_attr_mu);
    lds_allow_locked(fn, lds);
}

int occupancy_cached(const void *fn, size_t lds, int threads)
{
    std::lock_guard<std::mutex> lk(g_attr_mu);
    lds_allow_locked(fn, lds);
    const auto key = std::make_pair(fn, lds);
    auto it = g_occ.find(key);
    if (it != g_occ.end()) return it->second;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, threads, lds) != hipSuccess || n < 1) n = 1;
    g_occ[key] = n;
    return n;
}

}  // namespace dnnf90
