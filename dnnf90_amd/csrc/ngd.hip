// ngd.hip -- the natural-gradient (Gauss-Newton) step on the GPU.
//
// Reference: app/optimizer_module.f90:221-469 (ngd_apply_inv, primal route)
// and :472-522 (gauss_solve).  There: G = (1/N_b) sum_n j_n j_n^T, damping
// c = mu (ABS) or mu tr(G) (TRACE), u = (G + cI)^{-1} (b / N_b), solved by
// Gaussian elimination with partial pivoting.  Here G + cI is symmetric
// positive definite (c > 0), so it is factored by a blocked FP64 Cholesky,
// G + cI = L L^T, followed by the two triangular solves; for an SPD system
// the two routes agree to rounding.
//
//   gram        G = R R^T, R the transposed FP32 rows (gemm.hip)
//   k_damp      G <- G / nb + c I (lower triangle), b <- b / nb
//   k_potrf     Cholesky of one NB x NB diagonal block in LDS
//   k_trsm      L21 = A21 L11^-T for the panel under it
//   k_syrk      A22 -= L21 L21^T on v_mfma_f64_16x16x4f64, lower tiles only
//   k_trsv_*    blocked forward and backward substitution
#include <algorithm>

#include "kernels.hpp"

namespace dnnf90 {

typedef double dbl4 __attribute__((ext_vector_type(4)));

static constexpr int NB = 64;

// tr(G) of the raw Gram matrix (before scaling); one block, fixed order
__global__ void k_trace(const double *__restrict__ G, int n, long lda, double *__restrict__ out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += G[(long)i * lda + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// G <- G / nb + c I on the lower triangle; c = mu (ABS) or mu tr(G)/nb (TRACE)
__global__ void k_damp(double *__restrict__ G, int n, long lda, double nb, double mu, int mode,
                       const double *__restrict__ trace, double *__restrict__ b)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (idx > i) return;
    const double c = mode == 1 ? mu : mu * (*trace / nb);
    double v = G[(long)i * lda + idx] / nb;
    if (idx == i) {
        v += c;
        b[i] = b[i] / nb;
    }
    G[(long)i * lda + idx] = v;
}

// Cholesky of the diagonal block at (k, k), in place (lower); info > 0 marks
// the first non-positive pivot (global index + 1)
__global__ __launch_bounds__(256) void k_potrf(double *__restrict__ A, long lda, int k, int nb,
                                               int *__restrict__ info)
{
    __shared__ double S[NB][NB + 1];
    const int tid = threadIdx.x;
    for (int u = tid; u < nb * nb; u += 256) {
        const int r = u / nb, c = u % nb;
        S[r][c] = c <= r ? A[(long)(k + r) * lda + k + c] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        if (tid == 0) {
            const double d = S[j][j];
            if (!(d > 0.0) && *info == 0) *info = k + j + 1;
            S[j][j] = sqrt(d > 0.0 ? d : 1.0);
        }
        __syncthreads();
        const double djj = S[j][j];
        for (int i = j + 1 + tid; i < nb; i += 256) S[i][j] /= djj;
        __syncthreads();
        // trailing update of the block: S[i][c] -= S[i][j] S[c][j], j < c <= i
        const int m = nb - j - 1;
        for (int u = tid; u < m * m; u += 256) {
            const int i = j + 1 + u / m, c = j + 1 + u % m;
            if (c <= i) S[i][c] -= S[i][j] * S[c][j];
        }
        __syncthreads();
    }
    for (int u = tid; u < nb * nb; u += 256) {
        const int r = u / nb, c = u % nb;
        if (c <= r) A[(long)(k + r) * lda + k + c] = S[r][c];
    }
}

// rows i >= k+nb of the panel: x L11^T = a, one lane per row, L11 in LDS
__global__ __launch_bounds__(64) void k_trsm(double *__restrict__ A, long lda, int k, int nb, int n)
{
    // L is read by broadcast; X is [column][lane] so lane accesses are
    // conflict free; 2 x 32 KiB of LDS
    __shared__ double L[NB][NB];
    __shared__ double X[NB][64];
    const int tid = threadIdx.x;
    for (int u = tid; u < nb * nb; u += 64) {
        const int r = u / nb, c = u % nb;
        L[r][c] = c <= r ? A[(long)(k + r) * lda + k + c] : 0.0;
    }
    const int row0 = k + nb + blockIdx.x * 64;
    for (int u = tid; u < 64 * nb; u += 64) {
        const int r = u / nb, c = u % nb;
        X[c][r] = row0 + r < n ? A[(long)(row0 + r) * lda + k + c] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        double s = X[j][tid];
        for (int m = 0; m < j; ++m) s -= X[m][tid] * L[j][m];
        X[j][tid] = s / L[j][j];
    }
    __syncthreads();
    for (int u = tid; u < 64 * nb; u += 64) {
        const int r = u / nb, c = u % nb;
        if (row0 + r < n) A[(long)(row0 + r) * lda + k + c] = X[c][r];
    }
}

// trailing update A22 -= L21 L21^T over 64 x 64 tiles with ti >= tj; the
// panel width is NB (the last panel may be narrower: nb)
__global__ __launch_bounds__(256) void k_syrk(double *__restrict__ A, long lda, int k, int nb, int n)
{
    __shared__ double Pi[NB][64];  // [kk][row of tile i]; 2 x 32 KiB
    __shared__ double Pj[NB][64];
    // tile index t -> (ti, tj) with ti >= tj in the trailing matrix
    const long t = blockIdx.x;
    int ti = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) / 2.0);
    while ((long)(ti + 1) * (ti + 2) / 2 <= t) ++ti;
    while ((long)ti * (ti + 1) / 2 > t) --ti;
    const int tj = (int)(t - (long)ti * (ti + 1) / 2);
    const int base = k + nb;
    const int r0 = base + ti * 64, c0 = base + tj * 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int u = tid; u < 64 * nb; u += 256) {
        const int r = u / nb, kk = u % nb;
        Pi[kk][r] = r0 + r < n ? A[(long)(r0 + r) * lda + k + kk] : 0.0;
        Pj[kk][r] = c0 + r < n ? A[(long)(c0 + r) * lda + k + kk] : 0.0;
    }
    __syncthreads();
    const int wm = wave >> 1, wn = wave & 1;
    dbl4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) acc[i][jj] = dbl4{0, 0, 0, 0};
    for (int kk = 0; kk < nb; kk += 4) {
        const int kr = kk + (lane >> 4);
        double a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = kr < nb ? Pi[kr][wm * 32 + i * 16 + (lane & 15)] : 0.0;
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) b[jj] = kr < nb ? Pj[kr][wn * 32 + jj * 16 + (lane & 15)] : 0.0;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj)
                acc[i][jj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[jj], acc[i][jj], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gr = r0 + wm * 32 + i * 16 + (lane >> 4) + 4 * r;  // f64 MFMA result rows
                const int gc = c0 + wn * 32 + jj * 16 + (lane & 15);
                if (gr < n && gc <= gr) A[(long)gr * lda + gc] -= acc[i][jj][r];
            }
}

// forward substitution on one diagonal block: y_k = L_kk^-1 y_k (one wave)
__global__ void k_trsv_lower_diag(const double *__restrict__ A, long lda, int k, int nb, double *__restrict__ y)
{
    __shared__ double v[NB];
    const int t = threadIdx.x;
    if (t < nb) v[t] = y[k + t];
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        if (t == 0) v[j] /= A[(long)(k + j) * lda + k + j];
        __syncthreads();
        if (t > j && t < nb) v[t] -= A[(long)(k + t) * lda + k + j] * v[j];
        __syncthreads();
    }
    if (t < nb) y[k + t] = v[t];
}

// y_i -= sum_{m < nb} L[i][k+m] y[k+m] for i >= k+nb: one wave per row
__global__ __launch_bounds__(256) void k_trsv_lower_update(const double *__restrict__ A, long lda, int k, int nb,
                                                            int n, double *__restrict__ y)
{
    const int lane = threadIdx.x & 63;
    const int i = k + nb + blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    double s = lane < nb ? A[(long)i * lda + k + lane] * y[k + lane] : 0.0;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) y[i] -= s;
}

// backward substitution on one diagonal block: u_k = L_kk^-T u_k
__global__ void k_trsv_upper_diag(const double *__restrict__ A, long lda, int k, int nb, double *__restrict__ y)
{
    __shared__ double v[NB];
    const int t = threadIdx.x;
    if (t < nb) v[t] = y[k + t];
    __syncthreads();
    for (int j = nb - 1; j >= 0; --j) {
        if (t == 0) v[j] /= A[(long)(k + j) * lda + k + j];
        __syncthreads();
        if (t < j) v[t] -= A[(long)(k + j) * lda + k + t] * v[j];
        __syncthreads();
    }
    if (t < nb) y[k + t] = v[t];
}

// y_i -= sum_m L[k+m][i] y[k+m] for i < k: lanes along i (row reads coalesce)
__global__ void k_trsv_upper_update(const double *__restrict__ A, long lda, int k, int nb, double *__restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    double s = 0.0;
    for (int m = 0; m < nb; ++m) s += A[(long)(k + m) * lda + i] * y[k + m];
    y[i] -= s;
}

__global__ void k_clear_info(int *info) { *info = 0; }

int ngd_factor_solve(double *G, int n, long lda, double nb, double mu, int mode, double *b, double *trace,
                     int *info, hipStream_t s)
{
    hipLaunchKernelGGL(k_clear_info, dim3(1), dim3(1), 0, s, info);
    hipLaunchKernelGGL(k_trace, dim3(1), dim3(256), 0, s, G, n, lda, trace);
    hipLaunchKernelGGL(k_damp, dim3((unsigned)((n + 255) / 256), (unsigned)n), dim3(256), 0, s, G, n, lda, nb, mu,
                       mode, trace, b);
    for (int k = 0; k < n; k += NB) {
        const int nb_ = std::min(NB, n - k);
        hipLaunchKernelGGL(k_potrf, dim3(1), dim3(256), 0, s, G, lda, k, nb_, info);
        const int rest = n - k - nb_;
        if (rest > 0) {
            hipLaunchKernelGGL(k_trsm, dim3((unsigned)((rest + 63) / 64)), dim3(64), 0, s, G, lda, k, nb_, n);
            const long T = (rest + 63) / 64;
            hipLaunchKernelGGL(k_syrk, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, G, lda, k, nb_, n);
        }
    }
    for (int k = 0; k < n; k += NB) {
        const int nb_ = std::min(NB, n - k);
        hipLaunchKernelGGL(k_trsv_lower_diag, dim3(1), dim3(64), 0, s, G, lda, k, nb_, b);
        const int rest = n - k - nb_;
        if (rest > 0)
            hipLaunchKernelGGL(k_trsv_lower_update, dim3((unsigned)((rest + 3) / 4)), dim3(256), 0, s, G, lda, k,
                               nb_, n, b);
    }
    const int last = ((n - 1) / NB) * NB;
    for (int k = last; k >= 0; k -= NB) {
        const int nb_ = std::min(NB, n - k);
        hipLaunchKernelGGL(k_trsv_upper_diag, dim3(1), dim3(64), 0, s, G, lda, k, nb_, b);
        if (k > 0)
            hipLaunchKernelGGL(k_trsv_upper_update, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, s, G, lda, k,
                               nb_, b);
    }
    return 0;
}

}  // namespace dnnf90
