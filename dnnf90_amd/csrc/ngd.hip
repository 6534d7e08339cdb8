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
// Right-looking blocked factorization over panels of PB = 128 columns:
//
//   gram          G = R R^T, lower triangle, R the transposed FP32 rows (gemm.hip)
//   k_damp        G <- G / nb + c I (lower triangle), b <- b / nb
//   k_chol_diag   one workgroup: Cholesky of the PB x PB diagonal block in LDS,
//                 then its inverse Linv (kept for the panel and the solves;
//                 k_linv_transpose forms the transposes after the last panel)
//   k_chol_panel  L21 = A21 Linv^T for the rows under the block: an FP64
//                 MFMA product instead of a row-sequential triangular solve;
//                 also packs L21 row-contiguously for the update
//   k_chol_syrk   A22 -= L21 L21^T on v_mfma_f64_16x16x4f64, lower 64 x 64
//                 tiles, the panel streamed through LDS in 32-column chunks
//   k_chol_fwd / k_chol_bwd
//                 one launch per panel: y_rest -= L21 y_k (or the transposed
//                 update), and the workgroup that owns the next diagonal
//                 block solves it with that block's Linv
#include <algorithm>

#include "kernels.hpp"

namespace dnnf90 {

typedef double dbl4 __attribute__((ext_vector_type(4)));

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

static constexpr int PB = 128;   // panel width
static constexpr int PBL = PB + 1;

// sum_{m < len} x[m] y[m] over LDS rows, four partial sums
__device__ __forceinline__ double ldot(const double *__restrict__ x, const double *__restrict__ y, int len)
{
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int m = 0;
    for (; m + 4 <= len; m += 4) {
        s0 += x[m] * y[m];
        s1 += x[m + 1] * y[m + 1];
        s2 += x[m + 2] * y[m + 2];
        s3 += x[m + 3] * y[m + 3];
    }
    for (; m < len; ++m) s0 += x[m] * y[m];
    return (s0 + s1) + (s2 + s3);
}

// Cholesky of the diagonal block at (k, k) (nb <= PB rows) in LDS, and
// its inverse X = L^-1, written to Linv[PB][PB] (zero above the diagonal;
// rows past nb hold the identity).  The transposes of all the blocks'
// inverses are formed once the factorization is done (k_linv_transpose),
// off the panels' dependent chain.  Right-looking over 16-column steps:
// wave 0 factors the 16 x 16 diagonal block in registers (one row a lane;
// each pivot column goes to every lane through a 16-double LDS row) and
// forms that block's inverse X_kk by forward substitution (one column a
// lane, the rows of L read back from LDS); the rows below take
// L21 = A21 X_kk^T and the trailing lower blocks the rank-16 update, both
// on v_mfma_f64_16x16x4f64 (16 x 16 blocks, a wave a block).  The
// off-diagonal blocks of X follow by block columns, which are independent
// of each other: X_ij = -X_ii (sum_{j <= m < i} L_im X_mj) needs only the
// blocks of column j above it, so a wave walks down whole columns (columns
// j and 6 - j, and column 3, balance the four waves) on FP64 MFMA with no
// barrier between block rows.  X (below the diagonal) is kept transposed
// in the free upper triangle of S, its diagonal in dg.  The block of G
// itself is left as the trailing updates made it (the solves and the panel
// read Linv, LinvT and the rows below the block).  info > 0 marks the
// first non-positive pivot (index + 1).
static constexpr int DB = 16;         // step width inside the diagonal block
static constexpr int ND = PB / DB;    // steps

// X[R][C] of the inverse as S holds it (R > C: S[C][R]; R == C: dg[R])
__device__ __forceinline__ double xinv(const double (*S)[PBL], const double *dg, int R, int C)
{
    return R > C ? S[C][R] : (R == C ? dg[R] : 0.0);
}

// (a) of a step, wave 0: the 16 x 16 diagonal block at (o, o), lane
// r + 16 q holding columns q, q + 4, q + 8, q + 12 of row o + r (a quarter
// of the row's rank-1 updates a lane), then its inverse, lane b forming
// column b of X_kk (lanes 16..63 repeat lanes 0..15).  LDS operations of
// one wave complete in program order, so a value a lane writes is there
// for the reads after it.
__device__ __forceinline__ void diag16(double (*S)[PBL], double *dg, double *colb, int *info, int k, int o, int lane)
{
    const int r = lane & 15, q = lane >> 4;
    double d[4], il_[DB];
    int bad = -1;   // the block's first non-positive pivot
#pragma unroll
    for (int m = 0; m < 4; ++m) d[m] = S[o + r][o + 4 * m + q];
#pragma unroll
    for (int j = 0; j < DB; ++j) {
        const int jm = j >> 2, jq = j & 3;
        // column j as the earlier steps left it, one element a row, to every lane
        if (q == jq) colb[r] = d[jm];
        __builtin_amdgcn_wave_barrier();
        const double djj = colb[j], arj = colb[r];
        double acj[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) acj[m] = colb[4 * m + q];
        __builtin_amdgcn_wave_barrier();
        if (!(djj > 0.0) && bad < 0) bad = j;   // reported after the loop
        // 1/sqrt from v_rsq_f64 and two Newton steps (each doubles the
        // correct bits), l = d * (1/sqrt d): no IEEE sqrt and division
        // sequences on the column's dependent path
        const double dd = djj > 0.0 ? djj : 1.0;
        double il = __builtin_amdgcn_rsq(dd);
        il = il * (1.5 - 0.5 * dd * il * il);
        il = il * (1.5 - 0.5 * dd * il * il);
        il_[j] = il;
        const double l = dd * il;
        const double lrj = arj * il;   // L[o + r][o + j] (rows r > j)
        if (q == jq) d[jm] = r == j ? l : (r > j ? lrj : d[jm]);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int c = 4 * m + q;
            const double lcj = acj[m] * il;   // L[o + c][o + j]
            if (c > j && c <= r) d[m] -= lrj * lcj;
        }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int c = 4 * m + q;
        if (c <= r) S[o + r][o + c] = d[m];
    }
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < DB; ++j) dg[o + j] = il_[j];
        if (bad >= 0 && *info == 0) *info = k + o + bad + 1;
    }
    __builtin_amdgcn_wave_barrier();
    // X[a][b] = -X[a][a] sum_{b <= m < a} L[a][m] X[m][b] (row a of L read
    // back from LDS, the same address in every lane)
    const int b = r;
    double x[DB];
#pragma unroll
    for (int a = 0; a < DB; ++a) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < a; ++m) s += S[o + a][o + m] * x[m];
        x[a] = a < b ? 0.0 : (a == b ? il_[a] : -il_[a] * s);
    }
    if (lane < DB) {
#pragma unroll
        for (int a = 0; a < DB; ++a)
            if (a > b) S[o + b][o + a] = x[a];   // X[a][b] at S[C][R]
    }
}

// (c) of a step: trailing lower block u (row-major over the lower triangle
// of blocks below the step's column o, from t0) -= L(ib, kb) L(jb, kb)^T
__device__ __forceinline__ void trail16(double (*S)[PBL], int o, int t0, int u, int li, int lk)
{
    int ti = 0;
    while ((ti + 1) * (ti + 2) / 2 <= u) ++ti;
    const int tj = u - ti * (ti + 1) / 2;
    const int ra = t0 + DB * ti, rc = t0 + DB * tj;
    double av[4], bv[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        av[c] = S[ra + li][o + 4 * c + lk];
        bv[c] = S[rc + li][o + 4 * c + lk];
    }
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 4; ++c) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[c], bv[c], acc, 0, 0, 0);
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) S[ra + lk + 4 * rr][rc + li] -= acc[rr];
}

// Block (ib, jb) of X = L^-1 (jb < ib), one wave: Y = sum_{jb <= mb < ib}
// L_(ib,mb) X_(mb,jb) on v_mfma_f64_16x16x4f64 (A operand: lane (i = lane %
// 16, k = lane / 16); B: (k = lane / 16, j = lane % 16); accumulator element
// r: row lane / 16 + 4 r, column lane % 16 -- so Y's element c is the B
// operand of k-chunk c of the next product), then X_(ib,jb) = -X_(ib,ib) Y
// into its transposed slot S[C][R] (rows in block jb, columns in block ib).
__device__ __forceinline__ void xblock(double (*S)[PBL], const double *dg, int ib, int jb, int li, int lk)
{
    dbl4 y = {0.0, 0.0, 0.0, 0.0};
    for (int mb = jb; mb < ib; ++mb) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int kk = 4 * c + lk;
            const double a = S[DB * ib + li][DB * mb + kk];   // L[16 ib + li][16 mb + kk]
            const int Rb = DB * mb + kk, Cb = DB * jb + li;   // X[Rb][Cb]
            const double b = mb == jb ? xinv(S, dg, Rb, Cb) : S[Cb][Rb];
            y = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, y, 0, 0, 0);
        }
    }
    dbl4 z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double a = xinv(S, dg, DB * ib + li, DB * ib + 4 * c + lk);   // X_(ib,ib)[li][4 c + lk]
        z = __builtin_amdgcn_mfma_f64_16x16x4f64(a, y[c], z, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) S[DB * jb + li][DB * ib + lk + 4 * r] = -z[r];
}

__global__ __launch_bounds__(256) void k_chol_diag(const double *__restrict__ A, long lda, int k, int nb,
                                                   double *__restrict__ Linv, int *__restrict__ info)
{
    __shared__ double S[PB][PBL];    // L (lower), then X transposed (upper)
    __shared__ double dg[PB];        // 1 / L[j][j] = X[j][j]
    __shared__ double colb[DB];      // a step's pivot column, one element a row
    const int tid = threadIdx.x;
    // the block in one round: every load of a thread in flight together
    // (only the lower triangle is read; the upper one is written before use)
    {
        constexpr int NL = PB * PB / 256;
        double v[NL];
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            // unconditional loads (an element outside the triangle reads the
            // block's first one), so that none waits for another
            const int u = tid + 256 * i, r = u / PB, c = u % PB;
            const bool in = r < nb && c <= r;
            const double x = A[(long)(k + (in ? r : 0)) * lda + k + (in ? c : 0)];
            v[i] = in ? x : (r == c ? 1.0 : 0.0);
        }
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const int u = tid + 256 * i;
            S[u / PB][u % PB] = v[i];
        }
    }
    __syncthreads();
    const int lane = tid & 63, li = lane & 15, lk = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // Step kb: (b) the rows below the block, then (c) the trailing update,
    // in which wave 0 updates the next diagonal block first and factors it
    // ((a) of step kb + 1) while waves 1-3 update the other blocks.
    if (wave == 0) diag16(S, dg, colb, info, k, 0, lane);
    __syncthreads();
    for (int kb = 0; kb + 1 < ND; ++kb) {
        const int o = kb * DB, t0 = o + DB, nrb = ND - 1 - kb;
        // (b) the rows below: L21 = A21 X_kk^T, a 16-row block a wave turn
        // (A operand: lane (i = li, k = lk); B: (k = lk, j = li); result
        // element r at row lk + 4 r, column li)
        {
            double bx[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) bx[c] = xinv(S, dg, o + li, o + 4 * c + lk);   // X_kk[li][4 c + lk]
            for (int rb = wave; rb < nrb; rb += 4) {
                const int r0 = t0 + DB * rb;
                double av[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) av[c] = S[r0 + li][o + 4 * c + lk];
                dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int c = 0; c < 4; ++c) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[c], bx[c], acc, 0, 0, 0);
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) S[r0 + lk + 4 * rr][o + li] = acc[rr];
            }
        }
        __syncthreads();
        // (c) trailing lower blocks (ib >= jb > kb) -= L(ib, kb) L(jb, kb)^T;
        // block 0 is the next diagonal block
        if (wave == 0) {
            trail16(S, o, t0, 0, li, lk);
            __builtin_amdgcn_wave_barrier();
            diag16(S, dg, colb, info, k, t0, lane);
        } else {
            // block row kb of X (its blocks longest first), then the other
            // trailing blocks
            const int nblk = nrb * (nrb + 1) / 2, nitem = kb + nblk - 1;
            for (int it = wave - 1; it < nitem; it += 3) {
                if (it < kb) xblock(S, dg, kb, it, li, lk);
                else trail16(S, o, t0, it - kb + 1, li, lk);
            }
        }
        __syncthreads();
    }
    // the last block row of X (its diagonal block was factored in the last step)
    for (int jb = wave; jb < ND - 1; jb += 4) xblock(S, dg, ND - 1, jb, li, lk);
    __syncthreads();
    // X in its row order, two adjacent elements a thread at a time
    typedef double dbl2 __attribute__((ext_vector_type(2)));
    for (int u = 2 * tid; u < PB * PB; u += 512) {
        const int r = u / PB, c = u % PB;
        dbl2 v;
        v[0] = r == c ? dg[r] : (r > c ? S[c][r] : 0.0);               // X[r][c]
        v[1] = r == c + 1 ? dg[r] : (r > c + 1 ? S[c + 1][r] : 0.0);   // X[r][c + 1]
        *reinterpret_cast<dbl2 *>(Linv + u) = v;
    }
}

// LinvT[p] = Linv[p]^T for every diagonal block p (blockIdx.y), a 32 x 32
// tile a workgroup (blockIdx.x: tile row * 4 + tile column) through LDS
__global__ __launch_bounds__(256) void k_linv_transpose(const double *__restrict__ Linv, double *__restrict__ LinvT)
{
    __shared__ double t[32][33];
    const long base = (long)blockIdx.y * PB * PB;
    const int bi = blockIdx.x >> 2, bj = blockIdx.x & 3;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int q = 0; q < 32; q += 8) t[ty + q][tx] = Linv[base + (long)(32 * bi + ty + q) * PB + 32 * bj + tx];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 32; q += 8) LinvT[base + (long)(32 * bj + ty + q) * PB + 32 * bi + tx] = t[tx][ty + q];
}

// L21 = A21 Linv^T for PR rows per workgroup (rows k+nb.. of the panel's
// columns k..k+nb-1); the rows are staged in LDS (all of a thread's loads
// in flight together), Linv is read through L2 four k-steps ahead of its
// products; wave (b, h) forms rows 16 b .. +15 by columns 128 / NCG
// starting at h 128 / NCG (16 rows: each wave a quarter of the columns,
// so that a workgroup's chain of products is half that of 32 rows and the
// panels, mostly under one round of workgroups, finish sooner).  The
// result is written back into G and packed into P[row][ldp].
static constexpr int PR = 16;   // rows a workgroup

__global__ __launch_bounds__(256) void k_chol_panel(double *__restrict__ A, long lda, int k, int nb, int n,
                                                    const double *__restrict__ Linv, double *__restrict__ P,
                                                    int ldp)
{
    __shared__ double Ar[PR][PBL];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr int NCG = 4 / (PR / 16);   // column groups: the waves split rows, then columns
    const int rb = wave / NCG, hc = wave % NCG;
    const int r0 = k + nb + blockIdx.x * PR;
    {
        constexpr int NL = PR * PB / 256;
        double v[NL];
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const int u = tid + 256 * i, r = u / PB, c = u % PB;
            v[i] = (r0 + r < n && c < nb) ? A[(long)(r0 + r) * lda + k + c] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const int u = tid + 256 * i;
            Ar[u / PB][u % PB] = v[i];
        }
    }
    __syncthreads();
    constexpr int NJ = PB / (16 * NCG);   // 16-column blocks a wave
    dbl4 acc[NJ];
#pragma unroll
    for (int jb = 0; jb < NJ; ++jb) acc[jb] = dbl4{0, 0, 0, 0};
    const int q = lane >> 4, p = lane & 15;
    // the Linv^T operands of k-step kk + 4 R are loaded while step kk is
    // multiplied (a ring of R register sets, the loop unrolled R times)
    constexpr int R = 4;
    double bv[R][NJ];
    auto fetch = [&](double (&d)[NJ], int kk) {
#pragma unroll
        for (int jb = 0; jb < NJ; ++jb)
            d[jb] = kk < nb ? Linv[(long)(16 * (NJ * hc + jb) + p) * PB + kk + q] : 0.0;   // Linv^T[kk+q][col]
    };
#pragma unroll
    for (int st = 0; st < R; ++st) fetch(bv[st], 4 * st);
    for (int kk0 = 0; kk0 < nb; kk0 += 4 * R) {
#pragma unroll
        for (int st = 0; st < R; ++st) {
            const int kk = kk0 + 4 * st;
            if (kk < nb) {
                const double a = Ar[16 * rb + p][kk + q];
#pragma unroll
                for (int jb = 0; jb < NJ; ++jb)
                    acc[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv[st][jb], acc[jb], 0, 0, 0);
                fetch(bv[st], kk + 4 * R);
            }
        }
    }
#pragma unroll
    for (int jb = 0; jb < NJ; ++jb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gr = r0 + 16 * rb + q + 4 * r;   // f64 MFMA result rows
            const int c = 16 * (NJ * hc + jb) + p;
            if (gr < n && c < nb) {
                A[(long)gr * lda + k + c] = acc[jb][r];
                P[(long)gr * ldp + c] = acc[jb][r];
            }
        }
}

// A22 -= L21 L21^T over lower 64 x 64 tiles of the trailing matrix (rows
// and columns >= base), all of them (nct = 0) or those of the first nct
// tile columns; L21 packed in P[row][ldp], nb columns used (one panel, or
// two panels' L21 side by side).  The
// panel goes through LDS in 32-column chunks, stored transposed with rows
// padded to 65 doubles (the transposing stores of one wave then fall in
// distinct banks); the next chunk's values are loaded into registers while
// the current chunk is multiplied.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_chol_syrk(
    double *__restrict__ A, long lda, const double *__restrict__ P, int ldp, int base, int nb, int n, int nct)
{
    __shared__ double Pi[32][65];
    __shared__ double Pj[32][65];
    const long t = blockIdx.x;
    int ti, tj;
    if (nct > 0) {   // the first nct tile columns only, every tile row
        ti = (int)(t / nct);
        tj = (int)(t % nct);
        if (ti < tj) return;
    } else {         // the whole lower triangle of tiles
        ti = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) / 2.0);
        while ((long)(ti + 1) * (ti + 2) / 2 <= t) ++ti;
        while ((long)ti * (ti + 1) / 2 > t) --ti;
        tj = (int)(t - (long)ti * (ti + 1) / 2);
    }
    const int r0 = base + ti * 64, c0 = base + tj * 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    dbl4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) acc[i][jj] = dbl4{0, 0, 0, 0};
    // thread tid stages (row tid / 32 + 8 q, column tid % 32) of a chunk, q < 8
    const int kk_t = tid & 31, r_t = tid >> 5;
    double vi[8], vj[8];
    auto fetch = [&](int k0) {
        const bool ok = k0 + kk_t < nb;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int r = r_t + 8 * q;
            vi[q] = (ok && r0 + r < n) ? P[(long)(r0 + r) * ldp + k0 + kk_t] : 0.0;
            vj[q] = (ok && c0 + r < n) ? P[(long)(c0 + r) * ldp + k0 + kk_t] : 0.0;
        }
    };
    fetch(0);
    // the tile of A, loaded now so that the subtraction at the end waits for
    // nothing (unconditional loads: outside the triangle, the tile's corner)
    double cv[2][2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gr = r0 + wm * 32 + i * 16 + (lane >> 4) + 4 * r;
                const int gc = c0 + wn * 32 + jj * 16 + (lane & 15);
                const bool in = gr < n && gc <= gr;
                cv[i][jj][r] = A[(long)(in ? gr : r0) * lda + (in ? gc : c0)];
            }
    for (int k0 = 0; k0 < nb; k0 += 32) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            Pi[kk_t][r_t + 8 * q] = vi[q];
            Pj[kk_t][r_t + 8 * q] = vj[q];
        }
        __syncthreads();
        if (k0 + 32 < nb) fetch(k0 + 32);
#pragma unroll
        for (int kk = 0; kk < 32; kk += 4) {
            const int kr = kk + (lane >> 4);
            double a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = Pi[kr][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) b[jj] = Pj[kr][wn * 32 + jj * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int jj = 0; jj < 2; ++jj)
                    acc[i][jj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[jj], acc[i][jj], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gr = r0 + wm * 32 + i * 16 + (lane >> 4) + 4 * r;
                const int gc = c0 + wn * 32 + jj * 16 + (lane & 15);
                if (gr < n && gc <= gr) A[(long)gr * lda + gc] = cv[i][jj][r] - acc[i][jj][r];
            }
}

// y_blk <- M y_blk for the block at k (nb rows), M = Linv or Linv^T given
// row-major (Mr = Linv, or LinvT for the transposed solve): one wave per
// row, lanes along the row, SW waves.  A wave takes its rows eight at a time
// with all their loads issued together, then reduces each (the same
// products and order as a row at a time).
static constexpr int SW = 16;   // waves of the solve kernels

__device__ void blk_solve(const double *__restrict__ Mr, int k, int nb, double *__restrict__ y,
                          double *__restrict__ v)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    if (tid < PB) v[tid] = tid < nb ? y[k + tid] : 0.0;
    __syncthreads();
    const double v0 = v[lane], v1 = v[lane + 64];
    for (int r0 = wave; r0 < nb; r0 += 8 * nw) {   // rows r0, r0 + nw, ..., r0 + 7 nw
        double a0[8], a1[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int r = min(r0 + nw * q, PB - 1);   // M has PB rows
            a0[q] = Mr[r * PB + lane];
            a1[q] = Mr[r * PB + lane + 64];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            double s = a0[q] * v0 + a1[q] * v1;
            for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
            if (lane == 0 && r0 + nw * q < nb) y[k + r0 + nw * q] = s;
        }
    }
}

__global__ __launch_bounds__(256) void k_chol_solve_first(const double *__restrict__ Mr, int k, int nb,
                                                          double *__restrict__ y)
{
    __shared__ double v[PB];
    blk_solve(Mr, k, nb, y, v);
}

// forward step of panel k (y_k already solved): rows i >= k+nb get
// y_i -= sum_m L[i][k+m] y[k+m], one wave per row, lanes along the row, a
// wave's rows eight at a time with their loads in flight together (SW
// waves: PB rows in one round).  Workgroup 0 owns the next diagonal block
// [k+nb, k+nb+nb2) and then solves it with its Linv; workgroup b > 0 takes
// the PB rows after those of workgroup b - 1.
__global__ __launch_bounds__(64 * SW) void k_chol_fwd(const double *__restrict__ A, long lda, int k, int nb, int n,
                                                      const double *__restrict__ Linv_next, int nb2,
                                                      double *__restrict__ y)
{
    __shared__ double v[PB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = k + nb + (int)blockIdx.x * PB;
    const double y0 = lane < nb ? y[k + lane] : 0.0, y1 = lane + 64 < nb ? y[k + 64 + lane] : 0.0;
    for (int rb = wave; rb < PB; rb += 8 * SW) {   // rows rb, rb + SW, ..., rb + 7 SW of the range
        double a0[8], a1[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const double *Li = A + (long)min(r0 + rb + SW * q, n - 1) * lda + k;
            a0[q] = lane < nb ? Li[lane] : 0.0;
            a1[q] = lane + 64 < nb ? Li[lane + 64] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            double s = a0[q] * y0 + a1[q] * y1;
            for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
            const int i = r0 + rb + SW * q;
            if (lane == 0 && i < n) y[i] -= s;
        }
    }
    if (blockIdx.x == 0 && nb2 > 0) {
        __threadfence_block();
        __syncthreads();
        blk_solve(Linv_next, k + nb, nb2, y, v);
    }
}

// backward step of panel k (u_k already solved): rows i < k get
// y_i -= sum_m L[k+m][i] u[k+m] (lanes along i: row reads coalesce).  A
// workgroup takes PB rows; the m range is split in eight parts (two waves
// a part), each summed with eight loads in flight, and the parts meet in LDS
// in a fixed order.  Workgroup 0 owns the previous block [k-PB, k) and then
// solves it with Linv^T (LinvT_prev, row-major); workgroup b > 0 takes the
// PB rows below those of workgroup b - 1.
__global__ __launch_bounds__(64 * SW) void k_chol_bwd(const double *__restrict__ A, long lda, int k, int nb,
                                                      const double *__restrict__ LinvT_prev, double *__restrict__ y)
{
    constexpr int NPART = 64 * SW / PB;
    __shared__ double v[PB];
    __shared__ double uk[PB];
    __shared__ double red[NPART][PB];
    const int tid = threadIdx.x, t = tid % PB, part = tid / PB;
    if (tid < PB) uk[tid] = tid < nb ? y[k + tid] : 0.0;
    __syncthreads();
    const int per = (nb + NPART - 1) / NPART, m0 = min(nb, part * per), m1 = min(nb, m0 + per);
    const int i = k - PB * (1 + (int)blockIdx.x) + t;
    double s0 = 0.0, s1 = 0.0;
    if (i >= 0) {
        for (int m = m0; m < m1; m += 8) {
            double a[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) a[j] = m + j < m1 ? A[(long)(k + m + j) * lda + i] : 0.0;
#pragma unroll
            for (int j = 0; j < 8; j += 2) {
                s0 += a[j] * uk[m + j < m1 ? m + j : 0];
                s1 += a[j + 1] * uk[m + j + 1 < m1 ? m + j + 1 : 0];
            }
        }
    }
    red[part][t] = s0 + s1;
    __syncthreads();
    if (part == 0 && i >= 0) {
        double s = red[0][t];
#pragma unroll
        for (int q = 1; q < NPART; ++q) s += red[q][t];
        y[i] -= s;
    }
    if (blockIdx.x == 0) {
        __threadfence_block();
        __syncthreads();
        blk_solve(LinvT_prev, k - PB, PB, y, v);
    }
}

__global__ void k_clear_info(int *info) { *info = 0; }

// Panels in groups of NGD_GROUP: a panel of a group updates only the block
// columns of the group's later panels; the last panel of the group updates
// the trailing matrix beyond the group with every panel's L21 of the group
// at once (an update over NGD_GROUP x 128 columns): fewer passes over the
// trailing triangle, more products per tile of it.  Same operations on
// every element as one update a panel, summed in another order.  Where the
// matrix ends inside a group, each of its panels updates all that is left.
static constexpr int NGD_GROUP = 4;

long ngd_work_doubles(int n)
{
    const long np = (n + PB - 1) / PB;
    return 2 * np * PB * PB + (long)NGD_GROUP * n * PB;
}

int ngd_factor_solve(double *G, int n, long lda, double nb, double mu, int mode, double *b, double *trace,
                     int *info, double *work, hipStream_t s)
{
    const int np = (n + PB - 1) / PB;
    constexpr int LDP = NGD_GROUP * PB;
    double *Linv = work, *LinvT = work + (long)np * PB * PB, *P = work + 2 * (long)np * PB * PB;
    hipLaunchKernelGGL(k_clear_info, dim3(1), dim3(1), 0, s, info);
    hipLaunchKernelGGL(k_trace, dim3(1), dim3(256), 0, s, G, n, lda, trace);
    hipLaunchKernelGGL(k_damp, dim3((unsigned)((n + 255) / 256), (unsigned)n), dim3(256), 0, s, G, n, lda, nb, mu,
                       mode, trace, b);
    for (int p = 0; p < np; ++p) {
        const int k = p * PB, nb_ = std::min(PB, n - k), q = p % NGD_GROUP;
        double *Li = Linv + (long)p * PB * PB;
        hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(256), 0, s, G, lda, k, nb_, Li, info);
        const int rest = n - k - nb_;
        if (rest <= 0) break;
        double *Pq = P + q * PB;
        hipLaunchKernelGGL(k_chol_panel, dim3((unsigned)((rest + PR - 1) / PR)), dim3(256), 0, s, G, lda, k, nb_, n,
                           (const double *)Li, Pq, LDP);
        const long T = (rest + 63) / 64;
        const int later = (NGD_GROUP - 1 - q) * PB;   // columns of the group's later panels
        if (rest <= later)        // the matrix ends inside the group: all that is left, this panel
            hipLaunchKernelGGL(k_chol_syrk, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, G, lda,
                               (const double *)Pq, LDP, k + nb_, nb_, n, 0);
        else if (later > 0)       // the group's later block columns only
            hipLaunchKernelGGL(k_chol_syrk, dim3((unsigned)(T * (later / 64))), dim3(256), 0, s, G, lda,
                               (const double *)Pq, LDP, k + nb_, nb_, n, later / 64);
        else                      // beyond the group, with every panel of it
            hipLaunchKernelGGL(k_chol_syrk, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, G, lda,
                               (const double *)P, LDP, k + nb_, q * PB + nb_, n, 0);
    }
    hipLaunchKernelGGL(k_linv_transpose, dim3(16, (unsigned)np), dim3(256), 0, s, (const double *)Linv, LinvT);
    // L y = b, then L^T u = y (in place in b)
    hipLaunchKernelGGL(k_chol_solve_first, dim3(1), dim3(256), 0, s, (const double *)Linv, 0, std::min(PB, n), b);
    for (int p = 0; p + 1 < np; ++p) {
        const int k = p * PB, nb_ = PB, rest = n - k - nb_;
        const int nb2 = std::min(PB, rest);
        hipLaunchKernelGGL(k_chol_fwd, dim3((unsigned)((rest + PB - 1) / PB)), dim3(64 * SW), 0, s, (const double *)G,
                           lda, k, nb_, n, (const double *)(Linv + (long)(p + 1) * PB * PB), nb2, b);
    }
    const int kl = (np - 1) * PB;
    hipLaunchKernelGGL(k_chol_solve_first, dim3(1), dim3(256), 0, s, (const double *)(LinvT + (long)(np - 1) * PB * PB),
                       kl, n - kl, b);
    for (int p = np - 1; p >= 1; --p) {
        const int k = p * PB, nb_ = std::min(PB, n - k);
        hipLaunchKernelGGL(k_chol_bwd, dim3((unsigned)((k + PB - 1) / PB)), dim3(64 * SW), 0, s, (const double *)G,
                           lda, k, nb_, (const double *)(LinvT + (long)(p - 1) * PB * PB), b);
    }
    return 0;
}


// ---------------------------------------------------------------------------
// Dual (Gram) route, app/optimizer_module.f90:408-446: with G = J^T J / N_b
// and c the damping, (G + cI)^{-1} b1 = (b1 - J^T K^{-1} (J b1) / N_b) / c,
// K = J J^T / N_b + cI (nrow x nrow).  The rows arrive transposed,
// R[w][n] = J[n][w]; J is formed row-contiguous for the Gram product.

// Jm[n][w] = R[w][n] (n < nrow, w < nw), zero for nw <= w < ldj, widened
// to FP64 (exactly); 32 x 32 tiles through LDS so both sides are coalesced
__global__ __launch_bounds__(256) void k_rows_transpose(const float *__restrict__ R, long ldr, int nw, int nrow,
                                                        double *__restrict__ Jm, long ldj)
{
    __shared__ float t[32][33];
    const int w0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int w = w0 + r, n = n0 + tx;
        t[r][tx] = (w < nw && n < nrow) ? R[(long)w * ldr + n] : 0.f;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int n = n0 + r, w = w0 + tx;
        if (n < nrow && w < ldj) Jm[(long)n * ldj + w] = (double)t[tx][r];
    }
}

// out[r] = sum_c M[r][c] x[c] for the matrix M[nr][ld] (c < nc), FP64
// sums; one wave per row, lanes along the row
template <typename T>
__global__ __launch_bounds__(256) void k_rows_matvec(const T *__restrict__ M, long ld, int nr, int nc,
                                                     const double *__restrict__ x, double *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= nr) return;
    const T *row = M + (long)r * ld;
    double s = 0.0;
    for (int c = lane; c < nc; c += 64) s += (double)row[c] * x[c];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) out[r] = s;
}

// u[w] = (b[w] / nb - (R y)[w] / nb) / c with c = mu (ABS) or mu tr / nb
// (TRACE), tr the raw trace left by ngd_factor_solve; (R y) precomputed
__global__ void k_dual_combine(int nw, const double *__restrict__ b, const double *__restrict__ ry, double nb,
                               double mu, int mode, const double *__restrict__ trace, double *__restrict__ u)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const double c = mode == 1 ? mu : mu * (*trace / nb);
    u[w] = (b[w] / nb - ry[w] / nb) / c;
}

void launch_rows_transpose(const float *R, long ldr, int nw, int nrow, double *Jm, long ldj, hipStream_t s)
{
    dim3 grid((unsigned)((ldj + 31) / 32), (unsigned)((nrow + 31) / 32));
    hipLaunchKernelGGL(k_rows_transpose, grid, dim3(256), 0, s, R, ldr, nw, nrow, Jm, ldj);
}

template <typename T>
void launch_rows_matvec(const T *M, long ld, int nr, int nc, const double *x, double *out, hipStream_t s)
{
    hipLaunchKernelGGL((k_rows_matvec<T>), dim3((unsigned)((nr + 3) / 4)), dim3(256), 0, s, M, ld, nr, nc, x, out);
}
template void launch_rows_matvec<float>(const float *, long, int, int, const double *, double *, hipStream_t);
template void launch_rows_matvec<double>(const double *, long, int, int, const double *, double *, hipStream_t);

void launch_dual_combine(int nw, const double *b, const double *ry, double nb, double mu, int mode,
                         const double *trace, double *u, hipStream_t s)
{
    hipLaunchKernelGGL(k_dual_combine, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, s, nw, b, ry, nb, mu, mode,
                       trace, u);
}

}  // namespace dnnf90
