// kernels.hpp -- launch wrappers for the gfx950 kernels of the derivative-
// carrying forward pass and its adjoint.
//
// Device layout (all layers): a layer's plane is row-major
// [neuron][slot * Np + point], Np = padded point count, so the whole
// multi-index axis and the point axis form one contiguous column axis of
// length na*Np.  The linear maps are then single GEMMs over na*Np columns,
// and one lane per (point, neuron) reads its slots coalesced across points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dnnf90 {

enum Precision : int {
    PREC_FP64 = 0,    // everything in FP64 (drop-in parity mode)
    PREC_FP32 = 1,    // FP32 storage, FP32 MFMA linear maps, FP32 recursion
    PREC_BF16X3 = 2,  // FP32 storage, BF16 hi+lo split MFMA linear maps
    PREC_BF16X6 = 3,  // FP32 storage, BF16 three-way split MFMA linear maps
};

// one term of a table-driven sum, 16 bytes so a wave-uniform term is a
// single scalar load: operand entries (or slots) s, b and the coefficient
struct alignas(16) TermRec64 {
    int s, b;
    double c;
};
struct alignas(16) TermRec32 {
    int s, b;
    float c;
    int pad;
};

// forward schedule on the device (see FwdSchedule in tables.hpp)
struct DevFwdSched {
    const int *slot_group;   // [na+1]
    const int *grp_term;     // [ngroup+1]
    const int *grp_store;    // [ngroup]
    const TermRec64 *t64;    // [F] (entry of S_b, entry of B(a-b,q-1), C(a,b)/q)
    const TermRec32 *t32;    // [F]
    int na, k, nentry;
};

struct DevPairs {
    const int *start, *num;  // [na]
    const TermRec64 *p64;    // [P] (slot a, slot a-b, C(a,b))
    const TermRec32 *p32;    // [P]
    int na;
};

struct ActParams {
    int iact, nq;            // nq = k+1 derivatives beyond sigma itself
    const double *tpc;       // [(k+2)][(k+3)]
    int ldt;                 // k+3
};

// LDS bytes per thread needed by the forward/adjoint kernels (0 = global path)
int bell_fwd_threads(int nentry, int elem_bytes);
int bell_bwd_threads(int na, int elem_bytes);

template <typename T>
void launch_input(const double *x, int npt, int d0, int na, const int *ind_e1, T *T0,
                  long Np, hipStream_t s);

template <typename T>
void launch_bell_fwd(const T *S, T *Tout, T *H, int nrow, long Np, const DevFwdSched &sch,
                     const ActParams &act, T *scratch, hipStream_t s);

template <typename T>
void launch_bell_bwd(const T *Tb, const T *H, T *Sb, int nrow, long Np, const DevPairs &pr,
                     hipStream_t s);

// C[M][ncol] = A[M][Kd] * Bext[Kd][ncol]
//   trans_a == false: A = W[M][Kd] (ldw = Kd, bias first); Bext row 0 is the
//                     bias channel (1 on the first Np columns), rows 1.. = B
//   trans_a == true : A[m][k] = W[k][m+1] (ldw = M+1); Bext = B
template <typename T>
void launch_gemm_wx(int prec, bool trans_a, const T *W, int ldw, const T *B, T *C, int M, int Kd,
                    long ncol, long Np, hipStream_t s);

// partial[chunk][M][Kd] = sum over the chunk's columns of A[m][c] * Bext[k][c]
// (Bext row 0 = bias channel, rows 1.. = B)
template <typename T>
void launch_gemm_grad(int prec, const T *A, const T *B, int M, int Kd, long ncol, long Np,
                      double *partial, int nchunk, long chunk_cols, hipStream_t s);
// grad[len] += sum over chunks of partial[chunk][len], in a fixed order
void launch_grad_reduce(double *partial, int nchunk, long len, double *grad, hipStream_t s);
// up to RS_MAX such reductions in one launch (set: out = the sum instead of
// out += it); each segment's bits are launch_grad_reduce's, but for
// segments of 1024 chunks or more (64 slices instead of 16)
constexpr int RS_MAX = 16;
struct ReduceSegs {
    int n = 0;
    const double *partial[RS_MAX];
    double *out[RS_MAX];
    long len[RS_MAX];
    int nchunk[RS_MAX];
    int set[RS_MAX];
    void add(const double *p, int nc, long l, double *o, bool st = false)
    {
        partial[n] = p;
        nchunk[n] = nc;
        len[n] = l;
        out[n] = o;
        set[n] = st ? 1 : 0;
        ++n;
    }
};
void launch_reduce_segs(const ReduceSegs &r, hipStream_t s);

// K1 (trans_a false) / K6 on BF16 three-way split products (gemm_x6.hip);
// arguments as launch_gemm_wx
void launch_lin_x6(bool trans_a, const float *W, int ldw, const float *B, float *C, int M, int Kd, long ncol,
                   long Np, hipStream_t s);
// fixed-chunk split linear maps and the one-block split K5 on (1) or off
// (0), or the setting queried (-1); returns the previous setting
int set_lin_x6c(int on);
// column chunks the one-block split K5 takes for an M x Kd weight block
// (one workgroup per chunk and 128 x 128 tile), 0 where it does not apply
// (launch_grad_x6 then takes the tiled kernel, whose chunks grad_chunks
// gives); any: whatever the split-kernel switch says
int grad_x6s_wgs(int M, int Kd, bool any = false);
// K5 chunk partials on BF16 three-way split products (gemm_x6.hip);
// partial[chunk][M][Kd] as launch_gemm_grad, without the final reduction
void launch_grad_x6(const float *A, const float *B, int M, int Kd, long ncol, long Np, double *partial,
                    int nchunk, long chunk, hipStream_t s);

// K5 and K6 of one layer in one kernel on BF16 three-way split products:
// K5 chunk partials as launch_grad_x6 forms them (over its own chunks) and
// Tb[Kd-1][ncol] = W' A as launch_lin_x6(true, ...) gives it (the same
// bits), S-bar read once.
// grad_lint_x6_ok: whether it takes an M x Kd weight block (the switch on,
// M 32 or 64, Kd - 1 <= 64); chunks as grad_x6s_wgs gives them.
// set_grad_lint_x6: on (1) / off (0) / query (-1), returns the previous
bool grad_lint_x6_ok(int M, int Kd);
bool launch_grad_lint_x6(const float *A, const float *B, const float *W, float *Tb, int M, int Kd, long ncol, long Np,
                         double *partial, int nchunk, long chunk, hipStream_t s);
int set_grad_lint_x6(int on);
// column chunks the fused kernel takes for an M x Kd block (0: not
// applicable); any: whatever the switches say (work-space sizing)
int grad_lint_x6_wgs(int M, int Kd, bool any = false);

// column chunks of launch_gemm_grad for an M x Kd weight block over ncol
// columns of esz-byte elements (chunk: a multiple of 128 columns)
void grad_chunks(long ncol, int M, int Kd, int esz, int *nchunk, long *chunk);

// partial buffer length needed by launch_gemm_grad / launch_gram:
// (nchunk + grad_reduce_splits(nchunk)) * M * Kd doubles
__host__ __device__ inline int grad_reduce_splits(int nchunk) { return nchunk >= 64 ? 16 : (nchunk >= 16 ? 4 : 1); }

// NGD: G <- G/nb + cI, b <- b/nb, then G = L L^T (FP64 blocked Cholesky, in
// place, lower) and b <- (L L^T)^{-1} b; *info > 0 flags a non-positive
// pivot (1-based column); trace is one device double of scratch
int ngd_factor_solve(double *G, int n, long lda, double nb, double mu, int mode, double *b, double *trace,
                     int *info, double *work, hipStream_t s);
// NGD dual route helpers (ngd.hip): Jm[n][w] = R[w][n] widened to FP64 with
// zero columns up to ldj; out = M x with FP64 sums; u = (b - R y)/nb / c
void launch_rows_transpose(const float *R, long ldr, int nw, int nrow, double *Jm, long ldj, hipStream_t s);
template <typename T>
void launch_rows_matvec(const T *M, long ld, int nr, int nc, const double *x, double *out, hipStream_t s);
// G = R R^T over FP64 rows R[n][ld] on FP64 MFMA, tiles holding
// lower-triangle entries (the dual route's K, where the 1/c of the step
// amplifies any error of K)
void launch_gram64(const double *R, int n, long ld, double *G, hipStream_t s);
void launch_dual_combine(int nw, const double *b, const double *ry, double nb, double mu, int mode,
                         const double *trace, double *u, hipStream_t s);
// doubles of device work space ngd_factor_solve needs for order n (the
// inverses of the diagonal blocks, their transposes and a group of packed
// panels)
long ngd_work_doubles(int n);

// G[i][j] (+)= sum_n R[i][n] R[j][n] for the transposed FP32 rows R[nw][ld],
// on every tile that holds lower-triangle entries
// (FP32 MFMA, FP32 accumulation per chunk, FP64 reduction); with one chunk
// G is overwritten, otherwise added to
void launch_gram(const float *R, int nw, long ld, double *partial, int nchunk, long chunk_cols, double *G,
                 hipStream_t s);

// per-point rows: rows[(j*Kd + k)][n] = sum_a A[j][a*Np+n] * Bext[k][a*Np+n]
// K5 of a single point (column 0 of every slot), added to grad[M][Kd] in FP64
template <typename T>
void launch_grad_point1(const T *A, const T *B, int M, int Kd, int na, long Np, double *grad, hipStream_t s);
template <typename T>
void launch_point_rows(const T *A, const T *B, int M, int Kd, int na, long Np, float *rowsT,
                       long ldr, hipStream_t s);

// value-only path (bell.hip, misc.hip): Z = sigma(S), Zd = sigma'(S);
// Sbar = Tbar * Zd; D = coef (Z - y) on npt points, 0 on the padding
template <typename T>
void launch_act_value(const T *S, T *Z, T *Zd, long n, const ActParams &ap, hipStream_t s);
template <typename T>
void launch_mul_value(const T *Tb, const T *Zd, T *Sb, long n, hipStream_t s);
template <typename T>
void launch_seed_value(const T *Z, const double *y, double coef, int npt, long Np, T *D, hipStream_t s);

template <typename T>
void launch_seed_user(const double *seed, int npt, int nout, int na, T *Tb, long Np, hipStream_t s);

struct DevResidual {
    int nterm;
    const double *coeff;
    const int *slot;
    const int *nonlin;
    int distinct = 0;   // 1: the terms' slots are pairwise distinct and none is slot 0
};

// R_n = [r0_n +] sum_k c_k [T_0 *] T_slot (output 0); seed = fac*R*dR/dT on
// output 0 (unit_r: seed = fac*dR/dT, the Gauss-Newton row seed);
// rsq_part[block] = sum of R^2 over the block's points; rows > 0 zeroed;
// r0 (optional) starts R_n, r_out (optional) receives R_n
template <typename T>
void launch_seed_residual(const T *TL, int nout, int na, int npt, long Np, const DevResidual &r,
                          double fac, bool unit_r, T *Tb, double *rsq_part, hipStream_t s,
                          const double *r0 = nullptr, double *r_out = nullptr);

template <typename T>
void launch_output(const T *TL, int nout, int na, int npt, long Np, double *tout, hipStream_t s);

void launch_convert_f64_f32(const double *src, float *dst, long n, hipStream_t s);
void launch_reduce_sum(const double *part, int n, double *out, hipStream_t s);
void launch_adam(double *w, double *m, double *v, const double *g, long n, double eta,
                 double beta1, double beta2, double eps, double nbatch, double bc1,
                 double bc2, hipStream_t s);
void launch_fill(double *p, long n, double v, hipStream_t s);
void launch_axpy(double *y, const double *x, long n, double a, hipStream_t s);

// ---------------------------------------------------------------------------
// extended Kalman filter (kalman.hip)

// filter scalars, resident on the device so that a sequence of observations
// runs without a host round trip
struct KfScalars {
    double lam, lam0, gate, q;
    double lam_dd;   // lambda of the downdate still owed to P
    double xi;       // innovation of the current observation, scaled
    double denom;    // lambda + j.P j of the current observation
    double gs;       // gate scale of the current observation
    int pending;     // P owes the downdate (kv, pj, lam_dd)
    int ngated;
    int gated;       // the current observation tripped the gate
    int bad;         // a nonpositive denominator was met
};

// how k_kf_gain forms the innovation xi_in before xi = scale * xi_in
enum {
    KF_XI_GIVEN = 0,    // xi_in = v
    KF_XI_LINEAR = 1,   // xi_in = v - seed.T (target minus the linear prediction)
    KF_XI_RESID = 2,    // xi_in = v - R, R from the device
    KF_XI_IEKF = 3,     // xi = -scale v - j.(wprior - w)
};

void launch_kf_init_tiles(double *P, int nt, int nw, double p0, hipStream_t s);
void launch_kf_init_blocks(double *P, const long *poff, const int *bsz, int nb, double p0, hipStream_t s);
// dense pass: pending downdate, then (want) the shares of P (scale g) in part[nt][nt*64]
void launch_kf_tiles(double *P, int nt, int nw, const double *g, double scale, const double *kv,
                     const double *pj, const KfScalars *st, bool want, double *part, hipStream_t s);
void launch_kf_sum_part(const double *part, int nt, double *pj, hipStream_t s);
// decoupled pass: pending downdate and (want) pj = P (scale g), per block
void launch_kf_blocks(double *P, const long *poff, const int *bsz, const int *boff, int nb, int maxb,
                      const double *g, double scale, const double *kv, double *pj, const KfScalars *st, bool want,
                      hipStream_t s);
void launch_kf_gain(KfScalars *st, const double *g, double scale, const double *pj, int nw, int xmode, double v,
                    const double *seed, const double *T, int nsl, const double *rv, const double *wprior,
                    const double *w, bool final, hipStream_t s);
void launch_kf_apply(const KfScalars *st, double *pj, double *kv, double *w, const double *wprior, int nw,
                     bool iekf, hipStream_t s);
void launch_kf_clear(KfScalars *st, hipStream_t s);

// committee: member k (1-based) folded into mean and m2; then m2 <- std
void launch_comm_welford(const double *t, double *mean, double *m2, long n, int k, hipStream_t s);
void launch_comm_finish(double *m2, long n, int nmem, hipStream_t s);

// launch-time kernel attributes, shared by every host thread (thread-safe):
// dynamic LDS above 64 KiB must be requested per kernel; the limit of fn is
// raised to lds bytes and never lowered
void lds_allow(const void *fn, size_t lds);
// resident workgroups of `threads` threads a CU for fn with lds bytes of
// dynamic LDS (at least 1), queried once per (kernel, lds) and cached; raises
// the LDS limit first
int occupancy_cached(const void *fn, size_t lds, int threads = 256);

}  // namespace dnnf90
