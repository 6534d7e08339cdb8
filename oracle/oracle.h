/* oracle.h -- CPU FP64 restatement of DNNF90's derivative-carrying forward
 * pass, its analytical adjoint, the collocation seed and the NGD solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against (and the CPU baseline timed beside it).  Nothing in the product
 * (dnnf90_amd/) includes, links or calls it.
 *
 * Conventions (all 0-based here; the reference is 1-based):
 *   slot 0 is the zero multi-index (the value itself);
 *   a network of nlayer layers has dims[0..nlayer-1]; layer l >= 1 owns
 *   dims[l] rows of (dims[l-1]+1) doubles, bias first, exactly one "#l="
 *   block of nn_weight.dat (lib/c_api_module.f90:185-203); the weight and
 *   gradient vectors concatenate those blocks for l = 1..nlayer-1;
 *   points are row-major [npt][d0], outputs and seeds [npt][nout][na].
 */
#ifndef DNNF90_ORACLE_H
#define DNNF90_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* activation codes: 0 tanh, 1 sin, 2 erf (reference codes), 5 erf-GELU
 * (x * Phi(x); no counterpart in the reference) */
int ot_tables_build(int d0, int k, int nseed, const int *seeds);
int ot_tables_free(int h);
int ot_set_activation(int h, int iact);
int ot_na(int h);
int ot_nfq(int h);
int ot_npair(int h);
int ot_alpha(int h, int *alpha_out);                 /* [na][d0] */
int ot_deg(int h, int *deg_out);                     /* [na] */
int ot_ind_e1(int h, int *e1_out);                   /* [d0], -1 if absent */
int ot_pairs(int h, int *start, int *num, int *ia, int *id, double *binom);
int ot_fq(int h, int *start, int *num, int *ib, int *id, double *c); /* start/num [na][k+1] */
int ot_tpc(int h, double *tpc);                      /* [(k+2)][(k+3)] */
void ot_act_derivs(int h, double a, int nq, double *d);

int ot_nweights(int nlayer, const int *dims);

int ot_eval(int h, int nlayer, const int *dims, const double *w, int npt,
            const double *x, double *tout, int nthreads);
/* forward + adjoint of the given seeds, gradient ADDED to grad[nw] */
int ot_grad(int h, int nlayer, const int *dims, const double *w, int npt,
            const double *x, const double *seed, double *grad, int nthreads);
/* per-point gradient rows, rows[npt][nw] (overwritten) */
int ot_grad_rows(int h, int nlayer, const int *dims, const double *w, int npt,
                 const double *x, const double *seed, double *rows, int nthreads);
/* collocation step on output 0: R = sum_k c_k [T_0 *] T_slot_k;
 * seed = fac*R*dR/dT (app/pinn_module.f90:32-65); gradient ADDED to grad;
 * *rsq_out = sum_n R_n^2; tout (optional) receives the forward outputs */
int ot_pinn_step(int h, int nlayer, const int *dims, const double *w, int npt,
                 const double *x, int nterm, const double *coeff,
                 const int *slot, const int *nonlin, double fac,
                 double *grad, double *rsq_out, double *tout, int nthreads);
/* Gauss-Newton rows dR/dw of the same residual, rows[npt][nw] */
int ot_pinn_rows(int h, int nlayer, const int *dims, const double *w, int npt,
                 const double *x, int nterm, const double *coeff,
                 const int *slot, const int *nonlin, double *rows, int nthreads);
/* NGD primal solve: u = (J^T J/nb + c I)^{-1} (b/nb), c = mu (mode 1, ABS)
 * or mu*tr(J^T J/nb) (mode 0, TRACE); Gaussian elimination with partial
 * pivoting.  Returns 0, or -3 when the metric is numerically singular. */
int ot_ngd_solve(int nrow, int nw, const double *rows, const double *b,
                 int nb, double mu, int mode, double *u);

/* committee mean and sample standard deviation of every output slot over
 * nmem members w[nmem][nw] (lib/committee_module.f90:79-111, Welford) */
int ot_comm_eval(int h, int nlayer, const int *dims, int nmem, const double *w,
                 int npt, const double *x, double *tmean, double *tstd);

/* extended Kalman filter (lib/kalman_module.f90): create returns a handle
 * (or -1 on bad arguments); update_grad / iekf_iter move w[nw] in place and
 * return -2 on a nonpositive innovation denominator */
int ot_kf_create(int nlayer, const int *dims, double p0, double lam_ini, double lam0, int decoupled);
int ot_kf_set(int h, double gate, double q);
int ot_kf_free(int h);
int ot_kf_update_grad(int h, double *w, const double *g, double xi, double scale);
int ot_kf_iekf_begin(int h, const double *w);
int ot_kf_iekf_iter(int h, double *w, const double *g, double resid, double rnoise, int final);
long ot_kf_cov_len(int h);
int ot_kf_state(int h, double *P, double *lam, int *ngated);

#ifdef __cplusplus
}
#endif
#endif
