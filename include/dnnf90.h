/* dnnf90.h -- C interface of the MI355X-native derivative-carrying engine.
 *
 * Drop-in for the reference's lib/dnnf90.h (24 functions, same names,
 * signatures, argument meaning and return codes; implemented there in
 * lib/c_api_module.f90:100-528).  Every evaluation and gradient call runs on
 * the GPU (gfx950 HIP kernels); nothing falls back to the CPU.
 *
 * Conventions (lib/c_api_module.f90:17-24, 38, 44-62, 72-96):
 *   handles are > 0; 0 is success for routines without a handle result;
 *   -1 invalid argument, invalid or stale handle (a network slot reused
 *      after dnnf90_net_free is refused to its old dependents);
 *   -2 missing file or unknown activation code;
 *   -3 (this library) a HIP runtime failure; -4 (this library) device
 *      memory exhausted;
 *   arrays are caller-owned host arrays; at most 64 handles per registry;
 *   one work handle (wid / tid) and one accumulator (gid) per host thread
 *   make concurrent calls safe (each work handle owns its own HIP stream);
 *   create and free handles outside parallel regions.
 */
#ifndef DNNF90_H
#define DNNF90_H
#ifdef __cplusplus
extern "C" {
#endif

/* shared tables: build the "current" family (replaces
 * dnnf90_tables_init_dense/closure, lib/c_api_module.f90:100-125) */
int dnnf90_tables_init_dense(int d0, int k);
/* seeds: d0*nseed ints, seed s at seeds[s*d0 .. s*d0+d0) */
int dnnf90_tables_init_closure(int d0, int k, int nseed, const int *seeds);
/* carried slot count of the current family (lib/c_api_module.f90:128-130) */
int dnnf90_nderiv(void);
/* multi-index of slot ia (1-based) into a[0..d0) (lib/c_api_module.f90:133-141) */
int dnnf90_alpha(int ia, int *a);

/* independent table sets (lib/c_api_module.f90:407-474); building one also
 * makes it the current family, as there */
int dnnf90_tabset_dense(int d0, int k);
int dnnf90_tabset_closure(int d0, int k, int nseed, const int *seeds);
int dnnf90_tabset_nderiv(int tid);
int dnnf90_net_create_ts(int tid, int nlayer, const int *dims);

/* nn_weight.dat reader on the current family (lib/c_api_module.f90:145-166) */
int dnnf90_net_load(const char *filename);
/* activation of a table set: 0 TANH, 1 SIN, 2 ERF, 3 BESSEL (J_0),
 * 4 BESSEL1 (J_1) as in the reference (lib/c_api_module.f90:480-493), and
 * 5 erf-GELU x*Phi(x), which the reference lacks */
int dnnf90_tabset_activation(int tid, int iact);
int dnnf90_net_load_ts(int tid, const char *filename);
/* empty network on the current family (lib/c_api_module.f90:169-183) */
int dnnf90_net_create(int nlayer, const int *dims);
/* layer l (2..nlayer): ndim(l) rows of ndim(l-1)+1 doubles, bias first
 * (lib/c_api_module.f90:189-205) */
int dnnf90_net_set_layer(int nid, int l, const double *wrow);
int dnnf90_net_free(int nid);

/* evaluation: t[0..nderiv) receives the value and every carried mixed
 * derivative of output 1 at x[0..d0) (lib/c_api_module.f90:221-278) */
int dnnf90_work_create(int nid);
int dnnf90_work_free(int wid);
int dnnf90_eval(int nid, int wid, const double *x, double *t);

/* training (lib/c_api_module.f90:282-400): forward, adjoint of the seed
 * dL/dT_alpha (seed[0..nderiv), output 1) and accumulation into gid */
int dnnf90_twork_create(int nid);
int dnnf90_twork_free(int tid);
int dnnf90_grad_create(int nid);
int dnnf90_grad_free(int gid);
int dnnf90_grad_zero(int gid);
int dnnf90_grad_point(int nid, int tid, const double *x,
                      const double *seed, int gid);
/* Adam with the reference's formulas and bias correction
 * (lib/train_module.f90:656-683) */
int dnnf90_adam_step(int nid, int gid, double eta, double beta1,
                     double beta2, double eps, int nbatch, int istep);

#ifdef __cplusplus
}
#endif
#endif
