/* dnnf90_amd.h -- batched and device-pointer entries of the MI355X engine.
 *
 * The reference ABI (dnnf90.h) is per point and has no gradient read-out,
 * no split forward/backward and no multi-output entry (lib/dnnf90.h:40-54).
 * These entries add them for GPU throughput; every one has the same return
 * conventions as dnnf90.h.  Arrays are host arrays unless `flags` has
 * DNNF90_DEVICE_PTRS, in which case every array argument is a device
 * pointer (for example a torch tensor's data pointer) on the current device.
 *
 * Layouts: points x[npt][d0]; outputs and seeds [npt][nout][nderiv];
 * weights and gradients concatenate, for l = 2..nlayer, ndim(l) rows of
 * ndim(l-1)+1 doubles, bias first (one "#l=" block of nn_weight.dat).
 */
#ifndef DNNF90_AMD_H
#define DNNF90_AMD_H
#include "dnnf90.h"
#ifdef __cplusplus
extern "C" {
#endif

#define DNNF90_DEVICE_PTRS 1
#define DNNF90_NGD_DUAL 2     /* dnnf90_ngd_solve: Gram (dual) route */

/* arithmetic of a batched work space */
#define DNNF90_PREC_FP64 0   /* everything FP64: reference parity mode */
#define DNNF90_PREC_FP32 1   /* FP32 storage and recursion, FP32 MFMA; the
                                forward linear map accumulates in FP64 */
#define DNNF90_PREC_BF16X3 2 /* FP32 storage, BF16 hi+lo split MFMA (about
                                16 significant bits per operand) */
#define DNNF90_PREC_BF16X6 3 /* FP32 storage and recursion, linear maps on
                                BF16 three-way split MFMA (six products);
                                the forward map accumulates in FP64 */

/* message of the last failure on the calling thread ("" if none) */
const char *dnnf90_last_error(void);

/* table-set introspection (host only, no GPU needed): counts[0..3] =
 * nderiv, number of adjoint pairs P, number of forward terms F, k; then the
 * lists of lib/multi_index_bell_module.f90:267-332 with 0-based slots:
 * pair_start/num [nderiv], pair_ia/id/binom [P], fq_start/num
 * [nderiv][k+1], fq_ib/id/c [F] (any pointer may be NULL) */
int dnnf90_tabset_counts(int tid, int *counts);
int dnnf90_tabset_lists(int tid, int *pair_start, int *pair_num, int *pair_ia,
                        int *pair_id, double *pair_binom, int *fq_start,
                        int *fq_num, int *fq_ib, int *fq_id, double *fq_c);

/* Release a table-set handle.  No counterpart in the reference: its table
 * sets (lib/c_api_module.f90:61, rtab(MAXH)) live until the process ends, so a
 * host may create at most 64.  Networks built from the set keep their own copy
 * of the tables and stay valid.  Returns 0, or -1 for a bad handle. */
int dnnf90_tabset_free(int tid);

/* the slot-order file hod_alpha_order.dat of the family
 * (lib/multi_index_bell_module.f90:451-471, same text); -2 if it cannot be
 * written */
int dnnf90_tabset_write_order(int tid, const char *filename);

/* Faa di Bruno kernels generated per table family at run time (hipRTC,
 * straight-line, operands in registers) are used where the family is small
 * enough; otherwise, or with codegen off, the table-driven kernels run.
 * dnnf90_set_codegen switches this process-wide (also DNNF90_CODEGEN=0)
 * and returns the previous setting.  With the generated kernels each hidden
 * layer keeps its pre-activation planes S and the adjoint recomputes the
 * activation derivative planes h from them in registers, so the forward pass
 * writes one plane less per hidden layer (results identical bit for bit);
 * dnnf90_set_store_s(0) keeps h instead (the forward writes T and h), and
 * returns the previous setting.  A family whose kept-S kernels would spill
 * keeps h.  dnnf90_net_kernel_info fills info[6]: generated kernels active
 * (0 no, 1 keeping h, 2 keeping S), the VGPR counts and scratch bytes of
 * the forward and adjoint kernels in use, and the operand entries per lane
 * of the family. */
int dnnf90_set_codegen(int on);
int dnnf90_set_store_s(int on);
/* a work space holding one point (the per-point ABI, the Kalman filter)
 * runs the whole forward, and the whole adjoint with the weight gradient,
 * as one generated kernel each when the network is small (FP64, at most 16
 * layers, every width at most 256 and width x nderiv at most 1536, a table
 * family of at most 48 slots); dnnf90_set_point_kernels(0) uses the layer
 * kernels instead.  Returns the previous setting.  No counterpart in the
 * reference (its per-point path is the only one it has). */
int dnnf90_set_point_kernels(int on);
/* the BF16x6 linear maps whose reduction is whole 32-row chunks (K1 and K6
 * of layers 32, 64 or 128 wide) take kernels with the chunk count fixed at
 * compile time; dnnf90_set_split_kernels(0) takes the general kernel for
 * every map (results identical bit for bit).  Returns the previous setting.
 * No counterpart in the reference. */
int dnnf90_set_split_kernels(int on);
/* with the generated kernels and S kept: the first hidden layer forms its
 * S from the input plane and its weights in the Faa di Bruno kernels (no
 * linear map, no S plane), and the last hidden layer forms its Tbar from
 * the output layer's Sbar (no transposed map of the output layer);
 * dnnf90_set_edge_kernels(0) runs those maps as separate kernels (results
 * agree to rounding).  Returns the previous setting.  No counterpart in the
 * reference. */
int dnnf90_set_edge_kernels(int on);
int dnnf90_net_kernel_info(int nid, int prec, int *info);
/* the fused kernels a batched step of many points takes in precision
 * prec, info[5]: [0] the first hidden layer forms S from the input plane,
 * forward and adjoint; [1] its adjoint also forms that layer's weight
 * gradient; [2] the last hidden layer forms Tbar from the output layer's
 * Sbar; [3] that adjoint also forms the output layer's weight gradient (one
 * output row) -- 1 taken, 0 not; [4] the layers l (bit l, 1-based as the
 * ABI counts them) whose weight gradient and transposed map are one
 * kernel.  A step of one point, or one writing per-point rows, takes none
 * of the gradient-forming kernels.  No counterpart in the reference. */
int dnnf90_net_step_info(int nid, int prec, int *info);
/* BF16X6 work spaces: a hidden layer's weight gradient and transposed map
 * (Tbar of the layer below) in one kernel that reads Sbar once, for layers
 * of 32 or 64 neurons over at most 64 inputs.  Tbar is identical bit for
 * bit to the separate map's; the layer's gradient sums the same products
 * over other column chunks (agreement to FP32 rounding, about 1e-7
 * relative).  On by default; returns the previous setting.  No counterpart
 * in the reference. */
int dnnf90_set_fused_grad(int on);
/* FP32 work spaces (DNNF90_PREC_FP32) that keep their planes (not per-layer
 * recompute) and whose network has at least two hidden layers, all 64 wide,
 * one output row, at most 8 inputs and at most 16 layers take the
 * collocation step (dnnf90_residual_step_batch) on fused layer kernels:
 * point-blocked planes ([block][slot][4 points][neuron]), each hidden
 * layer's linear map inside its Faa di Bruno forward (S never stored; T and
 * h written; the map on BF16 three-way split products, moved into FP64
 * every 32 terms), and one adjoint kernel a layer that reads h, forms the
 * pair sums, the layer's weight gradient and Tbar of the layer below (Sbar
 * never stored; FP32 MFMA).
 * The work space's planes are then in that layout, so a dnnf90_backward_batch
 * must follow a dnnf90_forward_batch, not the step (refused otherwise, as
 * always).  dnnf90_set_fused_layers(0) takes the layer-by-layer kernels
 * instead (results agree to FP32 rounding).  On by default; returns the
 * previous setting.  No counterpart in the reference (its lib/ runs one
 * point at a time). */
int dnnf90_set_fused_layers(int on);
/* whether the fused layer kernels take this network's FP32 collocation
 * step (in a work space that keeps its planes), info[14]: [0] 1 taken; 0
 * not (shape or switch); -1 the family is too large to unroll, -2 hipRTC
 * failed, -3 module or kernel missing, -4 a kernel spills past its
 * allowance, -5 no occupancy; [1..6] registers and [7..12] scratch bytes a
 * lane of the six kernels (first hidden layer forward, forward, last hidden
 * layer forward with the output row, last hidden layer adjoint, adjoint,
 * first hidden layer adjoint); [13] the adjoint kernels' grid.  Compiles the
 * module on first use.  No counterpart in the reference. */
int dnnf90_net_fused_info(int nid, int *info);
/* Debug mode for reads of unwritten memory: batched work spaces created
 * while it is on fill their planes, operand scratch and gradient partials
 * with NaN (every byte 0xFF) at creation and again before every batched
 * forward, so a value that a step reads without having written it
 * propagates as NaN.  Results must not change (tested).  Off by default;
 * returns the previous setting.  No counterpart in the reference (whose
 * hardened build poisons its reals with signalling NaN, Makefile:254-258). */
int dnnf90_set_poison(int on);

int dnnf90_net_nweights(int nid);
int dnnf90_net_nlayer(int nid);
int dnnf90_net_dims(int nid, int *dims);
int dnnf90_net_nderiv(int nid);
/* inverse of dnnf90_net_set_layer: reads layer l (2..nlayer) */
int dnnf90_net_get_layer(int nid, int l, double *wrow);
/* whole weight vector (nweights doubles), read and write */
int dnnf90_net_get_weights(int nid, double *w);
int dnnf90_net_set_weights(int nid, const double *w);
/* nn_weight.dat writer (the format lib/net_module.f90:219-243 writes) */
int dnnf90_net_save(int nid, const char *filename);

/* gradient read-out and write (nweights doubles) */
int dnnf90_grad_get(int gid, double *g);
int dnnf90_grad_set(int gid, const double *g);

/* batched work space for up to maxpt points in arithmetic prec */
int dnnf90_bwork_create(int nid, int maxpt, int prec);
/* the same with a device-memory budget in bytes: when keeping the plane the
 * adjoint needs of every hidden layer (S, or h where the layers keep h; see
 * dnnf90_set_store_s) would pass the budget, the work space keeps the
 * layers' T only and the reverse pass recomputes that plane of each hidden
 * layer from the stored T of the layer below, into one scratch plane -- one
 * more linear map (and, for h, Faa di Bruno step) per hidden layer, results
 * identical bit for bit.  budget_bytes <= 0: keep everything
 * (dnnf90_bwork_create). */
int dnnf90_bwork_create_budget(int nid, int maxpt, int prec, double budget_bytes);
/* info[5]: recompute (0/1), device bytes held, bytes needed with every plane
 * kept, maxpt, prec */
int dnnf90_bwork_info(int bid, double *info);
int dnnf90_bwork_free(int bid);
/* launch on the caller's HIP stream instead of the work space's own
 * (stream is a hipStream_t; NULL restores the work space's own stream) */
int dnnf90_bwork_set_stream(int bid, void *stream);
/* wait for everything launched on the work space's stream */
int dnnf90_bwork_sync(int bid);
/* kernel timing: with profiling on (on = 1), every kernel category is
 * bracketed by HIP events on the work space's stream; on = 2 + c brackets
 * category c alone (so that one kernel can be timed inside otherwise
 * unmarked steps); on = 0 stops.  dnnf90_bwork_kernel_ms returns and
 * resets the accumulated milliseconds and launch counts of the nine
 * categories (ms[9], count[9]): 0 linear map (W*T), 1 Faa di Bruno
 * forward, 2 adjoint pair sums, 3 weight gradient chunk partials (or
 * per-point rows), 4 transposed linear map (W^T*Sbar), 5 input/seed/output
 * kernels, 6 fixed-order reduction of the gradient partials, 7 the Kalman
 * filter's covariance pass (downdate + P j), 8 its gain and weight step */
int dnnf90_bwork_profile(int bid, int on);
int dnnf90_bwork_kernel_ms(int bid, double *ms, long *count);

/* forward over npt points; tout [npt][nout][nderiv] (may be NULL); the
 * work space keeps every layer's state for a following backward */
int dnnf90_forward_batch(int nid, int bid, int npt, const double *x,
                         double *tout, int flags);
/* adjoint of seed [npt][nout][nderiv] (dL/dT) after a forward on the same
 * work space, npt being that forward's point count (the seed must hold that
 * many rows); the weight gradient is ADDED to gid */
int dnnf90_backward_batch(int nid, int bid, const double *seed, int gid,
                          int flags);
/* value-only batched step of a single-output network (net_grad_batch,
 * lib/train_module.f90:525-631): forward with sigma only, output seed
 * coef * (u - y) per point, deltas (W^T D) sigma', gradient blocks with the
 * bias added to gid.  x [npt][d0], y [npt].  Uses the work space's memory
 * with one slot; a following dnnf90_backward_batch needs a new forward. */
int dnnf90_grad_batch(int nid, int bid, int npt, const double *x, const double *y, double coef, int gid,
                      int flags);

/* fused collocation step on output 1: R = sum_k coeff[k] * [T_1 *] T_slot[k]
 * (slot 1-based; nonlin[k] != 0 multiplies by the value T_1), seed =
 * fac * R * dR/dT (app/pinn_module.f90:32-65), adjoint, gradient ADDED to
 * gid.  coeff, slot and nonlin are host arrays whatever the flags; *rsq
 * (host, may be NULL; when given the call waits for the result) receives
 * sum_n R_n^2 */
int dnnf90_residual_step_batch(int nid, int bid, int npt, const double *x,
                               int nterm, const double *coeff,
                               const int *slot, const int *nonlin,
                               double fac, int gid, double *rsq, int flags);
/* the network's outputs and their derivatives (the output row R was formed
 * from) at the points of the last dnnf90_residual_step_batch on this work
 * space, tout[npt][nout * nderiv] as dnnf90_forward_batch writes them (host
 * or, with DNNF90_DEVICE_PTRS, device); -1 if no such step left them (a
 * forward, another call or a one-point step since).  No counterpart in the
 * reference. */
int dnnf90_bwork_step_output(int bid, double *tout, int flags);
/* Gauss-Newton rows dR_n/dw of the same residual (seed dR/dT), written
 * transposed as FP32 rows_t[w][n], w < nweights (device pointer always).
 * The kernels write whole 64-point blocks, the rows of the padding points
 * as zeros, so the row stride must be at least npt rounded up to a
 * multiple of 64 (-1 otherwise); dnnf90_ngd_solve further needs ld to be a
 * multiple of 16. */
int dnnf90_residual_rows_batch(int nid, int bid, int npt, const double *x,
                               int nterm, const double *coeff,
                               const int *slot, const int *nonlin,
                               float *rows_t, long ld, int flags);

/* NGD step (app/optimizer_module.f90:221-469, primal route): G = R R^T over
 * the first nrow columns of the transposed FP32 rows rows_t[nweights][ld]
 * (device; e.g. from dnnf90_residual_rows_batch; ld a multiple of 16, columns
 * nrow..ld-1 zero), formed on FP32 MFMA; damping c = mu
 * (mode 1, ABS) or mu*tr(G)/nb (mode 0, TRACE); u = (G/nb + cI)^{-1} (b/nb)
 * by an FP64 Cholesky on the GPU.  b and u hold nweights doubles (host, or
 * device with DNNF90_DEVICE_PTRS).  Returns -5 when G/nb + cI is not
 * positive definite (the reference stops on a singular pivot,
 * optimizer_module.f90:489-505).  Work space comes from bid.
 * flags & DNNF90_NGD_DUAL takes the Gram route of
 * app/optimizer_module.f90:408-446 instead: K = R^T R / nb + cI (nrow x
 * nrow), y = K^{-1} (R^T b) / nb, u = (b - R y) / nb / c -- the same step
 * to rounding, cheaper when nrow < nweights (needs mu > 0). */
int dnnf90_ngd_solve(int nid, int bid, int nrow, const float *rows_t, long ld,
                     const double *b, int nb, double mu, int mode, double *u,
                     int flags);

/* ---- extended Kalman filter over the weights (lib/kalman_module.f90) ----
 * The covariance lives on the device: dense, as the lower triangle of P
 * (nw (nw + 64) / 2 doubles, no size guard beyond device memory; the
 * reference stops above 4 GB, :133-141), or decoupled, one (fan-in + 1)^2
 * block per neuron (:71-83, fan-in <= 8191 here).  The row of an
 * observation is dy/dw in this library's weight order, which is the
 * reference's flatten order (:378-386).  Observations are processed in
 * order, entirely on the device; the weights move on the device and are
 * copied back to the network once per call.  The observation entries take
 * an FP64 batched work space bid of the same network and return -5 (where
 * the reference stops, :399-405) when a denominator lambda + j.Pj is not
 * positive; the filter then refuses further updates. */

/* kf_init (:107-167): p0 > 0, lam_ini and lam0 in (0, 1]; decoupled != 0
 * for the node-wise filter.  Returns a filter handle of network nid. */
int dnnf90_kf_create(int nid, double p0, double lam_ini, double lam0, int decoupled);
/* the innovation gate and the process noise q (kalman_t%gate :62, %q :70,
 * set after kf_init by app/lib_net_module.f90:474-475); 0 = off */
int dnnf90_kf_set(int kid, double gate, double q);
int dnnf90_kf_free(int kid);
/* kf_update_grad (:353-493): one update from the row held by the gradient
 * accumulator gid, residual xi = target - prediction, row and residual
 * scaled by scale */
int dnnf90_kf_update_grad(int nid, int kid, int gid, double xi, double scale);
/* kf_update_obs (:309-334) and kf_update (:201-218, a one-hot seed) over
 * nobs observations x[nobs][d0], seed[nobs][nout][na] (the reference's
 * form is nout = 1): xi = target[n] - pred[n], or, pred == NULL,
 * target[n] - seed.T(x) at the running weights.  x and seed are device
 * pointers with DNNF90_DEVICE_PTRS; target and pred are host arrays. */
int dnnf90_kf_update_obs(int nid, int kid, int bid, int nobs, const double *x,
                         const double *seed, const double *target,
                         const double *pred, int flags);
/* kf_update_resid_multi (:280-307) and kf_update_resid (:240-254): seed =
 * dR/dT of every output, xi = -resid[n], row and residual scaled by
 * 1/sqrt(rnoise[n]) (rnoise == NULL: 1).  resid, rnoise on the host. */
int dnnf90_kf_update_resid(int nid, int kid, int bid, int nobs, const double *x,
                           const double *seed, const double *resid,
                           const double *rnoise, int flags);
/* the trainer's collocation observation (app/sgd_batch_module.f90:547-550):
 * R = rsrc[n] + sum_k coeff_k [T_0 *] T_slot_k on output 0 (calc_pinn_residual,
 * app/pinn_module.f90:32-47; slots 1-based as in dnnf90_residual_step_batch;
 * rsrc may be NULL), formed on the device at the running weights, seed
 * dR/dT (set_pinn_seed with R = fac = 1), then kf_update_resid(x, seed, R). */
int dnnf90_kf_update_residual(int nid, int kid, int bid, int nobs, const double *x,
                              int nterm, const double *coeff, const int *slot,
                              const int *nonlin, const double *rsrc, int flags);
/* kf_iekf_begin (:512-528) and kf_iekf_iter (:532-648): resid = R at the
 * current weights (the caller re-evaluates it, e.g. with
 * dnnf90_forward_batch), seed = dR/dT [nout][na], rnoise > 0; final != 0
 * downdates the covariance and advances lambda. */
int dnnf90_kf_iekf_begin(int nid, int kid);
int dnnf90_kf_iekf_iter(int nid, int kid, int bid, const double *x, const double *seed,
                        double resid, double rnoise, int final, int flags);
/* The covariance in the reference's layout -- P(nw, nw) column-major, or
 * Pb(maxb, maxb, nb) column-major with zeros outside each block -- (P may
 * be NULL) and info[10] = lambda, lam0, gate, q, gated updates, nw, nb,
 * maxb (0, 0 for the dense filter), broken flag, bytes held on the device. */
int dnnf90_kf_get(int kid, double *P, double *info);

/* comm_eval (lib/committee_module.f90:79-111) over npt points: mean and
 * sample standard deviation (Welford, n - 1 denominator, 0 for a single
 * member) of every slot of every output over the nmem networks nids, which
 * share the shape and the table family; tmean, tstd [npt][nout][na].  bid is
 * a batched work space of nids[0] (its precision is the members'). */
int dnnf90_comm_eval_batch(int nmem, const int *nids, int bid, int npt, const double *x,
                           double *tmean, double *tstd, int flags);

/* dnnf90_grad_point without allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving or per-observation loops, harnesses that forbid
 * hipMalloc).  One point through the generated one-point kernels: x [d0],
 * seed [nderiv] (output 1; the others are seeded with zero) and grad [nw]
 * are DEVICE pointers, the gradient is ADDED to grad, everything is
 * enqueued on stream (a hipStream_t; NULL = the null stream) and nothing
 * waits.  The planes live in a fixed device array of 49152 doubles
 * (384 KiB) per GPU: networks whose planes need more, or that fall outside
 * the one-point kernels (more than 16 layers, a width above 256, width x
 * slots above 1536, a family above 48 slots or 96 operand entries), are
 * refused with -1.  The network must have been prepared by a work space
 * (dnnf90_twork_create or any other work-space creation); a weight change
 * since then is copied, nothing is allocated.  All calls on one GPU share
 * the array: calls that overlap on one GPU race. */
int dnnf90_grad_point_noalloc(int nid, const double *x, const double *seed, double *grad, void *stream);

#ifdef __cplusplus
}
#endif
#endif
