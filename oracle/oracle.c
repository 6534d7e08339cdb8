/* oracle.c -- CPU FP64 restatement of the DNNF90 hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed
 * CPU baseline.  The product never links it.
 *
 * Every routine follows a named reference routine loop for loop, in the same
 * accumulation order, so that on the same libm its results agree with the
 * reference to the last bit or two:
 *   table build ........ lib/multi_index_bell_module.f90:128-351
 *   tanh polynomials ... lib/multi_index_bell_module.f90:406-429, 852-867
 *   sin / erf sigma^(q)  lib/multi_index_bell_module.f90:745-755, 831-850
 *   forward ............ lib/train_module.f90:241-335
 *   adjoint ............ lib/train_module.f90:341-450
 *   collocation seed ... app/pinn_module.f90:32-65
 *   NGD primal solve ... app/optimizer_module.f90:221-469 (primal branch),
 *                        gauss_solve :472-522
 *   committee .......... lib/committee_module.f90:79-111
 *   Kalman filter ...... lib/kalman_module.f90:107-198, 353-659
 * The erf-GELU activation (code 5) has no reference counterpart; its
 * sigma^(q) is the closed form x*Phi^(q) + q*Phi^(q-1) with
 * Phi^(m) = (-1)^(m-1) He_(m-1)(x) phi(x).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OT_MAXH 32

typedef struct {
    int used, d0, k, na, iact;
    int *alpha, *deg, *ind_e1;
    int *pair_start, *pair_num, npair, *pair_ia, *pair_id;
    double *pair_binom;
    int *fq_start, *fq_num, nfq, *fq_ib, *fq_id; /* start/num: [na][k+1] */
    double *fq_c;
    double *tpc; /* [q][i], q = 0..k+1, i = 0..k+2 */
} ot_tab;

static ot_tab g_tab[OT_MAXH];

static double fact_tab[17];

/* ------------------------------------------------------------------ */
/* table build: multi_index_bell_module.f90:128-351                     */

static long n_multi(int d0, int k)
{ /* sum_p C(p+d0-1, p), :551-572 */
    long tot = 0;
    for (int p = 0; p <= k; ++p) {
        long c = 1;
        for (int i = 1; i <= p; ++i) c = c * (d0 - 1 + i) / i;
        tot += c;
    }
    return tot;
}

/* degree-p multi-indices in combinations-with-replacement order (:576-615) */
static void cwr(int d0, int p, int depth, int vmin, int *slots, int *list, long *cnt)
{
    for (int v = vmin; v < d0; ++v) {
        slots[depth] = v;
        if (depth == p - 1) {
            int *a = list + (*cnt) * d0;
            memset(a, 0, sizeof(int) * d0);
            for (int i = 0; i < p; ++i) a[slots[i]]++;
            (*cnt)++;
        } else {
            cwr(d0, p, depth + 1, v, slots, list, cnt);
        }
    }
}

static int lex_cmp(const int *a, const int *b, int d0)
{
    for (int i = 0; i < d0; ++i)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

static const int *g_sort_alpha;
static int g_sort_d0;
static int sort_cmp(const void *x, const void *y)
{
    int i = *(const int *)x, j = *(const int *)y;
    return lex_cmp(g_sort_alpha + i * g_sort_d0, g_sort_alpha + j * g_sort_d0, g_sort_d0);
}

static int find_alpha(const ot_tab *t, const int *lex, const int *a)
{ /* binary search over the lexicographic order (:520-538) */
    int lo = 0, hi = t->na - 1;
    while (lo <= hi) {
        int mid = (lo + hi) / 2, ic = lex[mid];
        int c = lex_cmp(t->alpha + ic * t->d0, a, t->d0);
        if (c == 0) return ic;
        if (c < 0) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

static void tab_free(ot_tab *t)
{
    free(t->alpha); free(t->deg); free(t->ind_e1);
    free(t->pair_start); free(t->pair_num); free(t->pair_ia); free(t->pair_id);
    free(t->pair_binom); free(t->fq_start); free(t->fq_num); free(t->fq_ib);
    free(t->fq_id); free(t->fq_c); free(t->tpc);
    memset(t, 0, sizeof(*t));
}

int ot_tables_build(int d0, int k, int nseed, const int *seeds)
{
    if (d0 < 1 || k < 1 || k > 15 || nseed < 0) return -1;
    for (int i = 0; i < d0 * nseed; ++i)
        if (seeds[i] < 0 || seeds[i] > k) return -1;
    int h = -1;
    for (int i = 0; i < OT_MAXH; ++i)
        if (!g_tab[i].used) { h = i; break; }
    if (h < 0) return -1;
    ot_tab *t = &g_tab[h];
    memset(t, 0, sizeof(*t));

    fact_tab[0] = 1.0;
    for (int i = 1; i <= 16; ++i) fact_tab[i] = fact_tab[i - 1] * (double)i;

    long nfull = n_multi(d0, k);
    int *afull = (int *)malloc(sizeof(int) * nfull * d0);
    long cnt = 0;
    int slots[16];
    for (int p = 0; p <= k; ++p) {
        if (p == 0) { memset(afull, 0, sizeof(int) * d0); cnt = 1; }
        else cwr(d0, p, 0, 0, slots, afull, &cnt);
    }
    /* closure filter (:164-180) */
    char *incl = (char *)malloc(nfull);
    for (long ia = 0; ia < nfull; ++ia) incl[ia] = nseed > 0 ? 0 : 1;
    for (int s = 0; s < nseed; ++s) {
        int ssum = 0;
        for (int v = 0; v < d0; ++v) ssum += seeds[s * d0 + v];
        if (ssum > k) { free(afull); free(incl); return -1; }
        for (long ia = 0; ia < nfull; ++ia) {
            int ok = 1;
            for (int v = 0; v < d0; ++v)
                if (afull[ia * d0 + v] > seeds[s * d0 + v]) { ok = 0; break; }
            if (ok) incl[ia] = 1;
        }
    }
    int na = 0;
    for (long ia = 0; ia < nfull; ++ia) na += incl[ia];
    t->d0 = d0; t->k = k; t->na = na; t->iact = 0;
    t->alpha = (int *)malloc(sizeof(int) * na * d0);
    t->deg = (int *)malloc(sizeof(int) * na);
    int c = 0;
    for (long ia = 0; ia < nfull; ++ia) {
        if (!incl[ia]) continue;
        int s = 0;
        for (int v = 0; v < d0; ++v) { t->alpha[c * d0 + v] = afull[ia * d0 + v]; s += afull[ia * d0 + v]; }
        t->deg[c++] = s;
    }
    free(afull); free(incl);

    int *lex = (int *)malloc(sizeof(int) * na);
    for (int i = 0; i < na; ++i) lex[i] = i;
    g_sort_alpha = t->alpha; g_sort_d0 = d0;
    qsort(lex, na, sizeof(int), sort_cmp);

    t->ind_e1 = (int *)malloc(sizeof(int) * d0);
    for (int v = 0; v < d0; ++v) t->ind_e1[v] = -1;
    for (int ia = 0; ia < na; ++ia)
        if (t->deg[ia] == 1)
            for (int v = 0; v < d0; ++v)
                if (t->alpha[ia * d0 + v] == 1) t->ind_e1[v] = ia;

    /* ind_diff and multi-binomials (:219-234) */
    int *idiff = (int *)malloc(sizeof(int) * (size_t)na * na);
    double *bin = (double *)malloc(sizeof(double) * (size_t)na * na);
    int *tmp = (int *)malloc(sizeof(int) * d0);
    for (int ia = 0; ia < na; ++ia) {
        const int *a = t->alpha + ia * d0;
        for (int ib = 0; ib < na; ++ib) {
            const int *b = t->alpha + ib * d0;
            int le = 1;
            for (int v = 0; v < d0; ++v) if (b[v] > a[v]) { le = 0; break; }
            idiff[(size_t)ia * na + ib] = -1;
            bin[(size_t)ia * na + ib] = 0.0;
            if (!le) continue;
            for (int v = 0; v < d0; ++v) tmp[v] = a[v] - b[v];
            idiff[(size_t)ia * na + ib] = find_alpha(t, lex, tmp);
            double r = 1.0;
            for (int v = 0; v < d0; ++v)
                r = r * fact_tab[a[v]] / (fact_tab[b[v]] * fact_tab[a[v] - b[v]]);
            bin[(size_t)ia * na + ib] = r;
        }
    }
    free(tmp); free(lex);

    /* adjoint pair lists (:267-297): for every b >= 2, all a with b <= a */
    t->pair_start = (int *)calloc(na, sizeof(int));
    t->pair_num = (int *)calloc(na, sizeof(int));
    int np = 0;
    for (int ib = 1; ib < na; ++ib)
        for (int ia = 1; ia < na; ++ia)
            if (idiff[(size_t)ia * na + ib] >= 0) t->pair_num[ib]++, np++;
    t->npair = np;
    t->pair_ia = (int *)malloc(sizeof(int) * (np + 1));
    t->pair_id = (int *)malloc(sizeof(int) * (np + 1));
    t->pair_binom = (double *)malloc(sizeof(double) * (np + 1));
    np = 0;
    for (int ib = 1; ib < na; ++ib) {
        t->pair_start[ib] = np;
        for (int ia = 1; ia < na; ++ia) {
            int id = idiff[(size_t)ia * na + ib];
            if (id >= 0) {
                t->pair_ia[np] = ia; t->pair_id[np] = id;
                t->pair_binom[np] = bin[(size_t)ia * na + ib];
                np++;
            }
        }
    }

    /* forward DP lists (:299-332): coefficient C(a,b)/q */
    int kq = k + 1;
    t->fq_start = (int *)calloc((size_t)na * kq, sizeof(int));
    t->fq_num = (int *)calloc((size_t)na * kq, sizeof(int));
    int nf = 0;
    for (int ia = 1; ia < na; ++ia) {
        int p = t->deg[ia];
        for (int q = 2; q <= p; ++q)
            for (int ib = 1; ib < na; ++ib)
                if (idiff[(size_t)ia * na + ib] >= 0 && p - t->deg[ib] >= q - 1) nf++;
    }
    t->nfq = nf;
    t->fq_ib = (int *)malloc(sizeof(int) * (nf + 1));
    t->fq_id = (int *)malloc(sizeof(int) * (nf + 1));
    t->fq_c = (double *)malloc(sizeof(double) * (nf + 1));
    nf = 0;
    for (int ia = 1; ia < na; ++ia) {
        int p = t->deg[ia];
        for (int q = 2; q <= p; ++q) {
            t->fq_start[ia * kq + q] = nf;
            for (int ib = 1; ib < na; ++ib) {
                int id = idiff[(size_t)ia * na + ib];
                if (id >= 0 && p - t->deg[ib] >= q - 1) {
                    t->fq_ib[nf] = ib; t->fq_id[nf] = id;
                    t->fq_c[nf] = bin[(size_t)ia * na + ib] / (double)q;
                    nf++;
                }
            }
            t->fq_num[ia * kq + q] = nf - t->fq_start[ia * kq + q];
        }
    }
    free(idiff); free(bin);

    /* tanh polynomials P_0 = t, P_{q+1} = (1-t^2) P_q' (:406-429) */
    int qmax = k + 1, ni = k + 3;
    t->tpc = (double *)calloc((size_t)(qmax + 1) * ni, sizeof(double));
    double *dc = (double *)malloc(sizeof(double) * ni);
    t->tpc[0 * ni + 1] = 1.0;
    for (int q = 0; q < qmax; ++q) {
        for (int i = 0; i <= qmax; ++i) dc[i] = (double)(i + 1) * t->tpc[q * ni + i + 1];
        dc[qmax + 1] = 0.0;
        for (int i = 0; i <= qmax + 1; ++i) {
            t->tpc[(q + 1) * ni + i] = dc[i];
            if (i >= 2) t->tpc[(q + 1) * ni + i] -= dc[i - 2];
        }
    }
    free(dc);
    t->used = 1;
    return h;
}

static ot_tab *tab_get(int h)
{
    if (h < 0 || h >= OT_MAXH || !g_tab[h].used) return NULL;
    return &g_tab[h];
}

int ot_tables_free(int h)
{
    ot_tab *t = tab_get(h);
    if (!t) return -1;
    tab_free(t);
    return 0;
}

int ot_set_activation(int h, int iact)
{
    ot_tab *t = tab_get(h);
    if (!t) return -1;
    if (!(iact == 0 || iact == 1 || iact == 2 || iact == 5)) return -2;
    t->iact = iact;
    return 0;
}

int ot_na(int h) { ot_tab *t = tab_get(h); return t ? t->na : -1; }
int ot_nfq(int h) { ot_tab *t = tab_get(h); return t ? t->nfq : -1; }
int ot_npair(int h) { ot_tab *t = tab_get(h); return t ? t->npair : -1; }

int ot_alpha(int h, int *out)
{
    ot_tab *t = tab_get(h);
    if (!t) return -1;
    memcpy(out, t->alpha, sizeof(int) * t->na * t->d0);
    return 0;
}

int ot_deg(int h, int *out)
{
    ot_tab *t = tab_get(h);
    if (!t) return -1;
    memcpy(out, t->deg, sizeof(int) * t->na);
    return 0;
}

int ot_ind_e1(int h, int *out)
{
    ot_tab *t = tab_get(h);
    if (!t) return -1;
    memcpy(out, t->ind_e1, sizeof(int) * t->d0);
    return 0;
}

int ot_pairs(int h, int *start, int *num, int *ia, int *id, double *binom)
{
    ot_tab *t = tab_get(h);
    if (!t) return -1;
    memcpy(start, t->pair_start, sizeof(int) * t->na);
    memcpy(num, t->pair_num, sizeof(int) * t->na);
    memcpy(ia, t->pair_ia, sizeof(int) * t->npair);
    memcpy(id, t->pair_id, sizeof(int) * t->npair);
    memcpy(binom, t->pair_binom, sizeof(double) * t->npair);
    return 0;
}

int ot_fq(int h, int *start, int *num, int *ib, int *id, double *c)
{
    ot_tab *t = tab_get(h);
    if (!t) return -1;
    memcpy(start, t->fq_start, sizeof(int) * t->na * (t->k + 1));
    memcpy(num, t->fq_num, sizeof(int) * t->na * (t->k + 1));
    memcpy(ib, t->fq_ib, sizeof(int) * t->nfq);
    memcpy(id, t->fq_id, sizeof(int) * t->nfq);
    memcpy(c, t->fq_c, sizeof(double) * t->nfq);
    return 0;
}

int ot_tpc(int h, double *out)
{
    ot_tab *t = tab_get(h);
    if (!t) return -1;
    memcpy(out, t->tpc, sizeof(double) * (t->k + 2) * (t->k + 3));
    return 0;
}

/* ------------------------------------------------------------------ */
/* sigma^(q), q = 0..nq: multi_index_bell_module.f90:723-867            */

static void act_derivs(const ot_tab *t, double a, int nq, double *d)
{
    int ni = t->k + 3;
    if (t->iact == 1) {
        const double pio2 = 1.57079632679489661923;
        for (int q = 0; q <= nq; ++q) d[q] = sin(a + (double)q * pio2);
    } else if (t->iact == 2) {
        const double c2spi = 1.12837916709551257390;
        double psi[20];
        d[0] = erf(a);
        if (nq < 1) return;
        psi[0] = exp(-a * a);
        psi[1] = 2.0 * a * psi[0];
        for (int n = 1; n <= nq - 1; ++n) psi[n + 1] = 2.0 * a * psi[n] - 2.0 * (double)n * psi[n - 1];
        for (int q = 1; q <= nq; ++q) d[q] = c2spi * ((q - 1) % 2 ? -1.0 : 1.0) * psi[q - 1];
    } else if (t->iact == 5) {
        /* erf-GELU x*Phi(x): sigma^(q) = x Phi^(q) + q Phi^(q-1) */
        const double rs2pi = 0.39894228040143267794; /* 1/sqrt(2 pi) */
        const double rs2 = 0.70710678118654752440;   /* 1/sqrt(2) */
        double psi[20], ph[20];
        psi[0] = rs2pi * exp(-0.5 * a * a);
        psi[1] = a * psi[0];
        for (int n = 1; n < nq; ++n) psi[n + 1] = a * psi[n] - (double)n * psi[n - 1];
        ph[0] = 0.5 * erfc(-a * rs2);
        for (int m = 1; m <= nq; ++m) ph[m] = ((m - 1) % 2 ? -1.0 : 1.0) * psi[m - 1];
        d[0] = a * ph[0];
        for (int q = 1; q <= nq; ++q) d[q] = a * ph[q] + (double)q * ph[q - 1];
    } else {
        double tt = tanh(a);
        for (int q = 0; q <= nq; ++q) {
            double v = t->tpc[q * ni + q + 1];
            for (int i = q; i >= 0; --i) v = v * tt + t->tpc[q * ni + i];
            d[q] = v;
        }
    }
}

void ot_act_derivs(int h, double a, int nq, double *d)
{
    ot_tab *t = tab_get(h);
    if (t) act_derivs(t, a, nq, d);
}

/* ------------------------------------------------------------------ */
/* per-point engine: train_module.f90:241-450                           */

int ot_nweights(int nlayer, const int *dims)
{
    int nw = 0;
    for (int l = 1; l < nlayer; ++l) nw += dims[l] * (dims[l - 1] + 1);
    return nw;
}

typedef struct {
    int nlayer, ndmax, na, k;
    const int *dims;
    const double *w;
    int woff[64];
    double *T;   /* [nlayer][ndmax+1][na]  (j = 0 is the bias channel) */
    double *h;   /* [nlayer][ndmax][na] */
    double *S, *Tb, *Sb, *acc; /* [ndmax][na], acc [ndmax][ndmax+1] */
    double *bqv; /* [ndmax][na][k+1] */
    double *dt;  /* [ndmax][k+2] */
} ot_work;

static void work_init(ot_work *wk, const ot_tab *t, int nlayer, const int *dims, const double *w)
{
    wk->nlayer = nlayer; wk->dims = dims; wk->w = w; wk->na = t->na; wk->k = t->k;
    int nd = 0, off = 0;
    for (int l = 0; l < nlayer; ++l) if (dims[l] > nd) nd = dims[l];
    wk->ndmax = nd;
    for (int l = 1; l < nlayer; ++l) { wk->woff[l] = off; off += dims[l] * (dims[l - 1] + 1); }
    size_t na = t->na;
    wk->T = (double *)calloc((size_t)nlayer * (nd + 1) * na, sizeof(double));
    wk->h = (double *)calloc((size_t)nlayer * nd * na, sizeof(double));
    wk->S = (double *)calloc((size_t)nd * na, sizeof(double));
    wk->Tb = (double *)calloc((size_t)nd * na, sizeof(double));
    wk->Sb = (double *)calloc((size_t)nd * na, sizeof(double));
    wk->acc = (double *)calloc((size_t)nd * (nd + 1), sizeof(double));
    wk->bqv = (double *)calloc((size_t)nd * na * (t->k + 1), sizeof(double));
    wk->dt = (double *)calloc((size_t)nd * (t->k + 2), sizeof(double));
    for (int l = 0; l < nlayer; ++l) wk->T[(size_t)l * (nd + 1) * na + 0] = 1.0; /* bias channel */
}

static void work_free(ot_work *wk)
{
    free(wk->T); free(wk->h); free(wk->S); free(wk->Tb); free(wk->Sb);
    free(wk->acc); free(wk->bqv); free(wk->dt);
}

#define TI(l, j, ia) wk->T[((size_t)(l) * (wk->ndmax + 1) + (j)) * na + (ia)]
#define HI(l, j, ia) wk->h[((size_t)(l) * wk->ndmax + (j)) * na + (ia)]

static void forward_core(const ot_tab *t, ot_work *wk, const double *x)
{
    const int na = t->na, K = t->k, kq = K + 1, nl = wk->nlayer;
    const int *dims = wk->dims;
    for (int j = 1; j <= dims[0]; ++j) {
        TI(0, j, 0) = x[j - 1];
        for (int ia = 1; ia < na; ++ia) TI(0, j, ia) = 0.0;
        if (t->ind_e1[j - 1] >= 0) TI(0, j, t->ind_e1[j - 1]) = 1.0;
    }
    for (int l = 1; l < nl; ++l) {
        const int nd = dims[l], ndm = dims[l - 1];
        const double *W = wk->w + wk->woff[l];
        double *S = wk->S;
        for (int ia = 0; ia < na; ++ia)
            for (int j = 0; j < nd; ++j) {
                double s = W[j * (ndm + 1) + 0] * TI(l - 1, 0, ia);
                for (int i = 1; i <= ndm; ++i) s = s + W[j * (ndm + 1) + i] * TI(l - 1, i, ia);
                S[j * na + ia] = s;
            }
        if (l < nl - 1) {
            for (int j = 0; j < nd; ++j) act_derivs(t, S[j * na + 0], K + 1, wk->dt + j * (K + 2));
            for (int j = 0; j < nd; ++j) {
                const double *dt = wk->dt + j * (K + 2);
                double *bq = wk->bqv + (size_t)j * na * kq;
                TI(l, j + 1, 0) = dt[0];
                HI(l, j, 0) = dt[1];
                for (int ia = 1; ia < na; ++ia) {
                    int p = t->deg[ia];
                    double sa = S[j * na + ia];
                    bq[ia * kq + 1] = sa;
                    double ttv = dt[1] * sa, htv = dt[2] * sa;
                    for (int q = 2; q <= p; ++q) {
                        double bsv = 0.0;
                        int st = t->fq_start[ia * kq + q], nn = t->fq_num[ia * kq + q];
                        for (int it = st; it < st + nn; ++it)
                            bsv = bsv + t->fq_c[it] * S[j * na + t->fq_ib[it]] * bq[t->fq_id[it] * kq + q - 1];
                        bq[ia * kq + q] = bsv;
                        ttv = ttv + dt[q] * bsv;
                        htv = htv + dt[q + 1] * bsv;
                    }
                    TI(l, j + 1, ia) = ttv;
                    HI(l, j, ia) = htv;
                }
            }
        } else {
            for (int j = 0; j < nd; ++j)
                for (int ia = 0; ia < na; ++ia) TI(l, j + 1, ia) = S[j * na + ia];
        }
    }
}

/* adjoint from the seeded Tb of the output layer; grad[nw] += */
static void backward_core(const ot_tab *t, ot_work *wk, double *grad)
{
    const int na = t->na, nl = wk->nlayer;
    const int *dims = wk->dims;
    for (int l = nl - 1; l >= 1; --l) {
        const int nd = dims[l], ndm = dims[l - 1];
        const double *W = wk->w + wk->woff[l];
        double *Tb = wk->Tb, *Sb = wk->Sb;
        if (l == nl - 1) {
            for (int j = 0; j < nd; ++j)
                for (int ia = 0; ia < na; ++ia) Sb[j * na + ia] = Tb[j * na + ia];
        } else {
            for (int j = 0; j < nd; ++j) {
                double b = Tb[j * na + 0] * HI(l, j, 0);
                for (int ia = 1; ia < na; ++ia) b = b + Tb[j * na + ia] * HI(l, j, ia);
                Sb[j * na + 0] = b;
            }
            for (int ib = 1; ib < na; ++ib)
                for (int j = 0; j < nd; ++j) {
                    double b = 0.0;
                    for (int ip = t->pair_start[ib]; ip < t->pair_start[ib] + t->pair_num[ib]; ++ip)
                        b = b + t->pair_binom[ip] * Tb[j * na + t->pair_ia[ip]] * HI(l, j, t->pair_id[ip]);
                    Sb[j * na + ib] = b;
                }
        }
        /* gradient block: acc(j,i) = sum_a Sb(j,a) T(i,a,l-1), :415-422 */
        double *acc = wk->acc;
        for (int j = 0; j < nd; ++j)
            for (int i = 0; i <= ndm; ++i) acc[j * (ndm + 1) + i] = 0.0;
        for (int ia = 0; ia < na; ++ia)
            for (int i = 0; i <= ndm; ++i) {
                double tv = TI(l - 1, i, ia);
                for (int j = 0; j < nd; ++j) acc[j * (ndm + 1) + i] = acc[j * (ndm + 1) + i] + Sb[j * na + ia] * tv;
            }
        double *g = grad + wk->woff[l];
        for (int j = 0; j < nd; ++j)
            for (int i = 0; i <= ndm; ++i) g[j * (ndm + 1) + i] = g[j * (ndm + 1) + i] + acc[j * (ndm + 1) + i];
        /* Tb of layer l-1, :430-444 (skipped for the first hidden layer) */
        if (l > 1) {
            for (int ia = 0; ia < na; ++ia)
                for (int i = 1; i <= ndm; ++i) {
                    double s = 0.0;
                    for (int j = 0; j < nd; ++j) s = s + W[j * (ndm + 1) + i] * Sb[j * na + ia];
                    Tb[(i - 1) * na + ia] = s;
                }
        }
    }
}

static void seed_output(const ot_tab *t, ot_work *wk, const double *seed)
{
    int nout = wk->dims[wk->nlayer - 1], na = t->na;
    for (int j = 0; j < nout; ++j)
        for (int ia = 0; ia < na; ++ia) wk->Tb[j * na + ia] = seed[j * na + ia];
}

static int nthreads_eff(int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    return nthreads < 1 ? 1 : nthreads;
#else
    (void)nthreads;
    return 1;
#endif
}

static void copy_out(const ot_tab *t, ot_work *wk, double *tout)
{
    int L = wk->nlayer - 1, nout = wk->dims[L], na = t->na;
    for (int j = 0; j < nout; ++j)
        for (int ia = 0; ia < na; ++ia) tout[j * na + ia] = TI(L, j + 1, ia);
}

int ot_eval(int h, int nlayer, const int *dims, const double *w, int npt,
            const double *x, double *tout, int nthreads)
{
    ot_tab *t = tab_get(h);
    if (!t || nlayer < 2 || nlayer > 63 || dims[0] != t->d0) return -1;
    int nt = nthreads_eff(nthreads), nout = dims[nlayer - 1];
#pragma omp parallel num_threads(nt)
    {
        ot_work wk;
        work_init(&wk, t, nlayer, dims, w);
#pragma omp for schedule(static)
        for (int n = 0; n < npt; ++n) {
            forward_core(t, &wk, x + (size_t)n * t->d0);
            copy_out(t, &wk, tout + (size_t)n * nout * t->na);
        }
        work_free(&wk);
    }
    return 0;
}

/* one accumulator per thread, reduced in thread order afterwards
 * (tools/example_train.f90:128-170) */
int ot_grad(int h, int nlayer, const int *dims, const double *w, int npt,
            const double *x, const double *seed, double *grad, int nthreads)
{
    ot_tab *t = tab_get(h);
    if (!t || nlayer < 2 || nlayer > 63 || dims[0] != t->d0) return -1;
    int nt = nthreads_eff(nthreads), nout = dims[nlayer - 1];
    int nw = ot_nweights(nlayer, dims);
    double *part = (double *)calloc((size_t)nt * nw, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        ot_work wk;
        work_init(&wk, t, nlayer, dims, w);
#pragma omp for schedule(static)
        for (int n = 0; n < npt; ++n) {
            forward_core(t, &wk, x + (size_t)n * t->d0);
            seed_output(t, &wk, seed + (size_t)n * nout * t->na);
            backward_core(t, &wk, part + (size_t)tid * nw);
        }
        work_free(&wk);
    }
    for (int i = 0; i < nt; ++i)
        for (int p = 0; p < nw; ++p) grad[p] += part[(size_t)i * nw + p];
    free(part);
    return 0;
}

int ot_grad_rows(int h, int nlayer, const int *dims, const double *w, int npt,
                 const double *x, const double *seed, double *rows, int nthreads)
{
    ot_tab *t = tab_get(h);
    if (!t || nlayer < 2 || nlayer > 63 || dims[0] != t->d0) return -1;
    int nt = nthreads_eff(nthreads), nout = dims[nlayer - 1];
    int nw = ot_nweights(nlayer, dims);
#pragma omp parallel num_threads(nt)
    {
        ot_work wk;
        work_init(&wk, t, nlayer, dims, w);
#pragma omp for schedule(static)
        for (int n = 0; n < npt; ++n) {
            double *r = rows + (size_t)n * nw;
            memset(r, 0, sizeof(double) * nw);
            forward_core(t, &wk, x + (size_t)n * t->d0);
            seed_output(t, &wk, seed + (size_t)n * nout * t->na);
            backward_core(t, &wk, r);
        }
        work_free(&wk);
    }
    return 0;
}

/* R = sum_k c_k [T_0 *] T_slot (calc_pinn_residual, app/pinn_module.f90:32-47) */
static double pinn_residual(const double *tv, int nterm, const double *coeff,
                            const int *slot, const int *nonlin)
{
    double R = 0.0;
    for (int k = 0; k < nterm; ++k) {
        if (nonlin[k]) R = R + coeff[k] * tv[0] * tv[slot[k]];
        else R = R + coeff[k] * tv[slot[k]];
    }
    return R;
}

/* seed = fac*R*dR/dT (set_pinn_seed, app/pinn_module.f90:50-65) */
static void pinn_seed(const double *tv, double R, double fac, int na, int nterm,
                      const double *coeff, const int *slot, const int *nonlin, double *seed)
{
    for (int ia = 0; ia < na; ++ia) seed[ia] = 0.0;
    for (int k = 0; k < nterm; ++k) {
        if (nonlin[k]) {
            seed[slot[k]] = seed[slot[k]] + fac * R * coeff[k] * tv[0];
            seed[0] = seed[0] + fac * R * coeff[k] * tv[slot[k]];
        } else {
            seed[slot[k]] = seed[slot[k]] + fac * R * coeff[k];
        }
    }
}

int ot_pinn_step(int h, int nlayer, const int *dims, const double *w, int npt,
                 const double *x, int nterm, const double *coeff,
                 const int *slot, const int *nonlin, double fac,
                 double *grad, double *rsq_out, double *tout, int nthreads)
{
    ot_tab *t = tab_get(h);
    if (!t || nlayer < 2 || nlayer > 63 || dims[0] != t->d0) return -1;
    for (int k = 0; k < nterm; ++k)
        if (slot[k] < 0 || slot[k] >= t->na) return -1;
    int nt = nthreads_eff(nthreads), na = t->na, nout = dims[nlayer - 1];
    int nw = ot_nweights(nlayer, dims);
    double *part = (double *)calloc((size_t)nt * nw, sizeof(double));
    double *rsq = (double *)calloc(nt, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        ot_work wk;
        work_init(&wk, t, nlayer, dims, w);
        double *tv = (double *)calloc((size_t)nout * na, sizeof(double));
        double *sd = (double *)calloc((size_t)nout * na, sizeof(double));
#pragma omp for schedule(static)
        for (int n = 0; n < npt; ++n) {
            forward_core(t, &wk, x + (size_t)n * t->d0);
            copy_out(t, &wk, tv);
            if (tout) memcpy(tout + (size_t)n * nout * na, tv, sizeof(double) * nout * na);
            double R = pinn_residual(tv, nterm, coeff, slot, nonlin);
            rsq[tid] += R * R;
            pinn_seed(tv, R, fac, na, nterm, coeff, slot, nonlin, sd);
            seed_output(t, &wk, sd);
            backward_core(t, &wk, part + (size_t)tid * nw);
        }
        free(tv); free(sd);
        work_free(&wk);
    }
    double r = 0.0;
    for (int i = 0; i < nt; ++i) {
        r += rsq[i];
        for (int p = 0; p < nw; ++p) grad[p] += part[(size_t)i * nw + p];
    }
    if (rsq_out) *rsq_out = r;
    free(part); free(rsq);
    return 0;
}

int ot_pinn_rows(int h, int nlayer, const int *dims, const double *w, int npt,
                 const double *x, int nterm, const double *coeff,
                 const int *slot, const int *nonlin, double *rows, int nthreads)
{
    ot_tab *t = tab_get(h);
    if (!t || nlayer < 2 || nlayer > 63 || dims[0] != t->d0) return -1;
    int nt = nthreads_eff(nthreads), na = t->na, nout = dims[nlayer - 1];
    int nw = ot_nweights(nlayer, dims);
#pragma omp parallel num_threads(nt)
    {
        ot_work wk;
        work_init(&wk, t, nlayer, dims, w);
        double *tv = (double *)calloc((size_t)nout * na, sizeof(double));
        double *sd = (double *)calloc((size_t)nout * na, sizeof(double));
#pragma omp for schedule(static)
        for (int n = 0; n < npt; ++n) {
            double *r = rows + (size_t)n * nw;
            memset(r, 0, sizeof(double) * nw);
            forward_core(t, &wk, x + (size_t)n * t->d0);
            copy_out(t, &wk, tv);
            /* row seed dR/dT: set_pinn_seed(t, 1, 1) (sgd_batch_module.f90:742) */
            pinn_seed(tv, 1.0, 1.0, na, nterm, coeff, slot, nonlin, sd);
            seed_output(t, &wk, sd);
            backward_core(t, &wk, r);
        }
        free(tv); free(sd);
        work_free(&wk);
    }
    return 0;
}

/* Gaussian elimination with partial pivoting, A row-major (optimizer_module.f90:472-522) */
static int gauss_solve(int n, double *A, double *b)
{
    double anorm = 0.0;
    for (size_t i = 0; i < (size_t)n * n; ++i) if (fabs(A[i]) > anorm) anorm = fabs(A[i]);
    double ptol = 1e-13 * anorm;
    for (int k = 0; k < n - 1; ++k) {
        int ip = k;
        double pmax = fabs(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[(size_t)i * n + k]) > pmax) { pmax = fabs(A[(size_t)i * n + k]); ip = i; }
        if (pmax <= ptol) return -3;
        if (ip != k) {
            for (int j = k; j < n; ++j) {
                double r = A[(size_t)k * n + j];
                A[(size_t)k * n + j] = A[(size_t)ip * n + j];
                A[(size_t)ip * n + j] = r;
            }
            double r = b[k]; b[k] = b[ip]; b[ip] = r;
        }
        for (int i = k + 1; i < n; ++i) {
            double fac = A[(size_t)i * n + k] / A[(size_t)k * n + k];
            for (int j = k; j < n; ++j) A[(size_t)i * n + j] -= fac * A[(size_t)k * n + j];
            b[i] -= fac * b[k];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = 0.0;
        for (int j = i + 1; j < n; ++j) s += A[(size_t)i * n + j] * b[j];
        b[i] = (b[i] - s) / A[(size_t)i * n + i];
    }
    return 0;
}

int ot_ngd_solve(int nrow, int nw, const double *rows, const double *b,
                 int nb, double mu, int mode, double *u)
{
    if (nrow < 1 || nw < 1 || nb < 1) return -1;
    double *G = (double *)calloc((size_t)nw * nw, sizeof(double));
    /* G = sum_n j_n j_n^T (:310-318) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nw; ++i)
        for (int ib = 0; ib < nrow; ++ib) {
            double ri = rows[(size_t)ib * nw + i];
            if (ri == 0.0) continue;
            for (int j = 0; j < nw; ++j) G[(size_t)i * nw + j] += ri * rows[(size_t)ib * nw + j];
        }
    double trace = 0.0;
    for (int ib = 0; ib < nrow; ++ib) {
        double s = 0.0;
        for (int i = 0; i < nw; ++i) s += rows[(size_t)ib * nw + i] * rows[(size_t)ib * nw + i];
        trace += s;
    }
    for (size_t i = 0; i < (size_t)nw * nw; ++i) G[i] /= (double)nb;
    for (int i = 0; i < nw; ++i) u[i] = b[i] / (double)nb;
    trace /= (double)nb;
    double c = (mode == 1) ? mu : mu * trace;
    for (int i = 0; i < nw; ++i) G[(size_t)i * nw + i] += c;
    int rc = gauss_solve(nw, G, u);
    free(G);
    return rc;
}

/* ------------------------------------------------------------------ */
/* committee mean and spread: lib/committee_module.f90:79-111          */

int ot_comm_eval(int h, int nlayer, const int *dims, int nmem, const double *w,
                 int npt, const double *x, double *tmean, double *tstd)
{
    ot_tab *t = tab_get(h);
    if (!t || nmem < 1 || nlayer < 2 || dims[0] != t->d0) return -1;
    const int nw = ot_nweights(nlayer, dims), nout = dims[nlayer - 1], na = t->na;
    const size_t len = (size_t)npt * nout * na;
    double *tk = (double *)malloc(sizeof(double) * (len ? len : 1));
    for (size_t i = 0; i < len; ++i) tmean[i] = 0.0, tstd[i] = 0.0;
    for (int k = 1; k <= nmem; ++k) {
        ot_eval(h, nlayer, dims, w + (size_t)(k - 1) * nw, npt, x, tk, 0);
        /* Welford: delta, then M2 += delta^2 (k-1)/k, then mean += delta/k */
        for (size_t i = 0; i < len; ++i) {
            const double d = tk[i] - tmean[i];
            tstd[i] = tstd[i] + d * d * (double)(k - 1) / (double)k;
            tmean[i] = tmean[i] + d / (double)k;
        }
    }
    for (size_t i = 0; i < len; ++i) tstd[i] = nmem > 1 ? sqrt(tstd[i] / (double)(nmem - 1)) : 0.0;
    free(tk);
    return 0;
}

/* ------------------------------------------------------------------ */
/* extended Kalman filter: lib/kalman_module.f90                        */
/* The weight vector is the flatten order of kf_update_grad (:378-386):
 * l ascending, neuron j, then bias and fan-in i -- this file's layout.
 * Dense P is kept column-major as the reference's P(nw,nw); the decoupled
 * filter keeps one maxb x maxb column-major block per neuron, Pb(:,:,ib). */

#define OT_KF_MAX 16

typedef struct {
    int used, nw, decoupled, nb, maxb, ngated;
    double lam, lam0, gate, q;
    int *bsz, *boff;
    double *P, *jf, *pj, *kv, *wprior, *wcur;
} ot_kf;

static ot_kf g_kf[OT_KF_MAX];

static ot_kf *kf_get(int h) { return (h >= 1 && h <= OT_KF_MAX && g_kf[h - 1].used) ? &g_kf[h - 1] : NULL; }

/* kf_init :107-167 and kf_block_map :171-185 (the size guard of :133-141
 * is the caller's business here) */
int ot_kf_create(int nlayer, const int *dims, double p0, double lam_ini, double lam0, int decoupled)
{
    if (nlayer < 2 || p0 <= 0.0 || lam_ini <= 0.0 || lam_ini > 1.0 || lam0 <= 0.0 || lam0 > 1.0) return -1;
    int h = 0;
    for (int i = 0; i < OT_KF_MAX; ++i)
        if (!g_kf[i].used) { h = i + 1; break; }
    if (!h) return -1;
    ot_kf *kf = &g_kf[h - 1];
    memset(kf, 0, sizeof(*kf));
    kf->used = 1;
    kf->nw = ot_nweights(nlayer, dims);
    kf->decoupled = decoupled != 0;
    const int nw = kf->nw;
    if (!kf->decoupled) {
        kf->P = (double *)calloc((size_t)nw * nw, sizeof(double));
        for (int i = 0; i < nw; ++i) kf->P[(size_t)i * nw + i] = p0;
    } else {
        for (int l = 1; l < nlayer; ++l) {
            kf->nb += dims[l];
            if (dims[l - 1] + 1 > kf->maxb) kf->maxb = dims[l - 1] + 1;
        }
        kf->bsz = (int *)malloc(sizeof(int) * kf->nb);
        kf->boff = (int *)malloc(sizeof(int) * kf->nb);
        int ib = 0, m = 0;
        for (int l = 1; l < nlayer; ++l)
            for (int j = 0; j < dims[l]; ++j) {
                kf->bsz[ib] = dims[l - 1] + 1;
                kf->boff[ib] = m;
                m += kf->bsz[ib++];
            }
        const size_t bb = (size_t)kf->maxb * kf->maxb;
        kf->P = (double *)calloc(bb * kf->nb, sizeof(double));
        for (int b = 0; b < kf->nb; ++b)
            for (int j = 0; j < kf->bsz[b]; ++j) kf->P[b * bb + (size_t)j * kf->maxb + j] = p0;
    }
    kf->jf = (double *)calloc(nw, sizeof(double));
    kf->pj = (double *)calloc(nw, sizeof(double));
    kf->kv = (double *)calloc(nw, sizeof(double));
    kf->lam = lam_ini;
    kf->lam0 = lam0;
    return h;
}

/* the gate and q fields (:62, :70), set after kf_init as
 * app/lib_net_module.f90:474-475 does */
int ot_kf_set(int h, double gate, double q)
{
    ot_kf *kf = kf_get(h);
    if (!kf) return -1;
    kf->gate = gate;
    kf->q = q;
    return 0;
}

int ot_kf_free(int h)
{
    ot_kf *kf = kf_get(h);
    if (!kf) return -1;
    free(kf->bsz); free(kf->boff); free(kf->P); free(kf->jf); free(kf->pj); free(kf->kv);
    free(kf->wprior); free(kf->wcur);
    memset(kf, 0, sizeof(*kf));
    return 0;
}

/* pj = P jf: per block (:389-396) or dsymv_like (:650-659) */
static void kf_pj(ot_kf *kf)
{
    const int nw = kf->nw;
    if (kf->decoupled) {
        const size_t bb = (size_t)kf->maxb * kf->maxb;
        for (int b = 0; b < kf->nb; ++b) {
            const int n = kf->bsz[b], o = kf->boff[b];
            const double *Pb = kf->P + b * bb;
            for (int i = 0; i < n; ++i) {
                double s = 0.0;
                for (int r = 0; r < n; ++r) s += Pb[(size_t)i * kf->maxb + r] * kf->jf[o + r];
                kf->pj[o + i] = s;
            }
        }
    } else {
        for (int i = 0; i < nw; ++i) {
            double s = 0.0;
            for (int r = 0; r < nw; ++r) s += kf->P[(size_t)i * nw + r] * kf->jf[r];
            kf->pj[i] = s;
        }
    }
}

/* P = sym((P - kv pj^T)/lam) + q I, per block (:447-464) or dense (:466-487) */
static void kf_downdate(ot_kf *kf)
{
    const int dec = kf->decoupled;
    const int nblk = dec ? kf->nb : 1;
    const size_t ld = dec ? (size_t)kf->maxb : (size_t)kf->nw;
    for (int b = 0; b < nblk; ++b) {
        const int n = dec ? kf->bsz[b] : kf->nw, o = dec ? kf->boff[b] : 0;
        double *P = kf->P + (dec ? b * ld * ld : 0);
        for (int i = 0; i < n; ++i)
            for (int r = 0; r < n; ++r)
                P[i * ld + r] = (P[i * ld + r] - kf->kv[o + r] * kf->pj[o + i]) / kf->lam;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                /* P(i,j) = 0.5 (P(i,j) + P(j,i)); P(j,i) = P(i,j) */
                P[j * ld + i] = 0.5 * (P[j * ld + i] + P[i * ld + j]);
                P[i * ld + j] = P[j * ld + i];
            }
        if (kf->q > 0.0)
            for (int i = 0; i < n; ++i) P[i * ld + i] = P[i * ld + i] + kf->q;
    }
}

/* kf_update_grad :353-493: w[nw] is moved in place; g is dy/dw.
 * Returns 0, or -2 for a nonpositive innovation denominator (where the
 * reference stops, :399-405). */
int ot_kf_update_grad(int h, double *w, const double *g, double xi_in, double scale)
{
    ot_kf *kf = kf_get(h);
    if (!kf) return -1;
    const int nw = kf->nw;
    double xi = scale * xi_in;
    for (int m = 0; m < nw; ++m) kf->jf[m] = scale * g[m];
    kf_pj(kf);
    double dot = 0.0;
    for (int m = 0; m < nw; ++m) dot += kf->jf[m] * kf->pj[m];
    double denom = kf->lam + dot;
    if (denom <= 0.0) return -2;
    if (kf->gate > 0.0) {
        const double jpj = denom - kf->lam;
        if (xi * xi > kf->gate * kf->gate * denom) {
            const double s2 = kf->gate * kf->gate * kf->lam / (xi * xi - kf->gate * kf->gate * jpj);
            const double s = sqrt(s2);
            xi = s * xi;
            for (int m = 0; m < nw; ++m) kf->jf[m] = s * kf->jf[m];
            for (int m = 0; m < nw; ++m) kf->pj[m] = s * kf->pj[m];
            denom = kf->lam + s2 * jpj;
            kf->ngated++;
        }
    }
    for (int m = 0; m < nw; ++m) kf->kv[m] = kf->pj[m] / denom;
    for (int m = 0; m < nw; ++m) w[m] = w[m] + kf->kv[m] * xi;
    kf_downdate(kf);
    kf->lam = kf->lam0 * kf->lam + 1.0 - kf->lam0;
    return 0;
}

/* kf_iekf_begin :512-528 */
int ot_kf_iekf_begin(int h, const double *w)
{
    ot_kf *kf = kf_get(h);
    if (!kf) return -1;
    if (!kf->wprior) {
        kf->wprior = (double *)malloc(sizeof(double) * kf->nw);
        kf->wcur = (double *)malloc(sizeof(double) * kf->nw);
    }
    memcpy(kf->wprior, w, sizeof(double) * kf->nw);
    return 0;
}

/* kf_iekf_iter :532-648: g = dR/dw at the current weights w, resid = R there */
int ot_kf_iekf_iter(int h, double *w, const double *g, double resid, double rnoise, int final)
{
    ot_kf *kf = kf_get(h);
    if (!kf || !kf->wprior || rnoise <= 0.0) return -1;
    const int nw = kf->nw;
    const double scale = 1.0 / sqrt(rnoise);
    for (int m = 0; m < nw; ++m) {
        kf->jf[m] = scale * g[m];
        kf->wcur[m] = w[m];
    }
    kf_pj(kf);
    double dot = 0.0;
    for (int m = 0; m < nw; ++m) dot += kf->jf[m] * kf->pj[m];
    const double denom = kf->lam + dot;
    if (denom <= 0.0) return -2;
    for (int m = 0; m < nw; ++m) kf->kv[m] = kf->pj[m] / denom;
    double c = 0.0;
    for (int m = 0; m < nw; ++m) c += kf->jf[m] * (kf->wprior[m] - kf->wcur[m]);
    const double xi = -scale * resid - c;
    for (int m = 0; m < nw; ++m) w[m] = kf->wprior[m] + kf->kv[m] * xi;
    if (final) {
        kf_downdate(kf);
        kf->lam = kf->lam0 * kf->lam + 1.0 - kf->lam0;
    }
    return 0;
}

/* covariance size in doubles: nw*nw (dense) or maxb*maxb*nb (blocks) */
long ot_kf_cov_len(int h)
{
    ot_kf *kf = kf_get(h);
    if (!kf) return -1;
    return kf->decoupled ? (long)kf->maxb * kf->maxb * kf->nb : (long)kf->nw * kf->nw;
}

/* P (column-major, as the reference holds it), lam and the gated count */
int ot_kf_state(int h, double *P, double *lam, int *ngated)
{
    ot_kf *kf = kf_get(h);
    if (!kf) return -1;
    if (P) memcpy(P, kf->P, sizeof(double) * (size_t)ot_kf_cov_len(h));
    if (lam) *lam = kf->lam;
    if (ngated) *ngated = kf->ngated;
    return 0;
}
