// Host-side exercise of the library under AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_host_sanitizers.py builds the
// library's host code with both and runs this on the CPU): table families
// of every shape the tables accept, the Faa di Bruno kernel source generated
// for each family small enough to unroll, the handle registries filled to
// their 64 slots and reused, the current-family queries, networks built,
// written, read back and refused, and the argument checks of the entries
// that need no device.  Exit status = number of failed checks.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dnnf90.h"
#include "dnnf90_amd.h"
#include "../../dnnf90_amd/csrc/codegen.hpp"

static int nfail = 0;
#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            ++nfail;                                               \
        }                                                          \
    } while (0)

// the generated kernel source of the family, FP64 and FP32
static void source(int d0, int k, int nseed, const int *seeds)
{
    dnnf90::TableSet t;
    std::string err;
    CHECK(t.build(d0, k, nseed, seeds, &err));
    dnnf90::FwdSchedule s;
    s.build(t);
    if (!dnnf90::gen_bell_usable(t, s)) return;
    for (int fp64 = 0; fp64 < 2; ++fp64) {
        const std::string src = dnnf90::gen_bell_source(t, s, fp64 != 0);
        CHECK(src.find("dnnf90_bell_fwd") != std::string::npos);
        CHECK(src.find("dnnf90_bell_bwds") != std::string::npos);
    }
    // the fused layer kernels of the family (FP32 collocation step)
    if (dnnf90::gen_fused_usable(t, s)) {
        const std::string src = dnnf90::gen_fused_source(t, s);
        for (const char *k : {"dnnf90_fz_fwd_in", "dnnf90_fz_fwd", "dnnf90_fz_fwd_last", "dnnf90_fz_bwd_last",
                              "dnnf90_fz_bwd", "dnnf90_fz_bwd_in"})
            CHECK(src.find(k) != std::string::npos);
    }
}

static void lists(int tid)
{
    int c[4];
    CHECK(dnnf90_tabset_counts(tid, c) == 0);
    int na = c[0], np = c[1], nf = c[2], k = c[3];
    int *ps = (int *)malloc(sizeof(int) * na), *pn = (int *)malloc(sizeof(int) * na);
    int *pia = (int *)malloc(sizeof(int) * (np + 1)), *pid = (int *)malloc(sizeof(int) * (np + 1));
    double *pb = (double *)malloc(sizeof(double) * (np + 1));
    int *fs = (int *)malloc(sizeof(int) * na * (k + 1)), *fn = (int *)malloc(sizeof(int) * na * (k + 1));
    int *fib = (int *)malloc(sizeof(int) * (nf + 1)), *fid = (int *)malloc(sizeof(int) * (nf + 1));
    double *fc = (double *)malloc(sizeof(double) * (nf + 1));
    CHECK(dnnf90_tabset_lists(tid, ps, pn, pia, pid, pb, fs, fn, fib, fid, fc) == 0);
    for (int i = 0; i < np; ++i) CHECK(pia[i] >= 0 && pia[i] < na && pid[i] >= 0 && pid[i] < na && pb[i] >= 1.0);
    for (int i = 0; i < nf; ++i) CHECK(fib[i] >= 0 && fib[i] < na && fid[i] >= 0 && fid[i] < na);
    free(ps); free(pn); free(pia); free(pid); free(pb); free(fs); free(fn); free(fib); free(fid); free(fc);
}

int main(int argc, char **argv)
{
    const char *tmp = argc > 1 ? argv[1] : "/tmp";
    char path[4096];
    /* table families: dense and closure, every d0 and k the tables take */
    for (int d0 = 1; d0 <= 6; ++d0)
        for (int k = 1; k <= (d0 == 1 ? 15 : 7 - d0 / 2); ++k) {
            int tid = dnnf90_tabset_dense(d0, k);
            CHECK(tid > 0);
            lists(tid);
            source(d0, k, 0, NULL);
            CHECK(dnnf90_tabset_free(tid) == 0);
        }
    int seeds[] = {7, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 1, 3, 2, 2, 0};
    int tc = dnnf90_tabset_closure(4, 7, 4, seeds);
    CHECK(tc > 0 && dnnf90_tabset_nderiv(tc) == 41);
    lists(tc);
    source(4, 7, 4, seeds);
    int kdv[] = {0, 1, 1, 0, 3, 0};
    source(2, 3, 3, kdv);
    snprintf(path, sizeof(path), "%s/hod_alpha_order.dat", tmp);
    CHECK(dnnf90_tabset_write_order(tc, path) == 0);
    /* bad table arguments */
    CHECK(dnnf90_tabset_dense(0, 3) == -1 && dnnf90_tabset_dense(2, 16) == -1 && dnnf90_tabset_dense(2, 0) == -1);
    int bad[] = {-1, 0};
    CHECK(dnnf90_tabset_closure(2, 3, 1, bad) == -1);
    CHECK(dnnf90_tabset_activation(tc, 9) == -2 && dnnf90_tabset_activation(999, 0) == -1);
    /* the registry: 64 slots, then refusals, then reuse */
    int h[70], nh = 0;
    for (int i = 0; i < 70; ++i) {
        int t = dnnf90_tabset_dense(2, 2);
        if (t > 0) h[nh++] = t;
        else CHECK(t == -1);
    }
    CHECK(nh == 63);   /* tc holds one slot */
    for (int i = 0; i < nh; ++i) CHECK(dnnf90_tabset_free(h[i]) == 0);
    CHECK(dnnf90_tabset_free(h[0]) == -1);
    /* the current family */
    CHECK(dnnf90_tables_init_closure(4, 7, 4, seeds) == 0 && dnnf90_nderiv() == 41);
    int a[8];
    for (int ia = 1; ia <= 41; ++ia) CHECK(dnnf90_alpha(ia, a) == 0);
    CHECK(dnnf90_alpha(0, a) == -1 && dnnf90_alpha(42, a) == -1);
    /* networks: shapes, weights, files */
    int dims[] = {4, 9, 130, 2};
    int nid = dnnf90_net_create_ts(tc, 4, dims);
    CHECK(nid > 0);
    int nw = dnnf90_net_nweights(nid);
    CHECK(nw == 9 * 5 + 130 * 10 + 2 * 131);
    double *w = (double *)malloc(sizeof(double) * nw), *w2 = (double *)malloc(sizeof(double) * nw);
    for (int i = 0; i < nw; ++i) w[i] = sin(0.37 * i) * 0.5;
    CHECK(dnnf90_net_set_weights(nid, w) == 0 && dnnf90_net_get_weights(nid, w2) == 0);
    CHECK(memcmp(w, w2, sizeof(double) * nw) == 0);
    CHECK(dnnf90_net_set_layer(nid, 3, w) == 0 && dnnf90_net_set_layer(nid, 9, w) == -1);
    snprintf(path, sizeof(path), "%s/nn_weight.dat", tmp);
    CHECK(dnnf90_net_save(nid, path) == 0);
    int nid2 = dnnf90_net_load_ts(tc, path);
    CHECK(nid2 > 0 && dnnf90_net_nweights(nid2) == nw);
    int bad_dims[] = {3, 4, 1};
    CHECK(dnnf90_net_create_ts(tc, 3, bad_dims) == -1);   /* input width != d0 */
    CHECK(dnnf90_net_create_ts(tc, 1, dims) == -1);
    snprintf(path, sizeof(path), "%s/missing.dat", tmp);
    CHECK(dnnf90_net_load_ts(tc, path) == -2);
    snprintf(path, sizeof(path), "%s/garbage.dat", tmp);
    FILE *f = fopen(path, "w");
    fprintf(f, "0\nfunc 0\nActivation_out 0\n3\n4\n5\n");   /* truncated */
    fclose(f);
    CHECK(dnnf90_net_load_ts(tc, path) < 0);
    /* entries refused on bad handles before any device work */
    double x[4] = {0}, t[82] = {0};
    CHECK(dnnf90_eval(nid, 999, x, t) == -1 && dnnf90_eval(0, 1, x, t) == -1);
    CHECK(dnnf90_grad_point(nid, 999, x, t, 1) == -1);
    CHECK(dnnf90_forward_batch(nid, 999, 1, x, t, 0) == -1);
    CHECK(dnnf90_kf_create(999, 1.0, 0.99, 0.999, 0) == -1);
    CHECK(dnnf90_grad_point_noalloc(999, x, t, w, NULL) == -1);
    CHECK(dnnf90_net_free(nid) == 0 && dnnf90_net_free(nid) == -1 && dnnf90_net_free(nid2) == 0);
    CHECK(dnnf90_tabset_free(tc) == 0);
    free(w); free(w2);
    if (nfail) fprintf(stderr, "%d checks failed\n", nfail);
    else printf("host sanitizer driver: all checks passed\n");
    return nfail;
}
