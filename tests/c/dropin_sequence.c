/*
 * A C caller of libnpbcg that keeps the reference benchmark's own call
 * sequence (GMAP/NPB-GPU CUDA/CG/cg.cu main):
 *   :406-449  generator (tran = seed, one discarded draw), column shift, x = 1,
 *             q = z = r = p = 0;
 *   :458-482  one untimed iteration through the serial-signature conj_grad on
 *             host arrays (here npbcg_compat_conj_grad, the exact signature),
 *             then norm_temp1 = x.z, norm_temp2 = 1/sqrt(z.z), x = norm_temp2 z
 *             on the host;
 *   :485-490  x = 1 again, setup_gpu (here npbcg_create);
 *   :500-519  NITER times: conj_grad_gpu (npbcg_conj_grad_device),
 *             gpu_kernel_ten (npbcg_norms), zeta = shift + 1/norm_temp1,
 *             gpu_kernel_eleven (npbcg_scale_x);
 *   :642      release_gpu (npbcg_destroy).
 * Output, one line each, values as C99 hex floats so nothing is lost:
 *   warmup <rnorm> <norm_temp1> <norm_temp2>
 *   it <k> <rnorm> <zeta>          (k = 1 .. niter)
 *   done <status>
 * Usage: dropin_sequence n nonzer niter shift seed
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "npbcg.h"

static int fail(const char *where, int rc) {
    printf("error %s %d %s\n", where, rc, npbcg_last_error());
    printf("done %d\n", rc);
    return 1;
}

int main(int argc, char **argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: %s n nonzer niter shift seed\n", argv[0]);
        return 2;
    }
    const int n = atoi(argv[1]);
    const int nonzer = atoi(argv[2]);
    const int niter = atoi(argv[3]);
    const double shift = atof(argv[4]);
    const double seed = atof(argv[5]);
    if (n < 1 || nonzer < 0 || niter < 1) return 2;

    const long cap = (long)n * (nonzer + 1) * (nonzer + 1);
    int *rowstr = malloc(sizeof(int) * (size_t)(n + 1));
    int *colidx = malloc(sizeof(int) * (size_t)cap);
    double *a = malloc(sizeof(double) * (size_t)cap);
    double *v[5];
    for (int i = 0; i < 5; i++) v[i] = calloc((size_t)n + 2, sizeof(double));
    if (!rowstr || !colidx || !a || !v[0] || !v[1] || !v[2] || !v[3] || !v[4]) return 3;
    double *x = v[0], *z = v[1], *p = v[2], *q = v[3], *r = v[4];

    const long nnz = npbcg_makea(n, nonzer, 0.1, shift, seed, cap, rowstr, colidx, a, NULL);
    if (nnz < 0) return fail("npbcg_makea", (int)nnz);
    for (int i = 0; i < n + 1; i++) x[i] = 1.0;

    /* cg.cu:458-482, the untimed iteration on host arrays */
    double rnorm = 0.0;
    int rc = npbcg_compat_set_n(n);
    if (rc != NPBCG_OK) return fail("npbcg_compat_set_n", rc);
    npbcg_compat_conj_grad(colidx, rowstr, x, z, a, p, q, r, &rnorm);
    if (isnan(rnorm)) return fail("npbcg_compat_conj_grad", NPBCG_ERR_DEVICE);
    double norm_temp1 = 0.0, norm_temp2 = 0.0;
    for (int j = 0; j < n; j++) {
        norm_temp1 += x[j] * z[j];
        norm_temp2 += z[j] * z[j];
    }
    norm_temp2 = 1.0 / sqrt(norm_temp2);
    for (int j = 0; j < n; j++) x[j] = norm_temp2 * z[j];
    printf("warmup %a %a %a\n", rnorm, norm_temp1, norm_temp2);
    npbcg_compat_release();

    /* cg.cu:485-490: x = 1, setup_gpu */
    for (int i = 0; i < n + 1; i++) x[i] = 1.0;
    npbcg_ctx *ctx = NULL;
    rc = npbcg_create(&ctx, n, rowstr, colidx, a, NULL);
    if (rc != NPBCG_OK) return fail("npbcg_create", rc);
    rc = npbcg_set_vector(ctx, NPBCG_VEC_X, x);
    if (rc != NPBCG_OK) return fail("npbcg_set_vector", rc);

    /* cg.cu:500-519, the timed loop */
    for (int it = 1; it <= niter; it++) {
        rc = npbcg_conj_grad_device(ctx, &rnorm);
        if (rc != NPBCG_OK) return fail("npbcg_conj_grad_device", rc);
        rc = npbcg_norms(ctx, &norm_temp1, &norm_temp2);
        if (rc != NPBCG_OK) return fail("npbcg_norms", rc);
        norm_temp2 = 1.0 / sqrt(norm_temp2);
        const double zeta = shift + 1.0 / norm_temp1;
        printf("it %d %a %a\n", it, rnorm, zeta);
        rc = npbcg_scale_x(ctx, norm_temp2);
        if (rc != NPBCG_OK) return fail("npbcg_scale_x", rc);
    }
    npbcg_destroy(ctx);
    printf("done 0\n");
    for (int i = 0; i < 5; i++) free(v[i]);
    free(rowstr);
    free(colidx);
    free(a);
    return 0;
}
