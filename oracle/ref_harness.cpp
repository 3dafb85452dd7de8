/*
 * Harness around the reference's own serial CG code — TEST INFRASTRUCTURE.
 *
 * Built only by oracle/Makefile (target `ref`) in a container that has the
 * reference tree, with outputs only under oracle/_ref/.  It compiles the
 * reference translation unit HIP/CG/cg.cpp as it lies (its `main` renamed)
 * so the file-static makea / conj_grad become callable, and exposes them
 * through a small C ABI used to pin the oracle and to write golden vectors.
 * Problem sizes are compile-time macros there (npbparams.hpp, produced by the
 * reference's own setparams tool), so one library is built per NPB class.
 * Nothing here launches a GPU kernel: only the serial functions are called.
 */
#define main npb_reference_main
#include REF_CG_SOURCE
#undef main

#include <cstring>

extern "C" {

/* The compile-time class parameters (setparams.cpp:751-785). */
void ref_params(int *na, int *nonzer, int *niter, double *shift,
                double *rcond)
{
    *na = NA;
    *nonzer = NONZER;
    *niter = NITER;
    *shift = SHIFT;
    *rcond = RCOND;
}

/* Steps of cg.cu main :366-438 with tran set to `seed`: decomposition
 * globals, RNG init with one discarded draw, makea, the column shift.
 * Copies the CSR (0-based) out; returns nnz or -1 if `cap` is too small. */
long ref_makea(double seed, int *rowstr_out, int *colidx_out, double *a_out,
               long cap)
{
    firstrow = 0;
    lastrow = NA - 1;
    firstcol = 0;
    lastcol = NA - 1;
    naa = NA;
    nzz = NZ;
    tran = seed;
    amult = 1220703125.0;
    (void)randlc(&tran, amult);
    makea(naa, nzz, a, colidx, rowstr, firstrow, lastrow, firstcol, lastcol,
          arow, (int(*)[NONZER + 1])(void *)acol,
          (double(*)[NONZER + 1])(void *)aelt, iv);
    for (int j = 0; j < lastrow - firstrow + 1; j++)
        for (int k = rowstr[j]; k < rowstr[j + 1]; k++)
            colidx[k] = colidx[k] - firstcol;
    long nnz = rowstr[NA];
    if (nnz > cap)
        return -1;
    std::memcpy(rowstr_out, rowstr, sizeof(int) * (NA + 1));
    std::memcpy(colidx_out, colidx, sizeof(int) * nnz);
    std::memcpy(a_out, a, sizeof(double) * nnz);
    return nnz;
}

/* The reference's serial conj_grad (cg.cu:653-801) on the matrix left by
 * the last ref_makea call.  x, z, p, q, r need NA+1 entries. */
void ref_conj_grad(double *x_io, double *z_io, double *p_io, double *q_io,
                   double *r_io, double *rnorm)
{
    std::memcpy(x, x_io, sizeof(double) * (NA + 1));
    conj_grad(colidx, rowstr, x, z, a, p, q, r, rnorm);
    std::memcpy(z_io, z, sizeof(double) * (NA + 1));
    std::memcpy(p_io, p, sizeof(double) * (NA + 1));
    std::memcpy(q_io, q, sizeof(double) * (NA + 1));
    std::memcpy(r_io, r, sizeof(double) * (NA + 1));
}

/* The serial inverse-power loop on the matrix left by ref_makea: x = 1
 * (cg.cu:441-443), then `niter` times conj_grad followed by the serial
 * norm / normalise step of :470-481 and zeta = SHIFT + 1/(x.z) (:514).
 * Writes the per-iteration zeta and rnorm, and the final x (NA entries). */
void ref_power(int niter, double *zeta_out, double *rnorm_out, double *x_out)
{
    for (int i = 0; i < NA + 1; i++)
        x[i] = 1.0;
    for (int it = 0; it < niter; it++) {
        double rnorm, t1 = 0.0, t2 = 0.0;
        conj_grad(colidx, rowstr, x, z, a, p, q, r, &rnorm);
        for (int j = 0; j < lastcol - firstcol + 1; j++) {
            t1 = t1 + x[j] * z[j];
            t2 = t2 + z[j] * z[j];
        }
        t2 = 1.0 / sqrt(t2);
        zeta_out[it] = SHIFT + 1.0 / t1;
        rnorm_out[it] = rnorm;
        for (int j = 0; j < lastcol - firstcol + 1; j++)
            x[j] = t2 * z[j];
    }
    if (x_out)
        std::memcpy(x_out, x, sizeof(double) * NA);
}

} /* extern "C" */
