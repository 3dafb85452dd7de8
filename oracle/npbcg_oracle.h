/*
 * CPU oracle for the NPB CG conj_grad path — TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's serial algorithm
 * (GMAP/NPB-GPU, CUDA/CG/cg.cu and CUDA/common/c_randdp.cpp).  It is the
 * checker the GPU library is compared against.  Only tests/, the smoke
 * check and bench.py's cpu_baseline leg may load it; the product library
 * (npb-gpu_amd/) never links, calls or falls back to it.
 *
 * Every function cites the reference lines it follows.  Parity status of
 * each function (what pins it) is listed next to its declaration.
 */
#ifndef NPBCG_ORACLE_H
#define NPBCG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* 46-bit LCG x <- a*x mod 2^46, returned as 2^-46*x.
 * Follows CUDA/common/c_randdp.cpp:84-114 (randlc).
 * Pinned by: exact 128-bit integer closed form a^k*x0 mod 2^46 (tests). */
double orc_randlc(double *x, double a);

/* Generate the NPB CG matrix in CSR form (0-based).
 * Follows the RNG initialisation at CUDA/CG/cg.cu:406-408 (tran=seed, one
 * discarded draw), makea :1303-1368, sprnvc :1798-1833, vecset :1841-1857,
 * icnvrt :1272-1274 and sparse :1611-1785 (insertion with duplicate merge,
 * diagonal shift, geometric row scaling, compaction).
 * rowstr has n+1 entries; colidx/a need nz_cap entries, where the
 * reference uses nz_cap = n*(nonzer+1)^2 (cg.cu:81).
 * Returns nnz (= rowstr[n]) or a negative error code:
 *   -1 invalid argument, -2 nz_cap exceeded (cg.cu:1670-1674),
 *   -3 internal insertion error (cg.cu:1748-1751), -4 out of memory.
 * *ratio_out (may be NULL) receives pow(rcond, 1/n) (cg.cu:1695).
 * Pinned by: NPB zeta constants S/W/A/B/C (cg.cu:373-393) and
 * rowstr/colidx/a hashes of the reference's own makea (tests/golden). */
int orc_makea(int n, int nonzer, double rcond, double shift, double seed,
              long nz_cap, int *rowstr, int *colidx, double *a,
              double *ratio_out);

/* One call of the reference's serial conj_grad, CUDA/CG/cg.cu:653-801:
 * 25 unpreconditioned CG iterations on A z = x from z = 0, then
 * r = A z and rnorm = ||x - A z||.  All sums run in index order.
 * x, z, p, q, r hold at least n+1 entries (the init loop writes index n,
 * cg.cu:671-676).  cgitmax is 25 in the reference; it is a parameter here.
 * Pinned by: NPB zeta constants, the reference's per-iteration
 * (rnorm, zeta) trajectories (tests/golden), dense solves at tiny n. */
void orc_conj_grad(int n, int cgitmax, const int *colidx, const int *rowstr,
                   const double *x, double *z, const double *a, double *p,
                   double *q, double *r, double *rnorm);

/* The inverse-power outer loop, serial form: CUDA/CG/cg.cu:458-482 for the
 * norm/normalise step and :500-520 for the loop and the zeta formula.
 * x (n+1 entries, in/out) is the start vector; z, p, q, r are workspace.
 * zeta_out/rnorm_out (niter entries each, may be NULL) get the
 * per-iteration values printed at cg.cu:515-516.
 * Pinned by: NPB zeta constants; reference trajectories (tests/golden);
 * dense eigenvalue check at tiny n (zeta -> mu* + rcond). */
void orc_power(int n, int niter, int cgitmax, double shift,
               const int *colidx, const int *rowstr, const double *a,
               double *x, double *z, double *p, double *q, double *r,
               double *zeta_out, double *rnorm_out);

/* Row-ordered CSR product y = A x, the loop at CUDA/CG/cg.cu:703-709.
 * Pinned by: dense matrix-vector products at tiny n (tests). */
void orc_spmv(int n, const int *rowstr, const int *colidx, const double *a,
              const double *x, double *y);

/* Plain FP64 CG on A x = b from x = 0 until the FP64 true relative residual
 * ||b - A x|| / ||b|| <= rtol or maxit iterations.  The iteration is the one
 * of cg.cu:689-772 (same update order); the stopping rule is this build's
 * (the reference always runs 25 steps).  Every `check` iterations and when
 * the recursive residual meets rtol, the true residual is recomputed.
 * Returns the iteration count; *relres gets the final true relative
 * residual.  work needs 4*n doubles.
 * Pinned by: dense solve at tiny n (tests); true residual is recomputed. */
int orc_cg_solve(int n, const int *rowstr, const int *colidx, const double *a,
                 const double *b, double *x, double rtol, int maxit,
                 double *relres, double *work);

/* Right-hand side for the solve-to-tolerance configurations:
 * b[i] = 2*u_i - 1 with u_i successive orc_randlc draws from tran = seed+1
 * (no discarded draw).  The reference has no such vector; this is the
 * build's own definition, recorded in the golden fixtures. */
/* Diagonal-preconditioned CG; work holds 5n doubles. */
int orc_pcg_solve(int n, const int *rowstr, const int *colidx, const double *a,
                  const double *minv, const double *b, double *x, double rtol,
                  int maxit, double *relres, double *work);
void orc_make_rhs(int n, double seed, double *b);

#ifdef __cplusplus
}
#endif
#endif
