/*
 * CPU oracle for the NPB CG conj_grad path — TEST INFRASTRUCTURE ONLY.
 * Plain-C restatement of the reference's serial algorithm; the checker, never
 * the thing measured or shipped.  See npbcg_oracle.h for what pins each
 * function.  Reference citations are file:line in GMAP/NPB-GPU.
 */
#include "npbcg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* 2^23, 2^-23, 2^46, 2^-46 as exact doubles (c_randdp.cpp:43-50). */
static const double P23 = 8388608.0;
static const double R23 = 1.0 / 8388608.0;
static const double P46 = 70368744177664.0;
static const double R46 = 1.0 / 70368744177664.0;

/* c_randdp.cpp:84-114.  The product a*x mod 2^46 is formed from 23-bit
 * halves so every intermediate is an exact FP64 integer. */
double orc_randlc(double *x, double a)
{
    double a_hi = (double)(int)(R23 * a);
    double a_lo = a - P23 * a_hi;
    double x_hi = (double)(int)(R23 * (*x));
    double x_lo = (*x) - P23 * x_hi;
    double cross = a_hi * x_lo + a_lo * x_hi;          /* < 2^47, exact */
    double cross_mod = cross - P23 * (double)(int)(R23 * cross);
    double full = P23 * cross_mod + a_lo * x_lo;        /* < 2^47, exact */
    *x = full - P46 * (double)(int)(R46 * full);
    return R46 * (*x);
}

static const double NPB_AMULT = 1220703125.0; /* 5^13, cg.cu:407 */

/* sprnvc, cg.cu:1798-1833: nzwant distinct positions in [1, n] drawn as
 * (int)(nn1*u)+1 (icnvrt, :1272-1274), each paired with the value drawn just
 * before it.  Positions above n and repeats are rejected. */
static void draw_sparse_vector(double *tran, int n, int nzwant, int nn1,
                               double *val, int *pos)
{
    int have = 0;
    while (have < nzwant) {
        double v = orc_randlc(tran, NPB_AMULT);
        double u = orc_randlc(tran, NPB_AMULT);
        int p = (int)(nn1 * u) + 1;
        if (p > n)
            continue;
        int dup = 0;
        for (int t = 0; t < have; t++)
            if (pos[t] == p) { dup = 1; break; }
        if (dup)
            continue;
        val[have] = v;
        pos[have] = p;
        have++;
    }
}

/* vecset, cg.cu:1841-1857: set position p to v, appending if absent. */
static void set_vector_entry(double *val, int *pos, int *count, int p, double v)
{
    int found = 0;
    for (int t = 0; t < *count; t++)
        if (pos[t] == p) { val[t] = v; found = 1; }
    if (!found) {
        val[*count] = v;
        pos[*count] = p;
        (*count)++;
    }
}

int orc_makea(int n, int nonzer, double rcond, double shift, double seed,
              long nz_cap, int *rowstr, int *colidx, double *a,
              double *ratio_out)
{
    /* nonzer <= n: sprnvc (cg.cu:1798-1833) draws nonzer distinct positions
       in 1..n, which terminates for any nonzer up to n. */
    if (n < 1 || nonzer < 0 || nonzer > n || nz_cap < 1 || !rowstr ||
        !colidx || !a)
        return -1;
    const int w = nonzer + 1; /* entries per outer-product vector */

    /* cg.cu:406-408: tran = seed, then one discarded draw. */
    double tran = seed;
    (void)orc_randlc(&tran, NPB_AMULT);

    /* makea, cg.cu:1322-1346. */
    int nn1 = 1;
    do { nn1 *= 2; } while (nn1 < n);

    int *vlen = (int *)malloc(sizeof(int) * (size_t)n);
    int *vcol = (int *)malloc(sizeof(int) * (size_t)n * w);
    double *vval = (double *)malloc(sizeof(double) * (size_t)n * w);
    int *dupcount = (int *)malloc(sizeof(int) * (size_t)n);
    if (!vlen || !vcol || !vval || !dupcount) {
        free(vlen); free(vcol); free(vval); free(dupcount);
        return -4;
    }
    for (int i = 0; i < n; i++) {
        double val[64];
        int pos[64];
        int cnt = nonzer;
        if (w > 64) { free(vlen); free(vcol); free(vval); free(dupcount); return -1; }
        draw_sparse_vector(&tran, n, nonzer, nn1, val, pos);
        set_vector_entry(val, pos, &cnt, i + 1, 0.5);
        vlen[i] = cnt;
        for (int t = 0; t < cnt; t++) {
            vcol[(size_t)i * w + t] = pos[t] - 1;
            vval[(size_t)i * w + t] = val[t];
        }
    }

    /* sparse, cg.cu:1641-1663: per-row slot counts (one slot per triple
     * landing in the row), prefix-summed into row starts. */
    for (int j = 0; j <= n; j++)
        rowstr[j] = 0;
    for (int i = 0; i < n; i++)
        for (int t = 0; t < vlen[i]; t++)
            rowstr[vcol[(size_t)i * w + t] + 1] += vlen[i];
    for (int j = 1; j <= n; j++)
        rowstr[j] += rowstr[j - 1];
    if ((long)rowstr[n] - 1 > nz_cap) {  /* cg.cu:1670 compares nza > nz */
        free(vlen); free(vcol); free(vval); free(dupcount);
        return -2;
    }
    /* cg.cu:1681-1687: clear slots, mark them empty with column -1. */
    for (int j = 0; j < n; j++) {
        for (int k = rowstr[j]; k < rowstr[j + 1]; k++) {
            a[k] = 0.0;
            colidx[k] = -1;
        }
        dupcount[j] = 0;
    }

    /* cg.cu:1694-1756: accumulate the scaled outer products.  Row j of
     * vector i receives aelt[i][r]*(size*aelt[i][c]); the diagonal of the
     * i-th vector also gets rcond - shift.  Columns are kept sorted by
     * insertion; a repeated column adds onto the existing slot. */
    double ratio = pow(rcond, 1.0 / (double)n);
    if (ratio_out)
        *ratio_out = ratio;
    double size = 1.0;
    for (int i = 0; i < n; i++) {
        const int *ci = vcol + (size_t)i * w;
        const double *vi = vval + (size_t)i * w;
        for (int s = 0; s < vlen[i]; s++) {
            int j = ci[s];
            double scale = size * vi[s];
            for (int t = 0; t < vlen[i]; t++) {
                int jcol = ci[t];
                double va = vi[t] * scale;
                if (jcol == j && j == i)
                    va = va + rcond - shift;
                int k, placed = 0;
                for (k = rowstr[j]; k < rowstr[j + 1]; k++) {
                    if (colidx[k] > jcol) {
                        /* open a hole at k by moving occupied slots right */
                        for (int kk = rowstr[j + 1] - 2; kk >= k; kk--) {
                            if (colidx[kk] > -1) {
                                a[kk + 1] = a[kk];
                                colidx[kk + 1] = colidx[kk];
                            }
                        }
                        colidx[k] = jcol;
                        a[k] = 0.0;
                        placed = 1;
                        break;
                    } else if (colidx[k] == -1) {
                        colidx[k] = jcol;
                        placed = 1;
                        break;
                    } else if (colidx[k] == jcol) {
                        dupcount[j]++;
                        placed = 1;
                        break;
                    }
                }
                if (!placed) {
                    free(vlen); free(vcol); free(vval); free(dupcount);
                    return -3;
                }
                a[k] = a[k] + va;
            }
        }
        size = size * ratio;
    }

    /* cg.cu:1763-1784: squeeze out the unused slots of every row. */
    for (int j = 1; j < n; j++)
        dupcount[j] += dupcount[j - 1];
    for (int j = 0; j < n; j++) {
        int dst0 = (j > 0) ? rowstr[j] - dupcount[j - 1] : 0;
        int dst1 = rowstr[j + 1] - dupcount[j];
        int src = rowstr[j];
        for (int k = dst0; k < dst1; k++, src++) {
            a[k] = a[src];
            colidx[k] = colidx[src];
        }
    }
    for (int j = 1; j <= n; j++)
        rowstr[j] -= dupcount[j - 1];

    free(vlen); free(vcol); free(vval); free(dupcount);
    return rowstr[n];
}

void orc_spmv(int n, const int *rowstr, const int *colidx, const double *a,
              const double *x, double *y)
{
    for (int j = 0; j < n; j++) {
        double s = 0.0;
        for (int k = rowstr[j]; k < rowstr[j + 1]; k++)
            s = s + a[k] * x[colidx[k]];
        y[j] = s;
    }
}

void orc_conj_grad(int n, int cgitmax, const int *colidx, const int *rowstr,
                   const double *x, double *z, const double *a, double *p,
                   double *q, double *r, double *rnorm)
{
    /* cg.cu:671-676: initialise through index n inclusive. */
    for (int j = 0; j <= n; j++) {
        q[j] = 0.0;
        z[j] = 0.0;
        r[j] = x[j];
        p[j] = r[j];
    }
    double rho = 0.0;                       /* cg.cu:684-686 */
    for (int j = 0; j < n; j++)
        rho = rho + r[j] * r[j];

    for (int it = 1; it <= cgitmax; it++) { /* cg.cu:689-772 */
        orc_spmv(n, rowstr, colidx, a, p, q);
        double d = 0.0;
        for (int j = 0; j < n; j++)
            d = d + p[j] * q[j];
        double alpha = rho / d;
        double rho0 = rho;
        for (int j = 0; j < n; j++) {
            z[j] = z[j] + alpha * p[j];
            r[j] = r[j] - alpha * q[j];
        }
        rho = 0.0;
        for (int j = 0; j < n; j++)
            rho = rho + r[j] * r[j];
        double beta = rho / rho0;
        for (int j = 0; j < n; j++)
            p[j] = r[j] + beta * p[j];
    }

    /* cg.cu:782-800: r = A z, rnorm = ||x - r||. */
    orc_spmv(n, rowstr, colidx, a, z, r);
    double sum = 0.0;
    for (int j = 0; j < n; j++) {
        double d = x[j] - r[j];
        sum = sum + d * d;
    }
    *rnorm = sqrt(sum);
}

void orc_power(int n, int niter, int cgitmax, double shift,
               const int *colidx, const int *rowstr, const double *a,
               double *x, double *z, double *p, double *q, double *r,
               double *zeta_out, double *rnorm_out)
{
    for (int it = 0; it < niter; it++) {
        double rnorm;
        orc_conj_grad(n, cgitmax, colidx, rowstr, x, z, a, p, q, r, &rnorm);
        /* cg.cu:470-476: x.z and z.z in one index-ordered pass. */
        double xz = 0.0, zz = 0.0;
        for (int j = 0; j < n; j++) {
            xz = xz + x[j] * z[j];
            zz = zz + z[j] * z[j];
        }
        double inv_norm = 1.0 / sqrt(zz);
        double zeta = shift + 1.0 / xz;     /* cg.cu:514 */
        if (zeta_out) zeta_out[it] = zeta;
        if (rnorm_out) rnorm_out[it] = rnorm;
        for (int j = 0; j < n; j++)         /* cg.cu:479-481 */
            x[j] = inv_norm * z[j];
    }
}

static double dot_n(int n, const double *u, const double *v)
{
    double s = 0.0;
    for (int j = 0; j < n; j++)
        s = s + u[j] * v[j];
    return s;
}

int orc_cg_solve(int n, const int *rowstr, const int *colidx, const double *a,
                 const double *b, double *x, double rtol, int maxit,
                 double *relres, double *work)
{
    double *r = work, *p = work + n, *q = work + 2 * (size_t)n,
           *t = work + 3 * (size_t)n;
    const int check = 25;
    double bnorm = sqrt(dot_n(n, b, b));
    for (int j = 0; j < n; j++) {
        x[j] = 0.0;
        r[j] = b[j];
        p[j] = b[j];
    }
    if (bnorm == 0.0) {
        *relres = 0.0;
        return 0;
    }
    double rho = dot_n(n, r, r);
    int it = 0;
    double true_rel = 1.0;
    while (it < maxit) {
        orc_spmv(n, rowstr, colidx, a, p, q);
        double d = dot_n(n, p, q);
        double alpha = rho / d;
        double rho0 = rho;
        for (int j = 0; j < n; j++) {
            x[j] = x[j] + alpha * p[j];
            r[j] = r[j] - alpha * q[j];
        }
        rho = dot_n(n, r, r);
        it++;
        if (sqrt(rho) <= rtol * bnorm || it % check == 0) {
            orc_spmv(n, rowstr, colidx, a, x, t);
            double s = 0.0;
            for (int j = 0; j < n; j++) {
                double e = b[j] - t[j];
                s = s + e * e;
            }
            true_rel = sqrt(s) / bnorm;
            if (true_rel <= rtol)
                break;
        }
        double beta = rho / rho0;
        for (int j = 0; j < n; j++)
            p[j] = r[j] + beta * p[j];
    }
    *relres = true_rel;
    return it;
}

/* Diagonal-preconditioned CG (PCG, M^-1 = diag(minv)): z = M^-1 r,
 * rho = r.z, p = z + beta p.  The true residual is checked when
 * |rho| <= rtol^2 |rho_0| (the preconditioned norm test) and every 25 steps;
 * the loop ends when ||b - A x|| <= rtol ||b||.  work holds 5n doubles. */
int orc_pcg_solve(int n, const int *rowstr, const int *colidx, const double *a,
                  const double *minv, const double *b, double *x, double rtol,
                  int maxit, double *relres, double *work)
{
    double *r = work, *p = work + n, *q = work + 2 * (size_t)n,
           *t = work + 3 * (size_t)n, *z = work + 4 * (size_t)n;
    const int check = 25;
    double bnorm = sqrt(dot_n(n, b, b));
    for (int j = 0; j < n; j++) {
        x[j] = 0.0;
        r[j] = b[j];
        z[j] = minv[j] * b[j];
        p[j] = z[j];
    }
    if (bnorm == 0.0) {
        *relres = 0.0;
        return 0;
    }
    double rho = dot_n(n, r, z);
    const double rho_first = fabs(rho);
    int it = 0;
    double true_rel = 1.0;
    while (it < maxit) {
        orc_spmv(n, rowstr, colidx, a, p, q);
        double d = dot_n(n, p, q);
        double alpha = rho / d;
        double rho0 = rho;
        for (int j = 0; j < n; j++) {
            x[j] = x[j] + alpha * p[j];
            r[j] = r[j] - alpha * q[j];
            z[j] = minv[j] * r[j];
        }
        rho = dot_n(n, r, z);
        it++;
        if (fabs(rho) <= rtol * rtol * rho_first || it % check == 0) {
            orc_spmv(n, rowstr, colidx, a, x, t);
            double s = 0.0;
            for (int j = 0; j < n; j++) {
                double e = b[j] - t[j];
                s = s + e * e;
            }
            true_rel = sqrt(s) / bnorm;
            if (true_rel <= rtol)
                break;
        }
        double beta = rho / rho0;
        for (int j = 0; j < n; j++)
            p[j] = z[j] + beta * p[j];
    }
    *relres = true_rel;
    return it;
}

void orc_make_rhs(int n, double seed, double *b)
{
    double t = seed + 1.0;
    for (int j = 0; j < n; j++)
        b[j] = 2.0 * orc_randlc(&t, NPB_AMULT) - 1.0;
}
