/*
 * ORACLE — test infrastructure only.
 *
 * Plain-C CPU restatement of the dense linear algebra on the MOLSCAT
 * log-derivative path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker; the product path never links it.
 *
 * Contents (all matrices column-major, leading dimension lda):
 *
 *  orc_syminv_inline   SYMINV for N <= 30: in-situ Wilkinson-Reinsch
 *                      L.D.L^T inverse without pivoting, KOUNT = number of
 *                      negative pivots, N+1 on an exact zero pivot.
 *                      Follows source_code/syminv.f:21-96.
 *  orc_dsytf2_l        Bunch-Kaufman L.D.L^T (lower), the published LAPACK
 *                      DSYTF2 algorithm (LAPACK is an external, unpinned
 *                      dependency of the reference: source_code/Makefile:12;
 *                      called through DSYTRF at source_code/syminl-alloc.f:21).
 *  orc_dsyneg_l        inertia of the BK factor, source_code/dsyneg.f:47-73.
 *  orc_dsytri_l        inverse from the BK factor, published LAPACK DSYTRI
 *                      ('L'), called at source_code/syminl-alloc.f:34.
 *  orc_syminv          SYMINV dispatch (source_code/syminv.f:11-14):
 *                      N <= 30 inline, else SYMINL = DSYTRF+DSYNEG+DSYTRI
 *                      (source_code/syminl-alloc.f:17-41).  On exit the
 *                      lower triangle holds the inverse (the inline path
 *                      also fills the upper triangle, as the reference does).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define A_(i, j) a[(size_t)(i) + (size_t)(j) * (size_t)lda]

/* source_code/syminv.f:21-96 (0-based indices here). */
int orc_syminv_inline(double *a, int lda, int n)
{
    int kount = 0;
    double x = A_(0, 0);
    if (x < 0.0) kount++;
    if (x == 0.0) return n + 1;
    A_(0, 0) = 1.0 / x;
    if (n == 1) return kount;
    /* form L */
    for (int i = 1; i < n; i++) {
        for (int j = 1; j < i; j++) {
            double s = A_(i, j);
            for (int k = 0; k < j; k++) s -= A_(i, k) * A_(j, k);
            A_(i, j) = s;
        }
        double d = A_(i, i);
        for (int k = 0; k < i; k++) {
            double y = A_(i, k);
            double z = y * A_(k, k);
            A_(i, k) = z;
            d -= y * z;
        }
        if (d < 0.0) kount++;
        if (d == 0.0) return n + 1;
        A_(i, i) = 1.0 / d;
    }
    /* invert L */
    for (int i = 0; i < n - 1; i++) {
        A_(i + 1, i) = -A_(i + 1, i);
        for (int j = i + 1; j < n - 1; j++) {
            double s = -A_(j + 1, i);
            for (int k = i + 1; k <= j; k++) s -= A_(j + 1, k) * A_(k, i);
            A_(j + 1, i) = s;
        }
    }
    /* form the inverse */
    for (int i = 0; i < n - 1; i++) {
        double s = A_(i, i);
        for (int k = i + 1; k < n; k++) {
            double y = A_(k, i);
            double z = y * A_(k, k);
            A_(k, i) = z;
            s += y * z;
        }
        A_(i, i) = s;
        for (int j = i + 1; j < n - 1; j++) {
            double t = A_(j, i);
            for (int k = j + 1; k < n; k++) t += A_(k, j) * A_(k, i);
            A_(j, i) = t;
        }
    }
    /* mirror into the upper triangle */
    for (int i = 1; i < n; i++)
        for (int j = 0; j < i; j++) A_(j, i) = A_(i, j);
    return kount;
}

static void swap_d(double *p, double *q)
{
    double t = *p;
    *p = *q;
    *q = t;
}

/* Published LAPACK DSYTF2, UPLO='L'.  ipiv is 1-based as in LAPACK:
 * ipiv[k] = kp (1x1 pivot, rows k and kp interchanged) or
 * ipiv[k] = ipiv[k+1] = -kp (2x2 pivot, rows k+1 and kp interchanged).
 * Returns INFO (0, or the 1-based index of the first exactly zero pivot). */
int orc_dsytf2_l(double *a, int lda, int n, int *ipiv)
{
    const double alpha = (1.0 + sqrt(17.0)) / 8.0;
    int info = 0;
    int k = 0; /* 0-based */
    while (k < n) {
        int kstep = 1, kp;
        double absakk = fabs(A_(k, k));
        int imax = k;
        double colmax = 0.0;
        if (k < n - 1) {
            imax = k + 1;
            colmax = fabs(A_(k + 1, k));
            for (int i = k + 2; i < n; i++)
                if (fabs(A_(i, k)) > colmax) { colmax = fabs(A_(i, k)); imax = i; }
        }
        if (fmax(absakk, colmax) == 0.0 || isnan(absakk)) {
            if (info == 0) info = k + 1;
            kp = k;
        } else {
            if (absakk >= alpha * colmax) {
                kp = k;
            } else {
                /* largest off-diagonal in row/column imax of the trailing matrix */
                int jmax = k;
                double rowmax = fabs(A_(imax, k));
                for (int j = k + 1; j < imax; j++)
                    if (fabs(A_(imax, j)) > rowmax) { rowmax = fabs(A_(imax, j)); jmax = j; }
                for (int j = imax + 1; j < n; j++)
                    if (fabs(A_(j, imax)) > rowmax) { rowmax = fabs(A_(j, imax)); jmax = j; }
                (void)jmax;
                if (absakk >= alpha * colmax * (colmax / rowmax)) {
                    kp = k;
                } else if (fabs(A_(imax, imax)) >= alpha * rowmax) {
                    kp = imax;
                } else {
                    kp = imax;
                    kstep = 2;
                }
            }
            int kk = k + kstep - 1;
            if (kp != kk) {
                for (int i = kp + 1; i < n; i++) swap_d(&A_(i, kk), &A_(i, kp));
                for (int j = kk + 1; j < kp; j++) swap_d(&A_(j, kk), &A_(kp, j));
                swap_d(&A_(kk, kk), &A_(kp, kp));
                if (kstep == 2) swap_d(&A_(k + 1, k), &A_(kp, k));
            }
            if (kstep == 1) {
                if (k < n - 1) {
                    double d11 = 1.0 / A_(k, k);
                    for (int j = k + 1; j < n; j++) {
                        double xj = A_(j, k);
                        if (xj != 0.0) {
                            double t = -d11 * xj;
                            for (int i = j; i < n; i++) A_(i, j) += A_(i, k) * t;
                        }
                    }
                    for (int i = k + 1; i < n; i++) A_(i, k) *= d11;
                }
            } else {
                if (k < n - 2) {
                    double d21 = A_(k + 1, k);
                    double d11 = A_(k + 1, k + 1) / d21;
                    double d22 = A_(k, k) / d21;
                    double t = 1.0 / (d11 * d22 - 1.0);
                    d21 = t / d21;
                    for (int j = k + 2; j < n; j++) {
                        double wk = d21 * (d11 * A_(j, k) - A_(j, k + 1));
                        double wkp1 = d21 * (d22 * A_(j, k + 1) - A_(j, k));
                        for (int i = j; i < n; i++)
                            A_(i, j) = A_(i, j) - A_(i, k) * wk - A_(i, k + 1) * wkp1;
                        A_(j, k) = wk;
                        A_(j, k + 1) = wkp1;
                    }
                }
            }
        }
        if (kstep == 1) {
            ipiv[k] = kp + 1;
        } else {
            ipiv[k] = -(kp + 1);
            ipiv[k + 1] = -(kp + 1);
        }
        k += kstep;
    }
    return info;
}

/* source_code/dsyneg.f:47-73 (UPLO='L' branch). */
int orc_dsyneg_l(const double *a, int lda, int n, const int *ipiv)
{
    int inert = 0;
    double t = 0.0;
    for (int k = n - 1; k >= 0; k--) {
        double d = A_(k, k);
        if (ipiv[k] <= 0) {
            if (t == 0.0) {
                t = fabs(A_(k, k - 1));
                d = (d / t) * A_(k - 1, k - 1) - t;
            } else {
                d = t;
                t = 0.0;
            }
        }
        if (d < 0.0) inert++;
    }
    return inert;
}

/* y = A_trail * x for the symmetric trailing block A(s:n-1, s:n-1), lower storage. */
static void symv_lower(const double *a, int lda, int s, int n, const double *x, double *y)
{
    int m = n - s;
    for (int i = 0; i < m; i++) y[i] = 0.0;
    for (int j = 0; j < m; j++) {
        double xj = x[j];
        double acc = 0.0;
        y[j] += A_(s + j, s + j) * xj;
        for (int i = j + 1; i < m; i++) {
            double aij = A_(s + i, s + j);
            y[i] += aij * xj;
            acc += aij * x[i];
        }
        y[j] += acc;
    }
}

/* Published LAPACK DSYTRI, UPLO='L'.  Returns INFO. */
int orc_dsytri_l(double *a, int lda, int n, const int *ipiv)
{
    for (int info = n - 1; info >= 0; info--)
        if (ipiv[info] > 0 && A_(info, info) == 0.0) return info + 1;
    double *work = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *tmp = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    int k = n - 1;
    while (k >= 0) {
        int kstep;
        if (ipiv[k] > 0) {
            A_(k, k) = 1.0 / A_(k, k);
            if (k < n - 1) {
                int m = n - 1 - k;
                for (int i = 0; i < m; i++) work[i] = A_(k + 1 + i, k);
                symv_lower(a, lda, k + 1, n, work, tmp);
                double dot = 0.0;
                for (int i = 0; i < m; i++) {
                    A_(k + 1 + i, k) = -tmp[i];
                    dot += work[i] * A_(k + 1 + i, k);
                }
                A_(k, k) -= dot;
            }
            kstep = 1;
        } else {
            double t = fabs(A_(k, k - 1));
            double ak = A_(k - 1, k - 1) / t;
            double akp1 = A_(k, k) / t;
            double akkp1 = A_(k, k - 1) / t;
            double d = t * (ak * akp1 - 1.0);
            A_(k - 1, k - 1) = akp1 / d;
            A_(k, k) = ak / d;
            A_(k, k - 1) = -akkp1 / d;
            if (k < n - 1) {
                int m = n - 1 - k;
                double dot;
                for (int i = 0; i < m; i++) work[i] = A_(k + 1 + i, k);
                symv_lower(a, lda, k + 1, n, work, tmp);
                dot = 0.0;
                for (int i = 0; i < m; i++) {
                    A_(k + 1 + i, k) = -tmp[i];
                    dot += work[i] * A_(k + 1 + i, k);
                }
                A_(k, k) -= dot;
                dot = 0.0;
                for (int i = 0; i < m; i++) dot += A_(k + 1 + i, k) * A_(k + 1 + i, k - 1);
                A_(k, k - 1) -= dot;
                for (int i = 0; i < m; i++) work[i] = A_(k + 1 + i, k - 1);
                symv_lower(a, lda, k + 1, n, work, tmp);
                dot = 0.0;
                for (int i = 0; i < m; i++) {
                    A_(k + 1 + i, k - 1) = -tmp[i];
                    dot += work[i] * A_(k + 1 + i, k - 1);
                }
                A_(k - 1, k - 1) -= dot;
            }
            kstep = 2;
        }
        int kp = abs(ipiv[k]) - 1;
        if (kp != k) {
            for (int i = kp + 1; i < n; i++) swap_d(&A_(i, k), &A_(i, kp));
            for (int j = k + 1; j < kp; j++) swap_d(&A_(j, k), &A_(kp, j));
            swap_d(&A_(k, k), &A_(kp, kp));
            if (kstep == 2) swap_d(&A_(k, k - 1), &A_(kp, k - 1));
        }
        k -= kstep;
    }
    free(work);
    free(tmp);
    return 0;
}

/* SYMINV dispatch: source_code/syminv.f:11-14 and source_code/syminl-alloc.f:17-41.
 * Returns KOUNT (number of negative eigenvalues), or N+1 when the inline
 * path meets an exact zero pivot, or -INFO (<0) when DSYTRF/DSYTRI report
 * a singular factor (the reference STOPs there). */
int orc_syminv(double *a, int lda, int n)
{
    if (n <= 30) return orc_syminv_inline(a, lda, n);
    int *ipiv = (int *)malloc(sizeof(int) * (size_t)n);
    int info = orc_dsytf2_l(a, lda, n, ipiv);
    if (info != 0) { free(ipiv); return -info; }
    int inert = orc_dsyneg_l(a, lda, n, ipiv);
    info = orc_dsytri_l(a, lda, n, ipiv);
    free(ipiv);
    if (info != 0) return -info;
    return inert;
}

/* Batched helper for the tests: nmat matrices stored back to back. */
void orc_syminv_batched(double *a, int lda, int n, int nmat, int *kount)
{
    for (int m = 0; m < nmat; m++)
        kount[m] = orc_syminv(a + (size_t)m * (size_t)lda * (size_t)n, lda, n);
}
