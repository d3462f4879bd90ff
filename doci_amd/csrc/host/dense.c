/* dense.c -- the small dense eigensolvers of the subspace steps.
 *
 * The reference calls LAPACK dsyev for the m x m Davidson problem
 * (src/davidson.c:81,88) and dsygv for the 2 x 2 CG problem (src/CG.c:104).
 * Here: a cyclic Jacobi (Golub & Van Loan 8.5) for the symmetric problem, a
 * Cholesky reduction for the generalized m x m problem, and for 2 x 2 the
 * same sequence of steps as dsygv (Cholesky, reduction, one plane rotation
 * by the dlaev2 formulas, ascending sort, back-substitution) so that
 * eigenvector signs, which steer the CG recurrence, agree with LAPACK's.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

/* Symmetric A (m x m, full storage, column-major, lda) is destroyed.
 * Eigenvalues ascending in w, orthonormal eigenvectors in Z's columns. */
void dh_jacobi(int m, double *A, int lda, double *w, double *Z, int ldz)
{
    double *V = calloc((size_t)m * m, sizeof(double));
    int *ord = malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    for (int i = 0; i < m; ++i) V[i + (size_t)i * m] = 1.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0, all = 0;
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                const double a = A[i + (size_t)j * lda];
                all += a * a;
                if (i != j) off += a * a;
            }
        if (off == 0 || off <= 1e-32 * all) break;
        for (int p = 0; p + 1 < m; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = A[p + (size_t)q * lda];
                if (apq == 0) continue;
                const double tau = (A[q + (size_t)q * lda] - A[p + (size_t)p * lda]) / (2 * apq);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1 + tau * tau));
                const double c = 1 / sqrt(1 + t * t), s = t * c;
                for (int k = 0; k < m; ++k) { /* columns p, q */
                    double *ak = A + k;
                    const double x = ak[(size_t)p * lda], y = ak[(size_t)q * lda];
                    ak[(size_t)p * lda] = c * x - s * y;
                    ak[(size_t)q * lda] = s * x + c * y;
                }
                for (int k = 0; k < m; ++k) { /* rows p, q */
                    double *ak = A + (size_t)k * lda;
                    const double x = ak[p], y = ak[q];
                    ak[p] = c * x - s * y;
                    ak[q] = s * x + c * y;
                }
                for (int k = 0; k < m; ++k) {
                    double *vk = V + k;
                    const double x = vk[(size_t)p * m], y = vk[(size_t)q * m];
                    vk[(size_t)p * m] = c * x - s * y;
                    vk[(size_t)q * m] = s * x + c * y;
                }
            }
    }
    for (int i = 0; i < m; ++i) ord[i] = i;
    for (int i = 1; i < m; ++i) /* stable insertion sort by eigenvalue */
        for (int j = i; j > 0; --j) {
            const double a = A[ord[j] + (size_t)ord[j] * lda], b = A[ord[j - 1] + (size_t)ord[j - 1] * lda];
            if (!(a < b)) break;
            const int t = ord[j];
            ord[j] = ord[j - 1];
            ord[j - 1] = t;
        }
    for (int j = 0; j < m; ++j) {
        w[j] = A[ord[j] + (size_t)ord[j] * lda];
        for (int i = 0; i < m; ++i) Z[i + (size_t)j * ldz] = V[i + (size_t)ord[j] * m];
    }
    free(V);
    free(ord);
}

/* c = G^-1 b for SPD G (upper triangle, column-major, leading dimension
 * ld) by Cholesky; returns b^T c, or -1 when G is not positive definite. */
double dh_gram_solve(int m, const double *G, int ld, const double *b, double *c)
{
    double *Lc = calloc((size_t)m * m, sizeof(double));
    double r = -1;
    if (!Lc) return -1;
    for (int j = 0; j < m; ++j) {
        double d = G[j + (size_t)j * ld];
        for (int k = 0; k < j; ++k) d -= Lc[j + (size_t)k * m] * Lc[j + (size_t)k * m];
        if (!(d > 0)) goto out;
        d = sqrt(d);
        Lc[j + (size_t)j * m] = d;
        for (int i = j + 1; i < m; ++i) {
            double v = G[j + (size_t)i * ld];
            for (int k = 0; k < j; ++k) v -= Lc[i + (size_t)k * m] * Lc[j + (size_t)k * m];
            Lc[i + (size_t)j * m] = v / d;
        }
    }
    for (int i = 0; i < m; ++i) { /* Lc z = b */
        double v = b[i];
        for (int k = 0; k < i; ++k) v -= Lc[i + (size_t)k * m] * c[k];
        c[i] = v / Lc[i + (size_t)i * m];
    }
    for (int i = m - 1; i >= 0; --i) { /* Lc^T c = z */
        double v = c[i];
        for (int k = i + 1; k < m; ++k) v -= Lc[k + (size_t)i * m] * c[k];
        c[i] = v / Lc[i + (size_t)i * m];
    }
    r = 0;
    for (int i = 0; i < m; ++i) r += b[i] * c[i];
out:
    free(Lc);
    return r;
}

/* S y = w G y for symmetric S and SPD G (upper triangles used, column-major
 * with leading dimension ld); Y has G-orthonormal columns (Y^T G Y = I).
 * Returns DOCI_E_BREAKDOWN when G is not positive definite. */
int dh_gen_eig(int m, const double *S, const double *G, int ld, double *w, double *Y, int ldy)
{
    const size_t mm = (size_t)m * m;
    double *Lc = calloc(mm, sizeof(double)), *C = calloc(mm, sizeof(double));
    double *T = calloc(mm, sizeof(double)), *Z = calloc(mm, sizeof(double));
    int rc = DOCI_OK;
    /* Cholesky G = Lc Lc^T (lower) from the upper triangle */
    for (int j = 0; j < m && rc == DOCI_OK; ++j) {
        double d = G[j + (size_t)j * ld];
        for (int k = 0; k < j; ++k) d -= Lc[j + (size_t)k * m] * Lc[j + (size_t)k * m];
        if (!(d > 0)) {
            rc = doci_fail(DOCI_E_BREAKDOWN, "subspace Gram matrix not positive definite");
            break;
        }
        d = sqrt(d);
        Lc[j + (size_t)j * m] = d;
        for (int i = j + 1; i < m; ++i) {
            double v = G[j + (size_t)i * ld]; /* G(i,j) from upper: G(j,i) */
            for (int k = 0; k < j; ++k) v -= Lc[i + (size_t)k * m] * Lc[j + (size_t)k * m];
            Lc[i + (size_t)j * m] = v / d;
        }
    }
    if (rc == DOCI_OK) {
        /* T = Lc^-1 S (full symmetric S) */
        for (int c = 0; c < m; ++c)
            for (int i = 0; i < m; ++i) {
                double v = i <= c ? S[i + (size_t)c * ld] : S[c + (size_t)i * ld];
                for (int k = 0; k < i; ++k) v -= Lc[i + (size_t)k * m] * T[k + (size_t)c * m];
                T[i + (size_t)c * m] = v / Lc[i + (size_t)i * m];
            }
        /* C = T Lc^-T, i.e. solve Lc C^T = T^T row by row; C symmetric */
        for (int r = 0; r < m; ++r)
            for (int i = 0; i < m; ++i) {
                double v = T[r + (size_t)i * m];
                for (int k = 0; k < i; ++k) v -= Lc[i + (size_t)k * m] * C[r + (size_t)k * m];
                C[r + (size_t)i * m] = v / Lc[i + (size_t)i * m];
            }
        for (int j = 0; j < m; ++j) /* symmetrise */
            for (int i = 0; i < j; ++i) {
                const double a = 0.5 * (C[i + (size_t)j * m] + C[j + (size_t)i * m]);
                C[i + (size_t)j * m] = C[j + (size_t)i * m] = a;
            }
        dh_jacobi(m, C, m, w, Z, m);
        /* Y = Lc^-T Z */
        for (int c = 0; c < m; ++c)
            for (int i = m - 1; i >= 0; --i) {
                double v = Z[i + (size_t)c * m];
                for (int k = i + 1; k < m; ++k) v -= Lc[k + (size_t)i * m] * Y[k + (size_t)c * ldy];
                Y[i + (size_t)c * ldy] = v / Lc[i + (size_t)i * m];
            }
    }
    free(Lc);
    free(C);
    free(T);
    free(Z);
    return rc;
}

/* dlaev2 formulas: eigenpairs of [[a,b],[b,c]]; rt1 has the larger
 * magnitude and (cs, sn) is its unit eigenvector. */
static void eig2(double a, double b, double c, double *rt1, double *rt2, double *cs, double *sn)
{
    const double sm = a + c, df = a - c, adf = fabs(df), tb = 2 * b, ab = fabs(tb);
    const double acmx = fabs(a) > fabs(c) ? a : c, acmn = fabs(a) > fabs(c) ? c : a;
    double rt;
    if (adf > ab) rt = adf * sqrt(1 + (ab / adf) * (ab / adf));
    else if (adf < ab) rt = ab * sqrt(1 + (adf / ab) * (adf / ab));
    else rt = ab * sqrt(2.0);
    int s1;
    if (sm < 0) {
        *rt1 = 0.5 * (sm - rt);
        s1 = -1;
        *rt2 = (acmx / *rt1) * acmn - (b / *rt1) * b;
    } else if (sm > 0) {
        *rt1 = 0.5 * (sm + rt);
        s1 = 1;
        *rt2 = (acmx / *rt1) * acmn - (b / *rt1) * b;
    } else {
        *rt1 = 0.5 * rt;
        *rt2 = -0.5 * rt;
        s1 = 1;
    }
    const int s2 = df >= 0 ? 1 : -1;
    const double cv = df >= 0 ? df + rt : df - rt;
    if (fabs(cv) > ab) {
        const double ct = -tb / cv;
        *sn = 1 / sqrt(1 + ct * ct);
        *cs = ct * *sn;
    } else if (ab == 0) {
        *cs = 1;
        *sn = 0;
    } else {
        const double tn = -cv / tb;
        *cs = 1 / sqrt(1 + tn * tn);
        *sn = tn * *cs;
    }
    if (s1 == s2) {
        const double t = *cs;
        *cs = -*sn;
        *sn = t;
    }
}

/* 2 x 2 generalized problem, column-major upper triangles A, B.
 * Z's columns are the B-normalised eigenvectors for w[0] <= w[1]. */
int dh_sygv2(const double *A, const double *B, double *w, double *Z)
{
    if (!(B[0] > 0)) return doci_fail(DOCI_E_BREAKDOWN, "2x2 Gram not positive definite");
    const double u11 = sqrt(B[0]), u12 = B[2] / u11;
    const double r22 = B[3] - u12 * u12;
    if (!(r22 > 0)) return doci_fail(DOCI_E_BREAKDOWN, "2x2 Gram not positive definite");
    const double u22 = sqrt(r22);
    /* C = U^-T A U^-1 in the order of LAPACK's dsygs2 (itype 1, upper) */
    const double c11 = A[0] / (u11 * u11);
    const double half = -0.5 * c11;
    double c12 = A[2] * (1 / u11);
    c12 = c12 + half * u12;
    double c22 = A[3] + c12 * (-u12) + u12 * (-c12);
    c12 = c12 + half * u12;
    c12 = c12 / u22;
    c22 = c22 / (u22 * u22);
    double d0 = c11, d1 = c22, y00 = 1, y10 = 0, y01 = 0, y11 = 1;
    if (c12 != 0 && !(fabs(c12) <= sqrt(fabs(c11)) * sqrt(fabs(c22)) * 0x1p-53)) {
        double r1, r2, cs, sn;
        eig2(c11, c12, c22, &r1, &r2, &cs, &sn);
        d0 = r1;
        d1 = r2;
        y00 = cs;
        y10 = sn;
        y01 = -sn;
        y11 = cs;
    }
    if (d1 < d0) {
        double t = d0;
        d0 = d1;
        d1 = t;
        t = y00, y00 = y01, y01 = t;
        t = y10, y10 = y11, y11 = t;
    }
    w[0] = d0;
    w[1] = d1;
    Z[1] = y10 / u22;
    Z[0] = (y00 - Z[1] * u12) / u11;
    Z[3] = y11 / u22;
    Z[2] = (y01 - Z[3] * u12) / u11;
    return DOCI_OK;
}
