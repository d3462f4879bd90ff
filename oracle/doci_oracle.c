/*
 * doci_oracle.c -- CPU restatement of klgunst/doci's sparse-eigensolve path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library; the product never
 * does.  Each function cites the reference file:line range it restates.
 *
 * Pinning (what checks each function beyond its own formula):
 *   orc_read_fcidump, orc_transform_integrals, orc_exchange_table,
 *   orc_diagonal, orc_sigma  -> bit-exact against oracle/_ref (the reference's
 *                               own src/doci.c + src/read_fcidump.c built by
 *                               oracle/Makefile) on the reference's test
 *                               FCIDUMPs and a synthetic L=24 case.
 *   orc_lex_masks, orc_rank  -> brute-force enumeration, closed form.
 *   orc_connections          -> Slater-Condon rule + brute force; the sigma
 *                               built from it matches _ref bit-exactly.
 *   orc_davidson             -> tests/coHF.out and tests/lifHF.out trajectories
 *                               and energies; dense eigh on small FCIDUMPs.
 *   orc_jacobi_eig           -> LAPACK dsyev (scipy) eigenvalues.
 *   orc_sygv2                -> LAPACK dsygv (scipy, driver "gv") including
 *                               eigenvector signs.
 *   orc_ray_q                -> per-iteration |g|^2 and Rayleigh quotients,
 *                               iteration counts and energies of the
 *                               reference's own CG/CGP runs
 *                               (tests/golden/solvers.json); a planted
 *                               Fletcher-Reeves or stopping-test mistake
 *                               (orc_ray_q_planted) fails them.
 *   orc_davidson (deflation) -> the same fixtures' Davidson runs with
 *                               DAVIDSON_MAX_VEC 4 and 6.
 */
#include "doci_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* FCIDUMP (src/read_fcidump.c:18-82, whitespace-blind compare :88-110)      */
/* ------------------------------------------------------------------------ */

/* Compare a line with a keyword ignoring every whitespace character of the
 * line (read_fcidump.c:88-110). */
static int line_is(const char *line, const char *kw)
{
    const char *p = line;
    for (; *kw; ++kw) {
        while (isspace((unsigned char)*p)) ++p;
        if (*p != *kw) return 0;
        ++p;
    }
    while (isspace((unsigned char)*p)) ++p;
    return *p == '\0';
}

int orc_read_fcidump(const char *path, unsigned *L, unsigned *Na, unsigned *Nb,
                     double *core, double **one, double **two)
{
    char buf[255];
    int nel = 0, ms2 = 0, line = 1;
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    if (!fgets(buf, sizeof buf, fp) ||
        sscanf(buf, " &FCI NORB= %u , NELEC= %d , MS2= %d , ", L, &nel, &ms2) != 3) {
        fclose(fp);
        return -2;
    }
    *Na = (unsigned)((nel + ms2) / 2);
    *Nb = (unsigned)((nel - ms2) / 2);
    *core = 0.0;
    const size_t l = *L, l2 = l * l, l3 = l2 * l;
    *one = calloc(l2, sizeof(double));
    *two = calloc(l2 * l2, sizeof(double));
    while (fgets(buf, sizeof buf, fp)) {
        ++line;
        if (line_is(buf, "&END") || line_is(buf, "/END") || line_is(buf, "/")) break;
    }
    while (fgets(buf, sizeof buf, fp)) {
        double v;
        int i, j, k, m;
        ++line;
        if (sscanf(buf, " %lf %d %d %d %d ", &v, &i, &j, &k, &m) != 5) {
            fclose(fp);
            return -(1000 + line);
        }
        double *dst;
        if (k != 0) {
            if (i < 1 || j < 1 || m < 1 || (size_t)i > l || (size_t)j > l ||
                (size_t)k > l || (size_t)m > l) { fclose(fp); return -(1000 + line); }
            dst = *two + (size_t)(m - 1) * l3 + (size_t)(k - 1) * l2 +
                  (size_t)(j - 1) * l + (size_t)(i - 1);
        } else if (i != 0) {
            if (i < 1 || j < 1 || (size_t)i > l || (size_t)j > l) {
                fclose(fp); return -(1000 + line);
            }
            dst = *one + (size_t)(j - 1) * l + (size_t)(i - 1);
        } else {
            dst = core;
        }
        if (fabs(*dst) >= 1e-14)
            fprintf(stderr, "Doubly inputted value at line %d, hope you don't mind\n", line);
        *dst = v;
    }
    fclose(fp);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Integral tables (src/doci.c:630-683)                                      */
/* ------------------------------------------------------------------------ */

void orc_transform_integrals(double *one, double *two, unsigned Lu)
{
    const size_t L = Lu, L2 = L * L, L3 = L2 * L;
    /* one-body symmetrisation, doci.c:636-638 */
    for (size_t i = 0; i < L; ++i)
        for (size_t j = 0; j <= i; ++j) one[i * L + j] = one[j * L + i];
    /* eight-fold unpacking of every non-zero (ij|kl), i>=j, i>=k, k>=l, doci.c:640-655 */
    for (size_t i = 0; i < L; ++i)
        for (size_t j = 0; j <= i; ++j)
            for (size_t k = 0; k <= i; ++k)
                for (size_t l = 0; l <= k; ++l) {
                    const double v = two[i + L * j + L2 * k + L3 * l];
                    if (v != 0) {
                        two[k + L * l + L2 * i + L3 * j] = v;
                        two[j + L * i + L2 * l + L3 * k] = v;
                        two[l + L * k + L2 * j + L3 * i] = v;
                        two[j + L * i + L2 * k + L3 * l] = v;
                        two[l + L * k + L2 * i + L3 * j] = v;
                        two[i + L * j + L2 * l + L3 * k] = v;
                        two[k + L * l + L2 * j + L3 * i] = v;
                    }
                }
    /* h'_kl = h_kl - 1/2 sum_j (kj|jl), j ascending, doci.c:657-664 */
    for (size_t k = 0; k < L; ++k)
        for (size_t l = 0; l < L; ++l) {
            double t = 0;
            for (size_t j = 0; j < L; ++j) t += two[k + j * L + j * L2 + l * L3];
            one[k + l * L] -= 0.5 * t;
        }
}

void orc_exchange_table(const double *two, unsigned Lu, double *K)
{
    const size_t L = Lu, L2 = L * L, L3 = L2 * L;
    for (size_t i = 0; i < L; ++i)
        for (size_t k = 0; k < L; ++k)
            K[i * L + k] = (i == k) ? 0.0 : two[i + k * L + i * L2 + k * L3];
}

/* ------------------------------------------------------------------------ */
/* Addressing (src/doci.c:384-548, 687-724, 980-1013)                        */
/* ------------------------------------------------------------------------ */

uint64_t orc_binom(unsigned k, unsigned n)
{
    if (k > n) return 0;
    uint64_t r = 1;
    for (unsigned d = 1; d <= k; ++d) {
        r *= n--;
        r /= d;
    }
    return r;
}

void orc_lex_masks(unsigned N, unsigned L, uint64_t n, uint64_t *masks)
{
    int occ[65];
    for (unsigned i = 0; i < N; ++i) occ[i] = (int)i;
    occ[N] = (int)L;
    for (uint64_t r = 0; r < n; ++r) {
        if (r) {
            /* bump the highest pair that can move, pack the rest after it */
            for (int j = (int)N - 1; j >= 0; --j)
                if (occ[j] + 1 != occ[j + 1]) {
                    ++occ[j];
                    for (unsigned k = (unsigned)j + 1; k < N; ++k) occ[k] = occ[k - 1] + 1;
                    break;
                }
        }
        uint64_t m = 0;
        for (unsigned k = 0; k < N; ++k) m |= 1ull << occ[k];
        masks[r] = m;
    }
}

/* rank = C(L,N) - 1 - sum_e C(L-1-c_e, N-e) over the sorted occupied c_e. */
uint64_t orc_rank(uint64_t mask, unsigned N, unsigned L)
{
    uint64_t s = 0;
    unsigned e = 0;
    for (unsigned p = 0; p < L; ++p)
        if (mask >> p & 1) {
            s += orc_binom(N - e, L - 1 - p);
            ++e;
        }
    return orc_binom(N, L) - 1 - s;
}

static void occ_of(uint64_t mask, unsigned L, int *occ)
{
    int e = 0;
    for (unsigned p = 0; p < L; ++p)
        if (mask >> p & 1) occ[e++] = (int)p;
}

/* Unrank by walking the lexicographic tree with binomial counts. */
static uint64_t unrank(uint64_t r, unsigned N, unsigned L)
{
    uint64_t m = 0;
    unsigned p = 0;
    for (unsigned e = 0; e < N; ++e) {
        for (;; ++p) {
            const uint64_t block = orc_binom(N - e - 1, L - 1 - p);
            if (r < block) break;
            r -= block;
        }
        m |= 1ull << p;
        ++p;
    }
    return m;
}

/* ------------------------------------------------------------------------ */
/* Diagonal (src/doci.c:550-628), exact addition order                       */
/* ------------------------------------------------------------------------ */

/* D at rank `r` (or at ranks[r] when ranks != NULL). */
static void diagonal_at(const double *one, const double *two, double core, unsigned Nu,
                        unsigned Lu, int64_t nrow, const uint64_t *ranks, double *D, int threads)
{
    const size_t L = Lu, L2 = L * L, L3 = L2 * L;
    const int N = (int)Nu;
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
    for (int64_t r = 0; r < nrow; ++r) {
        int occ[66];
        occ_of(unrank(ranks ? ranks[r] : (uint64_t)r, Nu, Lu), Lu, occ);
        occ[N] = (int)L;
        double t = core;
        for (int i = N - 1; i >= 0; --i) {
            const size_t o = (size_t)occ[i];
            size_t i1 = o * (L + 1), i2 = o * (1 + L3);
            int c = 0;
            t += 2 * one[i1];
            for (size_t q = 0; q < L; ++q) {
                const int is = occ[c] == (int)q;
                c += is;
                t += 2 * is * two[i1];
                t += (1 - is) * two[i2];
                i1 += L2 + L3;
                i2 += L + L2;
            }
        }
        D[r] = t;
    }
}

void orc_diagonal(const double *one, const double *two, double core,
                  unsigned Nu, unsigned Lu, uint64_t n, double *D, int threads)
{
    diagonal_at(one, two, core, Nu, Lu, (int64_t)n, NULL, D, threads);
}

void orc_diagonal_rows(const double *one, const double *two, double core, unsigned N,
                       unsigned L, const uint64_t *ranks, int64_t count, double *D)
{
    diagonal_at(one, two, core, N, L, count, ranks, D, 1);
}

/* ------------------------------------------------------------------------ */
/* Connections and sigma (src/doci.c:164-238, 851-978)                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    unsigned N, L;
    uint64_t C[65][66]; /* C[x][j] = binom(j, x) */
} btab;

static void btab_init(btab *b, unsigned N, unsigned L)
{
    b->N = N;
    b->L = L;
    for (unsigned x = 0; x <= L; ++x)
        for (unsigned j = 0; j <= N + 1 && j < 66; ++j) b->C[x][j] = orc_binom(j, x);
}

/* Lower-triangle connections of I (occupied list occ, rank r): for the hole at
 * pair index h from N-1 down to 0 and every empty target t below it, t
 * descending, J = I - {c_h} + {t}.  rank(J) - rank(I) =
 *   C(L-1-c_h, N-h) - C(L-1-t, N-j) + sum_{e=j}^{h-1} [C(L-1-c_e, N-e) - C(L-1-c_e, N-e-1)]
 * with j = #{occupied below t}; the sum grows as t passes occupied orbitals. */
#define WALK_LOWER(B, occ, r, BODY)                                              \
    do {                                                                         \
        const int N_ = (int)(B)->N, L_ = (int)(B)->L;                            \
        for (int h = N_ - 1; h >= 0; --h) {                                      \
            const int o_ = (occ)[h];                                             \
            const int64_t base_ = (int64_t)(B)->C[L_ - 1 - o_][N_ - h];          \
            int64_t s_ = 0;                                                      \
            int j_ = h;                                                          \
            for (int t_ = o_ - 1; t_ >= 0; --t_) {                               \
                if (j_ > 0 && (occ)[j_ - 1] == t_) {                             \
                    --j_;                                                        \
                    s_ += (int64_t)(B)->C[L_ - 1 - t_][N_ - j_] -                \
                          (int64_t)(B)->C[L_ - 1 - t_][N_ - j_ - 1];             \
                    continue;                                                    \
                }                                                                \
                const uint64_t J = (uint64_t)((int64_t)(r) + base_ -             \
                                              (int64_t)(B)->C[L_ - 1 - t_][N_ - j_] + s_); \
                const int kidx = o_ * L_ + t_;                                   \
                BODY                                                             \
            }                                                                    \
        }                                                                        \
    } while (0)

int orc_connections(uint64_t rank, unsigned N, unsigned L, uint64_t *Jout,
                    uint32_t *kout, int cap)
{
    btab b;
    btab_init(&b, N, L);
    int occ[66];
    occ_of(unrank(rank, N, L), L, occ);
    int cnt = 0;
    WALK_LOWER(&b, occ, rank, {
        if (cnt < cap) {
            Jout[cnt] = J;
            kout[cnt] = (uint32_t)kidx;
        }
        ++cnt;
    });
    return cnt;
}

void orc_sigma(const double *K, const double *D, unsigned N, unsigned L,
               uint64_t n, const double *c, double *sigma, int threads)
{
    if (threads < 1) threads = 1;
    btab b;
    btab_init(&b, N, L);
    double *priv = threads > 1 ? calloc(n * (size_t)(threads - 1), sizeof(double)) : NULL;
#pragma omp parallel num_threads(threads)
    {
#ifdef _OPENMP
        const int t = omp_get_thread_num();
#else
        const int t = 0;
#endif
        double *acc = t ? priv + n * (size_t)(t - 1) : sigma;
#pragma omp for schedule(static)
        for (int64_t r = 0; r < (int64_t)n; ++r) sigma[r] = c[r] * D[r];
        const uint64_t lo = (uint64_t)t * n / (uint64_t)threads;
        const uint64_t hi = (uint64_t)(t + 1) * n / (uint64_t)threads;
        int occ[66];
        if (lo < hi) occ_of(unrank(lo, N, L), L, occ);
        for (uint64_t r = lo; r < hi; ++r) {
            if (r != lo) {
                /* lexicographic successor of occ (doci.c:980-1013) */
                int j = (int)N - 1;
                while (occ[j] - j == (int)(L - N)) --j;
                ++occ[j];
                for (; j < (int)N - 1; ++j) occ[j + 1] = occ[j] + 1;
            }
            WALK_LOWER(&b, occ, r, {
                const double v = K[kidx];
                acc[r] += v * c[J];
                acc[J] += v * c[r];
            });
        }
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t r = 0; r < (int64_t)n; ++r)
            for (int i = 0; i < threads - 1; ++i) sigma[r] += priv[(size_t)r + (size_t)i * n];
    }
    free(priv);
}

/* ------------------------------------------------------------------------ */
/* Dense helpers                                                             */
/* ------------------------------------------------------------------------ */

/* Cyclic Jacobi eigensolver for a symmetric m x m matrix given by its upper
 * triangle (column-major, lda); eigenvalues ascending in w, eigenvectors in
 * the columns of Z.  Stands in for LAPACK dsyev (davidson.c:81,88). */
void orc_jacobi_eig(int m, const double *A, int lda, double *w, double *Z, int ldz)
{
    double *a = malloc(sizeof(double) * (size_t)(m * m));
    double *v = malloc(sizeof(double) * (size_t)(m * m));
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
            const double x = i <= j ? A[i + j * lda] : A[j + i * lda];
            a[i + j * m] = x;
            v[i + j * m] = i == j;
        }
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                tot += a[i + j * m] * a[i + j * m];
                if (i != j) off += a[i + j * m] * a[i + j * m];
            }
        if (off <= 1e-30 * tot || off == 0) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = a[p + q * m];
                if (apq == 0) continue;
                const double app = a[p + p * m], aqq = a[q + q * m];
                const double theta = (aqq - app) / (2 * apq);
                const double t = (theta >= 0 ? 1 : -1) / (fabs(theta) + sqrt(theta * theta + 1));
                const double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < m; ++k) {
                    const double akp = a[k + p * m], akq = a[k + q * m];
                    a[k + p * m] = c * akp - s * akq;
                    a[k + q * m] = s * akp + c * akq;
                }
                for (int k = 0; k < m; ++k) {
                    const double apk = a[p + k * m], aqk = a[q + k * m];
                    a[p + k * m] = c * apk - s * aqk;
                    a[q + k * m] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; ++k) {
                    const double vkp = v[k + p * m], vkq = v[k + q * m];
                    v[k + p * m] = c * vkp - s * vkq;
                    v[k + q * m] = s * vkp + c * vkq;
                }
            }
    }
    int *idx = malloc(sizeof(int) * (size_t)m);
    for (int i = 0; i < m; ++i) idx[i] = i;
    for (int i = 1; i < m; ++i)
        for (int j = i; j > 0 && a[idx[j] * (m + 1)] < a[idx[j - 1] * (m + 1)]; --j) {
            const int t = idx[j];
            idx[j] = idx[j - 1];
            idx[j - 1] = t;
        }
    for (int j = 0; j < m; ++j) {
        w[j] = a[idx[j] * (m + 1)];
        for (int i = 0; i < m; ++i) Z[i + j * ldz] = v[i + idx[j] * m];
    }
    free(idx);
    free(a);
    free(v);
}

/* Eigen-decomposition of [[a,b],[b,c]] following LAPACK dlaev2's published
 * steps: rt1 is the eigenvalue of larger magnitude, (cs1, sn1) its unit
 * eigenvector. */
static void laev2(double a, double b, double c, double *rt1, double *rt2,
                  double *cs1, double *sn1)
{
    const double sm = a + c, df = a - c, adf = fabs(df), tb = b + b, ab = fabs(tb);
    double acmx, acmn, rt, cs, ct, tn;
    int sgn1, sgn2;
    if (fabs(a) > fabs(c)) { acmx = a; acmn = c; } else { acmx = c; acmn = a; }
    if (adf > ab) rt = adf * sqrt(1 + (ab / adf) * (ab / adf));
    else if (adf < ab) rt = ab * sqrt(1 + (adf / ab) * (adf / ab));
    else rt = ab * sqrt(2.0);
    if (sm < 0) { *rt1 = 0.5 * (sm - rt); sgn1 = -1; *rt2 = (acmx / *rt1) * acmn - (b / *rt1) * b; }
    else if (sm > 0) { *rt1 = 0.5 * (sm + rt); sgn1 = 1; *rt2 = (acmx / *rt1) * acmn - (b / *rt1) * b; }
    else { *rt1 = 0.5 * rt; *rt2 = -0.5 * rt; sgn1 = 1; }
    if (df >= 0) { cs = df + rt; sgn2 = 1; } else { cs = df - rt; sgn2 = -1; }
    if (fabs(cs) > ab) { ct = -tb / cs; *sn1 = 1 / sqrt(1 + ct * ct); *cs1 = ct * *sn1; }
    else if (ab == 0) { *cs1 = 1; *sn1 = 0; }
    else { tn = -cs / tb; *cs1 = 1 / sqrt(1 + tn * tn); *sn1 = tn * *cs1; }
    if (sgn1 == sgn2) { tn = *cs1; *cs1 = -*sn1; *sn1 = tn; }
}

/* 2x2 A z = lambda B z (upper triangles, column-major) following LAPACK
 * dsygv(itype=1, 'V', 'U'): Cholesky B = U^T U, C = U^-T A U^-1, 2x2
 * rotation, ascending sort, z = U^-1 y.  Stands in for CG.c:104.
 * Returns 0, or 1 when B is not positive definite. */
int orc_sygv2(const double *A, const double *B, double *w, double *Z)
{
    double u11 = B[0], u12 = B[2], u22 = B[3];
    if (u11 <= 0) return 1;
    u11 = sqrt(u11);
    u12 = u12 / u11;
    u22 = u22 - u12 * u12;
    if (u22 <= 0) return 1;
    u22 = sqrt(u22);
    double a11 = A[0], a12 = A[2], a22 = A[3];
    a11 = a11 / (u11 * u11);
    a12 = a12 * (1 / u11);
    const double ct = -0.5 * a11;
    a12 = a12 + ct * u12;
    a22 = a22 + a12 * (-u12) + u12 * (-a12);
    a12 = a12 + ct * u12;
    a12 = a12 / u22;
    a22 = a22 / (u22 * u22);
    double d1 = a11, d2 = a22, z11 = 1, z21 = 0, z12 = 0, z22 = 1;
    const double eps = 1.1102230246251565e-16;
    if (!(fabs(a12) <= sqrt(fabs(d1)) * sqrt(fabs(d2)) * eps) && a12 != 0) {
        double rt1, rt2, c, s;
        laev2(a11, a12, a22, &rt1, &rt2, &c, &s);
        d1 = rt1; d2 = rt2;
        z11 = c; z21 = s; z12 = -s; z22 = c;
    }
    if (d2 < d1) {
        double t = d1; d1 = d2; d2 = t;
        t = z11; z11 = z12; z12 = t;
        t = z21; z21 = z22; z22 = t;
    }
    w[0] = d1;
    w[1] = d2;
    /* back-substitution U x = y, column by column */
    double x2 = z21 / u22, x1 = (z11 - x2 * u12) / u11;
    Z[0] = x1; Z[1] = x2;
    x2 = z22 / u22; x1 = (z12 - x2 * u12) / u11;
    Z[2] = x1; Z[3] = x2;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Davidson (src/davidson.c:33-255)                                          */
/* ------------------------------------------------------------------------ */

/* Sums over n entries in ORC_CHUNKS fixed chunks added in chunk order: the
 * result does not depend on the thread count or on the order in which
 * threads finish (an OpenMP reduction clause combines in arrival order,
 * which makes rounding-sensitive runs differ from one call to the next). */
#define ORC_CHUNKS 256
#define CHUNK_LO(n, c) ((uint64_t)(n) * (uint64_t)(c) / ORC_CHUNKS)

static double dot(const double *a, const double *b, uint64_t n)
{
    double part[ORC_CHUNKS];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < ORC_CHUNKS; ++c) {
        double s = 0;
        for (uint64_t i = CHUNK_LO(n, c); i < CHUNK_LO(n, c + 1); ++i) s += a[i] * b[i];
        part[c] = s;
    }
    double s = 0;
    for (int c = 0; c < ORC_CHUNKS; ++c) s += part[c];
    return s;
}

int orc_davidson(double *result, double *energy, int max_vec, int keep,
                 double tol, const double *K, const double *D, unsigned N,
                 unsigned L, uint64_t n, int max_its, int threads,
                 double *trace_res, double *trace_e, int trace_cap, int *its_out)
{
    return orc_davidson_timed(result, energy, max_vec, keep, tol, K, D, N, L, n, max_its, threads,
                              trace_res, trace_e, NULL, trace_cap, its_out);
}

static double wall_now(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

int orc_davidson_timed(double *result, double *energy, int max_vec, int keep,
                       double tol, const double *K, const double *D, unsigned N,
                       unsigned L, uint64_t n, int max_its, int threads,
                       double *trace_res, double *trace_e, double *trace_t,
                       int trace_cap, int *its_out)
{
    const double t_start = wall_now();
    const size_t mv = (size_t)max_vec;
    double *S = calloc(mv * mv, sizeof(double));
    double *Y = calloc(mv * mv, sizeof(double));
    double *w = calloc(mv, sizeof(double));
    double *V = malloc(sizeof(double) * n * mv);
    double *W = malloc(sizeof(double) * n * mv);
    double *t = malloc(sizeof(double) * n);
    memcpy(t, result, sizeof(double) * n);
    double res = tol * 10;
    int its = 0, m = 0;
    *energy = 0;
    while (res > tol && its++ < max_its) {
        /* new_search_vector: modified Gram-Schmidt, davidson.c:133-176 */
        for (int i = 0; i < m; ++i) {
            const double a = -dot(V + (size_t)i * n, t, n);
            const double *vi = V + (size_t)i * n;
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < (int64_t)n; ++k) t[k] += a * vi[k];
        }
        const double nrm = 1 / sqrt(dot(t, t, n));
        double *vm = V + (size_t)m * n;
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < (int64_t)n; ++k) vm[k] = t[k] * nrm;
        orc_sigma(K, D, N, L, n, vm, W + (size_t)m * n, threads);
        /* expand_submatrix, davidson.c:178-193 */
        for (int i = 0; i <= m; ++i) S[(size_t)m * mv + i] = dot(V + (size_t)i * n, W + (size_t)m * n, n);
        ++m;
        orc_jacobi_eig(m, S, max_vec, w, Y, max_vec);
        if (m == max_vec) {
            /* deflate, davidson.c:229-255 */
            double *nv = malloc(sizeof(double) * n * (size_t)keep);
            for (int pass = 0; pass < 2; ++pass) {
                double *X = pass ? W : V;
#pragma omp parallel for schedule(static)
                for (int64_t k = 0; k < (int64_t)n; ++k)
                    for (int c = 0; c < keep; ++c) {
                        double s = 0;
                        for (int j = 0; j < max_vec; ++j) s += X[(size_t)k + (size_t)j * n] * Y[j + c * max_vec];
                        nv[(size_t)k + (size_t)c * n] = s;
                    }
                memcpy(X, nv, sizeof(double) * n * (size_t)keep);
            }
            free(nv);
            for (int c = 0; c < keep; ++c)
                for (int i = 0; i <= c; ++i)
                    S[(size_t)c * mv + i] = dot(V + (size_t)i * n, W + (size_t)c * n, n);
            m = keep;
            orc_jacobi_eig(m, S, max_vec, w, Y, max_vec);
        }
        /* calculate_residue, davidson.c:195-212 */
        const double th = w[0];
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < (int64_t)n; ++k) {
            double x = 0, r = 0;
            for (int j = 0; j < m; ++j) x += V[(size_t)k + (size_t)j * n] * Y[j];
            for (int j = 0; j < m; ++j) r += W[(size_t)k + (size_t)j * n] * Y[j];
            r -= th * x;
            result[k] = x;
            t[k] = r;
        }
        res = sqrt(dot(t, t, n));
        *energy = th;
        if (trace_res && its - 1 < trace_cap) {
            trace_res[its - 1] = res;
            trace_e[its - 1] = th;
            if (trace_t) trace_t[its - 1] = wall_now() - t_start;
        }
        /* create_new_vec_t, davidson.c:214-227 */
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < (int64_t)n; ++k) {
            const double d = fabs(D[k] - th);
            t[k] = d > 1e-12 ? t[k] / d : t[k] / 1e-12;
        }
    }
    *energy = w[0];
    *its_out = its;
    free(S); free(Y); free(w); free(V); free(W); free(t);
    return its >= max_its;
}

/* ------------------------------------------------------------------------ */
/* Rayleigh-quotient CG (src/CG.c:13-164)                                     */
/* ------------------------------------------------------------------------ */

int orc_ray_q(double *vec, double *energy, double tol, int max_its,
              int precond, const double *K, const double *D, unsigned N,
              unsigned L, uint64_t n, int threads, double *trace_res,
              double *trace_e, int trace_cap, int *its_out)
{
    return orc_ray_q_planted(vec, energy, tol, max_its, precond, K, D, N, L, n, threads,
                             trace_res, trace_e, trace_cap, its_out, 0);
}

/* orc_ray_q with a mistake planted on purpose (tests show that the fixture
 * bars reject it): mistake 1 = Fletcher-Reeves ratio of the unpreconditioned
 * norms |g|^2 instead of the preconditioned ones (CG.c:92); 2 = stopping test
 * on the preconditioned norm instead of |g|^2 (CG.c:90); 0 = none. */
int orc_ray_q_planted(double *vec, double *energy, double tol, int max_its,
                      int precond, const double *K, const double *D, unsigned N,
                      unsigned L, uint64_t n, int threads, double *trace_res,
                      double *trace_e, int trace_cap, int *its_out, int mistake)
{
    double *x = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n);
    double *Ax = malloc(sizeof(double) * n), *Ap = malloc(sizeof(double) * n);
    double *g = malloc(sizeof(double) * n), *pinv = NULL;
    if (precond) {
        pinv = malloc(sizeof(double) * n);
        for (uint64_t k = 0; k < n; ++k) pinv[k] = fabs(D[k]) > 1e-12 ? 1 / D[k] : 1 / 1e-12;
    }
    memcpy(x, vec, sizeof(double) * n);
    orc_sigma(K, D, N, L, n, x, Ax, threads);
    double rho = dot(Ax, x, n);
    for (uint64_t k = 0; k < n; ++k) g[k] = 2 * (Ax[k] - rho * x[k]);
    double normq = dot(g, g, n);
    if (pinv)
        for (uint64_t k = 0; k < n; ++k) g[k] *= pinv[k];
    double ng = dot(g, g, n), ngp = ng;
    double nq_prev = normq; /* for mistake 1 */
    int its = 0;
    while ((mistake == 2 ? ng : normq) > tol && its < max_its) {
        if (its != 0) {
            const double q = mistake == 1 ? normq / nq_prev : ng / ngp;
            for (uint64_t k = 0; k < n; ++k) p[k] = q * p[k] - g[k];
        } else {
            for (uint64_t k = 0; k < n; ++k) p[k] = -g[k];
        }
        orc_sigma(K, D, N, L, n, p, Ap, threads);
        double A2[4], B2[4], lam[2], Z[4];
        A2[0] = dot(x, Ax, n); A2[1] = dot(p, Ax, n); A2[2] = dot(x, Ap, n); A2[3] = dot(p, Ap, n);
        B2[0] = dot(x, x, n);  B2[1] = dot(p, x, n);  B2[2] = dot(x, p, n);  B2[3] = dot(p, p, n);
        orc_sygv2(A2, B2, lam, Z);
        for (uint64_t k = 0; k < n; ++k) {
            x[k] = Z[0] * x[k] + Z[1] * p[k];
            Ax[k] = Z[0] * Ax[k] + Z[1] * Ap[k];
        }
        rho = lam[0];
        for (uint64_t k = 0; k < n; ++k) g[k] = 2 * (Ax[k] - rho * x[k]);
        nq_prev = normq;
        normq = dot(g, g, n);
        if (pinv)
            for (uint64_t k = 0; k < n; ++k) g[k] *= pinv[k];
        ngp = ng;
        ng = dot(g, g, n);
        if (trace_res && its < trace_cap) {
            trace_res[its] = ng;
            trace_e[its] = rho;
        }
        ++its;
    }
    memcpy(vec, x, sizeof(double) * n);
    *energy = rho;
    *its_out = its;
    free(x); free(p); free(Ax); free(Ap); free(g); free(pinv);
    return 1;
}
