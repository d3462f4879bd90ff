/* fcidump.c -- FCIDUMP reader (drop-in for src/read_fcidump.c:18-110),
 * canonical writer and the seeded synthetic-integral generator. */
#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

/* A line equals a keyword when every non-whitespace character matches in
 * order (the reference's whitespace-blind compare, read_fcidump.c:88-110). */
static int same_ignoring_ws(const char *line, const char *kw)
{
    for (; *kw; ++kw, ++line) {
        while (isspace((unsigned char)*line)) ++line;
        if (*line != *kw) return 0;
    }
    while (isspace((unsigned char)*line)) ++line;
    return *line == '\0';
}

int dh_read_fcidump(const char *path, unsigned *L, unsigned *Na, unsigned *Nb, double *core,
                    double **one_out, double **two_out, int *errline)
{
    char line[255];
    int nelec, ms2, lineno = 1;
    FILE *fp = fopen(path, "r");
    if (!fp) return doci_fail(DOCI_E_IO, "cannot open %s", path);
    /* header: read_fcidump.c:37-45 */
    if (!fgets(line, sizeof line, fp) ||
        sscanf(line, " &FCI NORB= %u , NELEC= %d , MS2= %d , ", L, &nelec, &ms2) != 3) {
        fclose(fp);
        return doci_fail(DOCI_E_FORMAT, "Error in reading %s. File is wrongly formatted.", path);
    }
    if (*L == 0 || *L > DK_MAX_L) {
        fclose(fp);
        return doci_fail(DOCI_E_ORBITALS, "NORB = %u outside 1..%d", *L, DK_MAX_L);
    }
    *Na = (unsigned)((nelec + ms2) / 2);
    *Nb = (unsigned)((nelec - ms2) / 2);
    *core = 0;
    const size_t l = *L, l2 = l * l;
    double *one = calloc(l2, sizeof(double));
    double *two = calloc(l2 * l2, sizeof(double));
    if (!one || !two) {
        free(one);
        free(two);
        fclose(fp);
        return doci_fail(DOCI_E_NOMEM, "FCIDUMP tables for L=%u", *L);
    }
    /* skip the namelist up to &END, /END or / (read_fcidump.c:53-58) */
    while (fgets(line, sizeof line, fp)) {
        ++lineno;
        if (same_ignoring_ws(line, "&END") || same_ignoring_ws(line, "/END") ||
            same_ignoring_ws(line, "/"))
            break;
    }
    /* "value i j k l" lines (read_fcidump.c:60-80) */
    while (fgets(line, sizeof line, fp)) {
        double v;
        int i, j, k, m;
        ++lineno;
        int bad = sscanf(line, " %lf %d %d %d %d ", &v, &i, &j, &k, &m) != 5;
        double *dst = NULL;
        if (!bad) {
            const int li = (int)l;
            if (k != 0) {
                bad = i < 1 || j < 1 || m < 1 || i > li || j > li || k < 1 || k > li || m > li;
                if (!bad)
                    dst = two + (size_t)(m - 1) * l2 * l + (size_t)(k - 1) * l2 +
                          (size_t)(j - 1) * l + (size_t)(i - 1);
            } else if (i != 0) {
                bad = i < 1 || j < 1 || i > li || j > li;
                if (!bad) dst = one + (size_t)(j - 1) * l + (size_t)(i - 1);
            } else {
                dst = core;
            }
        }
        if (bad) {
            free(one);
            free(two);
            fclose(fp);
            if (errline) *errline = lineno;
            return doci_fail(DOCI_E_LINE,
                             "Whilst reading the integrals, an error occured, wrong formatting at "
                             "line %d!",
                             lineno);
        }
        if (fabs(*dst) >= 1e-14)
            fprintf(stderr, "Doubly inputted value at line %d, hope you don't mind\n", lineno);
        *dst = v;
    }
    fclose(fp);
    *one_out = one;
    *two_out = two;
    return DOCI_OK;
}

int doci_read_fcidump(const char *path, unsigned *L, unsigned *Na, unsigned *Nb, double *core,
                      double **one, double **two, int *errline)
{
    return dh_read_fcidump(path, L, Na, Nb, core, one, two, errline);
}

void doci_free(void *p) { free(p); }

void read_fcidump(char *filename, unsigned int *NORB, unsigned int *NALPHA, unsigned int *NBETA,
                  double *core_energy, double **one_p_int, double **two_p_int)
{
    int line = 0;
    const int rc = dh_read_fcidump(filename, NORB, NALPHA, NBETA, core_energy, one_p_int,
                                   two_p_int, &line);
    if (rc == DOCI_OK) return;
    if (rc == DOCI_E_IO)
        fprintf(stderr, "Error reading fcidump file\n");
    else
        fprintf(stderr, "%s\n", doci_last_error());
    exit(EXIT_FAILURE);
}

/* ------------------------------------------------------------------------ */

int doci_write_fcidump(const char *path, unsigned L, unsigned N, double core, const double *one,
                       const double *two)
{
    FILE *fp = fopen(path, "w");
    if (!fp) return doci_fail(DOCI_E_IO, "cannot write %s", path);
    const size_t l = L, l2 = l * l, l3 = l2 * l;
    fprintf(fp, " &FCI NORB=%u,NELEC=%u,MS2=0,\n  ORBSYM=", L, 2 * N);
    for (unsigned i = 0; i < L; ++i) fprintf(fp, "1,");
    fprintf(fp, "\n  ISYM=1,\n &END\n");
    /* canonical two-body entries i>=j, k>=l, (ij)>=(kl), 1-based */
    for (size_t i = 1; i <= l; ++i)
        for (size_t j = 1; j <= i; ++j)
            for (size_t k = 1; k <= i; ++k)
                for (size_t m = 1; m <= k; ++m) {
                    if (k == i && m > j) break;
                    const double v = two[(m - 1) * l3 + (k - 1) * l2 + (j - 1) * l + (i - 1)];
                    if (fabs(v) < 1e-12) continue;
                    fprintf(fp, "%.17e %4zu %4zu %4zu %4zu\n", v, i, j, k, m);
                }
    for (size_t i = 1; i <= l; ++i)
        for (size_t j = 1; j <= i; ++j) {
            const double v = one[(j - 1) * l + (i - 1)];
            if (fabs(v) < 1e-12) continue;
            fprintf(fp, "%.17e %4zu %4zu %4d %4d\n", v, i, j, 0, 0);
        }
    fprintf(fp, "%.17e %4d %4d %4d %4d", core, 0, 0, 0, 0);
    return fclose(fp) == 0 ? DOCI_OK : doci_fail(DOCI_E_IO, "write %s", path);
}

/* ------------------------------------------------------------------------ */
/* splitmix64 -> 53-bit uniform -> Box-Muller (cosine branch)                */

typedef struct {
    uint64_t s;
} rng_t;

static uint64_t splitmix64(rng_t *r)
{
    uint64_t z = (r->s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static double uniform53(rng_t *r) { return (double)(splitmix64(r) >> 11) * 0x1.0p-53; }

static double normal(rng_t *r)
{
    const double u1 = 1.0 - uniform53(r); /* (0, 1] */
    const double u2 = uniform53(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

int doci_gen_integrals(unsigned L, uint64_t seed, double *core, double *one, double *two)
{
    if (L < 2 || L > DK_MAX_L) return doci_fail(DOCI_E_ARG, "generator needs 2 <= L <= 64");
    const size_t l = L, l2 = l * l, l3 = l2 * l;
    rng_t r = {seed};
    memset(one, 0, sizeof(double) * l2);
    memset(two, 0, sizeof(double) * l2 * l2);
    double *h = calloc(l2, sizeof(double));
    const size_t npair = l * (l + 1) / 2, nP = 2 * l;
    double *B = calloc(npair * nP, sizeof(double)); /* B[pair * nP + P] */
    if (!h || !B) {
        free(h);
        free(B);
        return doci_fail(DOCI_E_NOMEM, "generator scratch");
    }
    for (size_t i = 0; i < l; ++i)
        h[i * l + i] = -12.0 + 11.0 * (double)i / (double)(l - 1) + 0.05 * normal(&r);
    for (size_t i = 0; i < l; ++i)
        for (size_t j = i + 1; j < l; ++j) {
            const double v = 0.1 * exp(-(double)(j - i) / 4.0) * normal(&r);
            h[i * l + j] = h[j * l + i] = v;
        }
    /* pair index of (i <= j) in row-major order */
    for (size_t P = 1; P <= nP; ++P) {
        const double sp = 0.3 / sqrt((double)P);
        size_t pi = 0;
        for (size_t i = 0; i < l; ++i)
            for (size_t j = i; j < l; ++j, ++pi)
                B[pi * nP + (P - 1)] = sp * exp(-(double)(j - i) / 3.0) * normal(&r);
    }
    /* pair index lookup for unordered (a,b) */
#define PIDX(a, b) ((a) <= (b) ? (a) * l - (a) * ((a) - 1) / 2 + ((b) - (a)) \
                               : (b) * l - (b) * ((b) - 1) / 2 + ((a) - (b)))
    /* raw layout, canonical entries only (0-based i>=j, k>=m, (ij)>=(km)) */
    for (size_t i = 0; i < l; ++i)
        for (size_t j = 0; j <= i; ++j)
            for (size_t k = 0; k <= i; ++k)
                for (size_t m = 0; m <= k; ++m) {
                    if (k == i && m > j) break;
                    const double *b1 = B + PIDX(i, j) * nP, *b2 = B + PIDX(k, m) * nP;
                    double v = 0;
                    for (size_t P = 0; P < nP; ++P) v += b1[P] * b2[P];
                    if (fabs(v) < 1e-12) continue;
                    two[m * l3 + k * l2 + j * l + i] = v;
                }
#undef PIDX
    for (size_t i = 0; i < l; ++i)
        for (size_t j = 0; j <= i; ++j) {
            const double v = h[i * l + j];
            if (fabs(v) >= 1e-12) one[j * l + i] = v;
        }
    *core = 10.0;
    free(h);
    free(B);
    return DOCI_OK;
}

int doci_gen_vector(int64_t n, uint64_t seed, double *out)
{
    if (n < 1) return doci_fail(DOCI_E_ARG, "vector length");
    rng_t r = {seed};
    double s = 0;
    for (int64_t i = 0; i < n; ++i) {
        out[i] = normal(&r);
        s += out[i] * out[i];
    }
    s = 1.0 / sqrt(s);
    for (int64_t i = 0; i < n; ++i) out[i] *= s;
    return DOCI_OK;
}
