/* operator.c -- the device-backed DOCI operator (drop-in for src/doci.c:69-238).
 *
 * fill_doci_data builds the integral tables on the host exactly as the
 * reference does (transform, exchange table), uploads them once, and
 * computes the diagonal on the GPU.  doci_matvec keeps the reference's
 * host-pointer contract; the device-resident solvers call dk_sigma directly.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "internal.h"

static __thread char g_err[512];

const char *doci_last_error(void) { return g_err; }

int doci_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

double dh_now(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

int dh_is_op(const void *p) { return p && ((const doci_op *)p)->magic == DOCI_OP_MAGIC; }

void doci_op_free(doci_op *op)
{
    if (!op) return;
    /* buffers go back to the process-wide cache: nothing may still use them */
    if (op->stream) hipStreamSynchronize(op->stream);
    dh_dfree(op->d_tables);
    dh_dfree(op->d_D);
    dh_dfree(op->d_red);
    dh_dfree(op->d_out);
    dh_dfree(op->d_idx);
    dh_dfree(op->d_io[0]);
    dh_dfree(op->d_io[1]);
    dh_dfree(op->pool);
    dh_slabs_free(op);
    dh_stars_free(op);
    if (op->h_out) hipHostFree(op->h_out);
    if (op->stream) hipStreamDestroy(op->stream);
    free(op->K);
    free(op->J);
    free(op->X);
    free(op->h);
    free(op->binom);
    free(op->D_host);
    op->magic = 0;
    free(op);
}

/* (ij|kl) as it stands after the reference's in-place eight-fold unpacking
 * (src/doci.c:640-655), without unpacking the tensor.  The unpacking visits
 * the tuples with i>=j, i>=k, k>=l in loop order (i, j, k, l ascending) and
 * copies each non-zero one to its eight index permutations, so every entry
 * of a permutation orbit ends up with the first non-zero value among the
 * orbit's loop-visited tuples (a later visited tuple of the same orbit reads
 * that copy back); an orbit with none keeps its raw entries. */
static double eri_unpacked(const double *two, size_t L, size_t i, size_t j, size_t k, size_t l)
{
    const size_t L2 = L * L, L3 = L2 * L;
    const size_t pm[8][4] = {{i, j, k, l}, {k, l, i, j}, {j, i, l, k}, {l, k, j, i},
                             {j, i, k, l}, {l, k, i, j}, {i, j, l, k}, {k, l, j, i}};
    /* the loop-visited permutations in visiting (lexicographic) order, read
     * until the first non-zero one: for a symmetric tensor that is a single
     * read instead of one per permutation (the reads are cache misses) */
    int ord[8], nv = 0;
    for (int q = 0; q < 8; ++q) {
        const size_t *t = pm[q];
        if (!(t[0] >= t[1] && t[0] >= t[2] && t[2] >= t[3])) continue;
        int pos = nv; /* insertion by lexicographic order (equal tuples: either order) */
        while (pos > 0) {
            const size_t *u = pm[ord[pos - 1]];
            int e = 0;
            while (e < 3 && t[e] == u[e]) ++e;
            if (!(t[e] < u[e])) break;
            ord[pos] = ord[pos - 1];
            --pos;
        }
        ord[pos] = q;
        ++nv;
    }
    for (int r = 0; r < nv; ++r) {
        const size_t *t = pm[ord[r]];
        const double v = two[t[0] + L * t[1] + L2 * t[2] + L3 * t[3]];
        if (v != 0) return v;
    }
    return two[i + L * j + L2 * k + L3 * l]; /* no visited tuple is non-zero: the raw entry */
}

/* The DOCI tables of raw FCIDUMP integrals (one[L^2], two[L^4] in the
 * reference's layout): K, J, X (src/doci.c:667-683 and the diagonal's
 * (pp|qq), (pq|qp)), h[p] = h'_pp = h_pp - 1/2 sum_j (pj|jp) summed with j
 * ascending (src/doci.c:657-664), each bit-identical to the reference's full
 * transform followed by the lookups, in O(L^2) instead of O(L^4). */
int dh_build_tables(unsigned L, unsigned N, const double *one_raw, const double *two_raw,
                    double *K, double *J, double *X, double *h, uint32_t *binom)
{
    const size_t l = L;
    for (size_t p = 0; p < l; ++p)
        for (size_t q = 0; q < l; ++q) {
            J[p * l + q] = eri_unpacked(two_raw, l, p, p, q, q);                   /* (pp|qq) */
            X[p * l + q] = eri_unpacked(two_raw, l, p, q, q, p);                   /* (pq|qp) */
            K[p * l + q] = p == q ? 0.0 : eri_unpacked(two_raw, l, p, q, p, q);    /* (pq|pq) */
        }
    for (size_t p = 0; p < l; ++p) {
        double acc = 0;
        for (size_t j = 0; j < l; ++j) acc += X[p * l + j];
        h[p] = one_raw[p * (l + 1)] - 0.5 * acc;
    }
    if (binom) { /* C(x, j) by Pascal's rule (exact: every entry used is < 2^31) */
        const int N2 = (int)N + 2;
        for (size_t x = 0; x <= l; ++x)
            for (int j = 0; j < N2; ++j)
                binom[x * N2 + j] = j == 0 ? 1u
                                    : x == 0 ? 0u
                                             : binom[(x - 1) * N2 + j - 1] + binom[(x - 1) * N2 + j];
    }
    return DOCI_OK;
}

int doci_integral_tables(unsigned L, const double *one, const double *two, double *K, double *J,
                         double *X, double *h)
{
    if (L == 0 || L > DK_MAX_L) return doci_fail(DOCI_E_ORBITALS, "Too many orbitals (L=%u).", L);
    if (!one || !two || !K || !J || !X || !h) return doci_fail(DOCI_E_ARG, "null argument");
    return dh_build_tables(L, 1, one, two, K, J, X, h, NULL);
}

/* Allocates an operator with empty host tables for (L, N). */
static int op_alloc(unsigned L, unsigned N, double core, doci_op **out)
{
    *out = NULL;
    if (L == 0 || L > DK_MAX_L) return doci_fail(DOCI_E_ORBITALS, "Too many orbitals (L=%u).", L);
    if (N == 0 || N > L) return doci_fail(DOCI_E_DIM, "invalid number of pairs N=%u", N);
    const uint64_t n = dh_binom(N, L);
    if (n >= (1ull << 31))
        return doci_fail(DOCI_E_DIM, "C(%u,%u) = %llu exceeds 2^31-1", L, N, (unsigned long long)n);
    const size_t l = L, l2 = l * l;
    doci_op *op = calloc(1, sizeof *op);
    if (!op) return doci_fail(DOCI_E_NOMEM, "host tables");
    op->magic = DOCI_OP_MAGIC;
    op->L = L;
    op->N = N;
    op->n = (int64_t)n;
    op->core = core;
    const int N2 = (int)N + 2;
    op->K = malloc(sizeof(double) * l2);
    op->J = malloc(sizeof(double) * l2);
    op->X = malloc(sizeof(double) * l2);
    op->h = malloc(sizeof(double) * l);
    op->binom = malloc(sizeof(uint32_t) * (l + 1) * (size_t)N2);
    if (!op->K || !op->J || !op->X || !op->h || !op->binom) {
        doci_op_free(op);
        return doci_fail(DOCI_E_NOMEM, "host tables");
    }
    *out = op;
    return DOCI_OK;
}

static int op_device_setup(doci_op *op);

/* Takes ownership of nothing: one/two are transformed in a private copy. */
int doci_op_from_integrals(unsigned L, unsigned N, double core, const double *one_raw,
                           const double *two_raw, doci_op **out)
{
    doci_op *op = NULL;
    int rc = op_alloc(L, N, core, &op);
    if (rc != DOCI_OK) return rc;
    if (dh_build_tables(L, N, one_raw, two_raw, op->K, op->J, op->X, op->h, op->binom) != DOCI_OK) {
        doci_op_free(op);
        return doci_fail(DOCI_E_NOMEM, "host tables");
    }
    rc = op_device_setup(op);
    if (rc != DOCI_OK) {
        doci_op_free(op);
        return rc;
    }
    *out = op;
    return DOCI_OK;
}

int dh_op_from_tables(unsigned L, unsigned N, double core, const double *K, const double *J,
                      const double *X, const double *h, doci_op **out)
{
    doci_op *op = NULL;
    int rc = op_alloc(L, N, core, &op);
    if (rc != DOCI_OK) return rc;
    const size_t l = L, l2 = l * l;
    const int N2 = (int)N + 2;
    memcpy(op->K, K, sizeof(double) * l2);
    memcpy(op->J, J, sizeof(double) * l2);
    memcpy(op->X, X, sizeof(double) * l2);
    memcpy(op->h, h, sizeof(double) * l);
    for (size_t x = 0; x <= l; ++x)
        for (int j = 0; j < N2; ++j) op->binom[x * N2 + j] = (uint32_t)dh_binom((unsigned)j, (unsigned)x);
    rc = op_device_setup(op);
    if (rc != DOCI_OK) {
        doci_op_free(op);
        return rc;
    }
    *out = op;
    return DOCI_OK;
}

/* Uploads the host tables, computes the diagonal. */
static int op_device_setup(doci_op *op)
{
    const unsigned L = op->L, N = op->N;
    const uint64_t n = (uint64_t)op->n;
    const double core = op->core;
    const size_t l = L, l2 = l * l;
    const int N2 = (int)N + 2;
    /* device side */
    int rc = DOCI_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return doci_fail(DOCI_E_HIP, "no HIP device available");
#define TRY(expr)                                                                     \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            rc = doci_fail(DOCI_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));       \
            goto fail;                                                                \
        }                                                                             \
    } while (0)
    TRY(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking));
    const size_t tb = sizeof(double) * (3 * l2 + l) + sizeof(uint32_t) * (l + 1) * (size_t)N2;
    TRY(dh_dmalloc(&op->d_tables, tb));
    {
        char *b = (char *)op->d_tables;
        double *dK = (double *)b, *dJ = dK + l2, *dX = dJ + l2, *dh = dX + l2;
        uint32_t *db = (uint32_t *)(dh + l);
        TRY(hipMemcpyAsync(dK, op->K, sizeof(double) * l2, hipMemcpyHostToDevice, op->stream));
        TRY(hipMemcpyAsync(dJ, op->J, sizeof(double) * l2, hipMemcpyHostToDevice, op->stream));
        TRY(hipMemcpyAsync(dX, op->X, sizeof(double) * l2, hipMemcpyHostToDevice, op->stream));
        TRY(hipMemcpyAsync(dh, op->h, sizeof(double) * l, hipMemcpyHostToDevice, op->stream));
        TRY(hipMemcpyAsync(db, op->binom, sizeof(uint32_t) * (l + 1) * (size_t)N2,
                           hipMemcpyHostToDevice, op->stream));
        op->dt.L = L;
        op->dt.N = N;
        op->dt.n = (uint32_t)n;
        op->dt.core = core;
        op->dt.K = dK;
        op->dt.J = dJ;
        op->dt.X = dX;
        op->dt.h = dh;
        op->dt.binom = db;
    }
    TRY(dh_dmalloc((void **)&op->d_D, sizeof(double) * n));
    TRY(dh_dmalloc((void **)&op->d_red, sizeof(double) * dk_red_scratch_doubles(DK_MAX_VEC + 8)));
    TRY(dh_dmalloc((void **)&op->d_out, sizeof(double) * (DK_MAX_VEC + 8)));
    TRY(dh_dmalloc((void **)&op->d_idx, sizeof(int64_t)));
    TRY(hipHostMalloc((void **)&op->h_out, sizeof(double) * (DK_MAX_VEC + 8), 0));
    TRY((hipError_t)dk_diagonal(&op->dt, op->d_D, op->stream));
    TRY(hipStreamSynchronize(op->stream));
#undef TRY
    /* star / slab tables are built on first use of those kernels */
    op->sigma_kind = n >= DOCI_STAR_MIN_N ? 0 : 1;
    return DOCI_OK;
fail:
    return rc;
}

int doci_op_from_fcidump(const char *path, doci_op **out, int *errline)
{
    unsigned L, Na, Nb;
    double core, *one = NULL, *two = NULL;
    *out = NULL;
    int rc = dh_read_fcidump(path, &L, &Na, &Nb, &core, &one, &two, errline);
    if (rc != DOCI_OK) return rc;
    if (Na != Nb) {
        free(one);
        free(two);
        return doci_fail(DOCI_E_SPIN,
                         "Number of up and down electrons not equal in FCIDUMP-file.\n"
                         "This is obligatory for DOCI.");
    }
    rc = doci_op_from_integrals(L, Na, core, one, two, out);
    free(one);
    free(two);
    return rc;
}

int64_t doci_op_dim(const doci_op *op) { return op->n; }
unsigned doci_op_L(const doci_op *op) { return op->L; }
unsigned doci_op_N(const doci_op *op) { return op->N; }
void *doci_op_stream(doci_op *op) { return (void *)op->stream; }

int doci_dev_copy(void *dst, const void *src, int64_t n)
{
    DKCHK(dk_convert(src, 0, dst, 0, n, 1.0, 0));
    HIPCHK(hipDeviceSynchronize());
    return DOCI_OK;
}

int doci_op_diagonal(doci_op *op, double *out)
{
    HIPCHK(hipMemcpyAsync(out, op->d_D, sizeof(double) * op->n, hipMemcpyDeviceToHost, op->stream));
    HIPCHK(hipStreamSynchronize(op->stream));
    return DOCI_OK;
}

/* Builds the tables kernel `kind` needs, once. */
static int ensure_tables(doci_op *op, int kind)
{
    if ((kind == 0 || kind == 2) && !op->stars) return dh_stars_build(op, 0, 0, 0);
    if ((kind == 3 || kind == 4) && !op->slabs) return dh_slabs_build(op, 0, 0, 0);
    return DOCI_OK;
}

int doci_op_sigma_dev(doci_op *op, const void *in, void *out, int prec, double shift)
{
    if (!in || !out || in == out) return doci_fail(DOCI_E_ARG, "sigma needs distinct buffers");
    const int rc = ensure_tables(op, op->sigma_kind);
    if (rc != DOCI_OK) return rc;
    switch (op->sigma_kind) {
    case 0:
    case 2: return dh_stars_sigma(op, in, out, prec, shift, op->sigma_kind == 2);
    case 3:
    case 4: return dh_slabs_sigma(op, in, out, prec, shift);
    default: break;
    }
    DKCHK(dk_sigma(&op->dt, op->d_D, shift, in, out, prec, op->stream));
    return DOCI_OK;
}

int doci_op_sigma_dev_noalloc(doci_op *op, const void *in, void *out, int prec, double shift)
{
    if (!dh_is_op(op)) return doci_fail(DOCI_E_ARG, "not an operator");
    if (!in || !out || in == out) return doci_fail(DOCI_E_ARG, "sigma needs distinct buffers");
    if (prec != 0 && prec != 1) return doci_fail(DOCI_E_ARG, "prec must be 0 (FP64) or 1 (FP32)");
    return dh_stars_sigma_static(op, in, out, prec, shift);
}

int doci_op_set_kernel(doci_op *op, int kind)
{
    if (kind < 0 || kind > 4) return doci_fail(DOCI_E_ARG, "kernel kind must be 0..4");
    const int rc = ensure_tables(op, kind);
    if (rc != DOCI_OK) return rc;
    op->sigma_kind = kind;
    return DOCI_OK;
}

int doci_op_kernel(const doci_op *op) { return op->sigma_kind; }

int64_t doci_op_check_violations(doci_op *op)
{
    const int64_t a = op->stars ? dh_stars_violations(op) : 0;
    if (a < 0) return a;
    const int64_t b = op->slabs ? dh_slabs_violations(op) : 0;
    return b < 0 ? b : a + b;
}

int doci_op_star_split(doci_op *op, unsigned *a, unsigned *b, unsigned *ngroups,
                       unsigned *nstargroups, size_t *zslots)
{
    const int rc = ensure_tables(op, 0);
    if (rc != DOCI_OK) return rc;
    dh_stars_info(op, a, b, ngroups, nstargroups, zslots);
    return DOCI_OK;
}

int doci_op_slab_split(doci_op *op, unsigned *a, unsigned *b, unsigned *ngroups)
{
    const int rc = ensure_tables(op, 3);
    if (rc != DOCI_OK) return rc;
    dh_slabs_info(op, a, b, ngroups);
    return op->slabs ? DOCI_OK : doci_fail(DOCI_E_ARG, "no slab tables");
}

static int ensure_io(doci_op *op)
{
    for (int k = 0; k < 2; ++k)
        if (!op->d_io[k]) HIPCHK(dh_dmalloc(&op->d_io[k], sizeof(double) * op->n));
    return DOCI_OK;
}

int doci_op_sigma(doci_op *op, const double *in, double *out)
{
    int rc = ensure_io(op);
    if (rc) return rc;
    const size_t b = sizeof(double) * op->n;
    HIPCHK(hipMemcpyAsync(op->d_io[0], in, b, hipMemcpyHostToDevice, op->stream));
    rc = doci_op_sigma_dev(op, op->d_io[0], op->d_io[1], 0, 0.0);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, op->d_io[1], b, hipMemcpyDeviceToHost, op->stream));
    HIPCHK(hipStreamSynchronize(op->stream));
    return DOCI_OK;
}

/* Copies nvals reduction results from d_out to the host (synchronises). */
int dh_fetch(dh_ctx *c, int nvals, double *host)
{
    HIPCHK(hipMemcpyAsync(c->h_out, c->d_out, sizeof(double) * nvals, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    memcpy(host, c->h_out, sizeof(double) * nvals);
    return DOCI_OK;
}

static int apply_op(dh_ctx *c, const void *in, void *out, int prec, double shift)
{
    return doci_op_sigma_dev((doci_op *)c->user, in, out, prec, shift);
}

void dh_ctx_of_op(doci_op *op, dh_ctx *c)
{
    memset(c, 0, sizeof *c);
    c->n = op->n;
    c->stream = op->stream;
    c->d_D = op->d_D;
    c->d_red = op->d_red;
    c->d_out = op->d_out;
    c->h_out = op->h_out;
    c->d_idx = op->d_idx;
    c->apply = apply_op;
    c->user = op;
    c->fp32_ok = 1;
    c->pool = &op->pool;
    c->pool_bytes = &op->pool_bytes;
}

/* ------------------------------------------------------------------------ */
/* drop-in entry points (include/doci.h)                                    */

void fill_doci_data(char *dumpfil, void **dat)
{
    doci_op *op = NULL;
    int line = 0;
    const int rc = doci_op_from_fcidump(dumpfil, &op, &line);
    if (rc != DOCI_OK) {
        if (rc == DOCI_E_IO)
            fprintf(stderr, "Error reading fcidump file\n");
        else if (rc == DOCI_E_SPIN) {
            fprintf(stderr, "%s\n", doci_last_error());
            printf("Exiting...\n");
        } else if (rc == DOCI_E_ORBITALS) {
            fprintf(stderr, "Too many orbitals.\n");
            printf("Exiting...\n");
        } else
            fprintf(stderr, "%s\n", doci_last_error());
        exit(EXIT_FAILURE);
    }
    printf("----------\n"
           "\t** executing DOCI (1 GPU)\n"
           "\t** N = (%u,%u)\n"
           "\t** L = %u\n"
           "\t** basis size : %lld\n"
           "----------\n",
           op->N, op->N, op->L, (long long)op->n);
    *dat = op;
}

double *doci_get_diagonal(void *dat)
{
    doci_op *op = dat;
    if (!op->D_host) {
        op->D_host = malloc(sizeof(double) * op->n);
        if (!op->D_host || doci_op_diagonal(op, op->D_host) != DOCI_OK) {
            fprintf(stderr, "doci_get_diagonal: %s\n", doci_last_error());
            exit(EXIT_FAILURE);
        }
    }
    return op->D_host;
}

int doci_get_basis_size(void *dat) { return (int)((doci_op *)dat)->n; }

void doci_cleanup_data(void **dat)
{
    doci_op_free((doci_op *)*dat);
    *dat = NULL;
}

void doci_matvec(double *vec, double *result, void *dat)
{
    if (doci_op_sigma((doci_op *)dat, vec, result) != DOCI_OK) {
        fprintf(stderr, "doci_matvec: %s\n", doci_last_error());
        exit(EXIT_FAILURE);
    }
}
