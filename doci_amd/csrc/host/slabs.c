/* slabs.c -- slab and group tables of the slab-blocked sigma kernel.
 *
 * B is the largest set of top orbitals whose biggest slab, C(b, m) over the
 * m that occur, fits the kernel's LDS staging (emax determinants).  Slabs
 * with the same m are packed into groups (whole slabs, at most gmax), and
 * groups are ordered by the rank of their first determinant so that a
 * contiguous run of groups covers a contiguous region of the vector. */
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define SLAB_LDS_BUDGET (64 * 1024)
#define SLAB_GMAX 16
#define SLAB_EMAX 4096

typedef struct {
    dk_slabs d;
    void *dmem;
    uint32_t *d_bad; /* violation counter of the bounds-checked kernel */
    size_t lds[2];
} dh_slabs;

typedef struct {
    uint32_t start, first;
    uint32_t word; /* nslab | m << 16 */
} grp_t;

static int cmp_grp(const void *x, const void *y)
{
    const grp_t *a = x, *b = y;
    return a->start < b->start ? -1 : a->start > b->start;
}

void dh_slabs_free(doci_op *op)
{
    dh_slabs *s = op->slabs;
    if (!s) return;
    dh_dfree(s->dmem);
    dh_dfree(s->d_bad);
    free(s);
    op->slabs = NULL;
}

int dh_slabs_build(doci_op *op, unsigned emax_req, unsigned gmax, int xcd_runs)
{
    const unsigned L = op->L, N = op->N;
    const uint64_t n = (uint64_t)op->n;
    unsigned emax = emax_req ? emax_req : SLAB_EMAX;
    if (!gmax) gmax = SLAB_GMAX;
    /* LDS: K table, binomials, FP64 staging of emax determinants and the
     * per-slab neighbour tables of gmax slabs; b is the largest split whose
     * biggest slab fits emax, and gmax shrinks until the tables fit too. */
    const size_t fixed = 8 * (size_t)L * L + 4 * (size_t)(L + 1) * (N + 2) + 64;
    if (fixed + 8 * (size_t)emax > SLAB_LDS_BUDGET)
        emax = (unsigned)((SLAB_LDS_BUDGET - fixed) / 8 / 2);
    unsigned b = L;
    for (; b >= 1; --b) {
        const unsigned a = L - b;
        const unsigned mlo = N > a ? N - a : 0, mhi = N < b ? N : b;
        uint64_t mx = 0;
        for (unsigned m = mlo; m <= mhi; ++m)
            if (dh_binom(m, b) > mx) mx = dh_binom(m, b);
        if (mx <= emax) break;
    }
    if (b < 1) return doci_fail(DOCI_E_ARG, "no slab split fits the LDS budget");
    const unsigned a = L - b;
    const unsigned mlo = N > a ? N - a : 0, mhi = N < b ? N : b;
    uint64_t nslab = 0;
    unsigned aamax = 0, abmax = 1, bamax = 1;
    for (unsigned m = mlo; m <= mhi; ++m) {
        const unsigned k = N - m;
        nslab += dh_binom(k, a);
        if (k * (a - k) > aamax) aamax = k * (a - k);
        if (k > abmax) abmax = k;
        if (a - k > bamax) bamax = a - k;
    }
    while (gmax > 1 && fixed + 8 * (size_t)emax +
                               (size_t)gmax * (6 * (size_t)aamax + 5 * (size_t)(abmax + bamax) + 48) >
                           SLAB_LDS_BUDGET)
        --gmax;
    if (aamax == 0) aamax = 1;
    uint32_t *start = malloc(sizeof(uint32_t) * nslab);
    uint64_t *amask = malloc(sizeof(uint64_t) * nslab);
    grp_t *grp = malloc(sizeof(grp_t) * nslab);
    uint32_t *gw = NULL;
    int rc = DOCI_OK;
    if (!start || !amask || !grp) {
        rc = doci_fail(DOCI_E_NOMEM, "slab tables");
        goto out;
    }
    uint64_t is = 0, ng = 0;
    int idx[64];
    for (unsigned m = mlo; m <= mhi; ++m) {
        const unsigned k = N - m;
        const uint64_t S = dh_binom(m, b);
        /* all k-subsets of A in lexicographic order */
        for (unsigned e = 0; e < k; ++e) idx[e] = (int)e;
        uint32_t gfirst = (uint32_t)is, gcount = 0;
        for (;;) {
            uint64_t msk = 0, sa = 0;
            for (unsigned e = 0; e < k; ++e) {
                msk |= 1ull << idx[e];
                sa += dh_binom(N - e, L - 1 - (unsigned)idx[e]);
            }
            if (gcount == gmax || (gcount + 1) * S > emax) {
                grp[ng].first = gfirst;
                grp[ng].start = start[gfirst];
                grp[ng].word = gcount | (m << 16);
                ++ng;
                gfirst = (uint32_t)is;
                gcount = 0;
            }
            start[is] = (uint32_t)(n - S - sa);
            amask[is] = msk;
            ++is;
            ++gcount;
            /* next combination */
            int e = (int)k - 1;
            while (e >= 0 && idx[e] == (int)(a - k) + e) --e;
            if (e < 0) break;
            ++idx[e];
            for (int f = e + 1; f < (int)k; ++f) idx[f] = idx[f - 1] + 1;
        }
        grp[ng].first = gfirst;
        grp[ng].start = start[gfirst];
        grp[ng].word = gcount | (m << 16);
        ++ng;
    }
    qsort(grp, ng, sizeof(grp_t), cmp_grp);
    gw = malloc(sizeof(uint32_t) * 2 * ng);
    if (!gw) {
        rc = doci_fail(DOCI_E_NOMEM, "slab groups");
        goto out;
    }
    for (uint64_t g = 0; g < ng; ++g) {
        gw[2 * g] = grp[g].first;
        gw[2 * g + 1] = grp[g].word;
    }
    dh_slabs *s = calloc(1, sizeof *s);
    if (!s) {
        rc = doci_fail(DOCI_E_NOMEM, "slab descriptor");
        goto out;
    }
    const size_t b_amask = sizeof(uint64_t) * nslab, b_start = sizeof(uint32_t) * nslab,
                 b_grp = sizeof(uint32_t) * 2 * ng;
    hipError_t e = dh_dmalloc(&s->dmem, b_amask + b_start + b_grp);
    if (e != hipSuccess) {
        free(s);
        rc = doci_fail(DOCI_E_NOMEM, "slab tables on device: %s", hipGetErrorString(e));
        goto out;
    }
    char *dm = s->dmem;
    e = hipMemcpy(dm, amask, b_amask, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dm + b_amask, start, b_start, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dm + b_amask + b_start, gw, b_grp, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        dh_dfree(s->dmem);
        free(s);
        rc = doci_fail(DOCI_E_HIP, "slab table upload: %s", hipGetErrorString(e));
        goto out;
    }
    s->d.a = a;
    s->d.b = b;
    s->d.ngroup = (uint32_t)ng;
    s->d.emax = emax;
    s->d.gmax = gmax;
    s->d.xcd_runs = (uint32_t)xcd_runs;
    s->d.aamax = aamax;
    s->d.abmax = abmax;
    s->d.bamax = bamax;
    s->d.slab_amask = (const uint64_t *)dm;
    s->d.slab_start = (const uint32_t *)(dm + b_amask);
    s->d.group = (const uint32_t *)(dm + b_amask + b_start);
    s->lds[0] = dk_sigma_slab_lds(&op->dt, &s->d, 0);
    s->lds[1] = dk_sigma_slab_lds(&op->dt, &s->d, 1);
    if (s->lds[0] > SLAB_LDS_BUDGET) {
        rc = doci_fail(DOCI_E_ARG, "slab kernel LDS %zu exceeds budget", s->lds[0]);
        dh_dfree(s->dmem);
        free(s);
        goto out;
    }
    e = dh_dmalloc((void **)&s->d_bad, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(s->d_bad, 0, sizeof(uint32_t));
    if (e != hipSuccess) {
        rc = doci_fail(DOCI_E_HIP, "slab counter: %s", hipGetErrorString(e));
        dh_dfree(s->dmem);
        free(s);
        goto out;
    }
    op->slabs = s;
out:
    free(start);
    free(amask);
    free(grp);
    free(gw);
    return rc;
}

int dh_slabs_sigma(doci_op *op, const void *in, void *out, int prec, double shift)
{
    dh_slabs *s = op->slabs;
    DKCHK(dk_sigma_slab_checked(&op->dt, &s->d, op->d_D, shift, in, out, prec,
                                op->sigma_kind == 4 ? s->d_bad : NULL, op->stream));
    return DOCI_OK;
}

int doci_op_tune(doci_op *op, const char *key, int value)
{
    dh_slabs *s = op->slabs;
    if (!strcmp(key, "star_pass")) return dh_stars_set_pass(op, value);
    if (!strcmp(key, "star_lds") || !strcmp(key, "star_sgmax") || !strcmp(key, "star_emax")) {
        /* rebuild the star tables with one limit changed (0 = default) */
        static const char *keys[3] = {"star_lds", "star_sgmax", "star_emax"};
        unsigned cur[3];
        dh_stars_limits(op, &cur[0], &cur[1], &cur[2]);
        for (int i = 0; i < 3; ++i)
            if (!strcmp(key, keys[i])) cur[i] = (unsigned)value;
        if (value < 0 || (cur[1] && cur[1] > 0xffff)) return doci_fail(DOCI_E_ARG, "bad %s", key);
        dh_stars_free(op);
        return dh_stars_build(op, cur[0], cur[1], cur[2]);
    }
    if (!s) { /* built on first use */
        const int rc = dh_slabs_build(op, 0, 0, 0);
        if (rc != DOCI_OK) return rc;
        s = op->slabs;
    }
    if (!strcmp(key, "xcd_runs")) {
        s->d.xcd_runs = value != 0;
        return DOCI_OK;
    }
    if (!strcmp(key, "emax") || !strcmp(key, "gmax")) {
        const unsigned emax = !strcmp(key, "emax") ? (unsigned)value : s->d.emax;
        const unsigned gmax = !strcmp(key, "gmax") ? (unsigned)value : s->d.gmax;
        const int runs = (int)s->d.xcd_runs;
        if (emax < 16 || gmax < 1 || gmax > 0xffff) return doci_fail(DOCI_E_ARG, "bad %s", key);
        dh_slabs_free(op);
        return dh_slabs_build(op, emax, gmax, runs);
    }
    return doci_fail(DOCI_E_ARG, "unknown tuning key %s", key);
}

int doci_op_time_sigma(doci_op *op, int prec, int reps, double *ms)
{
    const size_t es = prec ? sizeof(float) : sizeof(double);
    void *x = NULL, *y = NULL;
    hipEvent_t e0, e1;
    int rc = DOCI_OK;
    if (reps < 1) return doci_fail(DOCI_E_ARG, "reps");
    HIPCHK(dh_dmalloc(&x, es * op->n));
    HIPCHK(dh_dmalloc(&y, es * op->n));
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipMemsetAsync(x, 0, es * op->n, op->stream));
    rc = doci_op_sigma_dev(op, x, y, prec, 0.0); /* warm-up */
    if (!rc) HIPCHK(hipEventRecord(e0, op->stream));
    for (int r = 0; !rc && r < reps; ++r) rc = doci_op_sigma_dev(op, x, y, prec, 0.0);
    if (!rc) {
        float t = 0;
        HIPCHK(hipEventRecord(e1, op->stream));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        *ms = t / reps;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamSynchronize(op->stream);
    dh_dfree(x);
    dh_dfree(y);
    return rc;
}

int64_t dh_slabs_violations(doci_op *op)
{
    dh_slabs *s = op->slabs;
    uint32_t v = 0;
    if (!s) return 0;
    HIPCHK(hipStreamSynchronize(op->stream));
    HIPCHK(hipMemcpy(&v, s->d_bad, sizeof v, hipMemcpyDeviceToHost));
    return (int64_t)v;
}

void dh_slabs_info(const doci_op *op, unsigned *a, unsigned *b, unsigned *ngroup)
{
    const dh_slabs *s = op->slabs;
    *a = s ? s->d.a : 0;
    *b = s ? s->d.b : 0;
    *ngroup = s ? s->d.ngroup : 0;
}
