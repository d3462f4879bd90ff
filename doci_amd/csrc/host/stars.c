/* stars.c -- tables of the two-pass (star / main) sigma.
 *
 * b is the largest number of top orbitals (<= 16) for which the biggest
 * centre slab C(b, N-j) fits the star pass's LDS (FP64 value + FP64
 * accumulator per element) and the biggest slab fits the main pass's; a
 * larger b means fewer A pairs per determinant and so fewer Z slots to write
 * and read.  Slabs are stored by k (pairs in A), then in lexicographic order
 * of their A-part, so a slab's Z offset follows from its index in the sector.
 * Main-pass groups pack whole slabs of one sector and are ordered by rank;
 * star groups pack whole stars of one j. */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define STAR_LDS_DEFAULT (52 * 1024) /* three star-pass workgroups per CU */
#define MAIN_LDS_MAX (64 * 1024)
#define STAR_BMAX 16
#define STAR_PASS_DEFAULT 0 /* column star pass (tuning key "star_pass") */
#define STAR_SGMAX_DEFAULT 64
#define MAIN_EMAX_DEFAULT 4096  /* raised to the biggest slab */
#define MAIN_GSLABS 256         /* slabs per main-pass group, at most */

/* Device tables of one shape (L, N, limits).  They do not depend on the
 * integrals, so every operator of that shape shares them; shapes are built
 * once per process and kept until it ends. */
typedef struct star_shape {
    unsigned L, N;
    unsigned lim[3]; /* requested LDS budget, stars per group, main emax (0 = default) */
    dk_stars d;
    void *dmem;
    void *dtab;      /* B-mask, hop and insert tables */
    size_t zelems;
    struct star_shape *next;
} star_shape;

static star_shape *g_shapes;
static pthread_mutex_t g_shapes_lock = PTHREAD_MUTEX_INITIALIZER;

/* Per operator: the shared shape, its own slot buffer and violation counter. */
typedef struct {
    const star_shape *sh;
    void *Z;         /* slot buffer, FP64-sized so both precisions fit */
    uint32_t *d_bad;
    int pass;        /* star-pass variant (dk_stars.pass) */
} dh_stars;

typedef struct {
    uint32_t start, first, word;
} sgrp_t;

static int cmp_sgrp(const void *x, const void *y)
{
    const sgrp_t *a = x, *b = y;
    return a->start < b->start ? -1 : a->start > b->start;
}

/* LDS doubles a main-pass group of nsl slabs of S stages: sets of 4 (2, 1)
 * slabs side by side at a stride of 5 (3, 1) doubles (main_spl, main_stride
 * of the kernels) */
static uint64_t own_doubles(uint64_t nsl, uint64_t S)
{
    const uint64_t spl = (nsl & 3) == 0 ? 4 : (nsl & 1) == 0 ? 2 : 1;
    const uint64_t str = spl == 1 ? 1 : spl + 1;
    return nsl / spl * S * str;
}

static int valid_sector(unsigned k, unsigned N, unsigned a, unsigned b)
{
    return k <= N && k <= a && N - k <= b;
}

static unsigned nslot(unsigned k, unsigned m, unsigned a) { return k + ((m >= 1 && k < a) ? 1 : 0); }

static size_t star_lds_bytes(unsigned L, unsigned N, unsigned sgmax, unsigned ymax, unsigned selems)
{
#define A8(x) (((size_t)(x) + 7) & ~(size_t)7)
    /* + 64 zero / scratch centre entries; member tables with one more row of
     * absent members (the buffer-addressed column pass) */
    return A8(8 * (size_t)L * L) + A8(4 * (size_t)(L + 1) * (N + 2)) + 2 * A8(8 * ((size_t)selems + 64)) +
           2 * A8(4 * ((size_t)sgmax + 1) * ymax) + 2 * A8(4 * (size_t)sgmax) +
           A8((size_t)sgmax * (ymax > 8 ? ymax : 8));
#undef A8
}

void dh_stars_free(doci_op *op)
{
    dh_stars *s = op->stars;
    if (!s) return;
    dh_dfree(s->d_bad);
    dh_dfree(s->Z);
    free(s);
    op->stars = NULL;
}

/* lexicographic enumeration of the k-subsets of {0..a-1} */
static int next_comb(int *idx, unsigned k, unsigned a)
{
    int e = (int)k - 1;
    while (e >= 0 && idx[e] == (int)(a - k) + e) --e;
    if (e < 0) return 0;
    ++idx[e];
    for (int f = e + 1; f < (int)k; ++f) idx[f] = idx[f - 1] + 1;
    return 1;
}

/* lexicographic rank of the m-subset idx[0..m) of {0..b-1} */
static uint32_t lex_rank(const int *idx, unsigned m, unsigned b)
{
    uint64_t s = 0;
    for (unsigned e = 0; e < m; ++e) s += dh_binom(m - e, b - 1 - (unsigned)idx[e]);
    return (uint32_t)(dh_binom(m, b) - 1 - s);
}

/* The B-mask, hop and insert tables of dk_stars, uploaded to s->dtab. */
/* Rows of the hop table of m-pair masks: the m(b-m) hops, padded to a
 * multiple of 8 so that the main pass loads them 8 at a time unguarded. */
static size_t hop_rows(unsigned m, unsigned b) { return ((size_t)m * (b - m) + 7) / 8 * 8; }

static int build_b_tables(star_shape *s, unsigned b)
{
    size_t nmask = 0, nhop = 0, nins = 0;
    for (unsigned m = 0; m <= b; ++m) {
        const size_t S = dh_binom(m, b);
        s->d.bmask_off[m] = (uint32_t)nmask;
        s->d.hop_off[m] = (uint32_t)nhop;
        s->d.ins_off[m] = (uint32_t)nins;
        nmask += S;
        nhop += S * hop_rows(m, b);
        if (m < b) nins += S * b;
    }
    const size_t bytes_m = sizeof(uint16_t) * nmask, bytes_h = sizeof(uint32_t) * nhop,
                 bytes_i = sizeof(uint16_t) * nins;
    const size_t o_h = (bytes_m + 255) & ~(size_t)255, o_i = (o_h + bytes_h + 255) & ~(size_t)255;
    char *h = malloc(o_i + bytes_i + 1);
    if (!h) return doci_fail(DOCI_E_NOMEM, "B tables");
    uint16_t *bm = (uint16_t *)h, *ins = (uint16_t *)(h + o_i);
    uint32_t *hop = (uint32_t *)(h + o_h);
    int idx[17], jdx[17];
    for (unsigned m = 0; m <= b; ++m) {
        const size_t S = dh_binom(m, b);
        for (unsigned e = 0; e < m; ++e) idx[e] = (int)e;
        size_t r = 0;
        do {
            unsigned mask = 0;
            for (unsigned e = 0; e < m; ++e) mask |= 1u << idx[e];
            bm[s->d.bmask_off[m] + r] = (uint16_t)mask;
            size_t hh = 0;
            for (unsigned eo = 0; eo < m; ++eo) {
                const int o = idx[eo];
                for (unsigned t = 0; t < b; ++t) {
                    if ((mask >> t) & 1) continue;
                    /* beta - o + t, sorted */
                    unsigned q = 0;
                    for (unsigned e = 0; e < m; ++e)
                        if (idx[e] != o) jdx[q++] = idx[e];
                    unsigned pos = q;
                    while (pos > 0 && jdx[pos - 1] > (int)t) {
                        jdx[pos] = jdx[pos - 1];
                        --pos;
                    }
                    jdx[pos] = (int)t;
                    const uint32_t J = lex_rank(jdx, m, b);
                    hop[s->d.hop_off[m] + hh * S + r] = J | (uint32_t)(8 * (o * b + t)) << 16;
                    ++hh;
                }
            }
            for (; hh < hop_rows(m, b); ++hh) /* padding: a hop onto itself with K = 0 */
                hop[s->d.hop_off[m] + hh * S + r] = (uint32_t)r | (uint32_t)(8 * b * b) << 16;
            if (m < b)
                for (unsigned t = 0; t < b; ++t) {
                    uint16_t v = 0xffff;
                    if (!((mask >> t) & 1)) {
                        unsigned q = 0, done = 0;
                        for (unsigned e = 0; e < m; ++e) {
                            if (!done && idx[e] > (int)t) {
                                jdx[q++] = (int)t;
                                done = 1;
                            }
                            jdx[q++] = idx[e];
                        }
                        if (!done) jdx[q++] = (int)t;
                        v = (uint16_t)lex_rank(jdx, m + 1, b);
                    }
                    ins[s->d.ins_off[m] + (size_t)t * S + r] = v;
                }
            ++r;
        } while (m > 0 && next_comb(idx, m, b));
    }
    hipError_t e = hipMalloc(&s->dtab, o_i + bytes_i + 1);
    if (e == hipSuccess) e = hipMemcpy(s->dtab, h, o_i + bytes_i, hipMemcpyHostToDevice);
    free(h);
    if (e != hipSuccess) {
        s->dtab = NULL;
        return doci_fail(DOCI_E_HIP, "B tables on device: %s", hipGetErrorString(e));
    }
    s->d.bmask = (const uint16_t *)s->dtab;
    s->d.hop = (const uint32_t *)((char *)s->dtab + o_h);
    s->d.ins = (const uint16_t *)((char *)s->dtab + o_i);
    return DOCI_OK;
}

static int build_shape(const dk_tables *dt, unsigned lds_budget, unsigned sgmax, unsigned main_emax,
                       star_shape **out)
{
    const size_t STAR_LDS_BUDGET = lds_budget ? lds_budget : STAR_LDS_DEFAULT;
    const unsigned STAR_SGMAX = sgmax ? sgmax : STAR_SGMAX_DEFAULT;
    const unsigned MAIN_EMAX = main_emax ? main_emax : MAIN_EMAX_DEFAULT;
    const unsigned L = dt->L, N = dt->N;
    const uint64_t n = (uint64_t)dt->n;
    unsigned b = L < STAR_BMAX ? L : STAR_BMAX, a = 0, ymax = 0, selems = 0;
    for (; b >= 1; --b) {
        a = L - b;
        uint64_t mx = 0; /* biggest centre slab */
        ymax = a; /* member tables: a entries per star (the column pass indexes them by orbital) */
        for (unsigned j = 0; j < N; ++j) {
            if (!valid_sector(j + 1, N, a, b)) continue; /* members' sector */
            const uint64_t st = valid_sector(j, N, a, b) ? dh_binom(N - j, b) : 0;
            if (st > mx) mx = st;
        }
        uint64_t maxS = 0; /* biggest slab: the main pass stages whole slabs */
        for (unsigned k = 0; k <= N; ++k)
            if (valid_sector(k, N, a, b) && dh_binom(N - k, b) > maxS) maxS = dh_binom(N - k, b);
        if (maxS < MAIN_EMAX) maxS = MAIN_EMAX;
        if (ymax > 64) continue;
        if (star_lds_bytes(L, N, STAR_SGMAX, ymax, (unsigned)mx) <= STAR_LDS_BUDGET &&
            8 * (size_t)b * b + 4 * (size_t)(b + 1) * (N + 2) + 12 * (size_t)maxS +
                    8 * (size_t)MAIN_GSLABS + 64 <= MAIN_LDS_MAX) {
            /* the rest of the budget stages as many centres as fit */
            selems = (unsigned)((STAR_LDS_BUDGET - star_lds_bytes(L, N, STAR_SGMAX, ymax, 0)) / 16);
            if (selems < mx) selems = (unsigned)mx;
            break;
        }
    }
    if (b < 1) return doci_fail(DOCI_E_ARG, "no star split fits the LDS budget");
    uint64_t emax = MAIN_EMAX;
    for (unsigned k = 0; k <= N; ++k)
        if (valid_sector(k, N, a, b) && dh_binom(N - k, b) > emax) emax = dh_binom(N - k, b);

    /* sector offsets in Z and slab counts */
    uint32_t *zoff = calloc(a + 2, sizeof(uint32_t));
    uint64_t nslab = 0, nstar = 0, zacc = 0;
    for (unsigned k = 0; k <= a; ++k) {
        zoff[k] = (uint32_t)zacc;
        if (valid_sector(k, N, a, b)) {
            nslab += dh_binom(k, a);
            zacc += dh_binom(k, a) * nslot(k, N - k, a) * dh_binom(N - k, b);
        }
    }
    zoff[a + 1] = (uint32_t)zacc;
    for (unsigned j = 0; j < N; ++j)
        if (valid_sector(j + 1, N, a, b)) nstar += dh_binom(j, a);
    int rc = DOCI_OK;
    if (zacc >= (1ull << 32)) {
        free(zoff);
        return doci_fail(DOCI_E_DIM, "slot buffer exceeds 2^32 entries");
    }
    uint32_t *sstart = malloc(sizeof(uint32_t) * nslab), *szb = malloc(sizeof(uint32_t) * nslab);
    sgrp_t *grp = malloc(sizeof(sgrp_t) * (nslab + 1));
    uint64_t *samask = malloc(sizeof(uint64_t) * (nstar + 1));
    /* member table: for star rho and orbital y < a not in rho, the first
     * rank and the Z position of its slot of member slab rho + y */
    uint32_t *mtab = malloc(sizeof(uint32_t) * 2 * (nstar * a + 1));
    uint64_t secbase[DK_MAX_L + 2] = {0}; /* first slab of each k-sector */
    uint64_t cb[65][65];                  /* cb[n][k] = C(n, k), n <= 64 */
    for (unsigned x = 0; x <= 64; ++x)
        for (unsigned q = 0; q <= 64; ++q) cb[x][q] = q > x ? 0 : (q == 0 || q == x) ? 1 : cb[x - 1][q - 1] + cb[x - 1][q];
    uint32_t *sgw = malloc(sizeof(uint32_t) * 2 * (nstar + 1)), *gw = NULL;
    star_shape *s = NULL;
    int idx[64];
    if (!sstart || !szb || !grp || !samask || !sgw || !mtab) {
        rc = doci_fail(DOCI_E_NOMEM, "star tables");
        goto out;
    }
    /* slabs and main-pass groups: as many whole slabs of one sector as fit
     * emax determinants, at most MAIN_GSLABS, rounded down to a multiple of 4
     * (or to 2): the main pass walks a slab's hops once for a set of 4 (2)
     * slabs of the group */
    uint64_t is = 0, ng = 0, gslabs = 1, owns = 0;
    for (unsigned k = 0; k <= a; ++k) {
        if (!valid_sector(k, N, a, b)) continue;
        const unsigned m = N - k;
        const uint64_t S = dh_binom(m, b), ns = nslot(k, m, a);
        uint64_t pg = emax / S;
        if (pg > MAIN_GSLABS) pg = MAIN_GSLABS;
        pg = pg >= 4 ? pg & ~(uint64_t)3 : pg >= 2 ? 2 : 1;
        if (pg > gslabs) gslabs = pg;
        secbase[k] = is;
        for (unsigned e = 0; e < k; ++e) idx[e] = (int)e;
        uint32_t gfirst = (uint32_t)is, gcount = 0;
        uint64_t li = 0;
        do {
            uint64_t sa = 0;
            for (unsigned e = 0; e < k; ++e) sa += dh_binom(N - e, L - 1 - (unsigned)idx[e]);
            if (gcount == pg) {
                const uint64_t ow = own_doubles(gcount, S);
                if (ow > owns) owns = ow;
                grp[ng].first = gfirst;
                grp[ng].start = sstart[gfirst];
                grp[ng].word = gcount | (m << 16);
                ++ng;
                gfirst = (uint32_t)is;
                gcount = 0;
            }
            sstart[is] = (uint32_t)(n - S - sa);
            szb[is] = (uint32_t)(zoff[k] + li * ns * S);
            ++is;
            ++li;
            ++gcount;
        } while (next_comb(idx, k, a));
        {
            const uint64_t ow = own_doubles(gcount, S);
            if (ow > owns) owns = ow;
        }
        grp[ng].first = gfirst;
        grp[ng].start = sstart[gfirst];
        grp[ng].word = gcount | (m << 16);
        ++ng;
    }
    qsort(grp, ng, sizeof(sgrp_t), cmp_sgrp);
    gw = malloc(sizeof(uint32_t) * 2 * ng);
    if (!gw) {
        rc = doci_fail(DOCI_E_NOMEM, "main groups");
        goto out;
    }
    for (uint64_t g = 0; g < ng; ++g) {
        gw[2 * g] = grp[g].first;
        gw[2 * g + 1] = grp[g].word;
    }
    /* stars and star groups */
    uint64_t ist = 0, nsg = 0;
    uint32_t jgroup[DK_MAX_L + 1];
    for (unsigned j = 0; j < N; ++j) {
        jgroup[j] = (uint32_t)nsg;
        if (!valid_sector(j + 1, N, a, b)) continue;
        const uint64_t stride = valid_sector(j, N, a, b) ? dh_binom(N - j, b) : 0;
        for (unsigned e = 0; e < j; ++e) idx[e] = (int)e;
        uint32_t gfirst = (uint32_t)ist, gcount = 0;
        do {
            uint64_t msk = 0;
            for (unsigned e = 0; e < j; ++e) msk |= 1ull << idx[e];
            if (gcount == STAR_SGMAX || (gcount + 1) * stride > selems) {
                sgw[2 * nsg] = gfirst;
                sgw[2 * nsg + 1] = gcount | (j << 16);
                ++nsg;
                gfirst = (uint32_t)ist;
                gcount = 0;
            }
            /* members rho + y: slab index secbase[j + 1] + lexicographic rank
             * of rho + y, C(a,k) - 1 - sum_e C(a-1-x_e, k-e) over its sorted
             * orbitals x_e (k = j + 1): rho's orbitals below y keep their
             * place e, those above move to e + 1 (prefix and suffix sums) */
            const unsigned k = j + 1;
            const uint64_t Sm = cb[b][N - k];
            uint64_t pre[65], suf[65];
            pre[0] = 0;
            for (unsigned e = 0; e < j; ++e) pre[e + 1] = pre[e] + cb[a - 1 - idx[e]][k - e];
            suf[j] = 0;
            for (unsigned e = j; e-- > 0;) suf[e] = suf[e + 1] + cb[a - 1 - idx[e]][k - 1 - e];
            for (unsigned y = 0, pos = 0; y < a; ++y) {
                uint32_t *mt = mtab + 2 * (ist * a + y);
                if ((msk >> y) & 1) {
                    mt[0] = mt[1] = 0; /* not a member */
                    ++pos;
                    continue;
                }
                const uint64_t li = cb[a][k] - 1 - (pre[pos] + cb[a - 1 - y][k - pos] + suf[pos]);
                const uint64_t si = secbase[k] + li;
                mt[0] = sstart[si];
                mt[1] = (uint32_t)(szb[si] + pos * Sm);
            }
            samask[ist++] = msk;
            ++gcount;
        } while (next_comb(idx, j, a));
        sgw[2 * nsg] = gfirst;
        sgw[2 * nsg + 1] = gcount | (j << 16);
        ++nsg;
    }
    s = calloc(1, sizeof *s);
    if (!s) {
        rc = doci_fail(DOCI_E_NOMEM, "star descriptor");
        goto out;
    }
    {
        const size_t b_am = sizeof(uint64_t) * nstar, b_ss = sizeof(uint32_t) * nslab,
                     b_zb = sizeof(uint32_t) * nslab, b_g = sizeof(uint32_t) * 2 * ng,
                     b_sg = sizeof(uint32_t) * 2 * nsg, b_zo = sizeof(uint32_t) * (a + 2);
        const size_t o_mt = (b_am + b_ss + b_zb + b_g + b_sg + b_zo + 7) & ~(size_t)7,
                     b_mt = sizeof(uint32_t) * 2 * nstar * a;
        hipError_t e = hipMalloc(&s->dmem, o_mt + b_mt);
        char *d = s->dmem;
        if (e == hipSuccess) e = hipMemcpy(d, samask, b_am, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d + b_am, sstart, b_ss, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d + b_am + b_ss, szb, b_zb, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d + b_am + b_ss + b_zb, gw, b_g, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(d + b_am + b_ss + b_zb + b_g, sgw, b_sg, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(d + b_am + b_ss + b_zb + b_g + b_sg, zoff, b_zo, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d + o_mt, mtab, b_mt, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            rc = doci_fail(DOCI_E_HIP, "star tables on device: %s", hipGetErrorString(e));
            goto out;
        }
        s->d.a = a;
        s->d.b = b;
        s->d.ngroup = (uint32_t)ng;
        s->d.emax = (uint32_t)emax;
        s->d.gslabs = (uint32_t)gslabs;
        s->d.owns = (uint32_t)owns;
        s->d.nsgroup = (uint32_t)nsg;
        s->d.sgmax = STAR_SGMAX;
        jgroup[N] = (uint32_t)nsg;
        memcpy(s->d.jgroup, jgroup, sizeof(uint32_t) * (N + 1));
        s->d.selems = selems;
        s->d.ymax = ymax;
        s->d.pass = STAR_PASS_DEFAULT;
        s->d.zsec_max = 0;
        for (unsigned k = 0; k <= a; ++k)
            if (zoff[k + 1] - zoff[k] > s->d.zsec_max) s->d.zsec_max = zoff[k + 1] - zoff[k];
        s->d.star_amask = (const uint64_t *)d;
        s->d.slab_start = (const uint32_t *)(d + b_am);
        s->d.slab_zbase = (const uint32_t *)(d + b_am + b_ss);
        s->d.group = (const uint32_t *)(d + b_am + b_ss + b_zb);
        s->d.sgroup = (const uint32_t *)(d + b_am + b_ss + b_zb + b_g);
        s->d.zoff = (const uint32_t *)(d + b_am + b_ss + b_zb + b_g + b_sg);
        s->d.mtab = (const uint32_t *)(d + o_mt);
        s->zelems = zacc + dk_star_kp_doubles(a, b); /* slots, then packed K */
        s->lim[0] = lds_budget;
        s->lim[1] = sgmax;
        s->lim[2] = main_emax;
    }
    rc = build_b_tables(s, b);
    if (rc != DOCI_OK) goto out;
    if (dk_sigma_star_lds(dt, &s->d, 0, 0) > STAR_LDS_BUDGET ||
        dk_sigma_star_lds(dt, &s->d, 0, 1) > MAIN_LDS_MAX) {
        rc = doci_fail(DOCI_E_ARG, "star kernels exceed the LDS budget");
        goto out;
    }
    s->L = L;
    s->N = N;
    *out = s;
    s = NULL;
out:
    if (s) {
        if (s->dmem) hipFree(s->dmem);
        if (s->dtab) hipFree(s->dtab);
        free(s);
    }
    free(zoff);
    free(sstart);
    free(szb);
    free(grp);
    free(samask);
    free(mtab);
    free(sgw);
    free(gw);
    return rc;
}

/* The shared shape of (L, N, limits), built on first request. */
static int shape_of(const dk_tables *dt, unsigned lds_budget, unsigned sgmax, unsigned main_emax,
                    const star_shape **out)
{
    const star_shape *sh = NULL;
    int rc = DOCI_OK;
    pthread_mutex_lock(&g_shapes_lock);
    for (const star_shape *c = g_shapes; c; c = c->next)
        if (c->L == dt->L && c->N == dt->N && c->lim[0] == lds_budget && c->lim[1] == sgmax &&
            c->lim[2] == main_emax) {
            sh = c;
            break;
        }
    if (!sh) {
        star_shape *ns = NULL;
        rc = build_shape(dt, lds_budget, sgmax, main_emax, &ns);
        if (rc == DOCI_OK) {
            ns->next = g_shapes;
            g_shapes = ns;
            sh = ns;
        }
    }
    pthread_mutex_unlock(&g_shapes_lock);
    *out = sh;
    return rc;
}

int dh_star_shape(const dk_tables *dt, const dk_stars **st, size_t *zelems)
{
    const star_shape *sh = NULL;
    const int rc = shape_of(dt, 0, 0, 0, &sh);
    if (rc != DOCI_OK) return rc;
    *st = &sh->d;
    *zelems = sh->zelems ? sh->zelems : 1;
    return DOCI_OK;
}

int dh_stars_build(doci_op *op, unsigned lds_budget, unsigned sgmax, unsigned main_emax)
{
    const star_shape *sh = NULL;
    const int rc = shape_of(&op->dt, lds_budget, sgmax, main_emax, &sh);
    if (rc != DOCI_OK) return rc;
    dh_stars *s = calloc(1, sizeof *s);
    if (!s) return doci_fail(DOCI_E_NOMEM, "star descriptor");
    s->sh = sh;
    s->pass = (int)sh->d.pass;
    hipError_t e = dh_dmalloc((void **)&s->d_bad, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(s->d_bad, 0, sizeof(uint32_t));
    if (e != hipSuccess) {
        dh_dfree(s->d_bad);
        free(s);
        return doci_fail(DOCI_E_HIP, "star counter: %s", hipGetErrorString(e));
    }
    op->stars = s;
    return DOCI_OK;
}

int dh_stars_sigma(doci_op *op, const void *in, void *out, int prec, double shift, int checked)
{
    dh_stars *s = op->stars;
    if (!s->Z) { /* the slot buffer is allocated on first use and kept */
        hipError_t e = dh_dmalloc(&s->Z, sizeof(double) * (s->sh->zelems ? s->sh->zelems : 1));
        if (e != hipSuccess) {
            s->Z = NULL;
            return doci_fail(DOCI_E_NOMEM, "sigma slot buffer (%zu entries): %s", s->sh->zelems,
                             hipGetErrorString(e));
        }
    }
    dk_stars d = s->sh->d;
    d.pass = (uint32_t)s->pass;
    DKCHK(dk_sigma_star(&op->dt, &d, op->d_D, shift, in, out, s->Z, prec,
                        checked ? s->d_bad : NULL, op->stream));
    return DOCI_OK;
}

int dh_stars_sigma_static(doci_op *op, const void *in, void *out, int prec, double shift)
{
    dh_stars *s = op->stars;
    if (!s)
        return doci_fail(DOCI_E_ARG, "star tables not built yet (doci_op_set_kernel(op, 0) or a "
                                     "first doci_op_sigma_dev builds them; this entry point "
                                     "allocates nothing)");
    const size_t need = s->sh->zelems ? s->sh->zelems : 1, cap = dk_static_slot_doubles();
    if (need > cap)
        return doci_fail(DOCI_E_ARG, "sigma slot buffer of %zu FP64 entries exceeds the static "
                                     "buffer of %zu (L=%u, N=%u)", need, cap, op->L, op->N);
    void *Z = NULL;
    DKCHK(dk_static_slot_buffer(&Z));
    dk_stars d = s->sh->d;
    d.pass = (uint32_t)s->pass;
    DKCHK(dk_sigma_star(&op->dt, &d, op->d_D, shift, in, out, Z, prec, NULL, op->stream));
    return DOCI_OK;
}

int dh_stars_set_pass(doci_op *op, int pass)
{
    if (!op->stars) {
        const int rc = dh_stars_build(op, 0, 0, 0);
        if (rc != DOCI_OK) return rc;
    }
    if (pass < 0 || pass > 4) return doci_fail(DOCI_E_ARG, "star_pass must be 0 to 4");
    ((dh_stars *)op->stars)->pass = pass;
    return DOCI_OK;
}

int64_t dh_stars_violations(doci_op *op)
{
    dh_stars *s = op->stars;
    uint32_t v = 0;
    if (!s) return doci_fail(DOCI_E_ARG, "no star tables");
    HIPCHK(hipStreamSynchronize(op->stream));
    HIPCHK(hipMemcpy(&v, s->d_bad, sizeof v, hipMemcpyDeviceToHost));
    return (int64_t)v;
}

void dh_stars_limits(const doci_op *op, unsigned *lds_budget, unsigned *sgmax, unsigned *main_emax)
{
    const dh_stars *s = op->stars;
    *lds_budget = s ? s->sh->lim[0] : 0;
    *sgmax = s ? s->sh->lim[1] : 0;
    *main_emax = s ? s->sh->lim[2] : 0;
}

void dh_stars_info(const doci_op *op, unsigned *a, unsigned *b, unsigned *ngroup,
                   unsigned *nsgroup, size_t *zelems)
{
    const dh_stars *s = op->stars;
    *a = s ? s->sh->d.a : 0;
    *b = s ? s->sh->d.b : 0;
    *ngroup = s ? s->sh->d.ngroup : 0;
    *nsgroup = s ? s->sh->d.nsgroup : 0;
    *zelems = s ? s->sh->zelems : 0;
}
