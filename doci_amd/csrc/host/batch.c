/* batch.c -- many independent DOCI Davidson solves on one GPU.
 *
 * The reference solves one problem per process (src/executable.c:41-90).
 * Here a batch holds any number of instances of any (L, N) that run the
 * Davidson iteration of src/davidson.c:72-99 in lock step: one iteration
 * is a fixed sequence of launches over all running instances (kernels in
 * batch.hip), recorded once in a HIP graph and replayed while the set of
 * running instances and the basis-size bucket stay the same.  Instances can
 * be queued at any time; each step first admits queued instances in FIFO
 * order while their device footprint fits the memory budget, runs one
 * iteration, then retires finished ones (continuous batching).
 *
 * Mixed-precision instances run an FP32-storage phase (operator shifted by
 * min D, stopping at switch_tol) and then restart in FP64 from their Ritz
 * vector, exactly like the single-instance mixed solve; the two phases are
 * separate launch groups, so FP32 and FP64 instances run side by side.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

/* a correction whose part orthogonal to an instance's basis is below this
 * fraction of its norm squared is projected out explicitly (an ORTHO pass);
 * any other joins the basis as it is */
#define DK_BATCH_DEP_TOL 0.05

typedef struct {
    int state;              /* 0 queued, 1 running, 2 finished */
    int want_mixed;
    int prec;               /* running phase: 1 FP32 storage, 0 FP64 */
    int slot;               /* device instance slot while running, else -1 */
    int fresh;              /* admitted after the status copy in flight was queued */
    unsigned L, N;
    uint32_t n;
    double core;
    char *tab;              /* transformed tables (device layout) until admission */
    void *dmain;            /* tables, D and vector slots */
    double *dx;             /* FP64 Ritz vector (kept until release) */
    size_t bytes;           /* device bytes held while running (accounted in mem_used) */
    size_t xbytes;          /* bytes of dx (accounted until it is freed) */
    const dk_stars *st;     /* large instances: shared star tables of the shape */
    void *Z;                /* and their slot buffer (NULL for lane-sigma instances) */
    size_t zbytes;
    dk_binst bi;
    doci_batch_result res;
} brec;

typedef struct {
    int n;                  /* running instances of this precision */
    uint32_t *h_list, *d_list;
    size_t list_cap;
    uint32_t ntile;
    uint32_t *h_tiles, *d_tiles;
    size_t tiles_cap;
    double *d_part;
    size_t part_cap;
    hipGraph_t graph;
    hipGraphExec_t gexec;
    int gM;                 /* bucket of the captured graph, 0 = none */
    int fresh;              /* lists rebuilt since the last iteration */
    uint32_t maxL, maxN;
} bgroup;

struct doci_batch {
    doci_batch_opts o;
    hipStream_t stream;
    int cap;
    dk_binst *d_inst;
    dk_binst *h_inst; /* pinned mirror of d_inst, uploaded whole by rebuild() */
    dk_bstate *d_state;
    dk_bstat *h_stat;       /* pinned */
    int *slot_rec;          /* device slot -> record, -1 free */
    brec *rec;
    int nrec, rec_cap;
    int qhead;              /* first record that may still be queued */
    bgroup grp[2];          /* [0] FP64, [1] FP32 storage */
    int dirty;
    size_t mem_used;
    double t0;
    int tile_rows, tuned;
    int iterations, graph_launches;
    char *h_stage;          /* pinned staging of instance tables (reused every step) */
    size_t stage_cap, stage_used;
    char *d_stage;          /* its device copy (same offsets) */
    size_t stage_sent;      /* bytes of h_stage already copied to d_stage */
    dk_badmit *h_adm;       /* pinned: admissions not yet launched, then their tiles */
    char *d_adm;            /* device copy of h_adm */
    size_t adm_cap;         /* bytes of h_adm / d_adm */
    double *d_apart;        /* 2 doubles per admission tile */
    size_t apart_cap;       /* tiles */
    int npend;              /* admissions in h_adm */
    uint32_t pend_tiles;
    uint32_t pend_maxL, pend_maxN;
    dk_brestart *h_rst;     /* pinned: FP64 restarts of this drain, then their tiles */
    char *d_rst;            /* device copy of h_rst */
    size_t rst_cap;         /* bytes of h_rst / d_rst */
    int nrst;               /* restarts in h_rst */
    uint32_t rst_tiles;
    int inflight;           /* an iteration (and its status copy) is queued */
    void **chunk;           /* device arena chunks */
    int nchunk, chunk_cap;
    struct frange { char *p; size_t bytes; } *frees; /* free ranges of the arena */
    int nfree, free_cap;
    size_t arena_bytes;
};

#define ALIGN256(x) (((size_t)(x) + 255) & ~(size_t)255)
#define DH_XCDS 8 /* XCDs (each with its own L2) of an MI355X */

/* DOCI_BATCH_TRACE=1 prints admission / retirement / step timings to stderr. */
static int trace_on(void)
{
    static int v = -1;
    if (v < 0) {
        const char *e = getenv("DOCI_BATCH_TRACE");
        v = e && *e && *e != '0';
    }
    return v;
}

static int arena_grow(doci_batch *b, size_t bytes);

void doci_batch_defaults(doci_batch_opts *o)
{
    memset(o, 0, sizeof *o);
    o->tol = 1e-8;
    o->max_its = 200;
    o->keep = 2;
    o->max_vec = 30;
    o->switch_tol = 1e-5;
    o->use_graphs = 1;
    o->tile_rows = 0;
    o->mem_budget = 0;
    o->max_instances = 1024;
    o->keep_vectors = 1;
}

/* Device memory arena.  Instance memory is carved from a few large
 * hipMalloc'd chunks (grown on demand, kept until the batch is freed) by a
 * first-fit allocator with coalescing, so that admitting and retiring
 * instances while the GPU runs never calls the driver's allocator (whose
 * frees synchronise the device).  Ranges are 256-byte aligned. */
#define ARENA_CHUNK ((size_t)8 << 30)

static int arena_grow(doci_batch *b, size_t bytes)
{
    size_t sz = bytes > ARENA_CHUNK ? bytes : ARENA_CHUNK;
    if (b->arena_bytes + sz > b->o.mem_budget) sz = bytes; /* stay within the budget if possible */
    if (b->arena_bytes + sz > b->o.mem_budget) return -1;
    void *p = NULL;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    if (b->nchunk == b->chunk_cap || b->nfree == b->free_cap) {
        const int nc = b->chunk_cap ? 2 * b->chunk_cap : 16, nf = b->free_cap ? 2 * b->free_cap : 256;
        void **c = realloc(b->chunk, sizeof(void *) * nc);
        if (c) b->chunk = c, b->chunk_cap = nc;
        struct frange *f = realloc(b->frees, sizeof *f * nf);
        if (f) b->frees = f, b->free_cap = nf;
        if (!c || !f || b->nchunk == b->chunk_cap || b->nfree == b->free_cap) {
            hipFree(p);
            return -1;
        }
    }
    b->chunk[b->nchunk] = p;
    b->frees[b->nfree].p = p;
    b->frees[b->nfree].bytes = sz;
    ++b->nfree;
    ++b->nchunk;
    b->arena_bytes += sz;
    return 0;
}

static void *balloc(doci_batch *b, size_t bytes)
{
    bytes = ALIGN256(bytes);
    for (int pass = 0; pass < 2; ++pass) {
        for (int i = 0; i < b->nfree; ++i)
            if (b->frees[i].bytes >= bytes) {
                char *p = b->frees[i].p;
                b->frees[i].p = p + bytes;
                b->frees[i].bytes -= bytes;
                if (!b->frees[i].bytes) b->frees[i] = b->frees[--b->nfree];
                return p;
            }
        if (pass == 0 && arena_grow(b, bytes) != 0) return NULL;
    }
    return NULL;
}

static void bfree(doci_batch *b, void *vp, size_t bytes)
{
    if (!vp) return;
    char *p = vp;
    bytes = ALIGN256(bytes);
    /* merge with adjacent free ranges (chunks are never released while the
     * batch lives, so a range may also span two neighbouring chunks) */
    for (int i = 0; i < b->nfree;) {
        struct frange *f = &b->frees[i];
        if (f->p + f->bytes == p) {
            p = f->p;
            bytes += f->bytes;
            *f = b->frees[--b->nfree];
            continue;
        }
        if (p + bytes == f->p) {
            bytes += f->bytes;
            *f = b->frees[--b->nfree];
            continue;
        }
        ++i;
    }
    if (b->nfree == b->free_cap) {
        const int nf = b->free_cap ? 2 * b->free_cap : 256;
        struct frange *f = realloc(b->frees, sizeof *f * nf);
        if (!f) return; /* the range is lost to the arena until the batch is freed */
        b->frees = f;
        b->free_cap = nf;
    }
    b->frees[b->nfree].p = p;
    b->frees[b->nfree].bytes = bytes;
    ++b->nfree;
}

static void arena_free(doci_batch *b)
{
    for (int c = 0; c < b->nchunk; ++c) hipFree(b->chunk[c]);
    free(b->chunk);
    free(b->frees);
    b->chunk = NULL;
    b->frees = NULL;
    b->nchunk = b->nfree = b->chunk_cap = b->free_cap = 0;
    b->arena_bytes = 0;
}

static size_t table_bytes(unsigned L, unsigned N)
{
    const size_t l = L;
    return ALIGN256(sizeof(double) * (3 * l * l + l) + sizeof(uint32_t) * (l + 1) * (N + 2));
}

static size_t slot_stride(uint32_t n) { return ALIGN256(sizeof(double) * (size_t)n); }

static size_t main_bytes(const doci_batch_opts *o, unsigned L, unsigned N, uint32_t n)
{
    const size_t P = (size_t)(o->max_vec + o->keep);
    return table_bytes(L, N) + slot_stride(n) /* D */ + 2 * P * slot_stride(n);
}

/* Large instances get the star-kernel sigma (as the single-instance
 * operator does from this size on); their slot buffer is below (N + 1) n
 * entries. */
static int uses_star(uint32_t n) { return n >= DOCI_STAR_MIN_N; }

/* An instance gets at most this many tiles: large instances use long tiles
 * (few partial sums to reduce, long coalesced streams per workgroup). */
#define MAX_TILES_PER_INSTANCE 2048u

/* Rows per tile of an instance: the batch's tile size, or more so that the
 * instance has at most MAX_TILES_PER_INSTANCE tiles.  Always a multiple of
 * 512: the streaming kernels read whole 16-byte vectors from every tile's
 * start. */
static uint32_t rows_for(uint32_t n, uint32_t base)
{
    uint32_t cap = (n + MAX_TILES_PER_INSTANCE - 1) / MAX_TILES_PER_INSTANCE;
    cap = (cap + 511) / 512 * 512;
    return base > cap ? base : cap;
}

size_t doci_batch_footprint(const doci_batch_opts *o, unsigned L, unsigned N)
{
    const uint64_t n = dh_binom(N, L);
    size_t b = main_bytes(o, L, N, (uint32_t)n) + slot_stride((uint32_t)n);
    if (uses_star((uint32_t)n)) b += sizeof(double) * (size_t)(N + 1) * n;
    return b;
}

double doci_batch_clock(const doci_batch *b) { return dh_now() - b->t0; }

int doci_batch_create(const doci_batch_opts *o_in, doci_batch **out)
{
    *out = NULL;
    doci_batch_opts o;
    if (o_in)
        o = *o_in;
    else
        doci_batch_defaults(&o);
    if (o.max_vec < 2 || o.max_vec > DK_BMAX_VEC - 2 || o.keep < 1 || o.keep > DK_BKEEP_MAX ||
        o.keep >= o.max_vec || o.max_instances < 1 || !(o.tol > 0) || o.max_its < 0)
        return doci_fail(DOCI_E_ARG, "batch options out of range (max_vec 2..%d, keep 1..%d)",
                         DK_BMAX_VEC - 2, DK_BKEEP_MAX);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return doci_fail(DOCI_E_HIP, "no HIP device available");
    /* the batch budgets and reserves free device memory: hand the buffers
     * cached for single solves back to the driver first */
    dh_dcache_trim();
    doci_batch *b = calloc(1, sizeof *b);
    if (!b) return doci_fail(DOCI_E_NOMEM, "batch");
    if (o.tile_rows > 0) o.tile_rows = (o.tile_rows + 511) / 512 * 512; /* see rows_for */
    b->o = o;
    b->cap = o.max_instances;
    b->t0 = dh_now();
    b->tile_rows = o.tile_rows;
    b->tuned = o.tile_rows > 0;
    int rc = DOCI_OK;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_inst, sizeof(dk_binst) * b->cap);
    if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_inst, sizeof(dk_binst) * b->cap, 0);
    if (e == hipSuccess) memset(b->h_inst, 0, sizeof(dk_binst) * b->cap);
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_state, sizeof(dk_bstate) * b->cap);
    if (e == hipSuccess) e = hipMemset(b->d_state, 0, sizeof(dk_bstate) * b->cap);
    if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_stat, sizeof(dk_bstat) * b->cap, 0);
    if (e == hipSuccess) memset(b->h_stat, 0, sizeof(dk_bstat) * b->cap);
    if (e != hipSuccess) rc = doci_fail(DOCI_E_HIP, "batch setup: %s", hipGetErrorString(e));
    b->slot_rec = malloc(sizeof(int) * b->cap);
    if (rc == DOCI_OK && !b->slot_rec) rc = doci_fail(DOCI_E_NOMEM, "batch");
    if (rc != DOCI_OK) {
        doci_batch_free(b);
        return rc;
    }
    for (int i = 0; i < b->cap; ++i) b->slot_rec[i] = -1;

    if (!o.mem_budget) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) b->o.mem_budget = (size_t)(0.9 * (double)fr);
    }
    if (o.mem_reserve) {
        /* in 32 GiB chunks (halved when the driver refuses one) */
        const size_t want = o.mem_reserve < b->o.mem_budget ? o.mem_reserve : b->o.mem_budget;
        size_t chunk = (size_t)32 << 30;
        while (b->arena_bytes < want && chunk >= ARENA_CHUNK) {
            const size_t left = want - b->arena_bytes;
            if (arena_grow(b, ALIGN256(left < chunk ? left : chunk)) != 0) chunk /= 2;
        }
    }
    *out = b;
    return DOCI_OK;
}

static void group_free(bgroup *g)
{
    if (g->gexec) hipGraphExecDestroy(g->gexec);
    if (g->graph) hipGraphDestroy(g->graph);
    if (g->d_list) hipFree(g->d_list);
    if (g->d_tiles) hipFree(g->d_tiles);
    if (g->d_part) hipFree(g->d_part);
    if (g->h_list) hipHostFree(g->h_list);
    if (g->h_tiles) hipHostFree(g->h_tiles);
    memset(g, 0, sizeof *g);
}

/* called with the stream drained */
static void free_x(doci_batch *b, brec *r)
{
    bfree(b, r->dx, r->xbytes);
    r->dx = NULL;
    b->mem_used -= r->xbytes;
    r->xbytes = 0;
}

/* called with the stream drained */
static void rec_free_main(doci_batch *b, brec *r)
{
    bfree(b, r->dmain, r->bytes - r->zbytes);
    r->dmain = NULL;
    bfree(b, r->Z, r->zbytes);
    r->Z = NULL;
    r->zbytes = 0;
    b->mem_used -= r->bytes;
    r->bytes = 0;
}

void doci_batch_free(doci_batch *b)
{
    if (!b) return;
    if (b->stream) hipStreamSynchronize(b->stream);
    for (int i = 0; i < b->nrec; ++i) {
        brec *r = &b->rec[i];
        rec_free_main(b, r);
        free_x(b, r);
        free(r->tab);
    }
    free(b->rec);
    arena_free(b);
    for (int g = 0; g < 2; ++g) group_free(&b->grp[g]);
    if (b->d_inst) hipFree(b->d_inst);
    if (b->h_inst) hipHostFree(b->h_inst);
    if (b->d_state) hipFree(b->d_state);
    if (b->h_stat) hipHostFree(b->h_stat);
    if (b->h_stage) hipHostFree(b->h_stage);
    if (b->d_stage) hipFree(b->d_stage);
    if (b->h_adm) hipHostFree(b->h_adm);
    if (b->d_adm) hipFree(b->d_adm);
    if (b->d_apart) hipFree(b->d_apart);
    if (b->h_rst) hipHostFree(b->h_rst);
    if (b->d_rst) hipFree(b->d_rst);
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->stream) hipStreamDestroy(b->stream);
    free(b->slot_rec);
    free(b);
}

static int check_instance(const doci_batch *b, unsigned L, unsigned N)
{
    if (L == 0 || L > DK_MAX_L) return doci_fail(DOCI_E_ORBITALS, "Too many orbitals (L=%u).", L);
    if (N == 0 || N > L) return doci_fail(DOCI_E_DIM, "invalid number of pairs N=%u", N);
    const uint64_t n = dh_binom(N, L);
    if (n >= (1ull << 31))
        return doci_fail(DOCI_E_DIM, "C(%u,%u) = %llu exceeds 2^31-1", L, N, (unsigned long long)n);
    if (doci_batch_footprint(&b->o, L, N) > b->o.mem_budget)
        return doci_fail(DOCI_E_NOMEM, "instance L=%u N=%u needs more than the memory budget", L, N);
    return DOCI_OK;
}

static int reserve_records(doci_batch *b, int count)
{
    if (b->nrec + count <= b->rec_cap) return DOCI_OK;
    int nc = b->rec_cap ? 2 * b->rec_cap : 64;
    while (nc < b->nrec + count) nc *= 2;
    brec *nr = realloc(b->rec, sizeof(brec) * nc);
    if (!nr) return doci_fail(DOCI_E_NOMEM, "batch records");
    b->rec = nr;
    b->rec_cap = nc;
    return DOCI_OK;
}

/* The request intake of one record: only the O(L^2) tables of the raw
 * integrals are kept (thread-safe: touches nothing but r). */
static int rec_tables(brec *r, unsigned L, unsigned N, const double *one, const double *two)
{
    const size_t l = L, l2 = l * l;
    r->tab = malloc(table_bytes(L, N));
    if (!r->tab) return DOCI_E_NOMEM;
    double *K = (double *)r->tab, *J = K + l2, *X = J + l2, *h = X + l2;
    if (dh_build_tables(L, N, one, two, K, J, X, h, (uint32_t *)(h + l)) != DOCI_OK) {
        free(r->tab);
        r->tab = NULL;
        return DOCI_E_NOMEM;
    }
    return DOCI_OK;
}

static void rec_fill(doci_batch *b, brec *r, unsigned L, unsigned N, double core, int precision)
{
    r->L = L;
    r->N = N;
    r->n = (uint32_t)dh_binom(N, L);
    r->core = core;
    r->want_mixed = precision != 0;
    r->slot = -1;
    r->res.t_submit = doci_batch_clock(b);
}

int doci_batch_add(doci_batch *b, unsigned L, unsigned N, double core, const double *one,
                   const double *two, int precision, int *id)
{
    int rc = check_instance(b, L, N);
    if (rc == DOCI_OK) rc = reserve_records(b, 1);
    if (rc != DOCI_OK) return rc;
    brec *r = &b->rec[b->nrec];
    memset(r, 0, sizeof *r);
    if (rec_tables(r, L, N, one, two) != DOCI_OK) return doci_fail(DOCI_E_NOMEM, "instance tables");
    rec_fill(b, r, L, N, core, precision);
    *id = b->nrec++;
    return DOCI_OK;
}

/* doci_batch_add_many: the intake of the instances split over threads, each
 * building the tables of a contiguous share of them. */
typedef struct {
    brec *rec;
    const unsigned *L, *N;
    const double *const *one, *const *two;
    int k0, k1, rc;
} intake_job;

static void *intake_worker(void *arg)
{
    intake_job *j = arg;
    for (int k = j->k0; k < j->k1 && j->rc == DOCI_OK; ++k)
        j->rc = rec_tables(&j->rec[k], j->L[k], j->N[k], j->one[k], j->two[k]);
    return NULL;
}

#define INTAKE_THREADS 16

int doci_batch_add_many(doci_batch *b, int count, const unsigned *L, const unsigned *N,
                        const double *core, const double *const *one, const double *const *two,
                        int precision, int *ids)
{
    if (count < 0 || (count > 0 && (!L || !N || !core || !one || !two || !ids)))
        return doci_fail(DOCI_E_ARG, "doci_batch_add_many: bad arguments");
    for (int k = 0; k < count; ++k) {
        const int rc = check_instance(b, L[k], N[k]);
        if (rc != DOCI_OK) return rc;
    }
    int rc = reserve_records(b, count);
    if (rc != DOCI_OK) return rc;
    brec *r = &b->rec[b->nrec];
    memset(r, 0, sizeof(brec) * (size_t)count);
    /* one thread per 16 instances, at most INTAKE_THREADS (the caller's
     * thread takes the first share) */
    int nt = (count + 15) / 16;
    if (nt > INTAKE_THREADS) nt = INTAKE_THREADS;
    if (nt < 1) nt = 1;
    intake_job job[INTAKE_THREADS];
    pthread_t th[INTAKE_THREADS];
    int started[INTAKE_THREADS] = {0};
    for (int t = 0; t < nt; ++t) {
        job[t] = (intake_job){r, L, N, one, two, (int)((int64_t)count * t / nt),
                              (int)((int64_t)count * (t + 1) / nt), DOCI_OK};
        if (t > 0) started[t] = pthread_create(&th[t], NULL, intake_worker, &job[t]) == 0;
    }
    intake_worker(&job[0]);
    for (int t = 1; t < nt; ++t) {
        if (started[t]) pthread_join(th[t], NULL);
        else intake_worker(&job[t]); /* no thread: this one does the share */
    }
    rc = DOCI_OK;
    for (int t = 0; t < nt; ++t)
        if (job[t].rc != DOCI_OK) rc = job[t].rc;
    if (rc != DOCI_OK) {
        for (int k = 0; k < count; ++k) free(r[k].tab);
        return doci_fail(DOCI_E_NOMEM, "instance tables");
    }
    for (int k = 0; k < count; ++k) {
        rec_fill(b, &r[k], L[k], N[k], core[k], precision);
        ids[k] = b->nrec++;
    }
    return DOCI_OK;
}

static int flush_admissions(doci_batch *b);

/* Room for `bytes` in the pinned staging buffer (and its device copy),
 * launching the pending admissions and waiting for the copies that read the
 * buffer before it is reused or grown. */
static int stage_reserve(doci_batch *b, size_t bytes, char **p)
{
    if (b->stage_used + bytes > b->stage_cap) {
        const int rc = flush_admissions(b);
        if (rc != DOCI_OK) return rc;
        HIPCHK(hipStreamSynchronize(b->stream));
        b->stage_used = b->stage_sent = 0;
        if (bytes > b->stage_cap) {
            size_t nc = b->stage_cap ? 2 * b->stage_cap : (1u << 20);
            while (nc < bytes) nc *= 2;
            if (b->h_stage) hipHostFree(b->h_stage);
            if (b->d_stage) hipFree(b->d_stage);
            b->h_stage = b->d_stage = NULL;
            b->stage_cap = 0;
            HIPCHK(hipHostMalloc((void **)&b->h_stage, nc, 0));
            HIPCHK(hipMalloc((void **)&b->d_stage, nc));
            b->stage_cap = nc;
        }
    }
    *p = b->h_stage + b->stage_used;
    b->stage_used += ALIGN256(bytes);
    return DOCI_OK;
}

/* Grows the pinned admission list (after the copies that read it) to hold
 * `na` descriptors and `ntile` tiles. */
static int adm_reserve(doci_batch *b, int na, uint32_t ntile)
{
    const size_t need = sizeof(dk_badmit) * (size_t)na + 2 * sizeof(uint32_t) * ntile;
    if (need > b->adm_cap) {
        HIPCHK(hipStreamSynchronize(b->stream));
        size_t nc = b->adm_cap ? 2 * b->adm_cap : (64u << 10);
        while (nc < need) nc *= 2;
        dk_badmit *h = NULL;
        HIPCHK(hipHostMalloc((void **)&h, nc, 0));
        if (b->h_adm) {
            memcpy(h, b->h_adm, sizeof(dk_badmit) * (size_t)b->npend);
            hipHostFree(b->h_adm);
        }
        b->h_adm = h;
        if (b->d_adm) hipFree(b->d_adm);
        b->d_adm = NULL;
        HIPCHK(hipMalloc((void **)&b->d_adm, nc));
        b->adm_cap = nc;
    }
    if (ntile > b->apart_cap) {
        HIPCHK(hipStreamSynchronize(b->stream));
        size_t nc = b->apart_cap ? 2 * b->apart_cap : 4096;
        while (nc < ntile) nc *= 2;
        if (b->d_apart) hipFree(b->d_apart);
        b->d_apart = NULL;
        HIPCHK(hipMalloc((void **)&b->d_apart, sizeof(double) * 2 * nc));
        b->apart_cap = nc;
    }
    return DOCI_OK;
}

/* Launches the pending admissions: one copy of the staged tables, one of
 * the descriptors and tiles, and two kernels for all of them (tables to the
 * instance memory, diagonal, guess at argmin D, start state). */
static int flush_admissions(doci_batch *b)
{
    if (!b->npend) return DOCI_OK;
    const int na = b->npend;
    const uint32_t nt = b->pend_tiles;
    int rc = adm_reserve(b, na, nt);
    if (rc != DOCI_OK) return rc;
    uint32_t *tiles = (uint32_t *)(b->h_adm + na);
    uint32_t q = 0;
    for (int a = 0; a < na; ++a)
        for (uint32_t k = 0; k < b->h_adm[a].ntiles; ++k) {
            tiles[2 * q] = (uint32_t)a;
            tiles[2 * q + 1] = k;
            ++q;
        }
    if (b->stage_used > b->stage_sent)
        HIPCHK(hipMemcpyAsync(b->d_stage + b->stage_sent, b->h_stage + b->stage_sent,
                              b->stage_used - b->stage_sent, hipMemcpyHostToDevice, b->stream));
    b->stage_sent = b->stage_used;
    const size_t ab = sizeof(dk_badmit) * (size_t)na + 2 * sizeof(uint32_t) * nt;
    HIPCHK(hipMemcpyAsync(b->d_adm, b->h_adm, ab, hipMemcpyHostToDevice, b->stream));
    DKCHK(dk_batch_admit((const dk_badmit *)b->d_adm, (uint32_t)na,
                         (const uint32_t *)(b->d_adm + sizeof(dk_badmit) * (size_t)na), nt,
                         b->d_apart, b->pend_maxL, b->pend_maxN, b->stream));
    /* the copies read h_adm asynchronously: the next admissions wait for
     * them (drain or stage_reserve synchronises the stream first) */
    b->npend = 0;
    b->pend_tiles = 0;
    b->pend_maxL = b->pend_maxN = 0;
    return DOCI_OK;
}

/* Builds the instance's tables and diagonal on the device (stream-ordered,
 * no synchronisation) and starts its first phase in device slot `slot`.
 * Large instances also get the shared star tables of their shape and a
 * slot buffer, for the star-kernel sigma. */
static int admit(doci_batch *b, int ri, int slot)
{
    brec *r = &b->rec[ri];
    const unsigned L = r->L, N = r->N;
    const uint32_t n = r->n;
    const size_t l = L, l2 = l * l, tb = table_bytes(L, N), vs = slot_stride(n);
    const double tt0 = dh_now();
    dk_binst *bi = &r->bi;
    memset(bi, 0, sizeof *bi);
    bi->L = L;
    bi->N = N;
    bi->n = n;
    char *h = NULL;
    int rc = stage_reserve(b, tb, &h);
    if (rc != DOCI_OK) return rc;
    memcpy(h, r->tab, tb);
    const size_t mb = main_bytes(&b->o, L, N, n);
    r->dmain = balloc(b, mb);
    r->dx = balloc(b, sizeof(double) * (size_t)n);
    if (!r->dmain || !r->dx) {
        bfree(b, r->dmain, mb);
        bfree(b, r->dx, sizeof(double) * (size_t)n);
        r->dmain = NULL;
        r->dx = NULL;
        return doci_fail(DOCI_E_NOMEM, "instance memory (L=%u N=%u)", L, N);
    }
    char *d = r->dmain;
    double *dK = (double *)d, *dJ = dK + l2, *dX = dJ + l2, *dh = dX + l2;
    uint32_t *db = (uint32_t *)(dh + l);
    const dk_tables dt = {L, N, n, r->core, db, dK, dJ, dX, dh};
    double *D = (double *)(d + tb);
    r->zbytes = 0;
    bi->lane_sigma = 1;
    const double tt1 = dh_now();
    if (uses_star(n)) {
        size_t zelems = 0;
        rc = dh_star_shape(&dt, &r->st, &zelems);
        if (rc == DOCI_OK) {
            r->zbytes = sizeof(double) * zelems;
            r->Z = balloc(b, r->zbytes);
            if (!r->Z) rc = doci_fail(DOCI_E_NOMEM, "slot buffer (L=%u N=%u)", L, N);
        }
        if (rc != DOCI_OK) {
            bfree(b, r->dmain, mb);
            bfree(b, r->dx, sizeof(double) * (size_t)n);
            r->dmain = NULL;
            r->dx = NULL;
            r->zbytes = 0;
            return rc;
        }
        bi->lane_sigma = 0;
    }
    const double tt2 = dh_now();
    r->bytes = mb + r->zbytes;
    r->xbytes = sizeof(double) * (size_t)n;
    b->mem_used += r->bytes + r->xbytes;
    bi->binom = db;
    bi->K = dK;
    bi->D = D;
    bi->slots = d + tb + vs;
    bi->slot_stride = vs;
    /* the Ritz vector is written only where it is read: kept vectors, or the
     * FP64 restart of a mixed-precision instance */
    bi->x = (b->o.keep_vectors || r->want_mixed) ? r->dx : NULL;
    r->prec = r->want_mixed ? 1 : 0;
    /* the device copy of the descriptor comes with the rebuild that puts
     * the instance in the launch lists (admission marks the batch dirty):
     * no kernel reads d_inst[slot] before that.  Tables, diagonal, guess and
     * start state are made by the step's admission launches
     * (flush_admissions), from the tables staged here. */
    rc = adm_reserve(b, b->npend + 1, 0);
    if (rc != DOCI_OK) return rc;
    dk_badmit *ad = &b->h_adm[b->npend++];
    memset(ad, 0, sizeof *ad);
    const size_t so = (size_t)(h - b->h_stage); /* same offset in the device copy */
    double *sK = (double *)(b->d_stage + so), *sJ = sK + l2, *sX = sJ + l2, *sh = sX + l2;
    ad->src = (dk_tables){L, N, n, r->core, (uint32_t *)(sh + l), sK, sJ, sX, sh};
    ad->dst = d;
    ad->tb = (uint32_t)tb;
    ad->tile0 = b->pend_tiles;
    ad->ntiles = (n + DK_ADMIT_ROWS - 1) / DK_ADMIT_ROWS;
    ad->prec = r->prec;
    ad->shift_min_d = r->prec;
    ad->max_its = b->o.max_its;
    ad->keep = b->o.keep;
    ad->max_vec = b->o.max_vec;
    ad->D = D;
    ad->t = bi->slots;
    ad->state = b->d_state + slot;
    ad->tol = r->prec ? b->o.switch_tol : b->o.tol;
    ad->otol = DK_BATCH_DEP_TOL;
    b->pend_tiles += ad->ntiles;
    if (L > b->pend_maxL) b->pend_maxL = L;
    if (N > b->pend_maxN) b->pend_maxN = N;
    free(r->tab);
    r->tab = NULL;
    if (trace_on())
        fprintf(stderr, "[batch] admit L=%u N=%u n=%u: memory %.2f ms, star tables %.2f ms, rest %.2f ms\n",
                L, N, n, 1e3 * (tt1 - tt0), 1e3 * (tt2 - tt1), 1e3 * (dh_now() - tt2));
    r->fresh = b->inflight;
    if (!b->inflight) memset(&b->h_stat[slot], 0, sizeof(dk_bstat));
    r->slot = slot;
    r->state = 1;
    r->res.status = 1;
    r->res.t_start = doci_batch_clock(b);
    b->slot_rec[slot] = ri;
    b->dirty = 1;
    return DOCI_OK;
}

static int free_slot(const doci_batch *b)
{
    for (int s = 0; s < b->cap; ++s)
        if (b->slot_rec[s] < 0) return s;
    return -1;
}

/* FIFO admission while the footprint fits the budget. */
static int admit_queued(doci_batch *b)
{
    int running = 0;
    for (int s = 0; s < b->cap; ++s) running += b->slot_rec[s] >= 0;
    while (b->qhead < b->nrec && b->rec[b->qhead].state != 0) ++b->qhead;
    for (int ri = b->qhead; ri < b->nrec; ++ri) {
        brec *r = &b->rec[ri];
        if (r->state != 0) continue;
        const size_t need = doci_batch_footprint(&b->o, r->L, r->N);
        if (b->mem_used + need > b->o.mem_budget) break;
        const int slot = free_slot(b);
        if (slot < 0) break;
        const int rc = admit(b, ri, slot);
        if (rc == DOCI_E_NOMEM && running > 0) break; /* fragmented: retry when others finish */
        if (rc != DOCI_OK) return rc;
        ++running;
    }
    return flush_admissions(b);
}

static int grow(void **h, void **d, size_t *cap, size_t need, size_t elem)
{
    if (need <= *cap) return DOCI_OK;
    size_t nc = *cap ? *cap : 64;
    while (nc < need) nc *= 2;
    if (h) {
        /* pinned: the copies of these lists are enqueued asynchronously (a
         * copy from pageable memory would hold the host for its duration) */
        void *nh = NULL;
        if (hipHostMalloc(&nh, elem * nc, 0) != hipSuccess || !nh)
            return doci_fail(DOCI_E_NOMEM, "batch lists");
        if (*h) {
            memcpy(nh, *h, elem * *cap);
            hipHostFree(*h);
        }
        *h = nh;
    }
    if (*d) hipFree(*d);
    *d = NULL;
    if (hipMalloc(d, elem * nc) != hipSuccess) return doci_fail(DOCI_E_NOMEM, "batch lists");
    *cap = nc;
    return DOCI_OK;
}

static void drop_graph(bgroup *g)
{
    if (g->gexec) hipGraphExecDestroy(g->gexec);
    if (g->graph) hipGraphDestroy(g->graph);
    g->gexec = NULL;
    g->graph = NULL;
    g->gM = 0;
}

/* Launch order of the tiles of group p (h_tiles, g->ntile launch positions).
 * Workgroup w of a launch runs on XCD w % DH_XCDS (round-robin dispatch), and
 * each XCD has its own L2.  Instances with the star-kernel sigma stream their
 * vectors: their tiles come first, in order, spread over every XCD.  Each
 * lane-sigma instance gathers at random from its own vectors: all its tiles
 * go to one XCD, so those vectors are cached in one L2 only.  Instances are
 * dealt to the XCDs largest first, each to the least loaded, and position
 * base + DH_XCDS q + x takes the q-th tile of XCD x.  Padding positions carry
 * tile index 0xffffffff, which every tile kernel skips; partial sums are
 * indexed by tile0 + tile, not by position. */
static int xcd_order(doci_batch *b, bgroup *g, int p)
{
    uint32_t pos = 0;
    for (int s = 0; s < b->cap; ++s) { /* star-sigma instances, in order */
        const int ri = b->slot_rec[s];
        if (ri < 0 || b->rec[ri].prec != p || b->rec[ri].bi.lane_sigma) continue;
        for (uint32_t k = 0; k < b->rec[ri].bi.ntiles; ++k) {
            g->h_tiles[2 * pos] = (uint32_t)s;
            g->h_tiles[2 * pos + 1] = k;
            ++pos;
        }
    }
    int *xs = malloc(sizeof(int) * (size_t)b->cap), *ord = malloc(sizeof(int) * (size_t)b->cap);
    if (!xs || !ord) {
        free(xs);
        free(ord);
        return doci_fail(DOCI_E_NOMEM, "batch lists");
    }
    int nl = 0, pad_slot = -1;
    for (int s = 0; s < b->cap; ++s) {
        const int ri = b->slot_rec[s];
        if (ri < 0 || b->rec[ri].prec != p) continue;
        pad_slot = s;
        if (b->rec[ri].bi.lane_sigma) ord[nl++] = s;
    }
    /* largest first (insertion sort: a few hundred instances at most per rebuild) */
    for (int i = 1; i < nl; ++i) {
        const int v = ord[i];
        const uint32_t tv = b->rec[b->slot_rec[v]].bi.ntiles;
        int j = i;
        while (j > 0 && b->rec[b->slot_rec[ord[j - 1]]].bi.ntiles < tv) {
            ord[j] = ord[j - 1];
            --j;
        }
        ord[j] = v;
    }
    uint32_t load[DH_XCDS] = {0}, maxload = 0;
    for (int i = 0; i < nl; ++i) {
        int x = 0;
        for (int y = 1; y < DH_XCDS; ++y)
            if (load[y] < load[x]) x = y;
        xs[ord[i]] = x;
        load[x] += b->rec[b->slot_rec[ord[i]]].bi.ntiles;
        if (load[x] > maxload) maxload = load[x];
    }
    const uint32_t base = nl ? (pos + DH_XCDS - 1) / DH_XCDS * DH_XCDS : pos;
    for (; pos < base + DH_XCDS * maxload; ++pos) { /* padding, overwritten below where used */
        g->h_tiles[2 * pos] = (uint32_t)(pad_slot < 0 ? 0 : pad_slot);
        g->h_tiles[2 * pos + 1] = 0xffffffffu;
    }
    uint32_t next[DH_XCDS] = {0};
    for (int i = 0; i < nl; ++i) {
        const int s = ord[i], x = xs[s];
        for (uint32_t k = 0; k < b->rec[b->slot_rec[s]].bi.ntiles; ++k) {
            const uint32_t q = base + DH_XCDS * next[x]++ + (uint32_t)x;
            g->h_tiles[2 * q] = (uint32_t)s;
            g->h_tiles[2 * q + 1] = k;
        }
    }
    free(xs);
    free(ord);
    g->ntile = pos;
    return DOCI_OK;
}

/* Instance lists and tile lists of both precision groups. */
static int rebuild(doci_batch *b)
{
    const uint32_t rows = (uint32_t)b->tile_rows;
    for (int p = 0; p < 2; ++p) {
        bgroup *g = &b->grp[p];
        drop_graph(g);
        int cnt = 0;
        size_t ntile = 0;
        g->maxL = g->maxN = 0;
        for (int s = 0; s < b->cap; ++s) {
            const int ri = b->slot_rec[s];
            if (ri < 0 || b->rec[ri].prec != p) continue;
            ++cnt;
            const uint32_t rr = rows_for(b->rec[ri].n, rows);
            ntile += (b->rec[ri].n + rr - 1) / rr;
        }
        /* launch positions: the tiles, plus padding of the XCD order below
         * (at most 7 + 8 x the tiles of the largest lane-sigma instance) */
        size_t tmax = 0;
        for (int s = 0; s < b->cap; ++s) {
            const int ri = b->slot_rec[s];
            if (ri < 0 || b->rec[ri].prec != p) continue;
            const uint32_t rr = rows_for(b->rec[ri].n, rows);
            if ((b->rec[ri].n + rr - 1) / rr > tmax) tmax = (b->rec[ri].n + rr - 1) / rr;
        }
        const size_t nlaunch = ntile + 7 + DH_XCDS * tmax;
        int rc = grow((void **)&g->h_list, (void **)&g->d_list, &g->list_cap, (size_t)cnt + 1,
                      sizeof(uint32_t));
        if (rc == DOCI_OK)
            rc = grow((void **)&g->h_tiles, (void **)&g->d_tiles, &g->tiles_cap, 2 * nlaunch + 2,
                      sizeof(uint32_t));
        if (rc == DOCI_OK) rc = grow(NULL, (void **)&g->d_part, &g->part_cap, ntile * DK_BPV + 1, sizeof(double));
        if (rc != DOCI_OK) return rc;
        uint32_t t = 0;
        g->n = 0;
        for (int s = 0; s < b->cap; ++s) {
            const int ri = b->slot_rec[s];
            if (ri < 0 || b->rec[ri].prec != p) continue;
            brec *r = &b->rec[ri];
            g->h_list[g->n++] = (uint32_t)s;
            r->bi.tile0 = t;
            r->bi.tile_rows = rows_for(r->n, rows);
            r->bi.ntiles = (r->n + r->bi.tile_rows - 1) / r->bi.tile_rows;
            t += r->bi.ntiles;
            if (r->bi.lane_sigma && r->L > g->maxL) g->maxL = r->L;
            if (r->bi.lane_sigma && r->N > g->maxN) g->maxN = r->N;
            b->h_inst[s] = r->bi;
        }
        rc = xcd_order(b, g, p);
        if (rc != DOCI_OK) return rc;
        g->fresh = 1;
        if (g->n) {
            HIPCHK(hipMemcpyAsync(g->d_list, g->h_list, sizeof(uint32_t) * g->n,
                                  hipMemcpyHostToDevice, b->stream));
            HIPCHK(hipMemcpyAsync(g->d_tiles, g->h_tiles, sizeof(uint32_t) * 2 * g->ntile,
                                  hipMemcpyHostToDevice, b->stream));
        }
    }
    /* every running instance's descriptor in one copy (slots without an
     * instance carry stale entries that no launch reads) */
    HIPCHK(hipMemcpyAsync(b->d_inst, b->h_inst, sizeof(dk_binst) * (size_t)b->cap,
                          hipMemcpyHostToDevice, b->stream));
    /* the host lists are sources of async copies: they are rewritten only by
     * the next rebuild, which runs after the stream has drained */
    b->dirty = 0;
    return DOCI_OK;
}

/* One Davidson iteration of every instance of group p (basis bucket M). */
static int enqueue_iteration(doci_batch *b, int p, int M)
{
    bgroup *g = &b->grp[p];
    hipStream_t s = b->stream;
    DKCHK(dk_batch_small(DK_BPH_A0, g->d_list, g->n, b->d_inst, b->d_state, g->d_part, s));
    DKCHK(dk_batch_stream(DK_BOP_ORTHO, p, M, b->d_inst, b->d_state, g->d_tiles, g->ntile, g->d_part, s));
    DKCHK(dk_batch_small(DK_BPH_B, g->d_list, g->n, b->d_inst, b->d_state, g->d_part, s));
    if (g->maxL)
        DKCHK(dk_batch_sigma_lane(p, g->maxL, g->maxN, b->d_inst, b->d_state, g->d_tiles, g->ntile, s));
    for (int k = 0; k < g->n; ++k) { /* large instances: star kernel, operands read on the device */
        const brec *r = &b->rec[b->slot_rec[g->h_list[k]]];
        if (r->Z) {
            const size_t l2 = (size_t)r->L * r->L;
            const double *dK = r->bi.K;
            const dk_tables dt = {r->L, r->N, r->n, r->core, r->bi.binom, dK, dK + l2, dK + 2 * l2,
                                  dK + 3 * l2};
            DKCHK(dk_sigma_star_io(&dt, r->st, r->bi.D, &b->d_state[g->h_list[k]].sio, r->Z, p, s));
        }
    }
    DKCHK(dk_batch_stream(DK_BOP_SCOL, p, M, b->d_inst, b->d_state, g->d_tiles, g->ntile, g->d_part, s));
    DKCHK(dk_batch_small(DK_BPH_D, g->d_list, g->n, b->d_inst, b->d_state, g->d_part, s));
    DKCHK(dk_batch_stream(DK_BOP_DEFLATE, p, M, b->d_inst, b->d_state, g->d_tiles, g->ntile, g->d_part, s));
    DKCHK(dk_batch_small(DK_BPH_E, g->d_list, g->n, b->d_inst, b->d_state, g->d_part, s));
    DKCHK(dk_batch_stream(DK_BOP_RESID, p, M, b->d_inst, b->d_state, g->d_tiles, g->ntile, g->d_part, s));
    DKCHK(dk_batch_small(DK_BPH_A, g->d_list, g->n, b->d_inst, b->d_state, g->d_part, s));
    return DOCI_OK;
}

static int run_iteration(doci_batch *b, int p, int M)
{
    bgroup *g = &b->grp[p];
    /* a group that changed since the last iteration runs with direct
     * launches; it is captured once it stays the same for an iteration, so
     * that a stream of arrivals does not pay a capture every step */
    if (!b->o.use_graphs || g->fresh) {
        g->fresh = 0;
        return enqueue_iteration(b, p, M);
    }
    if (!g->gexec || g->gM != M) {
        drop_graph(g);
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_iteration(b, p, M);
        hipGraph_t graph = NULL;
        const hipError_t e = hipStreamEndCapture(b->stream, &graph);
        if (rc != DOCI_OK) {
            if (graph) hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return doci_fail(DOCI_E_HIP, "graph capture: %s", hipGetErrorString(e));
        g->graph = graph;
        HIPCHK(hipGraphInstantiate(&g->gexec, graph, NULL, NULL, 0));
        g->gM = M;
    }
    HIPCHK(hipGraphLaunch(g->gexec, b->stream));
    ++b->graph_launches;
    return DOCI_OK;
}

static int bucket_of(int m) { return m <= 8 ? 8 : m <= 16 ? 16 : DK_BMAX_VEC; }

/* Tuned tile sizes, process-wide, by the shape of the admitted set (lane-
 * and star-sigma instance counts, lane determinants in units of 2^16, the
 * largest lane L and N): a later batch of the same set skips the timing. */
#define TUNE_CACHE 16
static struct {
    uint64_t key[5];
    int rows;
} g_tune[TUNE_CACHE];
static int g_ntune;
static pthread_mutex_t g_tune_lock = PTHREAD_MUTEX_INITIALIZER;

static void tune_key(const doci_batch *b, uint64_t key[5])
{
    memset(key, 0, sizeof(uint64_t) * 5);
    for (int s = 0; s < b->cap; ++s) {
        const int ri = b->slot_rec[s];
        if (ri < 0) continue;
        const brec *r = &b->rec[ri];
        if (!r->bi.lane_sigma) {
            ++key[1];
            continue;
        }
        ++key[0];
        key[2] += r->n;
        if (r->L > key[3]) key[3] = r->L;
        if (r->N > key[4]) key[4] = r->N;
    }
    key[2] >>= 16;
}

/* Times the batched lane sigma and one streaming pass for each tile size
 * while every running instance is still at its start (m = 0): the sigma
 * then writes only a W slot not yet in use and the pass only reads t. */
static int autotune(doci_batch *b)
{
    static const int cand[] = {512, 1024, 2048, 4096};
    uint64_t key[5];
    tune_key(b, key);
    pthread_mutex_lock(&g_tune_lock);
    for (int i = 0; i < g_ntune; ++i)
        if (!memcmp(g_tune[i].key, key, sizeof key)) {
            b->tile_rows = g_tune[i].rows;
            pthread_mutex_unlock(&g_tune_lock);
            b->tuned = 1;
            b->dirty = 1;
            return DOCI_OK;
        }
    pthread_mutex_unlock(&g_tune_lock);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 1e30;
    int best_rows = 1024;
    for (size_t c = 0; c < sizeof cand / sizeof cand[0]; ++c) {
        b->tile_rows = cand[c];
        int rc = rebuild(b);
        if (rc != DOCI_OK) return rc;
        float ms = 0;
        for (int rep = 0; rep < 3; ++rep) {
            HIPCHK(hipEventRecord(e0, b->stream));
            for (int p = 0; p < 2; ++p) {
                bgroup *g = &b->grp[p];
                if (!g->n) continue;
                if (g->maxL)
                    DKCHK(dk_batch_sigma_lane(p, g->maxL, g->maxN, b->d_inst, b->d_state, g->d_tiles,
                                              g->ntile, b->stream));
                DKCHK(dk_batch_stream(DK_BOP_ORTHO, p, 8, b->d_inst, b->d_state, g->d_tiles,
                                      g->ntile, g->d_part, b->stream));
            }
            HIPCHK(hipEventRecord(e1, b->stream));
            HIPCHK(hipEventSynchronize(e1));
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, e0, e1));
            if (rep > 0) ms += t; /* first repetition warms up */
        }
        if (ms < best) {
            best = ms;
            best_rows = cand[c];
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    b->tile_rows = best_rows;
    b->tuned = 1;
    b->dirty = 1;
    pthread_mutex_lock(&g_tune_lock);
    const int slot = g_ntune < TUNE_CACHE ? g_ntune++ : (int)(key[2] % TUNE_CACHE);
    memcpy(g_tune[slot].key, key, sizeof key);
    g_tune[slot].rows = best_rows;
    pthread_mutex_unlock(&g_tune_lock);
    return DOCI_OK;
}

/* Grows the pinned restart list to `nr` descriptors and `ntile` tiles,
 * keeping the descriptors recorded so far.  Called from drain, after the
 * stream has synchronised: the previous list's copy has completed. */
static int rst_reserve(doci_batch *b, int nr, uint32_t ntile)
{
    const size_t need = sizeof(dk_brestart) * (size_t)nr + 2 * sizeof(uint32_t) * ntile;
    if (need <= b->rst_cap) return DOCI_OK;
    size_t nc = b->rst_cap ? 2 * b->rst_cap : (64u << 10);
    while (nc < need) nc *= 2;
    dk_brestart *h = NULL;
    HIPCHK(hipHostMalloc((void **)&h, nc, 0));
    if (b->h_rst) {
        memcpy(h, b->h_rst, sizeof(dk_brestart) * (size_t)b->nrst);
        hipHostFree(b->h_rst);
    }
    b->h_rst = h;
    if (b->d_rst) hipFree(b->d_rst);
    b->d_rst = NULL;
    HIPCHK(hipMalloc((void **)&b->d_rst, nc));
    b->rst_cap = nc;
    return DOCI_OK;
}

/* The FP32 phase of a mixed-precision instance met the switch tolerance: its
 * FP64 phase restarts from the Ritz vector (recorded here, launched for all
 * of the drain's restarts by flush_restarts). */
static int restart_fp64(doci_batch *b, brec *r, const dk_bstat *st)
{
    int rc = rst_reserve(b, b->nrst + 1, 0);
    if (rc != DOCI_OK) return rc;
    r->res.iterations_fp32 = st->its;
    r->prec = 0;
    dk_brestart *q = &b->h_rst[b->nrst++];
    memset(q, 0, sizeof *q);
    q->x = r->bi.x;
    q->t = (double *)r->bi.slots; /* slot 0 holds t at a (re)start */
    q->state = b->d_state + r->slot;
    q->n = r->n;
    q->its0 = (uint32_t)st->its;
    q->max_its = b->o.max_its;
    q->keep = b->o.keep;
    q->max_vec = b->o.max_vec;
    q->tol = b->o.tol;
    q->otol = DK_BATCH_DEP_TOL;
    b->rst_tiles += (r->n + DK_ADMIT_ROWS - 1) / DK_ADMIT_ROWS;
    b->dirty = 1;
    return DOCI_OK;
}

/* One copy of the descriptors and tiles and one launch for every restart
 * recorded by this drain. */
static int flush_restarts(doci_batch *b)
{
    if (!b->nrst) return DOCI_OK;
    const int nr = b->nrst;
    const uint32_t nt = b->rst_tiles;
    int rc = rst_reserve(b, nr, nt);
    if (rc != DOCI_OK) return rc;
    uint32_t *tiles = (uint32_t *)(b->h_rst + nr);
    uint32_t q = 0;
    for (int i = 0; i < nr; ++i)
        for (uint32_t k = 0; k < (b->h_rst[i].n + DK_ADMIT_ROWS - 1) / DK_ADMIT_ROWS; ++k) {
            tiles[2 * q] = (uint32_t)i;
            tiles[2 * q + 1] = k;
            ++q;
        }
    const size_t bytes = sizeof(dk_brestart) * (size_t)nr + 2 * sizeof(uint32_t) * nt;
    HIPCHK(hipMemcpyAsync(b->d_rst, b->h_rst, bytes, hipMemcpyHostToDevice, b->stream));
    DKCHK(dk_batch_restart((const dk_brestart *)b->d_rst,
                           (const uint32_t *)(b->d_rst + sizeof(dk_brestart) * (size_t)nr), nt,
                           b->stream));
    b->nrst = 0;
    b->rst_tiles = 0;
    return DOCI_OK;
}

static void finish(doci_batch *b, brec *r, const dk_bstat *st)
{
    const double t0 = dh_now();
    r->res.status = st->status;
    r->res.iterations = st->its;
    if (r->prec) r->res.iterations_fp32 = st->its;
    r->res.energy = st->energy;
    r->res.residual = st->res;
    r->res.t_done = doci_batch_clock(b);
    r->state = 2;
    b->slot_rec[r->slot] = -1;
    r->slot = -1;
    rec_free_main(b, r);
    if (!b->o.keep_vectors) free_x(b, r);
    b->dirty = 1;
    if (trace_on())
        fprintf(stderr, "[batch] retire L=%u N=%u: %.2f ms\n", r->L, r->N, 1e3 * (dh_now() - t0));
}

/* Waits for the queued iteration and acts on the statuses it produced. */
static int drain(doci_batch *b)
{
    if (!b->inflight) return DOCI_OK;
    HIPCHK(hipStreamSynchronize(b->stream));
    b->inflight = 0;
    b->stage_used = b->stage_sent = 0; /* every staged copy has completed */
    for (int s = 0; s < b->cap; ++s) {
        const int ri = b->slot_rec[s];
        if (ri < 0) continue;
        brec *r = &b->rec[ri];
        if (r->fresh) { /* the copied status predates this instance */
            r->fresh = 0;
            memset(&b->h_stat[s], 0, sizeof(dk_bstat));
            continue;
        }
        const dk_bstat *st = &b->h_stat[s];
        if (st->status == DK_B_ACTIVE) continue;
        if (r->prec == 1 && st->status == DK_B_CONVERGED) {
            const int rc = restart_fp64(b, r, st);
            if (rc != DOCI_OK) return rc;
        } else {
            finish(b, r, st);
        }
    }
    return flush_restarts(b);
}

/* Pipelined: first finish the iteration queued by the previous call, then
 * queue the next iteration of the running instances, then admit queued
 * instances (their host-side set-up overlaps that iteration; they join the
 * next one), and return without waiting, so that host work between calls
 * (request intake) overlaps the GPU too. */
int doci_batch_step(doci_batch *b, int *running, int *waiting)
{
    const double ts0 = dh_now();
    int rc = drain(b);
    const double ts1 = dh_now();
    double ts_rb = ts1;
    if (rc != DOCI_OK) return rc;
    int nrun = 0;
    for (int s = 0; s < b->cap; ++s) nrun += b->slot_rec[s] >= 0;
    const int admit_first = nrun == 0; /* nothing to overlap with */
    if (admit_first) {
        rc = admit_queued(b);
        if (rc != DOCI_OK) return rc;
    }
    nrun = 0;
    for (int s = 0; s < b->cap; ++s) nrun += b->slot_rec[s] >= 0;
    if (nrun) {
        if (!b->tuned) {
            rc = autotune(b);
            if (rc != DOCI_OK) return rc;
        }
        if (b->dirty) {
            rc = rebuild(b);
            if (rc != DOCI_OK) return rc;
        }
        ts_rb = dh_now();
        for (int p = 0; p < 2; ++p) {
            bgroup *g = &b->grp[p];
            if (!g->n) continue;
            int mmax = 0;
            for (int k = 0; k < g->n; ++k) {
                const int m = b->h_stat[g->h_list[k]].m;
                if (m > mmax) mmax = m;
            }
            rc = run_iteration(b, p, bucket_of(mmax + 1));
            if (rc != DOCI_OK) return rc;
        }
        ++b->iterations;
        HIPCHK(hipMemcpy2DAsync(b->h_stat, sizeof(dk_bstat), b->d_state, sizeof(dk_bstate),
                                sizeof(dk_bstat), (size_t)b->cap, hipMemcpyDeviceToHost, b->stream));
        b->inflight = 1;
    }
    const double ts2 = dh_now();
    if (!admit_first) {
        rc = admit_queued(b);
        if (rc != DOCI_OK) return rc;
    }
    if (trace_on())
        fprintf(stderr, "[batch] step %d: drain %.2f ms, launch %.2f ms (rebuild %.2f), admit %.2f ms\n",
                b->iterations, 1e3 * (ts1 - ts0), 1e3 * (ts2 - ts1), 1e3 * (ts_rb - ts1),
                1e3 * (dh_now() - ts2));
    int nr = 0, nw = 0;
    for (int i = 0; i < b->nrec; ++i) {
        nr += b->rec[i].state == 1;
        nw += b->rec[i].state == 0;
    }
    if (running) *running = nr;
    if (waiting) *waiting = nw;
    return DOCI_OK;
}

int doci_batch_run(doci_batch *b)
{
    for (;;) {
        int nr = 0, nw = 0;
        const int rc = doci_batch_step(b, &nr, &nw);
        if (rc != DOCI_OK) return rc;
        if (!nr && !nw) return DOCI_OK;
        if (!nr && nw) return doci_fail(DOCI_E_NOMEM, "queued instances do not fit the memory budget");
    }
}

int doci_batch_result_of(doci_batch *b, int id, doci_batch_result *res, double *x)
{
    if (id < 0 || id >= b->nrec) return doci_fail(DOCI_E_ARG, "no instance %d", id);
    brec *r = &b->rec[id];
    if (b->inflight) HIPCHK(hipStreamSynchronize(b->stream)); /* h_stat is being written */
    if (res) {
        *res = r->res;
        if (r->state == 1) {
            const dk_bstat *st = &b->h_stat[r->slot];
            res->iterations = st->its;
            res->energy = st->energy;
            res->residual = st->res;
        }
    }
    if (x) {
        if (r->state != 2 || !r->dx) return doci_fail(DOCI_E_ARG, "instance %d has no eigenvector", id);
        HIPCHK(hipMemcpy(x, r->dx, sizeof(double) * r->n, hipMemcpyDeviceToHost));
    }
    return DOCI_OK;
}

int doci_batch_release(doci_batch *b, int id)
{
    if (id < 0 || id >= b->nrec) return doci_fail(DOCI_E_ARG, "no instance %d", id);
    brec *r = &b->rec[id];
    if (r->state != 2) return doci_fail(DOCI_E_ARG, "instance %d has not finished", id);
    if (b->inflight) HIPCHK(hipStreamSynchronize(b->stream));
    free_x(b, r);
    return DOCI_OK;
}

int doci_batch_stats(const doci_batch *b, int *tile_rows, int *iterations, int *graph_launches)
{
    if (tile_rows) *tile_rows = b->tile_rows;
    if (iterations) *iterations = b->iterations;
    if (graph_launches) *graph_launches = b->graph_launches;
    return DOCI_OK;
}
