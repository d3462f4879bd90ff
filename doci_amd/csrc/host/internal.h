/* internal.h -- shared declarations of the C host layer. */
#ifndef DOCI_INTERNAL_H
#define DOCI_INTERNAL_H

#include <stdint.h>
#include <stdio.h>

#include "../../../include/doci_amd.h"
#include "../dk.h"

#define DOCI_OP_MAGIC 0x444f4349414d4431ull /* "DOCIAMD1" */
/* Below this many determinants the lane kernel is the default sigma (no
 * decomposition tables to build); from it on, the star kernel. */
#define DOCI_STAR_MIN_N (1u << 20)

/* Process-wide cache of device buffers (devmem.c): released buffers are
 * kept for the next request they fit.  Release only what no stream still
 * uses. */
hipError_t dh_dmalloc(void **p, size_t bytes);
void dh_dfree(void *p);
void dh_dcache_trim(void);

struct doci_op {
    uint64_t magic;
    unsigned L, N;
    int64_t n;
    double core;
    /* host tables (transformed integrals) */
    double *K, *J, *X, *h;
    uint32_t *binom;
    double *D_host; /* lazily filled for doci_get_diagonal */
    /* device */
    hipStream_t stream;
    dk_tables dt;   /* device pointers */
    void *d_tables; /* one allocation holding the tables */
    double *d_D;
    double *d_red;  /* reduction scratch */
    double *d_out;  /* small reduction results */
    double *h_out;  /* pinned mirror of d_out */
    int64_t *d_idx;
    void *d_io[2];  /* staging for host-pointer sigma calls (lazy) */
    void *pool;        /* Davidson subspace kept between solves (grows, freed with op) */
    size_t pool_bytes;
    int sigma_kind;    /* 0 star (default), 1 lane, 2 star checked, 3 slab, 4 slab checked */
    void *slabs;       /* slab-kernel tables (slabs.c) */
    void *stars;       /* star-kernel tables (stars.c) */
    /* scratch for shifted sigma (lazy) */
};

/* error reporting */
int doci_fail(int code, const char *fmt, ...);
#define HIPCHK(expr)                                                              \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess)                                                     \
            return doci_fail(DOCI_E_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #expr, \
                             hipGetErrorString(e_));                              \
    } while (0)
#define DKCHK(expr)                                                               \
    do {                                                                          \
        int e_ = (expr);                                                          \
        if (e_ != 0)                                                              \
            return doci_fail(DOCI_E_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #expr, \
                             hipGetErrorString((hipError_t)e_));                  \
    } while (0)

/* fcidump.c */
int dh_read_fcidump(const char *path, unsigned *L, unsigned *Na, unsigned *Nb, double *core,
                    double **one, double **two, int *errline);

/* integrals.c */
uint64_t dh_binom(unsigned k, unsigned n);

/* operator.c */
int dh_is_op(const void *p);
/* K (zero diagonal), J, X (L x L), h' (L) and binomials ((L+1) x (N+2)) from
 * raw integrals, transformed as src/doci.c:630-683. */
int dh_build_tables(unsigned L, unsigned N, const double *one_raw, const double *two_raw,
                    double *K, double *J, double *X, double *h, uint32_t *binom);
/* Operator from tables already transformed (dh_build_tables). */
int dh_op_from_tables(unsigned L, unsigned N, double core, const double *K, const double *J,
                      const double *X, const double *h, doci_op **out);

/* dense.c: small dense eigensolvers */
void dh_jacobi(int m, double *A, int lda, double *w, double *Z, int ldz);
int dh_gen_eig(int m, const double *S, const double *G, int ld, double *w, double *Y, int ldy);
double dh_gram_solve(int m, const double *G, int ld, const double *b, double *c);
int dh_sygv2(const double *A, const double *B, double *w, double *Z);

/* Solver context: everything the device Davidson / CG need besides the
 * options.  `apply` computes out = (H - shift) in on device vectors. */
typedef struct dh_ctx {
    int64_t n;
    hipStream_t stream;
    const double *d_D;
    double *d_red, *d_out, *h_out;
    int64_t *d_idx;
    int (*apply)(struct dh_ctx *c, const void *in, void *out, int prec, double shift);
    void *user;
    int fp32_ok; /* apply supports prec 1 */
    void **pool;        /* optional persistent workspace (owned by an operator) */
    size_t *pool_bytes;
} dh_ctx;

void dh_ctx_of_op(doci_op *op, dh_ctx *c);
int dh_fetch(dh_ctx *c, int nvals, double *host);

/* davidson.c / cg.c: d_x holds the FP64 guess on entry, the unit Ritz
 * vector on exit.  Return the reference's code (Davidson 0/1, CG 1) or a
 * negative DOCI_E_*. */
int dh_davidson(dh_ctx *c, const doci_solve_opts *o, double *d_x, double *energy,
                doci_solve_info *info);
int dh_ray_q(dh_ctx *c, const doci_solve_opts *o, int precond, double *d_x, double *energy,
             doci_solve_info *info);

double dh_now(void);

/* min(D) of the context's diagonal (device argmin, then one value copied
 * back): the shift that keeps FP32-stored (H - shift) v entries small. */
int dh_diag_min(dh_ctx *c, double *dmin);

/* slabs.c */
int dh_slabs_build(doci_op *op, unsigned emax, unsigned gmax, int xcd_runs);
void dh_slabs_free(doci_op *op);
int dh_slabs_sigma(doci_op *op, const void *in, void *out, int prec, double shift);
void dh_slabs_info(const doci_op *op, unsigned *a, unsigned *b, unsigned *ngroup);
int64_t dh_slabs_violations(doci_op *op);

/* stars.c */
int dh_stars_build(doci_op *op, unsigned lds_budget, unsigned sgmax, unsigned main_emax);
void dh_stars_free(doci_op *op);
void dh_stars_limits(const doci_op *op, unsigned *lds_budget, unsigned *sgmax, unsigned *main_emax);
int dh_stars_sigma(doci_op *op, const void *in, void *out, int prec, double shift, int checked);
/* The star sigma with the slot buffer in the library's static device array
 * (no allocation; the star tables must exist). */
int dh_stars_sigma_static(doci_op *op, const void *in, void *out, int prec, double shift);
int64_t dh_stars_violations(doci_op *op);
int dh_stars_set_pass(doci_op *op, int pass);
/* The star tables of shape (dt->L, dt->N) with default limits (shared by
 * every user of that shape, built on first request) and the number of
 * slot-buffer entries a sigma of that shape needs. */
int dh_star_shape(const dk_tables *dt, const dk_stars **st, size_t *zelems);
void dh_stars_info(const doci_op *op, unsigned *a, unsigned *b, unsigned *ngroup,
                   unsigned *nsgroup, size_t *zelems);

#endif
