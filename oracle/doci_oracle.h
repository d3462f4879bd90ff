/*
 * doci_oracle.h -- CPU restatement of klgunst/doci's sparse-eigensolve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library (doci_amd) neither links nor calls anything here.
 *
 * Every function names the reference file:line range it restates.  Masks are
 * 64-bit (the reference's are 32-bit, src/doci.c:84-89) and every size is
 * size_t, so L=34 runs here although it cannot in the reference.
 */
#ifndef DOCI_ORACLE_H
#define DOCI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/read_fcidump.c:18-82.  Returns 0, or -1 (open), -2 (header),
 * -(1000+line) for a malformed integral line.  *one (L*L) and *two (L^4) are
 * calloc'd and owned by the caller (free with orc_free). */
int orc_read_fcidump(const char *path, unsigned *L, unsigned *Na, unsigned *Nb,
                     double *core, double **one, double **two);
void orc_free(void *p);

/* src/doci.c:630-665, in place. */
void orc_transform_integrals(double *one, double *two, unsigned L);
/* src/doci.c:667-683: K[i*L+k] = (ik|ik), K[i*L+i] = 0. */
void orc_exchange_table(const double *two, unsigned L, double *K);
/* src/doci.c:687-701 (64-bit result). */
uint64_t orc_binom(unsigned k, unsigned n);

/* src/doci.c:472-548: rank -> mask table in the reference's lexicographic
 * order of sorted occupied tuples (last pair fastest). */
void orc_lex_masks(unsigned N, unsigned L, uint64_t n, uint64_t *masks);
/* Closed-form lexicographic rank of a mask (pinned by brute force in tests). */
uint64_t orc_rank(uint64_t mask, unsigned N, unsigned L);

/* src/doci.c:550-628 with the exact addition order; one, two already
 * transformed.  Threads only split the rank range: every element is computed
 * with the same order, so the result does not depend on the thread count. */
void orc_diagonal(const double *one, const double *two, double core,
                  unsigned N, unsigned L, uint64_t n, double *D, int threads);
/* The same diagonal entries at `count` given ranks only (large n). */
void orc_diagonal_rows(const double *one, const double *two, double core, unsigned N,
                       unsigned L, const uint64_t *ranks, int64_t count, double *D);

/* src/doci.c:851-978: lower-triangle connections of determinant `rank`, in the
 * reference's enumeration order.  Writes up to cap (rank, K-index) pairs and
 * returns the count. */
int orc_connections(uint64_t rank, unsigned N, unsigned L, uint64_t *J,
                    uint32_t *kidx, int cap);

/* src/doci.c:164-238: sigma = D.c + lower-triangle walk scattered both ways,
 * into `threads` private accumulators summed in thread order (the reference's
 * layout; threads=1 gives its single-thread summation order). */
void orc_sigma(const double *K, const double *D, unsigned N, unsigned L,
               uint64_t n, const double *c, double *sigma, int threads);

/* src/davidson.c:33-128 (MGS, |D-theta| preconditioner with 1e-12 floor,
 * deflation at max_vec keeping `keep`).  The m x m eigensolve is a cyclic
 * Jacobi (the reference calls LAPACK dsyev).  trace (may be NULL) receives
 * per-iteration residual norms and energies, up to trace_cap entries each.
 * Returns its >= max_its like the reference; *its_out gets the count. */
int orc_davidson(double *result, double *energy, int max_vec, int keep,
                 double tol, const double *K, const double *D, unsigned N,
                 unsigned L, uint64_t n, int max_its, int threads,
                 double *trace_res, double *trace_e, int trace_cap,
                 int *its_out);
/* orc_davidson that also records, per iteration, the wall seconds since the
 * solve started (trace_t, may be NULL). */
int orc_davidson_timed(double *result, double *energy, int max_vec, int keep,
                       double tol, const double *K, const double *D, unsigned N,
                       unsigned L, uint64_t n, int max_its, int threads,
                       double *trace_res, double *trace_e, double *trace_t,
                       int trace_cap, int *its_out);

/* src/CG.c:13-164: Rayleigh-quotient CG, preconditioned by 1/D when
 * precond != 0.  The 2x2 generalized eigensolve follows LAPACK dsygv's
 * steps (Cholesky, reduction, 2x2 rotation, back-substitution).  Returns 1
 * always, like the reference. */
int orc_ray_q(double *vec, double *energy, double tol, int max_its,
              int precond, const double *K, const double *D, unsigned N,
              unsigned L, uint64_t n, int threads, double *trace_res,
              double *trace_e, int trace_cap, int *its_out);
int orc_ray_q_planted(double *vec, double *energy, double tol, int max_its,
              int precond, const double *K, const double *D, unsigned N,
              unsigned L, uint64_t n, int threads, double *trace_res,
              double *trace_e, int trace_cap, int *its_out, int mistake);

/* Dense helpers exposed for their own tests. */
void orc_jacobi_eig(int m, const double *A, int lda, double *w, double *Z,
                    int ldz);
int orc_sygv2(const double *A, const double *B, double *w, double *Z);

#ifdef __cplusplus
}
#endif
#endif
