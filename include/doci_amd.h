/*
 * doci_amd.h -- public C ABI of libdoci_amd.so, a DOCI ground-state solver
 * whose diagonal, sigma = H c and Davidson / CG iterations run resident in
 * the HBM of one AMD Instinct MI355X.
 *
 * Part 1 is the drop-in surface: the same names, signatures, argument
 * meaning and error behaviour (print to stderr, exit(EXIT_FAILURE)) as the
 * reference program klgunst/doci.  Part 2 is this library's own
 * device-resident API (doci_*), which returns error codes instead of exiting.
 */
#ifndef DOCI_AMD_H
#define DOCI_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================== */
/* Part 1: drop-in entry points                                             */
/* ======================================================================== */

/* Replaces include/read_fcidump.h:4 (src/read_fcidump.c:18-82).  *one_p_int
 * (L*L) and *two_p_int (L^4) are malloc'd; the caller frees them. */
void read_fcidump(char *filename, unsigned int *NORB, unsigned int *NALPHA,
                  unsigned int *NBETA, double *core_energy, double **one_p_int,
                  double **two_p_int);

/* Replaces include/doci.h:14 (src/doci.c:69-134).  Reads the FCIDUMP, builds
 * the integral and addressing tables on the GPU and computes the diagonal
 * there.  Limits: Na == Nb, L <= 64 (64-bit masks), C(L,N) < 2^31. */
void fill_doci_data(char *dumpfil, void **dat);

/* Replaces include/doci.h:12 (src/doci.c:136-140): host copy of the diagonal,
 * owned by dat (freed by doci_cleanup_data). */
double *doci_get_diagonal(void *dat);

/* Replaces include/doci.h:16 (src/doci.c:142-146). */
int doci_get_basis_size(void *dat);

/* Replaces include/doci.h:18 (src/doci.c:148-159); sets *dat = NULL. */
void doci_cleanup_data(void **dat);

/* Replaces include/doci.h:31 (src/doci.c:164-238): result = H vec for host
 * vectors of doci_get_basis_size(dat) doubles (copied through the GPU). */
void doci_matvec(double *vec, double *result, void *dat);

/* Replaces include/wrapper_solvers.h:36-38 (src/wrapper_solvers.c:10-24).
 * solver "D" (Davidson), "CG", "CGP" (CG with 1/diag preconditioner);
 * returns -1 for an unknown solver, Davidson 0/1 (converged / max_its hit),
 * CG and CGP 1.  When matvec == doci_matvec and dat came from
 * fill_doci_data the whole solve runs resident on the GPU and only vec and
 * *energy cross back; any other matvec is called on host vectors while the
 * solver's vector algebra runs on the GPU. */
int sparse_eigensolve(double *vec, int dims, double *energy,
                      void (*matvec)(double *, double *, void *), double *diagonal,
                      double *M, double tol, int max_its, char solver[],
                      int davidson_keep, int davidson_max_vec, void *dat);

/* Replaces include/davidson.h:25-27 (src/davidson.c:33-128). */
int davidson(double *result, double *energy, int max_vectors, int keep_deflate,
             double davidson_tol, void (*matvec)(double *, double *, void *),
             double *diagonal, int basis_size, int max_its, void *vdat);

/* Replaces include/CG.h:4-5 (src/CG.c:13-164): (matvec, M, vec, size, energy,
 * tol, max_its, precond, dat).  M must be NULL. */
int ray_q(void (*matvec)(double *, double *, void *), double *M, double *vec,
          int size, double *energy, double cg_tol, int max_its, double *precond,
          void *vdat);

/* ======================================================================== */
/* Part 2: device-resident API                                              */
/* ======================================================================== */

enum {
    DOCI_OK = 0,
    DOCI_E_IO = -1,          /* file cannot be opened */
    DOCI_E_FORMAT = -2,      /* FCIDUMP header malformed */
    DOCI_E_LINE = -3,        /* FCIDUMP integral line malformed (see errline) */
    DOCI_E_SPIN = -4,        /* Na != Nb */
    DOCI_E_ORBITALS = -5,    /* L > 64 */
    DOCI_E_DIM = -6,         /* C(L,N) >= 2^31 or N = 0 */
    DOCI_E_HIP = -7,         /* HIP runtime error / no GPU */
    DOCI_E_NOMEM = -8,       /* host or device allocation failed */
    DOCI_E_ARG = -9,         /* invalid argument */
    DOCI_E_SOLVER = -10,     /* unknown solver name */
    DOCI_E_BREAKDOWN = -11   /* subspace breakdown (zero or non-finite norm) */
};

typedef struct doci_op doci_op;

/* Last error message of the calling thread (static storage). */
const char *doci_last_error(void);

/* Operator from an FCIDUMP file or from raw integrals in the reference's
 * layout (one[(j-1)L+(i-1)] for "v i j 0 0", two[(l-1)L^3+(k-1)L^2+(j-1)L+(i-1)]
 * for "v i j k l"), which are transformed exactly as the reference does.
 * *errline receives the offending line for DOCI_E_LINE (may be NULL). */
int doci_op_from_fcidump(const char *path, doci_op **op, int *errline);
int doci_op_from_integrals(unsigned L, unsigned N, double core, const double *one,
                           const double *two, doci_op **op);
void doci_op_free(doci_op *op);
int64_t doci_op_dim(const doci_op *op);
unsigned doci_op_L(const doci_op *op);
unsigned doci_op_N(const doci_op *op);
/* Copies the diagonal (bit-identical to the reference's) into host memory. */
int doci_op_diagonal(doci_op *op, double *out);
/* Host-buffer sigma: out = H in, FP64. */
int doci_op_sigma(doci_op *op, const double *in, double *out);
/* Device-buffer sigma on the operator's stream.  prec 0: FP64 vectors;
 * prec 1: FP32 vectors with FP64 sums.  out = (H - shift) in. */
int doci_op_sigma_dev(doci_op *op, const void *in, void *out, int prec, double shift);
/* doci_op_sigma_dev without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * runs the star-kernel sigma (kernel kind 0) with its slot buffer in a
 * static device array of the library, 8,388,608 FP64 entries (64 MiB) per
 * GPU, held until the process ends; problems whose slot buffer is larger
 * (doci_op_star_split's zslots, e.g. L=24 / N=8 and the reference's test
 * inputs fit, L=34 / N=10 does not) are rejected with DOCI_E_ARG, as is an
 * operator whose star tables are not built yet (doci_op_set_kernel(op, 0) or
 * a first doci_op_sigma_dev builds them).  Every call on one GPU shares that
 * array: calls that overlap on one GPU race. */
int doci_op_sigma_dev_noalloc(doci_op *op, const void *in, void *out, int prec, double shift);
/* The operator's HIP stream (hipStream_t) and device. */
void *doci_op_stream(doci_op *op);

/* Sigma kernel of the operator: 0 = two-pass star kernel (default from
 * 2^20 determinants on; it sums a centre's hops with LDS atomics, so results
 * may differ between runs in the last bits), 1 = one lane per determinant
 * (default below 2^20 determinants), 2 = star kernel with every computed
 * index bounds-checked (violations counted, see doci_op_check_violations,
 * and the access skipped), 3 = one-pass slab kernel (bit-reproducible),
 * 4 = slab kernel bounds-checked.  The star and slab decomposition tables
 * are built on first use. */
int doci_op_set_kernel(doci_op *op, int kind);
/* The operator's current sigma kernel (kinds as above). */
int doci_op_kernel(const doci_op *op);
/* The star split: a low / b high orbitals, main-pass and star-pass
 * workgroups, and the FP64 entries of the slot buffer (slots, then the
 * packed K of the column star pass). */
int doci_op_star_split(doci_op *op, unsigned *a, unsigned *b, unsigned *ngroups,
                       unsigned *nstargroups, size_t *zslots);
/* The slab kernel's split: a low orbitals, b high orbitals, workgroups. */
int doci_op_slab_split(doci_op *op, unsigned *a, unsigned *b, unsigned *ngroups);
/* Index violations counted by kernel kinds 2 and 4 since the operator was built. */
int64_t doci_op_check_violations(doci_op *op);
/* Tuning knobs.  Slab kernel: "xcd_runs" (0: groups dispatched in rank
 * order, 1: contiguous runs per XCD), "emax" (determinants staged per
 * workgroup, rebuilds the slab split), "gmax" (slabs per workgroup).  Star
 * kernel: "star_pass" (0, the default: column pass with wave-uniform K; 1:
 * per-star vector products; 2: per-star products on the FP64 matrix cores;
 * 3: the column pass on the FP32 matrix cores for FP32 vectors, 4-term FP32
 * partial sums);
 * and (each rebuilds the star tables; 0 restores the default):
 * "star_lds" (LDS bytes of a star-pass workgroup), "star_sgmax" (stars per
 * star-pass workgroup), "star_emax" (determinants staged per main-pass
 * workgroup). */
int doci_op_tune(doci_op *op, const char *key, int value);
/* Average event-timed duration (ms) of `reps` sigma launches on scratch
 * vectors of precision prec (0 FP64, 1 FP32). */
int doci_op_time_sigma(doci_op *op, int prec, int reps, double *ms);

/* Plain device copy of n doubles through this library's streaming kernel
 * (synchronous); used to calibrate hardware byte counters. */
int doci_dev_copy(void *dst, const void *src, int64_t n);

/* Solver settings; doci_solve_defaults gives the reference's defaults
 * (include/options.h:4-7; src/executable.c:212-218). */
typedef struct {
    char solver[8];        /* "D", "CG", "CGP" */
    double tol;            /* 1e-8 */
    int max_its;           /* 200 */
    int davidson_keep;     /* 2 */
    int davidson_max_vec;  /* 30 */
    int precision;         /* 0 = FP64, 1 = mixed: FP32 storage of the subspace (Davidson)
                              or of x, p, Ax, Ap, g (CG) with FP64 sums and dots, then an
                              FP64 finish (FP64 sigma, residual and Rayleigh quotient) */
    double switch_tol;     /* mixed Davidson: residual at which FP64 takes over (1e-5) */
    int trace;             /* print "its INFO residue energy" lines */
} doci_solve_opts;

typedef struct {
    int iterations;        /* total iterations (all phases) */
    int iterations_fp32;   /* of which in FP32 storage */
    int sigma_calls;
    int converged;
    double residual;       /* final residual norm (Davidson) or |2r|^2 (CG) */
    double energy;
    double seconds;        /* wall time inside the solve */
    double seconds_sigma;  /* of which in sigma kernels (event-timed) */
    int sigma_calls_fp32;  /* sigma launches on FP32 vectors */
    double seconds_sigma_fp32; /* their event-timed total */
    double d_energy;       /* energy change of the last iteration (the reference's
                              "d_energy" of its summary line) */
} doci_solve_info;

void doci_solve_defaults(doci_solve_opts *o);

/* Ground state.  guess: host vector of dim doubles, or NULL for the HF guess
 * (unit vector at the first minimum of the diagonal).  x_out (may be NULL)
 * receives the unit eigenvector.  Returns the reference's solver code
 * (0/1 for Davidson, 1 for CG) or a negative DOCI_E_* error. */
int doci_solve(doci_op *op, const doci_solve_opts *o, const double *guess,
               double *x_out, double *energy, doci_solve_info *info);

/* ---- FCIDUMP reader/writer and synthetic integrals ---------------------- */

/* The DOCI tables of raw integrals one[L*L], two[L^4] in read_fcidump's
 * layout (host only, no GPU): K[p*L+q] = (pq|pq) with K[p*L+p] = 0,
 * J[p*L+q] = (pp|qq), X[p*L+q] = (pq|qp), h[p] = h_pp - 1/2 sum_j (pj|jp),
 * each bit-identical to transform_integrals (src/doci.c:630-665) followed by
 * init_Viikk (src/doci.c:667-683) and the diagonal's lookups
 * (src/doci.c:600-620), read in O(L^2) without unpacking two[]. */
int doci_integral_tables(unsigned L, const double *one, const double *two, double *K, double *J,
                         double *X, double *h);

/* read_fcidump's parser returning DOCI_E_IO / DOCI_E_FORMAT / DOCI_E_LINE
 * (with *errline) instead of exiting; free the tables with doci_free. */
int doci_read_fcidump(const char *path, unsigned *L, unsigned *Na, unsigned *Nb,
                      double *core, double **one, double **two, int *errline);
void doci_free(void *p);

/* Writes the canonical FCIDUMP (i>=j, k>=l, ij>=kl; "%.17e"; core line last;
 * no trailing blank line; entries with |v| < 1e-12 dropped). */
int doci_write_fcidump(const char *path, unsigned L, unsigned N, double core,
                       const double *one, const double *two);

/* Seeded synthetic integrals in the raw layout above (canonical entries
 * only, |v| < 1e-12 dropped): h_ii = -12 + 11 i/(L-1) + N(0,0.05),
 * h_ij ~ N(0, 0.1 e^{-|i-j|/4}), (ij|kl) = sum_{P=1..2L} B^P_ij B^P_kl with
 * B^P_ij ~ N(0, 0.3 e^{-|i-j|/3}/sqrt(P)); core = 10.  splitmix64 -> 53-bit
 * uniform -> Box-Muller, drawn in the order h_ii, h_ij (i<j), B^P (i<=j). */
int doci_gen_integrals(unsigned L, uint64_t seed, double *core, double *one,
                       double *two);

/* N(0,1) vector of n entries from the same generator, normalised in FP64. */
int doci_gen_vector(int64_t n, uint64_t seed, double *out);

/* ---- input file (src/executable.c:201-268) ------------------------------ */

typedef struct {
    char fcidump[255];
    char solver[20];
    int max_its;
    double tol;
    int davidson_keep;
    int davidson_max_vec;
    int hf_init;
    int precision;         /* DOCI_PRECISION = MIXED -> 1 (a key of this library) */
    char eigvec[255];      /* EIGENVECTOR_FILE = path (a key of this library): the CLI
                              writes the eigenvector there with doci_write_vector */
} doci_input;

/* Reads the reference's key = value input file with its defaults. */
int doci_read_inputfile(const char *path, doci_input *in);

/* Eigenvector file (no counterpart in the reference, which keeps the vector
 * in memory): a text header line "DOCI_VECTOR L N n\n" followed by n FP64
 * values, little-endian binary, in the determinant order of the basis
 * (lexicographic order of the sorted occupied orbitals, last fastest). */
int doci_write_vector(const char *path, unsigned L, unsigned N, int64_t n, const double *x);
/* Reads such a file; *x (malloc'd, free with doci_free) receives n values. */
int doci_read_vector(const char *path, unsigned *L, unsigned *N, int64_t *n, double **x);

/* ---- many independent solves on one GPU ----------------------------------
 * No counterpart in the reference (one solve per process there).  A batch
 * holds DOCI instances of any (L, N) that iterate Davidson in lock step:
 * every iteration is one fixed sequence of launches covering all running
 * instances (captured in a HIP graph), each instance advancing its own state
 * on the device.  Instances can be added while others run (serving): they
 * are admitted at the next iteration boundary, as long as the vector memory
 * of the running instances stays within `mem_budget`.  Results per instance
 * equal those of an independent solve within the solver tolerances. */
typedef struct {
    double tol;            /* 1e-8 */
    int max_its;           /* 200 */
    int keep;              /* 2 (at most 4) */
    int max_vec;           /* 30 (at most 30) */
    double switch_tol;     /* mixed instances: FP32 phase until this residual (1e-5) */
    int use_graphs;        /* capture each iteration's launches in a HIP graph (1) */
    int tile_rows;         /* rows per workgroup of the batched launches, rounded up to a
                              multiple of 512; 0 = autotune */
    size_t mem_budget;     /* bytes of device memory for running instances (0 = 90% free) */
    int max_instances;     /* capacity (1024) */
    int keep_vectors;      /* 1: keep a finished instance's eigenvector until
                              doci_batch_release (counted in the budget); 0: free it
                              at completion (energies only) (1) */
    size_t mem_reserve;    /* device bytes taken for the instance arena at creation
                              (0: grown on demand in 8 GiB chunks) */
} doci_batch_opts;

typedef struct doci_batch doci_batch;

typedef struct {
    int status;            /* 0 waiting, 1 running, 2 converged, 3 max_its, 4 breakdown */
    int iterations;        /* total (both phases) */
    int iterations_fp32;
    double energy;
    double residual;
    double t_submit, t_start, t_done; /* seconds on the batch's clock */
} doci_batch_result;

void doci_batch_defaults(doci_batch_opts *o);
int doci_batch_create(const doci_batch_opts *o, doci_batch **b);
void doci_batch_free(doci_batch *b);
/* Device bytes an instance of (L, N) holds while it runs. */
size_t doci_batch_footprint(const doci_batch_opts *o, unsigned L, unsigned N);
/* Queue an instance (raw integrals as in doci_op_from_integrals; copied).
 * precision 0 = FP64, 1 = mixed.  *id identifies it in later calls. */
int doci_batch_add(doci_batch *b, unsigned L, unsigned N, double core, const double *one,
                   const double *two, int precision, int *id);
/* The same for count instances at once (instance k: L[k], N[k], core[k],
 * one[k], two[k]); the host transform of their integrals runs on several
 * threads.  ids[k] receives instance k's id; nothing is queued on error. */
int doci_batch_add_many(doci_batch *b, int count, const unsigned *L, const unsigned *N,
                        const double *core, const double *const *one, const double *const *two,
                        int precision, int *ids);
/* Admit what fits, run one Davidson iteration of every running instance,
 * retire finished ones.  *running / *waiting (may be NULL) receive the
 * counts afterwards. */
int doci_batch_step(doci_batch *b, int *running, int *waiting);
/* Step until every queued instance has finished. */
int doci_batch_run(doci_batch *b);
/* Status of an instance; x (may be NULL) receives its unit eigenvector
 * (C(L,N) doubles) once finished. */
int doci_batch_result_of(doci_batch *b, int id, doci_batch_result *r, double *x);
/* Free a finished instance's device memory (its result stays readable). */
int doci_batch_release(doci_batch *b, int id);
/* Seconds on the batch's clock (the clock of t_submit / t_start / t_done). */
double doci_batch_clock(const doci_batch *b);
/* Rows per workgroup in use (after autotuning), iterations run, graph launches. */
int doci_batch_stats(const doci_batch *b, int *tile_rows, int *iterations, int *graph_launches);

#ifdef __cplusplus
}
#endif
#endif
