/*
 * dk.h -- the thin C ABI between the C host code and the HIP kernels.
 *
 * Every function takes device pointers, sizes and a stream, launches one or
 * more kernels on that stream and returns a hipError_t value (0 = success).
 * No function here synchronises unless its comment says so.
 */
#ifndef DOCI_DK_H
#define DOCI_DK_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DK_MAX_L 64
#define DK_MAX_VEC 64 /* most vectors one fused Davidson kernel addresses */

/* Device-resident integral and addressing tables of one DOCI operator. */
typedef struct {
    uint32_t L, N;
    uint32_t n;              /* C(L,N) < 2^31 */
    double core;
    const uint32_t *binom;   /* binom[x*(N+2)+j] = C(x,j) mod 2^32, x <= L */
    const double *K;         /* K[o*L+t] = (ot|ot), zero diagonal (off-diagonal H) */
    const double *J;         /* J[p*L+q] = (pp|qq) */
    const double *X;         /* X[p*L+q] = (pq|qp) */
    const double *h;         /* h[p] = h'_pp */
} dk_tables;

/* Diagonal, bit-identical to the reference's addition order. */
int dk_diagonal(const dk_tables *t, double *D, hipStream_t s);

/* sigma = (H - shift) c.  prec: 0 = FP64 vectors, 1 = FP32 vectors (FP64
 * sums).  shift = 0 gives H c exactly. */
int dk_sigma(const dk_tables *t, const double *D, double shift, const void *c,
             void *sigma, int prec, hipStream_t s);

/* Slab decomposition for the slab-blocked sigma: B = the top b orbitals, A
 * the other a = L - b.  Slab i holds the C(b, m) determinants whose A-part is
 * slab_amask[i], starting at rank slab_start[i].  Group g = slabs
 * [group[2g], group[2g] + (group[2g+1] & 0xffff)) which all have
 * m = group[2g+1] >> 16 pairs in B and together at most emax determinants
 * (at most gmax slabs).  aamax/abmax/bamax bound the per-slab neighbour
 * tables: max k(a-k), max k and max a-k over the groups' k = N - m. */
typedef struct {
    uint32_t a, b, ngroup, emax, gmax, aamax, abmax, bamax;
    uint32_t xcd_runs;       /* 1: deal contiguous runs of groups to each XCD; 0: in order */
    const uint32_t *slab_start;
    const uint64_t *slab_amask;
    const uint32_t *group;
} dk_slabs;

/* Dynamic LDS bytes the slab kernel needs (prec as for dk_sigma). */
size_t dk_sigma_slab_lds(const dk_tables *t, const dk_slabs *sl, int prec);
/* sigma = (H - shift) c with the slab-blocked kernel. */
int dk_sigma_slab(const dk_tables *t, const dk_slabs *sl, const double *D, double shift,
                  const void *c, void *sigma, int prec, hipStream_t s);
/* The same with every computed index bounds-checked: violations are counted
 * in *bad (device) and the access skipped. */
int dk_sigma_slab_checked(const dk_tables *t, const dk_slabs *sl, const double *D, double shift,
                          const void *c, void *sigma, int prec, uint32_t *bad, hipStream_t s);

/* Star decomposition for the two-pass sigma (B = top b orbitals, A the
 * other a).  A slab is the set of C(b, m) determinants sharing an A-part
 * alpha (k = N - m pairs in A).  The star of rho (rho in A, j pairs) is the
 * centre slab rho (m = N - j pairs in B; absent when N - j > b) with the
 * member slabs rho + y, y an empty A orbital.
 * Star pass: one workgroup stages whole stars in LDS and writes, into the
 * slot buffer Z, for every member rho + x the sum of its hops that empty x
 * (to another A orbital: dense K_Y products; to a B orbital: the centre), and
 * for the centre the sum of its hops from B into A (the members).
 * Main pass: sigma = (D - shift) c + hops inside the slab + sum of the slab's
 * slots.  Slab alpha (k pairs in A) has nslot = k + [m >= 1 && k < a] slots
 * of C(b, m) values starting at slab_zbase; member slot = position of x in
 * alpha, centre slot = k.  zoff[k] = first Z index of the k-pair sector, whose
 * slabs are stored in lexicographic order of alpha. */
typedef struct {
    uint32_t a, b;
    uint32_t ngroup;            /* main-pass groups */
    uint32_t emax;              /* determinants staged per main-pass group */
    uint32_t gslabs;            /* slabs per main-pass group, at most */
    uint32_t owns;              /* LDS doubles of a main-pass group's staged slabs (padded layout) */
    uint32_t nsgroup;           /* star-pass groups */
    uint32_t sgmax;             /* stars per star group, at most */
    uint32_t selems;            /* LDS elements staged per star group, at most */
    uint32_t ymax;              /* member-table entries per star (= a) */
    uint32_t pass;              /* star pass: 0 column pass (wave-uniform K), 1 per-star vector,
                                   2 per-star FP64 matrix cores, 3 FP32 column pass on the matrix cores,
                                   4 column pass with 64-bit addresses (0 uses 32-bit buffer offsets
                                   where the vector and every member sector of Z fit them) */
    uint32_t zsec_max;          /* largest member sector of Z, zoff[k+1] - zoff[k] (elements) */
    const uint32_t *slab_start; /* per slab (order: k, then lexicographic alpha) */
    const uint32_t *slab_zbase;
    const uint32_t *group;      /* main groups: first slab, nslab | m << 16 */
    const uint64_t *star_amask; /* stars (order: j, then lexicographic rho) */
    const uint32_t *sgroup;     /* star groups: first star, nstar | j << 16 */
    const uint32_t *zoff;       /* [a + 2]: first Z index of each k-sector */
    const uint32_t *mtab;       /* per star and orbital y < a, pairs: first rank and Z position
                                   of the slot of member slab rho + y (0, 0 for y in rho) */
    uint32_t jgroup[DK_MAX_L + 1]; /* star groups of j: [jgroup[j], jgroup[j+1]), j < N */
    /* Tables over the masks beta of m pairs in B (b <= 16), r = lexB(beta):
     * bmask[bmask_off[m] + r] = beta;
     * hop[hop_off[m] + h * C(b,m) + r] = J | 8 (o * b + t) << 16 for the h-th
     *   hop beta - o + t (o in beta ascending, then t not in beta ascending),
     *   J = its lexB rank; rows h from m(b-m) up to the next multiple of 8
     *   are padding, r | (b * b) << 16 (index b * b names a zero K entry);
     * ins[ins_off[m] + t * C(b,m) + r] = lexB rank of beta + t among the
     *   (m+1)-pair masks, 0xffff when t is in beta (m < b). */
    const uint16_t *bmask;
    const uint32_t *hop;
    const uint16_t *ins;
    uint32_t bmask_off[17], hop_off[17], ins_off[17];
} dk_stars;

/* Doubles of packed K behind the slots (0 when the column pass does not apply). */
size_t dk_star_kp_doubles(uint32_t a, uint32_t b);
/* LDS bytes of the star pass and of the main pass (prec as for dk_sigma). */
size_t dk_sigma_star_lds(const dk_tables *t, const dk_stars *st, int prec, int pass);
/* sigma = (H - shift) c by the star pass then the main pass.  Z: slot buffer
 * of zoff[a+1] elements of the vector precision, followed (at byte offset
 * 8 * zoff[a+1]) by dk_star_kp_doubles(a, b) doubles of packed K that the
 * column star pass writes and reads.  bad: NULL, or a device counter for the
 * bounds-checked instantiation. */
int dk_sigma_star(const dk_tables *t, const dk_stars *st, const double *D, double shift,
                  const void *c, void *sigma, void *Z, int prec, uint32_t *bad, hipStream_t s);

/* Operands of a sigma launch read from device memory when the launch runs
 * (so a recorded launch follows operands chosen on the device). */
typedef struct {
    const void *c;
    void *sigma;
    double shift;
    int32_t active;          /* 0: the launch does nothing */
    int32_t pad;
} dk_sigma_io;
/* dk_sigma_star with c, sigma and shift taken from *io. */
/* The star sigma's slot buffer for callers that cannot allocate: a
 * file-scope device array of dk_static_slot_doubles() FP64 entries, one per
 * device; dk_static_slot_buffer gives the current device's copy's address
 * (looked up on the first call per device). */
size_t dk_static_slot_doubles(void);
int dk_static_slot_buffer(void **addr);
int dk_sigma_star_io(const dk_tables *t, const dk_stars *st, const double *D,
                     const dk_sigma_io *io, void *Z, int prec, hipStream_t s);

/* ---- batched multi-instance Davidson -----------------------------------
 * Many independent DOCI instances iterate in lock step; every launch covers
 * all of them.  Streaming launches run one workgroup per tile (a row block
 * of one instance, listed in `tiles`; each instance has its own tile size,
 * so that large instances use few, long tiles) and write one row of
 * DK_BPV partial sums per tile; "small" launches run one 64-lane workgroup
 * per instance, reduce that instance's tile partials in a fixed order and
 * advance its Davidson state (dk_bstate, device resident).  Each instance's
 * vectors live in slots of `slot_stride` bytes: V slots [0, P) and W slots
 * [P, 2P), P = max_vec + keep. */
#define DK_BMAX_VEC 32
#define DK_BKEEP_MAX 4
#define DK_BPV (DK_BMAX_VEC + 2)

typedef struct {             /* constant per instance */
    uint32_t L, N, n;
    uint32_t tile0, ntiles;  /* its tiles; tile k writes partial-sum row tile0 + k */
    uint32_t tile_rows;      /* rows per tile (ntiles = ceil(n / tile_rows)) */
    uint32_t lane_sigma;     /* 1: sigma by the batched lane kernel */
    uint32_t pad;
    const uint32_t *binom;   /* (L+1) x (N+2) */
    const double *K;         /* L x L, zero diagonal */
    const double *D;         /* diagonal, n */
    char *slots;
    uint64_t slot_stride;
    double *x;               /* FP64 Ritz vector, n */
} dk_binst;

enum { DK_B_IDLE = 0, DK_B_ACTIVE = 1, DK_B_CONVERGED = 2, DK_B_MAXITS = 3, DK_B_BREAKDOWN = 4 };
enum { DK_BF_FIRST = 1, DK_BF_ORTHO = 2, DK_BF_DEFLATE = 4 };

typedef struct {             /* summary copied to the host every iteration */
    int32_t status, m, its, flags;
    double energy, res;
} dk_bstat;

typedef struct {             /* mutable per instance */
    dk_bstat st;
    int32_t max_its, keep, max_vec, P;
    int32_t tslot;           /* V slot holding t (the next basis vector) */
    int32_t sin, sout;       /* sigma input (V slot) and output (W slot) */
    int32_t vslot[DK_BMAX_VEC], wslot[DK_BMAX_VEC];
    int32_t dv[DK_BKEEP_MAX], dw[DK_BKEEP_MAX]; /* deflation targets */
    uint64_t vused, wused;   /* slot occupancy bitmasks */
    double tol, otol, shift, theta;  /* theta: lowest eigenvalue of the projected H - shift;
                                        otol: a correction whose part orthogonal to the
                                        basis is below otol of its norm squared is
                                        projected out explicitly */
    double s[DK_BMAX_VEC];   /* T_j = s_j * stored T_j (same for W_j), |T_j| = 1 */
    double g[DK_BMAX_VEC];   /* projection coefficients on stored T (explicit projection) */
    double y[DK_BMAX_VEC];   /* residual coefficients on stored T / W */
    double defl[DK_BKEEP_MAX][DK_BMAX_VEC];
    /* The basis T (raw corrections) = Q R with Q orthonormal (never formed):
     * R upper triangular, R[i*DK_BMAX_VEC+j] for i <= j */
    double R[DK_BMAX_VEC * DK_BMAX_VEC];
    double A[DK_BMAX_VEC * DK_BMAX_VEC]; /* projected matrix Q^T (H - shift) Q in its eigenbasis */
    double Y[DK_BMAX_VEC * DK_BMAX_VEC]; /* eigenvectors in Q, column k = Y[j*DK_BMAX_VEC+k] */
    dk_sigma_io sio;         /* star-kernel sigma operands of this iteration */
} dk_bstate;

enum { DK_BPH_A = 0, DK_BPH_B = 1, DK_BPH_C = 2, DK_BPH_D = 3, DK_BPH_E = 4, DK_BPH_A0 = 5 };
enum { DK_BOP_ORTHO = 0, DK_BOP_ORTHO2 = 1, DK_BOP_SCOL = 2, DK_BOP_DEFLATE = 3,
       DK_BOP_RESID = 4 };

/* Streaming step `op` over all tiles; M = largest basis size it may meet
 * (picks the instantiation).  part: ntile x DK_BPV doubles. */
int dk_batch_stream(int op, int prec, int M, const dk_binst *inst, dk_bstate *state,
                    const uint32_t *tiles, uint32_t ntile, double *part, hipStream_t s);
/* Small step `phase` for the ninst instances listed in ilist (device). */
int dk_batch_small(int phase, const uint32_t *ilist, uint32_t ninst, const dk_binst *inst,
                   dk_bstate *state, const double *part, hipStream_t s);
/* sigma of every active lane-sigma instance: W[sout] = (H - shift) V[sin]. */
int dk_batch_sigma_lane(int prec, uint32_t maxL, uint32_t maxN, const dk_binst *inst,
                        dk_bstate *state, const uint32_t *tiles, uint32_t ntile, hipStream_t s);
/* Admission of many instances at once.  Each admitted instance's tables
 * (the layout of the instance memory: K, J, X (L x L FP64), h (L FP64),
 * binomials ((L+1) x (N+2) u32), tb bytes) sit in a device staging copy at
 * src; dk_batch_admit copies them to dst, computes the diagonal D
 * (bit-identical to dk_diagonal), clears the guess slot t, and per instance
 * sets t = e_argmin(D) (first index of the minimum) and the start state of
 * its first phase (precision prec, shift = min D when shift_min_d, else 0):
 * two launches for all of them.  Rows are covered by
 * tiles of DK_ADMIT_ROWS; the tiles of admitted instance a are
 * [tile0, tile0 + ntiles) of the partial buffer (2 doubles per tile). */
#define DK_ADMIT_ROWS 256
typedef struct {
    dk_tables src;           /* L, N, n, core and the tables in the staging copy */
    void *dst;               /* instance table memory (tb bytes, 8-aligned) */
    uint32_t tb;
    uint32_t tile0, ntiles;
    int32_t prec, shift_min_d, max_its, keep, max_vec;
    double *D;
    void *t;                 /* guess slot (prec 0: FP64, 1: FP32) */
    dk_bstate *state;
    double tol, otol;
} dk_badmit;
/* adm: na descriptors, tiles: (admitted index, tile) pairs, part: 2 doubles
 * per tile; all device memory. */
int dk_batch_admit(const dk_badmit *adm, uint32_t na, const uint32_t *tiles, uint32_t ntile,
                   double *part, uint32_t maxL, uint32_t maxN, hipStream_t s);

/* FP64 restart of many mixed-precision instances at once: the FP32 phase's
 * Ritz vector x (stored FP64) becomes the guess t of the FP64 phase, and the
 * start state is that of an FP64 first phase without shift, continuing the
 * iteration count from its0: one launch for all of them.  tiles: (restart index, tile) pairs of DK_ADMIT_ROWS rows. */
typedef struct {
    const double *x;
    double *t;
    dk_bstate *state;
    uint32_t n, its0;
    int32_t max_its, keep, max_vec, pad;
    double tol, otol;
} dk_brestart;
int dk_batch_restart(const dk_brestart *rs, const uint32_t *tiles, uint32_t ntile,
                     hipStream_t s);

/* ---- streaming vector kernels (deterministic two-level reductions) -------
 * Vectors are arrays of n elements; `vecs` lists device pointers.  prec as
 * above for the stored vectors; reductions and coefficients are FP64.
 * `red` is a device scratch of dk_red_scratch_doubles(nvals) doubles;
 * results land in out_dev (device) in the order documented per function. */
size_t dk_red_scratch_doubles(int nvals);

/* out[0] = <y, y>, out[1+j] = <X_j, y> for j < m. */
int dk_multidot(int m, void *const *X, const void *y, int64_t n, int prec,
                double *red, double *out_dev, hipStream_t s);

/* y <- y - sum_j g_j X_j, then out[0] = <y_new, y_new>, out[1+j] = <X_j, y_new>. */
int dk_ortho(int m, void *const *X, const double *g, void *y, int64_t n,
             int prec, double *red, double *out_dev, hipStream_t s);

/* Davidson residual step (src/davidson.c:195-227 fused).  W holds
 * (H - shift) V, so with the Ritz value theta:
 *   x = sum_j a_j V_j, r = sum_j a_j W_j - (theta - shift) x,
 *   t = r / max-floor(|D - theta|, 1e-12)  written to `t`,
 *   out[0] = <r,r>, out[1] = <t,t>, out[2+j] = <V_j, t> (j < m), and when
 *   wdots (m <= DK_RESIDUAL_WDOTS_MAX) also out[2+m+j] = <W_j, t>.
 * Dots with t use t as stored (in the storage precision).  When x_out is not
 * NULL the Ritz vector x is also written there.  red / out_dev hold
 * 2 + 2m values. */
#define DK_RESIDUAL_WDOTS_MAX 32
int dk_residual(int m, void *const *V, void *const *W, const double *a,
                double theta, double shift, const double *D, void *t, void *x_out, int wdots,
                int64_t n, int prec, double *red, double *out_dev, hipStream_t s);

/* out = sum_j g_j X_j (out may not alias any X_j); out has precision
 * prec_out. */
int dk_lincomb(int m, void *const *X, const double *g, void *out, int64_t n,
               int prec, int prec_out, hipStream_t s);

/* Conversions and simple maps. */
int dk_convert(const void *in, int prec_in, void *out, int prec_out, int64_t n,
               double scale, hipStream_t s);

/* CG (src/CG.c:74-126 fused):
 * dk_cg_grad: g = 2(Ax - rho x); out[0] = <g,g>; if pinv: g *= pinv;
 *             out[1] = <g,g>.
 * dk_cg_dir:  p = q p - g (first: p = -g when first != 0).
 * dk_cg_gram: out = [x.Ax, p.Ax, x.Ap, p.Ap, x.x, p.x, p.p].
 * dk_cg_step: x = z0 x + z1 p; Ax = z0 Ax + z1 Ap; then dk_cg_grad on the
 *             new x, Ax (same outputs). */
int dk_cg_grad(const void *x, const void *Ax, double rho, const double *pinv,
               void *g, int64_t n, int prec, double *red, double *out_dev,
               hipStream_t s);
int dk_cg_dir(void *p, const void *g, double q, int first, int64_t n, int prec,
              hipStream_t s);
int dk_cg_gram(const void *x, const void *p, const void *Ax, const void *Ap,
               int64_t n, int prec, double *red, double *out_dev, hipStream_t s);
int dk_cg_step(void *x, const void *p, void *Ax, const void *Ap, double z0,
               double z1, double rho, const double *pinv, void *g, int64_t n,
               int prec, double *red, double *out_dev, hipStream_t s);
/* pinv = 1/D with the reference's 1e-12 floor (CG.c:57-63). */
int dk_precond_inv(const double *D, double *pinv, int64_t n, hipStream_t s);
/* index of the first minimum of D (executable.c:276-288); result in *idx_dev. */
int dk_argmin(const double *D, int64_t n, double *red, int64_t *idx_dev, hipStream_t s);

#ifdef __cplusplus
}
#endif
#endif
