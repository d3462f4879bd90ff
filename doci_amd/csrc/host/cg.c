/* cg.c -- device-resident Rayleigh-quotient CG (src/CG.c:13-164).
 *
 * Kept from the reference: x taken as given (assumed unit), g = 2(Ax - rho x)
 * with convergence on |g|^2 (before preconditioning), the optional 1/D
 * preconditioner with its 1e-12 floor, the Fletcher-Reeves ratio of the
 * preconditioned norms, the 2 x 2 generalized Rayleigh-Ritz on [x p] solved
 * with dsygv's conventions, the recurrence Ax <- z0 Ax + z1 Ap (no fresh
 * matvec), and the return value 1.  Each iteration is one sigma and three
 * fused streaming passes (direction, 2 x 4 Gram, update + gradient).
 *
 * Mixed precision (precision = 1): x, p, Ax, Ap and g are stored in FP32
 * (every dot product, the Gram matrix and the 2 x 2 problem in FP64) and the
 * operator is applied as H - shift with shift = min(D), so that the stored
 * Ax entries are small (the iteration is shift-invariant: rho moves by the
 * shift, g and the 2 x 2 eigenvectors do not).  When the FP32 iteration
 * meets the stopping test (or max_its), the solve finishes in FP64: x is
 * normalised in FP64, its FP64 sigma refreshes Ax, and rho, g and |g|^2 are
 * recomputed; if that FP64 |g|^2 is still above tol, FP64 iterations follow
 * (restarted with p = -g) until it is not.  Convergence and the returned
 * energy always come from FP64 quantities.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "internal.h"

typedef struct {
    void *x, *p, *Ax, *Ap, *g; /* vectors of the phase's precision */
    int prec;
} cg_vecs;

int dh_ray_q(dh_ctx *c, const doci_solve_opts *o, int precond, double *d_x, double *energy,
             doci_solve_info *info)
{
    const double t0 = dh_now();
    const int64_t n = c->n;
    const int mixed = o->precision == 1 && c->fp32_ok;
    const size_t vb = ((size_t)n * sizeof(double) + 255) & ~(size_t)255;
    const size_t vb32 = ((size_t)n * sizeof(float) + 255) & ~(size_t)255;
    char *blk = NULL, *blk32 = NULL;
    int rc = DOCI_OK;
    double shift = 0;
    hipEvent_t e0 = NULL, e1 = NULL;
    if (mixed) {
        rc = dh_diag_min(c, &shift);
        if (rc) return rc;
    }
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    hipError_t he = dh_dmalloc((void **)&blk, vb * (precond ? 5 : 4));
    if (he == hipSuccess && mixed) he = dh_dmalloc((void **)&blk32, vb32 * 5);
    if (he != hipSuccess) {
        dh_dfree(blk);
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return doci_fail(DOCI_E_NOMEM, "CG vectors: %s", hipGetErrorString(he));
    }
    double *pinv = precond ? (double *)(blk + 4 * vb) : NULL;
    cg_vecs v64 = {d_x, blk, blk + vb, blk + 2 * vb, blk + 3 * vb, 0};
    cg_vecs v32 = {blk32, blk32 + vb32, blk32 + 2 * vb32, blk32 + 3 * vb32, blk32 + 4 * vb32, 1};
    double out[8], rho = 0, normq = 0, ng = 0, ngp = 0, d_energy = 0;
    int its = 0, ran64 = 0;
    float ms;
#define DK(expr)                                                                            \
    do {                                                                                    \
        int e_ = (expr);                                                                    \
        if (e_) {                                                                           \
            rc = doci_fail(DOCI_E_HIP, "%s: %s", #expr, hipGetErrorString((hipError_t)e_)); \
            goto done;                                                                      \
        }                                                                                   \
    } while (0)
#define CK(expr)           \
    do {                   \
        rc = (expr);       \
        if (rc) goto done; \
    } while (0)
#define SIGMA(in, outv, prec, sh)                                \
    do {                                                         \
        DK(hipEventRecord(e0, c->stream));                       \
        CK(c->apply(c, in, outv, prec, sh));                     \
        DK(hipEventRecord(e1, c->stream));                       \
        info->sigma_calls++;                                     \
        if (prec) info->sigma_calls_fp32++;                      \
    } while (0)
#define SIGMA_TIME(prec)                                         \
    do {                                                         \
        DK(hipEventElapsedTime(&ms, e0, e1));                    \
        info->seconds_sigma += 1e-3 * ms;                        \
        if (prec) info->seconds_sigma_fp32 += 1e-3 * ms;         \
    } while (0)

    if (pinv) DK(dk_precond_inv(c->d_D, pinv, n, c->stream));
    if (o->trace) printf("\nIT\tINFO\tRESIDUE\tENERGY\n");
    for (int phase = mixed ? 0 : 1; phase < 2; ++phase) {
        /* phase 0: FP32 storage of (H - shift); phase 1: FP64 (the finish) */
        const cg_vecs *v = phase ? &v64 : &v32;
        const int prec = v->prec;
        const double sh = phase ? 0.0 : shift;
        if (phase == 0) {
            DK(dk_convert(d_x, 0, v32.x, 1, n, 1.0, c->stream));
        } else if (mixed) {
            /* x from the FP32 iteration, normalised in FP64 */
            void *X[1] = {v32.x};
            DK(dk_multidot(0, X, v32.x, n, 1, c->d_red, c->d_out, c->stream));
            CK(dh_fetch(c, 1, out));
            if (!(out[0] > 0) || !isfinite(out[0])) {
                rc = doci_fail(DOCI_E_BREAKDOWN, "CG: FP32 iterate has zero or invalid norm");
                goto done;
            }
            DK(dk_convert(v32.x, 1, d_x, 0, n, 1.0 / sqrt(out[0]), c->stream));
        }
        SIGMA(v->x, v->Ax, prec, sh);
        {
            void *X[1] = {v->x};
            DK(dk_multidot(1, X, v->Ax, n, prec, c->d_red, c->d_out, c->stream));
            CK(dh_fetch(c, 2, out));
            SIGMA_TIME(prec);
            if (phase == 1 && mixed) d_energy = rho - out[1]; /* rho: FP32 phase's, of H */
            rho = out[1]; /* of H - sh */
        }
        DK(dk_cg_grad(v->x, v->Ax, rho, pinv, v->g, n, prec, c->d_red, c->d_out, c->stream));
        CK(dh_fetch(c, 2, out));
        normq = out[0];
        ng = ngp = out[1];
        if (o->trace) printf("%d\t%d\t%e\t%f\n", its, 0, ng, rho + sh);
        ran64 = phase == 1;
        int first = 1; /* p = -g on the phase's first step (CG.c:96-98) */
        while (normq > o->tol && its < o->max_its) {
            DK(dk_cg_dir(v->p, v->g, first ? 0.0 : ng / ngp, first, n, prec, c->stream));
            first = 0;
            SIGMA(v->p, v->Ap, prec, sh);
            DK(dk_cg_gram(v->x, v->p, v->Ax, v->Ap, n, prec, c->d_red, c->d_out, c->stream));
            CK(dh_fetch(c, 7, out));
            SIGMA_TIME(prec);
            const double A2[4] = {out[0], out[1], out[2], out[3]};
            const double B2[4] = {out[4], out[5], out[5], out[6]};
            double lam[2], Z[4];
            CK(dh_sygv2(A2, B2, lam, Z));
            d_energy = rho - lam[0];
            rho = lam[0];
            DK(dk_cg_step(v->x, v->p, v->Ax, v->Ap, Z[0], Z[1], rho, pinv, v->g, n, prec,
                          c->d_red, c->d_out, c->stream));
            CK(dh_fetch(c, 2, out));
            normq = out[0];
            ngp = ng;
            ng = out[1];
            if (o->trace) printf("%d\t%d\t%e\t%f\n", its, 0, ng, rho + sh);
            ++its;
            if (!phase) info->iterations_fp32++;
        }
        if (!phase) rho += shift; /* energies from here on are of H */
    }
    const double el = dh_now() - t0;
    if (o->trace) {
        if (its <= o->max_its)
            printf("CG converged in %d iterations with %e residue and d_energy %e.\n", its, ng,
                   d_energy);
        else
            printf("CG didn't converge within %d iterations! d_energy is %e.\n", o->max_its,
                   d_energy);
        printf("elapsed time : %lf sec\n", el);
        printf("average time per matvec : %lf sec\n", el / (its + 1));
    }
    *energy = rho;
    info->iterations = its;
    info->residual = normq;
    info->energy = rho;
    info->d_energy = d_energy;
    info->converged = ran64 && normq <= o->tol;
    info->seconds = el;
    rc = 1;
done:
#undef DK
#undef CK
#undef SIGMA
#undef SIGMA_TIME
    hipStreamSynchronize(c->stream);
    dh_dfree(blk);
    dh_dfree(blk32);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}
