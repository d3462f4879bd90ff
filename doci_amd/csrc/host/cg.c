/* cg.c -- device-resident Rayleigh-quotient CG (src/CG.c:13-164).
 *
 * Kept from the reference: x taken as given (assumed unit), g = 2(Ax - rho x)
 * with convergence on |g|^2 (before preconditioning), the optional 1/D
 * preconditioner with its 1e-12 floor, the Fletcher-Reeves ratio of the
 * preconditioned norms, the 2 x 2 generalized Rayleigh-Ritz on [x p] solved
 * with dsygv's conventions, the recurrence Ax <- z0 Ax + z1 Ap (no fresh
 * matvec), and the return value 1.  Each iteration is one sigma and three
 * fused streaming passes (direction, 2 x 4 Gram, update + gradient).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "internal.h"

int dh_ray_q(dh_ctx *c, const doci_solve_opts *o, int precond, double *d_x, double *energy,
             doci_solve_info *info)
{
    const double t0 = dh_now();
    const int64_t n = c->n;
    const size_t vb = ((size_t)n * sizeof(double) + 255) & ~(size_t)255;
    char *blk = NULL;
    int rc = DOCI_OK;
    hipEvent_t e0 = NULL, e1 = NULL;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    hipError_t he = dh_dmalloc((void **)&blk, vb * (precond ? 5 : 4));
    if (he != hipSuccess) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return doci_fail(DOCI_E_NOMEM, "CG vectors: %s", hipGetErrorString(he));
    }
    double *p = (double *)blk, *Ax = (double *)(blk + vb), *Ap = (double *)(blk + 2 * vb);
    double *g = (double *)(blk + 3 * vb), *pinv = precond ? (double *)(blk + 4 * vb) : NULL;
    double out[8], rho, normq, ng, ngp, d_energy = 0;
    int its = 0;
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
#define SIGMA(in, outv)                                          \
    do {                                                         \
        DK(hipEventRecord(e0, c->stream));                       \
        CK(c->apply(c, in, outv, 0, 0.0));                       \
        DK(hipEventRecord(e1, c->stream));                       \
        info->sigma_calls++;                                     \
    } while (0)
#define SIGMA_TIME()                                             \
    do {                                                         \
        DK(hipEventElapsedTime(&ms, e0, e1));                    \
        info->seconds_sigma += 1e-3 * ms;                        \
    } while (0)

    if (pinv) DK(dk_precond_inv(c->d_D, pinv, n, c->stream));
    if (o->trace) printf("\nIT\tINFO\tRESIDUE\tENERGY\n");
    SIGMA(d_x, Ax);
    {
        void *X[1] = {d_x};
        DK(dk_multidot(1, X, Ax, n, 0, c->d_red, c->d_out, c->stream));
        CK(dh_fetch(c, 2, out));
        SIGMA_TIME();
        rho = out[1];
    }
    DK(dk_cg_grad(d_x, Ax, rho, pinv, g, n, 0, c->d_red, c->d_out, c->stream));
    CK(dh_fetch(c, 2, out));
    normq = out[0];
    ng = ngp = out[1];
    if (o->trace) printf("%d\t%d\t%e\t%f\n", its, 0, ng, rho);
    while (normq > o->tol && its < o->max_its) {
        DK(dk_cg_dir(p, g, its ? ng / ngp : 0.0, its == 0, n, 0, c->stream));
        SIGMA(p, Ap);
        DK(dk_cg_gram(d_x, p, Ax, Ap, n, 0, c->d_red, c->d_out, c->stream));
        CK(dh_fetch(c, 7, out));
        SIGMA_TIME();
        const double A2[4] = {out[0], out[1], out[2], out[3]};
        const double B2[4] = {out[4], out[5], out[5], out[6]};
        double lam[2], Z[4];
        CK(dh_sygv2(A2, B2, lam, Z));
        d_energy = rho - lam[0];
        rho = lam[0];
        DK(dk_cg_step(d_x, p, Ax, Ap, Z[0], Z[1], rho, pinv, g, n, 0, c->d_red, c->d_out,
                      c->stream));
        CK(dh_fetch(c, 2, out));
        normq = out[0];
        ngp = ng;
        ng = out[1];
        if (o->trace) printf("%d\t%d\t%e\t%f\n", its, 0, ng, rho);
        ++its;
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
    info->converged = normq <= o->tol;
    info->seconds = el;
    rc = 1;
done:
#undef DK
#undef CK
#undef SIGMA
#undef SIGMA_TIME
    hipStreamSynchronize(c->stream);
    dh_dfree(blk);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}
