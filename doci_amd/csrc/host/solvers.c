/* solvers.c -- solver dispatch (src/wrapper_solvers.c:10-24) and the
 * drop-in davidson / ray_q entry points, plus doci_solve. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

void doci_solve_defaults(doci_solve_opts *o)
{
    memset(o, 0, sizeof *o);
    strcpy(o->solver, "D");
    o->tol = 1e-8;
    o->max_its = 200;
    o->davidson_keep = 2;
    o->davidson_max_vec = 30;
    o->precision = 0;
    o->switch_tol = 1e-5;
    o->trace = 0;
}

static int solver_kind(const char *s)
{
    if (!strcmp(s, "D")) return 0;
    if (!strcmp(s, "CG")) return 1;
    if (!strcmp(s, "CGP")) return 2;
    return -1;
}

/* Runs the chosen solver on a context whose d_x holds the guess. */
static int run_solver(dh_ctx *c, const doci_solve_opts *o, double *d_x, double *energy,
                      doci_solve_info *info)
{
    switch (solver_kind(o->solver)) {
    case 0: return dh_davidson(c, o, d_x, energy, info);
    case 1: return dh_ray_q(c, o, 0, d_x, energy, info);
    case 2: return dh_ray_q(c, o, 1, d_x, energy, info);
    default: return doci_fail(DOCI_E_SOLVER, "no correct solver inputted!");
    }
}

/* Places the HF guess (unit vector at the first minimum of D) in d_x. */
static int hf_guess_dev(dh_ctx *c, double *d_x)
{
    int64_t idx = 0;
    const double one = 1.0;
    DKCHK(dk_argmin(c->d_D, c->n, c->d_red, c->d_idx, c->stream));
    HIPCHK(hipMemcpyAsync(&idx, c->d_idx, sizeof idx, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemsetAsync(d_x, 0, sizeof(double) * c->n, c->stream));
    HIPCHK(hipMemcpyAsync(d_x + idx, &one, sizeof one, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DOCI_OK;
}

int doci_solve(doci_op *op, const doci_solve_opts *o, const double *guess, double *x_out,
               double *energy, doci_solve_info *info)
{
    doci_solve_info local;
    if (!info) info = &local;
    memset(info, 0, sizeof *info);
    if (!dh_is_op(op) || !o || !energy) return doci_fail(DOCI_E_ARG, "doci_solve arguments");
    if (solver_kind(o->solver) < 0) return doci_fail(DOCI_E_SOLVER, "no correct solver inputted!");
    dh_ctx c;
    dh_ctx_of_op(op, &c);
    double *d_x = NULL;
    HIPCHK(dh_dmalloc((void **)&d_x, sizeof(double) * op->n));
    int rc;
    if (guess) {
        hipError_t e = hipMemcpyAsync(d_x, guess, sizeof(double) * op->n, hipMemcpyHostToDevice,
                                      c.stream);
        rc = e == hipSuccess ? DOCI_OK : doci_fail(DOCI_E_HIP, "guess upload: %s", hipGetErrorString(e));
    } else {
        rc = hf_guess_dev(&c, d_x);
    }
    if (rc == DOCI_OK) rc = run_solver(&c, o, d_x, energy, info);
    if (rc >= 0 && x_out) {
        hipError_t e = hipMemcpyAsync(x_out, d_x, sizeof(double) * op->n, hipMemcpyDeviceToHost,
                                      c.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
        if (e != hipSuccess) rc = doci_fail(DOCI_E_HIP, "result download: %s", hipGetErrorString(e));
    }
    hipStreamSynchronize(c.stream);
    dh_dfree(d_x);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* generic operator: a host matvec callback                                 */

typedef struct {
    void (*matvec)(double *, double *, void *);
    void *dat;
    double *h_in, *h_out;
} cb_user;

static int apply_cb(dh_ctx *c, const void *in, void *out, int prec, double shift)
{
    cb_user *u = c->user;
    if (prec != 0) return doci_fail(DOCI_E_ARG, "host matvec supports FP64 only");
    const size_t b = sizeof(double) * (size_t)c->n;
    HIPCHK(hipMemcpyAsync(u->h_in, in, b, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    u->matvec(u->h_in, u->h_out, u->dat);
    if (shift != 0)
        for (int64_t i = 0; i < c->n; ++i) u->h_out[i] -= shift * u->h_in[i];
    HIPCHK(hipMemcpyAsync(out, u->h_out, b, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DOCI_OK;
}

/* Solve with either the resident DOCI operator or a host callback.  vec is
 * the guess on entry and the eigenvector on exit (host). */
static int solve_dropin(double *vec, int dims, double *energy,
                        void (*matvec)(double *, double *, void *), double *diagonal,
                        const doci_solve_opts *o, void *dat, doci_solve_info *info)
{
    memset(info, 0, sizeof *info);
    const int kind = solver_kind(o->solver);
    if (kind < 0) return -1;
    dh_ctx c;
    cb_user u = {matvec, dat, NULL, NULL};
    double *d_Dtmp = NULL, *d_x = NULL;
    int own_ctx = 0, rc = DOCI_OK;
    if (matvec == doci_matvec && dh_is_op(dat) && ((doci_op *)dat)->n == dims) {
        doci_op *op = dat;
        dh_ctx_of_op(op, &c);
        if (diagonal && diagonal != op->D_host) { /* caller-supplied diagonal */
            HIPCHK(dh_dmalloc((void **)&d_Dtmp, sizeof(double) * dims));
            HIPCHK(hipMemcpy(d_Dtmp, diagonal, sizeof(double) * dims, hipMemcpyHostToDevice));
            c.d_D = d_Dtmp;
        }
    } else {
        memset(&c, 0, sizeof c);
        own_ctx = 1;
        c.n = dims;
        c.apply = apply_cb;
        c.user = &u;
        HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        HIPCHK(dh_dmalloc((void **)&c.d_red, sizeof(double) * dk_red_scratch_doubles(DK_MAX_VEC + 8)));
        HIPCHK(dh_dmalloc((void **)&c.d_out, sizeof(double) * (DK_MAX_VEC + 8)));
        HIPCHK(dh_dmalloc((void **)&c.d_idx, sizeof(int64_t)));
        HIPCHK(hipHostMalloc((void **)&c.h_out, sizeof(double) * (DK_MAX_VEC + 8), 0));
        HIPCHK(hipHostMalloc((void **)&u.h_in, sizeof(double) * dims, 0));
        HIPCHK(hipHostMalloc((void **)&u.h_out, sizeof(double) * dims, 0));
        HIPCHK(dh_dmalloc((void **)&d_Dtmp, sizeof(double) * dims));
        if (diagonal)
            HIPCHK(hipMemcpy(d_Dtmp, diagonal, sizeof(double) * dims, hipMemcpyHostToDevice));
        else
            HIPCHK(hipMemset(d_Dtmp, 0, sizeof(double) * dims));
        c.d_D = d_Dtmp;
    }
    HIPCHK(dh_dmalloc((void **)&d_x, sizeof(double) * dims));
    HIPCHK(hipMemcpy(d_x, vec, sizeof(double) * dims, hipMemcpyHostToDevice));
    rc = run_solver(&c, o, d_x, energy, info);
    if (rc >= 0) HIPCHK(hipMemcpy(vec, d_x, sizeof(double) * dims, hipMemcpyDeviceToHost));
    hipStreamSynchronize(c.stream);
    dh_dfree(d_x);
    dh_dfree(d_Dtmp);
    if (own_ctx) {
        dh_dfree(c.d_red);
        dh_dfree(c.d_out);
        dh_dfree(c.d_idx);
        hipHostFree(c.h_out);
        hipHostFree(u.h_in);
        hipHostFree(u.h_out);
        hipStreamDestroy(c.stream);
    }
    return rc;
}

static void dropin_opts(doci_solve_opts *o, const char *solver, double tol, int max_its, int keep,
                        int max_vec)
{
    doci_solve_defaults(o);
    snprintf(o->solver, sizeof o->solver, "%s", solver);
    o->tol = tol;
    o->max_its = max_its;
    o->davidson_keep = keep;
    o->davidson_max_vec = max_vec;
    const char *p = getenv("DOCI_PRECISION");
    if (p && (!strcmp(p, "mixed") || !strcmp(p, "MIXED") || !strcmp(p, "1"))) o->precision = 1;
    const char *t = getenv("DOCI_TRACE");
    o->trace = t && *t && strcmp(t, "0");
}

static int dropin_fail(int rc)
{
    fprintf(stderr, "%s\n", doci_last_error());
    exit(EXIT_FAILURE);
    return rc;
}

int davidson(double *result, double *energy, int max_vectors, int keep_deflate,
             double davidson_tol, void (*matvec)(double *, double *, void *), double *diagonal,
             int basis_size, int max_its, void *vdat)
{
    doci_solve_opts o;
    doci_solve_info info;
    dropin_opts(&o, "D", davidson_tol, max_its, keep_deflate, max_vectors);
    if (o.trace) printf("IT\tINFO\tRESIDUE\tENERGY\n");
    const int rc = solve_dropin(result, basis_size, energy, matvec, diagonal, &o, vdat, &info);
    if (rc < 0) return dropin_fail(rc);
    if (o.trace) {
        /* davidson.c:101-105: the loop counter ends one past max_its when
         * the limit stops the loop (info.iterations is that counter) */
        const int its = info.iterations;
        if (its <= max_its)
            printf("Davidson converged in %d iterations with %e residue and d_energy %e.\n", its,
                   info.residual, info.d_energy);
        else
            printf("Davidson didn't converge within %d iterations! d_energy is %e.\n", max_its,
                   info.d_energy);
        printf("elapsed time : %lf sec\n", info.seconds);
        printf("average time per matvec : %lf sec\n",
               info.iterations ? info.seconds / info.iterations : 0.0);
    } else {
        printf("\t\t ** DAVIDSON ITERATIONS : %d\n", info.iterations);
    }
    return rc;
}

int ray_q(void (*matvec)(double *, double *, void *), double *M, double *vec, int size,
          double *energy, double cg_tol, int max_its, double *precond, void *vdat)
{
    if (M != NULL) {
        fprintf(stderr, "%s:%d @ %s :: Wrong M!\n", __FILE__, __LINE__, __func__);
        exit(EXIT_FAILURE);
    }
    doci_solve_opts o;
    doci_solve_info info;
    dropin_opts(&o, precond ? "CGP" : "CG", cg_tol, max_its, 2, 30);
    o.trace = 1; /* the reference's CG always prints its trace (CG.c:70-72,88,127,130-139) */
    const int rc = solve_dropin(vec, size, energy, matvec, precond, &o, vdat, &info);
    return rc < 0 ? dropin_fail(rc) : rc;
}

int sparse_eigensolve(double *vec, int dims, double *energy,
                      void (*matvec)(double *, double *, void *), double *diagonal, double *M,
                      double tol, int max_its, char solver[], int davidson_keep,
                      int davidson_max_vec, void *dat)
{
    switch (solver_kind(solver)) {
    case 0:
        return davidson(vec, energy, davidson_max_vec, davidson_keep, tol, matvec, diagonal, dims,
                        max_its, dat);
    case 1: return ray_q(matvec, M, vec, dims, energy, tol, max_its, NULL, dat);
    case 2: return ray_q(matvec, M, vec, dims, energy, tol, max_its, diagonal, dat);
    default: return -1;
    }
}
