/* davidson.c -- device-resident Davidson (src/davidson.c:33-255).
 *
 * Semantics kept from the reference: the loop test `res > tol && its++ <
 * max_its`, the |D - theta| preconditioner with its 1e-12 floor, deflation
 * right after the subspace reaches max_vec keeping the `keep` lowest Ritz
 * vectors, and the returned code its >= max_its.
 *
 * MI355X form: the subspace V and its images W = (H - shift) V live in HBM;
 * each iteration is three streaming passes (sigma, one new column of V^T W,
 * fused Ritz/residual/preconditioner) and only O(m) scalars cross to the
 * host.  Vectors are stored unnormalised with a scale per vector
 * (V_j = s_j Vt_j), which removes the normalisation pass.  The reference
 * orthogonalises each correction against V (modified Gram-Schmidt); here a
 * correction joins the basis as it is and the Rayleigh-Ritz step solves
 * S y = theta G y with the Gram matrix G = V^T V measured in FP64 (the same
 * subspace, so the same Ritz pairs to rounding); a correction nearly
 * dependent on V is projected out explicitly first.  G also absorbs the
 * rounding of FP32 storage.
 *
 * Mixed precision (precision = 1): the subspace is stored in FP32 with FP64
 * sums and the operator applied as H - shift with shift = min(D) (so stored
 * W entries are small and FP32 rounding of W stays below the switch
 * tolerance); once the residual drops below switch_tol the solve restarts in
 * FP64 from the current Ritz vector and finishes with FP64 residuals and
 * Rayleigh quotients.  The FP32 phase also hands over when its residual has
 * stopped falling (not halved over the last four iterations): FP32 storage
 * bounds the residual it can reach (a few times 1e-8 on the reference's
 * inputs), and a switch_tol below that would otherwise spend every remaining
 * iteration in FP32.  The restart keeps the FP32 phase's last correction
 * (its preconditioned residual, as an FP64 vector) when that phase met
 * switch_tol: the first FP64 iteration computes the FP64 sigma of the Ritz
 * vector and then adds that correction to the basis, instead of recomputing
 * the residual of the same vector.  The FP64 phase always runs after an FP32
 * phase, also when the FP32 residual is already below tol (switch_tol <= tol,
 * or one iteration past both): an FP32 residual describes the rounded
 * operator, so convergence is only reported from an FP64 sigma, residual and
 * Rayleigh quotient.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#define DK_MAX_STALL 1024 /* FP32 phase iterations watched for a stalled residual */

typedef struct {
    int prec;
    size_t es;
    int nslot;
    void **Vs, **Ws; /* slot storage */
    void *block;
    int *basis;      /* slot index of V_j */
    int *freelist;
    int nfree;
    int pooled;      /* block belongs to the operator's persistent pool */
} dav_ws;

static void ws_free(dav_ws *w)
{
    if (w->block && !w->pooled) dh_dfree(w->block);
    free(w->Vs);
    free(w->Ws);
    free(w->basis);
    free(w->freelist);
    memset(w, 0, sizeof *w);
}

static int ws_alloc(dav_ws *w, dh_ctx *c, int64_t n, int prec, int nslot)
{
    memset(w, 0, sizeof *w);
    w->prec = prec;
    w->es = prec ? sizeof(float) : sizeof(double);
    w->nslot = nslot;
    const size_t vb = ((size_t)n * w->es + 255) & ~(size_t)255;
    w->Vs = calloc((size_t)nslot, sizeof(void *));
    w->Ws = calloc((size_t)nslot, sizeof(void *));
    w->basis = calloc((size_t)nslot, sizeof(int));
    w->freelist = calloc((size_t)nslot, sizeof(int));
    if (!w->Vs || !w->Ws || !w->basis || !w->freelist) {
        ws_free(w);
        return doci_fail(DOCI_E_NOMEM, "Davidson host workspace");
    }
    const size_t need = vb * 2 * (size_t)nslot;
    hipError_t e = hipSuccess;
    if (c->pool && *c->pool && *c->pool_bytes >= need) {
        w->block = *c->pool; /* reuse the operator's subspace memory */
        w->pooled = 1;
    } else if (c->pool) {
        dh_dfree(*c->pool);
        *c->pool = NULL;
        *c->pool_bytes = 0;
        e = dh_dmalloc(&w->block, need);
        if (e == hipSuccess) {
            *c->pool = w->block;
            *c->pool_bytes = need;
            w->pooled = 1;
        }
    } else {
        e = dh_dmalloc(&w->block, need);
    }
    if (e != hipSuccess) {
        w->block = NULL;
        ws_free(w);
        return doci_fail(DOCI_E_NOMEM, "Davidson subspace of %d x 2 vectors of %lld x %zu B: %s",
                         nslot, (long long)n, w->es, hipGetErrorString(e));
    }
    for (int k = 0; k < nslot; ++k) {
        w->Vs[k] = (char *)w->block + vb * (size_t)k;
        w->Ws[k] = (char *)w->block + vb * (size_t)(nslot + k);
        w->freelist[k] = nslot - 1 - k;
    }
    w->nfree = nslot;
    return DOCI_OK;
}

static int take_slot(dav_ws *w) { return w->freelist[--w->nfree]; }
static void give_slot(dav_ws *w, int k) { w->freelist[w->nfree++] = k; }

typedef struct {
    dh_ctx *c;
    const doci_solve_opts *o;
    doci_solve_info *info;
    int *its;
    hipEvent_t ev0, ev1;
} dav_run;

static int timed_apply(dav_run *r, const void *in, void *out, int prec, double shift)
{
    HIPCHK(hipEventRecord(r->ev0, r->c->stream));
    int rc = r->c->apply(r->c, in, out, prec, shift);
    if (rc) return rc;
    HIPCHK(hipEventRecord(r->ev1, r->c->stream));
    r->info->sigma_calls++;
    if (prec) r->info->sigma_calls_fp32++;
    return DOCI_OK;
}

static int add_sigma_time(dav_run *r, int prec)
{
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, r->ev0, r->ev1));
    r->info->seconds_sigma += 1e-3 * ms;
    if (prec) r->info->seconds_sigma_fp32 += 1e-3 * ms;
    return DOCI_OK;
}

/* One Davidson phase at storage precision `prec`; stops at residual <=
 * stop_tol.  d_x: FP64 guess in, FP64 Ritz vector out.  t_in (FP64, may be
 * NULL): the correction of the guess, already known (the preconditioned
 * residual of the previous phase), used instead of recomputing it, with
 * res_in its residual norm.  t_out (FP64, may be NULL): receives the
 * correction that the next iteration would have used. */
static int dav_phase(dav_run *r, int prec, double shift, double stop_tol, double *d_x,
                     const double *t_in, double res_in, double *t_out, double *energy,
                     double *res_out)
{
    dh_ctx *c = r->c;
    const doci_solve_opts *o = r->o;
    const int mv = o->davidson_max_vec, keep = o->davidson_keep, ld = mv;
    const int64_t n = c->n;
    const double dep_tol = 0.05;
    dav_ws w;
    int rc = ws_alloc(&w, c, n, prec, mv + keep);
    if (rc) return rc;
    double *S = calloc((size_t)mv * mv, sizeof(double)), *G = calloc((size_t)mv * mv, sizeof(double));
    double *Y = calloc((size_t)mv * mv, sizeof(double)), *th = calloc((size_t)mv, sizeof(double));
    double *s = calloc((size_t)mv + 1, sizeof(double)), *a = calloc((size_t)mv + 2, sizeof(double));
    double *red = calloc(2 * (size_t)mv + 8, sizeof(double)), *g = calloc((size_t)mv + 2, sizeof(double));
    void **Vp = calloc((size_t)mv + 1, sizeof(void *)), **Wp = calloc((size_t)mv + 1, sizeof(void *));
    /* P: the projected matrix of the shifted operator in the stored vectors,
     * P[i + j ld] = <V_i, W_j> (i <= j, unscaled); gam: the coefficients the
     * projection subtracted from t (none unless t was nearly dependent on V); bw: <W_j, t> from the residual */
    double *P = calloc((size_t)mv * mv, sizeof(double)), *gam = calloc((size_t)mv + 2, sizeof(double));
    double *bw = calloc((size_t)mv + 2, sizeof(double));
    double *bv = calloc((size_t)mv + 2, sizeof(double)), *cv = calloc((size_t)mv + 2, sizeof(double));
    int have_b = 0;
    if (!S || !G || !Y || !th || !s || !a || !red || !g || !Vp || !Wp || !P || !gam || !bw || !bv ||
        !cv) {
        rc = doci_fail(DOCI_E_NOMEM, "Davidson host arrays");
        goto done;
    }
#define CK(expr)                       \
    do {                               \
        rc = (expr);                   \
        if (rc) goto done;             \
    } while (0)
#define DK(expr)                                                                            \
    do {                                                                                    \
        int e_ = (expr);                                                                    \
        if (e_) {                                                                           \
            rc = doci_fail(DOCI_E_HIP, "%s: %s", #expr, hipGetErrorString((hipError_t)e_)); \
            goto done;                                                                      \
        }                                                                                   \
    } while (0)

    int m = 0;
    int tslot = take_slot(&w);
    double tt;
    /* initial vector: t = guess, its norm (davidson.c:68 then :144-146) */
    DK(dk_convert(d_x, 0, w.Vs[tslot], prec, n, 1.0, c->stream));
    DK(dk_multidot(0, NULL, w.Vs[tslot], n, prec, c->d_red, c->d_out, c->stream));
    CK(dh_fetch(c, 1, red));
    tt = red[0];
    double res = stop_tol * 10, theta = 0; /* davidson.c:45: start above the tolerance */
    const int start_its = *r->its;
    double hist[DK_MAX_STALL]; /* residuals of this phase's iterations */
    while (res > stop_tol && (*r->its)++ < o->max_its) {
        /* the new direction t (davidson.c:133-176 orthogonalises it): it joins
         * the basis as it is, the Rayleigh-Ritz step using the measured Gram
         * matrix G of the stored vectors, unless it is nearly dependent on V
         * (its part orthogonal to V below dep_tol of its norm squared); then
         * its projection onto V, G^-1 <V, t>, is subtracted (at most twice).
         * Consecutive corrections stay far from dependent (Gram condition
         * numbers of a few units over a whole solve), so in practice no
         * iteration pays a pass over V for this. */
        for (int j = 0; j < m; ++j) {
            Vp[j] = w.Vs[w.basis[j]];
            gam[j] = 0;
        }
        for (int pass = 0; m > 0 && pass < 2; ++pass) {
            for (int j = 0; j < m; ++j) bv[j] = s[j] * a[j];
            const double pn = dh_gram_solve(m, G, ld, bv, cv);
            if (pn < 0) {
                rc = doci_fail(DOCI_E_BREAKDOWN, "Davidson: subspace Gram matrix not positive definite");
                goto done;
            }
            if (!(tt > 0) || tt - pn > dep_tol * tt) break;
            for (int j = 0; j < m; ++j) {
                g[j] = s[j] * cv[j];
                gam[j] += g[j];
            }
            DK(dk_ortho(m, Vp, g, w.Vs[tslot], n, prec, c->d_red, c->d_out, c->stream));
            CK(dh_fetch(c, m + 1, red));
            tt = red[0];
            for (int j = 0; j < m; ++j) a[j] = red[1 + j];
        }
        if (!(tt > 0) || !isfinite(tt)) {
            rc = doci_fail(DOCI_E_BREAKDOWN, "Davidson: new direction has zero or invalid norm");
            goto done;
        }
        s[m] = 1.0 / sqrt(tt);
        for (int j = 0; j < m; ++j) G[j + (size_t)m * ld] = s[j] * s[m] * a[j];
        G[m + (size_t)m * ld] = s[m] * s[m] * tt;
        w.basis[m] = tslot;
        Vp[m] = w.Vs[tslot];
        /* W_m = (H - shift) V_m */
        CK(timed_apply(r, w.Vs[tslot], w.Ws[tslot], prec, shift));
        /* new column of V^T W (davidson.c:178-193).  With the residual's
         * <W_j, t> at hand, the entries j < m follow from the symmetry of H:
         * <V_j, W_m> = <W_j, V_m> = <W_j, t> - sum_i gam_i <W_j, V_i>, and
         * only <V_m, W_m> is streamed (two vectors instead of m + 2). */
        if (have_b) {
            void *Vm[1] = {Vp[m]};
            DK(dk_multidot(1, Vm, w.Ws[tslot], n, prec, c->d_red, c->d_out, c->stream));
            CK(dh_fetch(c, 2, red));
            CK(add_sigma_time(r, prec));
            for (int j = 0; j < m; ++j) {
                double v = bw[j];
                for (int i = 0; i < m; ++i)
                    v -= gam[i] * (i <= j ? P[i + (size_t)j * ld] : P[j + (size_t)i * ld]);
                P[j + (size_t)m * ld] = v;
            }
            P[m + (size_t)m * ld] = red[1];
        } else {
            DK(dk_multidot(m + 1, Vp, w.Ws[tslot], n, prec, c->d_red, c->d_out, c->stream));
            CK(dh_fetch(c, m + 2, red));
            CK(add_sigma_time(r, prec));
            for (int i = 0; i <= m; ++i) P[i + (size_t)m * ld] = red[1 + i];
        }
        have_b = 0;
        for (int i = 0; i <= m; ++i)
            S[i + (size_t)m * ld] = s[i] * s[m] * P[i + (size_t)m * ld] + shift * G[i + (size_t)m * ld];
        ++m;
        CK(dh_gen_eig(m, S, G, ld, th, Y, ld));
        if (m == mv) {
            /* deflation (davidson.c:229-255): keep the `keep` lowest Ritz vectors */
            int nb[DK_MAX_VEC];
            for (int j = 0; j < m; ++j) {
                Vp[j] = w.Vs[w.basis[j]];
                Wp[j] = w.Ws[w.basis[j]];
            }
            for (int k = 0; k < keep; ++k) {
                nb[k] = take_slot(&w);
                for (int j = 0; j < m; ++j) g[j] = s[j] * Y[j + (size_t)k * ld];
                DK(dk_lincomb(m, Vp, g, w.Vs[nb[k]], n, prec, prec, c->stream));
                DK(dk_lincomb(m, Wp, g, w.Ws[nb[k]], n, prec, prec, c->stream));
            }
            for (int j = 0; j < m; ++j) give_slot(&w, w.basis[j]);
            m = keep;
            for (int k = 0; k < keep; ++k) {
                w.basis[k] = nb[k];
                s[k] = 1.0;
                Vp[k] = w.Vs[nb[k]];
            }
            for (int k = 0; k < keep; ++k) {
                DK(dk_multidot(k + 1, Vp, w.Vs[nb[k]], n, prec, c->d_red, c->d_out, c->stream));
                CK(dh_fetch(c, k + 2, red));
                G[k + (size_t)k * ld] = red[0];
                for (int i = 0; i < k; ++i) G[i + (size_t)k * ld] = red[1 + i];
                DK(dk_multidot(k + 1, Vp, w.Ws[nb[k]], n, prec, c->d_red, c->d_out, c->stream));
                CK(dh_fetch(c, k + 2, red));
                for (int i = 0; i <= k; ++i) {
                    P[i + (size_t)k * ld] = red[1 + i];
                    S[i + (size_t)k * ld] = red[1 + i] + shift * G[i + (size_t)k * ld];
                }
            }
            CK(dh_gen_eig(m, S, G, ld, th, Y, ld));
        }
        theta = th[0];
        /* residual, preconditioner and the dots of the next correction with V */
        for (int j = 0; j < m; ++j) {
            Vp[j] = w.Vs[w.basis[j]];
            Wp[j] = w.Ws[w.basis[j]];
            g[j] = s[j] * Y[j];
        }
        tslot = take_slot(&w);
        if (t_in && m == 1 && *r->its == start_its + 1) {
            /* the guess's correction is given: its overlaps with V (same
             * subspace as recomputing it) */
            DK(dk_convert(t_in, 0, w.Vs[tslot], prec, n, 1.0, c->stream));
            DK(dk_multidot(m, Vp, w.Vs[tslot], n, prec, c->d_red, c->d_out, c->stream));
            CK(dh_fetch(c, m + 1, red));
            res = res_in;
            tt = red[0];
            for (int j = 0; j < m; ++j) a[j] = red[1 + j];
        } else {
            /* FP64 storage only: with FP32 storage the two ways of forming the
             * column differ by the rounding of the stored W (~1e-7 relative),
             * which moves the FP32 phase's iterates; its explicit dots are
             * half the bytes anyway */
            const int wd = prec == 0 && m <= DK_RESIDUAL_WDOTS_MAX;
            DK(dk_residual(m, Vp, Wp, g, theta, shift, c->d_D, w.Vs[tslot], NULL, wd, n, prec,
                           c->d_red, c->d_out, c->stream));
            CK(dh_fetch(c, wd ? 2 + 2 * m : m + 2, red));
            res = sqrt(red[0]);
            tt = red[1];
            for (int j = 0; j < m; ++j) a[j] = red[2 + j];
            for (int j = 0; wd && j < m; ++j) bw[j] = red[2 + m + j];
            have_b = wd;
        }
        r->info->d_energy = *energy - theta; /* davidson.c:92-93 */
        *energy = theta;
        if (prec) r->info->iterations_fp32++;
        if (o->trace) printf("%d\t%d\t%e\t%f\n", *r->its, 0, res, theta);
        /* FP32 storage: hand over to FP64 once the residual has stopped
         * falling (not halved over the last four iterations) */
        const int k = *r->its - start_its - 1;
        if (prec && k < DK_MAX_STALL) hist[k] = res;
        if (prec && k >= 4 && k < DK_MAX_STALL && res > 0.5 * hist[k - 4]) break;
    }
    if (t_out && m > 0 && res <= stop_tol) DK(dk_convert(w.Vs[tslot], prec, t_out, 0, n, 1.0, c->stream));
    /* Ritz vector x = V y (davidson.c:205) */
    if (m > 0) {
        for (int j = 0; j < m; ++j) {
            Vp[j] = w.Vs[w.basis[j]];
            g[j] = s[j] * Y[j];
        }
        DK(dk_lincomb(m, Vp, g, d_x, n, prec, 0, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    (void)start_its;
    *res_out = res;
    rc = DOCI_OK;
done:
#undef CK
#undef DK
    ws_free(&w);
    free(S);
    free(G);
    free(Y);
    free(th);
    free(s);
    free(a);
    free(red);
    free(g);
    free(Vp);
    free(Wp);
    free(P);
    free(gam);
    free(bw);
    free(bv);
    free(cv);
    return rc;
}

int dh_diag_min(dh_ctx *c, double *dmin)
{
    int64_t idx = 0;
    int e = dk_argmin(c->d_D, c->n, c->d_red, c->d_idx, c->stream);
    if (!e) e = hipMemcpyAsync(&idx, c->d_idx, sizeof idx, hipMemcpyDeviceToHost, c->stream);
    if (!e) e = hipStreamSynchronize(c->stream);
    if (!e) e = hipMemcpy(dmin, c->d_D + idx, sizeof *dmin, hipMemcpyDeviceToHost);
    return e ? doci_fail(DOCI_E_HIP, "argmin of the diagonal: %s", hipGetErrorString((hipError_t)e))
             : DOCI_OK;
}

int dh_davidson(dh_ctx *c, const doci_solve_opts *o, double *d_x, double *energy,
                doci_solve_info *info)
{
    if (o->davidson_max_vec < 2 || o->davidson_max_vec > DK_MAX_VEC || o->davidson_keep < 1 ||
        o->davidson_keep >= o->davidson_max_vec)
        return doci_fail(DOCI_E_ARG, "need 1 <= DAVIDSON_KEEP < DAVIDSON_MAX_VEC <= %d", DK_MAX_VEC);
    const double t0 = dh_now();
    int its = 0, rc;
    double res = 0, res32 = 0;
    double *d_t = NULL; /* FP64 correction handed from the FP32 phase to the FP64 phase */
    int t_ok = 0;
    dav_run r = {c, o, info, &its, NULL, NULL};
    HIPCHK(hipEventCreate(&r.ev0));
    HIPCHK(hipEventCreate(&r.ev1));
    *energy = 0;
    const int mixed = o->precision == 1 && c->fp32_ok;
    int ran64 = 0; /* an FP64 phase produced the final residual */
    if (mixed) {
        double shift = 0;
        rc = dh_diag_min(c, &shift);
        if (rc) goto out;
        const double sw = o->switch_tol > o->tol ? o->switch_tol : o->tol;
        /* the FP32 phase's last correction seeds the FP64 phase: its first
         * iteration computes the FP64 sigma of the Ritz vector and goes on
         * with that correction instead of a second residual of the same x */
        if (dh_dmalloc((void **)&d_t, sizeof(double) * (size_t)c->n) != hipSuccess) d_t = NULL;
        rc = dav_phase(&r, 1, shift, sw, d_x, NULL, 0.0, d_t, energy, &res);
        if (rc) goto out;
        res32 = res;
        t_ok = d_t && res <= sw; /* the phase wrote its last correction */
    }
    {
        /* FP64 phase (the whole solve when not mixed).  After an FP32 phase
         * that stopped at switch_tol it goes on with the handed-over
         * correction; after one that already reached tol it starts from the
         * Ritz vector alone, whose FP64 residual its first iteration checks. */
        const int seeded = t_ok && res32 > o->tol;
        if (its < o->max_its || its == 0) {
            rc = dav_phase(&r, 0, 0.0, o->tol, d_x, seeded ? d_t : NULL, res32, NULL, energy, &res);
            if (rc) goto out;
            ran64 = 1;
        }
    }
    rc = its >= o->max_its;
out:
    if (d_t) {
        hipStreamSynchronize(c->stream);
        dh_dfree(d_t);
    }
    hipEventDestroy(r.ev0);
    hipEventDestroy(r.ev1);
    info->iterations = its;
    info->residual = res;
    info->energy = *energy;
    info->converged = ran64 && res <= o->tol;
    info->seconds = dh_now() - t0;
    return rc;
}
