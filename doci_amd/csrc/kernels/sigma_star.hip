// sigma_star.hip -- sigma = (H - shift) c in two passes (star pass, main pass).
//
// Orbitals split into A = {0..a-1} and B = {a..L-1} (the b fastest-varying
// orbitals of the lexicographic order).  A slab = the C(b, m) determinants
// with one A-part alpha (k = N - m pairs in A); it is contiguous in rank,
// starting at n - C(b,m) - sum_e C(L-1-alpha_e, N-e), in the order lexB(beta).
//
// Star pass.  The star of rho (j pairs in A) is its centre slab rho (m = N-j
// pairs in B, absent when N - j > b) and the a - j member slabs rho + y.  One
// workgroup stages whole stars in LDS.  For member rho + x it sums every hop
// that empties x:  x -> y in A  (sum_y K_xy X_y[beta]: a small dense product
// over the staged members, same offset) and  x -> t in B  (K_xt times the
// centre at lexB(beta + t));  for the centre it sums every hop from B into A
// (K_oy times member y at lexB(beta - o)).  Each sum is one slot of Z.
// Every slab is staged k + 1 times per sigma (k stars as member, one as
// centre) with full-line coalesced loads, instead of being re-read once per
// neighbouring determinant.
//
// Main pass.  sigma = (D - shift) c + hops inside the slab (own slab staged in
// LDS) + the slab's Z slots (coalesced streams).
#include <hip/hip_runtime.h>

#include "../dk.h"
#include "addressing.hpp"

namespace {

constexpr int kBlock = 256;

__host__ __device__ __forceinline__ size_t a8(size_t x) { return (x + 7) & ~(size_t)7; }

// first rank of the slab with A-part `am` and mp pairs in B
__device__ __forceinline__ uint32_t slab_first(uint64_t am, int mp, int L, int N, int b, uint32_t n,
                                               const dk::Binom &C)
{
    uint32_t s = n - C(b, mp);
    int e = 0;
    while (am) {
        const int x = __ffsll((unsigned long long)am) - 1;
        am &= am - 1;
        s -= C(L - 1 - x, N - e);
        ++e;
    }
    return s;
}

// lexicographic index of the k-subset `am` among the k-subsets of A
__device__ __forceinline__ uint32_t lex_a(uint64_t am, int k, int a, const dk::Binom &C)
{
    uint32_t s = 0;
    int e = 0;
    while (am) {
        const int x = __ffsll((unsigned long long)am) - 1;
        am &= am - 1;
        s += C(a - 1 - x, k - e);
        ++e;
    }
    return C(a, k) - 1u - s;
}

__device__ __forceinline__ int nslot_of(int k, int m, int a) { return k + ((m >= 1 && k < a) ? 1 : 0); }

struct StarLds {
    double *K;
    uint32_t *bt;
    void *stage;
    uint32_t *mem_start, *mem_zb, *ctr_start, *ctr_zb;
    uint8_t *ycol, *mpos;
};

template <typename T>
__host__ __device__ size_t star_lds(uint32_t L, uint32_t N, const dk_stars &st, StarLds *p, char *base)
{
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *q = base ? base + off : nullptr;
        off += a8(bytes);
        return q;
    };
    char *K = take(sizeof(double) * L * L);
    char *bt = take(sizeof(uint32_t) * (L + 1) * (N + 2));
    char *stg = take(sizeof(T) * st.selems);
    char *ms = take(sizeof(uint32_t) * st.sgmax * st.ymax);
    char *mz = take(sizeof(uint32_t) * st.sgmax * st.ymax);
    char *cs = take(sizeof(uint32_t) * st.sgmax);
    char *cz = take(sizeof(uint32_t) * st.sgmax);
    char *yc = take(st.sgmax * st.ymax);
    char *mp = take(st.sgmax * st.ymax);
    if (p) {
        p->K = (double *)K;
        p->bt = (uint32_t *)bt;
        p->stage = stg;
        p->mem_start = (uint32_t *)ms;
        p->mem_zb = (uint32_t *)mz;
        p->ctr_start = (uint32_t *)cs;
        p->ctr_zb = (uint32_t *)cz;
        p->ycol = (uint8_t *)yc;
        p->mpos = (uint8_t *)mp;
    }
    return off;
}

template <typename T>
__host__ __device__ size_t main_lds(uint32_t L, uint32_t N, const dk_stars &st)
{
    return a8(sizeof(double) * L * L) + a8(sizeof(uint32_t) * (L + 1) * (N + 2)) +
           a8(sizeof(T) * st.emax);
}

template <typename T, bool kChecked, int YMAX, int BMAX>
__global__ __launch_bounds__(kBlock) void k_star(dk_tables t, dk_stars st, const T *__restrict__ c,
                                                 T *__restrict__ Z, uint32_t *bad)
{
    extern __shared__ double lds_raw[];
    const uint32_t sg = blockIdx.x;
    if (sg >= st.nsgroup) return;
    StarLds p;
    star_lds<T>(t.L, t.N, st, &p, (char *)lds_raw);
    const int L = (int)t.L, N = (int)t.N, a = (int)st.a, b = (int)st.b;
    const uint32_t n = t.n, ztot = st.zoff[a + 1];
    auto ok = [&](uint32_t idx, uint32_t lim) -> bool {
        if (!kChecked) return true;
        if (idx < lim) return true;
        atomicAdd(bad, 1u);
        return false;
    };
    for (int i = threadIdx.x; i < L * L; i += kBlock) p.K[i] = t.K[i];
    for (int i = threadIdx.x; i < (L + 1) * (N + 2); i += kBlock) p.bt[i] = t.binom[i];
    const uint32_t first = st.sgroup[2 * sg], w = st.sgroup[2 * sg + 1];
    const int nst = (int)(w & 0xffffu), j = (int)(w >> 16);
    __syncthreads();
    const dk::Binom C{p.bt, N + 2};
    const int mc = N - j;               // B pairs of the centre
    const bool has_ctr = mc <= b;
    const int mm = mc - 1;              // B pairs of the members
    const int Sm = (int)C(b, mm), Sc = has_ctr ? (int)C(b, mc) : 0;
    const int ny = a - j, stride = ny * Sm + Sc;
    const int km = j + 1;
    const uint32_t zm_sector = st.zoff[km], zc_sector = st.zoff[j];
    const int nsm = nslot_of(km, mm, a), nsc = nslot_of(j, mc, a);
    const uint64_t amask_all = a < 64 ? ((1ull << a) - 1) : ~0ull;

    // member / centre tables
    for (int task = threadIdx.x; task < nst * ny; task += kBlock) {
        const int s = task / ny, yi = task - s * ny;
        const uint64_t rho = st.star_amask[first + s];
        uint64_t z = ~rho & amask_all;
        for (int q = 0; q < yi; ++q) z &= z - 1;
        const int y = __ffsll((unsigned long long)z) - 1;
        const uint64_t mem = rho | (1ull << y);
        p.mem_start[s * ny + yi] = slab_first(mem, mm, L, N, b, n, C);
        p.mem_zb[s * ny + yi] = zm_sector + lex_a(mem, km, a, C) * (uint32_t)(nsm * Sm);
        p.ycol[s * ny + yi] = (uint8_t)y;
        p.mpos[s * ny + yi] = (uint8_t)__popcll(rho & ((1ull << y) - 1));
    }
    if (has_ctr)
        for (int s = threadIdx.x; s < nst; s += kBlock) {
            const uint64_t rho = st.star_amask[first + s];
            p.ctr_start[s] = slab_first(rho, mc, L, N, b, n, C);
            p.ctr_zb[s] = zc_sector + lex_a(rho, j, a, C) * (uint32_t)(nsc * Sc) + (uint32_t)(j * Sc);
        }
    __syncthreads();
    T *stage = (T *)p.stage;
    for (int e = threadIdx.x; e < nst * stride; e += kBlock) {
        const int s = e / stride, r = e - s * stride;
        uint32_t gi;
        if (r < ny * Sm) {
            const int yi = r / Sm;
            gi = p.mem_start[s * ny + yi] + (uint32_t)(r - yi * Sm);
        } else {
            gi = p.ctr_start[s] + (uint32_t)(r - ny * Sm);
        }
        stage[e] = ok(gi, n) ? c[gi] : (T)0;
    }
    __syncthreads();

    const double *K = p.K;
    // members: every hop that empties x, for each member rho + x
    for (int it = threadIdx.x; it < nst * Sm; it += kBlock) {
        const int s = it / Sm, bb = it - s * Sm;
        const T *X = stage + s * stride;
        const T *ctr = X + ny * Sm;
        const uint8_t *yc = p.ycol + s * ny;
        double xv[YMAX];
#pragma unroll
        for (int yi = 0; yi < YMAX; ++yi) xv[yi] = yi < ny ? (double)X[yi * Sm + bb] : 0.0;
        const uint64_t bm = dk::unrank((uint32_t)bb, b, mm, C);
        uint32_t offs[BMAX];
        if (has_ctr) { // lexB of beta + t in the centre (mm + 1 pairs)
            const uint32_t Sb = (uint32_t)Sm - 1u - (uint32_t)bb;
            uint32_t P1 = 0, P2 = 0;
            int jj = 0;
#pragma unroll
            for (int tl = 0; tl < BMAX; ++tl) {
                offs[tl] = 0;
                if (tl < b) {
                    if ((bm >> tl) & 1) {
                        P2 += C(b - 1 - tl, mm + 1 - jj);
                        P1 += C(b - 1 - tl, mm - jj);
                        ++jj;
                    } else {
                        offs[tl] = (uint32_t)Sc - 1u - P2 - C(b - 1 - tl, mm + 1 - jj) - (Sb - P1);
                    }
                }
            }
        }
        for (int xi = 0; xi < ny; ++xi) {
            const double *Kx = K + yc[xi] * L;
            double acc = 0;
#pragma unroll
            for (int yi = 0; yi < YMAX; ++yi)
                if (yi < ny) acc += Kx[yc[yi]] * xv[yi];
            if (has_ctr) {
#pragma unroll
                for (int tl = 0; tl < BMAX; ++tl)
                    if (tl < b && !((bm >> tl) & 1) && ok(offs[tl], (uint32_t)Sc))
                        acc += Kx[a + tl] * (double)ctr[offs[tl]];
            }
            const uint32_t zi = p.mem_zb[s * ny + xi] + (uint32_t)(p.mpos[s * ny + xi] * Sm + bb);
            if (ok(zi, ztot)) Z[zi] = (T)acc;
        }
    }
    // centre: every hop from a B orbital o of beta into an empty A orbital y
    if (has_ctr)
        for (int it = threadIdx.x; it < nst * Sc; it += kBlock) {
            const int s = it / Sc, bb = it - s * Sc;
            const T *X = stage + s * stride;
            const uint8_t *yc = p.ycol + s * ny;
            const uint64_t bm = dk::unrank((uint32_t)bb, b, mc, C);
            const uint32_t Sb = (uint32_t)Sc - 1u - (uint32_t)bb;
            uint32_t Q1 = 0, P1 = 0;
            int h = 0;
            double acc = 0;
            uint64_t rem = bm;
            while (rem) {
                const int o = __ffsll((unsigned long long)rem) - 1;
                rem &= rem - 1;
                const uint32_t off = (uint32_t)Sm - 1u - Q1 - (Sb - P1 - C(b - 1 - o, mc - h));
                const double *Ko = K + (a + o) * L;
                if (ok(off, (uint32_t)Sm))
                    for (int yi = 0; yi < ny; ++yi) acc += Ko[yc[yi]] * (double)X[yi * Sm + off];
                Q1 += C(b - 1 - o, mc - 1 - h);
                P1 += C(b - 1 - o, mc - h);
                ++h;
            }
            const uint32_t zi = p.ctr_zb[s] + (uint32_t)bb;
            if (ok(zi, ztot)) Z[zi] = (T)acc;
        }
}

template <typename T, bool kChecked>
__global__ __launch_bounds__(kBlock) void k_star_main(dk_tables t, dk_stars st,
                                                      const double *__restrict__ D, double shift,
                                                      const T *__restrict__ c,
                                                      const T *__restrict__ Z,
                                                      T *__restrict__ sigma, uint32_t *bad)
{
    extern __shared__ double lds_raw[];
    const uint32_t g = blockIdx.x;
    if (g >= st.ngroup) return;
    const int L = (int)t.L, N = (int)t.N, a = (int)st.a, b = (int)st.b;
    const uint32_t n = t.n, ztot = st.zoff[a + 1];
    auto ok = [&](uint32_t idx, uint32_t lim) -> bool {
        if (!kChecked) return true;
        if (idx < lim) return true;
        atomicAdd(bad, 1u);
        return false;
    };
    double *K = lds_raw;
    uint32_t *bt = (uint32_t *)((char *)lds_raw + a8(sizeof(double) * L * L));
    T *own = (T *)((char *)bt + a8(sizeof(uint32_t) * (L + 1) * (N + 2)));
    for (int i = threadIdx.x; i < L * L; i += kBlock) K[i] = t.K[i];
    for (int i = threadIdx.x; i < (L + 1) * (N + 2); i += kBlock) bt[i] = t.binom[i];
    const uint32_t first = st.group[2 * g], w = st.group[2 * g + 1];
    const int nsl = (int)(w & 0xffffu), m = (int)(w >> 16), k = N - m;
    __syncthreads();
    const dk::Binom C{bt, N + 2};
    const int S = (int)C(b, m), nel = nsl * S, ns = nslot_of(k, m, a);
    for (int e = threadIdx.x; e < nel; e += kBlock) {
        const int s = e / S, bi = e - s * S;
        const uint32_t gi = st.slab_start[first + s] + (uint32_t)bi;
        own[e] = ok(gi, n) ? c[gi] : (T)0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nel; e += kBlock) {
        const int s = e / S, bi = e - s * S;
        const uint32_t I = st.slab_start[first + s] + (uint32_t)bi;
        if (!ok(I, n)) continue;
        const T *os = own + s * S;
        double acc = (D[I] - shift) * (double)os[bi];
        const uint64_t bm = dk::unrank((uint32_t)bi, b, m, C);
        dk::for_each_hop((uint32_t)bi, bm, b, m, C, [&](int o, int tt, uint32_t J) {
            if (ok(J, (uint32_t)S)) acc += K[(a + o) * L + a + tt] * (double)os[J];
        });
        const uint32_t zb = st.slab_zbase[first + s] + (uint32_t)bi;
        double zs = 0;
        for (int q = 0; q < ns; ++q) {
            const uint32_t zi = zb + (uint32_t)(q * S);
            if (ok(zi, ztot)) zs += (double)Z[zi];
        }
        sigma[I] = (T)(acc + zs);
    }
}

template <typename T, bool kChecked>
int launch(const dk_tables *t, const dk_stars *st, const double *D, double shift, const T *c,
           T *sigma, T *Z, uint32_t *bad, hipStream_t s)
{
    const size_t ls = star_lds<T>(t->L, t->N, *st, nullptr, nullptr);
    if (st->nsgroup == 0)
        ; /* a = 0: no stars, every hop is inside the one slab */
    else if (st->b <= 16 && st->ymax <= 32)
        hipLaunchKernelGGL((k_star<T, kChecked, 32, 16>), dim3(st->nsgroup), dim3(kBlock), ls, s,
                           *t, *st, c, Z, bad);
    else if (st->b <= 16 && st->ymax <= 64)
        hipLaunchKernelGGL((k_star<T, kChecked, 64, 16>), dim3(st->nsgroup), dim3(kBlock), ls, s,
                           *t, *st, c, Z, bad);
    else
        return (int)hipErrorInvalidValue;
    int e = (int)hipGetLastError();
    if (e) return e;
    const size_t lm = main_lds<T>(t->L, t->N, *st);
    hipLaunchKernelGGL((k_star_main<T, kChecked>), dim3(st->ngroup), dim3(kBlock), lm, s, *t, *st,
                       D, shift, c, (const T *)Z, sigma, bad);
    return (int)hipGetLastError();
}

} // namespace

extern "C" size_t dk_sigma_star_lds(const dk_tables *t, const dk_stars *st, int prec, int pass)
{
    if (pass == 0)
        return prec ? star_lds<float>(t->L, t->N, *st, nullptr, nullptr)
                    : star_lds<double>(t->L, t->N, *st, nullptr, nullptr);
    return prec ? main_lds<float>(t->L, t->N, *st) : main_lds<double>(t->L, t->N, *st);
}

extern "C" int dk_sigma_star(const dk_tables *t, const dk_stars *st, const double *D, double shift,
                             const void *c, void *sigma, void *Z, int prec, uint32_t *bad,
                             hipStream_t s)
{
    if (st->ngroup == 0) return (int)hipErrorInvalidValue;
    if (prec == 0)
        return bad ? launch<double, true>(t, st, D, shift, (const double *)c, (double *)sigma,
                                          (double *)Z, bad, s)
                   : launch<double, false>(t, st, D, shift, (const double *)c, (double *)sigma,
                                           (double *)Z, bad, s);
    return bad ? launch<float, true>(t, st, D, shift, (const float *)c, (float *)sigma, (float *)Z,
                                     bad, s)
               : launch<float, false>(t, st, D, shift, (const float *)c, (float *)sigma, (float *)Z,
                                      bad, s);
}
