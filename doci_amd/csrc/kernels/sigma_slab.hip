// sigma_slab.hip -- sigma = (H - shift) c, slab-blocked for gfx950.
//
// The L orbitals are split into A = {0..a-1} and B = {a..L-1} (the b
// fastest-varying orbitals of the lexicographic order).  Every determinant
// is (alpha, beta) with alpha in A (k pairs) and beta in B (m = N - k pairs);
// for fixed alpha the C(b, m) determinants form one contiguous "slab" of the
// rank space, at offset lexB(beta) from start(alpha) =
// n - C(b,m) - sum_e C(L-1-alpha_e, N-e).  A pair hop o -> t of I is then one
// of four kinds:
//   BB  (o, t in B): same slab, new offset -> gathered from the slab staged in LDS;
//   AA  (o, t in A): slab start(alpha - o + t), same offset -> coalesced
//                    stream over the neighbour slab (no per-lane rank work);
//   AB  (o in A, t in B): slab start(alpha - o) with m+1 pairs, offset
//                    lexB_{m+1}(beta + t) -> gather inside that slab;
//   BA  (o in B, t in A): slab start(alpha + t) with m-1 pairs, offset
//                    lexB_{m-1}(beta - o) -> gather inside that slab.
// A workgroup owns a group of slabs with the same m (whole slabs, so the
// result is a pull-form sum with no atomics).  Its prologue computes the
// neighbour-slab starts of every slab once into LDS; per element only the
// b-orbital walks remain.  Groups are ordered by rank and dealt to the 8
// XCDs in contiguous runs so that neighbouring slabs share an L2.
#include <hip/hip_runtime.h>

#include "../dk.h"
#include "addressing.hpp"

namespace {

constexpr int kBlock = 512;
constexpr int kXcd = 8;

struct Lds {
    double *K;
    uint32_t *bt;
    void *cown;
    uint32_t *aa_start, *ab_start, *ba_start;
    uint16_t *aa_ot;
    uint8_t *ab_o, *ba_t;
};

__device__ __forceinline__ size_t align8(size_t x) { return (x + 7) & ~(size_t)7; }

template <typename T>
__device__ __forceinline__ Lds carve(char *base, const dk_tables &t, const dk_slabs &sl)
{
    Lds p;
    size_t off = 0;
    p.K = (double *)(base + off);
    off += align8(sizeof(double) * t.L * t.L);
    p.cown = base + off;
    off += align8(sizeof(T) * sl.emax);
    p.bt = (uint32_t *)(base + off);
    off += align8(sizeof(uint32_t) * (t.L + 1) * (t.N + 2));
    p.aa_start = (uint32_t *)(base + off);
    off += align8(sizeof(uint32_t) * sl.gmax * sl.aamax);
    p.ab_start = (uint32_t *)(base + off);
    off += align8(sizeof(uint32_t) * sl.gmax * sl.abmax);
    p.ba_start = (uint32_t *)(base + off);
    off += align8(sizeof(uint32_t) * sl.gmax * sl.bamax);
    p.aa_ot = (uint16_t *)(base + off);
    off += align8(sizeof(uint16_t) * sl.gmax * sl.aamax);
    p.ab_o = (uint8_t *)(base + off);
    off += align8(sl.gmax * sl.abmax);
    p.ba_t = (uint8_t *)(base + off);
    return p;
}

// start of the slab of A-part `amask` (|amask| = N - mp pairs in A, mp in B)
__device__ __forceinline__ uint32_t slab_start(uint64_t amask, int mp, int L, int N, int b,
                                               uint32_t n, const dk::Binom &C)
{
    uint32_t s = n - C(b, mp);
    int e = 0;
    while (amask) {
        const int x = __ffsll((unsigned long long)amask) - 1;
        amask &= amask - 1;
        s -= C(L - 1 - x, N - e);
        ++e;
    }
    return s;
}

// kChecked: every computed global index is tested against n and every
// LDS-table index against its extent; a violation is counted in *bad and
// the access skipped (used to prove the index arithmetic before the
// unchecked kernel runs).
template <typename T, bool kChecked>
__global__ __launch_bounds__(kBlock) void k_sigma_slab(dk_tables t, dk_slabs sl,
                                                       const double *__restrict__ D, double shift,
                                                       const T *__restrict__ c,
                                                       T *__restrict__ sigma, uint32_t *bad)
{
    extern __shared__ double lds_raw[];
    // groups are dealt to XCDs in contiguous runs (workgroup i runs on XCD i % 8)
    const uint32_t per = (sl.ngroup + kXcd - 1) / kXcd;
    const uint32_t g = sl.xcd_runs ? (blockIdx.x % kXcd) * per + blockIdx.x / kXcd : blockIdx.x;
    if (g >= sl.ngroup) return; // uniform for the whole workgroup
    const int L = (int)t.L, N = (int)t.N, a = (int)sl.a, b = (int)sl.b;
    const uint32_t n = t.n;
    Lds p = carve<T>((char *)lds_raw, t, sl);
    auto ok = [&](uint32_t idx, uint32_t lim) -> bool {
        if (!kChecked) return true;
        if (idx < lim) return true;
        atomicAdd(bad, 1u);
        return false;
    };
    T *cown = (T *)p.cown;
    for (int i = threadIdx.x; i < L * L; i += kBlock) p.K[i] = t.K[i];
    for (int i = threadIdx.x; i < (L + 1) * (N + 2); i += kBlock) p.bt[i] = t.binom[i];
    const uint32_t first = sl.group[2 * g], gw = sl.group[2 * g + 1];
    const int nsl = (int)(gw & 0xffffu), m = (int)(gw >> 16), k = N - m;
    __syncthreads();
    const dk::Binom C{p.bt, N + 2};
    const int S = (int)C(b, m);
    const int nel = nsl * S;
    const int nA = a - k; // empty A orbitals

    // stage the own slabs
    for (int e = threadIdx.x; e < nel; e += kBlock) {
        const int s = e / S, bi = e - s * S;
        const uint32_t gi = sl.slab_start[first + s] + bi;
        cown[e] = ok(gi, n) ? c[gi] : (T)0;
    }
    // neighbour-slab tables: task (s, j), j < k: hole j of alpha (AA row, AB entry);
    // k <= j < a: the (j-k)-th empty A orbital (BA entry)
    for (int task = threadIdx.x; task < nsl * a; task += kBlock) {
        const int s = task / a, j = task - s * a;
        const uint64_t am = sl.slab_amask[first + s];
        const uint32_t st = sl.slab_start[first + s];
        if (j < k) {
            uint64_t r = am;
            for (int q = 0; q < j; ++q) r &= r - 1;
            const int o = __ffsll((unsigned long long)r) - 1;
            uint32_t *as = p.aa_start + s * sl.aamax + j * nA;
            uint16_t *ao = p.aa_ot + s * sl.aamax + j * nA;
            const uint32_t base = st + C(L - 1 - o, N - j);
            int jj = j;
            uint32_t sm = 0;
            for (int tt = o - 1; tt >= 0; --tt) {
                if ((am >> tt) & 1) {
                    --jj;
                    sm += C(L - 1 - tt, N - jj) - C(L - 1 - tt, N - jj - 1);
                } else if (ok((uint32_t)(tt - jj), (uint32_t)nA)) {
                    as[tt - jj] = base - C(L - 1 - tt, N - jj) + sm;
                    ao[tt - jj] = (uint16_t)(o | (tt << 8));
                }
            }
            jj = j + 1;
            sm = 0;
            for (int tt = o + 1; tt < a; ++tt) {
                if ((am >> tt) & 1) {
                    sm += C(L - 1 - tt, N - jj + 1) - C(L - 1 - tt, N - jj);
                    ++jj;
                } else if (ok((uint32_t)(tt - jj), (uint32_t)nA)) {
                    as[tt - jj] = base - C(L - 1 - tt, N - jj + 1) - sm;
                    ao[tt - jj] = (uint16_t)(o | (tt << 8));
                }
            }
            if (m + 1 <= b) {
                p.ab_start[s * sl.abmax + j] = slab_start(am & ~(1ull << o), m + 1, L, N, b, n, C);
                p.ab_o[s * sl.abmax + j] = (uint8_t)o;
            }
        } else if (m >= 1) {
            const int i = j - k;
            // i-th zero bit of am below a
            uint64_t z = ~am & ((a < 64) ? ((1ull << a) - 1) : ~0ull);
            for (int q = 0; q < i; ++q) z &= z - 1;
            const int tt = __ffsll((unsigned long long)z) - 1;
            p.ba_start[s * sl.bamax + i] = slab_start(am | (1ull << tt), m - 1, L, N, b, n, C);
            p.ba_t[s * sl.bamax + i] = (uint8_t)tt;
        }
    }
    __syncthreads();

    const double *K = p.K;
    const int nAA = k * nA;
    const uint32_t Cb_m1 = m + 1 <= b ? C(b, m + 1) : 0u, Cb_mm1 = m >= 1 ? C(b, m - 1) : 0u;
    for (int e = threadIdx.x; e < nel; e += kBlock) {
        const int s = e / S, bi = e - s * S;
        const uint32_t I = sl.slab_start[first + s] + (uint32_t)bi;
        if (!ok(I, n)) continue;
        const uint64_t bm = dk::unrank((uint32_t)bi, b, m, C);
        const T *own = cown + s * S;
        double acc = (D[I] - shift) * (double)own[bi];
        // BB: hops inside the slab
        dk::for_each_hop((uint32_t)bi, bm, b, m, C, [&](int o, int tt, uint32_t J) {
            if (ok(J, (uint32_t)S)) acc += K[(a + o) * L + a + tt] * (double)own[J];
        });
        // AA: neighbour slabs at the same offset (coalesced)
        {
            const uint32_t *as = p.aa_start + s * sl.aamax;
            const uint16_t *ao = p.aa_ot + s * sl.aamax;
            // four independent partial sums keep four neighbour-slab loads in flight
            double q0 = 0, q1 = 0, q2 = 0, q3 = 0;
            int h = 0;
            for (; h + 4 <= nAA; h += 4) {
                const uint32_t g0 = as[h] + bi, g1 = as[h + 1] + bi, g2 = as[h + 2] + bi,
                               g3 = as[h + 3] + bi;
                const uint16_t o0 = ao[h], o1 = ao[h + 1], o2 = ao[h + 2], o3 = ao[h + 3];
                const double v0 = ok(g0, n) ? (double)c[g0] : 0.0;
                const double v1 = ok(g1, n) ? (double)c[g1] : 0.0;
                const double v2 = ok(g2, n) ? (double)c[g2] : 0.0;
                const double v3 = ok(g3, n) ? (double)c[g3] : 0.0;
                q0 += K[(o0 & 0xff) * L + (o0 >> 8)] * v0;
                q1 += K[(o1 & 0xff) * L + (o1 >> 8)] * v1;
                q2 += K[(o2 & 0xff) * L + (o2 >> 8)] * v2;
                q3 += K[(o3 & 0xff) * L + (o3 >> 8)] * v3;
            }
            for (; h < nAA; ++h) {
                const uint16_t ot = ao[h];
                const uint32_t gi = as[h] + bi;
                if (ok(gi, n)) q0 += K[(ot & 0xff) * L + (ot >> 8)] * (double)c[gi];
            }
            acc += (q0 + q1) + (q2 + q3);
        }
        const uint32_t Sb = (uint32_t)S - 1u - (uint32_t)bi; // sum of beta's rank terms
        // AB: a pair leaves A for the empty B orbital tl
        if (m + 1 <= b && k > 0) {
            const uint32_t *abs_ = p.ab_start + s * sl.abmax;
            const uint8_t *abo = p.ab_o + s * sl.abmax;
            uint32_t P1 = 0, P2 = 0;
            int jj = 0;
            for (int tl = 0; tl < b; ++tl) {
                if ((bm >> tl) & 1) {
                    P2 += C(b - 1 - tl, m + 1 - jj);
                    P1 += C(b - 1 - tl, m - jj);
                    ++jj;
                } else {
                    const uint32_t off = Cb_m1 - 1u - P2 - C(b - 1 - tl, m + 1 - jj) - (Sb - P1);
                    for (int q = 0; q < k; ++q) {
                        const uint32_t gi = abs_[q] + off;
                        if (ok(gi, n)) acc += K[abo[q] * L + a + tl] * (double)c[gi];
                    }
                }
            }
        }
        // BA: the pair on B orbital tl moves into the empty A orbitals
        if (m >= 1 && nA > 0) {
            const uint32_t *bas = p.ba_start + s * sl.bamax;
            const uint8_t *bat = p.ba_t + s * sl.bamax;
            uint32_t Q1 = 0, P1 = 0;
            int h = 0;
            uint64_t rem = bm;
            while (rem) {
                const int tl = __ffsll((unsigned long long)rem) - 1;
                rem &= rem - 1;
                const uint32_t off = Cb_mm1 - 1u - Q1 - (Sb - P1 - C(b - 1 - tl, m - h));
                const double *Krow = K + (a + tl) * L;
                for (int q = 0; q < nA; ++q) {
                    const uint32_t gi = bas[q] + off;
                    if (ok(gi, n)) acc += Krow[bat[q]] * (double)c[gi];
                }
                Q1 += C(b - 1 - tl, m - 1 - h);
                P1 += C(b - 1 - tl, m - h);
                ++h;
            }
        }
        sigma[I] = (T)acc;
    }
}

template <typename T>
size_t lds_bytes(const dk_tables &t, const dk_slabs &sl)
{
    auto a8 = [](size_t x) { return (x + 7) & ~(size_t)7; };
    return a8(sizeof(double) * t.L * t.L) + a8(sizeof(T) * sl.emax) +
           a8(sizeof(uint32_t) * (t.L + 1) * (t.N + 2)) +
           a8(sizeof(uint32_t) * sl.gmax * sl.aamax) + a8(sizeof(uint32_t) * sl.gmax * sl.abmax) +
           a8(sizeof(uint32_t) * sl.gmax * sl.bamax) + a8(sizeof(uint16_t) * sl.gmax * sl.aamax) +
           a8((size_t)sl.gmax * sl.abmax) + a8((size_t)sl.gmax * sl.bamax);
}

} // namespace

extern "C" size_t dk_sigma_slab_lds(const dk_tables *t, const dk_slabs *sl, int prec)
{
    return prec ? lds_bytes<float>(*t, *sl) : lds_bytes<double>(*t, *sl);
}

extern "C" int dk_sigma_slab(const dk_tables *t, const dk_slabs *sl, const double *D,
                             double shift, const void *c, void *sigma, int prec, hipStream_t s)
{
    return dk_sigma_slab_checked(t, sl, D, shift, c, sigma, prec, NULL, s);
}

extern "C" int dk_sigma_slab_checked(const dk_tables *t, const dk_slabs *sl, const double *D,
                                     double shift, const void *c, void *sigma, int prec,
                                     uint32_t *bad, hipStream_t s)
{
    if (sl->ngroup == 0) return (int)hipErrorInvalidValue;
    const uint32_t per = (sl->ngroup + kXcd - 1) / kXcd;
    const dim3 grid(sl->xcd_runs ? per * kXcd : sl->ngroup);
    const size_t lds = prec ? lds_bytes<float>(*t, *sl) : lds_bytes<double>(*t, *sl);
    if (prec == 0 && !bad)
        hipLaunchKernelGGL((k_sigma_slab<double, false>), grid, dim3(kBlock), lds, s, *t, *sl, D,
                           shift, (const double *)c, (double *)sigma, bad);
    else if (prec == 0)
        hipLaunchKernelGGL((k_sigma_slab<double, true>), grid, dim3(kBlock), lds, s, *t, *sl, D,
                           shift, (const double *)c, (double *)sigma, bad);
    else if (!bad)
        hipLaunchKernelGGL((k_sigma_slab<float, false>), grid, dim3(kBlock), lds, s, *t, *sl, D,
                           shift, (const float *)c, (float *)sigma, bad);
    else
        hipLaunchKernelGGL((k_sigma_slab<float, true>), grid, dim3(kBlock), lds, s, *t, *sl, D,
                           shift, (const float *)c, (float *)sigma, bad);
    return (int)hipGetLastError();
}
