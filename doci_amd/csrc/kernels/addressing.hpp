// addressing.hpp -- device-side determinant addressing for DOCI.
//
// A determinant of N pairs in L orbitals is a 64-bit occupation mask (bit p =
// orbital p).  Ranks follow the reference's lexicographic order of sorted
// occupied tuples with the highest pair varying fastest (src/doci.c:472-548):
//   rank(c_0 < ... < c_{N-1}) = C(L,N) - 1 - sum_e C(L-1-c_e, N-e).
// All rank arithmetic is modulo 2^32; every true rank is < 2^31.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dk {

// Binomial table layout shared by every kernel: bt[x * N2 + j] = C(x, j),
// x in [0, L], j in [0, N+1], N2 = N + 2.
struct Binom {
    const uint32_t *bt;
    int N2;
    __device__ __forceinline__ uint32_t operator()(int x, int j) const { return bt[x * N2 + j]; }
};

// Rank -> mask by walking the lexicographic tree (O(L) table lookups).
__device__ __forceinline__ uint64_t unrank(uint32_t r, int L, int N, const Binom &C)
{
    uint64_t m = 0;
    int p = 0;
    for (int e = 0; e < N; ++e) {
        const int need = N - e - 1;
        for (;;) {
            const uint32_t blk = C(L - 1 - p, need);
            if (r < blk) break;
            r -= blk;
            ++p;
        }
        m |= 1ull << p;
        ++p;
    }
    return m;
}

// Visit every single pair hop J = I - {o} + {t} of determinant I (rank r,
// mask m): o runs over occupied orbitals ascending, t over empty orbitals,
// first downwards from o-1 then upwards from o+1.  rank(J) is updated in O(1)
// per step:
//   t < o: rank(J) = r + C(L-1-o, N-h) - C(L-1-t, N-j) + s_dn,
//          s_dn = sum over occupied c_e in (t, o) of C(L-1-c_e, N-e) - C(L-1-c_e, N-e-1)
//   t > o: rank(J) = r + C(L-1-o, N-h) - C(L-1-t, N-j+1) - s_up,
//          s_up = sum over occupied c_e in (o, t) of C(L-1-c_e, N-e+1) - C(L-1-c_e, N-e)
// where h is o's pair index and j = #{occupied orbitals of I below t}.
template <typename F>
__device__ __forceinline__ void for_each_hop(uint32_t r, uint64_t m, int L, int N,
                                             const Binom &C, F &&f)
{
    uint64_t rem = m;
    int h = 0;
    while (rem) {
        const int o = __ffsll((unsigned long long)rem) - 1;
        rem &= rem - 1;
        const uint32_t base = r + C(L - 1 - o, N - h);
        int j = h;
        uint32_t s = 0;
        for (int t = o - 1; t >= 0; --t) {
            if ((m >> t) & 1) {
                --j;
                s += C(L - 1 - t, N - j) - C(L - 1 - t, N - j - 1);
            } else {
                f(o, t, base - C(L - 1 - t, N - j) + s);
            }
        }
        j = h + 1;
        s = 0;
        for (int t = o + 1; t < L; ++t) {
            if ((m >> t) & 1) {
                s += C(L - 1 - t, N - j + 1) - C(L - 1 - t, N - j);
                ++j;
            } else {
                f(o, t, base - C(L - 1 - t, N - j + 1) - s);
            }
        }
        ++h;
    }
}

// The same walk without per-lane branches, for lanes that gather: f(o, t, J,
// hop) is called for every t != o of every occupied o (L-1 calls per o, the
// same count in every lane), with hop = (t is empty) and J = rank(I - o + t)
// when hop (J is meaningless otherwise).  Both binomials of a step are read
// whether or not t is occupied, so the loop has no divergent branch and the
// caller's gathers of consecutive steps can be in flight together.
template <typename F>
__device__ __forceinline__ void for_each_hop_pred(uint32_t r, uint64_t m, int L, int N,
                                                  const Binom &C, F &&f)
{
    uint64_t rem = m;
    int h = 0;
    while (rem) {
        const int o = __ffsll((unsigned long long)rem) - 1;
        rem &= rem - 1;
        const uint32_t base = r + C(L - 1 - o, N - h);
        int j = h;
        uint32_t s = 0;
#pragma unroll 4
        for (int t = o - 1; t >= 0; --t) {
            const int occ = (int)((m >> t) & 1);
            j -= occ;
            const uint32_t c1 = C(L - 1 - t, N - j), c0 = C(L - 1 - t, N - j - 1);
            f(o, t, base - c1 + s, occ == 0);
            s += occ ? c1 - c0 : 0u;
        }
        j = h + 1;
        s = 0;
#pragma unroll 4
        for (int t = o + 1; t < L; ++t) {
            const int occ = (int)((m >> t) & 1);
            const uint32_t c1 = C(L - 1 - t, N - j + 1), c0 = C(L - 1 - t, N - j);
            f(o, t, base - c1 - s, occ == 0);
            s += occ ? c1 - c0 : 0u;
            j += occ;
        }
        ++h;
    }
}

// Hop ranks in separable form, for determinants with N <= NM pairs and
// L - N <= KM empty orbitals (every array index below is a compile-time
// constant, so the arrays stay in registers).  With o_h the h-th occupied
// orbital, t_k the k-th empty one and j = t_k - k the occupied orbitals below
// t_k, the two O(1) updates above telescope into
//   t_k < o_h: rank(J) = Pd_h - Td_k,   t_k > o_h: rank(J) = Pu_h - Tu_k,
//   Pd_h = r + C(L-1-o_h, N-h) + Qd_h,       Td_k = C(L-1-t_k, N-j) + Qd_j,
//   Pu_h = r + C(L-1-o_h, N-h) + Qu_{h+1},   Tu_k = C(L-1-t_k, N-j+1) + Qu_j,
// with the prefix sums over the occupied orbitals
//   Qd_e = sum_{e'<e} C(L-1-o_e', N-e') - C(L-1-o_e', N-e'-1),
//   Qu_e = sum_{e'<e} C(L-1-o_e', N-e'+1) - C(L-1-o_e', N-e').
// f(o, t, J, hop) is called for the N(L-N) hops (hop = true), o ascending,
// then t ascending, in groups of G empty orbitals with no branch inside a
// group, so that a caller's G gathers are in flight together; the calls of a
// group past the last empty orbital have hop = false, t = 0 and a J that
// may lie outside the vector (read it with vec_load and discard it).
template <int NM, int KM, int G, typename F>
__device__ __forceinline__ void for_each_hop_sep(uint32_t r, uint64_t m, int L, int N,
                                                 const Binom &C, F &&f)
{
    int o[NM];
    uint32_t Pd[NM], Pu[NM], wd[NM], wu[NM];
    uint64_t rem = m;
    uint32_t qd = 0, qu = 0;
#pragma unroll
    for (int h = 0; h < NM; ++h) {
        if (h < N) {
            const int oh = __ffsll((unsigned long long)rem) - 1;
            rem &= rem - 1;
            o[h] = oh;
            const int x = L - 1 - oh;
            const uint32_t cm = C(x, N - h - 1), c0 = C(x, N - h), cp = C(x, N - h + 1);
            Pd[h] = r + c0 + qd;
            wd[h] = c0 - cm;
            wu[h] = cp - c0;
            qd += wd[h];
            qu += wu[h];
            Pu[h] = r + c0 + qu;
        }
    }
    const int nk = L - N;
    int t[KM];
    uint32_t Td[KM], Tu[KM];
    uint64_t z = ~m & (L < 64 ? (1ull << L) - 1 : ~0ull);
#pragma unroll
    for (int k = 0; k < KM; ++k) {
        if (k < nk) {
            const int tk = __ffsll((unsigned long long)z) - 1;
            z &= z - 1;
            t[k] = tk;
            const int j = tk - k, x = L - 1 - tk;
            uint32_t sd = 0, su = 0;
#pragma unroll
            for (int h = 0; h < NM; ++h)
                if (h < N) {
                    const bool below = o[h] < tk;
                    sd += below ? wd[h] : 0u;
                    su += below ? wu[h] : 0u;
                }
            Td[k] = C(x, N - j) + sd;
            Tu[k] = C(x, N - j + 1) + su;
        } else {
            t[k] = 0;
            Td[k] = Tu[k] = 0;
        }
    }
#pragma unroll
    for (int h = 0; h < NM; ++h)
        if (h < N) {
#pragma unroll
            for (int k0 = 0; k0 < KM; k0 += G)
                if (k0 < nk) {
#pragma unroll
                    for (int k = k0; k < k0 + G; ++k)
                        f(o[h], t[k], t[k] < o[h] ? Pd[h] - Td[k] : Pu[h] - Tu[k], k < nk);
                }
        }
}

// A vector addressed through a buffer resource: a load at any 32-bit byte
// offset is safe (past the vector it returns 0), so a walk may issue loads
// for steps it then discards without computing an in-range index for them.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t vec_rsrc(const void *base, uint64_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0,
                                             (int)(bytes < 0xffffffffull ? bytes : 0xffffffffull),
                                             0x00020000);
}
template <typename T>
__device__ __forceinline__ double vec_load(__amdgpu_buffer_rsrc_t r, uint32_t i)
{
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(i * 8u), 0, 0));
    else
        return (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(i * 4u), 0, 0));
}

// Diagonal entry of determinant mask m (src/doci.c:589-624), the reference's
// exact addition order: start at core, occupied orbitals in descending order
// add 2 h'_o, then per q ascending 2 is J_oq then (1 - is) X_oq.  One of the
// two per-q terms is an exact zero that leaves the sum unchanged, so a single
// add of the other gives the same bits.  XJ[2 (o L + q) + is] holds X_oq for
// is = 0 and 2 J_oq (exact) for is = 1.
__device__ __forceinline__ double diag_of(uint64_t m, int L, double core, const double *XJ,
                                          const double *h)
{
    double acc = core;
    uint64_t rem = m;
    while (rem) {
        const int o = 63 - __clzll((long long)rem);
        rem ^= 1ull << o;
        acc += 2 * h[o];
        const double *row = XJ + 2 * o * L;
        uint64_t mq = m; // bit 0: q occupied
#pragma unroll 8
        for (int q = 0; q < L; ++q) {
            acc += row[2 * q + (int)(mq & 1)];
            mq >>= 1;
        }
    }
    return acc;
}

} // namespace dk
