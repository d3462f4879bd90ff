#pragma once
// star_kernels.hpp -- the kernels of sigma = (H - shift) c in two passes (star
// pass, main pass); included by sigma_star.hip (launcher, C API) and by the
// translation units that instantiate the column pass at exact row counts.
//
// Orbitals split into A = {0..a-1} and B = {a..L-1} (the b fastest-varying
// orbitals of the lexicographic order).  A slab = the C(b, m) determinants
// with one A-part alpha (k = N - m pairs in A); it is contiguous in rank,
// starting at n - C(b,m) - sum_e C(L-1-alpha_e, N-e), in the order lexB(beta).
//
// Star pass.  The star of rho (j pairs in A) is its centre slab rho (m = N-j
// pairs in B, absent when N - j > b) and the a - j member slabs rho + y.  A
// lane per (star, member offset beta) loads the members' values at beta into
// registers (coalesced across lanes).  For member rho + x it sums every hop
// that empties x:  x -> y in A  (sum_y K_xy X_y[beta]: a small dense product
// in registers) and  x -> t in B  (K_xt times the centre, staged in LDS, at
// lexB(beta + t));  at the same centre offsets it accumulates in LDS the
// centre's hops from B into A.  Each sum is one slot of Z.  Every slab is
// read k + 1 times per sigma (k stars as member, one as centre) with
// full-line coalesced loads, instead of once per neighbouring determinant.
//
// Main pass.  sigma = (D - shift) c + hops inside the slab (own slab staged in
// LDS) + the slab's Z slots (coalesced streams).
#include <hip/hip_runtime.h>

#include "../dk.h"
#include "addressing.hpp"

namespace {

constexpr int kBlock = 256;
constexpr int kCtrPad = 64;
// Byte offset of an absent member (or a lane past the last column) in the
// buffer-addressed column pass: beyond every buffer's range, so the
// hardware range check returns 0 for a load and drops a store.  Buffers must
// stay below kBufLimit bytes (checked by the launcher), and an offset within
// a slab (< 2^16 elements of at most 8 bytes) added to kOob does not wrap.
constexpr uint32_t kOob = 0xfff00000u;
constexpr uint64_t kBufLimit = 0xffe00000ull;

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
    double *acc;   // centre accumulators (hops from B into A), per star
    void *ctr;     // staged centre slabs, per star
    uint32_t *mem_start, *mem_zb, *ctr_start, *ctr_zb;
    uint8_t *ycol;
};

// Star pass LDS: K, binomials, per star the centre slab (values + FP64
// accumulator, together `selems` elements per group) and member tables.
// kWithK = false (the column passes, which read K through scalar loads):
// no K table, so that more workgroups fit a CU.
template <typename T, bool kWithK = true>
__host__ __device__ size_t star_lds(uint32_t L, uint32_t N, const dk_stars &st, StarLds *p, char *base)
{
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *q = base ? base + off : nullptr;
        off += a8(bytes);
        return q;
    };
    char *K = take(kWithK ? sizeof(double) * L * L : 0);
    char *bt = take(sizeof(uint32_t) * (L + 1) * (N + 2));
    // centres: + kCtrPad entries behind the staged ones (zero values / scratch
    // accumulators for lanes without a centre hop); member tables: + one row
    // of absent members (the buffer-addressed column pass)
    char *acc = take(sizeof(double) * (st.selems + kCtrPad));
    char *ctr = take(sizeof(T) * (st.selems + kCtrPad));
    char *ms = take(sizeof(uint32_t) * (st.sgmax + 1) * st.ymax);
    char *mz = take(sizeof(uint32_t) * (st.sgmax + 1) * st.ymax);
    char *cs = take(sizeof(uint32_t) * st.sgmax);
    char *cz = take(sizeof(uint32_t) * st.sgmax);
    char *yc = take(st.sgmax * (st.ymax > 8 ? st.ymax : 8)); // also the stars' A-parts (column pass)
    if (p) {
        p->K = (double *)K;
        p->bt = (uint32_t *)bt;
        p->acc = (double *)acc;
        p->ctr = ctr;
        p->mem_start = (uint32_t *)ms;
        p->mem_zb = (uint32_t *)mz;
        p->ctr_start = (uint32_t *)cs;
        p->ctr_zb = (uint32_t *)cz;
        p->ycol = (uint8_t *)yc;
    }
    return off;
}

// Main pass LDS: K restricted to B x B, binomials C(x, j) for x <= b, the
// staged slabs (FP64 in either precision), the group's slab starts and slot
// bases.
template <typename T>
__host__ __device__ size_t main_lds(uint32_t L, uint32_t N, const dk_stars &st)
{
    return a8(sizeof(double) * (st.b * st.b + 1)) + a8(sizeof(uint32_t) * (st.b + 1) * (N + 2)) +
           a8(sizeof(T) * st.owns) + 2 * a8(sizeof(uint32_t) * st.gslabs);
}

// One lane per (star, member offset beta).  The lane loads the member slabs'
// values at beta (coalesced across lanes) into registers and produces, for
// every member rho + x, the sum of the hops that empty x:
//   sum_y K_xy X_y[beta]            (A -> A, a dense |Y| x |Y| product)
// + sum_t K_xt ctr[lexB(beta + t)]  (A -> B, centre staged in LDS)
// and adds to the centre accumulator at lexB(beta + t) the hops of the
// centre from B orbital t into A:  sum_y K_ty X_y[beta].
// The products run over YMAX >= |Y| members and BMAX >= b orbitals with zero
// padding, so that they are straight-line code (no per-term branches) whose
// LDS reads the compiler can batch.  Groups [g0, g0 + gridDim.x) all have
// a - j <= YMAX.
// What both star-pass kernels know about their star group after set-up.
struct StarCtx {
    int L, N, a, b;
    uint32_t n, ztot;
    uint32_t zm_sector, zm_len; // first Z index and length of the members' sector
    int nst, j, mc, mm, Sm, Sc, ny;
    bool has_ctr;
    const uint16_t *ins; // insert table of the members' B-pair count
};

// Stages K, the binomials, the member tables and the centre slabs of star
// group sg in LDS (zeroing the centre accumulators); ends with a barrier.
// kByY (the column pass): no K, and the member tables indexed by orbital,
// entry s * a + y for every y < a (zero for y in rho), copied from the
// host-built st.mtab instead of computed by member number s * ny + yi.
// kBuf (with kByY): the member tables hold byte offsets instead, the first
// rank times sizeof(T) and the slot position relative to the members' Z
// sector times sizeof(T), kOob for y in rho and in the extra row nst; the
// kCtrPad centre values behind the staged ones are zero.
template <typename T, bool kChecked, int NT = kBlock, bool kByY = false, bool kBuf = false, typename OK>
__device__ __forceinline__ StarCtx star_setup(const dk_tables &t, const dk_stars &st, uint32_t sg,
                                              const T *__restrict__ c, StarLds &p, OK &&ok)
{
    StarCtx x;
    x.L = (int)t.L;
    x.N = (int)t.N;
    x.a = (int)st.a;
    x.b = (int)st.b;
    x.n = t.n;
    x.ztot = st.zoff[x.a + 1];
    const int L = x.L, N = x.N, a = x.a, b = x.b;
    const uint32_t n = x.n;
    if (!kByY)
        for (int i = threadIdx.x; i < L * L; i += NT) p.K[i] = t.K[i];
    for (int i = threadIdx.x; i < (L + 1) * (N + 2); i += NT) p.bt[i] = t.binom[i];
    const uint32_t first = st.sgroup[2 * sg], w = st.sgroup[2 * sg + 1];
    const int nst = (int)(w & 0xffffu), j = (int)(w >> 16);
    __syncthreads();
    const dk::Binom C{p.bt, N + 2};
    const int mc = N - j;               // B pairs of the centre
    const bool has_ctr = mc <= b;
    const int mm = mc - 1;              // B pairs of the members
    const int Sm = (int)C(b, mm), Sc = has_ctr ? (int)C(b, mc) : 0;
    const int ny = a - j;
    x.nst = nst;
    x.j = j;
    x.mc = mc;
    x.mm = mm;
    x.Sm = Sm;
    x.Sc = Sc;
    x.ny = ny;
    x.has_ctr = has_ctr;
    x.ins = st.ins + st.ins_off[has_ctr ? mm : 0];
    const int km = j + 1;
    const uint32_t zm_sector = st.zoff[km], zc_sector = st.zoff[j];
    x.zm_sector = zm_sector;
    x.zm_len = st.zoff[km + 1] - zm_sector;
    const int nsm = nslot_of(km, mm, a), nsc = nslot_of(j, mc, a);
    const uint64_t amask_all = a < 64 ? ((1ull << a) - 1) : ~0ull;

    // centre slab starts first: the centre loads below are then in flight
    // while the member tables are built (kByY: the stars' A-parts go to LDS
    // too, for the column loop)
    uint64_t *srho = (uint64_t *)p.ycol;
    for (int s = threadIdx.x; s < nst; s += NT) {
        const uint64_t rho = st.star_amask[first + s];
        if (kByY) srho[s] = rho;
        if (has_ctr) {
            p.ctr_start[s] = slab_first(rho, mc, L, N, b, n, C);
            p.ctr_zb[s] = zc_sector + lex_a(rho, j, a, C) * (uint32_t)(nsc * Sc) + (uint32_t)(j * Sc);
        }
    }
    __syncthreads();
    constexpr int CPT = 8; // centre loads in flight per lane
    const int nce = has_ctr ? nst * Sc : 0;
    T cv[CPT];
    auto ctr_load = [&](int e0) {
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int e = e0 + u * NT + (int)threadIdx.x;
            cv[u] = (T)0;
            if (e < nce) {
                const int s = e / Sc;
                const uint32_t gi = p.ctr_start[s] + (uint32_t)(e - s * Sc);
                if (ok(gi, n)) cv[u] = c[gi];
            }
        }
    };
    T *ctrv = (T *)p.ctr;
    auto ctr_put = [&](int e0) {
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int e = e0 + u * NT + (int)threadIdx.x;
            if (e < nce) {
                ctrv[e] = cv[u];
                p.acc[e] = 0.0;
            }
        }
    };
    ctr_load(0);

    // member tables: first rank, Z position of this star's slot, orbital y
    if (kByY && kBuf) { // byte offsets, kOob for absent members
        // eight table entries per lane in flight, loaded without branches
        // (entries of the extra row read entry 0 and are not used)
        const uint2 *mt = (const uint2 *)st.mtab + (size_t)first * a;
        constexpr int kML = 8;
        const int ntask = (nst + 1) * a;
        for (int t0 = threadIdx.x; t0 < ntask; t0 += kML * NT) {
            uint2 v[kML];
#pragma unroll
            for (int q = 0; q < kML; ++q) {
                const int task = t0 + q * NT;
                v[q] = mt[task < nst * a ? task : 0];
            }
#pragma unroll
            for (int q = 0; q < kML; ++q) {
                const int task = t0 + q * NT;
                if (task < ntask) {
                    const int s = task / a, y = task - s * a;
                    const bool mem = s < nst && !((srho[s < nst ? s : 0] >> y) & 1);
                    p.mem_start[task] = mem ? v[q].x * (uint32_t)sizeof(T) : kOob;
                    p.mem_zb[task] = mem ? (v[q].y - zm_sector) * (uint32_t)sizeof(T) : kOob;
                }
            }
        }
        if (threadIdx.x < kCtrPad) ((T *)p.ctr)[st.selems + threadIdx.x] = (T)0;
    } else if (kByY) { // built on the host (st.mtab), contiguous for the group's stars
        const uint2 *mt = (const uint2 *)st.mtab + (size_t)first * a;
        for (int task = threadIdx.x; task < nst * a; task += NT) {
            const uint2 v = mt[task];
            p.mem_start[task] = v.x;
            p.mem_zb[task] = v.y;
        }
    } else
    for (int task = threadIdx.x; task < nst * ny; task += NT) {
        const int s = task / ny, yi = task - s * ny;
        const uint64_t rho = st.star_amask[first + s];
        uint64_t z = ~rho & amask_all;
        for (int q = 0; q < yi; ++q) z &= z - 1;
        const int y = __ffsll((unsigned long long)z) - 1;
        const uint64_t mem = rho | (1ull << y);
        const int pos = __popcll(rho & ((1ull << y) - 1));
        p.mem_start[s * ny + yi] = slab_first(mem, mm, L, N, b, n, C);
        p.mem_zb[s * ny + yi] =
            zm_sector + lex_a(mem, km, a, C) * (uint32_t)(nsm * Sm) + (uint32_t)(pos * Sm);
        p.ycol[s * ny + yi] = (uint8_t)y;
    }
    ctr_put(0);
    for (int e0 = CPT * NT; e0 < nce; e0 += CPT * NT) {
        ctr_load(e0);
        ctr_put(e0);
    }
    __syncthreads();
    return x;
}

// Writes the centre accumulators to their Z slots (after a barrier).
template <typename T, int NT = kBlock, typename OK>
__device__ __forceinline__ void star_finish(const StarCtx &x, StarLds &p, T *__restrict__ Z, OK &&ok)
{
    if (!x.has_ctr) return;
    __syncthreads();
    for (int e = threadIdx.x; e < x.nst * x.Sc; e += NT) {
        const int s = e / x.Sc;
        const uint32_t zi = p.ctr_zb[s] + (uint32_t)(e - s * x.Sc);
        if (ok(zi, x.ztot)) Z[zi] = (T)p.acc[e];
    }
}

template <typename T, bool kChecked, int YMAX, int BMAX>
__global__ __launch_bounds__(kBlock, (YMAX <= 24 ? 3 : 2)) void k_star(dk_tables t, dk_stars st, uint32_t g0,
                                                    const T *__restrict__ c, T *__restrict__ Z,
                                                    uint32_t *bad, const dk_sigma_io *io)
{
    extern __shared__ double lds_raw[];
    const uint32_t sg = g0 + blockIdx.x;
    if (sg >= st.nsgroup) return;
    if (io) {
        if (!io->active) return;
        c = (const T *)io->c;
    }
    StarLds p;
    star_lds<T>(t.L, t.N, st, &p, (char *)lds_raw);
    auto ok = [&](uint32_t idx, uint32_t lim) -> bool {
        if (!kChecked) return true;
        if (idx < lim) return true;
        atomicAdd(bad, 1u);
        return false;
    };
    const StarCtx x = star_setup<T, kChecked>(t, st, sg, c, p, ok);
    const int L = x.L, a = x.a, b = x.b;
    const uint32_t n = x.n, ztot = x.ztot;
    const int nst = x.nst, Sm = x.Sm, Sc = x.Sc, ny = x.ny, mm = x.mm;
    const bool has_ctr = x.has_ctr;
    if (kChecked && (ny > YMAX || b > BMAX)) {
        if (threadIdx.x == 0) atomicAdd(bad, 1u);
        return;
    }
    T *ctrv = (T *)p.ctr;

    const double *K = p.K;
    for (int it = threadIdx.x; it < nst * Sm; it += kBlock) {
        const int s = it / Sm, bb = it - s * Sm;
        const uint8_t *yc = p.ycol + s * ny;
        const uint32_t *ms = p.mem_start + s * ny;
        // members beyond ny read member 0 again and count as zero
        double xv[YMAX];
        int yr[YMAX];
#pragma unroll
        for (int yi = 0; yi < YMAX; ++yi) {
            const int q = yi < ny ? yi : 0;
            yr[yi] = yc[q];
            const uint32_t gi = ms[q] + (uint32_t)bb;
            const double v = ok(gi, n) ? (double)c[gi] : 0.0;
            xv[yi] = yi < ny ? v : 0.0;
        }
        // centre values at lexB(beta + t): zero where t is in beta, t >= b or
        // there is no centre; orbital rows past b read row b - 1
        double cv[BMAX];
        uint32_t offs[BMAX];
        bool live[BMAX];
        {
            // lexB(beta + t) from the insert table (coalesced across lanes)
            const uint16_t *ip = st.ins + st.ins_off[has_ctr ? mm : 0] + bb;
#pragma unroll
            for (int tl = 0; tl < BMAX; ++tl) {
                const uint32_t off = (has_ctr && tl < b) ? ip[(size_t)tl * Sm] : 0xffffu;
                live[tl] = off != 0xffffu && ok(off, (uint32_t)Sc);
                offs[tl] = live[tl] ? off : 0;
                cv[tl] = live[tl] ? (double)ctrv[s * Sc + offs[tl]] : 0.0;
            }
        }
        // the centre's hops from B orbital tl into A, landing on the members at beta
        if (has_ctr) {
            double *acc = p.acc + s * Sc;
#pragma unroll
            for (int tl = 0; tl < BMAX; ++tl) {
                const double *Kt = K + (a + (tl < b ? tl : b - 1)) * L;
                double u0 = 0, u1 = 0;
#pragma unroll
                for (int yi = 0; yi < YMAX; yi += 2) {
                    u0 += Kt[yr[yi]] * xv[yi];
                    if (yi + 1 < YMAX) u1 += Kt[yr[yi + 1]] * xv[yi + 1];
                }
                if (live[tl])
                    __hip_atomic_fetch_add(&acc[offs[tl]], u0 + u1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        for (int xi = 0; xi < ny; ++xi) {
            const double *Kx = K + yc[xi] * L;
            double a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
            for (int yi = 0; yi < YMAX; yi += 2) {
                a0 += Kx[yr[yi]] * xv[yi];
                if (yi + 1 < YMAX) a1 += Kx[yr[yi + 1]] * xv[yi + 1];
            }
#pragma unroll
            for (int tl = 0; tl < BMAX; ++tl) a2 += Kx[a + (tl < b ? tl : b - 1)] * cv[tl];
            const uint32_t zi = p.mem_zb[s * ny + xi] + (uint32_t)bb;
            if (ok(zi, ztot)) Z[zi] = (T)((a0 + a1) + a2);
        }
    }
    star_finish<T>(x, p, Z, ok);
}

// Star pass on the matrix cores.  A wave takes one star and 16 member
// offsets beta at a time and forms, with v_mfma_f64_16x16x4f64 (A 16x4,
// B 4x16, FP64 accumulate),
//   x rows:  Z_x[beta] = sum_y K_xy X_y[beta] + sum_t K_xt ctr[lexB(beta+t)]
//   t rows:  u_t[beta] = sum_y K_ty X_y[beta]  -> centre accumulator at lexB(beta+t)
// The B operands (member values X_y[beta], centre values at lexB(beta+t))
// are loaded once per 16 offsets and shared by all row tiles; the A operands
// (K entries) come from LDS, one read per lane per product.  Lane l holds
// A[l % 16][l / 16], B[l / 16][l % 16] and D[4 r + l / 16][l % 16] in its
// r-th accumulator register (layout checked on gfx950).
// XT: row tiles of 16 members (members <= 16 XT).
typedef double dk_d4 __attribute__((ext_vector_type(4)));

template <typename T, bool kChecked, int XT>
__global__ __launch_bounds__(kBlock) void k_star_mfma(dk_tables t, dk_stars st, uint32_t g0,
                                                      const T *__restrict__ c, T *__restrict__ Z,
                                                      uint32_t *bad, const dk_sigma_io *io)
{
    constexpr int YB = 4 * XT;  // 4-member blocks of the products
    constexpr int TB = 4;       // 4-orbital blocks of B (b <= 16)
    constexpr int kWaves = kBlock / 64;
    extern __shared__ double lds_raw[];
    const uint32_t sg = g0 + blockIdx.x;
    if (sg >= st.nsgroup) return;
    if (io) {
        if (!io->active) return;
        c = (const T *)io->c;
    }
    StarLds p;
    star_lds<T>(t.L, t.N, st, &p, (char *)lds_raw);
    auto ok = [&](uint32_t idx, uint32_t lim) -> bool {
        if (!kChecked) return true;
        if (idx < lim) return true;
        atomicAdd(bad, 1u);
        return false;
    };
    const StarCtx x = star_setup<T, kChecked>(t, st, sg, c, p, ok);
    const int L = x.L, a = x.a, b = x.b, ny = x.ny, Sm = x.Sm, Sc = x.Sc;
    if (kChecked && (ny > 16 * XT || b > 4 * TB)) {
        if (threadIdx.x == 0) atomicAdd(bad, 1u);
        return;
    }
    const T *ctrv = (const T *)p.ctr;
    const double *K = p.K;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int ny4 = (ny + 3) >> 2, ups = (Sm + 15) >> 4;
    for (int u = wave; u < x.nst * ups; u += kWaves) {
        const int s = u / ups;
        const int beta = (u - s * ups) * 16 + li;
        const bool bok = beta < Sm;
        const uint8_t *yc = p.ycol + s * ny;
        const uint32_t *ms = p.mem_start + s * ny;
        // B operands: member values and centre values at this lane's (k, beta)
        double xb[YB], gb[TB];
#pragma unroll
        for (int kb = 0; kb < YB; ++kb) {
            xb[kb] = 0.0;
            const int y = 4 * kb + lk;
            if (kb < ny4 && y < ny && bok) {
                const uint32_t gi = ms[y] + (uint32_t)beta;
                if (ok(gi, x.n)) xb[kb] = (double)c[gi];
            }
        }
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) {
            gb[tb] = 0.0;
            const int tt = 4 * tb + lk;
            if (x.has_ctr && tt < b && bok) {
                const uint32_t off = x.ins[(size_t)tt * Sm + beta];
                if (off != 0xffffu && ok(off, (uint32_t)Sc)) gb[tb] = (double)ctrv[s * Sc + off];
            }
        }
        // column orbital of each A-operand k index, per member block
        int ykb[YB];
#pragma unroll
        for (int kb = 0; kb < YB; ++kb) {
            const int y = 4 * kb + lk;
            ykb[kb] = (kb < ny4 && y < ny) ? yc[y] : -1;
        }
        // x rows: members rho + x
#pragma unroll
        for (int xt = 0; xt < XT; ++xt) {
            if (16 * xt >= ny) break; // uniform
            const int xr = 16 * xt + li;
            const int xo = xr < ny ? yc[xr] : -1;
            dk_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kb = 0; kb < YB; ++kb) {
                if (kb >= ny4) break; // uniform
                const double av = (xo >= 0 && ykb[kb] >= 0) ? K[xo * L + ykb[kb]] : 0.0;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, xb[kb], acc, 0, 0, 0);
            }
            if (x.has_ctr) {
#pragma unroll
                for (int tb = 0; tb < TB; ++tb) {
                    const int tt = 4 * tb + lk;
                    const double av = (xo >= 0 && tt < b) ? K[xo * L + a + tt] : 0.0;
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, gb[tb], acc, 0, 0, 0);
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int xw = 16 * xt + 4 * r + lk;
                if (xw < ny && bok) {
                    const uint32_t zi = p.mem_zb[s * ny + xw] + (uint32_t)beta;
                    if (ok(zi, x.ztot)) Z[zi] = (T)acc[r];
                }
            }
        }
        // t rows: the centre's hops from B orbital t into A
        if (x.has_ctr) {
            dk_d4 acc = {0.0, 0.0, 0.0, 0.0};
            const int tr = li;
#pragma unroll
            for (int kb = 0; kb < YB; ++kb) {
                if (kb >= ny4) break; // uniform
                const double av = (tr < b && ykb[kb] >= 0) ? K[(a + tr) * L + ykb[kb]] : 0.0;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, xb[kb], acc, 0, 0, 0);
            }
            double *accl = p.acc + s * Sc;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int tt = 4 * r + lk;
                if (tt < b && bok) {
                    const uint32_t off = x.ins[(size_t)tt * Sm + beta];
                    if (off != 0xffffu && ok(off, (uint32_t)Sc))
                        __hip_atomic_fetch_add(&accl[off], acc[r], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        }
    }
    star_finish<T>(x, p, Z, ok);
}

// Column star pass.  With the products written over all of A instead of
// over each star's members, the K entries no longer depend on the star:
// for a column (star rho, member offset beta), with X_y = c[rho + y][beta]
// (zero for y in rho) and C_t = ctr[lexB(beta + t)] (zero for t in beta),
//   x rows:  Z_x = sum_{y in A} K_xy X_y + sum_{t in B} K_xt C_t   (x not in rho)
//   t rows:  u_t = sum_{y in A} K_ty X_y  -> centre accumulator at lexB(beta + t)
// so every lane of a wave runs the same FMA sequence with K as a wave-uniform
// operand (scalar loads from the packed table KP, built per call by k_kpack),
// whatever star its column belongs to.  A wave takes 64 consecutive columns
// of the group; an orbital y that is in rho for every lane's star is skipped
// by the whole wave (a wave inside one star does (a - j) column passes).
// AMAX >= a and BMAX >= b pad the row sets; KP is zero in the padding and on
// the diagonal (x = y is not a hop).
//   KP[y * (AMAX + BMAX) + r]       r < AMAX: K_{r y};  r >= AMAX: K_{a + r - AMAX, y}
//   KP[AMAX * (AMAX + BMAX) + q * AMAX + x] = K_{x, a + q}
__host__ __device__ constexpr int col_amax(int a) { return 8 * ((a + 7) / 8); }
__host__ __device__ constexpr int col_bmax(int b) { return b <= 12 ? 12 : 16; }
// (plus 16 doubles of zero padding: the column pass prefetches 16 entries of
// whatever column comes next)
__host__ __device__ constexpr size_t col_kp_doubles(int am, int bm) { return (size_t)am * (am + bm) + (size_t)bm * am + 16; }
constexpr int kColAMax = 32; // larger A sets use the per-star kernel

template <int AMAX, int BMAX>
__global__ __launch_bounds__(kBlock) void k_kpack(dk_tables t, int a, int b, double *__restrict__ KP)
{
    constexpr int R = AMAX + BMAX;
    const int L = (int)t.L;
    for (int i = threadIdx.x; i < AMAX * R; i += kBlock) {
        const int y = i / R, r = i - y * R;
        double v = 0.0;
        if (y < a) {
            if (r < AMAX) {
                if (r < a && r != y) v = t.K[r * L + y];
            } else if (r - AMAX < b) {
                v = t.K[(a + r - AMAX) * L + y];
            }
        }
        KP[i] = v;
    }
    for (int i = threadIdx.x; i < BMAX * AMAX + 16; i += kBlock) {
        const int q = i / AMAX, xo = i - q * AMAX;
        KP[AMAX * R + i] = (q < b && xo < a) ? t.K[xo * L + a + q] : 0.0;
    }
}

// An integer laundered through an empty asm tied to a loop-varying value:
// loads at offsets derived from it cannot be hoisted out of the loop.  The
// asm has no side effects and touches no memory, so such loads of read-only
// kernel arguments stay scalar.
__device__ __forceinline__ int opaque(int v, int tie)
{
    v = __builtin_amdgcn_readfirstlane(v);
    asm("" : "+s"(v) : "s"(__builtin_amdgcn_readfirstlane(tie)));
    return v;
}

// out[r] += k[r] * v for r < NR, k in global memory read through scalar
// registers kKC entries at a time: kbuf holds the first kKC entries of k on
// entry, and on exit the first kKC of kn (the next call's k), which are
// loaded while the last chunk's FMAs run.  The scheduling barriers keep one
// chunk in flight beside the one in use.
constexpr int kKC = 8;
// skip: a row whose K entry is a structural zero (the diagonal x = y of an
// A column), left out when it is a constant after unrolling; -1 for none.
template <int NR, int NOUT>
__device__ __forceinline__ void col_fma(const double *k, const double *kn, double v, double (&out)[NOUT],
                                        double (&kbuf)[kKC], int skip = -1)
{
    constexpr int NCH = (NR + kKC - 1) / kKC;
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
        double knx[kKC];
        const bool last = ch + 1 == NCH;
        const double *src = last ? kn : k + kKC * (ch + 1);
        const int cnt = last ? kKC : (NR - kKC * (ch + 1) < kKC ? NR - kKC * (ch + 1) : kKC);
#pragma unroll
        for (int i = 0; i < kKC; ++i) knx[i] = i < cnt ? src[i] : 0.0;
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < kKC; ++i)
            if (kKC * ch + i < NR && kKC * ch + i != skip)
                out[kKC * ch + i] = __builtin_fma(kbuf[i], v, out[kKC * ch + i]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < kKC; ++i) kbuf[i] = knx[i];
    }
}

template <typename T>
__device__ __forceinline__ T load_global(const T *p)
{
    return *(const __attribute__((address_space(1))) T *)p;
}

template <typename T, bool kChecked, int AMAX, int BMAX>
__global__ __launch_bounds__(kBlock, 3) void k_star_col(dk_tables t, dk_stars st, const T *__restrict__ c,
                                                        T *__restrict__ Z, uint32_t *bad,
                                                        const dk_sigma_io *io,
                                                        const double *__restrict__ KP)
{
    constexpr int R = AMAX + BMAX;
    extern __shared__ double lds_raw[];
    const uint32_t sg = blockIdx.x;
    if (sg >= st.nsgroup) return;
    if (io) {
        if (!io->active) return;
        c = (const T *)io->c;
    }
    StarLds p;
    star_lds<T, false>(t.L, t.N, st, &p, (char *)lds_raw);
    auto ok = [&](uint32_t idx, uint32_t lim) -> bool {
        if (!kChecked) return true;
        if (idx < lim) return true;
        atomicAdd(bad, 1u);
        return false;
    };
    const StarCtx x = star_setup<T, kChecked, kBlock, true>(t, st, sg, c, p, ok);
    // star-group constants (wave-uniform; some were read from LDS)
    const int a = (int)st.a, b = (int)st.b;
    const int Sm = __builtin_amdgcn_readfirstlane(x.Sm);
    const int Sc = __builtin_amdgcn_readfirstlane(x.Sc);
    const int nst = __builtin_amdgcn_readfirstlane(x.nst);
    const bool has_ctr = __builtin_amdgcn_readfirstlane((int)x.has_ctr) != 0;
    if (a > AMAX || b > BMAX) { // uniform: the launcher picks AMAX >= a, BMAX >= b
        if (kChecked && threadIdx.x == 0) atomicAdd(bad, 1u);
        return;
    }
    const int ncol = nst * Sm;
    const uint64_t amask = a < 64 ? ((1ull << a) - 1) : ~0ull;
    const int lane = (int)(threadIdx.x & 63);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (int c0 = wave * 64; c0 < ncol; c0 += kBlock) {
        // K addresses recomputed per column block (scalar adds) instead of
        // hoisted out of the loop into spilled registers
        const double *kp = KP + opaque(0, c0);
        const int col = c0 + lane;
        const bool live = col < ncol;
        const int s = live ? col / Sm : 0;
        const int beta = live ? col - s * Sm : 0;
        const uint64_t rho = live ? ((const uint64_t *)p.ycol)[s] : amask; // staged by star_setup
        // orbitals in rho for every lane's star: skipped by the whole wave
        uint32_t clo = (uint32_t)rho, chi = (uint32_t)(rho >> 32);
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            clo &= (uint32_t)__shfl_xor((int)clo, o);
            chi &= (uint32_t)__shfl_xor((int)chi, o);
        }
        const uint64_t common = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)chi) << 32) |
                                (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)clo);
        // member tables of this lane's star, indexed by orbital
        const uint32_t *ms = p.mem_start + s * a;
        const uint32_t *mz = p.mem_zb + s * a;
        // member values X_y, all loads in flight at once, coalesced across the
        // lanes of one star.  Lanes without member y (y in rho, y >= a, or no
        // column) read a zero instead: KP[0] is a zero diagonal entry, which
        // reads as 0 in either precision.
        const T *zero = (const T *)kp;
        const uint64_t rhox = live ? (rho | ~amask) : ~0ull;
        // Orbitals outside the wave's set ym (in every lane's rho, or >= a)
        // are skipped by the whole wave; for FP64 their loads too (for FP32
        // the branches cost more registers than the loads they save).
        const uint64_t ym = ~common & amask; // the wave's orbitals y
        T v[AMAX];
#pragma unroll
        for (int y = 0; y < AMAX; ++y) {
            v[y] = (T)0;
            if (sizeof(T) == 4 || ((ym >> y) & 1)) {
                const bool mem = !((rhox >> y) & 1);
                const uint32_t gi = ms[y] + (uint32_t)beta; // read for every lane: the table is in LDS
                v[y] = load_global(mem && ok(gi, x.n) ? c + gi : zero);
            }
        }
        // the centre's insert offsets lexB(beta + t), loaded while the member
        // values arrive and the A products run
        uint32_t offs[BMAX];
        if (has_ctr) {
            // per-block copies of b and Sm: comparisons and offsets derived
            // from them are recomputed here instead of held across the loop
            const int bq = opaque(b, c0), Smq = opaque(Sm, c0);
#pragma unroll
            for (int q = 0; q < BMAX; ++q) {
                const uint16_t *row = x.ins + (size_t)(q * Smq); // uniform row base
                offs[q] = q < bq ? (uint32_t)row[(uint32_t)beta] : 0xffffu;
            }
        }
        // out[r]: row r < AMAX is member row x = r, row AMAX + q the centre row
        // of B orbital a + q
        double out[R];
#pragma unroll
        for (int r = 0; r < R; ++r) out[r] = 0.0;
        const double *kctr = kp + AMAX * R; // centre columns
        double kbuf[kKC];
        {
            const double *k0 = ym ? kp + (__ffsll((unsigned long long)ym) - 1) * R : kctr;
#pragma unroll
            for (int i = 0; i < kKC; ++i) kbuf[i] = k0[i];
        }
#pragma unroll
        for (int y = 0; y < AMAX; ++y) {
            if ((ym >> y) & 1) {
                const uint64_t rest = ym & ~((2ull << y) - 1);
                const double *kn = rest ? kp + (__ffsll((unsigned long long)rest) - 1) * R : kctr;
                col_fma<R>(kp + y * R, kn, (double)v[y], out, kbuf);
            }
        }
        // hops A -> B through the centre values, and the centre's hops B -> A:
        // the centre values first (all loads in flight together), the
        // accumulator updates, then the products
        if (has_ctr) {
            const int bq = opaque(b, c0);
            const T *ctrv = (const T *)p.ctr + s * Sc;
            double *cacc = p.acc + s * Sc;
            uint32_t on = 0;
            double cv[BMAX];
#pragma unroll
            for (int q = 0; q < BMAX; ++q) {
                const bool o = live && offs[q] != 0xffffu && ok(offs[q], (uint32_t)Sc);
                on |= (uint32_t)o << q;
                const double val = (double)ctrv[o ? offs[q] : 0];
                cv[q] = o ? val : 0.0;
            }
#pragma unroll
            for (int q = 0; q < BMAX; ++q)
                if ((on >> q) & 1)
                    __hip_atomic_fetch_add(&cacc[offs[q]], out[AMAX + q], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
            for (int q = 0; q < BMAX; ++q)
                if (q < bq)
                    col_fma<AMAX>(kctr + q * AMAX, q + 1 < bq ? kctr + (q + 1) * AMAX : kp, cv[q], out, kbuf);
        }
        // member slots
#pragma unroll
        for (int xo = 0; xo < AMAX; ++xo) {
            if (xo < a && !((common >> xo) & 1)) {
                if (live && !((rho >> xo) & 1)) {
                    const uint32_t zi = mz[xo] + (uint32_t)beta;
                    if (ok(zi, x.ztot)) Z[zi] = (T)out[xo];
                }
            }
        }
    }
    star_finish<T>(x, p, Z, ok);
}

// AND of a 64-bit value over the 64 lanes of a wave (all lanes active), by
// DPP row shifts and row broadcasts; the result is read from lane 63.
__device__ __forceinline__ uint32_t wave_and32(uint32_t v)
{
    // row_shr:1, 2, 4, 8 (lanes without a source keep the AND identity)
    v &= (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x111, 0xf, 0xf, false);
    v &= (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x112, 0xf, 0xf, false);
    v &= (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x114, 0xf, 0xf, false);
    v &= (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x118, 0xf, 0xf, false);
    // row_bcast:15 into rows 1 and 3, row_bcast:31 into rows 2 and 3
    v &= (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x142, 0xa, 0xf, false);
    v &= (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x143, 0xc, 0xf, false);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
__device__ __forceinline__ uint64_t wave_and64(uint64_t v)
{
    return ((uint64_t)wave_and32((uint32_t)(v >> 32)) << 32) | wave_and32((uint32_t)v);
}

// Buffer loads and stores at a 32-bit byte offset (range-checked by the
// hardware: an offset past the buffer reads 0 and drops a store).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
template <typename T>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, T v)
{
    if constexpr (sizeof(T) == 8) {
        typedef unsigned int u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)off, 0, 0);
    } else {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)off, 0, 0);
    }
}

// Column star pass, buffer-addressed (the default where the vector and every
// members' sector of Z are below kBufLimit bytes).  The same products as
// k_star_col; what differs is the addressing:
// * member values, insert offsets and slot stores go through buffer
//   resources with 32-bit byte offsets (the vector; the group's insert-table
//   rows, the row as a scalar offset; the members' Z sector).  An absent
//   member (y in rho; every member of a lane past the last column, which
//   reads the member-table row nst) has the offset kOob, so its load returns
//   0 and its slot store is dropped by the range check: no per-load address
//   selects and no exec-mask changes around the stores;
// * a lane whose centre value does not exist (t in beta, or past the last
//   column) reads and accumulates at its own entry behind the staged centres
//   (value 0, accumulator never read), so the centre reads and the LDS
//   atomics are unconditional;
// * col / Sm from a float reciprocal and one correction step (col < 2^24);
// * the orbitals in every lane's star: the AND of the A-parts of the stars
//   the wave's columns touch, a scalar loop over the first and last lane's.
template <typename T, bool kChecked, int AMAX, int BMAX, bool kExact>
__global__ __launch_bounds__(kBlock, 3) void k_star_colb(dk_tables t, dk_stars st, const T *__restrict__ c,
                                                         T *__restrict__ Z, uint32_t *bad,
                                                         const dk_sigma_io *io,
                                                         const double *__restrict__ KP)
{
    constexpr int R = AMAX + BMAX;
    extern __shared__ double lds_raw[];
    const uint32_t sg = blockIdx.x;
    if (sg >= st.nsgroup) return;
    if (io) {
        if (!io->active) return;
        c = (const T *)io->c;
    }
    StarLds p;
    star_lds<T, false>(t.L, t.N, st, &p, (char *)lds_raw);
    auto ok = [&](uint32_t idx, uint32_t lim) -> bool {
        if (!kChecked) return true;
        if (idx < lim) return true;
        atomicAdd(bad, 1u);
        return false;
    };
    const StarCtx x = star_setup<T, kChecked, kBlock, true, true>(t, st, sg, c, p, ok);
    const int a = (int)st.a, b = (int)st.b;
    const int Sm = __builtin_amdgcn_readfirstlane(x.Sm);
    const int Sc = __builtin_amdgcn_readfirstlane(x.Sc);
    const int nst = __builtin_amdgcn_readfirstlane(x.nst);
    const bool has_ctr = __builtin_amdgcn_readfirstlane((int)x.has_ctr) != 0;
    if (a > AMAX || b > BMAX || (kExact && a != AMAX)) { // uniform: the launcher picks AMAX >= a, BMAX >= b
        if (kChecked && threadIdx.x == 0) atomicAdd(bad, 1u);
        return;
    }
    const uint32_t cbytes = x.n * (uint32_t)sizeof(T), zbytes = x.zm_len * (uint32_t)sizeof(T);
    const __amdgpu_buffer_rsrc_t rc = buf_rsrc(c, cbytes);
    const __amdgpu_buffer_rsrc_t rz = buf_rsrc(Z + x.zm_sector, zbytes);
    const __amdgpu_buffer_rsrc_t ri = buf_rsrc(x.ins, has_ctr ? (uint32_t)(2 * b * Sm) : 0u);
    const int ncol = nst * Sm;
    const float inv = 1.0f / (float)Sm;
    const uint64_t amask = a < 64 ? ((1ull << a) - 1) : ~0ull;
    const int lane = (int)(threadIdx.x & 63);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const T *ctrv = (const T *)p.ctr;
    const int dummy = (int)st.selems + lane;
    for (int c0 = wave * 64; c0 < ncol; c0 += kBlock) {
        const double *kp = KP + opaque(0, c0);
        const int col = c0 + lane;
        const bool live = col < ncol;
        const int colc = live ? col : ncol - 1;
        int sq = (int)((float)colc * inv);
        int beta = colc - (int)__umul24(sq, Sm);
        if (beta < 0) {
            --sq;
            beta += Sm;
        } else if (beta >= Sm) {
            ++sq;
            beta -= Sm;
        }
        // orbitals in rho for every lane's star (AND over the wave, lanes past
        // the last column repeat its star): skipped by the whole wave
        const uint64_t rho_l = ((const uint64_t *)p.ycol)[sq];
        const uint64_t common = wave_and64(rho_l) & amask;
        if (kChecked) { // the same AND by lane shuffles
            uint32_t clo = (uint32_t)rho_l, chi = (uint32_t)(rho_l >> 32);
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                clo &= (uint32_t)__shfl_xor((int)clo, o);
                chi &= (uint32_t)__shfl_xor((int)chi, o);
            }
            if (((((uint64_t)chi << 32) | clo) & amask) != common) atomicAdd(bad, 1u);
        }
        const uint64_t ym = ~common & amask; // the wave's orbitals y
        const int s = live ? sq : nst;       // row nst: every member absent
        const uint32_t bb = live ? (uint32_t)beta * (uint32_t)sizeof(T) : 0u;
        const uint32_t *ms = p.mem_start + __umul24(s, a);
        const uint32_t *mz = p.mem_zb + __umul24(s, a);
        // member values, all loads in flight at once (an orbital common to
        // the wave has the offset kOob in every lane: no memory access)
        T v[AMAX];
#pragma unroll
        for (int y = 0; y < AMAX; ++y) {
            // (padded rows y >= a: no member, no table entry)
            const uint32_t o = (kExact || y < a) ? ms[y] + bb : kOob;
            if (kChecked && o < kOob) ok(o, cbytes);
            v[y] = buf_load<T>(rc, o);
        }
        // the centre's insert offsets lexB(beta + t), loaded while the member
        // values arrive and the A products run
        uint32_t offs[BMAX];
        if (has_ctr) {
            const int bq = opaque(b, c0), Smq = opaque(Sm, c0);
            const int vo = 2 * beta;
#pragma unroll
            for (int q = 0; q < BMAX; ++q)
                offs[q] = q < bq ? (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(ri, vo, 2 * q * Smq, 0)
                                 : 0xffffu;
        }
        double out[R];
#pragma unroll
        for (int r = 0; r < R; ++r) out[r] = 0.0;
        const double *kctr = kp + AMAX * R; // centre columns
        double kbuf[kKC];
        {
            const double *k0 = ym ? kp + (__ffsll((unsigned long long)ym) - 1) * R : kctr;
#pragma unroll
            for (int i = 0; i < kKC; ++i) kbuf[i] = k0[i];
        }
#pragma unroll
        for (int y = 0; y < AMAX; ++y) {
            if ((ym >> y) & 1) {
                const uint64_t rest = ym & ~((2ull << y) - 1);
                const double *kn = rest ? kp + (__ffsll((unsigned long long)rest) - 1) * R : kctr;
                col_fma<R>(kp + y * R, kn, (double)v[y], out, kbuf, y);
            }
        }
        if (has_ctr) {
            const int bq = opaque(b, c0);
            const int cb = (int)__umul24(s, Sc);
            int ci[BMAX];
#pragma unroll
            for (int q = 0; q < BMAX; ++q) {
                const bool o = live && offs[q] != 0xffffu && ok(offs[q], (uint32_t)Sc);
                ci[q] = o ? cb + (int)offs[q] : dummy;
            }
            double cv[BMAX];
#pragma unroll
            for (int q = 0; q < BMAX; ++q) cv[q] = (double)ctrv[ci[q]];
            double *acc = p.acc;
#pragma unroll
            for (int q = 0; q < BMAX; ++q)
                __hip_atomic_fetch_add(&acc[ci[q]], out[AMAX + q], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
            for (int q = 0; q < BMAX; ++q)
                if (q < bq)
                    col_fma<AMAX>(kctr + q * AMAX, q + 1 < bq ? kctr + (q + 1) * AMAX : kp, cv[q], out, kbuf);
        }
        // member slots (absent members: dropped by the range check)
#pragma unroll
        for (int xo = 0; xo < AMAX; ++xo) {
            if (xo < a && !((common >> xo) & 1)) {
                const uint32_t o = mz[xo] + bb;
                if (kChecked && o < kOob) ok(o, zbytes);
                buf_store<T>(rz, o, (T)out[xo]);
            }
        }
    }
    star_finish<T>(x, p, Z, ok);
}

// FP32 star pass on the matrix cores.  The column formulation of
// k_star_col (products over all of A, K independent of the star) with
// v_mfma_f32_16x16x4f32 tiles: a wave takes 16 columns at a time; the B
// operands are the columns' member values X_y (k index y = 4 kb + lane / 16)
// and centre values; the A operands are K in FP32, the same for every column,
// so they stay in registers for the whole kernel.  Each 4-term product step
// and the sum over the steps accumulate in FP32 (the slot is stored in FP32);
// the main pass sums a determinant's slots in FP64.  Lane l holds A[l % 16]
// [l / 16], B[l / 16][l % 16] and, in result register r, D[4 (l / 16) + r]
// [l % 16] (layout checked on gfx950).  YK: 4-orbital steps over A (a <= 4 YK).
constexpr int kBlockM = 512;
typedef float dk_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int m32_row(int l, int r) { return 4 * (l >> 4) + r; }

template <bool kChecked, int YK>
__global__ __launch_bounds__(kBlockM) void k_star_m32(dk_tables t, dk_stars st, const float *__restrict__ c,
                                                      float *__restrict__ Z, uint32_t *bad,
                                                      const dk_sigma_io *io)
{
    constexpr int XT = (4 * YK + 15) / 16; // row tiles of members
    constexpr int kW = kBlockM / 64;
    extern __shared__ double lds_raw[];
    const uint32_t sg = blockIdx.x;
    if (sg >= st.nsgroup) return;
    if (io) {
        if (!io->active) return;
        c = (const float *)io->c;
    }
    StarLds p;
    star_lds<float>(t.L, t.N, st, &p, (char *)lds_raw);
    auto ok = [&](uint32_t idx, uint32_t lim) -> bool {
        if (!kChecked) return true;
        if (idx < lim) return true;
        atomicAdd(bad, 1u);
        return false;
    };
    const StarCtx x = star_setup<float, kChecked, kBlockM>(t, st, sg, c, p, ok);
    const int L = (int)t.L, a = (int)st.a, b = (int)st.b;
    const int ny = __builtin_amdgcn_readfirstlane(x.ny);
    const int Sm = __builtin_amdgcn_readfirstlane(x.Sm);
    const int Sc = __builtin_amdgcn_readfirstlane(x.Sc);
    const int nst = __builtin_amdgcn_readfirstlane(x.nst);
    const bool has_ctr = __builtin_amdgcn_readfirstlane((int)x.has_ctr) != 0;
    if (a > 4 * YK || b > 16) { // uniform: the launcher picks 4 YK >= a
        if (kChecked && threadIdx.x == 0) atomicAdd(bad, 1u);
        return;
    }
    const int lane = (int)(threadIdx.x & 63), li = lane & 15, lk = lane >> 4;
    // A operands: K in FP32 (zero outside A x A, B x A, A x B and on the diagonal)
    float ax[XT][YK], axc[XT][4], at[YK];
#pragma unroll
    for (int xt = 0; xt < XT; ++xt) {
        const int xo = 16 * xt + li;
#pragma unroll
        for (int kb = 0; kb < YK; ++kb) {
            const int y = 4 * kb + lk;
            ax[xt][kb] = (xo < a && y < a && xo != y) ? (float)p.K[xo * L + y] : 0.f;
        }
#pragma unroll
        for (int tb = 0; tb < 4; ++tb) {
            const int tt = 4 * tb + lk;
            axc[xt][tb] = (xo < a && tt < b) ? (float)p.K[xo * L + a + tt] : 0.f;
        }
    }
#pragma unroll
    for (int kb = 0; kb < YK; ++kb) {
        const int y = 4 * kb + lk;
        at[kb] = (li < b && y < a) ? (float)p.K[(a + li) * L + y] : 0.f;
    }
    const uint32_t first = st.sgroup[2 * sg];
    const int ncol = nst * Sm, nunit = (ncol + 15) / 16;
    const uint64_t amask = a < 64 ? ((1ull << a) - 1) : ~0ull;
    const float *ctrv = (const float *)p.ctr;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    for (int u = wave; u < nunit; u += kW) {
        const int col = 16 * u + li;
        const bool live = col < ncol;
        const int s = live ? col / Sm : 0;
        const int beta = live ? col - s * Sm : 0;
        const uint64_t rho = live ? st.star_amask[first + s] : amask;
        const uint64_t rhox = rho | ~amask; // orbitals without a member
        const uint32_t *ms = p.mem_start + s * ny;
        const uint32_t *mz = p.mem_zb + s * ny;
        // B operands: member values of this lane's column at y = 4 kb + lane / 16
        float xb[YK];
#pragma unroll
        for (int kb = 0; kb < YK; ++kb) {
            const int y = 4 * kb + lk;
            xb[kb] = 0.f;
            if (live && !((rhox >> y) & 1)) {
                const int yi = y - __popcll(rho & ((1ull << y) - 1));
                const uint32_t gi = ms[yi] + (uint32_t)beta;
                if (ok(gi, x.n)) xb[kb] = c[gi];
            }
        }
        // centre values at lexB(beta + t), t = 4 tb + lane / 16
        float cb[4];
#pragma unroll
        for (int tb = 0; tb < 4; ++tb) {
            const int tt = 4 * tb + lk;
            cb[tb] = 0.f;
            if (has_ctr && live && tt < b) {
                const uint32_t off = x.ins[(size_t)tt * Sm + beta];
                if (off != 0xffffu && ok(off, (uint32_t)Sc)) cb[tb] = ctrv[s * Sc + off];
            }
        }
        dk_f4 dx[XT];
#pragma unroll
        for (int xt = 0; xt < XT; ++xt) {
            dx[xt] = dk_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < YK; ++kb)
                dx[xt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ax[xt][kb], xb[kb], dx[xt], 0, 0, 0);
#pragma unroll
            for (int tb = 0; tb < 4; ++tb)
                dx[xt] = __builtin_amdgcn_mfma_f32_16x16x4f32(axc[xt][tb], cb[tb], dx[xt], 0, 0, 0);
        }
        // member slots: this lane's column, rows m32_row(lane, r) of each tile
#pragma unroll
        for (int xt = 0; xt < XT; ++xt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int xo = 16 * xt + m32_row(lane, r);
                if (live && !((rhox >> xo) & 1)) {
                    const int xi = xo - __popcll(rho & ((1ull << xo) - 1));
                    const uint32_t zi = mz[xi] + (uint32_t)beta;
                    if (ok(zi, x.ztot)) Z[zi] = dx[xt][r];
                }
            }
        // the centre's hops from B orbital t into A, into its accumulator
        if (has_ctr) {
            dk_f4 dt = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < YK; ++kb)
                dt = __builtin_amdgcn_mfma_f32_16x16x4f32(at[kb], xb[kb], dt, 0, 0, 0);
            double *cacc = p.acc + s * Sc;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int tt = m32_row(lane, r);
                if (live && tt < b) {
                    const uint32_t off = x.ins[(size_t)tt * Sm + beta];
                    if (off != 0xffffu && ok(off, (uint32_t)Sc))
                        __hip_atomic_fetch_add(&cacc[off], (double)dt[r], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
        }
    }
    star_finish<float, kBlockM>(x, p, Z, ok);
}

// Main pass over one set of SPL slabs of a group, lane per offset bi: the
// hop table entry of (h, bi) and its K are read once for all SPL slabs, whose
// values sit side by side in LDS (own[(set * S + bi) * main_stride(SPL) + u],
// FP64; the stride is padded to an odd number of doubles for SPL > 1, so that
// the lanes' random offsets J spread over the banks as for a single slab), so
// a hop costs one table load, one K read and an address per set plus one LDS
// read and one FMA per slab.  Per determinant the sums are added in the same
// order as for a single slab: ((D - shift) c + hops in table order) + slots.
__host__ __device__ constexpr int main_stride(int spl) { return spl == 1 ? 1 : spl + 1; }
// slabs per set of a main-pass group of nsl slabs (groups hold a multiple of
// 4 or 2 slabs except at the end of a sector)
__host__ __device__ constexpr int main_spl(int nsl) { return (nsl & 3) == 0 ? 4 : (nsl & 1) == 0 ? 2 : 1; }

template <int SPL, typename T, bool kChecked, typename OK>
__device__ __forceinline__ void main_sets(const double *__restrict__ D, double shift,
                                          const T *__restrict__ Z, T *__restrict__ sigma,
                                          const T *own, const uint32_t *gstart,
                                          const uint32_t *gzb, const char *Kbb,
                                          __amdgpu_buffer_rsrc_t hrsrc, int S, int nsets, int ns,
                                          int nhp, uint32_t n, uint32_t ztot, uint32_t kbytes, OK &&ok)
{
    constexpr int kZS = SPL == 4 ? 8 : 12; // slots per slab loaded before the hops
    constexpr int kHB = 8;                  // hop-table rows loaded together
    for (int w = threadIdx.x; w < nsets * S; w += kBlock) {
        const int set = w / S, bi = w - set * S;
        constexpr int STR = main_stride(SPL);
        const T *ob = own + (size_t)set * S * STR;
        // the first hop-table rows, then the slots and the diagonal (HBM),
        // all in flight together; each later chunk of rows is loaded while
        // the one before is used
        const uint32_t voff = (uint32_t)bi * 4u;
        uint32_t hv[kHB];
#pragma unroll
        for (int i = 0; i < kHB; ++i)
            hv[i] = __builtin_amdgcn_raw_buffer_load_b32(hrsrc, (int)voff, i * S * 4, 0);
        uint32_t I[SPL];
        double dI[SPL];
        T zv[SPL][kZS];
#pragma unroll
        for (int u = 0; u < SPL; ++u) {
            const int s = set * SPL + u;
            I[u] = gstart[s] + (uint32_t)bi;
            const uint32_t zb = gzb[s] + (uint32_t)bi;
            // loads without per-slot branches, so that all of them are in
            // flight together: slots past ns read the last one again, a slab
            // without slots (a = 0) reads Z[0] (every slot buffer has an
            // entry), and neither is summed
#pragma unroll
            for (int q = 0; q < kZS; ++q) {
                const uint32_t zi = ns > 0 ? zb + (uint32_t)((q < ns ? q : ns - 1) * S) : 0u;
                zv[u][q] = (ns == 0 || ok(zi, ztot)) ? Z[zi] : (T)0;
            }
            dI[u] = ok(I[u], n) ? D[I[u]] : 0.0;
        }
        double acc[SPL];
#pragma unroll
        for (int u = 0; u < SPL; ++u) acc[u] = 0.0;
        for (int h0 = 0; h0 < nhp; h0 += kHB) {
            uint32_t hn[kHB];
            if (h0 + kHB < nhp) { // uniform
#pragma unroll
                for (int i = 0; i < kHB; ++i)
                    hn[i] = __builtin_amdgcn_raw_buffer_load_b32(hrsrc, (int)voff, (h0 + kHB + i) * S * 4, 0);
            }
            // four hops at a time: every LDS read issued before the products
#pragma unroll
            for (int i0 = 0; i0 < kHB; i0 += 4) {
                double Kv[4];
                T ov[4][SPL];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    // J, and the byte offset of K in Kbb (which starts the LDS)
                    const uint32_t J = hv[i0 + i] & 0xffffu, kb = hv[i0 + i] >> 16;
                    const bool g = ok(J, (uint32_t)S) && ok(kb, kbytes);
                    const T *o = (const T *)((const char *)ob + (g ? J : 0u) * ((uint32_t)sizeof(T) * STR));
                    Kv[i] = g ? *(const double *)(Kbb + kb) : 0.0;
#pragma unroll
                    for (int u = 0; u < SPL; ++u) ov[i][u] = o[u];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int u = 0; u < SPL; ++u) acc[u] += Kv[i] * (double)ov[i][u];
            }
#pragma unroll
            for (int i = 0; i < kHB; ++i) hv[i] = hn[i];
        }
#pragma unroll
        for (int u = 0; u < SPL; ++u) {
            const int s = set * SPL + u;
            double zs = 0;
#pragma unroll
            for (int q = 0; q < kZS; ++q) zs += q < ns ? (double)zv[u][q] : 0.0;
            const uint32_t zb = gzb[s] + (uint32_t)bi;
            for (int q = kZS; q < ns; ++q) {
                const uint32_t zi = zb + (uint32_t)(q * S);
                if (ok(zi, ztot)) zs += (double)Z[zi];
            }
            if (ok(I[u], n)) sigma[I[u]] = (T)(((dI[u] - shift) * (double)ob[bi * STR + u] + acc[u]) + zs);
        }
    }
}

template <typename T, bool kChecked>
__global__ __launch_bounds__(kBlock) void k_star_main(dk_tables t, dk_stars st,
                                                      const double *__restrict__ D, double shift,
                                                      const T *__restrict__ c,
                                                      const T *__restrict__ Z,
                                                      T *__restrict__ sigma, uint32_t *bad,
                                                      const dk_sigma_io *io)
{
    extern __shared__ double lds_raw[];
    const uint32_t g = blockIdx.x;
    if (g >= st.ngroup) return;
    if (io) {
        if (!io->active) return;
        c = (const T *)io->c;
        sigma = (T *)io->sigma;
        shift = io->shift;
    }
    const int L = (int)t.L, N = (int)t.N, a = (int)st.a, b = (int)st.b;
    const uint32_t n = t.n, ztot = st.zoff[a + 1];
    auto ok = [&](uint32_t idx, uint32_t lim) -> bool {
        if (!kChecked) return true;
        if (idx < lim) return true;
        atomicAdd(bad, 1u);
        return false;
    };
    double *Kbb = lds_raw; // Kbb[o * b + t] = K between B orbitals o and t; Kbb[b * b] = 0
    uint32_t *bt = (uint32_t *)((char *)lds_raw + a8(sizeof(double) * (b * b + 1)));
    T *own = (T *)((char *)bt + a8(sizeof(uint32_t) * (b + 1) * (N + 2)));
    // the group's slab starts and slot bases
    uint32_t *gstart = (uint32_t *)((char *)own + a8(sizeof(T) * st.owns));
    uint32_t *gzb = gstart + a8(sizeof(uint32_t) * st.gslabs) / sizeof(uint32_t);
    for (int i = threadIdx.x; i <= b * b; i += kBlock) {
        const int o = i / b, tt = i - o * b;
        Kbb[i] = i < b * b ? t.K[(a + o) * L + a + tt] : 0.0;
    }
    for (int i = threadIdx.x; i < (b + 1) * (N + 2); i += kBlock) bt[i] = t.binom[i];
    const uint32_t first = st.group[2 * g], w = st.group[2 * g + 1];
    const int nsl = (int)(w & 0xffffu), m = (int)(w >> 16), k = N - m;
    if (kChecked && (uint32_t)nsl > st.gslabs) {
        if (threadIdx.x == 0) atomicAdd(bad, 1u);
        return;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nsl; i += kBlock) {
        gstart[i] = st.slab_start[first + i];
        gzb[i] = st.slab_zbase[first + i];
    }
    __syncthreads();
    const dk::Binom C{bt, N + 2};
    const int S = (int)C(b, m), nel = nsl * S, ns = nslot_of(k, m, a);
    const int spl = main_spl(nsl), str = main_stride(spl);
    const int lsp = spl == 4 ? 2 : spl == 2 ? 1 : 0;
    if (kChecked && (uint32_t)(nel / spl * str) > st.owns) {
        if (threadIdx.x == 0) atomicAdd(bad, 1u);
        return;
    }
    // stage the slabs as FP64, the values of a set side by side (element e:
    // u = e % spl, then bi, then set; at r * str + u); eight loads in flight,
    // issued without branches (elements past nel load the group's first
    // value and are not stored)
    constexpr int kSL = 8;
    for (int e0 = threadIdx.x; e0 < nel; e0 += kSL * kBlock) {
        T v[kSL];
#pragma unroll
        for (int q = 0; q < kSL; ++q) {
            const int e = e0 + q * kBlock;
            const int ec = e < nel ? e : 0;
            const int u = ec & (spl - 1), r = ec >> lsp, set = r / S, bi = r - set * S;
            const uint32_t gi = gstart[set * spl + u] + (uint32_t)bi;
            v[q] = ok(gi, n) ? c[gi] : (T)0;
        }
#pragma unroll
        for (int q = 0; q < kSL; ++q) {
            const int e = e0 + q * kBlock;
            if (e < nel) own[(e >> lsp) * str + (e & (spl - 1))] = v[q];
        }
    }
    __syncthreads();
    constexpr int kHB = 8;
    const int nhp = (m * (b - m) + kHB - 1) / kHB * kHB;
    // the hop table of m-pair masks, nhp rows of S entries (loads past it return 0)
    const __amdgpu_buffer_rsrc_t hrsrc = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(st.hop + st.hop_off[m]), (short)0, nhp * S * 4, 0x00020000);
    const int nsets = nsl >> lsp;
    const char *Kb = (const char *)Kbb;
    const uint32_t kbytes = (uint32_t)(8 * (b * b + 1));
    if (spl == 4)
        main_sets<4, T, kChecked>(D, shift, Z, sigma, own, gstart, gzb, Kb, hrsrc, S, nsets, ns, nhp,
                                  n, ztot, kbytes, ok);
    else if (spl == 2)
        main_sets<2, T, kChecked>(D, shift, Z, sigma, own, gstart, gzb, Kb, hrsrc, S, nsets, ns, nhp,
                                  n, ztot, kbytes, ok);
    else
        main_sets<1, T, kChecked>(D, shift, Z, sigma, own, gstart, gzb, Kb, hrsrc, S, nsets, ns, nhp,
                                  n, ztot, kbytes, ok);
}

} // namespace

// Column star pass at the exact row counts AMAX = a, BMAX = 13 of the default
// star splits (a = 12..21, every shape from L = 25 to L = 34), instead of a
// padded to a multiple of 8 and b to 16: fewer FMAs, loads, stores and
// registers.  Instantiated in sigma_star_exact_lo.hip (a = 12..16) and
// sigma_star_exact_hi.hip (a = 17..21), separate translation units that
// compile in parallel.  Each launches k_kpack and k_star_colb (buf != 0) or
// k_star_col on stream s and returns 0, or returns -1 when it has no
// instantiation for (a, b).
__host__ __device__ constexpr bool col_exact_shape(int a, int b) { return b == 13 && a >= 12 && a <= 21; }
int dk_col_exact_lo(int prec, int checked, int buf, int a, int b, dim3 grid, size_t lds, hipStream_t s,
                    const dk_tables &t, const dk_stars &st, const void *c, void *Z, uint32_t *bad,
                    const dk_sigma_io *io, double *kp);
int dk_col_exact_hi(int prec, int checked, int buf, int a, int b, dim3 grid, size_t lds, hipStream_t s,
                    const dk_tables &t, const dk_stars &st, const void *c, void *Z, uint32_t *bad,
                    const dk_sigma_io *io, double *kp);
