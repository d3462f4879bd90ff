// batch.hip -- many independent DOCI Davidson solves in lock step.
//
// The Davidson iteration of src/davidson.c:72-99 is split into streaming
// steps (one workgroup per tile = a row block of one instance; all
// instances in one launch) and small steps (one 64-lane workgroup per
// instance) that reduce the tile partials in a fixed order and advance that
// instance's state:
//
//   small A0  (instances just started) loop test with the initial residual
//   ORTHO     flagged instances only: t -= sum g_j T_j (explicit projection
//             of a nearly dependent correction) or |t| of a first t, with
//             the norm and overlaps
//   small B   after ORTHO: t joins the basis (its scale and column of R)
//   sigma     W_m = (H - shift) T_m
//   SCOL      new column T_j^T W_m of the projected matrix
//   small D   bordered update of the projected eigenproblem (cyclic Jacobi,
//             parallel ordering), lowest Ritz pair; deflation plan when the
//             basis reaches max_vec (src/davidson.c:83-89)
//   DEFLATE   T'_k = sum_j (R^-1 Y)_jk T_j, W'_k likewise (flagged instances only)
//   small E   basis = the kept Ritz vectors
//   RESID     x = T y, r = W y - theta x, t = r / |D - E| with the 1e-12
//             floor (src/davidson.c:214-227), x written in FP64
//   small A   residual norm, loop test `res > tol && its++ < max_its`;
//             the new correction t joins the basis T as it is (its
//             overlaps with the basis from RESID give the new column of R)
//             unless it is nearly dependent on it, which flags ORTHO
//
// The basis T holds the raw corrections (src/davidson.c:133-176
// orthogonalises each one; the same subspace here), factored T = Q R with
// Q orthonormal and never formed: R comes from the overlaps RESID already
// computes, and the projected matrix is kept in Q, bordered each iteration
// with the new column of T^T (H - shift) T transformed by R.  This spares
// the Gram-Schmidt pass over the basis in every iteration.
//
// Every partial sum is written and reduced within one iteration's launch
// sequence, so the tile layout may change between iterations (instances
// joining or leaving).
// Every launch reads each instance's state from device memory, so the same
// launch sequence serves every iteration and can be captured in a HIP graph.
// Vectors are stored unnormalised with a scale per basis vector (V_j =
// s_j * stored_j), which saves a normalisation pass.
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>

#include "../dk.h"
#include "addressing.hpp"

namespace {

constexpr int kBlock = 256;
static_assert(kBlock == DK_ADMIT_ROWS, "one admission / restart row per thread");
constexpr int kWaves = kBlock / 64;
constexpr int kSmall = 64;
constexpr int LDY = DK_BMAX_VEC;

template <typename T>
__device__ __forceinline__ T *slot_ptr(const dk_binst &in, int slot)
{
    return (T *)(in.slots + (uint64_t)slot * in.slot_stride);
}

template <int NV>
__device__ __forceinline__ void tile_partials(double (&acc)[NV], int nv, double *out)
{
    __shared__ double sh[NV][kWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        if (k < nv) {
            double v = acc[k];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0) sh[k][wave] = v;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nv; k += kBlock) {
        double s = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) s += sh[k][w];
        out[k] = s;
    }
}

// ---- streaming steps -------------------------------------------------------
// Each lane streams VW consecutive rows per step with one vector load per
// vector (16 bytes, narrower when M per-row values are kept); tiles start at
// multiples of 512 rows and slots are 256-byte aligned, so only an
// instance's last tile has a scalar tail.
template <typename T, int VW> struct Vec {
    typedef T type __attribute__((ext_vector_type(VW)));
};
template <typename T> struct Vec<T, 1> {
    typedef T type;
};
template <typename T, int VW>
__device__ __forceinline__ double el(const typename Vec<T, VW>::type &v, int e)
{
    if constexpr (VW == 1) return (double)v;
    else return (double)v[e];
}
template <typename T, int VW>
__device__ __forceinline__ void set_el(typename Vec<T, VW>::type &v, int e, T x)
{
    if constexpr (VW == 1) v = x;
    else v[e] = x;
}
template <typename T, int VW>
__device__ __forceinline__ typename Vec<T, VW>::type ldv(const T *p, uint32_t q)
{
    return ((const typename Vec<T, VW>::type *)p)[q];
}
template <typename T, int VW>
__device__ __forceinline__ void stv(T *p, uint32_t q, const typename Vec<T, VW>::type &v)
{
    ((typename Vec<T, VW>::type *)p)[q] = v;
}
template <typename T, int M>
constexpr int vw_for()
{
    constexpr int w = 16 / (int)sizeof(T);
    return M <= 16 ? w : (M <= 32 ? (w >= 2 ? w / 2 : 1) : (w >= 4 ? w / 4 : 1));
}

// Vectors are loaded in groups of kGrp, all loads of a group issued
// together: a group past the instance's m is skipped (a uniform branch per
// group), slots of a group past m repeat the instance's last vector with
// coefficient 0 (see k_bstream).
constexpr int kGrp = 4;

// Rows [r0, r1) of one tile: steps of VW rows (chunk index q = row / VW),
// then the scalar tail.
#define DK_ROWS(STEP)                                                                              \
    do {                                                                                           \
        const uint32_t rv = r0 + (r1 - r0) / VW * VW;                                              \
        for (uint32_t q = r0 / VW + threadIdx.x; q < rv / VW; q += kBlock) STEP(VW, q);            \
        for (uint32_t q = rv + threadIdx.x; q < r1; q += kBlock) STEP(1, q);                       \
    } while (0)

template <typename T, int M, int W>
__device__ __forceinline__ void ortho_rows(const T *const (&V)[M], int m, const double (&g)[M], T *t,
                                           uint32_t q, double (&acc)[M + 1])
{
    typedef typename Vec<T, W>::type VT;
    VT xv[M];
    VT tv = ldv<T, W>(t, q);
    double td[W];
#pragma unroll
    for (int e = 0; e < W; ++e) td[e] = el<T, W>(tv, e);
#pragma unroll
    for (int j0 = 0; j0 < M; j0 += kGrp)
        if (j0 < m) {
#pragma unroll
            for (int j = j0; j < j0 + kGrp; ++j) xv[j] = ldv<T, W>(V[j], q);
#pragma unroll
            for (int j = j0; j < j0 + kGrp; ++j)
#pragma unroll
                for (int e = 0; e < W; ++e) td[e] -= g[j] * el<T, W>(xv[j], e);
        }
#pragma unroll
    for (int e = 0; e < W; ++e) {
        set_el<T, W>(tv, e, (T)td[e]);
        td[e] = el<T, W>(tv, e);
        acc[0] += td[e] * td[e];
    }
    if (m > 0) stv<T, W>(t, q, tv);
#pragma unroll
    for (int j0 = 0; j0 < M; j0 += kGrp)
        if (j0 < m)
#pragma unroll
            for (int j = j0; j < j0 + kGrp; ++j)
#pragma unroll
                for (int e = 0; e < W; ++e) acc[1 + j] += el<T, W>(xv[j], e) * td[e];
}

template <typename T, int M, int W>
__device__ __forceinline__ void scol_rows(const T *const (&Vp)[M + 1], int m, const T *Wm, uint32_t q,
                                          double (&acc)[M + 1])
{
    typedef typename Vec<T, W>::type VT;
    const VT wv = ldv<T, W>(Wm, q);
    {
        const VT vv = ldv<T, W>(Vp[0], q);
#pragma unroll
        for (int e = 0; e < W; ++e) acc[0] += el<T, W>(vv, e) * el<T, W>(wv, e);
    }
#pragma unroll
    for (int j0 = 1; j0 <= M; j0 += kGrp)
        if (j0 <= m) {
            VT vv[kGrp];
#pragma unroll
            for (int j = 0; j < kGrp; ++j) vv[j] = ldv<T, W>(Vp[j0 + j], q);
#pragma unroll
            for (int j = 0; j < kGrp; ++j)
#pragma unroll
                for (int e = 0; e < W; ++e) acc[j0 + j] += el<T, W>(vv[j], e) * el<T, W>(wv, e);
        }
}

template <typename T, int M, int W>
__device__ __forceinline__ void deflate_rows(const T *const (&V)[M], const T *const (&Wv)[M], int m,
                                             const double (&c)[M], T *dv, T *dw, uint32_t q)
{
    typedef typename Vec<T, W>::type VT;
    double a[W], b[W];
#pragma unroll
    for (int e = 0; e < W; ++e) a[e] = b[e] = 0;
#pragma unroll
    for (int j0 = 0; j0 < M; j0 += kGrp)
        if (j0 < m) {
            VT vv[kGrp], wv[kGrp];
#pragma unroll
            for (int j = 0; j < kGrp; ++j) {
                vv[j] = ldv<T, W>(V[j0 + j], q);
                wv[j] = ldv<T, W>(Wv[j0 + j], q);
            }
#pragma unroll
            for (int j = 0; j < kGrp; ++j)
#pragma unroll
                for (int e = 0; e < W; ++e) {
                    a[e] += c[j0 + j] * el<T, W>(vv[j], e);
                    b[e] += c[j0 + j] * el<T, W>(wv[j], e);
                }
        }
    VT ov, ow;
#pragma unroll
    for (int e = 0; e < W; ++e) {
        set_el<T, W>(ov, e, (T)a[e]);
        set_el<T, W>(ow, e, (T)b[e]);
    }
    stv<T, W>(dv, q, ov);
    stv<T, W>(dw, q, ow);
}

template <typename T, int M, int W>
__device__ __forceinline__ void resid_rows(const T *const (&V)[M], const T *const (&Wv)[M], int m,
                                           const double (&yc)[M], double theta, double E,
                                           const double *D, T *t, double *x, uint32_t q,
                                           double (&acc)[M + 2])
{
    typedef typename Vec<T, W>::type VT;
    VT vv[M];
    double xs[W], ws[W];
#pragma unroll
    for (int e = 0; e < W; ++e) xs[e] = ws[e] = 0;
#pragma unroll
    for (int j0 = 0; j0 < M; j0 += kGrp)
        if (j0 < m) {
            VT wv[kGrp];
#pragma unroll
            for (int j = j0; j < j0 + kGrp; ++j) {
                vv[j] = ldv<T, W>(V[j], q);
                wv[j - j0] = ldv<T, W>(Wv[j], q);
            }
#pragma unroll
            for (int j = j0; j < j0 + kGrp; ++j)
#pragma unroll
                for (int e = 0; e < W; ++e) {
                    xs[e] += yc[j] * el<T, W>(vv[j], e);
                    ws[e] += yc[j] * el<T, W>(wv[j - j0], e);
                }
        }
    VT tv;
    double td[W];
#pragma unroll
    for (int e = 0; e < W; ++e) {
        const uint32_t i = q * W + e;
        const double r = ws[e] - theta * xs[e];
        const double d = fabs(D[i] - E);
        const double tval = d > 1e-12 ? r / d : r / 1e-12;
        set_el<T, W>(tv, e, (T)tval);
        td[e] = el<T, W>(tv, e);
        if (x) x[i] = xs[e];
        acc[0] += r * r;
        acc[1] += td[e] * td[e];
    }
    stv<T, W>(t, q, tv);
#pragma unroll
    for (int j0 = 0; j0 < M; j0 += kGrp)
        if (j0 < m)
#pragma unroll
            for (int j = j0; j < j0 + kGrp; ++j)
#pragma unroll
                for (int e = 0; e < W; ++e) acc[2 + j] += el<T, W>(vv[j], e) * td[e];
}

template <typename T, int M, int OP>
__global__ __launch_bounds__(kBlock) void k_bstream(const dk_binst *__restrict__ insts,
                                                    dk_bstate *__restrict__ states,
                                                    const uint32_t *__restrict__ tiles,
                                                    double *__restrict__ part)
{
    const uint32_t ii = tiles[2 * blockIdx.x], ti = tiles[2 * blockIdx.x + 1];
    const dk_binst in = insts[ii];
    const dk_bstate &st = states[ii];
    if (st.st.status != DK_B_ACTIVE || ti >= in.ntiles) return; // ti = 0xffffffff: padding
    const int flags = st.st.flags;
    // explicit projection (and the norm of a first t): flagged instances only
    if ((OP == DK_BOP_ORTHO || OP == DK_BOP_ORTHO2) && !(flags & (DK_BF_ORTHO | DK_BF_FIRST))) return;
    if (OP == DK_BOP_DEFLATE && !(flags & DK_BF_DEFLATE)) return;
    const int m = st.st.m;
    const uint32_t r0 = ti * in.tile_rows;
    const uint32_t r1 = min(in.n, r0 + in.tile_rows);
    double *out = part + (size_t)(in.tile0 + ti) * DK_BPV; // the instance's rows, in tile order

    // slots past m repeat the last basis vector (t itself while the basis is
    // empty) with coefficient 0, so that every load is unconditional; their
    // sums are not stored
    const T *V[M];
#pragma unroll
    for (int j = 0; j < M; ++j)
        V[j] = m > 0 ? slot_ptr<T>(in, st.vslot[j < m ? j : m - 1]) : slot_ptr<T>(in, st.tslot);

    if constexpr (OP == DK_BOP_ORTHO || OP == DK_BOP_ORTHO2) {
        constexpr int VW = vw_for<T, M>();
        T *t = slot_ptr<T>(in, st.tslot);
        double g[M];
#pragma unroll
        for (int j = 0; j < M; ++j) g[j] = j < m ? st.g[j] : 0.0;
        double acc[M + 1];
#pragma unroll
        for (int k = 0; k <= M; ++k) acc[k] = 0;
#define DK_STEP(W_, q_) ortho_rows<T, M, W_>(V, m, g, t, q_, acc)
        DK_ROWS(DK_STEP);
#undef DK_STEP
        tile_partials<M + 1>(acc, m + 1, out);
    } else if constexpr (OP == DK_BOP_SCOL) {
        constexpr int VW = 16 / (int)sizeof(T);
        // column m: V_0..V_m (V_m is t itself) against W_m, the sigma just computed
        const T *Vp[M + 1];
#pragma unroll
        for (int j = 0; j <= M; ++j) Vp[j] = slot_ptr<T>(in, st.vslot[j <= m ? j : m]);
        const T *Wm = slot_ptr<T>(in, st.wslot[m]);
        double acc[M + 1];
#pragma unroll
        for (int k = 0; k <= M; ++k) acc[k] = 0;
#define DK_STEP(W_, q_) scol_rows<T, M, W_>(Vp, m, Wm, q_, acc)
        DK_ROWS(DK_STEP);
#undef DK_STEP
        tile_partials<M + 1>(acc, m + 1, out);
    } else if constexpr (OP == DK_BOP_DEFLATE) {
        constexpr int VW = vw_for<T, M>();
        const int keep = st.keep;
        const T *W[M];
#pragma unroll
        for (int j = 0; j < M; ++j) W[j] = slot_ptr<T>(in, st.wslot[j < m ? j : m - 1]);
        for (int k = 0; k < keep; ++k) {
            double c[M];
#pragma unroll
            for (int j = 0; j < M; ++j) c[j] = j < m ? st.defl[k][j] : 0.0;
            T *dv = slot_ptr<T>(in, st.dv[k]);
            T *dw = slot_ptr<T>(in, st.dw[k]);
#define DK_STEP(W_, q_) deflate_rows<T, M, W_>(V, W, m, c, dv, dw, q_)
            DK_ROWS(DK_STEP);
#undef DK_STEP
        }
    } else { // DK_BOP_RESID
        constexpr int VW = vw_for<T, M>();
        T *t = slot_ptr<T>(in, st.tslot);
        const double E = st.theta + st.shift, theta = st.theta;
        double yc[M];
        const T *W[M];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            yc[j] = j < m ? st.y[j] : 0.0;
            W[j] = slot_ptr<T>(in, st.wslot[j < m ? j : m - 1]);
        }
        double acc[M + 2];
#pragma unroll
        for (int k = 0; k < M + 2; ++k) acc[k] = 0;
#define DK_STEP(W_, q_) resid_rows<T, M, W_>(V, W, m, yc, theta, E, in.D, t, in.x, q_, acc)
        DK_ROWS(DK_STEP);
#undef DK_STEP
        tile_partials<M + 2>(acc, m + 2, out);
    }
}
#undef DK_ROWS

// ---- batched lane sigma ----------------------------------------------------
// hops whose gathers are issued together in the separable walk
constexpr int kHopGroup = 4;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bsigma_lane(const dk_binst *__restrict__ insts,
                                                        const dk_bstate *__restrict__ states,
                                                        const uint32_t *__restrict__ tiles)
{
    extern __shared__ double lds[];
    const uint32_t ii = tiles[2 * blockIdx.x], ti = tiles[2 * blockIdx.x + 1];
    const dk_binst in = insts[ii];
    const dk_bstate &st = states[ii];
    if (st.st.status != DK_B_ACTIVE || !in.lane_sigma || ti >= in.ntiles) return;
    const int L = (int)in.L, N = (int)in.N, N2 = N + 2;
    double *K = lds;
    uint32_t *bt = (uint32_t *)(K + L * L);
    for (int i = threadIdx.x; i < L * L; i += kBlock) K[i] = in.K[i];
    for (int i = threadIdx.x; i < (L + 1) * N2; i += kBlock) bt[i] = in.binom[i];
    __syncthreads();
    const T *c = slot_ptr<T>(in, st.sin);
    T *sigma = slot_ptr<T>(in, st.sout);
    const double shift = st.shift;
    const dk::Binom C{bt, N2};
    const uint32_t r0 = ti * in.tile_rows;
    const uint32_t r1 = min(in.n, r0 + in.tile_rows);
    // gathers through a buffer resource: 32-bit offsets, and the discarded
    // steps of a group need no in-range index
    const __amdgpu_buffer_rsrc_t rc = dk::vec_rsrc(c, (uint64_t)in.n * sizeof(T));
    for (uint32_t I = r0 + threadIdx.x; I < r1; I += kBlock) {
        const uint64_t msk = dk::unrank(I, L, N, C);
        double acc = (in.D[I] - shift) * (double)c[I];
        if (N <= 12 && L - N <= 16) { // uniform: one instance per workgroup
            dk::for_each_hop_sep<12, 16, kHopGroup>(I, msk, L, N, C,
                                                    [&](int o, int tt, uint32_t Jr, bool hop) {
                const double v = K[o * L + tt] * dk::vec_load<T>(rc, Jr);
                acc += hop ? v : 0.0;
            });
        } else {
            // branch-free walk: a step whose t is occupied reads c[I] and adds 0
            dk::for_each_hop_pred(I, msk, L, N, C, [&](int o, int tt, uint32_t Jr, bool hop) {
                const double v = K[o * L + tt] * (double)c[hop ? Jr : I];
                acc += hop ? v : 0.0;
            });
        }
        sigma[I] = (T)acc;
    }
}

// ---- small steps -------------------------------------------------------
// Sum value k over the instance's tiles, k < nv, in a fixed order.
__device__ void reduce_tiles(const dk_binst &in, const double *__restrict__ part, int nv,
                             double *res /* LDS, DK_BPV */)
{
    __shared__ double sh[DK_BPV][kSmall + 1];
    double acc[DK_BPV];
#pragma unroll
    for (int k = 0; k < DK_BPV; ++k) acc[k] = 0;
    // whole rows, loaded without per-value branches so that they are in
    // flight together (values k >= nv of a row are never written by the
    // streaming passes; their sums are not read)
    typedef double d2 __attribute__((ext_vector_type(2)));
    static_assert(DK_BPV % 2 == 0, "rows of 16-byte pairs");
    for (uint32_t t = threadIdx.x; t < in.ntiles; t += kSmall) {
        const d2 *p = (const d2 *)(part + (size_t)(in.tile0 + t) * DK_BPV);
        d2 v[DK_BPV / 2];
#pragma unroll
        for (int k = 0; k < DK_BPV / 2; ++k) v[k] = p[k];
#pragma unroll
        for (int k = 0; k < DK_BPV / 2; ++k) {
            acc[2 * k] += v[k].x;
            acc[2 * k + 1] += v[k].y;
        }
    }
#pragma unroll
    for (int k = 0; k < DK_BPV; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    if (threadIdx.x < nv) {
        double s = 0;
        for (int q = 0; q < kSmall; ++q) s += sh[threadIdx.x][q];
        res[threadIdx.x] = s;
    }
    __syncthreads();
}

// lowest free slot of the pool [lo, lo + count); bit k of `used` = slot lo + k
__device__ int take_slot(uint64_t &used, int lo, int count)
{
    for (int k = 0; k < count; ++k)
        if (!((used >> k) & 1)) {
            used |= 1ull << k;
            return lo + k;
        }
    return -1;
}

// Cyclic Jacobi on the leading n x n block of A (row-major, stride LDY),
// rotations accumulated into Y; pairs of a sweep run in parallel in the
// round-robin (tournament) order, so the disjoint rotations of one step
// commute.  A pair is rotated only while |a_pq| > eps sqrt(|a_pp a_qq|)
// (below that its effect on the eigenvalues is under the rounding of the
// diagonal); stops after a sweep without a rotation, or when the
// off-diagonal mass is below 1e-32 of the total.  (Without the threshold,
// off-diagonal entries at the rounding level can keep every sweep rotating
// until the sweep limit.)
__device__ void jacobi_par(double *A, double *Y, int n)
{
    __shared__ double cs[DK_BMAX_VEC / 2][2];
    __shared__ int pr[DK_BMAX_VEC / 2][2];
    __shared__ double red[2][kSmall];
    __shared__ int done, rotated;
    const int nn = n + (n & 1);
    const int npair = nn / 2;
    for (int sweep = 0; sweep < 40; ++sweep) {
        double off = 0, all = 0;
        for (int e = threadIdx.x; e < n * n; e += kSmall) {
            const int i = e / n, j = e - i * n;
            const double a = A[i * LDY + j];
            all += a * a;
            if (i != j) off += a * a;
        }
        red[0][threadIdx.x] = off;
        red[1][threadIdx.x] = all;
        __syncthreads();
        if (threadIdx.x == 0) {
            double o = 0, a = 0;
            for (int q = 0; q < kSmall; ++q) {
                o += red[0][q];
                a += red[1][q];
            }
            done = (o == 0 || o <= 1e-32 * a);
            rotated = 0;
        }
        __syncthreads();
        if (done) break;
        for (int step = 0; step < nn - 1; ++step) {
            if (threadIdx.x < npair) {
                // tournament pairing: index nn-1 fixed, the others rotate
                const int k = threadIdx.x;
                int p, q;
                if (k == 0) {
                    p = nn - 1;
                    q = step;
                } else {
                    p = (step + k) % (nn - 1);
                    q = (step - k + (nn - 1)) % (nn - 1);
                }
                if (p > q) {
                    const int z = p;
                    p = q;
                    q = z;
                }
                double c = 1, s = 0;
                if (q < n) {
                    const double apq = A[p * LDY + q], app = A[p * LDY + p], aqq = A[q * LDY + q];
                    if (apq != 0 && fabs(apq) > DBL_EPSILON * sqrt(fabs(app) * fabs(aqq))) {
                        const double tau = (aqq - app) / (2 * apq);
                        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1 + tau * tau));
                        c = 1 / sqrt(1 + t * t);
                        s = t * c;
                        if (s != 0) rotated = 1;
                    }
                }
                pr[k][0] = p;
                pr[k][1] = q;
                cs[k][0] = c;
                cs[k][1] = s;
            }
            __syncthreads();
            // rows p, q of A (A <- J^T A)
            for (int e = threadIdx.x; e < npair * n; e += kSmall) {
                const int k = e / n, col = e - k * n;
                const int p = pr[k][0], q = pr[k][1];
                if (q >= n || cs[k][1] == 0) continue;
                const double c = cs[k][0], s = cs[k][1];
                const double x = A[p * LDY + col], y = A[q * LDY + col];
                A[p * LDY + col] = c * x - s * y;
                A[q * LDY + col] = s * x + c * y;
            }
            __syncthreads();
            // columns p, q of A and Y (A <- A J, Y <- Y J)
            for (int e = threadIdx.x; e < npair * n; e += kSmall) {
                const int k = e / n, row = e - k * n;
                const int p = pr[k][0], q = pr[k][1];
                if (q >= n || cs[k][1] == 0) continue;
                const double c = cs[k][0], s = cs[k][1];
                double x = A[row * LDY + p], y = A[row * LDY + q];
                A[row * LDY + p] = c * x - s * y;
                A[row * LDY + q] = s * x + c * y;
                x = Y[row * LDY + p];
                y = Y[row * LDY + q];
                Y[row * LDY + p] = c * x - s * y;
                Y[row * LDY + q] = s * x + c * y;
            }
            __syncthreads();
        }
        if (!rotated) break; // read by all after the step's barrier, reset after the next one
    }
}

// The overlaps of a new direction t with the orthonormal Q of the basis
// (T = Q R): from d_j = <stored T_j, t> and tt = <t, t>, r = R^-T b with
// b_j = s_j d_j / |t|; returns rho^2 = 1 - |r|^2, the part of t/|t| outside
// the basis.  One thread.
__device__ double bs_overlaps(const dk_bstate &st, int m, double tt, const double *d, double *r)
{
    const double it = 1 / sqrt(tt);
    double nr = 0;
    for (int j = 0; j < m; ++j) {
        double v = st.s[j] * d[j] * it;
        for (int i = 0; i < j; ++i) v -= st.R[i * LDY + j] * r[i];
        r[j] = v / st.R[j * LDY + j];
        nr += r[j] * r[j];
    }
    return 1 - nr;
}

// t becomes basis vector m with scale 1/|t| and column m of R (r, rho);
// W_m goes to a free W slot, and the star-kernel sigma operands are set.
// One thread.
__device__ void bs_finalize(dk_bstate &st, const dk_binst &in, int m, double tt, const double *r,
                            double rho)
{
    st.s[m] = 1 / sqrt(tt);
    for (int j = 0; j < m; ++j) st.R[j * LDY + m] = r[j];
    st.R[m * LDY + m] = rho;
    st.vslot[m] = st.tslot;
    const int w = take_slot(st.wused, st.P, st.P);
    st.wslot[m] = w;
    st.sin = st.tslot;
    st.sout = w;
    if (w < 0) {
        st.st.status = DK_B_BREAKDOWN;
        return;
    }
    // operands of the star-kernel sigma (instances without lane sigma)
    st.sio.c = in.slots + (uint64_t)st.tslot * in.slot_stride;
    st.sio.sigma = in.slots + (uint64_t)w * in.slot_stride;
    st.sio.shift = st.shift;
    st.sio.active = 1;
}

__global__ __launch_bounds__(kSmall) void k_bsmall(int phase, const uint32_t *__restrict__ ilist,
                                                   const dk_binst *__restrict__ insts,
                                                   dk_bstate *__restrict__ states,
                                                   const double *__restrict__ part)
{
    __shared__ double red[DK_BPV], rbuf[DK_BMAX_VEC];
    __shared__ double A[DK_BMAX_VEC * LDY], Y[DK_BMAX_VEC * LDY];
    __shared__ int ctl;
    const uint32_t ii = ilist[blockIdx.x];
    const dk_binst in = insts[ii];
    dk_bstate &st = states[ii];
    if (st.st.status != DK_B_ACTIVE) return;
    const int m = st.st.m;
    const int flags = st.st.flags;

    if (phase == DK_BPH_A0 || phase == DK_BPH_A) {
        // A0 (start of an iteration) handles instances that have just been
        // started; A (end of an iteration) reduces the residual pass of
        // everything else, in the same launch sequence that wrote it.
        const bool first = flags & DK_BF_FIRST;
        if (first != (phase == DK_BPH_A0)) return;
        if (!first) reduce_tiles(in, part, m + 2, red);
        if (threadIdx.x == 0) {
            const double res = first ? st.tol * 10 : sqrt(red[0]); // src/davidson.c:45
            st.st.res = res;
            // src/davidson.c:72: res > tol && its++ < max_its
            bool cont = res > st.tol;
            if (cont) cont = st.st.its++ < st.max_its;
            ctl = 0;
            if (!cont)
                st.st.status = res > st.tol ? DK_B_MAXITS : DK_B_CONVERGED;
            else if (!first && (!(red[1] > 0) || !isfinite(red[0]) || !isfinite(red[1])))
                st.st.status = DK_B_BREAKDOWN;
            else
                ctl = 1;
            int nf = flags & ~DK_BF_FIRST;
            if (ctl && first) {
                nf |= DK_BF_ORTHO; // the ORTHO pass measures |t| (m = 0: no projection)
            } else if (ctl) {
                // t joins the basis as it is unless it is nearly dependent on it
                double *r = rbuf;
                const double tt = red[1];
                const double rho2 = bs_overlaps(st, m, tt, red + 2, r);
                if (rho2 > st.otol) {
                    bs_finalize(st, in, m, tt, r, sqrt(rho2));
                } else {
                    // explicit projection t -= T c, c = R^-1 R^-T <T, t>
                    const double nt = sqrt(tt);
                    for (int j = m - 1; j >= 0; --j) {
                        double v = r[j] * nt;
                        for (int i = j + 1; i < m; ++i) v -= st.R[j * LDY + i] * st.g[i];
                        st.g[j] = v / st.R[j * LDY + j];
                    }
                    for (int j = 0; j < m; ++j) st.g[j] *= st.s[j];
                    nf |= DK_BF_ORTHO;
                }
            }
            st.st.flags = nf;
        }
        return;
    }
    if (phase == DK_BPH_B) {
        // after an explicit projection (or the first t's norm): the
        // projected t joins the basis
        if (!(flags & DK_BF_ORTHO)) return;
        reduce_tiles(in, part, m + 1, red);
        if (threadIdx.x == 0) {
            double *r = rbuf;
            const double tt = red[0];
            const double rho2 = tt > 0 && isfinite(tt) ? bs_overlaps(st, m, tt, red + 1, r) : -1.0;
            if (!(rho2 > 0))
                st.st.status = DK_B_BREAKDOWN;
            else
                bs_finalize(st, in, m, tt, r, sqrt(rho2));
            st.st.flags = flags & ~DK_BF_ORTHO;
        }
        return;
    }
    if (phase == DK_BPH_D) {
        if (threadIdx.x == 0) st.sio.active = 0; // this iteration's sigma has run
        reduce_tiles(in, part, m + 1, red);
        const int mn = m + 1;
        // column m of T^T (H - shift) T (T_j = s_j * stored T_j)
        __shared__ double col[DK_BMAX_VEC], z[DK_BMAX_VEC], u[DK_BMAX_VEC], w[DK_BMAX_VEC],
            v[DK_BMAX_VEC], rr[DK_BMAX_VEC];
        for (int j = threadIdx.x; j <= m; j += kSmall) col[j] = st.s[j] * st.s[m] * red[j];
        for (int e = threadIdx.x; e < m * m; e += kSmall) {
            const int i = e / m, j = e - i * m;
            A[i * LDY + j] = st.A[i * LDY + j];
            Y[i * LDY + j] = st.Y[i * LDY + j];
        }
        for (int j = threadIdx.x; j < m; j += kSmall) rr[j] = st.R[j * LDY + m];
        __syncthreads();
        // in Q (T = Q R, R = [R0 r; 0 rho]): column m of Q^T (H - shift) Q is
        //   (R0^-T c - Ahat r) / rho  and  (c_m - 2 r.z + r.Ahat r) / rho^2,
        // z = R0^-T c, Ahat = Y A Y^T the previous projected matrix
        if (threadIdx.x == 0)
            for (int j = 0; j < m; ++j) {
                double a = col[j];
                for (int i = 0; i < j; ++i) a -= st.R[i * LDY + j] * z[i];
                z[j] = a / st.R[j * LDY + j];
            }
        __syncthreads();
        for (int k = threadIdx.x; k < m; k += kSmall) { // w = Y^T z, u = Y^T r
            double a = 0, b = 0;
            for (int j = 0; j < m; ++j) {
                a += Y[j * LDY + k] * z[j];
                b += Y[j * LDY + k] * rr[j];
            }
            w[k] = a;
            u[k] = b;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < m; k += kSmall) { // v = A u
            double a = 0;
            for (int j = 0; j < m; ++j) a += A[k * LDY + j] * u[j];
            v[k] = a;
        }
        __syncthreads();
        // border the rotated matrix: A' = [A b; b^T corner], b = Y^T (column in Q)
        const double rho = st.R[m * LDY + m];
        for (int k = threadIdx.x; k < m; k += kSmall) {
            const double b = (w[k] - v[k]) / rho;
            A[k * LDY + m] = b;
            A[m * LDY + k] = b;
            Y[k * LDY + m] = 0;
            Y[m * LDY + k] = 0;
        }
        if (threadIdx.x == 0) {
            double rz = 0, uv = 0;
            for (int j = 0; j < m; ++j) {
                rz += rr[j] * z[j];
                uv += u[j] * v[j];
            }
            A[m * LDY + m] = (col[m] - 2 * rz + uv) / (rho * rho);
            Y[m * LDY + m] = 1;
        }
        __syncthreads();
        jacobi_par(A, Y, mn);
        __shared__ int order[DK_BMAX_VEC];
        if (threadIdx.x == 0) {
            // ascending eigenvalues (selection order; ties keep the lower index)
            int used = 0;
            for (int r = 0; r < mn; ++r) {
                int best = -1;
                for (int k = 0; k < mn; ++k)
                    if (!((used >> k) & 1) && (best < 0 || A[k * LDY + k] < A[best * LDY + best]))
                        best = k;
                order[r] = best;
                used |= 1 << best;
            }
            st.theta = A[order[0] * LDY + order[0]];
            st.st.energy = st.theta + st.shift;
            st.st.m = mn;
            ctl = mn == st.max_vec;
            if (ctl) {
                for (int k = 0; k < st.keep; ++k) {
                    st.dv[k] = take_slot(st.vused, 0, st.P);
                    st.dw[k] = take_slot(st.wused, st.P, st.P);
                    if (st.dv[k] < 0 || st.dw[k] < 0) st.st.status = DK_B_BREAKDOWN;
                }
                st.st.flags |= DK_BF_DEFLATE;
            } else {
                st.tslot = take_slot(st.vused, 0, st.P);
                if (st.tslot < 0) st.st.status = DK_B_BREAKDOWN;
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < mn * mn; e += kSmall) {
            const int i = e / mn, j = e - i * mn;
            st.A[i * LDY + j] = A[i * LDY + j];
            st.Y[i * LDY + j] = Y[i * LDY + j];
        }
        // Ritz vectors on the stored vectors: R^-1 (Y column), times the scales
        const int nvec = ctl ? st.keep : 1;
        if (threadIdx.x < nvec) {
            const int k = threadIdx.x;
            double *out = ctl ? st.defl[k] : st.y;
            for (int j = mn - 1; j >= 0; --j) {
                double a = Y[j * LDY + order[k]];
                for (int i = j + 1; i < mn; ++i) a -= st.R[j * LDY + i] * out[i];
                out[j] = a / st.R[j * LDY + j];
            }
            for (int j = 0; j < mn; ++j) out[j] *= st.s[j];
        }
        // the kept Ritz values, for small E
        if (ctl)
            for (int k = threadIdx.x; k < st.keep; k += kSmall) st.g[k] = A[order[k] * LDY + order[k]];
        return;
    }
    if (phase == DK_BPH_E) {
        if (!(flags & DK_BF_DEFLATE)) return;
        if (threadIdx.x == 0) {
            const int keep = st.keep;
            for (int j = 0; j < m; ++j) {
                st.vused &= ~(1ull << st.vslot[j]);
                st.wused &= ~(1ull << (st.wslot[j] - st.P));
            }
            for (int k = 0; k < keep; ++k) {
                st.vslot[k] = st.dv[k];
                st.wslot[k] = st.dw[k];
                st.s[k] = 1;
                st.y[k] = k == 0 ? 1.0 : 0.0;
            }
            for (int i = 0; i < keep; ++i)
                for (int j = 0; j < keep; ++j) {
                    st.A[i * LDY + j] = i == j ? st.g[i] : 0.0;
                    st.Y[i * LDY + j] = i == j ? 1.0 : 0.0;
                    st.R[i * LDY + j] = i == j ? 1.0 : 0.0; // the kept Ritz vectors are orthonormal
                }
            st.theta = st.g[0];
            st.st.energy = st.theta + st.shift;
            st.st.m = keep;
            st.st.flags = flags & ~DK_BF_DEFLATE;
            st.tslot = take_slot(st.vused, 0, st.P);
            if (st.tslot < 0) st.st.status = DK_B_BREAKDOWN;
        }
        return;
    }
}

// ---- admission -------------------------------------------------------------
// The start state of an instance's first phase (one thread).
__device__ void start_state(dk_bstate *st, const double *__restrict__ D, const int64_t *__restrict__ idx,
                            int shift_min_d, double tol, double otol, int max_its, int keep,
                            int max_vec, int its0)
{
    st->st.status = DK_B_ACTIVE;
    st->st.m = 0;
    st->st.its = its0;
    st->st.flags = DK_BF_FIRST;
    st->st.res = tol * 10;
    st->max_its = max_its;
    st->keep = keep;
    st->max_vec = max_vec;
    st->P = max_vec + keep;
    st->tslot = 0;
    st->sin = 0;
    st->sout = max_vec + keep; // first W slot: a sigma probe before the first iteration writes there
    st->vused = 1;
    st->wused = 0;
    st->tol = tol;
    st->otol = otol;
    st->shift = shift_min_d ? D[*idx] : 0.0;
    st->theta = 0;
    st->st.energy = st->shift;
    st->sio.active = 0;
}

// ---- admission of many instances: two launches --------------------------
// First index of the minimum: (value, index) pairs, ties to the lower index.
__device__ __forceinline__ void amin2(double &v, double &ix, double v2, double ix2)
{
    if (v2 < v || (v2 == v && ix2 < ix)) {
        v = v2;
        ix = ix2;
    }
}

// Per tile of an admitted instance: its tables in LDS (from the staging
// copy), D for the tile's rows (the reference's exact order), the guess slot
// cleared, the tile's (min D, first index) in part; tile 0 also copies the
// tables to the instance memory.
__global__ __launch_bounds__(kBlock) void k_badmit_diag(const dk_badmit *__restrict__ adm,
                                                        const uint32_t *__restrict__ tiles,
                                                        double *__restrict__ part)
{
    extern __shared__ double lds[];
    __shared__ double sv[kBlock], si[kBlock];
    const uint32_t ai = tiles[2 * blockIdx.x], k = tiles[2 * blockIdx.x + 1];
    const dk_badmit &A = adm[ai];
    const dk_tables &t = A.src;
    const int L = (int)t.L, N = (int)t.N, N2 = N + 2, LL = L * L;
    double *XJ = lds, *h = XJ + 2 * LL;
    uint32_t *bt = (uint32_t *)(h + L);
    for (int i = threadIdx.x; i < LL; i += kBlock) {
        XJ[2 * i] = t.X[i];
        XJ[2 * i + 1] = 2 * t.J[i];
    }
    for (int i = threadIdx.x; i < L; i += kBlock) h[i] = t.h[i];
    for (int i = threadIdx.x; i < (L + 1) * N2; i += kBlock) bt[i] = t.binom[i];
    if (k == 0)
        for (uint32_t i = threadIdx.x; i < A.tb / 4; i += kBlock)
            ((uint32_t *)A.dst)[i] = ((const uint32_t *)t.K)[i];
    __syncthreads();
    const uint32_t I = k * DK_ADMIT_ROWS + threadIdx.x;
    double v = INFINITY, ix = 1e300;
    if (I < t.n) {
        const dk::Binom C{bt, N2};
        const double d = dk::diag_of(dk::unrank(I, L, N, C), L, t.core, XJ, h);
        A.D[I] = d;
        if (A.prec) ((float *)A.t)[I] = 0.f;
        else ((double *)A.t)[I] = 0.0;
        v = d;
        ix = (double)I;
    }
    sv[threadIdx.x] = v;
    si[threadIdx.x] = ix;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double a = sv[threadIdx.x], b = si[threadIdx.x];
            amin2(a, b, sv[threadIdx.x + w], si[threadIdx.x + w]);
            sv[threadIdx.x] = a;
            si[threadIdx.x] = b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * (A.tile0 + k)] = sv[0];
        part[2 * (A.tile0 + k) + 1] = si[0];
    }
}

// Per admitted instance: the argmin over its tiles' partials (a total order,
// so the merge order does not matter), t[argmin] = 1, the start state.
__global__ __launch_bounds__(kBlock) void k_badmit_start(const dk_badmit *__restrict__ adm,
                                                         const double *__restrict__ part)
{
    __shared__ double sv[kBlock], si[kBlock];
    __shared__ int64_t idx;
    const dk_badmit &A = adm[blockIdx.x];
    double v = INFINITY, ix = 1e300;
    for (uint32_t k = threadIdx.x; k < A.ntiles; k += kBlock)
        amin2(v, ix, part[2 * (A.tile0 + k)], part[2 * (A.tile0 + k) + 1]);
    sv[threadIdx.x] = v;
    si[threadIdx.x] = ix;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double a = sv[threadIdx.x], b = si[threadIdx.x];
            amin2(a, b, sv[threadIdx.x + w], si[threadIdx.x + w]);
            sv[threadIdx.x] = a;
            si[threadIdx.x] = b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        idx = (int64_t)si[0];
        if (A.prec) ((float *)A.t)[idx] = 1.f;
        else ((double *)A.t)[idx] = 1.0;
        start_state(A.state, A.D, &idx, A.shift_min_d, A.tol, A.otol, A.max_its, A.keep,
                    A.max_vec, 0);
    }
}

// FP64 restart of mixed-precision instances: t = x over the tile's rows;
// thread 0 of an instance's tile 0 writes its start state.
__global__ __launch_bounds__(kBlock) void k_brestart(const dk_brestart *__restrict__ rs,
                                                     const uint32_t *__restrict__ tiles)
{
    const dk_brestart &R = rs[tiles[2 * blockIdx.x]];
    const uint32_t k = tiles[2 * blockIdx.x + 1];
    if (k == 0 && threadIdx.x == 0)
        start_state(R.state, nullptr, nullptr, 0, R.tol, R.otol, R.max_its, R.keep, R.max_vec,
                    (int)R.its0);
    const uint32_t I = k * DK_ADMIT_ROWS + threadIdx.x;
    if (I < R.n) R.t[I] = R.x[I];
}

template <typename T, int M>
int stream_launch(int op, const dk_binst *inst, dk_bstate *state, const uint32_t *tiles,
                  uint32_t ntile, double *part, hipStream_t s)
{
    const dim3 g(ntile), b(kBlock);
    switch (op) {
    case DK_BOP_ORTHO:
        hipLaunchKernelGGL((k_bstream<T, M, DK_BOP_ORTHO>), g, b, 0, s, inst, state, tiles, part);
        break;
    case DK_BOP_ORTHO2:
        hipLaunchKernelGGL((k_bstream<T, M, DK_BOP_ORTHO2>), g, b, 0, s, inst, state, tiles, part);
        break;
    case DK_BOP_SCOL:
        hipLaunchKernelGGL((k_bstream<T, M, DK_BOP_SCOL>), g, b, 0, s, inst, state, tiles, part);
        break;
    case DK_BOP_DEFLATE:
        hipLaunchKernelGGL((k_bstream<T, M, DK_BOP_DEFLATE>), g, b, 0, s, inst, state, tiles, part);
        break;
    case DK_BOP_RESID:
        hipLaunchKernelGGL((k_bstream<T, M, DK_BOP_RESID>), g, b, 0, s, inst, state, tiles, part);
        break;
    default:
        return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

template <typename T>
int stream_bucket(int op, int M, const dk_binst *inst, dk_bstate *state, const uint32_t *tiles,
                  uint32_t ntile, double *part, hipStream_t s)
{
    if (M <= 8) return stream_launch<T, 8>(op, inst, state, tiles, ntile, part, s);
    if (M <= 16) return stream_launch<T, 16>(op, inst, state, tiles, ntile, part, s);
    if (M <= DK_BMAX_VEC) return stream_launch<T, DK_BMAX_VEC>(op, inst, state, tiles, ntile, part, s);
    return (int)hipErrorInvalidValue;
}

} // namespace

extern "C" int dk_batch_stream(int op, int prec, int M, const dk_binst *inst, dk_bstate *state,
                               const uint32_t *tiles, uint32_t ntile, double *part, hipStream_t s)
{
    if (ntile == 0) return 0;
    return prec ? stream_bucket<float>(op, M, inst, state, tiles, ntile, part, s)
                : stream_bucket<double>(op, M, inst, state, tiles, ntile, part, s);
}

extern "C" int dk_batch_small(int phase, const uint32_t *ilist, uint32_t ninst,
                              const dk_binst *inst, dk_bstate *state, const double *part,
                              hipStream_t s)
{
    if (ninst == 0) return 0;
    hipLaunchKernelGGL(k_bsmall, dim3(ninst), dim3(kSmall), 0, s, phase, ilist, inst, state, part);
    return (int)hipGetLastError();
}

extern "C" int dk_batch_sigma_lane(int prec, uint32_t maxL, uint32_t maxN, const dk_binst *inst,
                                   dk_bstate *state, const uint32_t *tiles, uint32_t ntile,
                                   hipStream_t s)
{
    if (ntile == 0) return 0;
    const size_t lds = sizeof(double) * maxL * maxL + sizeof(uint32_t) * (maxL + 1) * (maxN + 2);
    if (prec)
        hipLaunchKernelGGL(k_bsigma_lane<float>, dim3(ntile), dim3(kBlock), lds, s, inst, state,
                           tiles);
    else
        hipLaunchKernelGGL(k_bsigma_lane<double>, dim3(ntile), dim3(kBlock), lds, s, inst, state,
                           tiles);
    return (int)hipGetLastError();
}

extern "C" int dk_batch_restart(const dk_brestart *rs, const uint32_t *tiles, uint32_t ntile,
                                hipStream_t s)
{
    if (ntile == 0) return 0;
    hipLaunchKernelGGL(k_brestart, dim3(ntile), dim3(kBlock), 0, s, rs, tiles);
    return (int)hipGetLastError();
}

extern "C" int dk_batch_admit(const dk_badmit *adm, uint32_t na, const uint32_t *tiles,
                              uint32_t ntile, double *part, uint32_t maxL, uint32_t maxN,
                              hipStream_t s)
{
    if (na == 0) return 0;
    const size_t lds = sizeof(double) * (2 * maxL * maxL + maxL) +
                       sizeof(uint32_t) * (maxL + 1) * (maxN + 2);
    hipLaunchKernelGGL(k_badmit_diag, dim3(ntile), dim3(kBlock), lds, s, adm, tiles, part);
    hipLaunchKernelGGL(k_badmit_start, dim3(na), dim3(kBlock), 0, s, adm, part);
    return (int)hipGetLastError();
}
