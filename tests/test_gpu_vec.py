"""The fused Davidson vector kernels (dk_multidot, dk_ortho, dk_residual with
and without W dots, dk_lincomb) called directly through the library's
exported symbols, against numpy in FP64 on the same inputs.

Covers every compile-time vector count the kernels dispatch to (m rounded
up to 1, 2, 4, 6, 8, 12, 16, 24, 32, 64: m on and between the buckets,
where the slots past m repeat the last vector with coefficient 0), the
scalar tail (n not a multiple of the 16-byte group), and vectors longer
than one addressing window (2^26 lane steps: 2^27 FP64 / 2^28 FP32
elements), which no solve in the other tests reaches.

Bars: sums are FP64 in every kernel, so dots agree with numpy's to the
rounding of a different summation order (1e-12 relative to the sum of
absolute terms); stored FP32 outputs to one FP32 rounding of the FP64
value (2^-23 relative, plus the FP64 difference)."""
import ctypes as C

import numpy as np
import pytest

import doci_amd as da

pytestmark = pytest.mark.gpu

VP = C.c_void_p


def _hip():
    return C.CDLL("libamdhip64.so")


class Dev:
    """Device copies of numpy arrays (hipMalloc / hipMemcpy), freed on exit."""

    def __init__(self):
        self.hip = _hip()
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(VP), C.c_size_t(a.nbytes), 1) == 0
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(VP), p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.hip.hipDeviceSynchronize()
        for p in self.ptrs:
            self.hip.hipFree(p)


@pytest.fixture(scope="module")
def L():
    lib = da.lib()
    lib.dk_red_scratch_doubles.restype = C.c_size_t
    lib.dk_red_scratch_doubles.argtypes = [C.c_int]
    lib.dk_multidot.argtypes = [C.c_int, VP, VP, C.c_int64, C.c_int, VP, VP, VP]
    lib.dk_ortho.argtypes = [C.c_int, VP, VP, VP, C.c_int64, C.c_int, VP, VP, VP]
    lib.dk_residual.argtypes = [C.c_int, VP, VP, VP, C.c_double, C.c_double, VP, VP, VP, C.c_int,
                                C.c_int64, C.c_int, VP, VP, VP]
    lib.dk_lincomb.argtypes = [C.c_int, VP, VP, VP, C.c_int64, C.c_int, C.c_int, VP]
    for f in (lib.dk_multidot, lib.dk_ortho, lib.dk_residual, lib.dk_lincomb):
        f.restype = C.c_int
    return lib


def _ptrs(ps):
    return (VP * len(ps))(*[p.value for p in ps])


def _close(got, want, scale, bar):
    err = np.max(np.abs(np.asarray(got) - np.asarray(want)) / np.maximum(scale, 1e-300))
    assert err <= bar, err


def _check_all(L, n, m, prec, seed, wdots):
    rng = np.random.default_rng(seed)
    dt = np.float64 if prec == 0 else np.float32
    eps_out = 0.0 if prec == 0 else 2.0 ** -23
    X = [rng.standard_normal(n).astype(dt) for _ in range(m)]
    W = [rng.standard_normal(n).astype(dt) for _ in range(m)]
    y = rng.standard_normal(n).astype(dt)
    D = rng.uniform(-3, 3, n)
    D[::97] = 0.25  # |D - theta| below the preconditioner floor
    g = rng.standard_normal(m)
    X64 = [x.astype(np.float64) for x in X]
    W64 = [w.astype(np.float64) for w in W]
    y64 = y.astype(np.float64)
    nv = 2 + 2 * m
    with Dev() as dv:
        dX = [dv.put(x) for x in X]
        dW = [dv.put(w) for w in W]
        dD = dv.put(D)
        red = dv.put(np.zeros(L.dk_red_scratch_doubles(nv + 8)))
        out = dv.put(np.zeros(nv + 8))
        PX, PW = _ptrs(dX), _ptrs(dW)
        gc = (C.c_double * m)(*g)
        print(f"[case] n={n} m={m} prec={prec}", flush=True)

        # multidot: <y,y>, <X_j,y>
        dy = dv.put(y)
        assert L.dk_multidot(m, PX, dy, n, prec, red, out, None) == 0
        o = dv.get(out, m + 1, np.float64)
        want = [y64 @ y64] + [x @ y64 for x in X64]
        scale = [y64 @ y64] + [np.abs(x) @ np.abs(y64) for x in X64]
        _close(o, want, scale, 1e-12)

        # ortho: y -= sum g_j X_j, then <y,y>, <X_j,y>
        assert L.dk_ortho(m, PX, gc, dy, n, prec, red, out, None) == 0
        yo = dv.get(dy, n, dt).astype(np.float64)
        yw = y64.copy()
        for j in range(m):
            yw -= g[j] * X64[j]
        sc = np.abs(y64) + sum(abs(g[j]) * np.abs(X64[j]) for j in range(m))
        _close(yo, yw, sc, 1e-13 + 2 * eps_out)
        o = dv.get(out, m + 1, np.float64)
        _close(o, [yo @ yo] + [x @ yo for x in X64],
               [yo @ yo] + [np.abs(x) @ np.abs(yo) for x in X64], 1e-12)

        # residual (davidson.c:195-227): x = sum a_j V_j, r = sum a_j W_j - (theta - shift) x,
        # t = r / max(|D - theta|, 1e-12); <r,r>, <t,t>, <V_j,t> (and <W_j,t>)
        theta, shift = 0.25, -0.5
        xs = sum(g[j] * X64[j] for j in range(m))
        ws = sum(g[j] * W64[j] for j in range(m))
        r = ws - (theta - shift) * xs
        d = np.abs(D - theta)
        tw = np.where(d > 1e-12, r / np.where(d > 1e-12, d, 1.0), r / 1e-12)
        rs = sum(abs(g[j]) * (np.abs(W64[j]) + abs(theta - shift) * np.abs(X64[j])) for j in range(m))
        ts = rs / np.maximum(d, 1e-12)
        for wd in ((0, 1) if m <= 32 else (0,)):
            dt_ = dv.put(np.zeros(n, dtype=dt))
            dxo = dv.put(np.zeros(n, dtype=dt))
            assert L.dk_residual(m, PX, PW, gc, theta, shift, dD, dt_, dxo, wd, n, prec, red, out,
                                 None) == 0
            t = dv.get(dt_, n, dt).astype(np.float64)
            xo = dv.get(dxo, n, dt).astype(np.float64)
            _close(t, tw, ts, 1e-12 + 2 * eps_out)
            _close(xo, xs, sum(abs(g[j]) * np.abs(X64[j]) for j in range(m)), 1e-13 + 2 * eps_out)
            o = dv.get(out, 2 + (2 if wd else 1) * m, np.float64)
            _close(o[0], r @ r, rs @ rs, 1e-12)
            _close(o[1], t @ t, t @ t, 1e-12)
            _close(o[2:2 + m], [x @ t for x in X64], [np.abs(x) @ np.abs(t) for x in X64], 1e-12)
            if wd:
                _close(o[2 + m:2 + 2 * m], [w @ t for w in W64], [np.abs(w) @ np.abs(t) for w in W64],
                       1e-12)

        # lincomb into both output precisions
        for po in (0, 1):
            odt = np.float64 if po == 0 else np.float32
            dout = dv.put(np.zeros(n, dtype=odt))
            assert L.dk_lincomb(m, PX, gc, dout, n, prec, po, None) == 0
            lo = dv.get(dout, n, odt).astype(np.float64)
            _close(lo, xs, sum(abs(g[j]) * np.abs(X64[j]) for j in range(m)),
                   1e-13 + (2.0 ** -23 if po else 0.0) * 2)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("m", [1, 2, 3, 5, 6, 7, 8, 9, 12, 13, 16, 17, 24, 25, 32, 33, 64])
def test_vector_kernels_every_bucket(L, m, prec):
    # n = 4099: 1024 groups of 4 FP32 / 2049 of 2 FP64 and a scalar tail
    _check_all(L, 4099, m, prec, 100 + m, True)


@pytest.mark.parametrize("prec,n", [(0, (1 << 27) + (1 << 20) + 3), (1, (1 << 28) + (1 << 20) + 1)])
def test_vector_kernels_beyond_one_window(L, prec, n):
    """Longer than one window of 2^26 lane steps (the bases advance between
    windows; the tail follows the last one)."""
    _check_all(L, n, 2, prec, 7, True)


@pytest.mark.parametrize("prec", [0, 1])
def test_vector_kernels_empty_basis(L, prec):
    """m = 0: multidot and ortho return <y,y> only (their vector slots all
    repeat y, with coefficient 0), and ortho leaves y as it is."""
    n = 4099
    dt = np.float64 if prec == 0 else np.float32
    y = np.random.default_rng(3).standard_normal(n).astype(dt)
    y64 = y.astype(np.float64)
    with Dev() as dv:
        dy = dv.put(y)
        red = dv.put(np.zeros(L.dk_red_scratch_doubles(8)))
        out = dv.put(np.zeros(8))
        empty = (VP * 1)()
        print(f"[case] empty basis prec={prec}", flush=True)
        assert L.dk_multidot(0, empty, dy, n, prec, red, out, None) == 0
        _close(dv.get(out, 1, np.float64), [y64 @ y64], [y64 @ y64], 1e-12)
        assert L.dk_ortho(0, empty, (C.c_double * 1)(), dy, n, prec, red, out, None) == 0
        assert np.array_equal(dv.get(dy, n, dt), y)
        _close(dv.get(out, 1, np.float64), [y64 @ y64], [y64 @ y64], 1e-12)
