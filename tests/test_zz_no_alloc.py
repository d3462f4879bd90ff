"""doci_op_sigma_dev_noalloc (the star sigma with its slot buffer in the
library's static device array, no allocation) against doci_op_sigma_dev
(the same kernels with the operator's own slot buffer), through the C ABI,
on two operators in a row, the second smaller, in FP64 and FP32 storage.
Bar: the FP64 sigma bar (1e-12 relative, max-norm); the star pass sums a
centre's hops with LDS atomics, so the two entry points agree to rounding,
not bit for bit.  FP32 storage: 1e-6 (FP32 rounding of the output).  A
problem whose slot buffer exceeds the static array is rejected."""
import ctypes as C

import numpy as np
import pytest

import doci_amd as da

pytestmark = pytest.mark.gpu


def _hip():
    return C.CDLL("libamdhip64.so")


def _dev(nbytes):
    p = C.c_void_p()
    assert _hip().hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
    return p


def _run(op, entry, c, prec):
    """entry(op, d_in, d_out, prec, shift) on device copies of c."""
    hip = _hip()
    es = 8 if prec == 0 else 4
    h = np.ascontiguousarray(c, dtype=np.float64 if prec == 0 else np.float32)
    din, dout = _dev(es * op.n), _dev(es * op.n)
    try:
        assert hip.hipMemcpy(din, h.ctypes.data_as(C.c_void_p), C.c_size_t(es * op.n), 1) == 0
        rc = entry(op.handle, din, dout, prec, C.c_double(0.25))
        assert rc == 0, (rc, da.lib().doci_last_error())
        assert hip.hipDeviceSynchronize() == 0
        out = np.zeros(op.n, dtype=h.dtype)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), dout, C.c_size_t(es * op.n), 2) == 0
    finally:
        hip.hipFree(din)
        hip.hipFree(dout)
    return out.astype(np.float64)


def test_sigma_noalloc_matches_sigma_dev(fcidump_dir, syn24):
    lib = da.lib()
    # the larger problem first, so that the second call finds the array
    # holding the first call's slots
    for k, path in enumerate([str(syn24), str(fcidump_dir / "lif_631g_1.FCIDUMP")]):
        print(f"[case] no-alloc sigma call {k + 1}: {path.rsplit('/', 1)[-1]}", flush=True)
        with da.Operator.from_fcidump(path) as op:
            op.set_kernel("star")
            split = op.star_split()
            assert split["zslots"] <= 8 << 20, split
            c = da.gen_vector(op.n, 3 + k)
            for prec, bar in ((0, 1e-12), (1, 1e-6)):
                ref = _run(op, lib.doci_op_sigma_dev, c, prec)
                got = _run(op, lib.doci_op_sigma_dev_noalloc, c, prec)
                err = np.abs(got - ref).max() / np.abs(ref).max()
                print(f"[case]   n={op.n} prec={prec} rel err {err:.2e}", flush=True)
                assert err <= bar, (k, prec, err)


def test_sigma_noalloc_rejects_oversized_problem():
    """L=34 / N=10 needs about 0.94 G slot entries: a clear error, no launch."""
    print("[case] no-alloc sigma rejects L34", flush=True)
    with da.Operator.synthetic(34, 10, 2) as op:
        op.set_kernel("star")
        rc = da.lib().doci_op_sigma_dev_noalloc(op.handle, C.c_void_p(16), C.c_void_p(32), 0,
                                                C.c_double(0.0))
        assert rc == -9  # DOCI_E_ARG
        assert b"exceeds the static buffer" in da.lib().doci_last_error()
