"""doci_amd -- DOCI ground-state solver on one AMD Instinct MI355X.

Python front-end of libdoci_amd.so (C host layer + HIP kernels for gfx950).
The library is loaded from this package directory; there is no fallback:
importing works without a GPU, but every compute call goes through HIP and
raises DociError when the library or the device is missing.

The C entry points mirror klgunst/doci's (fill_doci_data, doci_matvec,
sparse_eigensolve, ...); see include/doci_amd.h.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libdoci_amd.so")
CLI_PATH = os.path.join(PKG, "bin", "doci")

_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class DociError(RuntimeError):
    pass


class SolveOpts(C.Structure):
    _fields_ = [("solver", C.c_char * 8), ("tol", C.c_double), ("max_its", C.c_int),
                ("davidson_keep", C.c_int), ("davidson_max_vec", C.c_int),
                ("precision", C.c_int), ("switch_tol", C.c_double), ("trace", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("iterations_fp32", C.c_int),
                ("sigma_calls", C.c_int), ("converged", C.c_int), ("residual", C.c_double),
                ("energy", C.c_double), ("seconds", C.c_double),
                ("seconds_sigma", C.c_double), ("sigma_calls_fp32", C.c_int),
                ("seconds_sigma_fp32", C.c_double), ("d_energy", C.c_double)]


class BatchOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_its", C.c_int), ("keep", C.c_int),
                ("max_vec", C.c_int), ("switch_tol", C.c_double), ("use_graphs", C.c_int),
                ("tile_rows", C.c_int), ("mem_budget", C.c_size_t), ("max_instances", C.c_int),
                ("keep_vectors", C.c_int), ("mem_reserve", C.c_size_t)]


class BatchResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("iterations_fp32", C.c_int),
                ("energy", C.c_double), ("residual", C.c_double), ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_done", C.c_double)]


class InputFile(C.Structure):
    _fields_ = [("fcidump", C.c_char * 255), ("solver", C.c_char * 20), ("max_its", C.c_int),
                ("tol", C.c_double), ("davidson_keep", C.c_int),
                ("davidson_max_vec", C.c_int), ("hf_init", C.c_int), ("precision", C.c_int),
                ("eigvec", C.c_char * 255)]


# Every symbol include/doci_amd.h declares (checked by the CPU test suite).
EXPORTS = [
    "read_fcidump", "fill_doci_data", "doci_get_diagonal", "doci_get_basis_size",
    "doci_cleanup_data", "doci_matvec", "sparse_eigensolve", "davidson", "ray_q",
    "doci_last_error", "doci_op_from_fcidump", "doci_op_from_integrals", "doci_op_free",
    "doci_op_dim", "doci_op_L", "doci_op_N", "doci_op_diagonal", "doci_op_sigma",
    "doci_op_sigma_dev", "doci_op_sigma_dev_noalloc", "doci_op_stream", "doci_solve_defaults", "doci_solve",
    "doci_write_fcidump", "doci_write_vector", "doci_read_vector", "doci_gen_integrals", "doci_gen_vector", "doci_read_inputfile",
    "doci_read_fcidump", "doci_integral_tables", "doci_free", "doci_dev_copy", "doci_op_set_kernel",
    "doci_op_slab_split", "doci_op_check_violations", "doci_op_tune", "doci_op_time_sigma",
    "doci_op_star_split", "doci_op_kernel", "doci_batch_defaults", "doci_batch_create", "doci_batch_free",
    "doci_batch_footprint", "doci_batch_add", "doci_batch_add_many", "doci_batch_step", "doci_batch_run",
    "doci_batch_result_of", "doci_batch_release", "doci_batch_clock", "doci_batch_stats",
]


def lib():
    """The loaded libdoci_amd.so (raises DociError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DociError(f"{LIB_PATH} is not built: run `python -m doci_amd.build`")
    l = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    l.doci_last_error.restype = C.c_char_p
    l.doci_op_from_fcidump.argtypes = [C.c_char_p, C.POINTER(vp), C.POINTER(C.c_int)]
    l.doci_op_from_integrals.argtypes = [C.c_uint, C.c_uint, C.c_double, _dp, _dp, C.POINTER(vp)]
    l.doci_op_free.argtypes = [vp]
    l.doci_op_dim.argtypes = [vp]
    l.doci_op_dim.restype = C.c_int64
    l.doci_op_L.argtypes = [vp]
    l.doci_op_L.restype = C.c_uint
    l.doci_op_N.argtypes = [vp]
    l.doci_op_N.restype = C.c_uint
    l.doci_op_diagonal.argtypes = [vp, _dp]
    l.doci_op_sigma.argtypes = [vp, _dp, _dp]
    l.doci_op_sigma_dev.argtypes = [vp, vp, vp, C.c_int, C.c_double]
    l.doci_op_sigma_dev_noalloc.argtypes = [vp, vp, vp, C.c_int, C.c_double]
    l.doci_op_stream.argtypes = [vp]
    l.doci_op_stream.restype = vp
    l.doci_solve_defaults.argtypes = [C.POINTER(SolveOpts)]
    l.doci_solve.argtypes = [vp, C.POINTER(SolveOpts), vp, vp, C.POINTER(C.c_double),
                             C.POINTER(SolveInfo)]
    l.doci_write_fcidump.argtypes = [C.c_char_p, C.c_uint, C.c_uint, C.c_double, _dp, _dp]
    l.doci_gen_integrals.argtypes = [C.c_uint, C.c_uint64, C.POINTER(C.c_double), _dp, _dp]
    l.doci_gen_vector.argtypes = [C.c_int64, C.c_uint64, _dp]
    l.doci_read_inputfile.argtypes = [C.c_char_p, C.POINTER(InputFile)]
    l.doci_write_vector.argtypes = [C.c_char_p, C.c_uint, C.c_uint, C.c_int64, _dp]
    l.doci_read_vector.argtypes = [C.c_char_p, C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                   C.POINTER(C.c_int64), C.POINTER(C.POINTER(C.c_double))]
    l.doci_read_fcidump.argtypes = [C.c_char_p, C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                    C.POINTER(C.c_uint), C.POINTER(C.c_double),
                                    C.POINTER(C.POINTER(C.c_double)),
                                    C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_int)]
    l.doci_free.argtypes = [vp]
    l.doci_dev_copy.argtypes = [vp, vp, C.c_int64]
    l.doci_op_set_kernel.argtypes = [vp, C.c_int]
    l.doci_op_slab_split.argtypes = [vp, C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                     C.POINTER(C.c_uint)]
    l.doci_op_check_violations.argtypes = [vp]
    l.doci_op_check_violations.restype = C.c_int64
    l.doci_op_tune.argtypes = [vp, C.c_char_p, C.c_int]
    l.doci_op_star_split.argtypes = [vp, C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                     C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                     C.POINTER(C.c_size_t)]
    l.doci_op_time_sigma.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
    l.doci_op_kernel.argtypes = [vp]
    l.doci_batch_defaults.argtypes = [C.POINTER(BatchOpts)]
    l.doci_batch_create.argtypes = [C.POINTER(BatchOpts), C.POINTER(vp)]
    l.doci_batch_free.argtypes = [vp]
    l.doci_batch_footprint.argtypes = [C.POINTER(BatchOpts), C.c_uint, C.c_uint]
    l.doci_batch_footprint.restype = C.c_size_t
    l.doci_batch_add.argtypes = [vp, C.c_uint, C.c_uint, C.c_double, _dp, _dp, C.c_int,
                                 C.POINTER(C.c_int)]
    l.doci_batch_add_many.argtypes = [vp, C.c_int, C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                      C.POINTER(C.c_double), C.POINTER(vp), C.POINTER(vp), C.c_int,
                                      C.POINTER(C.c_int)]
    l.doci_batch_step.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    l.doci_batch_run.argtypes = [vp]
    l.doci_batch_result_of.argtypes = [vp, C.c_int, C.POINTER(BatchResult), vp]
    l.doci_batch_release.argtypes = [vp, C.c_int]
    l.doci_batch_clock.argtypes = [vp]
    l.doci_batch_clock.restype = C.c_double
    l.doci_batch_stats.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    _lib = l
    return l


def _check(rc: int, what: str) -> int:
    if rc < 0:
        msg = lib().doci_last_error().decode(errors="replace")
        raise DociError(f"{what} failed ({rc}): {msg}")
    return rc


# ---------------------------------------------------------------------------
# host-side helpers (no GPU needed)
# ---------------------------------------------------------------------------

def gen_integrals(L: int, seed: int):
    """Seeded synthetic integrals (core, one[L*L], two[L^4]) in the raw
    FCIDUMP layout (canonical entries only)."""
    one = np.zeros(L * L)
    two = np.zeros(L ** 4)
    core = C.c_double()
    _check(lib().doci_gen_integrals(L, seed, C.byref(core), one, two), "doci_gen_integrals")
    return core.value, one, two


def write_fcidump(path: str, L: int, N: int, core: float, one, two) -> None:
    _check(lib().doci_write_fcidump(path.encode(), L, N, core, np.ascontiguousarray(one),
                                    np.ascontiguousarray(two)), "doci_write_fcidump")


def gen_fcidump(path: str, L: int, N: int, seed: int) -> None:
    core, one, two = gen_integrals(L, seed)
    write_fcidump(path, L, N, core, one, two)


def read_fcidump(path: str):
    """(L, Na, Nb, core, one[L*L], two[L^4]) as read (untransformed); raises
    DociError with the offending line number for malformed input."""
    L, Na, Nb = C.c_uint(), C.c_uint(), C.c_uint()
    core = C.c_double()
    one, two = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
    line = C.c_int(0)
    rc = lib().doci_read_fcidump(path.encode(), C.byref(L), C.byref(Na), C.byref(Nb),
                                 C.byref(core), C.byref(one), C.byref(two), C.byref(line))
    if rc < 0:
        err = DociError(f"doci_read_fcidump({path}) failed ({rc}): "
                        + lib().doci_last_error().decode(errors="replace"))
        err.code, err.line = rc, line.value
        raise err
    l = L.value
    o = np.ctypeslib.as_array(one, (l * l,)).copy()
    t = np.ctypeslib.as_array(two, (l ** 4,)).copy()
    lib().doci_free(C.cast(one, C.c_void_p))
    lib().doci_free(C.cast(two, C.c_void_p))
    return l, Na.value, Nb.value, core.value, o, t


def gen_vector(n: int, seed: int) -> np.ndarray:
    v = np.zeros(n)
    _check(lib().doci_gen_vector(n, seed, v), "doci_gen_vector")
    return v


def read_inputfile(path: str) -> dict:
    s = InputFile()
    _check(lib().doci_read_inputfile(path.encode(), C.byref(s)), "doci_read_inputfile")
    return dict(fcidump=s.fcidump.decode(), solver=s.solver.decode(), max_its=s.max_its,
                tol=s.tol, davidson_keep=s.davidson_keep,
                davidson_max_vec=s.davidson_max_vec, hf_init=s.hf_init, precision=s.precision,
                eigvec=s.eigvec.decode())


def read_vector(path: str):
    """(L, N, x) from an eigenvector file written by doci_write_vector."""
    L, N, n = C.c_uint(), C.c_uint(), C.c_int64()
    p = C.POINTER(C.c_double)()
    _check(lib().doci_read_vector(path.encode(), C.byref(L), C.byref(N), C.byref(n), C.byref(p)),
           "doci_read_vector")
    x = np.ctypeslib.as_array(p, (n.value,)).copy()
    lib().doci_free(C.cast(p, C.c_void_p))
    return L.value, N.value, x


def write_vector(path: str, L: int, N: int, x) -> None:
    x = np.ascontiguousarray(x, dtype=np.float64)
    _check(lib().doci_write_vector(path.encode(), L, N, x.size, x), "doci_write_vector")


def binom(k: int, n: int) -> int:
    from math import comb
    return comb(n, k)


# ---------------------------------------------------------------------------
# device-resident operator
# ---------------------------------------------------------------------------

@dataclass
class Result:
    energy: float
    code: int
    x: np.ndarray | None
    info: SolveInfo


class Operator:
    """A DOCI Hamiltonian resident on the GPU (tables, diagonal)."""

    def __init__(self, handle):
        self._h = handle
        self.n = int(lib().doci_op_dim(handle))
        self.L = int(lib().doci_op_L(handle))
        self.N = int(lib().doci_op_N(handle))
        names = {v: k for k, v in self.KERNELS.items()}
        self.kernel = names[int(lib().doci_op_kernel(handle))]  # default: by size

    @classmethod
    def from_fcidump(cls, path: str) -> "Operator":
        h = C.c_void_p()
        line = C.c_int(0)
        _check(lib().doci_op_from_fcidump(path.encode(), C.byref(h), C.byref(line)),
               f"doci_op_from_fcidump({path})")
        return cls(h)

    @classmethod
    def from_integrals(cls, L: int, N: int, core: float, one, two) -> "Operator":
        h = C.c_void_p()
        _check(lib().doci_op_from_integrals(L, N, core, np.ascontiguousarray(one, dtype=np.float64),
                                            np.ascontiguousarray(two, dtype=np.float64),
                                            C.byref(h)), "doci_op_from_integrals")
        return cls(h)

    @classmethod
    def synthetic(cls, L: int, N: int, seed: int) -> "Operator":
        core, one, two = gen_integrals(L, seed)
        return cls.from_integrals(L, N, core, one, two)

    @property
    def handle(self):
        return self._h

    def diagonal(self) -> np.ndarray:
        out = np.zeros(self.n)
        _check(lib().doci_op_diagonal(self._h, out), "doci_op_diagonal")
        return out

    def sigma(self, c: np.ndarray) -> np.ndarray:
        out = np.zeros(self.n)
        _check(lib().doci_op_sigma(self._h, np.ascontiguousarray(c, dtype=np.float64), out),
               "doci_op_sigma")
        return out

    def sigma_dev(self, d_in: int, d_out: int, prec: int = 0, shift: float = 0.0) -> None:
        """Device pointers (ints), on the operator's stream."""
        _check(lib().doci_op_sigma_dev(self._h, d_in, d_out, prec, shift), "doci_op_sigma_dev")

    KERNELS = {"star": 0, "lane": 1, "star-checked": 2, "slab": 3, "slab-checked": 4}

    def set_kernel(self, name: str) -> None:
        """Sigma kernel: "star" (default from 2^20 determinants), "lane"
        (default below), "slab" (bit-reproducible), or a "-checked" variant."""
        _check(lib().doci_op_set_kernel(self._h, self.KERNELS[name]), "doci_op_set_kernel")
        self.kernel = name

    def slab_split(self):
        a, b, g = C.c_uint(), C.c_uint(), C.c_uint()
        _check(lib().doci_op_slab_split(self._h, C.byref(a), C.byref(b), C.byref(g)),
               "doci_op_slab_split")
        return a.value, b.value, g.value

    def tune(self, key: str, value: int) -> None:
        _check(lib().doci_op_tune(self._h, key.encode(), int(value)), f"doci_op_tune({key})")

    def time_sigma(self, prec: int = 0, reps: int = 5) -> float:
        """Average sigma duration in ms (HIP events on the operator's stream)."""
        ms = C.c_double()
        _check(lib().doci_op_time_sigma(self._h, prec, reps, C.byref(ms)), "doci_op_time_sigma")
        return ms.value

    def star_split(self):
        a, b, g, sg, z = C.c_uint(), C.c_uint(), C.c_uint(), C.c_uint(), C.c_size_t()
        _check(lib().doci_op_star_split(self._h, C.byref(a), C.byref(b), C.byref(g),
                                        C.byref(sg), C.byref(z)), "doci_op_star_split")
        return dict(a=a.value, b=b.value, groups=g.value, star_groups=sg.value, zslots=z.value)

    def check_violations(self) -> int:
        return _check(lib().doci_op_check_violations(self._h), "doci_op_check_violations")

    def stream(self) -> int:
        return int(lib().doci_op_stream(self._h) or 0)

    def solve(self, solver: str = "D", tol: float = 1e-8, max_its: int = 200, keep: int = 2,
              max_vec: int = 30, precision: str = "fp64", switch_tol: float = 1e-5,
              guess: np.ndarray | None = None, want_x: bool = True, trace: bool = False) -> Result:
        o = SolveOpts()
        lib().doci_solve_defaults(C.byref(o))
        o.solver = solver.encode()
        o.tol, o.max_its, o.davidson_keep, o.davidson_max_vec = tol, max_its, keep, max_vec
        o.precision = 1 if precision == "mixed" else 0
        o.switch_tol = switch_tol
        o.trace = int(trace)
        info = SolveInfo()
        e = C.c_double()
        x = np.zeros(self.n) if want_x else None
        g = None if guess is None else np.ascontiguousarray(guess, dtype=np.float64)
        rc = lib().doci_solve(self._h, C.byref(o), None if g is None else g.ctypes.data,
                              None if x is None else x.ctypes.data, C.byref(e), C.byref(info))
        _check(rc, "doci_solve")
        return Result(e.value, rc, x, info)

    def close(self):
        if self._h:
            lib().doci_op_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# many independent solves (lock-step batch / continuous-batching serving)
# ---------------------------------------------------------------------------

class Batch:
    """Independent DOCI instances solved together on one GPU (see
    doci_batch_* in include/doci_amd.h).  add() queues an instance, step()
    admits what fits and runs one Davidson iteration of every running
    instance, run() steps until all queued instances have finished."""

    STATUS = {0: "waiting", 1: "running", 2: "converged", 3: "max_its", 4: "breakdown"}

    def __init__(self, tol=1e-8, max_its=200, keep=2, max_vec=30, switch_tol=1e-5,
                 use_graphs=True, tile_rows=0, mem_budget=0, max_instances=1024,
                 keep_vectors=True, mem_reserve=0):
        l = lib()
        o = BatchOpts()
        l.doci_batch_defaults(C.byref(o))
        o.tol, o.max_its, o.keep, o.max_vec, o.switch_tol = tol, max_its, keep, max_vec, switch_tol
        o.use_graphs, o.tile_rows, o.mem_budget = int(use_graphs), tile_rows, mem_budget
        o.max_instances = max_instances
        o.keep_vectors = int(keep_vectors)
        o.mem_reserve = int(mem_reserve)
        self.opts = o
        h = C.c_void_p()
        _check(l.doci_batch_create(C.byref(o), C.byref(h)), "doci_batch_create")
        self._h = h
        self.dims = {}

    def footprint(self, L: int, N: int) -> int:
        return int(lib().doci_batch_footprint(C.byref(self.opts), L, N))

    def add(self, L: int, N: int, core: float, one, two, precision: str = "fp64") -> int:
        i = C.c_int()
        _check(lib().doci_batch_add(self._h, L, N, core, np.ascontiguousarray(one, np.float64),
                                    np.ascontiguousarray(two, np.float64),
                                    1 if precision == "mixed" else 0, C.byref(i)),
               "doci_batch_add")
        self.dims[i.value] = binom(N, L)
        return i.value

    def add_synthetic(self, L: int, N: int, seed: int, precision: str = "fp64") -> int:
        core, one, two = gen_integrals(L, seed)
        return self.add(L, N, core, one, two, precision)

    def add_many(self, items, precision: str = "fp64") -> list:
        """Queues (L, N, core, one, two) tuples in one call (the library
        transforms their integrals on several threads).  Returns the ids."""
        items = list(items)
        k = len(items)
        if k == 0:
            return []
        arrs = [(np.ascontiguousarray(it[3], np.float64), np.ascontiguousarray(it[4], np.float64))
                for it in items]
        Ls = (C.c_uint * k)(*[int(it[0]) for it in items])
        Ns = (C.c_uint * k)(*[int(it[1]) for it in items])
        cores = (C.c_double * k)(*[float(it[2]) for it in items])
        ones = (C.c_void_p * k)(*[a[0].ctypes.data for a in arrs])
        twos = (C.c_void_p * k)(*[a[1].ctypes.data for a in arrs])
        ids = (C.c_int * k)()
        _check(lib().doci_batch_add_many(self._h, k, Ls, Ns, cores, ones, twos,
                                         1 if precision == "mixed" else 0, ids), "doci_batch_add_many")
        for it, i in zip(items, ids):
            self.dims[i] = binom(int(it[1]), int(it[0]))
        return list(ids)

    def add_pipelined(self, items, precision: str = "fp64", chunk: int = 64) -> list:
        """Queues many instances ((L, N, core, one, two) tuples) in chunks,
        stepping the engine after each chunk but the last, so that the host
        intake of a chunk (the integral transform, on several threads)
        overlaps the GPU's iterations of the chunks before it.  Returns the
        instance ids."""
        items = list(items)
        ids = []
        for c0 in range(0, len(items), chunk):
            ids += self.add_many(items[c0:c0 + chunk], precision=precision)
            if c0 + chunk < len(items):
                self.step()
        return ids

    def step(self):
        r, w = C.c_int(), C.c_int()
        _check(lib().doci_batch_step(self._h, C.byref(r), C.byref(w)), "doci_batch_step")
        return r.value, w.value

    def run(self) -> None:
        _check(lib().doci_batch_run(self._h), "doci_batch_run")

    def result(self, i: int, want_x: bool = False):
        r = BatchResult()
        x = np.zeros(self.dims[i]) if want_x else None
        _check(lib().doci_batch_result_of(self._h, i, C.byref(r),
                                          None if x is None else x.ctypes.data),
               "doci_batch_result_of")
        return r, x

    def release(self, i: int) -> None:
        _check(lib().doci_batch_release(self._h, i), "doci_batch_release")

    def clock(self) -> float:
        return float(lib().doci_batch_clock(self._h))

    def stats(self):
        t, it, g = C.c_int(), C.c_int(), C.c_int()
        lib().doci_batch_stats(self._h, C.byref(t), C.byref(it), C.byref(g))
        return dict(tile_rows=t.value, iterations=it.value, graph_launches=g.value)

    def close(self):
        if self._h:
            lib().doci_batch_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
