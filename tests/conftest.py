import gzip
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

SMALL = ["H2O_STO3G", "N2_STO3G", "lili_321g", "H2O_631G"]
LARGE = ["lif_631g_1", "co_631g_1"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running case")
    # build what is missing (a built tree skips this; it covers a bare checkout)
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"],
                       check=True)
    if not os.path.exists(os.path.join(ROOT, "doci_amd", "libdoci_amd.so")):
        from doci_amd.build import build
        build()


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN, "golden.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def golden_small():
    return dict(np.load(os.path.join(GOLDEN, "small.npz")))


@pytest.fixture(scope="session")
def fcidump_dir(tmp_path_factory):
    """The reference's test FCIDUMPs, unpacked from the fixtures."""
    d = tmp_path_factory.mktemp("fcidump")
    src = os.path.join(GOLDEN, "fcidump")
    for f in os.listdir(src):
        with gzip.open(os.path.join(src, f), "rb") as fi, open(d / f[:-3], "wb") as fo:
            shutil.copyfileobj(fi, fo)
    return d


@pytest.fixture(scope="session")
def syn24(tmp_path_factory):
    """Synthetic L=24, N=8, seed 1 FCIDUMP (735,471 determinants)."""
    import doci_amd as da
    p = tmp_path_factory.mktemp("syn") / "syn_L24_N8_s1.FCIDUMP"
    da.gen_fcidump(str(p), 24, 8, 1)
    return p


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def check_golden_array(arr, rec, key):
    """Exact comparison with a golden record (SHA-256 + sampled bit patterns)."""
    import hashlib
    idx = np.array(rec[key + "_idx"], dtype=np.int64)
    want = np.array(rec[key + "_bits"], dtype=np.uint64)
    got = bits(arr)[idx]
    assert np.array_equal(got, want), f"{key}: sampled entries differ"
    h = hashlib.sha256(np.ascontiguousarray(arr, dtype=np.float64).tobytes()).hexdigest()
    assert h == rec[key + "_sha256"], f"{key}: SHA-256 differs"
