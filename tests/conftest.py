import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "spams-devel_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build(with_ref=False)
    return oracle


def golden_files(prefix):
    return sorted(f for f in os.listdir(GOLDEN) if f.startswith(prefix) and f.endswith(".npz"))


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


def assert_csc_layout(A, pB, r, skip_cols=()):
    """The engine's CSC arrays equal the reference's SpMatrix arrays
    (linalg/linalg.h:4840-4862: column pointers pB, rows sorted ascending
    inside each column) bit for bit.  ``skip_cols`` lists columns whose
    reference entries are an artefact the engine does not reproduce (a
    masked signal with no observed row); the arrays are then compared with
    those columns taken out of both."""
    A = A.tocsc() if not hasattr(A, "indptr") else A
    ip = np.asarray(A.indptr, np.int64)
    ix = np.asarray(A.indices, np.int64)
    pB = np.asarray(pB, np.int64)
    r = np.asarray(r, np.int64)
    if skip_cols:
        keep = np.ones(len(pB) - 1, bool)
        keep[list(skip_cols)] = False

        def strip(ptr, idx):
            cnt = np.diff(ptr)[keep]
            rows = np.concatenate([idx[ptr[j]:ptr[j + 1]] for j in np.flatnonzero(keep)]
                                  or [np.zeros(0, np.int64)])
            return np.concatenate([[0], np.cumsum(cnt)]), rows
        ip, ix = strip(ip, ix)
        pB, r = strip(pB, r)
    assert ip.shape == pB.shape and np.array_equal(ip, pB), "column pointers differ"
    assert ix.shape == r.shape and np.array_equal(ix, r), "row indices differ"


def csc_equal(a, b):
    a = a.tocsc()
    b = b.tocsc()
    return (a.shape == b.shape and np.array_equal(a.indptr, b.indptr)
            and np.array_equal(a.indices, b.indices) and np.array_equal(a.data, b.data))
