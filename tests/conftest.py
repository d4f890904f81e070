import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gbx():
    import gridbench_amd
    return gridbench_amd


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle


@pytest.fixture(scope="session")
def gpu(gbx):
    if gbx.device_count() == 0:
        pytest.skip("no GPU in this environment")
    return True


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def round_storage(x, prec):
    """x rounded to nearest (ties to even) in the storage precision, as float64."""
    x = np.asarray(x, dtype=np.float64)
    if prec == "f64":
        return x
    if prec == "f16":
        return x.astype(np.float16).astype(np.float64)
    f = x.astype(np.float32)
    if prec == "f32":
        return f.astype(np.float64)
    b = f.view(np.uint32).astype(np.uint64)  # bf16: round the float32's low 16 bits away
    b = (((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return b.view(np.float32).astype(np.float64)


# Dhop's error with 16-bit storage and FP32 arithmetic comes from rounding the
# spinor, the links and the output once each: about sqrt(3) = 1.7 x the
# relative rounding error of storing the input field.  The bound allows 2.5 x;
# a kernel that also rounded its sums to 16 bits would exceed it.
ROUNDING_FACTOR_16 = 2.5


def storage_tol(prec, x):
    """Relative L2 bound of Dhop against the FP64 oracle for a field x (the
    exact input values) stored in prec: 1e-14 FP64, 1e-6 FP32, and for FP16 /
    BF16 ROUNDING_FACTOR_16 x the relative error of storing x itself
    (measured on up to 2M of its values)."""
    if prec == "f64":
        return 1e-14
    if prec == "f32":
        return 1e-6
    s = np.asarray(x, dtype=np.float64).reshape(-1)[:1 << 21]
    return ROUNDING_FACTOR_16 * rel_l2(round_storage(s, prec), s)


class Problem:
    """Seeded inputs of one lattice in both layouts, plus oracle helpers."""

    def __init__(self, gbx, orc, dims, Ls, nrhs=1, gseed=101, sseed=102):
        self.dims, self.Ls, self.nrhs = tuple(dims), Ls, nrhs
        self.V = int(np.prod(dims))
        self.links = gbx.gen_su3_links(dims, gseed)
        self.dbl = gbx.double_links(dims, self.links)
        self.psi = gbx.gen_spinor(dims, Ls, nrhs, sseed)  # [nrhs][V][Ls][4][3][2]
        self.gbx, self.orc = gbx, orc

    def oracle_lex(self, dag=False, rhs=0, dtype=np.float64, nthreads=8, psi=None):
        """Oracle Dhop of right-hand side rhs, returned in lexicographic layout."""
        O = self.orc
        nbr, prm = O.tables(self.dims, 1, self.Ls)
        U = O.pack_gauge(self.dims, 1, self.dbl).astype(dtype)
        src = self.psi[rhs] if psi is None else psi
        inp = O.pack_spinor(self.dims, 1, self.Ls, src).astype(dtype)
        out = O.hand(O.aligned_copy(U), O.aligned_copy(inp), nbr, prm, self.V, self.Ls, 1, dag=dag, nthreads=nthreads)
        return O.unpack_spinor(self.dims, 1, self.Ls, out.astype(np.float64)).reshape(self.V, self.Ls, 4, 3, 2)


@pytest.fixture
def problem(gbx, orc):
    def make(dims, Ls, nrhs=1, gseed=101, sseed=102):
        return Problem(gbx, orc, dims, Ls, nrhs, gseed, sseed)
    return make
