"""Edge cases of the device path against the oracle (the CPU restatement of
WilsonKernelsHandCpu.h:385-451): the largest minor extent a field takes
(Ls * nrhs = 4096), a minor extent that is no multiple of anything the
kernels tile by (4095 = 63 x 65), lattices with extent-1 and extent-2
dimensions (a site is its own neighbour, or both neighbours are one site),
and the limit itself.
"""
import numpy as np
import pytest

from conftest import rel_l2, storage_tol

pytestmark = pytest.mark.gpu



def apply(gbx, p, prec, cfg=None, dag=False):
    geom = gbx.Geometry(p.dims)
    U = gbx.GaugeField(geom, prec).upload(p.dbl)
    src = gbx.SpinorField(geom, p.Ls, p.nrhs, prec).upload(p.psi)
    dst = gbx.SpinorField(geom, p.Ls, p.nrhs, prec)
    gbx.dhop(U, dst, src, dag=dag, cfg=cfg)
    return dst.download()


@pytest.mark.parametrize("prec", ["f64", "f32", "f16", "bf16"])
def test_largest_minor_extent(gpu, gbx, problem, prec):
    p = problem((2, 2, 2, 4), 64, nrhs=64)  # Ls * nrhs = 4096
    got = apply(gbx, p, prec)
    for r in (0, 31, 63):
        err = rel_l2(got[r], p.oracle_lex(rhs=r))
        print(f"prec={prec} rhs={r} rel_l2={err:.3e}")
        assert err <= storage_tol(prec, p.psi)


def test_ragged_minor_extent(gpu, gbx, problem):
    p = problem((2, 2, 4, 2), 63, nrhs=65)  # Ls * nrhs = 4095
    for dag in (False, True):
        got = apply(gbx, p, "f64", dag=dag)
        for r in (0, 64):
            assert rel_l2(got[r], p.oracle_lex(rhs=r, dag=dag)) <= 1e-14


@pytest.mark.parametrize("prec", ["f16", "bf16"])
def test_mfma_largest_minor_extent(gpu, gbx, problem, prec):
    p = problem((2, 2, 2, 4), 64, nrhs=64)
    got = apply(gbx, p, prec, cfg=gbx.kcfg(1, 256, kind=1))
    for r in (0, 63):
        assert rel_l2(got[r], p.oracle_lex(rhs=r)) <= storage_tol(prec, p.psi)


@pytest.mark.parametrize("dims", [(1, 2, 4, 4), (4, 1, 1, 4), (2, 2, 2, 2), (1, 1, 1, 1)])
@pytest.mark.parametrize("dag", [False, True])
def test_tiny_extents(gpu, gbx, problem, dims, dag):
    p = problem(dims, 3)
    got = apply(gbx, p, "f64", dag=dag)[0]
    assert rel_l2(got, p.oracle_lex(dag=dag)) <= 1e-14


def test_minor_extent_limit(gpu, gbx):
    geom = gbx.Geometry((2, 2, 2, 2))
    with pytest.raises(gbx.GbxError):
        gbx.SpinorField(geom, 64, 65, "f64")
    with pytest.raises(gbx.GbxError):
        gbx.SpinorField(geom, 0, 1, "f64")
