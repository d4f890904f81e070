"""The slab-marching kernel (kernel kind 3, march.hip): a workgroup owns a
tile over three dimensions and marches along the fourth, slab by slab, with
the slabs in LDS.  It runs the VALU kernel's arithmetic in the VALU kernel's
leg order on the same loaded values, so its output must equal the VALU
kernel's bit for bit, for every tile, slab, segment, minor chunk and march
dimension -- and hence the oracle's (WilsonKernelsHandCpu.h:385-451) within
the storage tolerance.  Cases cover the slice-before and first-slice-after
reads at slab and segment edges, tiles that span a dimension (periodic
inside the tile), an extent-1 march dimension, multi-RHS minor chunks, both
spin bases and the dagger."""
import numpy as np
import pytest

from conftest import rel_l2, storage_tol

pytestmark = pytest.mark.gpu

EUNSUPPORTED, EINVAL = 8, 1

# (dims, Ls, nrhs, tile extents (x, y, z, t), march dim, segment, minor chunk)
CASES = [
    ((4, 4, 4, 8), 16, 1, (2, 2, 1, 4), 3, 0, 16),
    ((4, 4, 4, 8), 16, 1, (2, 2, 1, 4), 3, 4, 16),
    ((4, 4, 8, 4), 16, 1, (2, 2, 4, 1), 2, 0, 16),
    ((4, 6, 4, 8), 16, 1, (1, 1, 1, 4), 3, 0, 16),
    ((2, 4, 4, 8), 16, 1, (2, 2, 1, 4), 3, 8, 16),
    ((4, 4, 4, 4), 16, 4, (2, 2, 2, 1), 3, 0, 64),
    ((4, 4, 4, 4), 16, 8, (2, 2, 2, 1), 3, 2, 64),
    ((4, 4, 4, 4), 16, 4, (2, 1, 1, 2), 3, 0, 32),
    ((4, 4, 4, 8), 16, 2, (2, 2, 1, 4), 3, 0, 16),
    ((4, 4, 4, 2), 8, 4, (2, 2, 1, 2), 3, 0, 32),
    ((4, 4, 4, 1), 16, 4, (2, 2, 1, 1), 3, 0, 64),
]


def fields(gbx, p, prec, tdiag):
    g = gbx.Geometry(p.dims)
    U = gbx.GaugeField(g, prec).upload(p.dbl)
    x = gbx.SpinorField(g, p.Ls, p.nrhs, prec).upload(p.psi)
    if tdiag:
        x.set_basis(gbx.BASIS_TDIAG)
    y = gbx.SpinorField(g, p.Ls, p.nrhs, prec)
    return g, U, x, y


def chiral(gbx, f):
    f.set_basis(gbx.BASIS_CHIRAL)
    return f.download()


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("tdiag", [False, True])
@pytest.mark.parametrize("dag", [False, True])
def test_march_equals_valu_kernel(gpu, gbx, problem, case, prec, tdiag, dag):
    dims, Ls, nrhs, e, md, seg, mc = CASES[case]
    if prec == "f64" and (np.prod(e) // e[md]) * e[md] * mc > 256:
        pytest.skip("FP64 workgroups hold at most 256 threads")
    p = problem(dims, Ls, nrhs=nrhs)
    g, U, x, y = fields(gbx, p, prec, tdiag)
    cfg = gbx.march_cfg(e, md, mc, seg)
    print(f"case {CASES[case]} {prec} tdiag={tdiag} dag={dag}", flush=True)
    gbx.dhop(U, y, x, dag=dag, cfg=gbx.kcfg(1, 64, 0, 0))
    want = chiral(gbx, y)  # the basis change is deterministic: equal bits stay equal
    y.zero()
    gbx.dhop(U, y, x, dag=dag, cfg=cfg)
    assert y.basis == x.basis
    got = chiral(gbx, y)
    assert np.array_equal(got, want)
    tol = storage_tol(prec, p.psi)
    for r in range(nrhs):
        assert rel_l2(got[r], p.oracle_lex(dag=dag, rhs=r)) <= tol


def test_march_full_size_headline_shape(gpu, gbx, problem):
    """24^4 x Ls16 FP64 on the TDIAG basis: the march kernel on the 4x6x6x12
    traversal blocks the headline uses, against the VALU kernel bit for bit."""
    p = problem((24, 24, 24, 24), 16, gseed=201, sseed=202)
    g = gbx.Geometry(p.dims, block=(4, 6, 6, 12))
    U = gbx.GaugeField(g, "f64").upload(p.dbl)
    x = gbx.SpinorField(g, 16, 1, "f64").upload(p.psi).set_basis(gbx.BASIS_TDIAG)
    y = gbx.SpinorField(g, 16, 1, "f64")
    gbx.dhop(U, y, x, cfg=gbx.kcfg(4, 64, 1, 0, pf=2))
    want = chiral(gbx, y)
    for e, md, seg in (((2, 2, 1, 4), 3, 0), ((2, 2, 4, 1), 2, 12), ((2, 1, 1, 4), 3, 12)):
        y.zero()
        gbx.dhop(U, y, x, cfg=gbx.march_cfg(e, md, 16, seg))
        assert np.array_equal(chiral(gbx, y), want), (e, md, seg)


def test_march_rejects_what_it_does_not_cover(gpu, gbx, problem):
    p = problem((4, 4, 4, 8), 16)
    g, U, x, y = fields(gbx, p, "f64", False)
    for bad in (gbx.march_cfg((2, 2, 1, 4), 3, 16, 3),      # segment not a multiple of the slab
                gbx.march_cfg((2, 2, 1, 2), 3, 16),         # 2 x 16 minor: not whole wavefronts
                gbx.march_cfg((2, 2, 2, 4), 3, 16),         # 512 threads: above FP64's 256
                gbx.march_cfg((3, 2, 1, 4), 3, 16)):        # 3 does not divide 4
        with pytest.raises(gbx.GbxError) as e:
            gbx.dhop(U, y, x, cfg=bad)
        assert e.value.status == EINVAL
    g16, U16, x16, y16 = fields(gbx, p, "f16", False)
    with pytest.raises(gbx.GbxError) as e:
        gbx.dhop(U16, y16, x16, cfg=gbx.march_cfg((2, 2, 1, 4), 3, 16))
    assert e.value.status == EUNSUPPORTED
    U12 = gbx.GaugeField(g, "f64", comp12=True).upload(p.dbl)
    with pytest.raises(gbx.GbxError) as e:
        gbx.dhop(U12, y, x, cfg=gbx.march_cfg((2, 2, 1, 4), 3, 16))
    assert e.value.status == EUNSUPPORTED
    ge, go = gbx.Geometry.rb_pair(p.dims)
    Ue = gbx.GaugeField(ge, "f64").upload(p.dbl)
    xo = gbx.SpinorField(go, 16, 1, "f64")
    ye = gbx.SpinorField(ge, 16, 1, "f64")
    with pytest.raises(gbx.GbxError) as e:
        gbx.dhop(Ue, ye, xo, cfg=gbx.march_cfg((1, 2, 1, 4), 3, 16))
    assert e.value.status == EUNSUPPORTED


def test_tuner_searches_the_march_kernel(gpu, gbx, problem, tmp_path):
    p = problem((8, 8, 8, 8), 16)
    g, U, x, y = fields(gbx, p, "f64", True)
    cfg, us = gbx.tune(U, y, x, kinds="march", cache_path=str(tmp_path / "m.jsonl"))
    assert cfg.kind == 3 and us > 0
    gbx.dhop(U, y, x, cfg=cfg)
    assert rel_l2(chiral(gbx, y)[0], p.oracle_lex()) <= 1e-14
