"""LDS-tiled hopping kernel (kernel kind 2): a workgroup owns a 4D box of
sites; legs whose neighbour lies inside the box read the LDS copy of the
box, the others gather from memory.  Both paths load the same stored
chunks and run the same leg arithmetic in the same order as the VALU
kernel (kind 0), so the outputs must agree bit for bit, in every storage
precision; FP64 is also checked against the oracle.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

VALU = dict(spw=1, threads=64, xcd=0, nt=0)


def fields(gbx, p, prec, comp12=False, geom=None):
    geom = geom or gbx.Geometry(p.dims)
    U = gbx.GaugeField(geom, prec, comp12).upload(p.dbl)
    x = gbx.SpinorField(geom, p.Ls, p.nrhs, prec).upload(p.psi)
    y = gbx.SpinorField(geom, p.Ls, p.nrhs, prec)
    return geom, U, x, y


def chiral(gbx, f):
    f.set_basis(gbx.BASIS_CHIRAL)
    return f.download()


def both(gbx, U, x, y, dag, cfg):
    gbx.dhop(U, y, x, dag=dag, cfg=cfg)
    got = y.download()
    gbx.dhop(U, y, x, dag=dag, cfg=gbx.kcfg(**VALU))
    return got, y.download()


@pytest.mark.parametrize("prec", ["f64", "f32", "f16", "bf16"])
@pytest.mark.parametrize("tile", [(2, 2, 2, 2), (4, 2, 2, 2), (2, 2, 2, 4), (3, 3, 3, 1), (1, 1, 1, 1)])
@pytest.mark.parametrize("dag", [False, True])
def test_tile_bitexact_vs_valu(gpu, gbx, problem, prec, tile, dag):
    p = problem((12, 6, 6, 4), 4, nrhs=2)
    geom, U, x, y = fields(gbx, p, prec)
    got, ref = both(gbx, U, x, y, dag, gbx.kcfg(0, 256, 1, 1, kind=2, tile=tile))
    assert np.array_equal(got, ref)
    if prec == "f64":
        assert rel_l2(got[1], p.oracle_lex(rhs=1, dag=dag)) <= 1e-14


@pytest.mark.parametrize("mc,threads", [(0, 512), (4, 64), (2, 128), (3, 64)])
@pytest.mark.parametrize("comp12", [False, True])
def test_tile_minor_chunks_and_compressed_links(gpu, gbx, problem, mc, threads, comp12):
    """Minor chunks (a ragged last chunk for mc=3), several items per thread
    (threads < tile sites x chunk), 12-real links."""
    p = problem((4, 4, 4, 4), 4, nrhs=2)
    for prec in ("f64", "f32"):
        geom, U, x, y = fields(gbx, p, prec, comp12)
        got, ref = both(gbx, U, x, y, False, gbx.kcfg(0, threads, 1, 0, kind=2, mc=mc, tile=(2, 2, 2, 2)))
        assert np.array_equal(got, ref), (prec, mc, threads)


@pytest.mark.parametrize("dims,tile", [((2, 4, 4, 4), (2, 2, 2, 2)), ((4, 4, 2, 2), (4, 2, 2, 2)),
                                       ((2, 2, 2, 2), (2, 2, 2, 2)), ((1, 4, 4, 2), (1, 2, 2, 2))])
def test_tile_spanning_periodic_extents(gpu, gbx, problem, dims, tile):
    """A box extent equal to the lattice extent: the periodic neighbour is
    inside the box (both directions when the extent is 2, itself when 1)."""
    p = problem(dims, 3)
    geom, U, x, y = fields(gbx, p, "f64")
    for dag in (False, True):
        got, ref = both(gbx, U, x, y, dag, gbx.kcfg(0, 256, 0, 1, kind=2, tile=tile))
        assert np.array_equal(got, ref)
        assert rel_l2(got[0], p.oracle_lex(dag=dag)) <= 1e-14


def test_tile_any_slot_order(gpu, gbx, problem):
    """The box's sites need not be consecutive slots: multi-level and
    single-level site orders give the same result."""
    p = problem((8, 4, 4, 8), 2, nrhs=3)
    for geom in (gbx.Geometry(p.dims, levels=[(2, 2, 2, 2), (4, 4, 4, 4)]), gbx.Geometry(p.dims, block=(8, 1, 2, 4)),
                 gbx.Geometry(p.dims, block=(1, 1, 1, 1))):
        _, U, x, y = fields(gbx, p, "f32", geom=geom)
        got, ref = both(gbx, U, x, y, True, gbx.kcfg(0, 128, 1, 1, kind=2, tile=(2, 2, 2, 2)))
        assert np.array_equal(got, ref)


def test_tile_tuner_and_errors(gpu, gbx, problem, tmp_path):
    p = problem((4, 4, 4, 8), 4)
    geom, U, x, y = fields(gbx, p, "f64")
    best, us = gbx.tune(U, y, x, kinds="tile", cache_path=str(tmp_path / "t.jsonl"))
    assert best.kind == 2 and us > 0
    got, ref = both(gbx, U, x, y, False, best)
    assert np.array_equal(got, ref)
    again, _ = gbx.tune(U, y, x, kinds="tile", cache_path=str(tmp_path / "t.jsonl"))
    assert again.as_dict() == best.as_dict()
    for bad, status in ((gbx.kcfg(0, 256, kind=2, tile=(3, 2, 2, 2)), 1),   # 3 does not divide 4
                        (gbx.kcfg(0, 1024, kind=2, tile=(2, 2, 2, 2)), 1),  # FP64 launch bound is 512
                        (gbx.kcfg(0, 100, kind=2, tile=(2, 2, 2, 2)), 1),
                        (gbx.kcfg(0, 256, kind=2, tile=(4, 4, 4, 8)), 1)):  # LDS: 512 sites
        with pytest.raises(gbx.GbxError) as e:
            gbx.dhop(U, y, x, cfg=bad)
        assert e.value.status == status
    ge, go = gbx.Geometry.rb_pair(p.dims)
    Ue = gbx.GaugeField(ge, "f64").upload(p.dbl)
    xo = gbx.SpinorField(go, 4, 1, "f64").zero()
    with pytest.raises(gbx.GbxError) as e:
        gbx.dhop(Ue, gbx.SpinorField(ge, 4, 1, "f64"), xo, cfg=gbx.kcfg(0, 256, kind=2, tile=(2, 2, 2, 2)))
    assert e.value.status == 8  # GBX_EUNSUPPORTED


# Persistent form (pf=1): workgroups walk (box, minor chunk) units with two
# LDS buffers, the next unit's spinors and links staged by direct
# global->LDS loads while the current one computes.  Same loads, arithmetic
# and leg order as the VALU kernel: bit-identical outputs.
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("tile,mc,threads,wgpc", [((2, 2, 2, 2), 4, 256, 0), ((2, 2, 2, 2), 2, 64, 1),
                                                  ((4, 2, 2, 2), 2, 128, 2), ((3, 3, 3, 1), 8, 192, 0),
                                                  ((1, 1, 1, 1), 1, 64, 1), ((2, 2, 2, 4), 4, 256, 1)])
@pytest.mark.parametrize("tdiag", [False, True])
def test_tile_async_bitexact_vs_valu(gpu, gbx, problem, prec, tile, mc, threads, wgpc, tdiag):
    """(12, 6, 6, 8) x Ls 8 x 2 RHS: up to 1296 units, several per workgroup."""
    p = problem((12, 6, 6, 8), 8, nrhs=2)
    geom, U, x, y = fields(gbx, p, prec)
    if tdiag:
        x.set_basis(gbx.BASIS_TDIAG)
    print(f"async prec={prec} tile={tile} mc={mc} threads={threads} wgpc={wgpc} tdiag={tdiag}", flush=True)
    for dag in (False, True):
        y.zero()
        gbx.dhop(U, y, x, dag=dag, cfg=gbx.kcfg(0, threads, 1, 1, kind=2, mc=mc, pf=1, wgpc=wgpc, tile=tile))
        assert y.basis == x.basis
        got = chiral(gbx, y)  # the basis change is deterministic: equal bits stay equal
        gbx.dhop(U, y, x, dag=dag, cfg=gbx.kcfg(**VALU))
        assert np.array_equal(got, chiral(gbx, y)), dag
        if prec == "f64" and not tdiag:
            assert rel_l2(got[1], p.oracle_lex(rhs=1, dag=dag)) <= 1e-14


@pytest.mark.parametrize("dims,tile", [((2, 4, 4, 4), (2, 2, 2, 2)), ((2, 2, 2, 2), (2, 2, 2, 2)),
                                       ((1, 4, 4, 2), (1, 2, 2, 2))])
def test_tile_async_spanning_periodic_extents(gpu, gbx, problem, dims, tile):
    p = problem(dims, 4, nrhs=2)
    geom, U, x, y = fields(gbx, p, "f64")
    for dag in (False, True):
        got, ref = both(gbx, U, x, y, dag, gbx.kcfg(0, 128, 0, 0, kind=2, mc=4, pf=1, tile=tile))
        assert np.array_equal(got, ref)
        assert rel_l2(got[0], p.oracle_lex(dag=dag)) <= 1e-14


def test_tile_async_full_size_fp32_16rhs_tdiag(gpu, gbx):
    """24^4 x Ls 16 x 16 RHS FP32 in the TDIAG basis (8.2 GB per field):
    bit-identical to the VALU kernel on right-hand sides 0, 7 and 15."""
    dims = (24, 24, 24, 24)
    geom = gbx.Geometry(dims)
    U = gbx.GaugeField(geom, "f32").upload(gbx.gen_doubled_gauge(dims, 301))
    x = gbx.SpinorField(geom, 16, 16, "f32").fill_gaussian(1302).set_basis(gbx.BASIS_TDIAG)
    y = gbx.SpinorField(geom, 16, 16, "f32")
    gbx.dhop(U, y, x, cfg=gbx.kcfg(0, 512, 1, 1, kind=2, mc=32, pf=1, tile=(2, 2, 2, 2)))
    y.set_basis(gbx.BASIS_CHIRAL)
    got = [y.download_rhs(r) for r in (0, 7, 15)]
    gbx.dhop(U, y, x, cfg=gbx.kcfg(4, 256, 1, 0, mc=64, pf=1))
    y.set_basis(gbx.BASIS_CHIRAL)
    for r, g in zip((0, 7, 15), got):
        assert np.array_equal(g, y.download_rhs(r)), r
    for f in (U, x, y):
        f.close()


def test_tile_async_rejects_what_it_does_not_cover(gpu, gbx, problem):
    p = problem((4, 4, 4, 4), 4, nrhs=2)
    for prec, comp12, cfg, status in (
            ("f16", False, gbx.kcfg(0, 256, kind=2, mc=4, pf=1, tile=(2, 2, 2, 2)), 8),   # FP64 / FP32 only
            ("f32", True, gbx.kcfg(0, 256, kind=2, mc=4, pf=1, tile=(2, 2, 2, 2)), 8),    # 18-real links only
            ("f32", False, gbx.kcfg(0, 256, kind=2, mc=3, pf=1, tile=(2, 2, 2, 2)), 1),   # mc must divide M = 8
            ("f64", False, gbx.kcfg(0, 512, kind=2, mc=4, pf=1, tile=(2, 2, 2, 2)), 1),   # FP64 bound: 256 threads
            ("f32", False, gbx.kcfg(0, 1024, kind=2, mc=4, pf=1, tile=(2, 2, 2, 2)), 1),  # FP32 bound: 512
            ("f32", False, gbx.kcfg(0, 256, kind=2, mc=4, pf=2, tile=(2, 2, 2, 2)), 1),
            ("f64", False, gbx.kcfg(0, 256, kind=2, mc=8, pf=1, tile=(4, 4, 2, 2)), 1)):  # LDS: 2 x 64 sites x 8
        geom, U, x, y = fields(gbx, p, prec, comp12)
        with pytest.raises(gbx.GbxError) as e:
            gbx.dhop(U, y, x, cfg=cfg)
        assert e.value.status == status, (prec, comp12, cfg)
