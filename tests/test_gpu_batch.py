"""Batched application over uneven lattices in one persistent, graph-captured
launch (gbx_batch_*): every lattice's output equals its own single-lattice
application bit for bit, and matches the oracle; graph replays repeat it."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

LATTICES = [((4, 4, 4, 4), 4, 1), ((6, 4, 2, 8), 3, 2), ((8, 8, 8, 8), 8, 1), ((2, 4, 4, 2), 5, 3), ((4, 8, 4, 4), 2, 8)]


def build(gbx, prec, lattices, seed0=5000):
    objs = []
    for i, (dims, Ls, nrhs) in enumerate(lattices):
        g = gbx.Geometry(dims)
        dbl = gbx.gen_doubled_gauge(dims, seed0 + 2 * i)
        psi = gbx.gen_spinor(dims, Ls, nrhs, seed0 + 2 * i + 1)
        U = gbx.GaugeField(g, prec).upload(dbl)
        x = gbx.SpinorField(g, Ls, nrhs, prec).upload(psi)
        y = gbx.SpinorField(g, Ls, nrhs, prec)
        objs.append((g, U, x, y, dbl, psi))
    return objs


@pytest.mark.parametrize("prec", ["f32", "f16", "f64"])
@pytest.mark.parametrize("dag", [False, True])
def test_batch_equals_single_lattice_runs(gpu, gbx, prec, dag):
    objs = build(gbx, prec, LATTICES)
    b = gbx.Batch([o[1] for o in objs], [o[3] for o in objs], [o[2] for o in objs], dag=dag)
    items, wgs = b.info()
    assert items > 0 and wgs > 0
    b.run()
    got = [o[3].download() for o in objs]
    for (g, U, x, y, dbl, psi), gb in zip(objs, got):
        ref = gbx.SpinorField(g, x.Ls, x.nrhs, prec)
        gbx.dhop(U, ref, x, dag=dag, cfg=gbx.kcfg(1, 64, 0, 0))
        assert np.array_equal(ref.download(), gb)


def test_batch_matches_oracle(gpu, gbx, orc, problem):
    objs = build(gbx, "f32", LATTICES)
    b = gbx.Batch([o[1] for o in objs], [o[3] for o in objs], [o[2] for o in objs])
    b.run()
    for (g, U, x, y, dbl, psi), (dims, Ls, nrhs) in zip(objs, LATTICES):
        out = y.download()
        for r in range(nrhs):
            p = problem(dims, Ls)
            p.dbl, p.psi = dbl, psi[r:r + 1]
            err = rel_l2(out[r], p.oracle_lex())
            print(f"lattice {dims} Ls={Ls} rhs={r}: rel_l2 vs FP64 oracle {err:.2e}")
            assert err <= 1e-6


def test_graph_replays_are_stable(gpu, gbx):
    """Replay the captured graph several times with eager work between the
    replays; every replay reproduces the first."""
    objs = build(gbx, "f32", LATTICES[:3])
    b = gbx.Batch([o[1] for o in objs], [o[3] for o in objs], [o[2] for o in objs])
    b.run()
    first = [o[3].download() for o in objs]
    for k in range(4):
        for o in objs:  # eager work on the same buffers between replays
            o[3].zero()
            o[0].sync()
        b.run()
        for o, f in zip(objs, first):
            assert np.array_equal(o[3].download(), f)
    us = b.time(3)
    assert us > 0


def test_batch_argument_errors(gpu, gbx):
    objs = build(gbx, "f32", LATTICES[:2])
    other = build(gbx, "f16", LATTICES[:1])
    with pytest.raises(gbx.GbxError):  # mixed precisions
        gbx.Batch([objs[0][1], other[0][1]], [objs[0][3], other[0][3]], [objs[0][2], other[0][2]])
    with pytest.raises(gbx.GbxError) as e:  # an output aliasing another lattice's input
        gbx.Batch([objs[0][1], objs[0][1]], [objs[0][3], objs[0][2]], [objs[0][2], objs[0][2]])
    assert e.value.status in (3, 9)


def test_batch_rechecks_the_spin_basis_on_every_run(gpu, gbx):
    """A batch captured on TDIAG fields: reading an output (converted to the
    chiral basis in place) and running again restamps the output, so the next
    read is right; an input converted back to chiral is refused (the graph
    holds the TDIAG kernel) until it is converted again."""
    objs = build(gbx, "f32", LATTICES[:3])
    for o in objs:
        o[2].set_basis(gbx.BASIS_TDIAG)
    b = gbx.Batch([o[1] for o in objs], [o[3] for o in objs], [o[2] for o in objs])
    want = []
    for g, U, x, y, dbl, psi in objs:
        ref = gbx.SpinorField(g, x.Ls, x.nrhs, "f32")
        gbx.dhop(U, ref, x, cfg=gbx.kcfg(1, 64, 0, 0))
        want.append(ref.set_basis(gbx.BASIS_CHIRAL).download())
    for _ in range(2):
        b.run()
        for o, w in zip(objs, want):
            assert o[3].basis == gbx.BASIS_TDIAG
            assert np.array_equal(o[3].set_basis(gbx.BASIS_CHIRAL).download(), w)
    objs[1][2].set_basis(gbx.BASIS_CHIRAL)
    with pytest.raises(gbx.GbxError) as e:
        b.run()
    assert e.value.status == 9  # GBX_EMISMATCH
    with pytest.raises(gbx.GbxError) as e:
        b.time(1)
    assert e.value.status == 9
    # back in the captured basis (the round trip re-rounds the input, so the
    # reference is recomputed from the converted input)
    g, U, x, y = objs[1][:4]
    x.set_basis(gbx.BASIS_TDIAG)
    b.run()
    ref = gbx.SpinorField(g, x.Ls, x.nrhs, "f32")
    gbx.dhop(U, ref, x, cfg=gbx.kcfg(1, 64, 0, 0))
    assert np.array_equal(y.set_basis(gbx.BASIS_CHIRAL).download(), ref.set_basis(gbx.BASIS_CHIRAL).download())


def test_batch_checks_each_lattice_like_the_single_entries(gpu, gbx):
    """Ls and nrhs must match between a lattice's input and output (equal
    Ls * nrhs is not enough), and fields must be allocated."""
    objs = build(gbx, "f32", LATTICES[:1])
    g, U, x, y, dbl, psi = objs[0]
    assert (x.Ls, x.nrhs) == (4, 1)
    bad = gbx.SpinorField(g, 2, 2, "f32")  # Ls * nrhs = 4 as well
    with pytest.raises(gbx.GbxError) as e:
        gbx.Batch([U], [bad], [x])
    assert e.value.status == 9


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_batch_launch_shapes_and_tuner(gpu, gbx, prec):
    """Every launch shape of the persistent kernel (threads, gather depth,
    persistent workgroups per CU, points per work item, store kind) computes
    the same bits as the single-lattice kernel; the tuner keeps one of them."""
    objs = build(gbx, prec, LATTICES)
    b = gbx.Batch([o[1] for o in objs], [o[3] for o in objs], [o[2] for o in objs])
    want = []
    for g, U, x, y, dbl, psi in objs:
        ref = gbx.SpinorField(g, x.Ls, x.nrhs, prec)
        gbx.dhop(U, ref, x, cfg=gbx.kcfg(1, 64, 0, 0))
        want.append(ref.download())
    assert b.get_cfg().threads == 256
    for threads, pf, wgpc, points, nt in ((512, 2, 3, 1024, 0), (256, 2, 8, 256, 1), (512, 1, 2, 512, 1),
                                          (256, 1, 0, 64, 0)):
        b.set_cfg(gbx.kcfg(points, threads, 0, nt, pf=pf, wgpc=wgpc))
        for o in objs:
            o[3].zero()
        b.run()
        for o, w in zip(objs, want):
            assert np.array_equal(o[3].download(), w), (threads, pf, wgpc, points, nt)
    cfg, us = b.tune(nrep=2)
    assert us > 0 and cfg.threads in (256, 512) and cfg.pf in (1, 2)
    assert b.get_cfg().as_dict() == cfg.as_dict()
    for o in objs:
        o[3].zero()
    b.run()
    for o, w in zip(objs, want):
        assert np.array_equal(o[3].download(), w)
    for bad in (gbx.kcfg(256, 128, pf=1), gbx.kcfg(256, 256, pf=0), gbx.kcfg(32, 256, pf=1), gbx.kcfg(256, 256, pf=1, wgpc=17)):
        with pytest.raises(gbx.GbxError) as e:
            b.set_cfg(bad)
        assert e.value.status == 1
