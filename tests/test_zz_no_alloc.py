"""gbx_batch_run_noalloc (the batched application without device
allocation) through the C-ABI against the allocating entry
(gbx_batch_create + gbx_batch_run) on the same fields: bit for bit, on two
calls in a row, the second on a smaller input (fewer and smaller lattices),
so that the fixed scratch array's counter and lattice table left by the
first call cannot leak into the second."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIRST = [((8, 8, 8, 8), 8, 2), ((6, 4, 2, 8), 3, 2), ((4, 4, 4, 4), 4, 1), ((12, 12, 12, 12), 16, 1),
         ((2, 4, 4, 2), 5, 3), ((4, 8, 4, 4), 2, 8)]
SECOND = [((4, 4, 4, 4), 4, 1), ((2, 2, 4, 2), 6, 1)]


def fields(gbx, prec, lattices, seed0):
    objs = []
    for i, (dims, Ls, nrhs) in enumerate(lattices):
        g = gbx.Geometry(dims)
        U = gbx.GaugeField(g, prec).upload(gbx.gen_doubled_gauge(dims, seed0 + 2 * i))
        x = gbx.SpinorField(g, Ls, nrhs, prec).upload(gbx.gen_spinor(dims, Ls, nrhs, seed0 + 2 * i + 1))
        objs.append((g, U, x))
    return objs


def both(gbx, objs, prec, dag, tdiag):
    if tdiag:
        for o in objs:
            o[2].set_basis(gbx.BASIS_TDIAG)
    ya = [gbx.SpinorField(g, x.Ls, x.nrhs, prec) for g, U, x in objs]
    yb = [gbx.SpinorField(g, x.Ls, x.nrhs, prec) for g, U, x in objs]
    b = gbx.Batch([o[1] for o in objs], ya, [o[2] for o in objs], dag=dag)
    b.run()
    b.close()
    gbx.batch_run_noalloc([o[1] for o in objs], yb, [o[2] for o in objs], dag=dag)
    for a, c in zip(ya, yb):
        assert c.basis == a.basis
        assert np.array_equal(a.download() if not tdiag else a.set_basis(gbx.BASIS_CHIRAL).download(),
                              c.download() if not tdiag else c.set_basis(gbx.BASIS_CHIRAL).download())


@pytest.mark.parametrize("prec,tdiag", [("f32", False), ("f32", True), ("f64", True), ("f16", False)])
@pytest.mark.parametrize("dag", [False, True])
def test_noalloc_equals_batch_on_two_calls(gpu, gbx, prec, tdiag, dag):
    both(gbx, fields(gbx, prec, FIRST, 7000), prec, dag, tdiag)
    both(gbx, fields(gbx, prec, SECOND, 7100), prec, dag, tdiag)


def test_noalloc_rejects_more_lattices_than_its_table(gpu, gbx):
    objs = fields(gbx, "f32", [((2, 2, 2, 2), 2, 1)] * (gbx.NOALLOC_MAX_LATTICES + 1), 7200)
    ys = [gbx.SpinorField(g, x.Ls, x.nrhs, "f32") for g, U, x in objs]
    with pytest.raises(gbx.GbxError) as e:
        gbx.batch_run_noalloc([o[1] for o in objs], ys, [o[2] for o in objs])
    assert e.value.status == 1  # GBX_EINVAL
    # the largest table it takes
    n = gbx.NOALLOC_MAX_LATTICES
    gbx.batch_run_noalloc([o[1] for o in objs[:n]], ys[:n], [o[2] for o in objs[:n]])
    ref = gbx.SpinorField(objs[n - 1][0], 2, 1, "f32")
    gbx.dhop(objs[n - 1][1], ref, objs[n - 1][2], cfg=gbx.kcfg(1, 64, 0, 0))
    assert np.array_equal(ys[n - 1].download(), ref.download())


def test_noalloc_argument_errors(gpu, gbx):
    objs = fields(gbx, "f32", FIRST[:2], 7300)
    ys = [gbx.SpinorField(g, x.Ls, x.nrhs, "f32") for g, U, x in objs]
    with pytest.raises(gbx.GbxError) as e:  # an output aliasing another lattice's input
        gbx.batch_run_noalloc([o[1] for o in objs], [ys[0], objs[0][2]], [o[2] for o in objs])
    assert e.value.status in (3, 9)
    other = fields(gbx, "f16", FIRST[:1], 7400)[0]
    with pytest.raises(gbx.GbxError) as e:  # mixed precisions
        gbx.batch_run_noalloc([objs[0][1], other[1]], [ys[0], gbx.SpinorField(other[0], 8, 2, "f16")],
                              [objs[0][2], other[2]])
    assert e.value.status == 9
