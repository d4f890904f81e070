"""Fields in the GBX_BASIS_TDIAG spin basis (gamma_t diagonal).

A TDIAG field stores psi' = S psi, S = [[1, 1], [1, -1]] / sqrt(2) on the
spin pairs (0, 2) and (1, 3).  The hopping term commutes with that change
of basis up to rounding (S Dhop S^-1 is Dhop written with S's gammas), so a
TDIAG field's Dhop, set back to the chiral basis, must agree with the
oracle (the reference's algorithm, WilsonKernelsHandCpu.h:426-446) within
the storage precision's tolerance: 1e-14 for FP64, 1e-6 for FP32.  In the
TDIAG kernel the +-t legs gather half of each spinor and the x, y, z legs
use the projectors of the opposite sign, so a mistake in S, in a projector
sign or in the half a t leg reads fails these comparisons.

The reference stores chiral spinors only; it has no counterpart.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

EUNSUPPORTED, EMISMATCH, EINVAL = 8, 9, 1
TOL = {"f64": 1e-14, "f32": 1e-6}


def fields(gbx, p, prec, comp12=False, geom=None):
    geom = geom or gbx.Geometry(p.dims)
    U = gbx.GaugeField(geom, prec, comp12).upload(p.dbl)
    x = gbx.SpinorField(geom, p.Ls, p.nrhs, prec).upload(p.psi)
    y = gbx.SpinorField(geom, p.Ls, p.nrhs, prec)
    return geom, U, x, y


def chiral_download(gbx, f):
    f.set_basis(gbx.BASIS_CHIRAL)
    return f.download()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_set_basis_keeps_norms_and_round_trips(gpu, gbx, problem, prec):
    p = problem((4, 4, 4, 8), 4, nrhs=2)
    geom, U, x, y = fields(gbx, p, prec)
    n0 = x.norm2()
    assert x.basis == gbx.BASIS_CHIRAL
    x.set_basis(gbx.BASIS_TDIAG)
    assert x.basis == gbx.BASIS_TDIAG
    # S is orthogonal: the norm is that of the chiral values
    assert abs(x.norm2() / n0 - 1) <= (1e-14 if prec == "f64" else 1e-6)
    x.set_basis(gbx.BASIS_TDIAG)  # no-op: already in that basis
    back = chiral_download(gbx, x)
    want = p.psi.astype(np.float32).astype(np.float64) if prec == "f32" else p.psi
    assert rel_l2(back, want) <= (6e-16 if prec == "f64" else 3e-7)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("dag", [False, True])
@pytest.mark.parametrize("comp12", [False, True])
@pytest.mark.parametrize("spw,threads,pf", [(1, 64, 0), (4, 64, 1), (4, 64, 2), (16, 256, 2)])
def test_dhop_matches_oracle(gpu, gbx, problem, prec, dag, comp12, spw, threads, pf):
    p = problem((4, 4, 4, 8), 4, nrhs=2)
    geom, U, x, y = fields(gbx, p, prec, comp12)
    cfg = gbx.kcfg(spw, threads, 1, 1, 0, 0, pf=pf)
    gbx.dhop(U, y, x, dag=dag, cfg=cfg)
    ref = y.download()  # chiral kernel, same configuration
    x.set_basis(gbx.BASIS_TDIAG)
    gbx.dhop(U, y, x, dag=dag, cfg=cfg)
    assert y.basis == gbx.BASIS_TDIAG
    got = chiral_download(gbx, y)
    for r in range(p.nrhs):
        assert rel_l2(got[r], p.oracle_lex(dag=dag, rhs=r)) <= TOL[prec], (prec, dag, comp12, pf, r)
    assert rel_l2(got, ref) <= TOL[prec]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_default_config_axpby_and_xpay(gpu, gbx, problem, prec):
    """gbx_dhop with its default configuration, out = a in + b Dhop in and
    out = a x + b Dhop in against the same operations on chiral fields."""
    p = problem((4, 6, 4, 8), 4, nrhs=1)
    geom, U, x, y = fields(gbx, p, prec)
    z = gbx.SpinorField(geom, p.Ls, p.nrhs, prec).upload(p.psi[:, ::-1].copy())
    refs = []
    gbx.dhop(U, y, x)
    refs.append(y.download())
    gbx.dhop_axpby(U, y, x, 1.0, -0.12, dag=True)
    refs.append(y.download())
    gbx.dhop_xpay(U, y, z, x, 0.5, 0.25)
    refs.append(y.download())
    x.set_basis(gbx.BASIS_TDIAG)
    z.set_basis(gbx.BASIS_TDIAG)
    gbx.dhop(U, y, x)
    got = [chiral_download(gbx, y)]
    gbx.dhop_axpby(U, y, x, 1.0, -0.12, dag=True)
    got.append(chiral_download(gbx, y))
    gbx.dhop_xpay(U, y, z, x, 0.5, 0.25)
    got.append(chiral_download(gbx, y))
    for g, r in zip(got, refs):
        assert rel_l2(g, r) <= TOL[prec]
    assert rel_l2(got[0][0], p.oracle_lex()) <= TOL[prec]


def test_red_black_halves(gpu, gbx, problem):
    """DhopOE on a red-black pair: the odd half's TDIAG input, the even half's output."""
    p = problem((4, 4, 4, 8), 4)
    ge, go = gbx.Geometry.rb_pair(p.dims)
    even, odd = gbx.rb_split(p.dims, p.psi)
    U = gbx.GaugeField(ge, "f64").upload(p.dbl)
    xo = gbx.SpinorField(go, p.Ls, 1, "f64").upload(odd)
    ye = gbx.SpinorField(ge, p.Ls, 1, "f64")
    gbx.dhop(U, ye, xo)
    ref = ye.download()
    xo.set_basis(gbx.BASIS_TDIAG)
    gbx.dhop(U, ye, xo)
    assert rel_l2(chiral_download(gbx, ye), ref) <= 1e-14


@pytest.mark.parametrize("dims,grid", [((4, 4, 4, 8), (1, 1, 1, 2)), ((4, 4, 6, 6), (1, 1, 2, 2))])
def test_domain_parts_with_whole_faces(gpu, gbx, problem, dims, grid):
    """Whole-spinor faces are copies, so they move TDIAG values unchanged:
    the parts' outputs put together agree with the oracle."""
    p = problem(dims, 4)
    coords = gbx.grid_coords(grid)
    geoms = {c: gbx.Geometry.part_of(dims, grid, c) for c in coords}
    loc = gbx.part_split(dims, grid, p.psi)
    U = {c: gbx.GaugeField(g, "f64").upload(p.dbl) for c, g in geoms.items()}
    src = {c: gbx.SpinorField(g, p.Ls, 1, "f64").upload(loc[c]).set_basis(gbx.BASIS_TDIAG) for c, g in geoms.items()}
    dst = {c: gbx.SpinorField(g, p.Ls, 1, "f64") for c, g in geoms.items()}
    gbx.halo_exchange(src)
    for c in coords:
        gbx.dhop(U[c], dst[c], src[c])
    got = gbx.part_merge(dims, grid, {c: chiral_download(gbx, d) for c, d in dst.items()})
    assert rel_l2(got[0], p.oracle_lex()) <= 1e-14


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("dag", [False, True])
@pytest.mark.parametrize("tile,mc", [((2, 2, 2, 2), 0), ((4, 2, 2, 2), 4), ((3, 3, 3, 1), 2), ((2, 2, 2, 4), 0)])
def test_tiled_kernel_bitexact_vs_valu(gpu, gbx, problem, prec, dag, tile, mc):
    """The tiled kernel's t legs read half a spinor from LDS or memory; same
    arithmetic in the same order as the VALU kernel on TDIAG fields."""
    p = problem((12, 6, 6, 4), 4, nrhs=2)
    geom, U, x, y = fields(gbx, p, prec)
    x.set_basis(gbx.BASIS_TDIAG)
    gbx.dhop(U, y, x, dag=dag, cfg=gbx.kcfg(0, 256, 1, 1, kind=2, mc=mc, tile=tile))
    got = chiral_download(gbx, y)
    gbx.dhop(U, y, x, dag=dag, cfg=gbx.kcfg(1, 64, 0, 0))
    assert np.array_equal(got, chiral_download(gbx, y))
    assert rel_l2(got[1], p.oracle_lex(dag=dag, rhs=1)) <= TOL[prec]


def test_tuner_on_tdiag_fields(gpu, gbx, problem):
    p = problem((4, 4, 4, 8), 4, nrhs=2)
    geom, U, x, y = fields(gbx, p, "f32")
    x.set_basis(gbx.BASIS_TDIAG)
    best, us = gbx.tune(U, y, x)
    assert best.kind in (0, 2) and us > 0
    gbx.dhop(U, y, x)
    assert rel_l2(chiral_download(gbx, y)[1], p.oracle_lex(rhs=1)) <= 1e-6
    best, us = gbx.tune(U, y, x, kinds="tile")
    assert best.kind == 2
    with pytest.raises(gbx.GbxError) as e:
        gbx.tune(U, y, x, kinds="mfma")  # the matrix-core form takes 16-bit (hence chiral) fields
    assert e.value.status == EUNSUPPORTED


def test_entries_without_a_tdiag_form_refuse(gpu, gbx, problem):
    p = problem((4, 4, 4, 8), 4)
    geom, U, x, y = fields(gbx, p, "f64")
    x.set_basis(gbx.BASIS_TDIAG)
    host = np.empty_like(p.psi)
    z0 = gbx.SpinorField(geom, p.Ls, 1, "f64")
    z1 = gbx.SpinorField(geom, p.Ls, 1, "f64").upload(p.psi)

    def status(fn):
        with pytest.raises(gbx.GbxError) as e:
            fn()
        return e.value.status

    assert status(lambda: x.download()) == EUNSUPPORTED
    assert status(lambda: x.download_rhs(0)) == EUNSUPPORTED
    assert status(lambda: x.upload_rhs(0, host[0])) == EUNSUPPORTED
    assert status(lambda: gbx.dwf_mooee(y, x)) == EUNSUPPORTED
    assert status(lambda: gbx.dwf_apply(U, y, x)) == EUNSUPPORTED
    assert status(lambda: gbx.cg_solve(U, None, y, x, 0.12)) == EMISMATCH  # x chiral, b TDIAG
    ge, go = gbx.Geometry.rb_pair(p.dims)
    Ue, Uo = gbx.GaugeField(ge, "f64").upload(p.dbl), gbx.GaugeField(go, "f64").upload(p.dbl)
    even, odd = gbx.rb_split(p.dims, p.psi)
    be = gbx.SpinorField(ge, p.Ls, 1, "f64").upload(even).set_basis(gbx.BASIS_TDIAG)
    bo = gbx.SpinorField(go, p.Ls, 1, "f64").upload(odd).set_basis(gbx.BASIS_TDIAG)
    xe, xo = gbx.SpinorField(ge, p.Ls, 1, "f64"), gbx.SpinorField(go, p.Ls, 1, "f64")
    assert status(lambda: gbx.cg_solve_eo((Ue, Uo), None, (xe, xo), (be, bo), 0.12)) == EMISMATCH  # x chiral
    xe.set_basis(gbx.BASIS_TDIAG)
    xo.set_basis(gbx.BASIS_TDIAG)
    assert status(lambda: gbx.dwf_solve_eo((Ue, Uo), None, (xe, xo), (be, bo))) == EUNSUPPORTED
    assert status(lambda: gbx.Batch([U, U], [y, z0], [x, z1])) == EMISMATCH  # inputs in two bases
    assert status(lambda: y.axpy(1.0, x)) == EMISMATCH  # y chiral, x TDIAG
    assert status(lambda: x.set_basis(7)) == EINVAL
    h = gbx.SpinorField(geom, p.Ls, 1, "f16")
    assert status(lambda: h.set_basis(gbx.BASIS_TDIAG)) == EUNSUPPORTED
    assert status(lambda: h.copy_from(x)) == EUNSUPPORTED
    # a slot range of an output in the other basis
    assert status(lambda: gbx.dhop_range(U, y, x, 0, 8)) == EMISMATCH
    # copies take the source's basis; a whole upload or a fill leaves the field chiral
    z = gbx.SpinorField(geom, p.Ls, 1, "f64").copy_from(x)
    assert z.basis == gbx.BASIS_TDIAG
    z.upload(p.psi)
    assert z.basis == gbx.BASIS_CHIRAL
    z.set_basis(gbx.BASIS_TDIAG).fill_gaussian(3)
    assert z.basis == gbx.BASIS_CHIRAL
    # a projected face of a domain part
    pg = gbx.Geometry.part_of(p.dims, (1, 1, 1, 2), (0, 0, 0, 0))
    xp = gbx.SpinorField(pg, p.Ls, 1, "f64").upload(gbx.part_split(p.dims, (1, 1, 1, 2), p.psi)[(0, 0, 0, 0)])
    xp.set_basis(gbx.BASIS_TDIAG)
    buf = gbx.DeviceBuffer(xp.halo_bytes(3, proj=True))
    assert status(lambda: xp.halo_pack(3, 1, buf.ptr, proj=True)) == EUNSUPPORTED


def test_cg_on_tdiag_fields(gpu, gbx, problem):
    """Defect-correction CG (FP64 outer, FP32 inner, and FP64 alone) on
    TDIAG fields: the solution, set back to the chiral basis, is the chiral
    solve's; 16-bit inner fields cannot hold the basis."""
    p = problem((4, 4, 4, 8), 4)
    geom = gbx.Geometry(p.dims)
    U64 = gbx.GaugeField(geom, "f64").upload(p.dbl)
    U32 = gbx.GaugeField(geom, "f32").upload(p.dbl)
    U16 = gbx.GaugeField(geom, "f16").upload(p.dbl)
    b = gbx.SpinorField(geom, p.Ls, 1, "f64").upload(p.psi)
    x = gbx.SpinorField(geom, p.Ls, 1, "f64")
    x.zero()
    st = gbx.cg_solve(U64, U32, x, b, 0.12, tol=1e-10)
    want = x.download()
    b.set_basis(gbx.BASIS_TDIAG)
    for lo in (U32, None):
        x.zero()
        x.set_basis(gbx.BASIS_TDIAG)
        st = gbx.cg_solve(U64, lo, x, b, 0.12, tol=1e-10)
        assert st["true_resid"] <= 1e-10
        assert x.basis == gbx.BASIS_TDIAG
        assert rel_l2(chiral_download(gbx, x), want) <= 1e-8
    x.set_basis(gbx.BASIS_TDIAG)
    with pytest.raises(gbx.GbxError) as e:
        gbx.cg_solve(U64, U16, x, b, 0.12)
    assert e.value.status == EUNSUPPORTED


def test_even_odd_wilson_solve_on_tdiag_fields(gpu, gbx, problem):
    """Schur-complement solve of M x = b (Wilson) on TDIAG halves, FP32 inner
    and FP64 alone: the solution, set back to the chiral basis, is the
    chiral solve's."""
    p = problem((4, 4, 4, 8), 4)
    ge, go = gbx.Geometry.rb_pair(p.dims)
    U = [gbx.GaugeField(g, "f64").upload(p.dbl) for g in (ge, go)]
    U32 = [gbx.GaugeField(g, "f32").upload(p.dbl) for g in (ge, go)]
    even, odd = gbx.rb_split(p.dims, p.psi)
    b = [gbx.SpinorField(ge, p.Ls, 1, "f64").upload(even), gbx.SpinorField(go, p.Ls, 1, "f64").upload(odd)]
    x = [gbx.SpinorField(g, p.Ls, 1, "f64") for g in (ge, go)]
    for f in x:
        f.zero()
    gbx.cg_solve_eo(U, U32, x, b, 0.12, tol=1e-10)
    want = [f.download() for f in x]
    for f in b:
        f.set_basis(gbx.BASIS_TDIAG)
    for lo in (U32, None):
        for f in x:
            f.zero()
            f.set_basis(gbx.BASIS_TDIAG)
        st = gbx.cg_solve_eo(U, lo, x, b, 0.12, tol=1e-10)
        assert st["true_resid"] <= 1e-10
        for f, w in zip(x, want):
            assert rel_l2(chiral_download(gbx, f), w) <= 1e-8


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("dag", [False, True])
def test_batch_on_tdiag_fields(gpu, gbx, prec, dag):
    """The persistent batch launch on TDIAG inputs: each lattice's output is
    its own single-lattice TDIAG application bit for bit (same arithmetic)."""
    lattices = [((4, 4, 4, 4), 4, 1), ((6, 4, 2, 8), 3, 2), ((2, 4, 4, 2), 5, 3)]
    objs = []
    for i, (dims, Ls, nrhs) in enumerate(lattices):
        g = gbx.Geometry(dims)
        U = gbx.GaugeField(g, prec).upload(gbx.gen_doubled_gauge(dims, 7000 + 2 * i))
        x = gbx.SpinorField(g, Ls, nrhs, prec).upload(gbx.gen_spinor(dims, Ls, nrhs, 7001 + 2 * i))
        x.set_basis(gbx.BASIS_TDIAG)
        objs.append((g, U, x, gbx.SpinorField(g, Ls, nrhs, prec)))
    b = gbx.Batch([o[1] for o in objs], [o[3] for o in objs], [o[2] for o in objs], dag=dag)
    b.run()
    for g, U, x, y in objs:
        assert y.basis == gbx.BASIS_TDIAG
        ref = gbx.SpinorField(g, x.Ls, x.nrhs, prec)
        gbx.dhop(U, ref, x, dag=dag, cfg=gbx.kcfg(1, 64, 0, 0))
        assert np.array_equal(chiral_download(gbx, ref), chiral_download(gbx, y))
