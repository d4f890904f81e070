"""Even-odd (red-black) hopping term and the Schur-complement solver.

DhopEO / DhopOE are pinned by the whole-lattice oracle: on an input that is
zero on one parity, the oracle's Dhop restricted to the other parity is the
hop between the halves (every neighbour of a site has the other parity).  The
GPU halves also reproduce the GPU whole-lattice kernel bit for bit, since each
output site sums the same 8 legs in the same order.  The solver's answer is
checked with the oracle's M = 1 - kappa Dhop on the merged lattice.
"""
import numpy as np
import pytest

from conftest import rel_l2, storage_tol

pytestmark = pytest.mark.gpu

KAPPA = 0.12


def halves(gbx, p, prec="f64", comp12=False, block=None):
    ge, go = gbx.Geometry.rb_pair(p.dims, block)
    Ue = gbx.GaugeField(ge, prec, comp12).upload(p.dbl)
    Uo = gbx.GaugeField(go, prec, comp12).upload(p.dbl)
    return ge, go, Ue, Uo


def masked(gbx, p, keep_parity):
    """p.psi with the sites of the other parity set to zero."""
    e, o = gbx.rb_split(p.dims, p.psi)
    zero = np.zeros_like(e)
    return gbx.rb_merge(p.dims, e if keep_parity == 0 else zero, o if keep_parity == 1 else zero)


@pytest.mark.parametrize("dims,Ls", [((4, 4, 4, 4), 4), ((8, 4, 6, 2), 3), ((2, 2, 2, 2), 2)])
@pytest.mark.parametrize("dag", [False, True])
def test_dhop_eo_oe_vs_oracle(gpu, gbx, problem, dims, Ls, dag):
    p = problem(dims, Ls)
    ge, go, Ue, Uo = halves(gbx, p)
    pe, po = gbx.rb_split(p.dims, p.psi)
    xe = gbx.SpinorField(ge, Ls, 1, "f64").upload(pe)
    xo = gbx.SpinorField(go, Ls, 1, "f64").upload(po)
    ye = gbx.SpinorField(ge, Ls, 1, "f64")
    yo = gbx.SpinorField(go, Ls, 1, "f64")
    gbx.dhop(Ue, ye, xo, dag=dag)  # DhopEO
    gbx.dhop(Uo, yo, xe, dag=dag)  # DhopOE
    # oracle: Dhop of the odd-only field, even part; Dhop of the even-only field, odd part
    want_e = gbx.rb_split(p.dims, p.oracle_lex(dag=dag, psi=masked(gbx, p, 1)[0])[None])[0][0]
    want_o = gbx.rb_split(p.dims, p.oracle_lex(dag=dag, psi=masked(gbx, p, 0)[0])[None])[1][0]
    ee, eo = rel_l2(ye.download()[0], want_e), rel_l2(yo.download()[0], want_o)
    print(f"dims={dims} Ls={Ls} dag={dag}: DhopEO {ee:.2e} DhopOE {eo:.2e}")
    assert ee <= 1e-14 and eo <= 1e-14


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_halves_reproduce_whole_lattice_kernel(gpu, gbx, problem, prec):
    p = problem((8, 4, 4, 6), 4, nrhs=2)
    g = gbx.Geometry(p.dims)
    U = gbx.GaugeField(g, prec).upload(p.dbl)
    x = gbx.SpinorField(g, 4, 2, prec).upload(p.psi)
    y = gbx.SpinorField(g, 4, 2, prec)
    gbx.dhop(U, y, x)
    full = y.download()
    ge, go, Ue, Uo = halves(gbx, p, prec)
    pe, po = gbx.rb_split(p.dims, p.psi)
    xe = gbx.SpinorField(ge, 4, 2, prec).upload(pe)
    xo = gbx.SpinorField(go, 4, 2, prec).upload(po)
    ye = gbx.SpinorField(ge, 4, 2, prec)
    yo = gbx.SpinorField(go, 4, 2, prec)
    gbx.dhop(Ue, ye, xo)
    gbx.dhop(Uo, yo, xe)
    assert np.array_equal(gbx.rb_merge(p.dims, ye.download(), yo.download()), full)


@pytest.mark.parametrize("prec", ["f16", "bf16"])
def test_mfma_and_compressed_links_on_halves(gpu, gbx, problem, prec):
    p = problem((4, 4, 4, 8), 8, nrhs=2)
    tol = storage_tol(prec, p.psi)
    for comp12 in (False, True):
        ge, go, Ue, Uo = halves(gbx, p, prec, comp12)
        pe, po = gbx.rb_split(p.dims, p.psi)
        xo = gbx.SpinorField(go, 8, 2, prec).upload(po)
        ye = gbx.SpinorField(ge, 8, 2, prec)
        gbx.dhop(Ue, ye, xo, cfg=gbx.kcfg(2, 128, 1, 1, kind=1, pf=1))
        mf = ye.download()
        gbx.dhop(Ue, ye, xo, cfg=gbx.kcfg(4, 64, 1, 0))
        assert rel_l2(mf, ye.download()) <= tol
        for r in range(2):
            want = gbx.rb_split(p.dims, p.oracle_lex(rhs=r, psi=masked(gbx, p, 1)[r])[None])[0][0]
            assert rel_l2(mf[r], want) <= tol


def oracle_M(p, x_lex):
    """M x = x - kappa Dhop x on the whole lattice, via the oracle."""
    return x_lex - KAPPA * p.oracle_lex(psi=x_lex)


@pytest.mark.parametrize("lo", [None, "f32", "f16", "bf16"])
def test_eo_solver_reaches_1e10(gpu, gbx, problem, lo):
    p = problem((8, 8, 8, 8), 4, gseed=321, sseed=322)
    ge, go, Ue, Uo = halves(gbx, p)
    low = None if lo is None else (gbx.GaugeField(ge, lo).upload(p.dbl), gbx.GaugeField(go, lo).upload(p.dbl))
    be, bo = gbx.rb_split(p.dims, p.psi)
    b = (gbx.SpinorField(ge, 4, 1, "f64").upload(be), gbx.SpinorField(go, 4, 1, "f64").upload(bo))
    x = (gbx.SpinorField(ge, 4, 1, "f64").zero(), gbx.SpinorField(go, 4, 1, "f64").zero())
    st = gbx.cg_solve_eo((Ue, Uo), low, x, b, KAPPA, tol=1e-10, inner_tol=1e-3, max_outer=60)
    xl = gbx.rb_merge(p.dims, x[0].download(), x[1].download())[0]
    res = rel_l2(oracle_M(p, xl), p.psi[0])
    print(f"inner={lo} stats={st} oracle |b - M x| / |b| = {res:.2e}")
    assert st["true_resid"] <= 1e-10
    assert res <= 2e-10


def test_eo_solver_matches_normal_equation_solver(gpu, gbx, problem):
    """M x = b by the Schur complement and by CG on M^dag M x = M^dag b over
    the whole lattice give the same x; the Schur form needs fewer hopping
    applications (each Schur application is two half-lattice hops)."""
    p = problem((8, 8, 8, 8), 4, gseed=331, sseed=332)
    g = gbx.Geometry(p.dims)
    U = gbx.GaugeField(g, "f64").upload(p.dbl)
    U32 = gbx.GaugeField(g, "f32").upload(p.dbl)
    b = gbx.SpinorField(g, 4, 1, "f64").upload(p.psi)
    mdb = gbx.SpinorField(g, 4, 1, "f64")
    gbx.dhop_axpby(U, mdb, b, 1.0, -KAPPA, dag=True)  # M^dag b
    xf = gbx.SpinorField(g, 4, 1, "f64").zero()
    stf = gbx.cg_solve(U, U32, xf, mdb, KAPPA, tol=1e-11)
    ge, go, Ue, Uo = halves(gbx, p)
    le, lo_ = gbx.GaugeField(ge, "f32").upload(p.dbl), gbx.GaugeField(go, "f32").upload(p.dbl)
    be, bo = gbx.rb_split(p.dims, p.psi)
    bb = (gbx.SpinorField(ge, 4, 1, "f64").upload(be), gbx.SpinorField(go, 4, 1, "f64").upload(bo))
    x = (gbx.SpinorField(ge, 4, 1, "f64").zero(), gbx.SpinorField(go, 4, 1, "f64").zero())
    ste = gbx.cg_solve_eo((Ue, Uo), (le, lo_), x, bb, KAPPA, tol=1e-11)
    xe = gbx.rb_merge(p.dims, x[0].download(), x[1].download())
    d = rel_l2(xe, xf.download())
    full_hops = 4 * (stf["inner_iters"] + stf["outer_iters"] + 1)
    eo_hops = 2 * (2 * ste["inner_iters"] + ste["hi_applies"] + ste["outer_iters"]) + 5
    print(f"normal eq: {stf}  ({full_hops} whole-lattice hops)\neven-odd: {ste}  ({eo_hops} half-lattice hops)")
    print(f"rel difference of the solutions {d:.2e}")
    assert d <= 1e-8
    assert ste["inner_iters"] < stf["inner_iters"]


def test_red_black_errors(gpu, gbx, problem):
    with pytest.raises(gbx.GbxError) as e:
        gbx.Geometry.rb_pair((4, 4, 3, 4))
    assert e.value.status == 1
    p = problem((4, 4, 4, 4), 2)
    ge, go, Ue, Uo = halves(gbx, p)
    xe = gbx.SpinorField(ge, 2, 1, "f64").zero()
    ye = gbx.SpinorField(ge, 2, 1, "f64")
    with pytest.raises(gbx.GbxError) as e:  # input of the same parity
        gbx.dhop(Ue, ye, xe)
    assert e.value.status == 9
    ge2, go2, Ue2, Uo2 = halves(gbx, p)
    xo2 = gbx.SpinorField(go2, 2, 1, "f64").zero()
    with pytest.raises(gbx.GbxError) as e:  # odd half of another pair
        gbx.dhop(Ue, ye, xo2)
    assert e.value.status == 9
    xo = gbx.SpinorField(go, 2, 1, "f64").zero()
    with pytest.raises(gbx.GbxError) as e:
        gbx.dhop_axpby(Ue, ye, xo, 1.0, -0.1)
    assert e.value.status == 1
    gbx.dhop_xpay(Ue, ye, xe, xo, 1.0, -0.1)  # the form that is defined on halves
    with pytest.raises(gbx.GbxError) as e:
        gbx.cg_solve(Ue, None, xe, ye, KAPPA)
    assert e.value.status == 1
    with pytest.raises(gbx.GbxError) as e:
        xe.upload_ref(np.zeros(64 * 2 * 24 // 2), 1)
    assert e.value.status == 8


def test_batch_of_half_lattice_hops(gpu, gbx, problem):
    """The persistent batch kernel takes red-black halves too (DhopEO and
    DhopOE of two lattices in one launch) and matches the single launches."""
    objs = []
    for i, (dims, Ls) in enumerate((((4, 4, 4, 4), 4), ((8, 4, 2, 4), 3))):
        p = problem(dims, Ls, gseed=700 + i, sseed=710 + i)
        ge, go, Ue, Uo = halves(gbx, p, "f32")
        pe, po = gbx.rb_split(p.dims, p.psi)
        xe = gbx.SpinorField(ge, Ls, 1, "f32").upload(pe)
        xo = gbx.SpinorField(go, Ls, 1, "f32").upload(po)
        objs += [(Ue, gbx.SpinorField(ge, Ls, 1, "f32"), xo), (Uo, gbx.SpinorField(go, Ls, 1, "f32"), xe)]
    b = gbx.Batch([o[0] for o in objs], [o[1] for o in objs], [o[2] for o in objs])
    b.run()
    got = [o[1].download() for o in objs]
    for (U, y, x), g in zip(objs, got):
        ref = gbx.SpinorField(y.geom, y.Ls, 1, "f32")
        gbx.dhop(U, ref, x, cfg=gbx.kcfg(1, 64, 0, 0))
        assert np.array_equal(ref.download(), g)
