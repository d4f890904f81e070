"""Domain-wall (Shamir) fifth-dimension terms and the DWF operator
D = Mooee - Dhop / 2 with the reference's DomainWallFermion parameters
(mass 0.1, M5 1.0, TableGenerate.cc:47-50).

Grid's own DWF code is not in the reference, so which chiral projector hops
up in s follows the Shamir form written in include/gbx.h (parity with Grid's
convention unpinned).  What is checked:
  * Mooee and its closed-form inverse against the Ls x Ls matrices of that
    form, built here and inverted with numpy (both adjoints too);
  * D against that Mooee plus the oracle's (pinned) Dhop;
  * the DWF identities: Mooee^-1 Mooee = 1, <a, D b> = <D^dag a, b>,
    D^dag = Gamma5 D Gamma5 with Gamma5 = gamma5 x (s -> Ls-1-s), and
    det Mooee = (d^Ls + mass) per chirality, d = 5 - M5;
  * the even-odd blocks reproduce the whole-lattice D bit for bit, and the
    Schur-complement solve of D x = b meets its tolerance in the oracle's D.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

M5, MASS = 1.0, 0.1


def s_matrices(Ls, M5, mass):
    """(A+, A-): Mooee on the P+ (spins 0,1) and P- (spins 2,3) components."""
    d = 5.0 - M5
    Ap = d * np.eye(Ls)
    Am = d * np.eye(Ls)
    for s in range(1, Ls):
        Ap[s, s - 1] = -1.0
    for s in range(Ls - 1):
        Am[s, s + 1] = -1.0
    Ap[0, Ls - 1] = mass
    Am[Ls - 1, 0] = mass
    return Ap, Am


def mooee_np(psi, M5, mass, inv=False, dag=False):
    """psi: [..., Ls, 4, 3, 2] lexicographic."""
    Ls = psi.shape[-4]
    Ap, Am = s_matrices(Ls, M5, mass)
    if dag:
        Ap, Am = Ap.T, Am.T
    if inv:
        Ap, Am = np.linalg.inv(Ap), np.linalg.inv(Am)
    out = np.empty_like(psi)
    out[..., 0:2, :, :] = np.einsum("st,...tabc->...sabc", Ap, psi[..., 0:2, :, :])
    out[..., 2:4, :, :] = np.einsum("st,...tabc->...sabc", Am, psi[..., 2:4, :, :])
    return out


def gamma5_R5(psi):
    out = psi[..., ::-1, :, :, :].copy()
    out[..., 2:4, :, :] *= -1.0
    return out


def test_mooee_determinant_closed_form():
    for Ls, M5_, mass in ((8, 1.0, 0.1), (12, 1.8, 0.01), (5, 0.5, 1.0), (16, 1.0, 0.0)):
        d = 5.0 - M5_
        for A in s_matrices(Ls, M5_, mass):
            assert np.isclose(np.linalg.det(A), d ** Ls + mass, rtol=1e-12)


@pytest.mark.parametrize("Ls,M5_,mass", [(8, 1.0, 0.1), (12, 1.8, 0.02), (5, 0.6, 1.0), (16, 1.0, 0.1)])
@pytest.mark.parametrize("inv", [False, True])
@pytest.mark.parametrize("dag", [False, True])
def test_mooee_matches_matrix_form(gpu, gbx, problem, Ls, M5_, mass, inv, dag):
    p = problem((4, 4, 2, 4), Ls, nrhs=2)
    g = gbx.Geometry(p.dims)
    x = gbx.SpinorField(g, Ls, 2, "f64").upload(p.psi)
    y = gbx.SpinorField(g, Ls, 2, "f64")
    gbx.dwf_mooee(y, x, M5_, mass, inv=inv, dag=dag)
    err = rel_l2(y.download(), mooee_np(p.psi, M5_, mass, inv, dag))
    print(f"Ls={Ls} M5={M5_} mass={mass} inv={inv} dag={dag}: {err:.2e}")
    assert err <= 1e-14
    x32 = gbx.SpinorField(g, Ls, 2, "f32").upload(p.psi)
    y32 = gbx.SpinorField(g, Ls, 2, "f32")
    gbx.dwf_mooee(y32, x32, M5_, mass, inv=inv, dag=dag)
    want32 = mooee_np(p.psi.astype(np.float32).astype(np.float64), M5_, mass, inv, dag)
    assert rel_l2(y32.download(), want32) <= 1e-6


@pytest.mark.parametrize("prec,tol", [("f64", 1e-14), ("f32", 1e-6), ("f16", 2e-3)])
def test_mooee_inverse_in_place(gpu, gbx, problem, prec, tol):
    p = problem((4, 2, 2, 4), 12, nrhs=1)
    g = gbx.Geometry(p.dims)
    x = gbx.SpinorField(g, 12, 1, prec).upload(p.psi)
    start = x.download()
    for dag in (False, True):
        gbx.dwf_mooee(x, x, inv=False, dag=dag)
        gbx.dwf_mooee(x, x, inv=True, dag=dag)
        assert rel_l2(x.download(), start) <= tol


@pytest.mark.parametrize("dag", [False, True])
def test_dwf_operator_vs_oracle(gpu, gbx, problem, dag):
    p = problem((4, 4, 4, 4), 8)
    g = gbx.Geometry(p.dims)
    U = gbx.GaugeField(g, "f64").upload(p.dbl)
    x = gbx.SpinorField(g, 8, 1, "f64").upload(p.psi)
    y = gbx.SpinorField(g, 8, 1, "f64")
    gbx.dwf_apply(U, y, x, dag=dag)
    want = mooee_np(p.psi[0], M5, MASS, dag=dag) - 0.5 * p.oracle_lex(dag=dag)
    err = rel_l2(y.download()[0], want)
    print(f"dag={dag}: D vs Mooee + oracle Dhop {err:.2e}")
    assert err <= 1e-14


def test_dwf_adjoint_and_gamma5_hermiticity(gpu, gbx, problem):
    p = problem((4, 4, 2, 4), 6, nrhs=2)
    g = gbx.Geometry(p.dims)
    U = gbx.GaugeField(g, "f64").upload(p.dbl)
    a = gbx.SpinorField(g, 6, 1, "f64").upload(p.psi[0:1])
    b = gbx.SpinorField(g, 6, 1, "f64").upload(p.psi[1:2])
    Db = gbx.SpinorField(g, 6, 1, "f64")
    Dda = gbx.SpinorField(g, 6, 1, "f64")
    gbx.dwf_apply(U, Db, b)
    gbx.dwf_apply(U, Dda, a, dag=True)
    lhs, rhs = a.dot(Db), Dda.dot(b)
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    # D^dag a = Gamma5 D Gamma5 a
    g5a = gbx.SpinorField(g, 6, 1, "f64").upload(gamma5_R5(p.psi[0:1]))
    t = gbx.SpinorField(g, 6, 1, "f64")
    gbx.dwf_apply(U, t, g5a)
    assert rel_l2(gamma5_R5(t.download()), Dda.download()) <= 1e-15


def test_even_odd_blocks_reproduce_whole_lattice_operator(gpu, gbx, problem):
    p = problem((4, 4, 4, 6), 8)
    g = gbx.Geometry(p.dims)
    U = gbx.GaugeField(g, "f64").upload(p.dbl)
    x = gbx.SpinorField(g, 8, 1, "f64").upload(p.psi)
    y = gbx.SpinorField(g, 8, 1, "f64")
    gbx.dwf_apply(U, y, x)
    ge, go = gbx.Geometry.rb_pair(p.dims)
    Ue, Uo = gbx.GaugeField(ge, "f64").upload(p.dbl), gbx.GaugeField(go, "f64").upload(p.dbl)
    pe, po = gbx.rb_split(p.dims, p.psi)
    xe, xo = gbx.SpinorField(ge, 8, 1, "f64").upload(pe), gbx.SpinorField(go, 8, 1, "f64").upload(po)
    ye, yo = gbx.SpinorField(ge, 8, 1, "f64"), gbx.SpinorField(go, 8, 1, "f64")
    # D_ee x_e + D_eo x_o = Mooee x_e - DhopEO x_o / 2 (the same two kernels the whole-lattice D runs)
    gbx.dhop(Ue, ye, xo)
    gbx.dhop(Uo, yo, xe)
    me, mo = gbx.SpinorField(ge, 8, 1, "f64"), gbx.SpinorField(go, 8, 1, "f64")
    gbx.dwf_mooee(me, xe)
    gbx.dwf_mooee(mo, xo)
    me.axpy(-0.5, ye)
    mo.axpy(-0.5, yo)
    got = gbx.rb_merge(p.dims, me.download(), mo.download())
    assert rel_l2(got, y.download()) <= 1e-15


@pytest.mark.parametrize("lo", [None, "f32", "f16"])
def test_dwf_even_odd_solve(gpu, gbx, problem, lo):
    p = problem((8, 8, 8, 8), 8, gseed=351, sseed=352)
    ge, go = gbx.Geometry.rb_pair(p.dims)
    Ue, Uo = gbx.GaugeField(ge, "f64").upload(p.dbl), gbx.GaugeField(go, "f64").upload(p.dbl)
    low = None if lo is None else (gbx.GaugeField(ge, lo).upload(p.dbl), gbx.GaugeField(go, lo).upload(p.dbl))
    be, bo = gbx.rb_split(p.dims, p.psi)
    b = (gbx.SpinorField(ge, 8, 1, "f64").upload(be), gbx.SpinorField(go, 8, 1, "f64").upload(bo))
    x = (gbx.SpinorField(ge, 8, 1, "f64").zero(), gbx.SpinorField(go, 8, 1, "f64").zero())
    st = gbx.dwf_solve_eo((Ue, Uo), low, x, b, tol=1e-10, max_outer=60)
    xl = gbx.rb_merge(p.dims, x[0].download(), x[1].download())[0]
    Dx = mooee_np(xl, M5, MASS) - 0.5 * p.oracle_lex(psi=xl)
    res = rel_l2(Dx, p.psi[0])
    print(f"inner={lo} stats={st} oracle |b - D x| / |b| = {res:.2e}")
    assert st["true_resid"] <= 1e-10
    assert res <= 2e-10


def test_dwf_argument_errors(gpu, gbx, problem):
    p = problem((4, 4, 4, 4), 4)
    g = gbx.Geometry(p.dims)
    U = gbx.GaugeField(g, "f64").upload(p.dbl)
    x = gbx.SpinorField(g, 4, 1, "f64").zero()
    y = gbx.SpinorField(g, 4, 1, "f64")
    with pytest.raises(gbx.GbxError) as e:
        gbx.dwf_mooee(y, x, M5=5.0)
    assert e.value.status == 1
    x1 = gbx.SpinorField(g, 1, 1, "f64").zero()
    with pytest.raises(gbx.GbxError) as e:
        gbx.dwf_mooee(gbx.SpinorField(g, 1, 1, "f64"), x1)
    assert e.value.status == 1
    with pytest.raises(gbx.GbxError) as e:
        gbx.dwf_mooee(gbx.SpinorField(g, 4, 1, "f32"), x)
    assert e.value.status == 9
    ge, go = gbx.Geometry.rb_pair(p.dims)
    Ue = gbx.GaugeField(ge, "f64").upload(p.dbl)
    xo = gbx.SpinorField(go, 4, 1, "f64").zero()
    with pytest.raises(gbx.GbxError) as e:
        gbx.dwf_apply(Ue, gbx.SpinorField(ge, 4, 1, "f64"), xo)
    assert e.value.status == 1
