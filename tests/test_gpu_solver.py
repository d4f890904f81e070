"""Defect-correction CG for M^dag M x = b, M = 1 - kappa Dhop (gbx_cg_solve).

The solver's own residual is not trusted: the final x is downloaded and the
residual b - M^dag M x recomputed with the oracle on the CPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KAPPA = 0.12


def oracle_mdagm(p, x):
    """M^dag M x with M = 1 - kappa D, via the oracle (lexicographic arrays)."""
    Dx = p.oracle_lex(psi=x)
    t = x - KAPPA * Dx
    return t - KAPPA * p.oracle_lex(psi=t, dag=True)


def solve(gbx, p, lo_prec, tol=1e-10, inner_tol=1e-3, comp12=False, max_outer=60):
    geom = gbx.Geometry(p.dims)
    Uhi = gbx.GaugeField(geom, "f64").upload(p.dbl)
    Ulo = None if lo_prec is None else gbx.GaugeField(geom, lo_prec, comp12).upload(p.dbl)
    b = gbx.SpinorField(geom, p.Ls, 1, "f64").upload(p.psi)
    x = gbx.SpinorField(geom, p.Ls, 1, "f64").zero()
    st = gbx.cg_solve(Uhi, Ulo, x, b, KAPPA, tol=tol, inner_tol=inner_tol, max_outer=max_outer)
    return x.download()[0], st


@pytest.mark.parametrize("lo", [None, "f32", "f16", "bf16"])
def test_cg_reaches_1e10(gpu, gbx, problem, lo):
    p = problem((8, 8, 8, 8), 8, gseed=301, sseed=302)
    x, st = solve(gbx, p, lo)
    r = p.psi[0] - oracle_mdagm(p, x)
    rel = np.linalg.norm(r) / np.linalg.norm(p.psi[0])
    print(f"inner={lo} stats={st} oracle residual={rel:.3e}")
    assert st["true_resid"] <= 1e-10
    assert rel <= 2e-10
    if lo is not None:
        assert st["outer_iters"] >= 2 and st["lo_applies"] > st["hi_applies"]


def test_cg_compressed_fp16_inner(gpu, gbx, problem):
    p = problem((4, 4, 8, 8), 6, gseed=311, sseed=312)
    x, st = solve(gbx, p, "f16", comp12=True)
    r = p.psi[0] - oracle_mdagm(p, x)
    assert np.linalg.norm(r) / np.linalg.norm(p.psi[0]) <= 2e-10


def test_cg_reports_nonconvergence(gpu, gbx, problem):
    p = problem((4, 4, 4, 4), 4)
    with pytest.raises(gbx.GbxError) as e:
        solve(gbx, p, "f32", tol=1e-14, max_outer=1)
    assert e.value.status == 10  # GBX_ENOCONV


def test_cg_argument_errors(gpu, gbx, problem):
    p = problem((4, 4, 4, 4), 2)
    geom = gbx.Geometry(p.dims)
    U = gbx.GaugeField(geom, "f64").upload(p.dbl)
    U32 = gbx.GaugeField(geom, "f32").upload(p.dbl)
    b = gbx.SpinorField(geom, 2, 1, "f64").upload(p.psi)
    x = gbx.SpinorField(geom, 2, 1, "f64")
    with pytest.raises(gbx.GbxError) as e:
        gbx.cg_solve(U, U32, x, b, 0.2)  # |kappa| >= 1/8
    assert e.value.status == 1
    with pytest.raises(gbx.GbxError) as e:
        gbx.cg_solve(U32, None, x, b, 0.1)  # outer field must be FP64
    assert e.value.status == 1
    with pytest.raises(gbx.GbxError) as e:
        gbx.cg_solve(U, U32, b, b, 0.1)  # x aliases b
    assert e.value.status == 3


def test_blas_reductions(gpu, gbx, problem):
    p = problem((4, 6, 4, 4), 3, nrhs=2)
    geom = gbx.Geometry(p.dims)
    for prec, tol in (("f64", 1e-14), ("f32", 1e-6), ("f16", 2e-3), ("bf16", 2e-2)):
        x = gbx.SpinorField(geom, 3, 2, prec).upload(p.psi)
        y = gbx.SpinorField(geom, 3, 2, prec).upload(p.psi[::-1].copy())
        xs, ys = x.download(), y.download()  # values as stored
        c = lambda a: (a[..., 0] + 1j * a[..., 1]).reshape(-1)
        assert abs(x.norm2() - np.sum(xs ** 2)) <= 1e-12 * np.sum(xs ** 2)
        d = x.dot(y)
        want = np.vdot(c(xs), c(ys))
        assert abs(d - want) <= 1e-12 * np.linalg.norm(xs) * np.linalg.norm(ys)
        y.axpy(0.5, x)
        assert np.linalg.norm(y.download() - (ys + 0.5 * xs)) <= tol * np.linalg.norm(ys)
        x.scale(-2.0)
        assert np.array_equal(x.download(), -2.0 * xs)
        x.zero()
        assert not x.download().any()
        # deterministic reductions: the same bits every time
        z = gbx.SpinorField(geom, 3, 2, prec).upload(p.psi)
        assert len({z.norm2() for _ in range(5)}) == 1
