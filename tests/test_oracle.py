"""Pins the oracle (oracle/hand.c: the CPU restatement of the reference Hand
kernel, WilsonKernelsHandCpu.h:385-451) and the layout/table conventions.

Each check is chosen so that a plausible mistake (a swapped projector sign,
a wrong leg direction, a transposed link, a mis-placed lane) fails it:
  * bit-exact against the reference kernel compiled from its own sources
    (oracle/_ref, when built in this checkout) for Nsimd 1..16, FP64 and FP32;
  * bit-exact against the committed fixtures made from that build
    (tests/golden/hand_fixtures.json), so the pin holds where _ref is absent;
  * an independent gamma-matrix statement of the operator
    (sum_mu (1 - g_mu) U psi(x+mu) + (1 + g_mu) U^dag psi(x-mu)) in numpy;
  * closed forms on the free field (U = 1): plane waves and constants;
  * gamma5-hermiticity of the dagger kernel (bit-exact), adjointness,
    gauge covariance.
"""
import json
import os

import numpy as np
import pytest

from conftest import rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "hand_fixtures.json")))

# gamma matrices of the chiral basis read off the reference projectors
# (WilsonKernelsHandCpu.h:94-159) -- written independently of hand.c
GX = np.array([[0, 0, 0, 1j], [0, 0, 1j, 0], [0, -1j, 0, 0], [-1j, 0, 0, 0]])
GY = np.array([[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]], dtype=complex)
GZ = np.array([[0, 0, 1j, 0], [0, 0, 0, -1j], [-1j, 0, 0, 0], [0, 1j, 0, 0]])
GT = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], dtype=complex)
GAMMA = [GX, GY, GZ, GT]


def make_case(gbx, orc, dims, Ls, nsimd, dtype, gseed=101, sseed=102):
    dbl = gbx.double_links(dims, gbx.gen_su3_links(dims, gseed))
    psi = gbx.gen_spinor(dims, Ls, 1, sseed)[0]
    U = orc.aligned_copy(orc.pack_gauge(dims, nsimd, dbl).astype(dtype))
    inp = orc.aligned_copy(orc.pack_spinor(dims, nsimd, Ls, psi).astype(dtype))
    nbr, prm = orc.tables(dims, nsimd, Ls)
    return dbl, psi, U, inp, nbr, prm, int(np.prod(dims)) // nsimd


def lex_dhop_gamma(dims, dbl, psi, dag=False):
    """Independent numpy statement: D psi(x) = sum_mu (1 -/+ g_mu) U_mu(x) psi(x+mu)
    + (1 +/- g_mu) U_mu(x-mu)^dag psi(x-mu), lower signs for the dagger."""
    L = tuple(dims[::-1])  # numpy axes (t, z, y, x)
    Ls = psi.shape[1]
    p = (psi[..., 0] + 1j * psi[..., 1]).reshape(L + (Ls, 4, 3))
    Uf = (dbl[..., 0] + 1j * dbl[..., 1]).reshape(L + (8, 3, 3))
    out = np.zeros_like(p)
    for mu in range(4):
        ax = 3 - mu  # numpy axis of dimension mu
        g = GAMMA[mu]
        sgn = -1 if not dag else 1
        fwd = np.roll(p, -1, axis=ax)  # psi(x + mu)
        bwd = np.roll(p, +1, axis=ax)  # psi(x - mu)
        Ufw, Ubw = Uf[..., mu, :, :], Uf[..., mu + 4, :, :]
        t = np.einsum("...ij,...sbj->...sbi", Ufw, fwd)
        out += t + sgn * np.einsum("ab,...sbi->...sai", g, t)
        t = np.einsum("...ij,...sbj->...sbi", Ubw, bwd)
        out += t - sgn * np.einsum("ab,...sbi->...sai", g, t)
    out = out.reshape(-1, Ls, 4, 3)
    return np.stack([out.real, out.imag], axis=-1)


def oracle_lex(gbx, orc, dims, dbl, psi, dag=False):
    Ls = psi.shape[1]
    nbr, prm = orc.tables(dims, 1, Ls)
    out = orc.hand(orc.pack_gauge(dims, 1, dbl), orc.pack_spinor(dims, 1, Ls, psi), nbr, prm, int(np.prod(dims)), Ls, 1,
                   dag=dag, nthreads=4)
    return orc.unpack_spinor(dims, 1, Ls, out).reshape(-1, Ls, 4, 3, 2)


# ------------------------------------------------- the reference build
ref_built = pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_w16.so")),
                               reason="reference kernel not built here (make -C oracle ref needs the reference sources)")


@ref_built
@pytest.mark.parametrize("nsimd,dtype", [(1, np.float64), (2, np.float64), (4, np.float64), (8, np.float64),
                                         (2, np.float32), (4, np.float32), (8, np.float32), (16, np.float32)])
@pytest.mark.parametrize("dims,Ls", [((4, 4, 4, 8), 3), ((2, 4, 6, 8), 2), ((8, 8, 8, 8), 8)])
def test_oracle_bitexact_vs_reference_build(gbx, orc, nsimd, dtype, dims, Ls):
    if any(d % s for d, s in zip(dims, orc.simd_layout(nsimd))):
        pytest.skip("lattice not divisible by the SIMD split")
    _, _, U, inp, nbr, prm, nsite = make_case(gbx, orc, dims, Ls, nsimd, dtype)
    got = orc.hand(U, inp, nbr, prm, nsite, Ls, nsimd, nthreads=2)
    want = orc.ref_hand(U, inp, nbr, prm, nsite, Ls, nsimd)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


@ref_built
def test_oracle_bitexact_vs_reference_scalar_kernel(gbx, orc):
    """dslash_simple.h:19-150, an independent scalar statement in the reference."""
    _, _, U, inp, nbr, prm, nsite = make_case(gbx, orc, (4, 6, 4, 2), 4, 1, np.float64)
    got = orc.hand(U, inp, nbr, prm, nsite, 4, 1)
    assert np.array_equal(got.view(np.uint8), orc.ref_simple(U, inp, nbr, prm, nsite, 4).view(np.uint8))


@ref_built
def test_dagger_is_gamma5_sandwich_of_reference(gbx, orc):
    _, _, U, inp, nbr, prm, nsite = make_case(gbx, orc, (4, 4, 4, 4), 2, 4, np.float64)
    got = orc.hand(U, inp, nbr, prm, nsite, 2, 4, dag=True)
    want = orc.gamma5(orc.ref_hand(U, orc.gamma5(inp, 4), nbr, prm, nsite, 2, 4), 4)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


# ------------------------------------------------- committed fixtures
@pytest.mark.parametrize("case", FIX["cases"], ids=lambda c: f"{'x'.join(map(str, c['dims']))}_Ls{c['Ls']}_n{c['nsimd']}_{c['prec']}")
def test_oracle_matches_golden_fixtures(gbx, orc, case):
    dims, Ls, nsimd = tuple(case["dims"]), case["Ls"], case["nsimd"]
    dtype = np.float64 if case["prec"] == "f64" else np.float32
    _, _, U, inp, nbr, prm, nsite = make_case(gbx, orc, dims, Ls, nsimd, dtype, case["gauge_seed"], case["spinor_seed"])
    # inputs and tables are the ones the fixture was made from
    assert orc.sha256(nbr) == case["sha256_nbr"] and orc.sha256(prm) == case["sha256_prm"]
    assert orc.sha256(U) == case["sha256_U"] and orc.sha256(inp) == case["sha256_in"]
    out = orc.hand(U, inp, nbr, prm, nsite, Ls, nsimd)
    assert orc.sha256(out) == case["sha256_out"]
    assert [float(x).hex() for x in out[:64].astype(np.float64)] == case["first64_out"]
    assert float(np.sum(out.astype(np.float64) ** 2)).hex() == case["norm2_out"]
    dag = orc.hand(U, inp, nbr, prm, nsite, Ls, nsimd, dag=True)
    assert orc.sha256(dag) == case["sha256_out_dag"]
    if "out" in case:
        assert [float(x).hex() for x in out.astype(np.float64)] == case["out"]
        assert [float(x).hex() for x in dag.astype(np.float64)] == case["out_dag"]
    if "sha256_out_simple" in case:
        assert case["sha256_out_simple"] == case["sha256_out"]


def test_generator_matches_golden(gbx, orc):
    g = FIX["generator"]
    assert [float(x).hex() for x in gbx.gen_gaussian(7, 3, 0, 16)] == g["gaussian_seed7_stream3_first16"]
    assert [float(x).hex() for x in gbx.gen_gaussian(7, 3, 1000000, 8)] == g["gaussian_seed7_stream3_from1000000_8"]
    links = gbx.gen_su3_links((2, 2, 2, 2), 101)
    assert orc.sha256(links) == g["su3_links_2222_seed101_sha256"]
    assert orc.sha256(gbx.gen_spinor((4, 4, 4, 4), 4, 1, 102)) == g["spinor_4444_ls4_seed102_sha256"]


# ------------------------------------------------- independent statements
@pytest.mark.parametrize("dag", [False, True])
@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (2, 4, 6, 4), (3, 5, 2, 4)])
def test_oracle_vs_gamma_matrix_form(gbx, orc, dims, dag):
    dbl = gbx.double_links(dims, gbx.gen_su3_links(dims, 11))
    psi = gbx.gen_spinor(dims, 2, 1, 12)[0]
    got = oracle_lex(gbx, orc, dims, dbl, psi, dag)
    want = lex_dhop_gamma(dims, dbl, psi, dag)
    assert rel_l2(got, want) < 1e-14


def test_gamma_matrices_are_a_clifford_basis():
    for a in range(4):
        assert np.allclose(GAMMA[a], GAMMA[a].conj().T)
        for b in range(4):
            ac = GAMMA[a] @ GAMMA[b] + GAMMA[b] @ GAMMA[a]
            assert np.allclose(ac, 2 * np.eye(4) * (a == b))
    g5 = GX @ GY @ GZ @ GT
    assert np.allclose(g5, np.diag([1, 1, -1, -1]))


def unit_gauge(dims):
    V = int(np.prod(dims))
    d = np.zeros((V, 8, 3, 3, 2))
    for i in range(3):
        d[:, :, i, i, 0] = 1.0
    return d


@pytest.mark.parametrize("dag", [False, True])
def test_free_field_plane_wave(gbx, orc, dag):
    """U = 1, psi(x) = exp(i p.x) chi: D psi = sum_mu [2 cos p_mu -/+ 2 i g_mu sin p_mu] psi."""
    dims = (4, 6, 8, 4)
    n = (1, 2, 3, 1)
    p = [2 * np.pi * n[d] / dims[d] for d in range(4)]
    rng = np.random.default_rng(3)
    chi = rng.normal(size=(4, 3)) + 1j * rng.normal(size=(4, 3))
    coords = np.stack(np.meshgrid(*[np.arange(L) for L in dims], indexing="ij"), -1)  # x,y,z,t
    phase = np.exp(1j * (coords @ np.array(p))).transpose(3, 2, 1, 0).reshape(-1)  # lexicographic x fastest
    psi_c = phase[:, None, None, None] * chi[None, None]
    psi = np.stack([psi_c.real, psi_c.imag], -1)
    got = oracle_lex(gbx, orc, dims, unit_gauge(dims), psi, dag)
    sgn = -1 if not dag else 1
    K = sum(2 * np.cos(p[m]) * np.eye(4) + sgn * 2j * np.sin(p[m]) * GAMMA[m] for m in range(4))
    want_c = np.einsum("ab,xsbc->xsac", K, psi_c)
    assert rel_l2(got, np.stack([want_c.real, want_c.imag], -1)) < 1e-13


def test_constant_field_gives_8psi(gbx, orc):
    dims = (4, 4, 2, 6)
    rng = np.random.default_rng(5)
    one = rng.normal(size=(1, 1, 4, 3, 2))
    psi = np.broadcast_to(one, (int(np.prod(dims)), 3, 4, 3, 2)).copy()
    got = oracle_lex(gbx, orc, dims, unit_gauge(dims), psi)
    # (1 - g) + (1 + g) = 2 per direction; exact up to the rounding of the
    # projector sums (a + b) + (a - b)
    assert rel_l2(got, 8.0 * psi) < 1e-15


def test_adjointness(gbx, orc):
    """<chi, D psi> = <D^dag chi, psi> with a consistent doubled field."""
    dims = (4, 4, 6, 4)
    dbl = gbx.double_links(dims, gbx.gen_su3_links(dims, 21))
    psi = gbx.gen_spinor(dims, 3, 2, 22)
    a, b = psi[0], psi[1]
    Da = oracle_lex(gbx, orc, dims, dbl, a)
    Ddb = oracle_lex(gbx, orc, dims, dbl, b, dag=True)
    c = lambda x: x[..., 0] + 1j * x[..., 1]
    lhs = np.vdot(c(b), c(Da))
    rhs = np.vdot(c(Ddb), c(a))
    assert abs(lhs - rhs) / abs(lhs) < 1e-13


def test_adjointness_fails_for_inconsistent_backward_links(gbx, orc):
    """The dagger relation needs U[x][mu+4] = U_mu(x-mu)^dag: with independent
    backward links it must fail -- so the test above really checks the field."""
    dims = (4, 4, 4, 4)
    dbl = gbx.double_links(dims, gbx.gen_su3_links(dims, 21))
    other = gbx.double_links(dims, gbx.gen_su3_links(dims, 99))
    dbl[:, 4:] = other[:, 4:]
    psi = gbx.gen_spinor(dims, 2, 2, 22)
    c = lambda x: x[..., 0] + 1j * x[..., 1]
    lhs = np.vdot(c(psi[1]), c(oracle_lex(gbx, orc, dims, dbl, psi[0])))
    rhs = np.vdot(c(oracle_lex(gbx, orc, dims, dbl, psi[1], dag=True)), c(psi[0]))
    assert abs(lhs - rhs) / abs(lhs) > 1e-3


def test_gauge_covariance(gbx, orc):
    """D[U^Omega](Omega psi) = Omega D[U] psi with U^Omega_mu(x) = Omega(x) U_mu(x) Omega(x+mu)^dag."""
    dims = (4, 2, 4, 6)
    V = int(np.prod(dims))
    links = gbx.gen_su3_links(dims, 31)
    om = gbx.gen_su3_links(dims, 32)[:, 0]  # one SU(3) matrix per site
    c = lambda x: x[..., 0] + 1j * x[..., 1]
    Uc, Om = c(links), c(om)
    strides = [1, dims[0], dims[0] * dims[1], dims[0] * dims[1] * dims[2]]
    coords = np.stack(np.unravel_index(np.arange(V), dims[::-1])[::-1], -1)
    Ut = np.empty_like(Uc)
    for mu in range(4):
        cc = coords.copy()
        cc[:, mu] = (cc[:, mu] + 1) % dims[mu]
        xp = cc @ np.array(strides)
        Ut[:, mu] = Om @ Uc[:, mu] @ Om[xp].conj().transpose(0, 2, 1)
    re = lambda z: np.stack([z.real, z.imag], -1)
    dbl = gbx.double_links(dims, re(Uc))
    dblt = gbx.double_links(dims, re(Ut))
    psi = c(gbx.gen_spinor(dims, 2, 1, 33)[0])
    lhs = c(oracle_lex(gbx, orc, dims, dblt, re(np.einsum("xij,xsaj->xsai", Om, psi))))
    rhs = np.einsum("xij,xsaj->xsai", Om, c(oracle_lex(gbx, orc, dims, dbl, re(psi))))
    assert rel_l2(re(lhs), re(rhs)) < 1e-13


def test_dense_matrix_gamma5_hermiticity(gbx, orc):
    """Brute force on 2x2x2x2 x Ls1: the matrix of the dagger kernel is the
    conjugate transpose of the matrix of the kernel, and g5 D is hermitian."""
    dims = (2, 2, 2, 2)
    V = 16
    dbl = gbx.double_links(dims, gbx.gen_su3_links(dims, 41))
    n = V * 12
    Dm = np.zeros((n, n), complex)
    Dd = np.zeros((n, n), complex)
    for j in range(2 * n):
        e = np.zeros(n * 2)
        e[j] = 1.0
        col = j // 2
        ph = 1.0 if j % 2 == 0 else 1j
        psi = e.reshape(V, 1, 4, 3, 2)
        if j % 2:
            continue
        Dm[:, col] = (lambda o: (o[..., 0] + 1j * o[..., 1]).reshape(-1))(oracle_lex(gbx, orc, dims, dbl, psi)) / ph
        Dd[:, col] = (lambda o: (o[..., 0] + 1j * o[..., 1]).reshape(-1))(oracle_lex(gbx, orc, dims, dbl, psi, True)) / ph
    assert np.abs(Dd - Dm.conj().T).max() < 1e-14
    g5 = np.kron(np.eye(V), np.kron(np.diag([1, 1, -1, -1]), np.eye(3)))
    H = g5 @ Dm
    assert np.abs(H - H.conj().T).max() < 1e-14


# ------------------------------------------------- tables and layouts
@pytest.mark.parametrize("nsimd", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("dims,Ls", [((4, 4, 4, 4), 2), ((8, 4, 6, 2), 3), ((2, 2, 2, 2), 1)])
def test_table_invariants(orc, nsimd, dims, Ls):
    simd = orc.simd_layout(nsimd)
    if any(d % s for d, s in zip(dims, simd)):
        pytest.skip("not divisible")
    nbr, prm = orc.tables(dims, nsimd, Ls)
    nsite = int(np.prod(dims)) // nsimd
    n5 = nsite * Ls
    t = nbr.reshape(nsite, Ls, 8)
    assert (nbr < n5).all() and set(np.unique(prm)) <= {0, 1}
    s = np.arange(Ls)[None, :]
    assert (t % Ls == s[..., None]).all()  # the s coordinate is carried
    for mu in range(8):
        col = t[:, :, mu].reshape(-1)
        assert len(np.unique(col)) == n5  # a bijection
        back = t.reshape(-1, 8)[col, (mu + 4) % 8]
        assert (back == np.arange(n5)).all()  # -mu undoes +mu
    rd = [d // s for d, s in zip(dims, simd)]
    o = np.stack(np.unravel_index(np.arange(nsite), rd[::-1])[::-1], -1)
    p = prm.reshape(nsite, Ls, 8)[:, 0, :]
    for mu in range(8):
        d = mu % 4
        wrap = (o[:, d] == rd[d] - 1) if mu < 4 else (o[:, d] == 0)
        assert (p[:, mu] == (wrap & (simd[d] > 1))).all()


@pytest.mark.parametrize("nsimd", [1, 2, 4, 8, 16])
def test_library_tables_equal_oracle_tables(gbx, orc, nsimd):
    for dims, Ls in (((4, 4, 4, 4), 2), ((6, 4, 2, 8), 3), ((2, 2, 2, 2), 4)):
        if any(d % s for d, s in zip(dims, orc.simd_layout(nsimd))):
            continue
        a = orc.tables(dims, nsimd, Ls)
        b = gbx.table_generate(dims, nsimd, Ls)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        nsite = int(np.prod(dims)) // nsimd
        assert gbx.table_recognise(nsimd, nsite, Ls, b[0], b[1]) == dims


def test_table_recognise_rejects_modified_tables(gbx):
    nbr, prm = gbx.table_generate((4, 4, 4, 4), 2, 2)
    nsite = 128
    n2 = nbr.copy()
    n2[40], n2[48] = n2[48], n2[40]
    assert gbx.table_recognise(2, nsite, 2, n2, prm) is None
    p2 = prm.copy()
    p2[np.argmax(prm == 0)] = 1
    assert gbx.table_recognise(2, nsite, 2, nbr, p2) is None


@pytest.mark.parametrize("nsimd", [1, 2, 4, 8, 16])
def test_pack_unpack_roundtrip(gbx, orc, nsimd):
    dims, Ls = (4, 4, 4, 4), 3
    psi = gbx.gen_spinor(dims, Ls, 1, 5)[0]
    ref = orc.pack_spinor(dims, nsimd, Ls, psi)
    assert np.array_equal(orc.unpack_spinor(dims, nsimd, Ls, ref), psi.reshape(-1))


def test_lane_layout_matches_grid_virtual_nodes(orc):
    """Lane l of outer site o holds global x = o + rd * lane_coord (x fastest)."""
    dims, nsimd = (4, 4, 4, 4), 8
    V = 256
    lex = np.zeros((V, 1, 12, 2))
    lex[:, 0, 0, 0] = np.arange(V)  # tag each site with its lexicographic index
    ref = orc.pack_spinor(dims, nsimd, 1, lex).reshape(V // nsimd, 12, nsimd, 2)
    simd = orc.simd_layout(nsimd)
    rd = [d // s for d, s in zip(dims, simd)]
    for o in range(V // nsimd):
        oc = np.unravel_index(o, rd[::-1])[::-1]
        for l in range(nsimd):
            ic = np.unravel_index(l, simd[::-1])[::-1]
            x = [oc[d] + rd[d] * ic[d] for d in range(4)]
            g = x[0] + 4 * (x[1] + 4 * (x[2] + 4 * x[3]))
            assert ref[o, 0, l, 0] == g


@pytest.mark.parametrize("width", [1, 2, 8])
def test_rrii_layout_is_planar_re_im(orc, width):
    """RRII vector = its lanes' real parts, then their imaginary parts
    (bench.cc:105-106), checked through numpy's complex view of RIRI."""
    rng = np.random.default_rng(width)
    riri = rng.standard_normal(5 * width * 2)
    z = riri.view(np.complex128).reshape(5, width)
    rrii = orc.to_rrii(riri, width).reshape(5, 2, width)
    assert np.array_equal(rrii[:, 0], z.real) and np.array_equal(rrii[:, 1], z.imag)
    assert np.array_equal(orc.from_rrii(orc.to_rrii(riri, width), width), riri)


def test_expanded_lanes_hold_replicas(orc):
    """Expanded vectors: replica nn at lanes nn*DATA_SIMD + n (bench.cc:110-122);
    with identical replicas every group of DATA_SIMD lanes repeats the data."""
    rng = np.random.default_rng(1)
    reps = [rng.standard_normal(6 * 2 * 2) for _ in range(4)]
    ex = orc.expand_lanes(reps, 2)
    z = ex.view(np.complex128).reshape(6, 8)
    for nn in range(4):
        assert np.array_equal(z[:, 2 * nn:2 * nn + 2], reps[nn].view(np.complex128).reshape(6, 2))
        assert np.array_equal(orc.replica_lanes(ex, 8, 2, nn), reps[nn])
    same = orc.expand_lanes([reps[0]] * 4, 2).view(np.complex128).reshape(6, 8)
    assert all(np.array_equal(same[:, 0:2], same[:, 2 * k:2 * k + 2]) for k in range(4))


# ------------------------------------------------- generator
def test_su3_links_are_special_unitary(gbx):
    U = gbx.gen_su3_links((4, 4, 2, 2), 7)
    M = (U[..., 0] + 1j * U[..., 1]).reshape(-1, 3, 3)
    eye = np.einsum("nij,nkj->nik", M, M.conj())
    assert np.abs(eye - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-14


def test_doubled_links_are_backward_daggers(gbx):
    dims = (3, 4, 2, 5)
    links = gbx.gen_su3_links(dims, 8)
    dbl = gbx.double_links(dims, links)
    c = lambda x: x[..., 0] + 1j * x[..., 1]
    L, D = c(links), c(dbl)
    V = int(np.prod(dims))
    coords = np.stack(np.unravel_index(np.arange(V), dims[::-1])[::-1], -1)
    strides = np.array([1, dims[0], dims[0] * dims[1], dims[0] * dims[1] * dims[2]])
    assert np.array_equal(D[:, :4], L)
    for mu in range(4):
        cc = coords.copy()
        cc[:, mu] = (cc[:, mu] - 1) % dims[mu]
        assert np.array_equal(D[:, mu + 4], L[cc @ strides, mu].conj().transpose(0, 2, 1))


def test_gaussian_statistics_and_random_access(gbx):
    from scipy import stats
    x = gbx.gen_gaussian(123, 0, 0, 400000)
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1) < 0.01
    assert stats.kstest(x, "norm").pvalue > 1e-3
    y = gbx.gen_gaussian(123, 0, 1000, 500)
    assert np.array_equal(y, x[1000:1500])
    z = gbx.gen_gaussian(123, 1, 0, 1000)
    assert not np.array_equal(z, x[:1000])
    assert abs(np.corrcoef(x[:100000:2], x[1:100000:2])[0, 1]) < 0.02
