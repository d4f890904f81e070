"""Domain parts on one GPU: the lattice cut into slabs (gbx_geom_create_part),
each slab's faces exchanged into its neighbours' ghost slots
(gbx_halo_pack / gbx_halo_unpack, or their spin-projected forms), then Dhop
on every slab.  The slabs' outputs put together must be the whole lattice's
Dhop to the bit (same kernel, same configuration: a site's arithmetic does
not depend on where its neighbours live), and the oracle's within the
storage precision's tolerance.

The reference runs on one rank (TableGenerate.cc:84); the whole-lattice
application is the reference operator (WilsonKernelsHandCpu.h:426-446).
"""
import numpy as np
import pytest

from conftest import rel_l2, storage_tol

pytestmark = pytest.mark.gpu

GRIDS = [((4, 4, 4, 8), (1, 1, 1, 2)), ((4, 4, 4, 8), (1, 1, 1, 4)), ((4, 4, 4, 8), (1, 1, 1, 8)),
         ((8, 4, 4, 4), (2, 1, 1, 1)), ((4, 6, 4, 4), (1, 3, 1, 1)), ((4, 4, 4, 4), (1, 1, 2, 1)),
         ((4, 4, 6, 6), (1, 1, 2, 2)), ((4, 6, 6, 6), (2, 1, 3, 2)), ((4, 4, 4, 4), (2, 2, 2, 2))]


def whole_dhop(gbx, p, prec, dag, cfg=None):
    geom = gbx.Geometry(p.dims)
    U = gbx.GaugeField(geom, prec).upload(p.dbl)
    src = gbx.SpinorField(geom, p.Ls, p.nrhs, prec).upload(p.psi)
    dst = gbx.SpinorField(geom, p.Ls, p.nrhs, prec)
    gbx.dhop(U, dst, src, dag=dag, cfg=cfg)
    return dst.download()


class Parts:
    def __init__(self, gbx, p, grid, prec):
        self.gbx, self.p, self.grid = gbx, p, tuple(grid)
        self.coords = gbx.grid_coords(grid)
        self.geoms = {c: gbx.Geometry.part_of(p.dims, grid, c) for c in self.coords}
        self.U = {c: gbx.GaugeField(g, prec).upload(p.dbl) for c, g in self.geoms.items()}
        loc = gbx.part_split(p.dims, grid, p.psi)
        self.src = {c: gbx.SpinorField(g, p.Ls, p.nrhs, prec).upload(loc[c]) for c, g in self.geoms.items()}
        self.dst = {c: gbx.SpinorField(g, p.Ls, p.nrhs, prec) for c, g in self.geoms.items()}

    def exchange(self, proj=False, dag=False):
        bufs = self.gbx.halo_buffers(self.src) if proj else None
        self.gbx.halo_exchange(self.src, proj=proj, dag=dag, bufs=bufs)
        return bufs

    def each(self):
        return [(self.U[c], self.dst[c], self.src[c]) for c in self.coords]

    def merged(self):
        for g in self.geoms.values():
            g.sync()
        return self.gbx.part_merge(self.p.dims, self.grid, {c: d.download() for c, d in self.dst.items()})


@pytest.mark.parametrize("dims,grid", GRIDS)
@pytest.mark.parametrize("dag", [False, True])
@pytest.mark.parametrize("prec,proj", [("f64", False), ("f64", True), ("f32", False), ("f32", True)])
def test_parts_match_whole_lattice(gpu, gbx, problem, dims, grid, dag, prec, proj):
    p = problem(dims, 4)
    parts = Parts(gbx, p, grid, prec)
    c0 = parts.coords[0]
    cfg = gbx.dhop_get_cfg(parts.U[c0], parts.src[c0])
    cfg.kind = 0
    want = whole_dhop(gbx, p, prec, dag, cfg=cfg)
    bufs = parts.exchange(proj=proj, dag=dag)
    for U, d, s in parts.each():
        gbx.dhop(U, d, s, dag=dag, cfg=cfg)
    got = parts.merged()
    print(f"dims={dims} grid={grid} dag={dag} prec={prec} proj={proj} max|diff|={np.abs(got - want).max():.3e}")
    assert np.array_equal(got, want)
    err = rel_l2(got[0], p.oracle_lex(dag=dag))
    assert err <= (1e-14 if prec == "f64" else 1e-6)
    del bufs


@pytest.mark.parametrize("dims,grid", [((4, 4, 4, 8), (1, 1, 1, 2)), ((4, 8, 4, 12), (1, 2, 1, 2)),
                                       ((4, 6, 6, 6), (2, 1, 3, 2)), ((4, 4, 4, 4), (2, 2, 2, 2))])
@pytest.mark.parametrize("dag", [False, True])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_faces_from_projected_halves_in_place(gpu, gbx, problem, dims, grid, dag, prec):
    """Projected faces packed straight into the neighbours' half-slots and
    read there by the face launch (no unpack): the whole lattice's Dhop to
    the bit."""
    p = problem(dims, 4)
    parts = Parts(gbx, p, grid, prec)
    c0 = parts.coords[0]
    cfg = gbx.dhop_get_cfg(parts.U[c0], parts.src[c0])
    want = whole_dhop(gbx, p, prec, dag, cfg=cfg if cfg.kind == 0 else None)
    gbx.halo_exchange_in_place(parts.src, dag=dag)
    for U, d, s in parts.each():
        gbx.dhop_range(U, d, s, 0, d.geom.ninterior, dag=dag)
        gbx.dhop_faces_projected(U, d, s, dag=dag)
    got = parts.merged()
    assert np.array_equal(got, want)
    assert rel_l2(got[0], p.oracle_lex(dag=dag)) <= (1e-14 if prec == "f64" else 1e-6)


@pytest.mark.parametrize("grid", [(1, 1, 1, 2), (1, 2, 1, 2)])
@pytest.mark.parametrize("dag", [False, True])
def test_interior_then_faces(gpu, gbx, problem, grid, dag):
    """gbx_dhop_range over [0, ninterior) before the exchange and over
    [ninterior, V) after it gives the full application."""
    p = problem((4, 8, 4, 12), 4)
    parts = Parts(gbx, p, grid, "f64")
    g0 = parts.geoms[parts.coords[0]]
    assert 0 < g0.ninterior < g0.volume
    for U, d, s in parts.each():
        gbx.dhop_range(U, d, s, 0, d.geom.ninterior, dag=dag)  # ghosts not yet filled
    parts.exchange()
    for U, d, s in parts.each():
        gbx.dhop_range(U, d, s, d.geom.ninterior, d.geom.volume, dag=dag)
    got = parts.merged()
    ref = Parts(gbx, p, grid, "f64")
    ref.exchange()
    for U, d, s in ref.each():
        gbx.dhop(U, d, s, dag=dag)
    assert np.array_equal(got, ref.merged())
    assert rel_l2(got[0], p.oracle_lex(dag=dag)) <= 1e-14


def test_range_on_whole_lattice(gpu, gbx, problem):
    """Slot ranges that tile the lattice reproduce gbx_dhop; untouched slots stay as they were."""
    p = problem((4, 4, 4, 4), 3)
    geom = gbx.Geometry(p.dims)
    U = gbx.GaugeField(geom, "f64").upload(p.dbl)
    src = gbx.SpinorField(geom, p.Ls, 1, "f64").upload(p.psi)
    full = gbx.SpinorField(geom, p.Ls, 1, "f64")
    gbx.dhop(U, full, src)
    out = gbx.SpinorField(geom, p.Ls, 1, "f64").zero()
    cuts = [0, 1, 37, 100, 255, 256]
    for a, b in zip(cuts[:-1], cuts[1:]):
        gbx.dhop_range(U, out, src, a, b)
    gbx.dhop_range(U, out, src, 40, 40)  # empty range
    assert np.array_equal(out.download(), full.download())
    part = gbx.SpinorField(geom, p.Ls, 1, "f64").zero()
    gbx.dhop_range(U, part, src, 0, 100)
    nz = np.abs(part.download()[0]).reshape(p.V, -1).sum(axis=1) > 0
    assert nz.sum() == 100
    with pytest.raises(gbx.GbxError):
        gbx.dhop_range(U, out, src, 10, 257)
    with pytest.raises(gbx.GbxError):
        gbx.dhop_range(U, out, src, 20, 10)


@pytest.mark.parametrize("prec", ["f16", "bf16"])
def test_parts_mfma_multi_rhs(gpu, gbx, problem, prec):
    """The MFMA kernel (FP16 / BF16, 16 minor columns) on parts with whole-spinor halos."""
    p = problem((4, 4, 4, 8), 4, nrhs=4)
    parts = Parts(gbx, p, (1, 1, 2, 2), prec)
    cfg = gbx.kcfg(4, 256, kind=1)
    want = whole_dhop(gbx, p, prec, False, cfg=cfg)
    parts.exchange()
    for U, d, s in parts.each():
        gbx.dhop(U, d, s, cfg=cfg)
    got = parts.merged()
    assert np.array_equal(got, want)
    for r in range(p.nrhs):
        assert rel_l2(got[r], p.oracle_lex(rhs=r)) <= storage_tol(prec, p.psi)


def test_projected_face_is_half_and_matches_projection(gpu, gbx, problem, orc):
    """A spin-projected top face along T holds, for DhopDag's backward leg
    (projector 1 - gamma_t: h0 = p0 - p2, h1 = p1 - p3), exactly those sums;
    along Z for Dhop's forward leg from the bottom face (projector 1 - gamma_z:
    h0 = p0 - i p2, h1 = p1 + i p3)."""
    p = problem((4, 4, 4, 8), 2)
    parts = Parts(gbx, p, (1, 1, 2, 2), "f64")
    c0 = parts.coords[0]
    x = parts.src[c0]
    loc = gbx.part_split(p.dims, (1, 1, 2, 2), p.psi)[c0][0]      # [V_local][Ls][4][3][2], local dims 4 4 2 4
    ld = (4, 4, 2, 4)
    M = p.Ls * p.nrhs

    def cplx(a):
        return a[..., 0] + 1j * a[..., 1]

    for dim, face, dag in ((3, 1, True), (2, 0, False)):
        assert x.halo_bytes(dim, proj=True) * 2 == x.halo_bytes(dim)
        buf = gbx.DeviceBuffer(x.halo_bytes(dim, proj=True))
        x.halo_pack(dim, face, buf.ptr, proj=True, dag=dag)
        x.geom.sync()
        F = x.geom.face[dim]
        h = buf.download(np.float64).reshape(F, 6, M, 2)  # [f][complex (spin*3+colour)][m][re, im]
        others = [e for e in range(4) if e != dim]
        for f in range(F):
            c = [0, 0, 0, 0]
            r = f
            for e in others:
                c[e] = r % ld[e]
                r //= ld[e]
            c[dim] = ld[dim] - 1 if face == 1 else 0
            site = c[0] + ld[0] * (c[1] + ld[1] * (c[2] + ld[2] * c[3]))
            for s in range(p.Ls):
                sp = cplx(loc[site, s])                      # [4][3] complex
                if dim == 3:
                    want = np.concatenate([sp[0] - sp[2], sp[1] - sp[3]])
                else:
                    want = np.concatenate([sp[0] - 1j * sp[2], sp[1] + 1j * sp[3]])
                got = cplx(h[f, :, s, :])
                assert np.array_equal(got, want), (dim, f, s)


@pytest.mark.parametrize("proj", [False, True])
@pytest.mark.parametrize("dag", [False, True])
def test_pack_all_equals_face_by_face(gpu, gbx, problem, proj, dag):
    """The one-launch pack / unpack of every face gives the bytes of the
    face-by-face entries."""
    grid = (1, 2, 2, 2)
    p = problem((4, 4, 4, 6), 3)
    parts = Parts(gbx, p, grid, "f64")
    x = parts.src[parts.coords[3]]
    cut = [1, 2, 3]
    one = {(d, f): gbx.DeviceBuffer(x.halo_bytes(d, proj)) for d in cut for f in (0, 1)}
    each = {(d, f): gbx.DeviceBuffer(x.halo_bytes(d, proj)) for d in cut for f in (0, 1)}
    x.halo_pack_all({k: b.ptr for k, b in one.items()}, proj=proj, dag=dag)
    for (d, f), b in each.items():
        x.halo_pack(d, f, b.ptr, proj=proj, dag=dag)
    x.geom.sync()
    for k in one:
        assert np.array_equal(one[k].download(), each[k].download()), k
    # unpack: all at once into x, one by one into a copy, compare the ghost ranges
    y = gbx.SpinorField(x.geom, p.Ls, p.nrhs, "f64").copy_from(x)
    x.halo_unpack_all({(d, s): one[(d, 1 - s)].ptr for d in cut for s in (0, 1)}, proj=proj)
    for d in cut:
        for s in (0, 1):
            y.halo_unpack(d, s, one[(d, 1 - s)].ptr, proj=proj)
    x.geom.sync()
    for d in cut:
        for s in (0, 1):
            assert np.array_equal(ghosts(x, d, s), ghosts(y, d, s)), (d, s)


def ghosts(x, d, side):
    """The bytes of the ghost range (d, side) of x."""
    import ctypes as C
    from gridbench_amd._lib import lib
    nb = x.halo_bytes(d)
    out = np.empty(nb, np.uint8)
    assert lib.gbx_buffer_download(C.c_void_p(x.halo_ghost(d, side)), nb, out.ctypes.data_as(C.c_void_p)) == 0
    return out


def test_halo_errors(gpu, gbx, problem):
    p = problem((4, 4, 4, 8), 2)
    whole = gbx.SpinorField(gbx.Geometry(p.dims), 2, 1, "f64")
    with pytest.raises(gbx.GbxError):
        whole.halo_bytes(3)
    parts = Parts(gbx, p, (1, 1, 1, 2), "f16")
    c0 = parts.coords[0]
    x = parts.src[c0]
    buf = gbx.DeviceBuffer(x.halo_bytes(3))
    with pytest.raises(gbx.GbxError):
        x.halo_pack(3, 0, buf.ptr, proj=True)   # projected 16-bit faces
    with pytest.raises(gbx.GbxError):
        x.halo_pack(3, 2, buf.ptr)              # face index
    with pytest.raises(gbx.GbxError):
        x.halo_pack(2, 0, buf.ptr)              # dimension 2 is not cut
    with pytest.raises(gbx.GbxError):
        x.halo_pack(3, 0, x.halo_ghost(3, 0))   # overlaps the field
    with pytest.raises(gbx.GbxError):
        x.halo_unpack(3, 1, x.halo_ghost(3, 0))  # overlaps the field
    x.halo_unpack(3, 0, x.halo_ghost(3, 0))      # in place: nothing to copy
    with pytest.raises(gbx.GbxError):  # the tiled kernel takes whole lattices only
        gbx.dhop(parts.U[c0], parts.dst[c0], x, cfg=gbx.kcfg(0, 256, kind=2, tile=gbx.pack_tile((2, 2, 2, 2))))
    with pytest.raises(gbx.GbxError):
        gbx.Geometry.part_of(p.dims, (1, 1, 1, 3), (0, 0, 0, 0))
    with pytest.raises(gbx.GbxError):  # projected ghosts of 16-bit fields
        x.halo_ghost_proj(3, 0)
    with pytest.raises(gbx.GbxError):
        gbx.dhop_faces_projected(parts.U[c0], parts.dst[c0], x)
    with pytest.raises(gbx.GbxError):  # not a part
        gbx.dhop_faces_projected(gbx.GaugeField(whole.geom, "f64"), gbx.SpinorField(whole.geom, 2, 1, "f64"), whole)


@pytest.mark.parametrize("grid,proj", [((1, 1, 1, 2), False), ((1, 1, 1, 2), True), ((1, 2, 1, 2), True)])
def test_overlap_timing_entry(gpu, gbx, problem, grid, proj):
    """The timed schedule runs, and the interior it computes is the true Dhop
    (the face sites are stand-in values by design)."""
    p = problem((4, 8, 4, 12), 4)
    parts = Parts(gbx, p, grid, "f64")
    parts.exchange()
    for U, d, s in parts.each():
        gbx.dhop(U, d, s)
    c0 = parts.coords[0]
    want = parts.dst[c0].download()
    ov, se = gbx.halo_overlap_time(parts.U[c0], parts.dst[c0], parts.src[c0], proj=proj, nrep=3)
    assert ov > 0 and se > 0
    got = parts.dst[c0].download()
    g0 = parts.geoms[c0]
    ld = g0.dims
    cut = [d for d in range(4) if grid[d] > 1]
    for site in range(g0.volume):
        c = [site % ld[0], site // ld[0] % ld[1], site // (ld[0] * ld[1]) % ld[2], site // (ld[0] * ld[1] * ld[2])]
        if all(0 < c[d] < ld[d] - 1 for d in cut):
            assert np.array_equal(got[0, site], want[0, site])


@pytest.mark.parametrize("dims,grid", [((4, 4, 4, 8), (1, 1, 1, 2)), ((4, 4, 4, 4), (1, 2, 2, 2)),
                                       ((8, 4, 4, 4), (2, 1, 1, 2))])
@pytest.mark.parametrize("dag", [False, True])
@pytest.mark.parametrize("proj", [False, True])
def test_red_black_parts(gpu, gbx, problem, dims, grid, dag, proj):
    """Red-black halves of domain parts: DhopEO and DhopOE on every block,
    after the halves' faces are exchanged, equal the whole lattice's halves
    (gbx_geom_create_rb) bit for bit."""
    p = problem(dims, 4)
    cfg = gbx.kcfg(4, 64)
    ge, go = gbx.Geometry.rb_pair(dims)
    Ue, Uo = gbx.GaugeField(ge, "f64").upload(p.dbl), gbx.GaugeField(go, "f64").upload(p.dbl)
    pe, po = gbx.rb_split(dims, p.psi)
    want = {}
    for par, (U, gout, gin, src) in enumerate(((Ue, ge, go, po), (Uo, go, ge, pe))):
        x = gbx.SpinorField(gin, p.Ls, 1, "f64").upload(src)
        y = gbx.SpinorField(gout, p.Ls, 1, "f64")
        gbx.dhop(U, y, x, dag=dag, cfg=cfg)
        want[par] = y.download()
    ld = tuple(d // n for d, n in zip(dims, grid))
    loc = gbx.part_split(dims, grid, p.psi)
    halves = {c: gbx.Geometry.part_rb_pair(dims, grid, c) for c in gbx.grid_coords(grid)}
    for par in (0, 1):  # out on parity par, in on 1 - par
        U, x, y = {}, {}, {}
        for c, (he, ho) in halves.items():
            hout, hin = (he, ho) if par == 0 else (ho, he)
            U[c] = gbx.GaugeField(hout, "f64").upload(p.dbl)
            x[c] = gbx.SpinorField(hin, p.Ls, 1, "f64").upload(gbx.rb_split(ld, loc[c])[1 - par])
            y[c] = gbx.SpinorField(hout, p.Ls, 1, "f64")
        if proj:
            gbx.halo_exchange_in_place(x, dag=dag)
            for c in x:
                gbx.dhop_range(U[c], y[c], x[c], 0, y[c].geom.ninterior, dag=dag)
                gbx.dhop_faces_projected(U[c], y[c], x[c], dag=dag)
        else:
            gbx.halo_exchange(x)
            for c in x:
                gbx.dhop(U[c], y[c], x[c], dag=dag, cfg=cfg)
        for g in (h for pair in halves.values() for h in pair):
            g.sync()
        zero = np.zeros_like(gbx.rb_split(ld, loc[next(iter(loc))])[0])
        full = {c: (gbx.rb_merge(ld, y[c].download(), zero) if par == 0 else gbx.rb_merge(ld, zero, y[c].download()))
                for c in x}
        got = gbx.rb_split(dims, gbx.part_merge(dims, grid, full))[par]
        assert np.array_equal(got, want[par]), (par, np.abs(got - want[par]).max())


def test_solvers_reject_domain_parts(gpu, gbx):
    """The solvers run on whole lattices: a domain part's halves (whose ghosts
    the solver would never refresh) and a part are refused."""
    he, ho = gbx.Geometry.part_rb_pair((4, 4, 4, 8), (1, 1, 1, 2), (0, 0, 0, 0))
    U = (gbx.GaugeField(he, "f64"), gbx.GaugeField(ho, "f64"))
    x = (gbx.SpinorField(he, 2, 1, "f64"), gbx.SpinorField(ho, 2, 1, "f64"))
    b = (gbx.SpinorField(he, 2, 1, "f64"), gbx.SpinorField(ho, 2, 1, "f64"))
    with pytest.raises(gbx.GbxError) as e:
        gbx.cg_solve_eo(U, None, x, b, 0.1)
    assert e.value.status == 1
    with pytest.raises(gbx.GbxError):
        gbx.dwf_solve_eo(U, None, x, b)
    g = gbx.Geometry.part_of((4, 4, 4, 8), (1, 1, 1, 2), (0, 0, 0, 0))
    Ug = gbx.GaugeField(g, "f64")
    xg, bg = gbx.SpinorField(g, 2, 1, "f64"), gbx.SpinorField(g, 2, 1, "f64")
    with pytest.raises(gbx.GbxError):
        gbx.cg_solve(Ug, None, xg, bg, 0.1)


def test_cg_over_parts_matches_the_whole_lattice(gpu, gbx, problem):
    """A solver over domain parts: CG on M^dag M x = b (M = 1 - kappa Dhop)
    with every M and M^dag applied block by block after a face exchange,
    and every inner product summed over the blocks, as ranks would run it.
    Its solution solves the whole lattice's system to the tolerance, and
    agrees with the library's whole-lattice solver."""
    dims, grid, kappa = (4, 4, 4, 8), (1, 1, 2, 2), 0.12
    p = problem(dims, 4)
    b_lex = p.psi
    parts = Parts(gbx, p, grid, "f64")            # src = b, one field per block
    coords = parts.coords
    mk = lambda: {c: gbx.SpinorField(parts.geoms[c], p.Ls, 1, "f64").zero() for c in coords}
    b, x, r, q, t, Aq = parts.src, mk(), mk(), mk(), mk(), mk()

    def apply_M(out, inp, dag):  # out = inp - kappa Dhop(inp), inp's faces exchanged first
        gbx.halo_exchange(inp)
        for c in coords:
            gbx.dhop_xpay(parts.U[c], out[c], inp[c], inp[c], 1.0, -kappa, dag=dag)

    def dot(u, v):
        return sum(u[c].dot(v[c]).real for c in coords)

    for c in coords:
        r[c].copy_from(b[c])
        q[c].copy_from(b[c])
    rr, bb = dot(r, r), dot(b, b)
    it = 0
    while rr > 1e-24 * bb and it < 500:
        apply_M(t, q, False)
        apply_M(Aq, t, True)
        alpha = rr / dot(q, Aq)
        for c in coords:
            x[c].axpy(alpha, q[c])
            r[c].axpy(-alpha, Aq[c])
        rr_new = dot(r, r)
        for c in coords:
            q[c].scale(rr_new / rr).axpy(1.0, r[c])
        rr, it = rr_new, it + 1
    assert rr <= 1e-24 * bb, it
    for g in parts.geoms.values():
        g.sync()
    xs = gbx.part_merge(dims, grid, {c: x[c].download() for c in coords})
    # the whole lattice: residual of the assembled solution, and the library's own solver
    geom = gbx.Geometry(dims)
    U = gbx.GaugeField(geom, "f64").upload(p.dbl)
    X = gbx.SpinorField(geom, p.Ls, 1, "f64").upload(xs)
    B = gbx.SpinorField(geom, p.Ls, 1, "f64").upload(b_lex)
    T = gbx.SpinorField(geom, p.Ls, 1, "f64")
    R = gbx.SpinorField(geom, p.Ls, 1, "f64")
    gbx.dhop_xpay(U, T, X, X, 1.0, -kappa)
    gbx.dhop_xpay(U, R, T, T, 1.0, -kappa, dag=True)
    R.axpy(-1.0, B)
    assert np.sqrt(R.norm2() / B.norm2()) <= 1e-10
    X2 = gbx.SpinorField(geom, p.Ls, 1, "f64").zero()
    gbx.cg_solve(U, None, X2, B, kappa, tol=1e-12)
    X2.axpy(-1.0, X)
    assert np.sqrt(X2.norm2() / X.norm2()) <= 1e-9
