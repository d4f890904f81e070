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

from conftest import rel_l2

pytestmark = pytest.mark.gpu

SPLITS = [((4, 4, 4, 8), 3, 2), ((4, 4, 4, 8), 3, 4), ((4, 4, 4, 8), 3, 8), ((8, 4, 4, 4), 0, 2),
          ((4, 6, 4, 4), 1, 3), ((4, 4, 4, 4), 2, 2)]


def whole_dhop(gbx, p, prec, dag, cfg=None):
    geom = gbx.Geometry(p.dims)
    U = gbx.GaugeField(geom, prec).upload(p.dbl)
    src = gbx.SpinorField(geom, p.Ls, p.nrhs, prec).upload(p.psi)
    dst = gbx.SpinorField(geom, p.Ls, p.nrhs, prec)
    gbx.dhop(U, dst, src, dag=dag, cfg=cfg)
    return dst.download()


class Parts:
    def __init__(self, gbx, p, split, nparts, prec):
        self.gbx, self.p, self.split, self.n = gbx, p, split, nparts
        self.geoms = [gbx.Geometry.part_of(p.dims, split, nparts, k) for k in range(nparts)]
        self.U = [gbx.GaugeField(g, prec).upload(p.dbl) for g in self.geoms]
        loc = gbx.part_split(p.dims, split, nparts, p.psi)
        self.src = [gbx.SpinorField(g, p.Ls, p.nrhs, prec).upload(a) for g, a in zip(self.geoms, loc)]
        self.dst = [gbx.SpinorField(g, p.Ls, p.nrhs, prec) for g in self.geoms]

    def exchange(self, proj=False, dag=False):
        bufs = None
        if proj:
            nb = self.src[0].halo_bytes(proj=True)
            bufs = [(self.gbx.DeviceBuffer(nb), self.gbx.DeviceBuffer(nb)) for _ in range(self.n)]
        self.gbx.halo_exchange(self.src, proj=proj, dag=dag, bufs=bufs)
        return bufs

    def merged(self):
        for g in self.geoms:
            g.sync()
        return self.gbx.part_merge(self.p.dims, self.split, [d.download() for d in self.dst])


@pytest.mark.parametrize("dims,split,nparts", SPLITS)
@pytest.mark.parametrize("dag", [False, True])
@pytest.mark.parametrize("prec,proj", [("f64", False), ("f64", True), ("f32", False), ("f32", True)])
def test_parts_match_whole_lattice(gpu, gbx, problem, dims, split, nparts, dag, prec, proj):
    p = problem(dims, 4)
    parts = Parts(gbx, p, split, nparts, prec)
    cfg = gbx.dhop_get_cfg(parts.U[0], parts.src[0])
    cfg.kind = 0
    want = whole_dhop(gbx, p, prec, dag, cfg=cfg)
    bufs = parts.exchange(proj=proj, dag=dag)
    for U, d, s in zip(parts.U, parts.dst, parts.src):
        gbx.dhop(U, d, s, dag=dag, cfg=cfg)
    got = parts.merged()
    print(f"dims={dims} split={split} nparts={nparts} dag={dag} prec={prec} proj={proj} "
          f"max|diff|={np.abs(got - want).max():.3e}")
    assert np.array_equal(got, want)
    err = rel_l2(got[0], p.oracle_lex(dag=dag))
    assert err <= (1e-14 if prec == "f64" else 1e-6)
    del bufs


@pytest.mark.parametrize("dag", [False, True])
def test_interior_then_boundary(gpu, gbx, problem, dag):
    """gbx_dhop_range over [0, ninterior) before the exchange and over
    [ninterior, V) after it gives the full application."""
    p = problem((4, 4, 4, 12), 4)
    parts = Parts(gbx, p, 3, 2, "f64")
    g0 = parts.geoms[0]
    assert 0 < g0.ninterior < g0.volume
    for U, d, s in zip(parts.U, parts.dst, parts.src):
        gbx.dhop_range(U, d, s, 0, d.geom.ninterior, dag=dag)  # ghosts not yet filled
    parts.exchange()
    for U, d, s in zip(parts.U, parts.dst, parts.src):
        gbx.dhop_range(U, d, s, d.geom.ninterior, d.geom.volume, dag=dag)
    got = parts.merged()
    ref = Parts(gbx, p, 3, 2, "f64")
    ref.exchange()
    for U, d, s in zip(ref.U, ref.dst, ref.src):
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
    parts = Parts(gbx, p, 3, 2, prec)
    cfg = gbx.kcfg(4, 256, kind=1)
    want = whole_dhop(gbx, p, prec, False, cfg=cfg)
    parts.exchange()
    for U, d, s in zip(parts.U, parts.dst, parts.src):
        gbx.dhop(U, d, s, cfg=cfg)
    got = parts.merged()
    assert np.array_equal(got, want)
    for r in range(p.nrhs):
        assert rel_l2(got[r], p.oracle_lex(rhs=r)) <= (2e-3 if prec == "f16" else 1.6e-2)


def test_projected_face_is_half_and_matches_projection(gpu, gbx, problem, orc):
    """A spin-projected top face holds, for DhopDag's backward leg along T
    (projector 1 - gamma_t: h0 = p0 - p2, h1 = p1 - p3), exactly those sums."""
    p = problem((4, 4, 4, 8), 2)
    parts = Parts(gbx, p, 3, 2, "f64")
    x = parts.src[0]
    assert x.halo_bytes(proj=True) * 2 == x.halo_bytes()
    buf = gbx.DeviceBuffer(x.halo_bytes(proj=True))
    x.halo_pack(1, buf.ptr, proj=True, dag=True)
    x.geom.sync()
    M, face = p.Ls * p.nrhs, x.geom.face
    h = buf.download(np.float64).reshape(face, 6, M, 2)  # [f][complex (spin*3+colour)][m][re, im]
    loc = gbx.part_split(p.dims, 3, 2, p.psi)[0][0]      # [V_local][Ls][4][3][2]
    ld = (4, 4, 4, 4)
    for f in range(face):
        site = f + face * (ld[3] - 1)                      # local lexicographic index of face site f, t = 3
        for s in range(p.Ls):
            sp = loc[site, s]
            want = np.concatenate([sp[0] - sp[2], sp[1] - sp[3]])  # [6][2]
            assert np.array_equal(h[f, :, s, :], want)


def test_halo_errors(gpu, gbx, problem):
    p = problem((4, 4, 4, 8), 2)
    whole = gbx.SpinorField(gbx.Geometry(p.dims), 2, 1, "f64")
    with pytest.raises(gbx.GbxError):
        whole.halo_bytes()
    parts = Parts(gbx, p, 3, 2, "f16")
    buf = gbx.DeviceBuffer(parts.src[0].halo_bytes())
    with pytest.raises(gbx.GbxError):
        parts.src[0].halo_pack(0, buf.ptr, proj=True)
    with pytest.raises(gbx.GbxError):
        parts.src[0].halo_pack(2, buf.ptr)
    with pytest.raises(gbx.GbxError):
        parts.src[0].halo_pack(0, parts.src[0].halo_ghost(0))  # overlaps the field
    with pytest.raises(gbx.GbxError):
        parts.src[0].halo_unpack(1, parts.src[0].halo_ghost(0))  # overlaps the field
    parts.src[0].halo_unpack(0, parts.src[0].halo_ghost(0))      # in place: nothing to copy
    with pytest.raises(gbx.GbxError):  # the tiled kernel takes whole lattices only
        gbx.dhop(parts.U[0], parts.dst[0], parts.src[0], cfg=gbx.kcfg(0, 256, kind=2, tile=gbx.pack_tile((2, 2, 2, 2))))
    with pytest.raises(gbx.GbxError):
        gbx.Geometry.part_of(p.dims, 3, 3, 0)


@pytest.mark.parametrize("proj", [False, True])
def test_overlap_timing_entry(gpu, gbx, problem, proj):
    """The timed overlap schedule runs, and the interior it computes is the
    true Dhop (the face sites are stand-in values by design)."""
    p = problem((4, 4, 4, 12), 4)
    parts = Parts(gbx, p, 3, 2, "f64")
    parts.exchange()
    for U, d, s in zip(parts.U, parts.dst, parts.src):
        gbx.dhop(U, d, s)
    want = parts.dst[0].download()
    g0 = parts.geoms[0]
    ov, se = gbx.halo_overlap_time(parts.U[0], parts.dst[0], parts.src[0], proj=proj, nrep=3)
    assert ov > 0 and se > 0
    got = parts.dst[0].download()
    # interior sites (local t = 1 .. 4 of 6) are unchanged
    got_i = got[0].reshape(6, 4 * 4 * 4, -1)[1:5]
    want_i = want[0].reshape(6, 4 * 4 * 4, -1)[1:5]
    assert np.array_equal(got_i, want_i)
