"""Domain-part tables (gbx_part_tables, the host side of gbx_geom_create_part),
checked on the CPU against the periodic lattice itself: for every part, slot
and leg, the site a leg reads (a local slot, or a ghost slot filled from the
neighbouring part's face) must be the global neighbour x +/- mu.

The reference runs on one rank (TableGenerate.cc:84 asserts that every
stencil neighbour is local), so there is no reference decomposition to
compare with: the periodic neighbour relation is the ground truth.
"""
import numpy as np
import pytest

CASES = [
    ((4, 4, 4, 8), 3, 2),
    ((4, 4, 4, 8), 3, 4),
    ((4, 4, 4, 8), 3, 8),   # one layer per part: both faces are the same sites
    ((8, 4, 4, 4), 0, 2),
    ((4, 6, 4, 4), 1, 3),
    ((4, 4, 4, 4), 2, 2),   # two layers per part: no interior
    ((6, 4, 2, 10), 3, 5),
]


def coords(site, dims):
    c = []
    for d in dims:
        c.append(site % d)
        site //= d
    return c


def lex(c, dims):
    s, stride = 0, 1
    for x, d in zip(c, dims):
        s += x * stride
        stride *= d
    return s


def global_nbr(site, mu, dims):
    c = coords(site, dims)
    d = mu & 3
    c[d] = (c[d] + (1 if mu < 4 else -1)) % dims[d]
    return lex(c, dims)


@pytest.mark.parametrize("dims,split,nparts", CASES)
def test_legs_read_the_global_neighbour(gbx, dims, split, nparts):
    parts = [gbx.part_tables(dims, split, nparts, p) for p in range(nparts)]
    vol, face = parts[0]["vol"], parts[0]["face"]
    ld = [d // nparts if i == split else d for i, d in enumerate(dims)]
    assert vol * nparts == int(np.prod(dims)) and face * ld[split] == vol
    seen = np.zeros(int(np.prod(dims)), np.int64)
    for p, t in enumerate(parts):
        prev, nxt = parts[(p - 1) % nparts], parts[(p + 1) % nparts]
        sos, gs, nbr = t["site_of_slot"], t["gsite"], t["nbr"]
        assert sorted(sos.tolist()) == list(range(vol))
        for q in range(vol):
            c = coords(int(sos[q]), ld)
            c[split] += p * ld[split]
            assert gs[q] == lex(c, dims)
        seen[gs] += 1
        for q in range(vol):
            for mu in range(8):
                want = global_nbr(int(gs[q]), mu, dims)
                n = int(nbr[q, mu])
                if n < vol:
                    got = gs[n]
                elif n < vol + face:
                    got = prev["gsite"][prev["face_slots"][1][n - vol]]
                else:
                    assert n < vol + 2 * face
                    got = nxt["gsite"][nxt["face_slots"][0][n - vol - face]]
                assert got == want, (p, q, mu, n)
    assert (seen == 1).all()


@pytest.mark.parametrize("dims,split,nparts", CASES)
def test_interior_and_faces(gbx, dims, split, nparts):
    ld = [d // nparts if i == split else d for i, d in enumerate(dims)]
    for p in range(nparts):
        t = gbx.part_tables(dims, split, nparts, p)
        vol, face, nint = t["vol"], t["face"], t["ninterior"]
        assert nint == (vol - 2 * face if ld[split] > 2 else 0)
        ghost_legs = (t["nbr"] >= vol).sum(axis=1)
        assert (ghost_legs[:nint] == 0).all()
        # a face site has one leg across the cut (two when the part is one layer thick)
        assert (ghost_legs[nint:] == (2 if ld[split] == 1 else 1)).all()
        for side, layer in ((0, 0), (1, ld[split] - 1)):
            fs = t["face_slots"][side]
            for f, q in enumerate(fs):
                c = coords(int(t["site_of_slot"][q]), ld)
                assert c[split] == layer
                other = [c[d] for d in range(4) if d != split]
                assert lex(other, [ld[d] for d in range(4) if d != split]) == f


@pytest.mark.parametrize("args", [((4, 4, 4, 8), 4, 2, 0), ((4, 4, 4, 8), 3, 1, 0), ((4, 4, 4, 8), 3, 3, 0),
                                  ((4, 4, 4, 8), 3, 2, 2), ((4, 4, 4, 8), 3, 2, -1)])
def test_rejects_bad_decompositions(gbx, args):
    with pytest.raises(gbx.GbxError) as e:
        gbx.part_tables(*args)
    assert e.value.status == 1  # GBX_EINVAL
