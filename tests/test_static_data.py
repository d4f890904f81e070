"""Reader / writer of the reference's static_data blobs (TableGenerate.cc:93-138).

CPU tests: round trip at the blob's 9-digit print precision, the text layout
TableGenerate's fprintf calls produce, split files, errors, and that the
written arrays link against the reference's own extern declarations
(arch/sse/static_data.h) when the reference sources are present.  The GPU
test writes a blob made from seeded inputs with the oracle's Dhop as
Psi_cpp, then runs gridbench_amd/static_data_check (the reference's check,
bench.cc:180-199) on it.
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"


def small_blob(nsite=16, Ls=2, nsimd=2, seed=3):
    rng = np.random.default_rng(seed)
    U = rng.standard_normal(nsite * 144 * nsimd)
    Phi = rng.standard_normal(nsite * 24 * Ls * nsimd)
    Psi = rng.standard_normal(nsite * 24 * Ls * nsimd) * 1e3
    nbr = rng.integers(0, nsite * Ls, nsite * Ls * 8, dtype=np.uint64)
    prm = rng.integers(0, 2, nsite * Ls * 8).astype(np.uint8)
    return nsite, Ls, nsimd, U, Phi, Psi, nbr, prm


def printed(v):
    """The value a reader gets back from TableGenerate's %16.8le."""
    return np.array([float("%16.8e" % x) for x in v])


def test_roundtrip_at_print_precision(gbx, tmp_path):
    nsite, Ls, nsimd, U, Phi, Psi, nbr, prm = small_blob()
    cc, h = tmp_path / "static_data.cc", tmp_path / "static_data.h"
    gbx.static_data_write(cc, h, nsite, Ls, nsimd, U, Phi, Psi, nbr, prm)
    assert gbx.static_data_read_header(h) == (nsite, Ls)
    b = gbx.static_data_read(cc)
    assert np.array_equal(b["U"], printed(U))
    assert np.array_equal(b["Phi"], printed(Phi))
    assert np.array_equal(b["Psi_cpp"], printed(Psi))
    assert np.array_equal(b["nbr"], nbr) and np.array_equal(b["prm"], prm)
    assert np.max(np.abs(b["U"] - U) / np.abs(U)) < 5.1e-9


def test_text_layout(gbx, tmp_path):
    nsite, Ls, nsimd, U, Phi, Psi, nbr, prm = small_blob(nsite=2, Ls=1, nsimd=1)
    cc, h = tmp_path / "s.cc", tmp_path / "s.h"
    gbx.static_data_write(cc, h, nsite, Ls, nsimd, U, Phi, Psi, nbr, prm)
    text = cc.read_text()
    assert text.startswith("#include <stdint.h>\ndouble U_static[] = { \n ")
    heads = re.findall(r"(double|uint64_t|uint8_t) (\w+)\[\] = \{", text)
    assert heads == [("double", "U_static"), ("double", "Phi_static"), ("double", "Psi_cpp_static"),
                     ("uint64_t", "nbr_static"), ("uint8_t", "prm_static")]
    assert text.count("    0}; \n ") == 5 and text.endswith("    0}; \n ")
    vals = re.findall(r"\n +([-0-9.e+]+), ", text)
    assert len(vals) == 2 * 144 + 2 * 2 * 24
    assert all(re.fullmatch(r"-?\d\.\d{8}e[-+]\d\d", v) for v in vals)
    assert "    0x%x, \n" % prm[0] in text and "    0x%x, \n" % nbr[0] in text
    assert h.read_text() == ("#include <stdint.h>\nconst uint64_t nsite = 2 ; \nconst int Ls = 1 ; \n"
                             "extern double   U_static[] ; \nextern double   Phi_static[] ; \n"
                             "extern double   Psi_cpp_static[] ; \nextern uint64_t nbr_static[] ; \n"
                             "extern uint8_t  prm_static[] ; \n")


def test_split_files_and_errors(gbx, tmp_path):
    nsite, Ls, nsimd, U, Phi, Psi, nbr, prm = small_blob(nsite=4, Ls=2, nsimd=1)
    cc, h = tmp_path / "s.cc", tmp_path / "s.h"
    gbx.static_data_write(cc, h, nsite, Ls, nsimd, U, Phi, Psi, nbr, prm)
    text = cc.read_text()
    cut = text.index("double Phi_static")
    gauge, fermion = tmp_path / "static_data_gauge.cc", tmp_path / "static_data_fermion.cc"
    gauge.write_text(text[:cut])
    fermion.write_text("#include <stdint.h>\n" + text[cut:])
    b = gbx.static_data_read([gauge, fermion])
    assert np.array_equal(b["U"], printed(U)) and np.array_equal(b["nbr"], nbr)
    with pytest.raises(gbx.GbxError) as e:
        gbx.static_data_read(gauge)  # the fermion arrays are missing
    assert e.value.status == 1
    with pytest.raises(gbx.GbxError) as e:
        gbx.static_data_read([cc, gauge])  # U_static defined twice
    assert e.value.status == 1
    bad = tmp_path / "bad.cc"
    bad.write_text(text.replace("uint8_t prm_static[] = { \n     0x%x," % prm[0],
                                "uint8_t prm_static[] = { \n     0x2,"))
    with pytest.raises(gbx.GbxError) as e:
        gbx.static_data_read(bad)
    assert e.value.status == 4


def test_malformed_blobs_are_rejected(gbx, tmp_path):
    """Truncated arrays, negative or oversized integers and headers without
    values fail with GBX_EINVAL instead of being read past or wrapped."""
    nsite, Ls, nsimd, U, Phi, Psi, nbr, prm = small_blob(nsite=4, Ls=2, nsimd=1)
    cc, h = tmp_path / "s.cc", tmp_path / "s.h"
    gbx.static_data_write(cc, h, nsite, Ls, nsimd, U, Phi, Psi, nbr, prm)
    text = cc.read_text()
    bad = tmp_path / "bad.cc"
    for cut in (0, 10, len(text) // 3, len(text) // 2, len(text) - 5):
        bad.write_text(text[:cut])
        with pytest.raises(gbx.GbxError) as e:
            gbx.static_data_read(bad)
        assert e.value.status == 1, cut
    for arrays, status in (("uint64_t nbr_static[] = { 0, -1, 0 };", 1),
                           ("uint64_t nbr_static[] = { 0, 99999999999999999999999, 0 };", 1),
                           ("uint8_t prm_static[] = { 0, 300, 0 };", 4)):  # 4: an invalid table entry
        bad.write_text(text.split("uint64_t nbr_static")[0] + arrays + "\n")
        with pytest.raises(gbx.GbxError) as e:
            gbx.static_data_read(bad)
        assert e.value.status == status, arrays
    hb = tmp_path / "bad.h"
    for header in ("const uint64_t nsite ; const int Ls ;", "const uint64_t nsite = 0 ; const int Ls = 2 ;", ""):
        hb.write_text(header)
        with pytest.raises(gbx.GbxError) as e:
            gbx.static_data_read_header(hb)
        assert e.value.status == 1, header


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference sources not present")
def test_links_against_the_references_declarations(gbx, tmp_path):
    """The written arrays satisfy the extern declarations the reference's
    bench includes (arch/sse/static_data.h)."""
    nsite, Ls, nsimd, U, Phi, Psi, nbr, prm = small_blob(nsite=4, Ls=2, nsimd=1)
    cc, h = tmp_path / "static_data.cc", tmp_path / "static_data.h"
    gbx.static_data_write(cc, h, nsite, Ls, nsimd, U, Phi, Psi, nbr, prm)
    main = tmp_path / "main.cc"
    main.write_text('#include "arch/sse/static_data.h"\n'
                    "int main() { return (U_static[0] == 0.0) + (int)(nbr_static[0] > ~0ull) + prm_static[0] * 0"
                    " + (Phi_static[1] == Psi_cpp_static[1]); }\n")
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++14", "-I", REFERENCE, str(main), str(cc), "-o", str(exe)], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


@pytest.mark.gpu
@pytest.mark.parametrize("nsimd", [1, 2, 4])
def test_blob_of_seeded_inputs_passes_the_references_check(gpu, gbx, orc, problem, tmp_path, nsimd):
    p = problem((8, 8, 8, 8), 8, gseed=101, sseed=102)
    V, Ls = p.V, p.Ls
    nsite = V // nsimd
    nbr, prm = orc.tables(p.dims, nsimd, Ls)
    U = orc.pack_gauge(p.dims, nsimd, p.dbl)
    Phi = orc.pack_spinor(p.dims, nsimd, Ls, p.psi[0])
    Psi = orc.hand(orc.aligned_copy(U), orc.aligned_copy(Phi), nbr, prm, nsite, Ls, nsimd, nthreads=8)
    cc, h = tmp_path / "static_data.cc", tmp_path / "static_data.h"
    gbx.static_data_write(cc, h, nsite, Ls, nsimd, U, Phi, Psi, nbr, prm)
    exe = os.path.join(ROOT, "gridbench_amd", "static_data_check")
    for extra, tol in (([], "1e-7"), (["--f32"], "1e-6")):
        r = subprocess.run([exe, str(h), str(cc), "--tol", tol] + extra, capture_output=True, text=True, timeout=120)
        print(r.stdout.splitlines()[0], r.stdout.splitlines()[-1])
        assert r.returncode == 0, r.stdout + r.stderr
