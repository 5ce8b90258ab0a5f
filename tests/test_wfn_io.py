"""Restart / interoperability files (openlowdin_amd/wfn_io.py) against the
files openLOWDIN's own tests hold:

* test/H.APMO.READCOEFF.plainvec (tests/golden/, data): the coefficients of
  e-(H) cc-pVDZ UHF (multiplicity 2) + H_1 DZSNB.  Read as
  DensityMatrixSCFGuess does, they give the reference's HF energy
  (test/H.APMO.READCOEFF.py: -0.464889091726, 1e-8) already at the first
  iteration, and the SCF started from them stays there.  A transposed read
  (column-major for the formatted file) or a species mix-up fails.
* test/Gly.e+-readFchk.E-.fchk (tests/golden/, data): the electron
  orbitals of the glycine radical in 6-311G.  The shell types agree with
  this package's shell order, the orbitals are orthonormal in this
  package's overlap matrix (so basis, normalisation and order agree with
  the program that wrote them), and eta C_occ C_occ^T reproduces the file's
  own Total SCF Density.
* Round trips of plainvec / vec / lowdin.wfn, and the record framing of
  the binary files.
"""
import os
import struct

import numpy as np
import pytest

import oracle as O
from openlowdin_amd import system as S
from openlowdin_amd import wfn_io as W

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
READCOEFF = os.path.join(GOLD, "H.APMO.READCOEFF.")
FCHK = os.path.join(GOLD, "Gly.e+-readFchk.E-.fchk")
E_READCOEFF = -0.464889091726   # test/H.APMO.READCOEFF.py


def readcoeff_system():
    """test/H.APMO.READCOEFF.lowdin: e-(H) cc-pVDZ multiplicity 2 (UHF),
    H_1 DZSNB, both at the origin (no point charges)."""
    return S.build_system([("H", (0.0, 0.0, 0.0))], "CC-PVDZ",
                          quantum=[("H_1", "DZSNB", (0.0, 0.0, 0.0), S.PROTON_MASS)], open_shell=True)


def readcoeff_guess(system):
    out = []
    for sp in system.species:
        C, _ = W.read_coefficients(READCOEFF, sp.name, sp.nbasis)
        out.append(W.guess_density(C, sp.n_occupied, sp.eta))
    return out


def test_plainvec_fixture_layout():
    sysm = readcoeff_system()
    assert [(sp.name, sp.nbasis) for sp in sysm.species] == [("E-ALPHA", 5), ("E-BETA", 5), ("H_1", 2)]
    C, path = W.read_coefficients(READCOEFF, "E-ALPHA", 5)
    assert path.endswith("plainvec")
    # first row of the file: -5.38845703E-01 -1.26214243E+00 0 0 0 (row-major)
    assert C[0, 0] == pytest.approx(-0.538845703) and C[0, 1] == pytest.approx(-1.26214243)
    eps = W.read_orbital_energies(path, "E-ALPHA")
    assert eps.shape == (5,) and eps[0] == pytest.approx(-0.48101042894694890, abs=0)
    assert W.read_coefficients(READCOEFF, "E+", 5) is None   # failContinue: guess falls back


def test_read_coefficients_give_reference_energy():
    import scf
    sysm = readcoeff_system()
    hist = []
    E, _ = scf.run_scf(sysm, guess=readcoeff_guess(sysm), history=hist)
    assert abs(hist[0] - E_READCOEFF) < 1e-8, hist[0]
    assert abs(E - E_READCOEFF) < 1e-8
    assert len(hist) <= 4
    # the transposed reading is not a stationary point of the same energy
    sp = sysm.species[0]
    C, _ = W.read_coefficients(READCOEFF, sp.name, sp.nbasis)
    bad = list(readcoeff_guess(sysm))
    bad[0] = W.guess_density(C.T, sp.n_occupied, sp.eta)
    hist_bad = []
    scf.run_scf(sysm, guess=bad, history=hist_bad, max_iter=1)
    assert abs(hist_bad[0] - E_READCOEFF) > 1e-3


def test_plainvec_and_vec_round_trip(tmp_path):
    sysm = readcoeff_system()
    blocks = []
    for sp in sysm.species:
        C, path = W.read_coefficients(READCOEFF, sp.name, sp.nbasis)
        blocks.append((sp.name, C, W.read_orbital_energies(path, sp.name)))
    plain = str(tmp_path / "x.plainvec")
    W.write_vec(plain, blocks, binary=False)
    binf = str(tmp_path / "y.vec")
    W.write_vec(binf, blocks, binary=True)
    for name, C, eps in blocks:
        Cp, _ = W.read_coefficients(str(tmp_path / "x."), name, C.shape[0])
        assert np.array_equal(Cp, C)          # ES15.8 values re-written in ES15.8
        Cb, fb = W.read_coefficients(str(tmp_path / "y."), name, C.shape[0])
        assert fb.endswith(".vec") and np.array_equal(Cb, C)
        assert np.array_equal(W.read_orbital_energies(fb, name), eps)
        assert np.array_equal(W.read_orbital_energies(plain, name), eps)
    # plainvec is preferred over vec when both exist (DensityMatrixSCFGuess.f90:84-100)
    os.rename(binf, str(tmp_path / "x.vec"))
    _, f = W.read_coefficients(str(tmp_path / "x."), "E-ALPHA", 5)
    assert f.endswith("plainvec")


def test_binary_record_framing(tmp_path):
    """gfortran framing: 4-byte lengths around each record; labels of 30
    characters; the count as integer(8); values column-major."""
    A = np.arange(6.0).reshape(2, 3)
    path = str(tmp_path / "a.vec")
    W.write_vec(path, [("E-", A, np.array([1.5, 2.5]))], binary=True)
    raw = open(path, "rb").read()
    assert struct.unpack_from("<i", raw, 0)[0] == 30
    assert raw[4:34].decode() == "COEFFICIENTS".ljust(30)
    assert struct.unpack_from("<i", raw, 34)[0] == 30
    recs = W.read_records(path)
    assert [len(r) for r in recs[:4]] == [30, 30, 8, 48]
    assert struct.unpack("<q", recs[2])[0] == 6
    assert list(np.frombuffer(recs[3], "<f8")) == [0.0, 3.0, 1.0, 4.0, 2.0, 5.0]
    with open(path, "r+b") as fh:   # corrupt a trailer
        fh.seek(34)
        fh.write(struct.pack("<i", 31))
    with pytest.raises(W.WfnFormatError):
        W.read_records(path)


def test_wfn_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    blocks = {}
    n = 4
    for lab in W.WFN_SPECIES_ORDER:
        if lab == "EXTERNAL":
            continue
        if lab in ("ORBITALS", "REMOVED-ORBITALS"):
            blocks[lab] = rng.normal(size=n)
        elif lab == "EXCHANGE-CORRELATION-ENERGY":
            blocks[lab] = 0.0
        else:
            blocks[lab] = rng.normal(size=(n, n))
    path = str(tmp_path / "lowdin.wfn")
    W.write_wfn(path, [("E-", blocks)], {"TOTALENERGY": -76.0, "VIRIAL": 2.0})
    got = W.read_wfn(path)
    assert np.array_equal(W.matrix_from_binary(got[("FOCK", "E-")], n, n), blocks["FOCK"])
    assert got[("TOTALENERGY",)][0] == -76.0 and got[("VIRIAL",)][0] == 2.0
    order = [k[0] for k in got]
    assert order[:len(W.WFN_SPECIES_ORDER) - 1] == [x for x in W.WFN_SPECIES_ORDER if x != "EXTERNAL"]
    assert order[-5:] == W.WFN_SCALARS


GLY = [("N", (0.92914, 0.65159, 0.25796)), ("C", (0.08973, -0.44202, 0.82116)),
       ("C", (0.09416, -0.39610, 2.35537)), ("O", (0.79215, 0.47431, 2.87347)),
       ("O", (-0.60778, -1.24843, 2.89495)), ("H", (1.89448, 0.53627, 0.52036)),
       ("H", (0.48806, -1.38246, 0.47040)), ("H", (-0.91478, -0.32058, 0.44545))]


def test_fchk_orbitals_in_this_basis():
    """test/Gly.e+-readFchk: electrons of the glycine radical, 6-311G."""
    sp = S.build_system(GLY, "6-311G").species[0]
    arr = W.read_fchk_arrays(FCHK)
    types = arr["Shell types"]
    assert list(np.abs(types)) == [sh.l for sh in sp.shells]
    C = W.read_fchk_coefficients(FCHK, [sh.l for sh in sp.shells])
    assert C.shape == (sp.nbasis, sp.nbasis) == (74, 74)
    Smat = O.one_body(O.Basis(sp), 1)
    Ct = C.T @ Smat @ C
    assert np.max(np.abs(Ct - np.eye(74))) < 1e-6
    D = W.guess_density(C, 19, 2.0)
    tri = arr["Total SCF Density"]
    Dfile = np.zeros((74, 74))
    Dfile[np.tril_indices(74)] = tri
    Dfile = Dfile + np.tril(Dfile, -1).T
    assert np.max(np.abs(D - Dfile)) < 1e-6


def test_fchk_d_reorder(tmp_path):
    """Cartesian d rows go from xx, yy, zz, xy, xz, yz to xx, xy, xz, yy, yz, zz."""
    n = 7   # one s, one d shell
    Cl = np.arange(n * n, dtype=float).reshape(n, n) + 1.0   # wanted (this package's order)
    gauss = np.empty_like(Cl)
    gauss[0] = Cl[0]
    for k_l, k_g in enumerate(W.FCHK_D_TO_LOWDIN):
        gauss[1 + k_g] = Cl[1 + k_l]
    flat = gauss.reshape(-1, order="F")
    lines = ["title", "SP RHF", "%-40s%-4s%-6s%12d" % ("Alpha MO coefficients", "R", "N=", flat.size)]
    for c in range(0, flat.size, 5):
        lines.append("".join("%16.8E" % v for v in flat[c:c + 5]))
    path = str(tmp_path / "t.fchk")
    open(path, "w").write("\n".join(lines) + "\n")
    C = W.read_fchk_coefficients(path, [0, 2])
    assert np.array_equal(C, Cl)
    # gauss row 1 (xx), 4 (xy), 5 (xz), 2 (yy), 6 (yz), 3 (zz)
    assert np.array_equal(C[1:, 0], gauss[[1, 4, 5, 2, 6, 3], 0])
