"""The Fortran host layer (ISO_C_BINDING, built with amdflang) drives the
GPU build exactly as openLOWDIN's WaveFunction direct branches do; its raw
library results are checked against the oracle and its scaled species
matrices against the reference's caller semantics (q^2 scaling, the
one-particle skip, q_i q_j scaling and upper-to-lower symmetrisation)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from openlowdin_amd import system as S

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "openlowdin_amd", "fortran", "fock_host_driver")


def _write_job(d, system, dens):
    with open(os.path.join(d, "job.txt"), "w") as fh:
        fh.write("%d\n" % len(system.species))
        for k, sp in enumerate(system.species):
            fh.write("%d %d %.17g %.17g %.17g %d %d\n" % (k + 1, int(sp.is_electron), sp.charge, sp.kappa, sp.eta,
                                                         sp.n_particles, len(sp.shells)))
            for sh in sp.shells:
                fh.write("%d %d %.17g %.17g %.17g\n" % (sh.l, sh.nprim, *sh.origin))
                fh.write(" ".join("%.17g" % x for x in sh.exps) + "\n")
                fh.write(" ".join("%.17g" % x for x in sh.coefs) + "\n")
    for k, D in enumerate(dens):
        np.asfortranarray(D).ravel(order="F").tofile(os.path.join(d, "dens_%d.bin" % (k + 1)))


def _read(d, name, n):
    return np.fromfile(os.path.join(d, name), dtype=np.float64).reshape((n, n), order="F")


def test_driver_is_built():
    assert os.path.exists(DRIVER), "run __graft_entry__.build() (amdflang driver)"


@pytest.mark.gpu
def test_fortran_driver_fhf_anion(tmp_path):
    system = S.fhf_anion("6-31G")
    e, p = system.species
    rng = np.random.default_rng(23)
    Ce = rng.normal(size=(e.nbasis, 10)) / np.sqrt(e.nbasis)
    De = 2.0 * Ce @ Ce.T
    cp = rng.normal(size=p.nbasis) / np.sqrt(p.nbasis)
    Dp = np.outer(cp, cp)
    _write_job(str(tmp_path), system, [De, Dp])
    print("\n[case] fortran driver: FHF- 6-31G + proton", flush=True)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    be, bp = O.Basis(e), O.Basis(p)
    # raw results = the oracle (reference semantics) to 1e-10
    Ge = _read(str(tmp_path), "g_1.bin", e.nbasis)
    assert np.max(np.abs(Ge - O.compute_2body_direct(be, De, -0.5))) < 1e-10
    Gp = _read(str(tmp_path), "g_2.bin", p.nbasis)
    assert np.max(np.abs(Gp - O.compute_2body_direct(bp, Dp, -1.0))) < 1e-10
    c12 = _read(str(tmp_path), "c_1_2.bin", e.nbasis)
    ref12 = O.compute_coupling_direct(be, bp, 1, 2, Dp)
    assert np.max(np.abs(c12.ravel(order="F") - ref12.ravel())) < 1e-10
    c21 = _read(str(tmp_path), "c_2_1.bin", p.nbasis)
    ref21 = O.compute_coupling_direct(bp, be, 2, 1, De)
    assert np.max(np.abs(c21.ravel(order="F") - ref21.ravel())) < 1e-10
    # caller semantics: electrons have 20 particles (G scaled by q^2 = 1),
    # the single proton's G is not built; coupling scaled by q_e q_p = -1
    two1 = _read(str(tmp_path), "two_1.bin", e.nbasis)
    assert np.max(np.abs(two1 - Ge)) == 0.0
    two2 = _read(str(tmp_path), "two_2.bin", p.nbasis)
    assert np.max(np.abs(two2)) == 0.0
    cpl1 = _read(str(tmp_path), "coupling_1.bin", e.nbasis)
    up = np.triu(c12)
    expect = -(up + np.triu(up, 1).T)
    assert np.max(np.abs(cpl1 - expect)) < 1e-14
    cpl2 = _read(str(tmp_path), "coupling_2.bin", p.nbasis)
    # SCF_GHOST_SPECIES = the proton: the electrons' coupling leaves it out,
    # the proton still feels the electrons (WaveFunction.f90:1502-1503)
    with open(os.path.join(str(tmp_path), "ghost.txt"), "w") as fh:
        fh.write("2\n")
    print("[case] fortran driver: FHF- with the proton as ghost species", flush=True)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert np.max(np.abs(_read(str(tmp_path), "coupling_1.bin", e.nbasis))) == 0.0
    assert np.max(np.abs(_read(str(tmp_path), "coupling_2.bin", p.nbasis) - cpl2)) < 1e-14


@pytest.mark.gpu
def test_fortran_driver_g12_and_first_quarter(tmp_path):
    """The Fortran layer's G12 and first-quarter-transform wrappers (the
    bind(C) names of Libint2Interface.f90) against the oracle: the geminal
    Fock and coupling builds of FHF- and the transform on orbital row 3."""
    system = S.fhf_anion("6-31G")
    e, p = system.species
    rng = np.random.default_rng(29)
    Ce = rng.normal(size=(e.nbasis, 10)) / np.sqrt(e.nbasis)
    De = 2.0 * Ce @ Ce.T
    cp = rng.normal(size=p.nbasis) / np.sqrt(p.nbasis)
    Dp = np.outer(cp, cp)
    _write_job(str(tmp_path), system, [De, Dp])
    gexp = [4.0, 1.2, 0.3, 0.0]
    gcoef = [0.8, -0.35, 0.6, 0.05]
    with open(os.path.join(str(tmp_path), "g12.txt"), "w") as fh:
        fh.write("%d\n%s\n%s\n" % (len(gexp), " ".join("%.17g" % x for x in gexp),
                                   " ".join("%.17g" % x for x in gcoef)))
    with open(os.path.join(str(tmp_path), "it.txt"), "w") as fh:
        fh.write("3\n")
    print("\n[case] fortran driver: FHF- 6-31G + proton, G12 (4 terms) and first quarter p=3", flush=True)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    be, bp = O.Basis(e), O.Basis(p)
    g1 = _read(str(tmp_path), "g12_1.bin", e.nbasis)
    assert np.max(np.abs(g1 - O.compute_g12_direct(be, De, -0.5, gcoef, gexp))) < 1e-10
    g2 = _read(str(tmp_path), "g12_2.bin", p.nbasis)
    assert np.max(np.abs(g2 - O.compute_g12_direct(bp, Dp, -1.0, gcoef, gexp))) < 1e-10
    c12 = _read(str(tmp_path), "g12c_1_2.bin", e.nbasis)
    ref12 = O.compute_g12inter_direct(be, bp, 1, 2, Dp, gcoef, gexp)
    assert np.max(np.abs(c12.ravel(order="F") - ref12.ravel())) < 1e-10
    c21 = _read(str(tmp_path), "g12c_2_1.bin", p.nbasis)
    ref21 = O.compute_g12inter_direct(bp, be, 2, 1, De, gcoef, gexp)
    assert np.max(np.abs(c21.ravel(order="F") - ref21.ravel())) < 1e-10
    # first quarter: the Fortran array (n, n, n) holds the C image j n^2 + k n + l
    n = e.nbasis
    it = np.fromfile(os.path.join(str(tmp_path), "it_1.bin"), dtype=np.float64).reshape(n, n, n)
    ref = O.compute_2body_directIT(be, De, De, 3)
    assert np.max(np.abs(it - ref)) < 1e-10


@pytest.mark.gpu
def test_fortran_driver_stores_and_inter_transform(tmp_path):
    """The Fortran layer's MEMORY-path wrappers (intra- and inter-species
    integral stores, Libint2Interface.f90:1065-1179) and the inter-species
    first-quarter transform (:677-707): their arrays, ranks and by-reference
    p against the same library through its C ABI from Python (the stores bit
    for bit: each stored element is written once; the transform, which
    accumulates with atomics, to summation order), and the transform against
    the oracle."""
    from openlowdin_amd import libint2_interface as LI
    system = S.fhf_anion("6-31G")
    e, p = system.species
    rng = np.random.default_rng(31)
    Ce = rng.normal(size=(e.nbasis, 10)) / np.sqrt(e.nbasis)
    De = 2.0 * Ce @ Ce.T
    cp = rng.normal(size=p.nbasis) / np.sqrt(p.nbasis)
    Dp = np.outer(cp, cp)
    _write_job(str(tmp_path), system, [De, Dp])
    with open(os.path.join(str(tmp_path), "all.txt"), "w") as fh:
        fh.write("2\n")
    print("\n[case] fortran driver: FHF- 6-31G + proton, stores and inter-species transform p=2", flush=True)
    r = subprocess.run([DRIVER, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    ne, npr = e.nbasis, p.nbasis
    he, hp = LI.Libint2Interface(e, 1), LI.Libint2Interface(p, 2)
    he.init_2body_ints()
    hp.init_2body_ints()
    rd = lambda name, shape: np.fromfile(os.path.join(str(tmp_path), name), dtype=np.float64).reshape(shape)
    for h, D, n, sid in ((he, De, ne, 1), (hp, Dp, npr, 2)):
        got = rd("all_%d.bin" % sid, (n, n, n, n))
        ref = h.compute_2body_intraspecies_direct_all(D)
        print("   store %d: %d nonzero, max|d| %.2e" % (sid, np.count_nonzero(got), np.max(np.abs(got - ref))))
        assert np.count_nonzero(got) > 0 and np.array_equal(got, ref)
    got = rd("allc_1_2.bin", (ne, ne, npr, npr))
    ref = he.compute_2body_interspecies_direct_all(hp, De)
    assert np.count_nonzero(got) > 0 and np.array_equal(got, ref)
    be, bp = O.Basis(e), O.Basis(p)
    for (h, o, D, n, no, sid, oid, bt, bo) in ((he, hp, De, ne, npr, 1, 2, be, bp), (hp, he, Dp, npr, ne, 2, 1, bp, be)):
        got = rd("itc_%d_%d.bin" % (sid, oid), (n, no, no))
        # the transform accumulates with atomics: equal to summation order
        assert np.max(np.abs(got - h.compute_2body_interspecies_direct_IT(o, D, D, 2))) < 1e-13
        ref = O.compute_coupling_directIT(bt, bo, sid, oid, D, 2)
        print("   transform %d<-%d: max|d| vs oracle %.2e" % (sid, oid, np.max(np.abs(got - ref))))
        assert np.max(np.abs(got - ref)) < 1e-10
