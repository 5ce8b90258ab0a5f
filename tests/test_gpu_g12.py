"""GPU Gaussian-geminal (G12) Fock and coupling builds through the C ABI,
against the oracle and against an identity the oracle does not enter.

Geminals: the 26-term VHH-CCSDT interparticle potential (with its constant
g = 0 term) and the 27-term VOH-CCSDT expansion from openLOWDIN's own
TIP4P-dimer test input (tests/golden/g12_potentials.json), and a short
synthetic one.  Bars: every element within 1e-10 absolute of the oracle.
Every test prints its case before the GPU call.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from openlowdin_amd import system as S

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-10


def _li():
    from openlowdin_amd import libint2_interface as LI
    return LI


def _case(msg):
    print("\n[case] " + msg, flush=True)


def _density(n, nocc, rng, eta=2.0):
    C = rng.normal(size=(n, nocc)) / np.sqrt(n)
    return eta * C @ C.T


def _potential(name, comp):
    with open(os.path.join(HERE, "golden", "g12_potentials.json")) as fh:
        terms = json.load(fh)["potentials"][name][comp]
    return [t["coefficient"] for t in terms], [t["exponent"] for t in terms]


GEMINALS = {
    "VHH-CCSDT": lambda: _potential("VHH-CCSDT", "O-X0.5+Y0.5+"),
    "synthetic-3": lambda: ([1.3, -0.4, 0.25], [0.0, 0.6, 3.0]),
}

INTRA = {
    "water-6-31G**": lambda: S.water("6-31G**").species[0],
    "FHF-proton-spd": lambda: S.fhf_anion().species[1],
    "water-dimer-6-31G**": lambda: S.water_cluster(2, np.random.default_rng(11)).species[0],
}


@pytest.mark.parametrize("gname", list(GEMINALS))
@pytest.mark.parametrize("name", list(INTRA))
def test_g12_direct_matches_oracle(name, gname):
    LI = _li()
    sp = INTRA[name]()
    cs, gs = GEMINALS[gname]()
    rng = np.random.default_rng(4)
    D = _density(sp.nbasis, max(1, sp.n_occupied), rng, sp.eta)
    h = LI.Libint2Interface(sp, 1)
    b = O.Basis(sp)
    for factor in (0.0, -1.0):
        _case("g12_direct: %s (n=%d), %s (%d terms), factor %g" % (name, sp.nbasis, gname, len(cs), factor))
        G = h.compute_g12_intraspecies_direct(D, factor, cs, gs)
        Gr = O.compute_g12_direct(b, D, factor, cs, gs)
        err = np.max(np.abs(G - Gr))
        print("   max|dG| = %.3e" % err)
        assert err <= TOL


def test_g12_constant_geminal_is_overlap_identity():
    """With the single term c = 1, g = 0 the operator is 1: J = S tr(S D)
    and K = S D S, S the GPU's own overlap matrix (no oracle involved)."""
    LI = _li()
    sp = S.water("6-31G**").species[0]
    n = sp.nbasis
    rng = np.random.default_rng(8)
    D = _density(n, 5, rng)
    h = LI.Libint2Interface(sp, 1)
    Smat = h.one_body(1)
    _case("g12 identity: water 6-31G** (n=%d), c=1, g=0" % n)
    for factor in (0.0, -0.5):
        G = h.compute_g12_intraspecies_direct(D, factor, [1.0], [0.0])
        want = Smat * np.sum(Smat * D) + factor * Smat @ D @ Smat
        err = np.max(np.abs(G - want))
        print("   factor %g: max|G - (S tr(SD) + f SDS)| = %.3e" % (factor, err))
        assert err < 1e-10


@pytest.mark.parametrize("ebasis", ["6-31G", "6-311++G**"])
def test_g12_coupling_matches_oracle(ebasis):
    LI = _li()
    f = S.fhf_anion(ebasis)
    e, p = f.species
    rng = np.random.default_rng(6)
    De = _density(e.nbasis, 10, rng)
    Dp = _density(p.nbasis, 1, rng, eta=1.0)
    he, hp = LI.Libint2Interface(e, 1), LI.Libint2Interface(p, 2)
    be, bp = O.Basis(e), O.Basis(p)
    cs, gs = _potential("VOH-CCSDT", "O-X0.5+")
    _case("g12 coupling: FHF- %s / proton (n_e=%d, n_p=%d), VOH-CCSDT (%d terms)"
          % (ebasis, e.nbasis, p.nbasis, len(cs)))
    Be = he.compute_g12_interspecies_direct(hp, Dp, cs, gs)
    Ber = O.compute_g12inter_direct(be, bp, 1, 2, Dp, cs, gs)
    Bp = hp.compute_g12_interspecies_direct(he, De, cs, gs)
    Bpr = O.compute_g12inter_direct(bp, be, 2, 1, De, cs, gs)
    for got, ref in ((Be, Ber), (Bp, Bpr)):
        err = np.max(np.abs(got.ravel(order="F") - ref.ravel()))
        print("   max|dB| = %.3e" % err)
        assert err <= TOL


def test_g12_device_entry_and_errors():
    LI = _li()
    from openlowdin_amd import hip as H
    sp = S.water("6-31G**").species[0]
    n = sp.nbasis
    rng = np.random.default_rng(9)
    D = _density(n, 5, rng)
    cs, gs = GEMINALS["synthetic-3"]()
    h = LI.Libint2Interface(sp, 1)
    _case("g12 device entry: water 6-31G** (n=%d)" % n)
    G = h.compute_g12_intraspecies_direct(D, -0.5, cs, gs)
    dD = H.DeviceArray.from_host(np.ascontiguousarray(D))
    dG = H.DeviceArray((n * n,))
    h.g12_device(dD.ptr, dG.ptr, cs, gs, factor=-0.5)
    assert np.max(np.abs(dG.download().reshape(n, n) - G)) < 1e-13
    # a negative exponent is refused (host check), result left as NaN
    with pytest.raises(LI.LowdinFockError):
        h.compute_g12_intraspecies_direct(D, -0.5, [1.0], [-1.0])
