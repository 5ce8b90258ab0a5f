"""GPU parity of the direct Fock build and the coupling build against the
CPU oracle, through the C ABI of liblowdin_fock.so.

Bars (FP64 path): every Fock / coupling element within 1e-10 Eh absolute of
the oracle; shell-pair lists identical; Schwarz bounds within 1e-12
relative; the reference's num_ints_computed identical.  Every test prints
the case it runs before the GPU call.
"""
import zlib

import numpy as np
import pytest

import oracle as O
from openlowdin_amd import system as S

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _li():
    from openlowdin_amd import libint2_interface as LI
    return LI


def _density(n, nocc, rng, eta=2.0):
    C = rng.normal(size=(n, nocc)) / np.sqrt(n)
    return eta * C @ C.T


def _case(msg):
    print("\n[case] " + msg, flush=True)


SYSTEMS = {
    "water-6-31G**": lambda: S.water("6-31G**").species[0],
    "FHF-e-6-311++G**": lambda: S.fhf_anion().species[0],
    "FHF-proton-spd": lambda: S.fhf_anion().species[1],
    "butane-cc-pVDZ": lambda: S.alkane(4, "cc-pVDZ").species[0],
    "water-dimer-6-31G**": lambda: S.water_cluster(2, np.random.default_rng(11)).species[0],
}


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_pair_list_and_schwarz(name):
    LI = _li()
    sp = SYSTEMS[name]()
    _case("pair list + Schwarz: %s (n=%d)" % (name, sp.nbasis))
    h = LI.Libint2Interface(sp, 1)
    h.init_2body_ints()
    got = h.shellpair_list()
    b = O.Basis(sp)
    ref, norms = O.shellpair_list(b)
    assert got.shape == ref.shape and np.array_equal(got, ref)
    Qg = h.schwarz()
    Qr = O.schwarz(b)
    mask = np.zeros_like(Qr, dtype=bool)
    mask[ref[:, 0], ref[:, 1]] = True
    mask[ref[:, 1], ref[:, 0]] = True
    rel = np.abs(Qg[mask] - Qr[mask]) / np.maximum(Qr[mask], 1e-300)
    k = int(np.argmax(rel))
    print("   worst Q: rel %.3e at Q=%.3e" % (rel[k], Qr[mask][k]))
    assert rel.max() < 1e-12, rel.max()


@pytest.mark.parametrize("name", list(SYSTEMS))
@pytest.mark.parametrize("factor", [-0.5, -1.0, 0.0])
def test_2body_direct_matches_oracle(name, factor):
    LI = _li()
    sp = SYSTEMS[name]()
    rng = np.random.default_rng(zlib.crc32(("%s/%g" % (name, factor)).encode()))
    D = _density(sp.nbasis, max(1, sp.n_occupied), rng)
    _case("2body direct: %s factor %g (n=%d)" % (name, factor, sp.nbasis))
    h = LI.Libint2Interface(sp, 1)
    G = h.compute_2body_intraspecies_direct(D, factor)
    Gr, cnt = O.compute_2body_direct(O.Basis(sp), D, factor, return_counts=True)
    err = np.max(np.abs(G - Gr))
    st = h.stats()
    print("   max|dG| = %.3e  quartets %d/%d  eris %d/%d" % (err, st["quartets"], cnt["quartets"], st["eris"], cnt["eris"]))
    assert err <= TOL
    assert st["quartets"] == cnt["quartets"] and st["eris"] == cnt["eris"]
    assert np.array_equal(G, G.T)


def test_2body_repeat_is_stable():
    LI = _li()
    sp = S.water("6-31G**").species[0]
    rng = np.random.default_rng(3)
    D = _density(sp.nbasis, 5, rng)
    _case("2body repeat: water")
    h = LI.Libint2Interface(sp, 1)
    G1 = h.compute_2body_intraspecies_direct(D, -0.5)
    G2 = h.compute_2body_intraspecies_direct(D, -0.5)
    assert np.max(np.abs(G1 - G2)) < 1e-13


@pytest.mark.parametrize("ebasis", ["6-31G", "6-311++G**"])
def test_coupling_matches_oracle(ebasis):
    LI = _li()
    f = S.fhf_anion(ebasis)
    e, p = f.species
    rng = np.random.default_rng(5)
    De = _density(e.nbasis, 10, rng)
    Dp = _density(p.nbasis, 1, rng, eta=1.0)
    _case("coupling: FHF- %s / proton (n_e=%d, n_p=%d)" % (ebasis, e.nbasis, p.nbasis))
    he, hp = LI.Libint2Interface(e, 1), LI.Libint2Interface(p, 2)
    be, bp = O.Basis(e), O.Basis(p)
    # electrons (id 1) <- proton (id 2): non-permuted in the reference
    Be = he.compute_2body_interspecies_direct(hp, Dp)
    Ber = O.compute_coupling_direct(be, bp, 1, 2, Dp)
    # proton (id 2) <- electrons (id 1): permuted in the reference
    Bp = hp.compute_2body_interspecies_direct(he, De)
    Bpr = O.compute_coupling_direct(bp, be, 2, 1, De)
    for got, ref in ((Be, Ber), (Bp, Bpr)):
        img = got.ravel(order="F")
        err = np.max(np.abs(img - ref.ravel()))
        print("   max|dB| = %.3e" % err)
        assert err <= TOL
    # the fused two-direction call gives the same two matrices
    Ra, Rb = he.coupling_pair(hp, De, Dp)
    assert np.max(np.abs(Ra - Be)) <= 1e-12 and np.max(np.abs(Rb - Bp)) <= 1e-12


def test_linearity_large_alkane():
    """C20H42 / cc-pVDZ (n = 510), the full-size workload: G[D1 + D2] =
    G[D1] + G[D2] up to the eps-level screening, G symmetric, and the
    computed-quartet count never exceeds the unique quartet count."""
    LI = _li()
    sp = S.alkane(20, "cc-pVDZ").species[0]
    rng = np.random.default_rng(17)
    D1 = _density(sp.nbasis, 81, rng)
    D2 = _density(sp.nbasis, 81, rng)
    _case("linearity: C20H42 cc-pVDZ n=%d" % sp.nbasis)
    h = LI.Libint2Interface(sp, 1)
    G1 = h.compute_2body_intraspecies_direct(D1, -0.5)
    G2 = h.compute_2body_intraspecies_direct(D2, -0.5)
    G12 = h.compute_2body_intraspecies_direct(D1 + D2, -0.5)
    st = h.stats()
    print("   quartets %d eris %d kernel %.2f ms" % (st["quartets"], st["eris"], st["kernel_ms"]))
    assert np.array_equal(G12, G12.T)
    assert np.max(np.abs(G12 - G1 - G2)) < 1e-9
    npairs = len(h.shellpair_list())
    assert st["quartets"] <= npairs * (npairs + 1) // 2


def test_mixed_precision_error_and_fp64_identity():
    """Mixed precision: quartets with Dnorm*Q*Q < tau in FP32.  tau = 0 must
    reproduce the FP64 build exactly.  For tau in 1e-4..1e-7 the error of the
    two-electron energy 1/2 Tr(D G) and of G must shrink with tau (bounded by
    0.1 tau and 0.5 tau), stay below 1e-6 Eh from tau = 1e-6 down, and FP32
    must carry a sizeable share of the quartets (butane / cc-pVDZ)."""
    LI = _li()
    sp = S.alkane(4, "cc-pVDZ").species[0]
    rng = np.random.default_rng(29)
    D = _density(sp.nbasis, sp.n_occupied, rng)
    _case("mixed precision: butane cc-pVDZ (n=%d)" % sp.nbasis)
    h = LI.Libint2Interface(sp, 1)
    G64 = h.compute_2body_intraspecies_direct(D, -0.5)
    h.set_precision("mixed", 0.0)
    G0 = h.compute_2body_intraspecies_direct(D, -0.5)
    assert h.stats()["quartets_fp32"] == 0
    assert np.max(np.abs(G0 - G64)) < 1e-13
    rows = []
    for tau in (1e-4, 1e-5, 1e-6, 1e-7):
        h.set_precision("mixed", tau)
        Gm = h.compute_2body_intraspecies_direct(D, -0.5)
        st = h.stats()
        dE = 0.5 * float(np.sum(D * (Gm - G64)))
        dG = float(np.max(np.abs(Gm - G64)))
        print("   tau %.0e: fp32 quartets %d / %d, max|dG| %.2e, dE %.2e" %
              (tau, st["quartets_fp32"], st["quartets"], dG, dE), flush=True)
        rows.append((tau, st["quartets_fp32"], st["quartets"], dE, dG))
    for tau, q32, q, dE, dG in rows:
        assert q32 > 0
        assert abs(dE) < 0.1 * tau
        assert dG < 0.5 * tau
        if tau <= 1e-6:
            assert abs(dE) < 1e-6
    assert rows[0][1] > 0.3 * rows[0][2]
    h.set_precision("fp64")
