"""GPU integral store and first-quarter transforms (the MEMORY and
integral-transformation variants of the two-particle path) through the C
ABI, against the oracle.

Bars: store elements within 1e-12 absolute of the oracle, except that an
integral whose raw value lies within rounding (1e-9 relative) of the 1e-10
drop may be kept on one side and dropped on the other;
elements the store does not write keep the caller's values; transforms
within 1e-11 absolute.  Every test prints its case before the GPU call.
"""
import numpy as np
import pytest

import oracle as O
from openlowdin_amd import system as S

pytestmark = pytest.mark.gpu


def _li():
    from openlowdin_amd import libint2_interface as LI
    return LI


def _density(n, nocc, rng, eta=2.0):
    C = rng.normal(size=(n, nocc)) / np.sqrt(n)
    return eta * C @ C.T


def _case(msg):
    print("\n[case] " + msg, flush=True)


def _store_close(got, ref, init, nprod):
    """Store parity.  An element one side kept and the other left at the
    caller's value ``init`` is a drop-decision mismatch; it is allowed only
    where the kept integral's raw value (element / product of the four
    functions' norma, ``nprod``) lies within 1e-9 relative of the 1e-10
    drop threshold, i.e. within rounding of the cut.  Every other element
    agrees to 1e-12."""
    got, ref = got.reshape(-1), ref.reshape(-1)
    nprod = np.asarray(nprod, dtype=np.float64).reshape(-1)
    assert nprod.shape == got.shape
    mism = (got == init) != (ref == init)
    kept = np.where(got == init, ref, got)
    raw = np.abs(kept[mism] / nprod[mism])
    near = np.abs(raw - 1e-10) <= 1e-9 * 1e-10
    d = np.abs(got - ref)
    d[mism] = 0.0
    print("   max|dA| = %.3e over agreeing elements; drop-decision mismatches %d (near the cut %d)"
          % (d.max(), int(mism.sum()), int(near.sum())))
    assert near.all()
    assert d.max() <= 1e-12


def _nprod4(n1, n2, n3, n4):
    return np.einsum("i,j,k,l->ijkl", n1, n2, n3, n4)


INTRA = {
    "water-6-31G**": lambda: S.water("6-31G**").species[0],
    "water-dimer-6-31G**": lambda: S.water_cluster(2, np.random.default_rng(11)).species[0],
    "FHF-proton-spd": lambda: S.fhf_anion().species[1],
}


@pytest.mark.parametrize("name", list(INTRA))
def test_store_all_matches_oracle(name):
    LI = _li()
    sp = INTRA[name]()
    rng = np.random.default_rng(3)
    D = _density(sp.nbasis, max(1, sp.n_occupied), rng, sp.eta)
    n = sp.nbasis
    _case("2body_directAll: %s (n=%d, %d elements)" % (name, n, n ** 4))
    h = LI.Libint2Interface(sp, 1)
    h.init_2body_ints()
    init = np.full(n ** 4, 3.0)
    A = h.compute_2body_intraspecies_direct_all(D, init)
    b = O.Basis(sp)
    Ar = O.compute_2body_directAll(b, D, A=init)
    nb = O.norma(b)
    _store_close(A, Ar, 3.0, _nprod4(nb, nb, nb, nb))  # unwritten elements kept the caller's 3.0
    st = h.stats()
    assert st["quartets"] > 0 and st["kernel_ms"] > 0


@pytest.mark.parametrize("name", list(INTRA))
def test_first_quarter_transform_matches_oracle(name):
    LI = _li()
    sp = INTRA[name]()
    n = sp.nbasis
    rng = np.random.default_rng(5)
    D = _density(n, max(1, sp.n_occupied), rng, sp.eta)
    Cm = rng.normal(size=(n, n))
    h = LI.Libint2Interface(sp, 1)
    h.init_2body_ints()
    b = O.Basis(sp)
    A = O.compute_2body_directAll(b, D)
    for p in sorted({0, n // 2, n - 1}):
        _case("2body_directIT: %s (n=%d) p=%d" % (name, n, p))
        GG = h.compute_2body_intraspecies_direct_IT(D, Cm, p)
        ref = np.tensordot(Cm[p, :], A, axes=(0, 0))
        err = np.max(np.abs(GG - ref))
        print("   max|dGG| = %.3e" % err)
        assert err < 1e-11
        assert np.array_equal(GG, GG.transpose(0, 2, 1))


def test_first_quarter_transform_larger_cluster():
    """Four water molecules (n=100): every class up to (dd|dd) on many
    centres, transform against the oracle's."""
    LI = _li()
    sp = S.water_cluster(4, np.random.default_rng(2)).species[0]
    n = sp.nbasis
    rng = np.random.default_rng(9)
    D = _density(n, sp.n_occupied, rng)
    Cm = rng.normal(size=(n, n)) / np.sqrt(n)
    _case("2body_directIT: water tetramer (n=%d) p=3" % n)
    h = LI.Libint2Interface(sp, 1)
    h.init_2body_ints()
    GG = h.compute_2body_intraspecies_direct_IT(D, Cm, 3)
    ref = O.compute_2body_directIT(O.Basis(sp), D, Cm, 3)
    err = np.max(np.abs(GG - ref))
    print("   max|dGG| = %.3e" % err)
    assert err < 1e-11


@pytest.mark.parametrize("ebasis", ["6-31G", "6-311++G**"])
def test_coupling_store_and_transform(ebasis):
    LI = _li()
    f = S.fhf_anion(ebasis)
    e, p = f.species
    rng = np.random.default_rng(7)
    he, hp = LI.Libint2Interface(e, 1), LI.Libint2Interface(p, 2)
    be, bp = O.Basis(e), O.Basis(p)
    De = _density(e.nbasis, 10, rng)
    _case("coupling_directAll: FHF- %s (n_e=%d, n_p=%d)" % (ebasis, e.nbasis, p.nbasis))
    init = np.full(e.nbasis ** 2 * p.nbasis ** 2, -2.0)
    T = he.compute_2body_interspecies_direct_all(hp, De, init)
    Tr = init.reshape(T.shape).copy()
    Tt = O.compute_coupling_directAll(be, bp, 1, 2)
    keep = Tt != 0.0
    Tr[keep] = Tt[keep]
    ne, npn = O.norma(be), O.norma(bp)
    _store_close(T, Tr, -2.0, _nprod4(ne, ne, npn, npn).reshape(T.shape))
    for this, other, hthis, hother, it, io in ((be, bp, he, hp, 1, 2), (bp, be, hp, he, 2, 1)):
        n = this.n
        Cm = rng.normal(size=(n, n))
        for q in sorted({0, n - 1}):
            _case("coupling_directIT: FHF- %s, transformed species %d, p=%d" % (ebasis, it, q))
            GG = hthis.compute_2body_interspecies_direct_IT(hother, np.eye(n), Cm, q)
            ref = O.compute_coupling_directIT(this, other, it, io, Cm, q)
            err = np.max(np.abs(GG - ref))
            print("   max|dGG| = %.3e" % err)
            assert err < 1e-11


def test_coupling_store_refuses_permuted_different_sizes():
    """From the higher species id with n_lint != n_olint the reference's
    indices leave the result array: the call fails and the result is left
    as it was."""
    LI = _li()
    e, p = S.fhf_anion("6-31G").species
    he, hp = LI.Libint2Interface(e, 1), LI.Libint2Interface(p, 2)
    _case("coupling_directAll refused: permuted, n_p=%d != n_e=%d" % (p.nbasis, e.nbasis))
    init = np.full(e.nbasis ** 2 * p.nbasis ** 2, 5.0)
    with pytest.raises(LI.LowdinFockError):
        hp.compute_2body_interspecies_direct_all(he, np.eye(p.nbasis), init)


def test_coupling_store_permuted_equal_sizes():
    """From the higher species id with equal basis sizes (two species on one
    basis, as alpha and beta electrons) the reference's strides fit: the
    store is indexed with x of the lower species id."""
    LI = _li()
    sp = S.water("6-31G").species[0]
    n = sp.nbasis
    ha, hb = LI.Libint2Interface(sp, 1), LI.Libint2Interface(sp, 2)
    b = O.Basis(sp)
    _case("coupling_directAll permuted, equal sizes: water 6-31G (n=%d) as species 1 and 2" % n)
    init = np.full(n ** 4, 1.5)
    T = hb.compute_2body_interspecies_direct_all(ha, np.eye(n), init)
    Tr = init.reshape(n, n, n, n).copy()
    Tt = O.compute_coupling_directAll(b, b, 2, 1)
    keep = Tt != 0.0
    Tr[keep] = Tt[keep]
    nb = O.norma(b)
    _store_close(T, Tr, 1.5, _nprod4(nb, nb, nb, nb))
    # one species, every pair: the tensor is the full ERI tensor, eight-fold symmetric where kept
    assert np.max(np.abs(T - T.transpose(2, 3, 0, 1))[keep & keep.transpose(2, 3, 0, 1)]) < 1e-12


def test_store_device_entry_matches_host_entry():
    LI = _li()
    from openlowdin_amd import hip as H
    sp = S.water("6-31G**").species[0]
    n = sp.nbasis
    rng = np.random.default_rng(12)
    D = _density(n, 5, rng)
    Cm = rng.normal(size=(n, n))
    h = LI.Libint2Interface(sp, 1)
    h.init_2body_ints()
    _case("store_device: water 6-31G** (n=%d), IT p=2 and All" % n)
    host_it = h.compute_2body_intraspecies_direct_IT(D, Cm, 2)
    host_all = h.compute_2body_intraspecies_direct_all(D)
    dD = H.DeviceArray.from_host(np.ascontiguousarray(D))
    dc = H.DeviceArray.from_host(np.ascontiguousarray(Cm[2, :]))
    dGG = H.DeviceArray((n ** 3,))
    h.store_device(1, dGG.ptr, dD.ptr, dc.ptr)
    assert np.max(np.abs(dGG.download().reshape(n, n, n) - host_it)) < 1e-13
    dA = H.DeviceArray.from_host(np.zeros(n ** 4))
    h.store_device(0, dA.ptr, dD.ptr)
    assert np.max(np.abs(dA.download().reshape(n, n, n, n) - host_all)) == 0.0
