"""CPU checks of the boundary: liblowdin_fock.so loads and exports every
entry point that include/lowdin_fock.h declares (no compute calls without a
GPU), and the Python mirror's argument checks."""
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(REPO, "include", "lowdin_fock.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)) -
                  {"if", "defined", "sizeof"})


def test_library_exports_declared_symbols():
    import ctypes
    from openlowdin_amd import libint2_interface as LI
    if not os.path.exists(LI.LIB_PATH):
        pytest.fail("liblowdin_fock.so is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LI.LIB_PATH)
    names = _declared()
    assert "LibintInterface_compute_2body_direct" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_reference_symbol_names_present():
    """The hot-path names bound by Libint2Interface.f90:165-373."""
    names = _declared()
    for n in ["LibintInterface_new", "LibintInterface_del", "LibintInterface_add_particle",
              "LibintInterface_add_shell", "LibintInterface_init_2body_ints",
              "LibintInterface_compute_2body_direct", "LibintInterface_compute_coupling_direct"]:
        assert n in names


def test_scheme_other_than_libint_is_rejected():
    from openlowdin_amd import libint2_interface as LI
    with pytest.raises(ValueError):
        LI.get_direct_intra_repulsion_matrix(1, "RYS", np.eye(2), -0.5, None, {})
    with pytest.raises(ValueError):
        LI.get_direct_inter_repulsion_matrix(1, 2, "CUDINT", np.eye(2), None, {})


def test_generated_kernels_cover_all_classes():
    import json
    t = json.load(open(os.path.join(REPO, "openlowdin_amd", "csrc", "ops_table.json")))["classes"]
    assert len(t) == 21
    for k in t:
        assert os.path.exists(os.path.join(REPO, "openlowdin_amd", "csrc", "gen", "k_%s.hip" % k))


def test_store_and_g12_argument_errors_before_device_work():
    """Argument errors of the transform, inter-species store and G12 entries
    are found before the device is touched: the call records an error, a
    transform's or build's result is NaN, a store's result is untouched."""
    import ctypes
    from openlowdin_amd import libint2_interface as LI
    from openlowdin_amd import system as S
    sp = S.water("6-31G**").species[0]
    h = LI.Libint2Interface(sp, 1)
    n = h.nbasis
    with pytest.raises(LI.LowdinFockError, match="orbital index p out of range"):
        h.compute_2body_intraspecies_direct_IT(np.eye(n), np.eye(n), n)
    with pytest.raises(LI.LowdinFockError, match="orbital index p out of range"):
        h.compute_2body_intraspecies_direct_IT(np.eye(n), np.eye(n), -1)
    e, p = S.fhf_anion("6-31G").species
    he, hp = LI.Libint2Interface(e, 1), LI.Libint2Interface(p, 2)
    init = np.full(e.nbasis ** 2 * p.nbasis ** 2, 5.0)
    with pytest.raises(LI.LowdinFockError, match="different basis size"):
        hp.compute_2body_interspecies_direct_all(he, np.eye(p.nbasis), init)
    # G12: more geminal terms than the kernels take, or a negative exponent
    L = LI.lib()
    G = np.zeros((n, n), order="F")
    f = ctypes.c_double(-0.5)
    for cs, gs in (([1.0] * 65, [0.5] * 65), ([1.0], [-1.0])):
        c = np.array(cs)
        g = np.array(gs)
        D = np.asfortranarray(np.eye(n))
        L.libintinterface_compute_g12_direct(h.this, LI._dptr(D), LI._dptr(G), ctypes.byref(f), LI._dptr(c),
                                             LI._dptr(g), len(cs))
        assert L.lowdin_fock_last_error().decode()
        assert np.all(np.isnan(G))


REF_IFACE = "/root/reference/src/ints/Libint2Interface.f90"


@pytest.mark.skipif(not os.path.exists(REF_IFACE), reason="reference tree not present")
def test_library_exports_every_reference_binding():
    """Link-level drop-in: every symbol the reference's Fortran binds with
    bind(C, name=...) (Libint2Interface.f90:165-373) is exported by
    liblowdin_fock.so under that exact name, so openLOWDIN's Fortran links
    against this library unchanged."""
    import ctypes
    from openlowdin_amd import libint2_interface as LI
    text = open(REF_IFACE).read()
    names = sorted(set(re.findall(r'bind\s*\(\s*[cC]\s*,\s*name\s*=\s*"([^"]+)"', text)))
    assert len(names) >= 20 and "LibintInterface_compute_2body_direct" in names
    L = ctypes.CDLL(LI.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    declared = set(_declared())
    assert not [n for n in names if n not in declared]
