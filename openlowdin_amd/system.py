"""Molecular system model: quantum species, their contracted Gaussian shells,
and point charges.

This mirrors the parts of openLOWDIN's data model that the two-particle Fock
build consumes:

* species constants kappa/eta/charge/mass and the exchange factor
  kappa/eta (``src/core/MolecularSystem.f90:1398-1413``,
  ``lib/dataBases/constantsOfCoupling.lib``);
* one basis set per particle, shells in particle order then file contraction
  order, raw (unnormalised) file coefficients
  (``src/ints/Libint2Interface.f90:412-446``);
* Angstrom to bohr by dividing by 0.52917724924
  (``src/core/Units.f90``, ``src/core/InputManager.f90:560-562``).

Geometry builders for the five benchmark workloads live at the bottom:
water (6-31G**), the [F-H-F]- anion with a quantum proton, linear all-trans
alkanes, random water clusters and the served job mix.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

ANGSTROM = 0.52917724924  # bohr -> Angstrom factor used by openLOWDIN (Units.f90)

_BASIS_LIB: Optional[dict] = None


def basis_library() -> dict:
    global _BASIS_LIB
    if _BASIS_LIB is None:
        path = os.path.join(os.path.dirname(__file__), "basis", "library.json")
        with open(path) as fh:
            _BASIS_LIB = json.load(fh)
    return _BASIS_LIB


# canonical basis-file names for the spellings used in inputs
_BASIS_ALIASES = {
    "6-31G**": "6-31G.D.P", "6-31G(D,P)": "6-31G.D.P", "6-31G.D.P": "6-31G.D.P",
    "6-31G*": "6-31G.D", "6-31G.D": "6-31G.D",
    "CC-PVDZ": "CC-PVDZ", "6-311++G**": "6-311++G.D.P", "6-311++G.D.P": "6-311++G.D.P",
    "6-311G": "6-311G", "6-31G": "6-31G", "NAKAI-5-SPD": "NAKAI-5-SPD",
    "NAKAI-3-S": "NAKAI-3-S", "NAKAI-3-SP": "NAKAI-3-SP", "DZSPNB": "DZSPNB",
    "GRIBAKIN-10S": "GRIBAKIN-10S", "NAKAI-7-SP": "NAKAI-7-SP", "SHARON-E-6S2P": "SHARON-E-6S2P",
    "SHARON-E+6S2P": "SHARON-E+6S2P", "DZSPDN": "DZSPDN", "AUG-CC-PVDZ": "AUG-CC-PVDZ",
    "E+AUG-CC-PVDZ": "E+AUG-CC-PVDZ", "STO-3G": "STO-3G", "1S1P": "1S1P", "PSX-DZ": "PSX-DZ",
}

ATOMIC_NUMBER = {"H": 1, "HE": 2, "LI": 3, "BE": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "NE": 10}


@dataclass
class Shell:
    """One contracted Cartesian shell: angular momentum l, raw primitives."""
    l: int
    exps: np.ndarray
    coefs: np.ndarray
    origin: np.ndarray  # bohr

    @property
    def nprim(self) -> int:
        return len(self.exps)

    @property
    def size(self) -> int:
        return (self.l + 1) * (self.l + 2) // 2


@dataclass
class Species:
    """A quantum species (electrons, a quantum nucleus, positrons, ...)."""
    name: str
    is_electron: bool
    kappa: float
    eta: float
    charge: float
    mass: float
    n_particles: int
    shells: List[Shell] = field(default_factory=list)

    @property
    def nbasis(self) -> int:
        return sum(s.size for s in self.shells)

    @property
    def exchange_factor(self) -> float:
        """kappa/eta (MolecularSystem_getFactorOfExchangeIntegrals)."""
        return self.kappa / self.eta

    @property
    def n_occupied(self) -> int:
        return int(round(self.n_particles / self.eta))

    def arrays(self):
        """Flat arrays (l, nprim, poff, exps, coefs, xyz) of the shells."""
        ls = np.array([s.l for s in self.shells], dtype=np.int32)
        nprim = np.array([s.nprim for s in self.shells], dtype=np.int32)
        poff = np.zeros(len(self.shells), dtype=np.int32)
        if len(self.shells) > 1:
            poff[1:] = np.cumsum(nprim)[:-1]
        exps = np.concatenate([s.exps for s in self.shells]).astype(np.float64)
        coefs = np.concatenate([s.coefs for s in self.shells]).astype(np.float64)
        xyz = np.array([s.origin for s in self.shells], dtype=np.float64).reshape(-1)
        return ls, nprim, poff, exps, coefs, xyz


@dataclass
class MolecularSystem:
    species: List[Species]
    charges: List[Tuple[float, np.ndarray]]  # point charges (Z, xyz bohr)

    def point_charge_repulsion(self) -> float:
        e = 0.0
        for i in range(len(self.charges)):
            for j in range(i):
                zi, ri = self.charges[i]
                zj, rj = self.charges[j]
                e += zi * zj / np.linalg.norm(ri - rj)
        return e


def basis_shells(basis: str, symbol: str, origin_bohr: Sequence[float]) -> List[Shell]:
    key = _BASIS_ALIASES.get(basis.upper(), basis.upper())
    entry = basis_library()[key][symbol.upper() if symbol.upper() in basis_library()[key] else symbol]
    o = np.asarray(origin_bohr, dtype=np.float64)
    return [Shell(sh["l"], np.array(sh["exps"]), np.array(sh["coefs"]), o.copy()) for sh in entry]


# species constants: (kappa, eta, charge, mass) -- constantsOfCoupling.lib,
# elementalParticles.lib
_ELECTRON = (-1.0, 2.0, -1.0, 1.0)
_ELECTRON_SPIN = (-1.0, 1.0, -1.0, 1.0)
_POSITRON = (-1.0, 1.0, 1.0, 1.0)
PROTON_MASS = 1836.15267247


def build_system(atoms: Sequence[Tuple[str, Sequence[float]]], e_basis: str,
                 quantum: Sequence[Tuple[str, str, Sequence[float], float]] = (),
                 extra_electrons: int = 0, open_shell: bool = False,
                 angstrom: bool = True, add_particles: Optional[Dict[str, int]] = None) -> MolecularSystem:
    """Build a system: electrons in ``e_basis`` on every atom of ``atoms``
    (element, xyz); point charges for every atom not listed as quantum;
    ``quantum`` entries are (species symbol, basis, xyz, mass) quantum nuclei,
    one particle each, or positron basis centres (symbol E+, one particle
    each; a nucleus keeps its point charge beside them); entries with the
    same symbol form one species.  ``add_particles`` adds to a quantum
    species' particle count, the addParticles keyword of a GEOMETRY line
    (Particle.f90:157; three E+ centres with addParticles=-2 hold one
    positron).

    With ``open_shell`` the electrons are split into E-ALPHA / E-BETA species
    (eta = 1), as openLOWDIN's UHF does."""
    f = 1.0 / ANGSTROM if angstrom else 1.0
    # a quantum nucleus replaces the point charge at its position; positrons
    # (E+) sit beside the nuclei
    qpos = [np.asarray(x, float) * f for (sym, _, x, _) in quantum if not sym.upper().startswith("E+")]
    e_shells: List[Shell] = []
    charges = []
    nelec = extra_electrons
    for el, x in atoms:
        xb = np.asarray(x, float) * f
        e_shells += basis_shells(e_basis, el, xb)
        nelec += ATOMIC_NUMBER[el.upper()]
        if not any(np.allclose(xb, qp) for qp in qpos):
            charges.append((float(ATOMIC_NUMBER[el.upper()]), xb))
    species: List[Species] = []
    if open_shell:
        na = (nelec + 1) // 2
        nb = nelec - na
        species.append(Species("E-ALPHA", True, *_ELECTRON_SPIN, na, list(e_shells)))
        species.append(Species("E-BETA", True, *_ELECTRON_SPIN, nb, [Shell(s.l, s.exps, s.coefs, s.origin) for s in e_shells]))
    else:
        species.append(Species("E-", True, *_ELECTRON, nelec, e_shells))
    by_symbol: Dict[str, Species] = {}
    for sym, bas, x, mass in quantum:
        xb = np.asarray(x, float) * f
        if sym.upper().startswith("E+"):
            kap, eta, q, m = _POSITRON
        else:
            kap, eta, q, m = -1.0, 1.0, 1.0, mass
        if sym not in by_symbol:
            by_symbol[sym] = Species(sym, False, kap, eta, q, m, 0, [])
            species.append(by_symbol[sym])
        sp = by_symbol[sym]
        sp.n_particles += 1
        sp.shells += basis_shells(bas, sym, xb)
    for sym, extra in (add_particles or {}).items():
        if sym not in by_symbol or by_symbol[sym].n_particles + extra < 1:
            raise ValueError("add_particles: no quantum species %r or fewer than one particle" % sym)
        by_symbol[sym].n_particles += extra
    return MolecularSystem(species, charges)


def custom_species(name: str, shells: List[Shell], n_particles: int = 1, charge: float = 1.0,
                   mass: float = PROTON_MASS, is_electron: bool = False) -> Species:
    return Species(name, is_electron, -1.0, 1.0, charge, mass, n_particles, shells)


# ---------------------------------------------------------------- workloads
WATER_ANGSTROM = [("O", (0.0, 0.0, 0.0)), ("H", (0.7572, 0.5865, 0.0)), ("H", (-0.7572, 0.5865, 0.0))]


def water(basis: str = "6-31G**") -> MolecularSystem:
    """H2O, O at the origin, H at (+/-0.7572, 0.5865, 0) Angstrom."""
    return build_system(WATER_ANGSTROM, basis)


def even_tempered_proton_shells(origin_bohr, lmax: int = 2, nexp: int = 4,
                                a0: float = 4.0, ratio: float = 2.5) -> List[Shell]:
    """Even-tempered s,p,d proton basis: exponents a0*ratio^k, coefficient 1."""
    out = []
    for l in range(lmax + 1):
        for k in range(nexp):
            out.append(Shell(l, np.array([a0 * ratio ** k]), np.array([1.0]), np.asarray(origin_bohr, float)))
    return out


def fhf_anion(e_basis: str = "6-311++G**") -> MolecularSystem:
    """[F-H-F]-: F at (0,0,+/-1.14) Angstrom carrying the electron basis
    ``e_basis`` (20 electrons: 9 + 9 + the proton's 1 + the extra 1) and F
    point charges; one quantum proton ``H_1`` at the origin with the
    even-tempered s,p,d basis (exponents 4*2.5^k, k = 0..3)."""
    atoms = [("F", (0.0, 0.0, -1.14)), ("F", (0.0, 0.0, 1.14))]
    sysm = build_system(atoms, e_basis, quantum=[], extra_electrons=2)
    sysm.species.append(custom_species("H_1", even_tempered_proton_shells(np.zeros(3))))
    return sysm


def alkane_geometry(ncarbon: int, cc: float = 1.54, ch: float = 1.09,
                    theta_deg: float = 109.4712206) -> List[Tuple[str, Tuple[float, float, float]]]:
    """All-trans linear alkane CnH2n+2 (Angstrom).  Carbon k at
    (k d sin(t/2), 0, (k mod 2) d cos(t/2)).  Every carbon carries two
    hydrogens out of the C-C-C plane (+/-y), bisecting, at tetrahedral
    angles; a terminal carbon also carries one in-plane hydrogen where the
    chain would continue (so CH3 ends)."""
    th = math.radians(theta_deg)

    def cpos(k):
        return np.array([k * cc * math.sin(th / 2), 0.0, (k % 2) * cc * math.cos(th / 2)])

    def unit(v):
        return v / np.linalg.norm(v)

    atoms = [("C", tuple(cpos(k))) for k in range(ncarbon)]
    yax = np.array([0.0, 1.0, 0.0])
    for k in range(ncarbon):
        c = cpos(k)
        b = unit(unit(cpos(k - 1) - c) + unit(cpos(k + 1) - c))
        for sgn in (1.0, -1.0):
            d = math.cos(th / 2) * (-b) + sgn * math.sin(th / 2) * yax
            atoms.append(("H", tuple(c + ch * unit(d))))
        for j in (k - 1, k + 1):
            if j < 0 or j >= ncarbon:
                atoms.append(("H", tuple(c + ch * unit(cpos(j) - c))))
    return atoms


def alkane(ncarbon: int = 20, basis: str = "cc-pVDZ") -> MolecularSystem:
    return build_system(alkane_geometry(ncarbon), basis)


def _random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def water_cluster_geometry(nmol: int, rng, min_oo: float = 2.7):
    """(H2O)_n: rigid monomers, Haar-random rotation, uniform translation in a
    cube of edge 3.1 n^(1/3) Angstrom, rejecting O-O closer than 2.7 A."""
    edge = 3.1 * nmol ** (1.0 / 3.0)
    mono = np.array([x for _, x in WATER_ANGSTROM])
    placed = []
    atoms = []
    while len(placed) < nmol:
        R = _random_rotation(rng)
        t = rng.uniform(0.0, edge, size=3)
        if any(np.linalg.norm(t - o) < min_oo for o in placed):
            continue
        placed.append(t)
        xyz = mono @ R.T + t
        atoms += [("O", tuple(xyz[0])), ("H", tuple(xyz[1])), ("H", tuple(xyz[2]))]
    return atoms


def water_cluster(nmol: int, rng, basis: str = "6-31G**") -> MolecularSystem:
    return build_system(water_cluster_geometry(nmol, rng), basis)


BATCH_SEED = 20261202


def water_cluster_batch(nmolecules: int = 256, seed: int = BATCH_SEED, basis: str = "6-31G**"):
    """The batched workload: ``nmolecules`` clusters, n ~ U{1..8} each."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, 9, size=nmolecules)
    return [water_cluster(int(n), rng, basis) for n in sizes]


def load_water_batch(path: str, basis: str = "6-31G**"):
    """The batched workload from its committed geometry fixture (JSON,
    Angstrom): one MolecularSystem per cluster."""
    import json
    with open(path) as fh:
        data = json.load(fh)
    return [build_system([(sym, tuple(xyz)) for sym, xyz in g], basis) for g in data["clusters"]]


def with_quantum_proton(atoms, e_basis: str, h_index: int) -> MolecularSystem:
    """``atoms`` with the nucleus of hydrogen ``h_index`` (index into atoms)
    made a quantum proton ``H_1`` carrying the even-tempered s,p,d basis of
    the [F-H-F]- workload; the electrons keep ``e_basis`` on every atom."""
    sysm = build_system(atoms, e_basis)
    xyz = np.asarray(atoms[h_index][1], float) / ANGSTROM
    sysm.charges = [(z, r) for z, r in sysm.charges if not np.allclose(r, xyz)]
    sysm.species.append(custom_species("H_1", even_tempered_proton_shells(xyz)))
    return sysm


def serving_job(rng, proton_share: float = 0.2):
    """One request of the served workload: (H2O)_n with n ~ U{1..24} placed
    as the batched clusters, or C_nH_2n+2 with n ~ U{2..20} built as the
    alkane workload (equal odds); 6-31G** or cc-pVDZ (equal odds); with
    probability ``proton_share`` one hydrogen (chosen uniformly) is a quantum
    proton.  Returns (label, MolecularSystem)."""
    basis = "6-31G**" if rng.random() < 0.5 else "cc-pVDZ"
    if rng.random() < 0.5:
        n = int(rng.integers(1, 25))
        atoms = water_cluster_geometry(n, rng)
        label = "(H2O)%d/%s" % (n, basis)
    else:
        n = int(rng.integers(2, 21))
        atoms = alkane_geometry(n)
        label = "C%dH%d/%s" % (n, 2 * n + 2, basis)
    if rng.random() < proton_share:
        hs = [k for k, (sym, _) in enumerate(atoms) if sym == "H"]
        k = hs[int(rng.integers(len(hs)))]
        return label + "+H_1", with_quantum_proton(atoms, basis, k)
    return label, build_system(atoms, basis)
