"""Species matrices of the SCF: the caller side of the direct Fock build.

Mirrors the direct branches of openLOWDIN's
``WaveFunction_buildTwoParticlesMatrix`` (src/scf/WaveFunction.f90:817-1159)
and ``WaveFunction_buildCouplingMatrix`` (:1163-1535), plus the Fock
assembly and energies (:1768-2003), over :mod:`libint2_interface`:

* factor = (kappa/eta) * exact-exchange fraction (:858-862);
* G is built only for species with more than one particle unless
  ``build_for_one_particle`` is set (:867; CONTROL default .false.);
* G <- G * q^2 (:1118-1119); coupling <- coupling * q_i q_j (:1458), upper
  triangle copied to the lower one (:1488-1498), summed over the other
  species except the ghost species (:1502-1503).

A small multi-species SCF driver (:func:`run_scf`) iterates them with DIIS
on the orbital gradient; the Fock builds run on the GPU, the dense linear
algebra of the orbital update on the host.
"""
from __future__ import annotations

from typing import Dict, List, Optional

import numpy as np

from . import libint2_interface as LI


class WaveFunction:
    def __init__(self, system, species_index: int, objs: Dict[int, LI.Libint2Interface]):
        self.system = system
        self.sid = species_index + 1           # 1-based as in the reference
        self.sp = system.species[species_index]
        self.objs = objs
        self.exact_exchange_fraction = 1.0
        self.density = None
        self.two_particles = None
        self.coupling = None
        self.hartree: Dict[int, np.ndarray] = {}

    # WaveFunction_buildTwoParticlesMatrix, direct branch
    def build_two_particles_matrix(self, density: Optional[np.ndarray] = None, factor_in: Optional[float] = None,
                                   build_for_one_particle: bool = False) -> np.ndarray:
        D = self.density if density is None else density
        n = self.sp.nbasis
        factor = self.sp.exchange_factor * (self.exact_exchange_fraction if factor_in is None else factor_in)
        G = np.zeros((n, n), order="F")
        if self.sp.n_particles > 1 or build_for_one_particle:
            G = LI.get_direct_intra_repulsion_matrix(self.sid, "LIBINT", D, factor, self.system, self.objs)
            G = G * self.sp.charge ** 2
        if density is None:
            self.two_particles = G
        return G

    # WaveFunction_buildCouplingMatrix, direct branch
    def build_coupling_matrix(self, densities: List[np.ndarray], ghost: str = "") -> np.ndarray:
        n = self.sp.nbasis
        C = np.zeros((n, n), order="F")
        self.hartree = {}
        for k, other in enumerate(self.system.species):
            osid = k + 1
            if osid == self.sid:
                continue
            aux = LI.get_direct_inter_repulsion_matrix(self.sid, osid, "LIBINT", densities[k], self.system, self.objs)
            aux = aux * self.sp.charge * other.charge
            h = np.triu(aux) + np.triu(aux, 1).T     # upper -> lower (column-major view)
            self.hartree[osid] = h
            if other.name != ghost:
                C = C + h
        self.coupling = C
        return C


def one_body_matrices(system, objs):
    """H-core and overlap of every species on the GPU:
    H_i = T_i / m_i - q_i sum_C Z_C (-1/|r - C|) over the point charges."""
    hcore, overlap = [], []
    for k, sp in enumerate(system.species):
        h = LI._instance(objs, system, k + 1)
        S = h.one_body(1)
        T = h.one_body(2)
        V = h.one_body(3, system.charges)
        hcore.append(T / sp.mass - sp.charge * V)
        overlap.append(S)
    return hcore, overlap


def _sym_orth(S, thresh=1e-8):
    w, U = np.linalg.eigh(S)
    keep = w >= thresh
    return U[:, keep] / np.sqrt(w[keep])


def damping_step(Finit, Dinit, Fnew, Dnew, zero: float = 1e-12):
    """One step of the reference's damping (Convergence_dampingMethod,
    Convergence.f90:444-515).  With dD = Dnew - Dinit, the density effect
    -1/2 Tr(Finit dD) and the Fock-and-density effect Tr((Fnew - Finit) dD)
    give the damping factor a = density / Fock-and-density; no damping when
    the latter is not above the former or is below DOUBLE_ZERO_THRESHOLD
    (1e-12).  Returns (Fock matrix to diagonalise, new Finit, new Dinit)."""
    dD = Dnew - Dinit
    de = -0.5 * float(np.sum(Finit * dD.T))
    fde = float(np.sum((Fnew - Finit) * dD.T))
    if fde <= de or abs(fde) < zero:
        return Fnew, Fnew.copy(), Dnew.copy()
    a = de / fde
    F0 = Finit + a * (Fnew - Finit)
    return F0, F0.copy(), Dinit + a * dD


def run_scf(system, hcore: Optional[List[np.ndarray]] = None, overlap: Optional[List[np.ndarray]] = None,
            max_iter: int = 300,
            e_tol: float = 1e-11, d_tol: float = 1e-9, objs=None, build_for_one_particle: bool = False,
            verbose: bool = False, guess=None, history=None, convergence: str = "diis", ghost: str = "",
            overlap_threshold: float = 1e-8):
    """Multi-species HF with GPU Fock builds.  hcore/overlap per species
    (normalised Cartesian basis; computed on the GPU when not given).
    guess: starting densities (None entries: H-core guess); history: list
    receiving every iteration's energy.  convergence: "diis" (DIIS(8) over
    all species, converged when |dE| < e_tol and the orthonormal FDS - SDF
    error < d_tol) or "damping" (the reference's default, CONVERGENCE_METHOD=1
    with ITERATION_SCHEME=3: one damped update of every species per global
    iteration, Convergence.f90:444-515, converged when |dE| < e_tol, the
    ENERGY criterion, MultiSCF.f90:795-830).  ghost: name of a species whose
    coupling is left out of the other species' Fock matrices but kept in the
    energy (SCF_GHOST_SPECIES, WaveFunction.f90:1502-1503, 1980-1986).
    overlap_threshold: overlap eigenvalues kept by the canonical
    orthogonaliser (OVERLAP_EIGEN_THRESHOLD, default 1e-8, WaveFunction.f90:
    419).  Returns (energy, wavefunctions)."""
    objs = {} if objs is None else objs
    if hcore is None or overlap is None:
        hcore, overlap = one_body_matrices(system, objs)
    wfs = [WaveFunction(system, k, objs) for k in range(len(system.species))]
    X = [_sym_orth(S, overlap_threshold) for S in overlap]

    def density(k, F):
        w, Cp = np.linalg.eigh(X[k].T @ F @ X[k])
        Cm = X[k] @ Cp
        occ = Cm[:, : wfs[k].sp.n_occupied]
        return wfs[k].sp.eta * occ @ occ.T

    for k, wf in enumerate(wfs):
        if guess is not None and guess[k] is not None:
            wf.density = np.array(guess[k], dtype=np.float64)   # read coefficients (wfn_io)
        else:
            wf.density = density(k, hcore[k])
    e_pc = system.point_charge_repulsion()
    hist_F, hist_e = [], []
    e_old, energy = None, None
    for it in range(max_iter):
        dens = [wf.density for wf in wfs]
        Fs = []
        for k, wf in enumerate(wfs):
            G = wf.build_two_particles_matrix(build_for_one_particle=build_for_one_particle)
            C = wf.build_coupling_matrix(dens, ghost=ghost)
            Fs.append(hcore[k] + G + C)
        energy = e_pc
        for k, wf in enumerate(wfs):
            energy += np.sum(wf.density * (hcore[k] + 0.5 * wf.two_particles))
            for j in range(k + 2, len(wfs) + 1):
                energy += np.sum(wf.density * wf.hartree[j])
        errs = [X[k].T @ (F @ wfs[k].density @ overlap[k] - overlap[k] @ wfs[k].density @ F) @ X[k]
                for k, F in enumerate(Fs)]
        err = max(float(np.abs(e).max()) for e in errs)
        if verbose:
            print(it, energy, err)
        if history is not None:
            history.append(energy)
        if convergence == "damping":
            if e_old is not None and abs(energy - e_old) < e_tol:
                break
            e_old = energy
            if it == 0:
                damp = [(F.copy(), wf.density.copy()) for F, wf in zip(Fs, wfs)]
            for k, wf in enumerate(wfs):
                Fuse, Fi, Di = damping_step(damp[k][0], damp[k][1], Fs[k], wf.density)
                damp[k] = (Fi, Di)
                wf.density = density(k, Fuse)
            continue
        if convergence != "diis":
            raise ValueError("convergence must be 'diis' or 'damping'")
        if e_old is not None and abs(energy - e_old) < e_tol and err < d_tol:
            break
        e_old = energy
        hist_F.append(Fs)
        hist_e.append(np.concatenate([e.ravel() for e in errs]))
        if len(hist_F) > 8:
            hist_F.pop(0)
            hist_e.pop(0)
        m = len(hist_F)
        B = -np.ones((m + 1, m + 1))
        B[m, m] = 0.0
        for a in range(m):
            for b in range(m):
                B[a, b] = hist_e[a] @ hist_e[b]
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        try:
            cf = np.linalg.solve(B, rhs)[:m]
        except np.linalg.LinAlgError:
            cf = np.zeros(m)
            cf[-1] = 1.0
        for k, wf in enumerate(wfs):
            wf.density = density(k, sum(c * hist_F[i][k] for i, c in enumerate(cf)))
    return energy, wfs
