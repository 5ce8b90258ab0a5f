"""Multi-species Hartree-Fock driver around the CPU oracle.

TEST INFRASTRUCTURE ONLY (see oracle.py).  It exists so that the oracle's
Fock builds can be pinned against the total energies that openLOWDIN's own
tests hold, at their 1e-8 Eh tolerance.

It follows the reference's SCF contract (MultiSCF.f90:276-440, 730-1002;
WaveFunction.f90:817-1159, 1163-1535, 1768-2003):
  F_i = H_i + G_i + C_i,   H_i = T_i / m_i - q_i sum_C Z_C/|r - C| ...
  G_i = q_i^2 (J[D_i] + (kappa_i/eta_i) K[D_i])   (only if N_i > 1)
  C_i = sum_{j != i} q_i q_j J_{i<-j}[D_j]         (upper triangle, symmetrised)
  D_i = eta_i C_occ C_occ^T
  E   = E_point_charges + sum_i Tr(D_i (H_i + G_i/2)) + sum_{i<j} Tr(D_i J_{i<-j})
The iteration itself (DIIS on F D S - S D F over all species) differs from the
reference's damping; only the converged stationary point is compared.
"""
from __future__ import annotations

import numpy as np

import oracle as O


def _sym_orth(S: np.ndarray, thresh: float = 1e-8) -> np.ndarray:
    w, U = np.linalg.eigh(S)
    keep = w >= thresh
    return U[:, keep] @ np.diag(w[keep] ** -0.5) @ U[:, keep].T if keep.all() else U[:, keep] / np.sqrt(w[keep])


class SpeciesState:
    def __init__(self, sid, species, charges, overlap_threshold: float = 1e-8):
        self.sid = sid
        self.sp = species
        self.basis = O.Basis(species)
        S = O.one_body(self.basis, 1)
        T = O.one_body(self.basis, 2)
        V = O.one_body(self.basis, 3, charges)  # sum_C -Z_C/|r-C|
        self.S = S
        self.H = T / species.mass - species.charge * V
        self.X = _sym_orth(S, overlap_threshold)
        self.pairs, _ = O.shellpair_list(self.basis)
        self.Q = O.schwarz(self.basis)
        self.D = None
        self.G = np.zeros_like(S)
        self.hartree = {}
        self.nocc = int(round(species.n_particles / species.eta))

    def density_from_fock(self, F):
        Fp = self.X.T @ F @ self.X
        w, Cp = np.linalg.eigh(Fp)
        Cmo = self.X @ Cp
        occ = Cmo[:, : self.nocc]
        return self.sp.eta * occ @ occ.T, w


def build_two_particles(st: SpeciesState, build_for_one: bool = False, factor_in: float = 1.0):
    """WaveFunction_buildTwoParticlesMatrix, direct branch."""
    sp = st.sp
    if sp.n_particles > 1 or build_for_one:
        G = O.compute_2body_direct(st.basis, st.D, sp.exchange_factor * factor_in, st.pairs, st.Q)
        return G * sp.charge ** 2
    return np.zeros_like(st.S)


def build_coupling(states, i):
    """WaveFunction_buildCouplingMatrix, direct branch: returns (C_i, {j: J_ij})."""
    st = states[i]
    Ctot = np.zeros_like(st.S)
    hart = {}
    for j, other in enumerate(states):
        if j == i:
            continue
        B = O.compute_coupling_direct(st.basis, other.basis, st.sid, other.sid, other.D)
        B = B * st.sp.charge * other.sp.charge
        # the row-major image has the lower triangle; Fortran symmetrises
        # upper -> lower of its column-major view, i.e. copies this lower
        # triangle into the upper one.
        L = np.tril(B)
        Bs = L + np.tril(B, -1).T
        hart[j] = Bs
        Ctot += Bs
    return Ctot, hart


def damping_step(Finit, Dinit, Fnew, Dnew, zero=1e-12):
    """Convergence_dampingMethod (Convergence.f90:444-515): the damping
    factor a = dE_D / dE_FD from -1/2 Tr(F0 dD) and Tr(dF dD); returns the
    Fock matrix to diagonalise and the new (Finit, Dinit)."""
    dD = Dnew - Dinit
    de = -0.5 * float(np.sum(Finit * dD.T))
    fde = float(np.sum((Fnew - Finit) * dD.T))
    if fde <= de or abs(fde) < zero:
        return Fnew, Fnew.copy(), Dnew.copy()
    a = de / fde
    F0 = Finit + a * (Fnew - Finit)
    return F0, F0.copy(), Dinit + a * dD


def run_scf(system, max_iter: int = 300, e_tol: float = 1e-11, d_tol: float = 1e-9,
            build_for_one: bool = False, verbose: bool = False, guess=None, history=None,
            convergence: str = "diis", overlap_threshold: float = 1e-8):
    """guess: optional list of starting densities (None entries: H-core);
    history: optional list that receives the energy of every iteration.
    convergence: "diis" (DIIS(8) on X^T (F D S - S D F) X, converged when
    |dE| < e_tol and the error < d_tol) or "damping" (the reference's
    default: CONVERGENCE_METHOD=1, ITERATION_SCHEME=3, one damped update of
    every species per global iteration, converged when |dE| < e_tol, the
    ENERGY criterion; MultiSCF.f90:276-440, 795-830).  overlap_threshold:
    overlap eigenvalues kept by the canonical orthogonaliser
    (OVERLAP_EIGEN_THRESHOLD, default 1e-8, WaveFunction.f90:419)."""
    states = [SpeciesState(k + 1, sp, system.charges, overlap_threshold) for k, sp in enumerate(system.species)]
    for k, st in enumerate(states):
        if guess is not None and guess[k] is not None:
            st.D = np.array(guess[k], dtype=np.float64)
        else:
            st.D, _ = st.density_from_fock(st.H)
    e_pc = system.point_charge_repulsion()
    diis_F, diis_e = [], []
    e_old = None
    energy = None
    for it in range(max_iter):
        Fs, comps = [], {}
        for i, st in enumerate(states):
            st.G = build_two_particles(st, build_for_one)
            C, st.hartree = build_coupling(states, i)
            Fs.append(st.H + st.G + C)
        energy = e_pc
        for i, st in enumerate(states):
            energy += np.sum(st.D * (st.H + 0.5 * st.G))
        for i, st in enumerate(states):
            for j in range(i + 1, len(states)):
                energy += np.sum(st.D * st.hartree[j])
        errs = [st.X.T @ (F @ st.D @ st.S - st.S @ st.D @ F) @ st.X for st, F in zip(states, Fs)]
        err = max(np.abs(e).max() for e in errs)
        if verbose:
            print(it, energy, err)
        if history is not None:
            history.append(energy)
        if convergence == "damping":
            if e_old is not None and abs(energy - e_old) < e_tol:
                break
            e_old = energy
            if it == 0:
                damp = [(F.copy(), st.D.copy()) for st, F in zip(states, Fs)]
            for i, st in enumerate(states):
                Fuse, Fi, Di = damping_step(damp[i][0], damp[i][1], Fs[i], st.D)
                damp[i] = (Fi, Di)
                st.D, _ = st.density_from_fock(Fuse)
            continue
        if e_old is not None and abs(energy - e_old) < e_tol and err < d_tol:
            break
        e_old = energy
        diis_F.append(Fs)
        diis_e.append(np.concatenate([e.ravel() for e in errs]))
        if len(diis_F) > 8:
            diis_F.pop(0)
            diis_e.pop(0)
        m = len(diis_F)
        Bm = -np.ones((m + 1, m + 1))
        Bm[m, m] = 0.0
        for a in range(m):
            for b in range(m):
                Bm[a, b] = diis_e[a] @ diis_e[b]
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        try:
            cfs = np.linalg.solve(Bm, rhs)[:m]
        except np.linalg.LinAlgError:
            cfs = np.zeros(m)
            cfs[-1] = 1.0
        for i, st in enumerate(states):
            F = sum(c * diis_F[k][i] for k, c in enumerate(cfs))
            st.D, _ = st.density_from_fock(F)
    return energy, states
