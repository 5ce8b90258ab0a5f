"""Systems of openLOWDIN's own tests whose HF energy this path reproduces,
with the energies their checkers hold (each at 1e-8 there).

Each entry: builder, reference HF energy, SCF options (build_for_one:
buildTwoParticlesMatrixForOneParticle; overlap_threshold:
overlapEigenThreshold), tolerance (1e-9; the checker's 1e-8 where the
reference's run stopped further from convergence).  Quantum protons carry openLOWDIN's isotope mass,
atomic weight x Dalton - Z electron masses (src/core/Particle.f90:345,
1.0078250321 for H_1).  Geometries in angstrom as in the inputs."""
from openlowdin_amd import system as S

H1_MASS = 1.0078250321 * 1822.88850065855 - 1.0


def _with_eta(sysm, name, eta):
    """The eta keyword of a GEOMETRY line: particles per orbital of a species."""
    next(sp for sp in sysm.species if sp.name == name).eta = eta
    return sysm
H2O = [("O", (0.0, 0.0, -0.066575)), ("H", (0.0, 0.754175, 0.528381)), ("H", (0.0, -0.754174, 0.528382))]

CASES = {
    # test/H2O.APMO.UP2.lowdin: UHF electrons 6-311G, two quantum protons
    # (Nakai-7-sp) as separate species
    "H2O.APMO.UP2": (lambda: S.build_system(
        H2O, "6-311G", quantum=[("H-A_1", "NAKAI-7-SP", H2O[1][1], H1_MASS),
                                ("H-B_1", "NAKAI-7-SP", H2O[2][1], H1_MASS)], open_shell=True),
        -75.930658814297, {}, 1e-9),
    # test/CO.CISD-dipole.lowdin: UHF, 6-31G* (Cartesian d)
    "CO.CISD-dipole": (lambda: S.build_system(
        [("C", (0, 0, -0.914059)), ("O", (0, 0, 0.213941))], "6-31G.D", open_shell=True),
        -112.737336707933, {}, 1e-9),
    # test/PsH.FCI.lowdin: positronium hydride, UHF electrons (addParticles=1)
    # and one positron, two-particle matrices built for one particle
    "PsH.FCI": (lambda: S.build_system(
        [("H", (0, 0, 0))], "SHARON-E-6S2P", quantum=[("E+", "SHARON-E+6S2P", (0, 0, 0), 1.0)],
        extra_electrons=1, open_shell=True),
        -0.666783062050, {"build_for_one": True}, 1e-9),
    # test/H-e+H-.DD-CISD.lowdin: two H- (addParticles=2) and one positron on
    # two centres (addParticles=-1), UHF
    "H-e+H-.DD-CISD": (lambda: S.build_system(
        [("H", (0, 0, 0)), ("H", (0, 0, 3.37))], "AUG-CC-PVDZ",
        quantum=[("E+", "E+AUG-CC-PVDZ", (0, 0, 0), 1.0), ("E+", "E+AUG-CC-PVDZ", (0, 0, 3.37), 1.0)],
        extra_electrons=2, open_shell=True, add_particles={"E+": -1}),
        -1.165428966723, {}, 1e-9),
    # test/H3+.FCI.lowdin: two electrons (addParticles=-1), one quantum proton
    # (DZSPDN), overlapEigenThreshold=1E-6, UHF
    "H3+.FCI": (lambda: S.build_system(
        [("H", (0, 0.5, 0)), ("H", (0.44, -0.25, 0)), ("H", (-0.44, -0.25, 0))], "6-31G.D.P",
        quantum=[("H_1", "DZSPDN", (0, 0.5, 0), H1_MASS)], extra_electrons=-1, open_shell=True),
        -1.255739154579, {"overlap_threshold": 1e-6}, 1e-9),
    # test/HeH+.NOCI.lowdin: STO-3G electrons (addParticles=-1), quantum
    # proton in 1S1P, RHF, the input geometry (before its rotational scan)
    "HeH+.NOCI": (lambda: S.build_system(
        [("HE", (0, 0, 0)), ("H", (0, 0, 0.774))], "STO-3G",
        quantum=[("H_1", "1S1P", (0, 0, 0.774), H1_MASS)], extra_electrons=-1),
        -2.828229313778, {}, 1e-9),
    # test/Ps2F2.RHF-customEta.lowdin: two F- (addParticles=1) and two
    # positrons in PSX-DZ with eta=2 (one doubly occupied positronic orbital,
    # exchange factor kappa/eta = -1/2; Species.f90:153-158), RHF
    "Ps2F2.RHF-customEta": (lambda: _with_eta(S.build_system(
        [("F", (0, 0, -1.33)), ("F", (0, 0, 1.33))], "AUG-CC-PVDZ",
        quantum=[("E+", "PSX-DZ", (0, 0, -1.33), 1.0), ("E+", "PSX-DZ", (0, 0, 1.33), 1.0)], extra_electrons=2),
        "E+", 2.0), -199.213740198536, {}, 1e-8),
    # test/FHF-.NOCI-UHF.lowdin: the FHF- + H_1 (Nakai-3-s) system of the RHF
    # case with UHF electrons
    "FHF-.NOCI-UHF": (lambda: S.build_system(
        [("F", (0, 0, -1.140079)), ("H", (0, 0, 0)), ("F", (0, 0, 1.140079))], "6-31G",
        quantum=[("H_1", "NAKAI-3-S", (0, 0, 0), H1_MASS)], extra_electrons=1, open_shell=True),
        -199.396255393556, {}, 1e-9),
    # test/H2O.BOA.UP2.lowdin: UHF water, 6-311G, point-charge nuclei
    "H2O.BOA.UP2": (lambda: S.build_system(H2O, "6-311G", open_shell=True), -76.008843007653, {}, 1e-9),
}
