"""Extract the basis-set blocks this package needs from openLOWDIN's basis library.

openLOWDIN keeps one text file per basis set under ``lib/basis`` (DEMON2K-like
format, parsed by ``src/core/BasisSet.f90:115-253``): a header line
``O-<NAME> <SYMBOL> (<BASIS>) BASIS TYPE: n``, comment lines starting with
``#``, the number of contractions, then per contraction ``id l nprim``
followed by ``nprim`` lines ``exponent coefficient`` (raw, unnormalised
coefficients).  This script copies the numeric data of the listed
(basis, symbol) blocks into ``library.json`` next to it, so that the package
runs where the reference tree is absent.  It is a build step run by hand:

    python openlowdin_amd/basis/extract_basis.py /root/reference/lib/basis
"""
import json
import os
import sys

# (file name in lib/basis, particle symbols to keep)
WANTED = {
    "6-31G.D.P": ["H", "C", "N", "O", "F"],
    "CC-PVDZ": ["H", "C", "N", "O", "F"],
    "6-311++G.D.P": ["H", "C", "N", "O", "F"],
    "6-311G": ["H", "C", "N", "O", "F"],
    "6-31G": ["H", "C", "N", "O", "F"],
    "6-31G.D": ["H", "C", "N", "O", "F"],
    "NAKAI-5-SPD": ["H_1", "H-A_1", "H-B_1"],
    "NAKAI-3-S": ["H_1", "H-A_1", "H-B_1"],
    "NAKAI-3-SP": ["H_1"],
    "DZSPNB": ["H_1"],
    "DZSNB": ["H_1"],
    "GRIBAKIN-10S": ["E+"],
    "NAKAI-7-SP": ["H_1", "H-A_1", "H-B_1"],
    "SHARON-E-6S2P": ["H"],
    "SHARON-E+6S2P": ["E+"],
    "DZSPDN": ["H_1"],
    "AUG-CC-PVDZ": ["H", "F"],
    "PSX-DZ": ["E+"],
    "E+AUG-CC-PVDZ": ["E+"],
    "STO-3G": ["H", "HE"],
    "1S1P": ["H_1"],
}


def parse_block(lines, start):
    """Parse one element block whose header is lines[start]; return shells."""
    i = start + 1
    while lines[i].strip().startswith("#") or not lines[i].strip():
        i += 1
    ncontr = int(lines[i].split()[0])
    i += 1
    shells = []
    for _ in range(ncontr):
        while not lines[i].strip():
            i += 1
        _, l, nprim = (int(t) for t in lines[i].split()[:3])
        i += 1
        exps, coefs = [], []
        for _ in range(nprim):
            tok = lines[i].split()
            exps.append(float(tok[0].replace("D", "E")))
            coefs.append(float(tok[1].replace("D", "E")))
            i += 1
        shells.append({"l": l, "exps": exps, "coefs": coefs})
    return shells


def extract(basis_dir):
    lib = {}
    for fname, symbols in WANTED.items():
        with open(os.path.join(basis_dir, fname)) as fh:
            lines = fh.read().splitlines()
        entry = {}
        for k, line in enumerate(lines):
            tok = line.split()
            if len(tok) >= 2 and tok[0].startswith("O-") and tok[1] in symbols \
                    and tok[1] not in entry:
                entry[tok[1]] = parse_block(lines, k)
        missing = set(symbols) - set(entry)
        if missing:
            raise SystemExit("%s: missing %s" % (fname, sorted(missing)))
        lib[fname] = entry
    return lib


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/lib/basis"
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "library.json")
    with open(out, "w") as fh:
        json.dump(extract(src), fh, indent=1)
    print("wrote", out)
