"""Writes g12_potentials.json: the Gaussian-expansion interparticle and
external potentials that openLOWDIN's own G12 test input holds
(test/TIP4P-dimer-singlet-direct.lowdin, POTENTIAL blocks VOH-CCSDT and
VHH-CCSDT), as (exponent, coefficient) lists per named component.

Run in a container that has the reference tree:
    python tests/golden/make_g12_potentials.py [/path/to/reference]
The file is parsed as text; nothing of the reference is executed.
"""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(REF, "test", "TIP4P-dimer-singlet-direct.lowdin")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "g12_potentials.json")


def parse(path):
    lines = [l.split("#")[0].strip() for l in open(path)]
    pots, i = {}, 0
    while i < len(lines):
        if lines[i].upper().startswith("POTENTIAL "):
            name = lines[i].split()[1]
            i += 1
            comps = {}
            while not lines[i].upper().startswith("END POTENTIAL"):
                if not lines[i]:
                    i += 1
                    continue
                label = lines[i]
                count = int(lines[i + 1])
                i += 2
                terms = []
                for _ in range(count):
                    _idx, l = lines[i].split()
                    expo, coef = (float(x) for x in lines[i + 1].split())
                    origin = [float(x) for x in lines[i + 2].split()]
                    terms.append({"l": int(l), "exponent": expo, "coefficient": coef, "origin": origin})
                    i += 3
                comps[label] = terms
            pots[name] = comps
        i += 1
    return pots


if __name__ == "__main__":
    pots = parse(SRC)
    with open(OUT, "w") as fh:
        json.dump({"source": "test/TIP4P-dimer-singlet-direct.lowdin", "potentials": pots}, fh, indent=1)
    print("wrote", OUT, {k: {c: len(v) for c, v in d.items()} for k, d in pots.items()})
