"""Write the geometries of the batched workload as a fixture:
256 water clusters (H2O)_n, n ~ U{1..8}, from openlowdin_amd.system's
generator with seed 20261202 (rigid 6-31G** monomers, Haar-random rotation,
uniform translation in a cube of edge 3.1 n^(1/3) A, O-O >= 2.7 A).

    python tests/golden/make_water_batch.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from openlowdin_amd import system as S  # noqa: E402


def clusters(nmolecules=256, seed=S.BATCH_SEED):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, 9, size=nmolecules)
    return [S.water_cluster_geometry(int(n), rng) for n in sizes]


if __name__ == "__main__":
    geo = clusters()
    out = {"seed": S.BATCH_SEED, "units": "angstrom", "basis": "6-31G**",
           "clusters": [[[sym, [round(v, 12) for v in xyz]] for sym, xyz in g] for g in geo]}
    with open(os.path.join(HERE, "water_clusters_256.json"), "w") as fh:
        json.dump(out, fh, separators=(",", ":"))
    print("wrote %d clusters, %d atoms" % (len(geo), sum(len(g) for g in geo)))
