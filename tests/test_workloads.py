"""The workload generators against their committed fixtures (CPU)."""
import json
import os

import numpy as np

from openlowdin_amd import system as S

HERE = os.path.dirname(os.path.abspath(__file__))


def test_water_batch_fixture_matches_generator():
    with open(os.path.join(HERE, "golden", "water_clusters_256.json")) as fh:
        data = json.load(fh)
    assert data["seed"] == S.BATCH_SEED and len(data["clusters"]) == 256
    rng = np.random.default_rng(S.BATCH_SEED)
    sizes = rng.integers(1, 9, size=256)
    for n, g in zip(sizes, data["clusters"]):
        ref = S.water_cluster_geometry(int(n), rng)
        assert len(g) == 3 * n == len(ref)
        for (sym, xyz), (rsym, rxyz) in zip(g, ref):
            assert sym == rsym and np.allclose(xyz, rxyz, atol=1e-10)


def test_water_batch_geometry_rules():
    """Monomers rigid (the C0 water's O-H length), O-O >= 2.7 A, O inside the cube of
    edge 3.1 n^(1/3) A; sizes cover 1..8."""
    systems = S.load_water_batch(os.path.join(HERE, "golden", "water_clusters_256.json"))
    roh = float(np.linalg.norm(np.subtract(S.WATER_ANGSTROM[1][1], S.WATER_ANGSTROM[0][1])))
    sizes = set()
    with open(os.path.join(HERE, "golden", "water_clusters_256.json")) as fh:
        clusters = json.load(fh)["clusters"]
    for g in clusters:
        n = len(g) // 3
        sizes.add(n)
        O_ = np.array([x for s, x in g if s == "O"])
        H_ = np.array([x for s, x in g if s == "H"]).reshape(n, 2, 3)
        for k in range(n):
            assert np.allclose(np.linalg.norm(H_[k] - O_[k], axis=1), roh * np.ones(2), atol=1e-9)
        d = np.linalg.norm(O_[:, None] - O_[None], axis=-1) + np.eye(n) * 99
        assert d.min() >= 2.7
        assert (O_ >= 0).all() and (O_ <= 3.1 * n ** (1 / 3)).all()
    assert sizes == set(range(1, 9))
    assert all(s.species[0].nbasis == 25 * (len(c) // 3) for s, c in zip(systems, clusters))
