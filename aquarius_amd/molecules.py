"""Molecular workloads: the water monomer and seeded water clusters.

H2O geometry (angstrom): O (0, 0, 0), H (+-0.7572, 0, 0.5865).  The
reference's own H2O test (test/input.aq:3-11) uses the numbers
O (0, 0, 0.11726921), H (+-0.75698224, 0, -0.46907685) with `units bohr`.

Water clusters: n copies of the monomer, each rotated by a uniformly random
rotation (QR of a Gaussian 3x3 with the sign fixed so the distribution is the
Haar measure) about its oxygen, with oxygens drawn uniformly in a cube of edge
`box` angstrom and redrawn while closer than `min_oo` angstrom to an oxygen
already placed.  numpy's PCG64 generator seeded with `seed` makes the geometry
reproducible on every machine.
"""

import numpy as np

H2O_ANGSTROM = [("O", 0.0, 0.0, 0.0), ("H", 0.7572, 0.0, 0.5865), ("H", -0.7572, 0.0, 0.5865)]
H2O_REFERENCE_BOHR = [("O", 0.0, 0.0, 0.11726921), ("H", 0.75698224, 0.0, -0.46907685),
                      ("H", -0.75698224, 0.0, -0.46907685)]


def random_rotation(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q @ np.diag(np.sign(np.diag(r)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def water_cluster(seed, n=8, box=7.0, min_oo=2.7, max_draws=100000):
    """[(symbol, x, y, z)] in angstrom for n waters (see module docstring)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mono = np.array([a[1:] for a in H2O_ANGSTROM])
    oxygens = []
    atoms = []
    draws = 0
    while len(oxygens) < n:
        draws += 1
        if draws > max_draws:
            raise RuntimeError("could not place the waters; enlarge the box or lower min_oo")
        o = rng.uniform(0.0, box, size=3)
        if any(np.linalg.norm(o - p) < min_oo for p in oxygens):
            continue
        R = random_rotation(rng)
        oxygens.append(o)
        for (sym, *_), r in zip(H2O_ANGSTROM, mono):
            x = o + R @ r
            atoms.append((sym, float(x[0]), float(x[1]), float(x[2])))
    return atoms
