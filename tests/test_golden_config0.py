"""The linear-rotor fixture (tests/golden/config0_oracle.npz) against the live
oracle: the first steps re-derived here (CPU) must equal the stored ones, so
the fixture is the oracle's output and not a stale file; and the stored run
has the properties the mathematics fixes: unit norm at every step, <cos^2>
near 1/3 after the first (weak-field) step from |J=0>, populations frozen
once the field is zero, and the fixed-dimension and adaptive runs agreeing
to the adaptive rule's tolerance."""
import os

import numpy as np

from _fixtures import GOLDEN

NCHECK = 40


def test_fixture_prefix_equals_live_oracle():
    import sys
    sys.path.insert(0, GOLDEN)
    from make_config0_fixture import run
    g = np.load(os.path.join(GOLDEN, "config0_oracle.npz"))
    for mode, m in (("fixed", 12), ("adaptive", None)):
        psi, c2, k = run(m, nsteps=NCHECK)
        assert np.abs(psi - g[f"psi_{mode}"][:NCHECK]).max() <= 1e-13
        assert np.abs(c2 - g[f"cos2_{mode}"][:NCHECK]).max() <= 1e-13
        assert np.array_equal(k, g[f"kused_{mode}"][:NCHECK])


def test_fixture_invariants():
    g = np.load(os.path.join(GOLDEN, "config0_oracle.npz"))
    for mode in ("fixed", "adaptive"):
        psi, c2, k = g[f"psi_{mode}"], g[f"cos2_{mode}"], g[f"kused_{mode}"]
        assert psi.shape == (1000, 16)
        assert np.abs(np.linalg.norm(psi, axis=1) - 1.0).max() < 1e-12
        # <0,0|cos^2|0,0> = 1/3 and the first step's field is still weak
        assert abs(c2[0] - 1.0 / 3.0) < 1e-3
        # no field from step 249 on: Krylov skipped, populations frozen
        assert np.all(k[249:] == 0) and np.all(k[:249] > 0)
        pops = np.abs(psi[249:]) ** 2
        assert np.abs(pops - pops[0]).max() < 1e-13
    # fixed dimension 12 and the adaptive rule agree to the adaptive rule's
    # tolerance: it stops once the squared change of the solution is <= 1e-15,
    # which leaves ~1e-12 per field step (1.6e-10 after the 249 of this run)
    assert np.abs(g["psi_fixed"] - g["psi_adaptive"]).max() < 1e-9
