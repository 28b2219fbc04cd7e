"""Build tests/golden/camphor_j3.npz from the reference's r-camphor data files
(tests/benchmarks/data/r-camphor_rchm_files: cmirichmol energies and dipole /
polarizability matrix elements of R-camphor, J <= 20).

Stored (data only): the rotational energies for J <= 20 (J, species, energy),
and, for the J <= 3 states, the eigenvalues of H0 + lam * V(F) for the dipole
and the polarizability operators read from the files (via the package's
.rchm reader) at four field directions -- basis-invariant quantities an
independently built operator must reproduce.

    python tests/golden/make_camphor_fixture.py /root/reference
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from richmol_amd.trove import CarTensTrove  # noqa: E402
from oracle.cartens import OracleCarTens  # noqa: E402

FIELDS = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.3, -0.2, 0.9]])
LAM_MU, LAM_ALPHA, JCUT = 1e-2, 1e-4, 3


def main(ref_root):
    p = os.path.join(ref_root, "tests", "benchmarks", "data", "r-camphor_rchm_files")
    states = os.path.join(p, "camphor_energies_j0_j20.rchm")
    J, sym, E = [], [], []
    for line in open(states):
        w = line.split()
        try:
            J.append(int(w[0])); sym.append(w[2]); E.append(float(w[4]))
        except (IndexError, ValueError):
            continue
    filt = lambda **kw: kw["J"] <= JCUT if "J" in kw else True   # noqa: E731
    h0 = CarTensTrove(states, bra=filt, ket=filt)
    mu = CarTensTrove(states, matelem=os.path.join(p, "camphor_matelem_mu_j<j1>_j<j2>.rchm"), bra=filt, ket=filt)
    al = CarTensTrove(states, matelem=os.path.join(p, "camphor_matelem_alpha_j<j1>_j<j2>.rchm"), bra=filt, ket=filt)
    H0 = OracleCarTens.from_attrs(h0).tomat(form="full", cart="0", repres="dense")

    def spec(T, lam):
        out = []
        for F in FIELDS:
            O = OracleCarTens.from_attrs(T)
            O.field(list(F))
            out.append(np.linalg.eigvalsh(H0 + lam * O.tomat(form="full", repres="dense")))
        return np.array(out)

    np.savez_compressed(os.path.join(HERE, "camphor_j3.npz"), J=np.array(J), sym=np.array(sym), energy=np.array(E),
                        fields=FIELDS, lam_mu=LAM_MU, lam_alpha=LAM_ALPHA, jcut=JCUT,
                        eig_mu=spec(mu, LAM_MU), eig_alpha=spec(al, LAM_ALPHA))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
