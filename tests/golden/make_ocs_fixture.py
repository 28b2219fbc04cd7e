"""Write tests/golden/ocs_alignment.npz from the reference tree's OCS data.

Provenance (all data files of the reference's TDSE test, read as text):
  tests/benchmarks/data/alignment_ocs/matelem/ocs_energies_j0_j30.rchm
  tests/benchmarks/data/alignment_ocs/matelem/ocs_matelem_alpha_j<j1>_j<j2>.rchm
  tests/benchmarks/data/alignment_ocs/field.txt          (column 4: E_Z in V/cm)
  tests/benchmarks/data/alignment_ocs/pop_lanczos.txt    (richmol's own Lanczos run,
        populations of J=0..12 every 0.1 ps, 4 decimals; written by
        tests/benchmarks/data/alignment_ocs/test_ocs_alignment.py:18-89)
  tests/benchmarks/data/alignment_ocs/pop_reference.txt  (cmirichmol populations)

The basis is the one the reference test selects (tests/test_tdse.py:36-42):
even J <= 30, m = 0 (16 states, linear molecule, one K function per J).
Only the data that basis uses is kept: energies, and per (J1, J2, irrep) the
1x1 K value and the 1x1 M value of every Cartesian component, as read by
richmol_amd.trove (the .rchm reader).

Run: python tests/golden/make_ocs_fixture.py  (needs /root/reference)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from richmol_amd.trove import CarTensTrove  # noqa: E402

REF = "/root/reference/tests/benchmarks/data/alignment_ocs/"
NSTEP = 25000          # 250 ps at dt = 0.01 ps, the length of pop_lanczos.txt
JPOP = [0, 2, 4, 6, 8, 10, 12]


def filt(**kw):
    ok = True
    if "J" in kw:
        ok = ok and kw["J"] <= 30 and kw["J"] % 2 == 0
    if "m" in kw:
        ok = ok and kw["m"] == 0.0
    return ok


def main():
    states = REF + "matelem/ocs_energies_j0_j30.rchm"
    h0 = CarTensTrove(states, bra=filt, ket=filt)
    alpha = CarTensTrove(states, REF + "matelem/ocs_matelem_alpha_j<j1>_j<j2>.rchm", bra=filt, ket=filt)
    Js = [int(J) for J in h0.Jlist1]
    assert Js == list(range(0, 31, 2)), Js
    energies = np.array([h0.kmat[(J, J)][("A", "A")][0].toarray()[0, 0].real for J in h0.Jlist1])
    carts = sorted(alpha.cart)
    rows = []
    for (J1, J2), blk in sorted(alpha.kmat.items()):
        for sp, kd in blk.items():
            assert sp == ("A", "A")
            for ir, K in kd.items():
                assert K.shape == (1, 1)
                md = alpha.mmat[(J1, J2)][sp][ir]
                mv = [md[c].toarray()[0, 0] if c in md else 0.0 for c in carts]
                rows.append((J1, J2, ir, K.toarray()[0, 0], mv))
    blk_J = np.array([[r[0], r[1]] for r in rows], dtype=np.int32)
    blk_irrep = np.array([r[2] for r in rows], dtype=np.int32)
    blk_k = np.array([r[3] for r in rows], dtype=np.complex128)
    blk_m = np.array([r[4] for r in rows], dtype=np.complex128)

    ez = []
    with open(REF + "field.txt") as f:
        for i, line in enumerate(f):
            if i >= NSTEP:
                break
            ez.append(1e2 * float(line.split()[3]))
    ez = np.array(ez)

    lan = np.loadtxt(REF + "pop_lanczos.txt")
    assert lan.shape == (NSTEP // 10, 8)
    pop_lanczos = np.rint(lan[:, 1:] * 1e4).astype(np.int16)

    ref = []
    with open(REF + "pop_reference.txt") as f:
        next(f)
        for line in f:
            w = line.split()
            n = int(w[1])
            row = dict()
            for c in range(n):
                pop, J = float(w[2 + 6 * c]), int(w[4 + 6 * c])
                row[J] = pop
            ref.append([row.get(J, 0.0) for J in JPOP])
            if len(ref) == NSTEP // 10:
                break
    pop_reference = np.rint(np.array(ref) * 1e4).astype(np.int16)

    out = os.path.join(HERE, "ocs_alignment.npz")
    np.savez_compressed(out, J=np.array(Js, dtype=np.int32), energies=energies, carts=np.array(carts),
                        blk_J=blk_J, blk_irrep=blk_irrep, blk_k=blk_k, blk_m=blk_m, field_z=ez,
                        pop_lanczos=pop_lanczos, pop_reference=pop_reference,
                        jpop=np.array(JPOP, dtype=np.int32))
    print(out, os.path.getsize(out), "bytes;", len(rows), "blocks")


if __name__ == "__main__":
    main()
