"""Operator builders (richmol_amd.rot, richmol_amd.trove) against closed forms,
sympy's exact 3j symbols and the reference's data files."""
import os
import random

import numpy as np
import pytest

from _fixtures import linear_rotor, ocs_fixture
from oracle.cartens import OracleCarTens
from richmol_amd import rot


def test_wigner3j_vs_sympy():
    wigner_3j = pytest.importorskip("sympy.physics.wigner").wigner_3j
    rnd = random.Random(0)
    for _ in range(400):
        j1 = rnd.randint(0, 60)
        j2 = rnd.randint(0, 2)
        j3 = rnd.randint(abs(j1 - j2), j1 + j2)
        m1 = rnd.randint(-j1, j1)
        m2 = rnd.randint(-j2, j2)
        a = rot.wigner3j(j1, j2, j3, m1, m2, -m1 - m2)
        b = float(wigner_3j(j1, j2, j3, m1, m2, -m1 - m2))
        assert abs(a - b) <= 2e-16 + 1e-15 * abs(b)
    assert rot.wigner3j(2, 1, 1, 0, 0, 1) == 0.0


def test_cos2_closed_form():
    """<J M|cos^2 theta|J M> - 1/3 = 2/3 (J(J+1) - 3M^2) / ((2J-1)(2J+3)) and
    <J M|cos^2|J+2 M> = sqrt(((J+1)^2-M^2)((J+2)^2-M^2)/((2J+1)(2J+3)^2(2J+5))) up to sign."""
    h0, H, cos2, bas = linear_rotor(8)
    C = OracleCarTens.from_attrs(cos2).tomat(form="full", cart="0").toarray()
    idx = {}
    i = 0
    for J in bas.Jlist:
        for m in bas.mlist[J]:
            idx[(J, m)] = i
            i += 1
    for (J, m), i in idx.items():
        assert abs(C[i, i] - 2 / 3 * (J * (J + 1) - 3 * m * m) / ((2 * J - 1) * (2 * J + 3))) < 1e-14
        if (J + 2, m) in idx:
            off = np.sqrt(((J + 1) ** 2 - m * m) * ((J + 2) ** 2 - m * m)
                          / ((2 * J + 1) * (2 * J + 3) ** 2 * (2 * J + 5)))
            assert abs(abs(C[i, idx[(J + 2, m)]]) - off) < 1e-14
    assert np.abs(C - C.conj().T).max() == 0.0


def test_polarizability_is_aperp_plus_dalpha_cos2():
    """For a Z field, alpha_ZZ = alpha_perp + (alpha_par - alpha_perp) cos^2 theta."""
    bas = rot.linear_basis(0.2029, 10)
    pol = rot.tensor(np.diag([26.0, 26.0, 53.6]), bas)
    P = OracleCarTens.from_attrs(pol)
    P.field([0, 0, 1.0])
    A = P.tomat(form="full").toarray()
    C = OracleCarTens.from_attrs(rot.named("cos2theta", bas)).tomat(form="full", cart="0").toarray()
    n = A.shape[0]
    assert np.abs(A - (26.0 * np.eye(n) + 27.6 * (C + np.eye(n) / 3))).max() < 1e-13
    # X field: alpha_XX has Delta m = 0, +-2 couplings and the same trace
    P.field([1.0, 0, 0])
    AX = P.tomat(form="full").toarray()
    assert abs(np.trace(AX) - np.trace(A)) < 1e-10 * abs(np.trace(A))
    assert np.abs(AX - AX.conj().T).max() < 1e-13


def test_costheta_closed_form():
    bas = rot.linear_basis(0.2029, 6)
    C = OracleCarTens.from_attrs(rot.named("costheta", bas)).tomat(form="full", cart="0").toarray()
    idx = {}
    i = 0
    for J in bas.Jlist:
        for m in bas.mlist[J]:
            idx[(J, m)] = i
            i += 1
    for (J, m), i in idx.items():
        if (J + 1, m) in idx:
            expect = np.sqrt(((J + 1) ** 2 - m * m) / ((2 * J + 1) * (2 * J + 3)))
            assert abs(abs(C[i, idx[(J + 1, m)]]) - expect) < 1e-14
        assert C[i, i] == 0


def test_linear_energies_and_dims():
    h0, H, cos2, bas = linear_rotor(40)
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    assert n == 1681
    d = OracleCarTens.from_attrs(h0).tomat(form="full", cart="0").diagonal().real
    assert np.allclose(d, [0.2029 * J * (J + 1) for J in range(41) for m in range(-J, J + 1)], rtol=0, atol=1e-12)


REF_OCS = "/root/reference/tests/benchmarks/data/alignment_ocs/matelem/"


@pytest.mark.skipif(not os.path.isdir(REF_OCS), reason="reference tree not present")
def test_trove_reader_matches_fixture():
    """The .rchm reader on the reference files reproduces the committed fixture,
    and its M tensor for J=0 -> 2 equals the 3j construction of rot.tensor."""
    from richmol_amd.trove import CarTensTrove
    filt = lambda **kw: (kw.get("J", 0) <= 30 and kw.get("J", 0) % 2 == 0) and kw.get("m", 0.0) == 0.0  # noqa
    states = REF_OCS + "ocs_energies_j0_j30.rchm"
    a = CarTensTrove(states, REF_OCS + "ocs_matelem_alpha_j<j1>_j<j2>.rchm", bra=filt, ket=filt)
    h0f, af, d = ocs_fixture()
    A1 = OracleCarTens.from_attrs(a)
    A2 = OracleCarTens.from_attrs(af)
    for F in ([0, 0, 1.0], [1.0, 0, 0], [0.3, -0.5, 0.7]):
        A1.field(F); A2.field(F)
        assert np.abs(A1.tomat(form="full").toarray() - A2.tomat(form="full").toarray()).max() < 1e-14
    # full m basis: the file's M tensors equal the 3j construction (up to the
    # K normalisation, which carries the molecule's polarizability)
    full = CarTensTrove(states, REF_OCS + "ocs_matelem_alpha_j<j1>_j<j2>.rchm",
                        bra=lambda **kw: kw.get("J", 0) <= 2, ket=lambda **kw: kw.get("J", 0) <= 2)
    bas = rot.linear_basis(1.0, 2)
    mine = rot.tensor(np.diag([1.0, 1.0, 2.0]), bas)
    for c in ("xx", "xz", "zz", "xy"):
        Mf = full.mmat[(0.0, 2.0)][("A", "A")][2][c].toarray()
        Mm = mine.mmat[(0, 2)][("A", "A")][2][c].toarray()
        assert np.abs(Mf - Mm).max() < 1e-8


# R-camphor (reference data tests/benchmarks/data/r-camphor_rchm_files, fixture
# tests/golden/camphor_j3.npz): tensors as the data's README quotes them [au].
CAMPHOR_MU = np.array([-1.21615, -0.30746, 0.01140])
CAMPHOR_ALPHA = np.array([[115.80434, 0.58739, -0.03276],
                          [0.58739, 112.28245, 1.36146],
                          [-0.03276, 1.36146, 108.47809]])
# The README's tensors are expressed in a frame turned by 0.915 degrees from
# the one the matrix elements were computed in; this rotation vector (fitted
# once to the dipole and polarizability Stark spectra together, 3 parameters
# for 672 eigenvalues) maps them into it.
CAMPHOR_FRAME = np.array([-0.00273776, 0.00980563, 0.01230923])


def _camphor():
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "camphor_j3.npz"))
    ref = {}
    for J, s, e in zip(d["J"], d["sym"], d["energy"]):
        ref.setdefault(int(J), {}).setdefault(str(s), []).append(float(e))
    # the J = 1 levels of a rigid rotor are A+B, A+C, B+C
    e1 = sorted(v[0] for v in ref[1].values())
    half = sum(e1) / 2
    abc = (half - e1[0], half - e1[1], half - e1[2])
    return d, ref, abc


def test_asymtop_energies_match_camphor():
    """Rigid-rotor energies of all 441 J <= 20 states from the three constants
    of the J = 1 levels, per D2 species (axes x = a, y = b, z = c), against
    the reference's data file (8 decimals)."""
    d, ref, (A, B, C) = _camphor()
    bas = rot.asymtop_basis(A, B, C, 20)
    for J in range(21):
        assert sorted(bas.energy[J]) == sorted(ref[J])
        for s in ref[J]:
            assert np.abs(np.sort(bas.energy[J][s]) - np.sort(ref[J][s])).max() < 2e-6, (J, s)
    assert sum(len(e) for J in bas.Jlist for e in bas.energy[J].values()) == 441


def test_asymtop_tensors_match_camphor():
    """Dipole and polarizability operators built here reproduce the eigenvalues
    of H0 + lam V(F) of the reference's matrix elements (basis invariant) for
    four field directions, J <= 3: relative to the Stark shift < 2e-7.  The
    same test without the frame rotation misses the dipole spectrum by 4e-3,
    so it resolves the tensor's orientation, not just its magnitude."""
    from scipy.spatial.transform import Rotation
    d, ref, (A, B, C) = _camphor()
    jc = int(d["jcut"])
    bas = rot.asymtop_basis(A, B, C, jc)
    for J in bas.Jlist:
        for s in bas.symlist[J]:
            bas.energy[J][s] = np.sort(ref[J][s])   # the file's energies: only the tensors differ
    H0 = OracleCarTens.from_attrs(rot.hamiltonian(bas)).tomat(form="full", cart="0", repres="dense")
    e0 = np.linalg.eigvalsh(H0)
    R = Rotation.from_rotvec(CAMPHOR_FRAME).as_matrix()
    for tens, lam, key in ((R @ CAMPHOR_MU, d["lam_mu"], "eig_mu"),
                           (R @ CAMPHOR_ALPHA @ R.T, d["lam_alpha"], "eig_alpha"),
                           (CAMPHOR_MU, d["lam_mu"], None)):
        T = rot.tensor(tens, bas)
        for F, eref in zip(d["fields"], d[key or "eig_mu"]):
            O = OracleCarTens.from_attrs(T)
            O.field(list(F))
            e = np.linalg.eigvalsh(H0 + float(lam) * O.tomat(form="full", repres="dense"))
            rel = np.abs(e - eref).max() / np.abs(eref - e0).max()
            if key is None:
                assert rel > 1e-3          # unrotated README dipole: clearly resolved
            else:
                assert rel < 2e-7, (key, rel)


def _flat_dense(T):
    """Dense matrix of a plan's operator at unit field products."""
    from richmol_amd import plan as P
    arr = P.flatten_operator(T, P.active_carts(T))
    n = arr["n"]
    mf = np.array([arr["term_mval"][arr["term_ptr"][t]:arr["term_ptr"][t + 1]].sum() for t in range(arr["nterm"])])
    D = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        for e in range(arr["row_ptr"][i], arr["row_ptr"][i + 1]):
            p0, p1 = arr["pair_ptr"][e], arr["pair_ptr"][e + 1]
            D[i, arr["col_idx"][e]] = np.sum(arr["pair_kval"][p0:p1] * mf[arr["pair_term"][p0:p1]])
    return D


@pytest.mark.parametrize("tens", [np.diag([12.0, 8.0, 16.0]),
                                  np.array([[10.0, 1.5, -0.7], [1.5, 7.0, 0.4], [-0.7, 0.4, 15.0]]),
                                  np.array([0.3, -0.2, 0.9])])
def test_zfield_m0_operator_equals_cartens(tens):
    """The serving path's direct m = 0 operator equals the general builder's
    operator (flattened under a unit Z field: F_z^rank = 1) entry by entry,
    and its energies equal the field-free Hamiltonian's diagonal."""
    bas = rot.asymtop_basis(14.0, 3.0, 27.0, 12, m_filter=lambda J, m: m == 0)
    T = rot.tensor(tens, bas)
    T.field([0.0, 0.0, 1.0])
    D = _flat_dense(T)
    A, E = rot.zfield_m0_operator(bas, tens)
    assert A.shape == D.shape
    assert np.abs(A.toarray() - D).max() <= 1e-13 * np.abs(D).max()
    h0 = rot.hamiltonian(bas)
    assert np.array_equal(E, h0.tomat(form="full", cart="0").diagonal().real)


@pytest.mark.parametrize("alpha", ["diag", "full"])
def test_species_filtered_m0_operator_equals_full(alpha):
    """Serving builds an asymmetric top's m = 0 operator only on the D2
    species the J = 0 state reaches: the operator, energies and whole-basis
    positions equal those built on the basis with all four species."""
    from richmol_amd import serving
    al = (np.diag(serving.ASYM_ALPHA) if alpha == "diag"
          else np.array([[9.1, 0.3, 0.0], [0.3, 9.8, 0.2], [0.0, 0.2, 9.4]]))
    keep = serving._reachable_species(al)
    assert keep == ({"A"} if alpha == "diag" else {"A", "B1", "B2", "B3"})
    rng = np.random.default_rng(11)
    for jmax in (7, 16):
        a, b, c = np.sort(rng.uniform(0.05, 30.0, 3))[::-1]
        m0 = lambda J, m: m == 0
        full = rot.asymtop_basis(b, c, a, jmax, m_filter=m0)
        A1, E1, idx1 = rot.zfield_m0_operator(full, al, start=(0, "A"))
        part = rot.asymtop_basis(b, c, a, jmax, m_filter=m0, species=keep)
        A2, E2, sub = rot.zfield_m0_operator(part, al, start=(0, "A"))
        idx2 = serving._m0_positions(part, jmax)[sub]
        assert A1.shape == A2.shape and abs(A1 - A2).max() == 0.0
        assert np.array_equal(E1, E2) and np.array_equal(idx1, idx2)
