"""Watson's centrifugal-distortion terms (rot/solution.py:288-386) in the
asymmetric-top builder, pinned by what the mathematics fixes:

* the symmetric-top limit (Bx = By, no off-diagonal constants): the
  energies are the closed form
  E(J, K) = B J(J+1) + (C - B) K^2 - DJ J^2(J+1)^2 - DJK J(J+1) K^2 - DK K^4
            + HJ J^3(J+1)^3 + HJK J^2(J+1)^2 K^2 + HKJ J(J+1) K^4 + HK K^6;
* every A- and S-reduction term against an independent construction from
  explicit Jx, Jy, Jz matrices (J+- = Jx +- i Jy, products of matrices);
* the D2 species split keeps the spectrum of the whole J block;
* unknown constants are refused.
"""
import numpy as np
import pytest

from richmol_amd import rot


def _jxyz(J):
    """Explicit molecule-frame Jx, Jy, Jz over k = -J..J from the standard
    ladder J+ |k> = sqrt(J(J+1) - k(k+1)) |k+1>."""
    ks = np.arange(-J, J + 1, dtype=np.float64)
    jj = J * (J + 1.0)
    Jp = np.zeros((2 * J + 1, 2 * J + 1), dtype=np.complex128)
    for i, k in enumerate(ks[:-1]):
        Jp[i + 1, i] = np.sqrt(jj - k * (k + 1))
    Jm = Jp.conj().T
    return (Jp + Jm) / 2, (Jp - Jm) / 2j, np.diag(ks).astype(np.complex128)


def _brute(J, reduction, c):
    Jx, Jy, Jz = _jxyz(J)
    J2 = Jx @ Jx + Jy @ Jy + Jz @ Jz
    Jpl, Jmi = Jx + 1j * Jy, Jx - 1j * Jy
    P = lambda n: np.linalg.matrix_power(Jpl, n) + np.linalg.matrix_power(Jmi, n)
    Z2 = Jz @ Jz
    anti = lambda A, B: A @ B + B @ A
    g = lambda k: c.get(k, 0.0)
    H = (-g("DeltaJ") * J2 @ J2 - g("DeltaJK") * J2 @ Z2 - g("DeltaK") * Z2 @ Z2
         + g("HJ") * J2 @ J2 @ J2 + g("HJK") * J2 @ J2 @ Z2 + g("HKJ") * J2 @ Z2 @ Z2 + g("HK") * Z2 @ Z2 @ Z2)
    if reduction == "watson_a":
        H = H - 0.5 * anti(g("deltaJ") * J2 + g("deltaK") * Z2, P(2))
        H = H + 0.5 * anti(g("phiJ") * J2 @ J2 + g("phiJK") * J2 @ Z2 + g("phiK") * Z2 @ Z2, P(2))
    else:
        H = H + g("d1") * J2 @ P(2) + g("d2") * P(4) + g("h1") * J2 @ J2 @ P(2) + g("h2") * J2 @ P(4) + g("h3") * P(6)
    assert np.abs(H - H.conj().T).max() < 1e-9 * max(1.0, np.abs(H).max())
    return H.real


CONSTS_A = {"DeltaJ": 1.3e-3, "DeltaJK": -5.2e-3, "DeltaK": 2.9e-2, "deltaJ": 5.1e-4, "deltaK": 1.2e-3,
            "HJ": 5e-7, "HJK": -2e-6, "HKJ": -1e-5, "HK": 3e-5, "phiJ": 2e-7, "phiJK": -4e-7, "phiK": 6e-6}
CONSTS_S = {"DeltaJ": 1.3e-3, "DeltaJK": -5.2e-3, "DeltaK": 2.9e-2, "d1": -5.1e-4, "d2": -2.2e-5,
            "HJ": 5e-7, "HJK": -2e-6, "HKJ": -1e-5, "HK": 3e-5, "h1": 2e-7, "h2": 1e-8, "h3": 3e-9}


@pytest.mark.parametrize("J", [0, 1, 2, 3, 6, 11])
@pytest.mark.parametrize("reduction,consts", [("watson_a", CONSTS_A), ("watson_s", CONSTS_S)])
def test_terms_vs_explicit_matrices(J, reduction, consts):
    W = rot.watson_matrix(J, reduction, consts)
    ref = _brute(J, reduction, consts)
    assert np.abs(W - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("J", [0, 1, 4, 9])
def test_symmetric_top_closed_form(J):
    B, C = 0.9, 0.4
    c = {"DeltaJ": 1e-4, "DeltaJK": -3e-4, "DeltaK": 5e-4, "HJ": 1e-8, "HJK": -2e-8, "HKJ": 3e-8, "HK": -4e-8}
    bas = rot.asymtop_basis(B, B, C, J, Jmin=J, watson="watson_a", distortion=c)
    got = np.sort(np.concatenate([bas.energy[J][s] for s in bas.energy[J]]))
    jj = J * (J + 1.0)
    K = np.arange(-J, J + 1, dtype=np.float64)
    want = np.sort(B * jj + (C - B) * K ** 2 - c["DeltaJ"] * jj ** 2 - c["DeltaJK"] * jj * K ** 2
                   - c["DeltaK"] * K ** 4 + c["HJ"] * jj ** 3 + c["HJK"] * jj ** 2 * K ** 2
                   + c["HKJ"] * jj * K ** 4 + c["HK"] * K ** 6)
    assert np.abs(got - want).max() < 1e-12


@pytest.mark.parametrize("reduction,consts", [("watson_a", CONSTS_A), ("watson_s", CONSTS_S)])
def test_species_split_keeps_the_spectrum(reduction, consts):
    """Water-like constants with distortion: the four D2 blocks together have
    the eigenvalues of the whole J block (the terms couple even Delta k only
    and commute with k -> -k)."""
    A, Bc, Cc = 27.88, 14.51, 9.28
    bas = rot.asymtop_basis(Cc, A, Bc, 8, watson=reduction, distortion=consts)
    for J in bas.Jlist:
        H = rot.rigid_rotor_matrix(J, Cc, A, Bc) + rot.watson_matrix(J, reduction, consts)
        got = np.sort(np.concatenate([bas.energy[J][s] for s in bas.energy[J]]))
        assert np.abs(got - np.linalg.eigvalsh(H)).max() < 1e-10
        # the stored coefficient columns are orthonormal eigenvectors of the full block
        for s in bas.energy[J]:
            _, V = bas.kfun[J][s]
            assert np.abs(H @ V - V * bas.energy[J][s][None, :]).max() < 1e-10
            assert np.abs(V.conj().T @ V - np.eye(V.shape[1])).max() < 1e-12


def test_unknown_constants_refused():
    with pytest.raises(ValueError, match="not terms"):
        rot.watson_matrix(3, "watson_a", {"d1": 1e-3})
    with pytest.raises(ValueError, match="not available"):
        rot.watson_matrix(3, "watson_x", {})
