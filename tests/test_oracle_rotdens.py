"""Pins of the oracle's rotational densities (oracle/rotdens.py:
rot/wig.py:5-28, rotdens.py:10-283) by what the mathematics fixes -- the
reference's tests hold no density data:

* Wigner d^J(beta) from the Jy eigenvectors against the closed forms of
  J = 1 and of J = 2 entries, d(0) = 1, d(pi)_mk = (-1)^(J+m) delta_{m,-k},
  orthogonality;
* the symmetric-top functions on a product grid (Gauss-Legendre in cos beta,
  uniform alpha and gamma: exact for these band-limited functions)
  orthonormal over SO(3) for a linear rotor and an asymmetric top;
* linear rotor: psi_{J m} = c_J Y_J^m(beta, alpha) / sqrt(2 pi), c_J the
  state's coefficient over primitive k = 0;
* dens_grid: the diagonal is |psi|^2 sin(beta), and the off-diagonal blocks
  are Hermitian in the state pair.
"""
import numpy as np
import pytest

from oracle.rotdens import dens_grid, psi_grid, wig_d
from richmol_amd import rot

B = np.linspace(0.05, 3.1, 9)


def test_wigner_d_j1_closed_form():
    c, s = np.cos(B), np.sin(B)
    # rows/columns m, k = -1, 0, 1
    want = np.array([[(1 + c) / 2, s / np.sqrt(2), (1 - c) / 2],
                     [-s / np.sqrt(2), c, s / np.sqrt(2)],
                     [(1 - c) / 2, -s / np.sqrt(2), (1 + c) / 2]])
    d = wig_d(1, B)
    assert np.abs(d - want).max() < 1e-14


def test_wigner_d_j2_entries():
    c, s = np.cos(B), np.sin(B)
    d = wig_d(2, B)   # index m + 2
    assert np.abs(d[2, 2] - (3 * c ** 2 - 1) / 2).max() < 1e-14
    assert np.abs(d[4, 2] - np.sqrt(3.0 / 8.0) * s ** 2).max() < 1e-14
    assert np.abs(d[4, 4] - ((1 + c) / 2) ** 2).max() < 1e-14
    assert np.abs(d[4, 3] + (1 + c) * s / 2).max() < 1e-14


@pytest.mark.parametrize("J", [0, 1, 3, 6])
def test_wigner_d_special_values(J):
    d = wig_d(J, np.array([0.0, np.pi, 1.1]))
    n = 2 * J + 1
    assert np.abs(d[:, :, 0] - np.eye(n)).max() < 1e-13
    m = np.arange(-J, J + 1)
    want = np.zeros((n, n))
    want[np.arange(n), n - 1 - np.arange(n)] = (-1.0) ** (J + m)
    assert np.abs(d[:, :, 1] - want).max() < 1e-13
    assert np.abs(d[:, :, 2] @ d[:, :, 2].T - np.eye(n)).max() < 1e-13


def _grid(nb=12, na=16, ng=16):
    x, w = np.polynomial.legendre.leggauss(nb)
    return 2 * np.pi * np.arange(na) / na, np.arccos(x), 2 * np.pi * np.arange(ng) / ng, w


@pytest.mark.parametrize("kind", ["linear", "asym"])
def test_psi_grid_orthonormal_over_so3(kind):
    bas = rot.linear_basis(1.0, 3) if kind == "linear" else rot.asymtop_basis(9.28, 27.88, 14.51, 3)
    alpha, beta, gamma, w = _grid()
    psi = psi_grid(bas, alpha, beta, gamma)
    allpsi = np.concatenate([psi[J][s] for J in bas.Jlist for s in bas.symlist[J]], axis=-1)
    S = np.einsum("abgi,abgj,b->ij", np.conj(allpsi), allpsi, w) * (2 * np.pi / len(alpha)) ** 2
    assert np.abs(S - np.eye(S.shape[0])).max() < 1e-12


def test_linear_rotor_spherical_harmonics():
    from scipy.special import sph_harm_y
    bas = rot.linear_basis(1.0, 4)
    alpha, beta, gamma = np.array([0.0, 0.7, 2.5]), np.array([0.3, 0.9, 1.7, 2.9]), np.array([0.0, 1.3])
    psi = psi_grid(bas, alpha, beta, gamma)
    A, Bg = np.meshgrid(alpha, beta, indexing="ij")
    for J in bas.Jlist:
        c = bas.kfun[J]["A"][1][0, 0]
        for i, m in enumerate(bas.mlist[J]):
            Y = sph_harm_y(J, m, Bg, A) / np.sqrt(2 * np.pi)
            for g in range(len(gamma)):
                assert np.abs(psi[J]["A"][:, :, g, i] - c * Y).max() < 1e-13


def test_dens_grid_diagonal_and_hermitian():
    bas = rot.asymtop_basis(9.28, 27.88, 14.51, 2)
    alpha, beta, gamma = np.array([0.2, 1.9]), np.array([0.4, 2.2]), np.array([0.1, 3.0])
    psi = psi_grid(bas, alpha, beta, gamma)
    dd = dens_grid(bas, alpha, beta, gamma, diag_only=True)
    full = dens_grid(bas, alpha, beta, gamma)
    for J in bas.Jlist:
        for s in bas.symlist[J]:
            want = np.abs(psi[J][s]) ** 2 * np.sin(beta)[None, :, None, None]
            assert np.abs(dd[(J, J)][(s, s)] - want).max() < 1e-14
    for (J1, J2), blk in full.items():
        for (s1, s2), d in blk.items():
            back = full[(J2, J1)][(s2, s1)]
            assert np.abs(d - np.conj(np.swapaxes(back, -1, -2))).max() < 1e-14
