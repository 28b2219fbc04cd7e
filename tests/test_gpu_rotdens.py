"""Rotational densities on the device (rm_rotdens through
richmol_amd.rotdens.density) against the oracle's wave functions
(oracle/rotdens.py: |sum_n c_wn psi_n|^2 sin(beta)), per wavepacket and
summed over an ensemble, and at a larger size through a property the
mathematics fixes: on a product grid exact for these band-limited
functions (Gauss-Legendre in cos(beta), uniform alpha and gamma) the density
integrates to the squared norms of the rows.
"""
import numpy as np
import pytest

from richmol_amd import rot

pytestmark = pytest.mark.gpu


def _oracle_density(bas, rows, alpha, beta, gamma):
    from oracle.rotdens import psi_grid
    p = psi_grid(bas, alpha, beta, gamma)
    full = np.concatenate([p[J][s] for J in bas.Jlist for s in bas.symlist[J]], axis=-1)
    psi = np.einsum("abgn,wn->wabg", full, rows)
    return np.abs(psi) ** 2 * np.sin(beta)[None, None, :, None]


@pytest.mark.parametrize("kind", ["linear", "asym"])
def test_density_vs_oracle(device, kind):
    from richmol_amd import rotdens
    bas = rot.linear_basis(1.3, 5) if kind == "linear" else rot.asymtop_basis(9.28, 27.88, 14.51, 4)
    n = sum(len(bas.mlist[J]) * bas.kfun[J][s][1].shape[1] for J in bas.Jlist for s in bas.symlist[J])
    rng = np.random.default_rng(11)
    rows = rng.normal(size=(5, n)) + 1j * rng.normal(size=(5, n))
    alpha = np.linspace(0.0, 2 * np.pi, 9, endpoint=False)
    beta = np.linspace(0.05, 3.0, 7)
    gamma = np.linspace(0.0, 2 * np.pi, 6, endpoint=False)
    want = _oracle_density(bas, rows, alpha, beta, gamma)
    per = rotdens.density(bas, rows, alpha, beta, gamma, per_wavepacket=True).cpu().numpy()
    ens = rotdens.density(bas, rows, alpha, beta, gamma).cpu().numpy()
    scale = want.max()
    assert np.abs(per - want).max() <= 1e-12 * scale
    assert np.abs(ens - want.sum(axis=0)).max() <= 1e-12 * scale * len(rows)


def test_ensemble_density_normalisation_j40(device):
    """512 Boltzmann-sampled |J, m> states of the J <= 40 linear rotor with
    random phases mixed by a unitary rotation inside each m: the ensemble
    density integrates to the number of rows (each row normalised) to 1e-10
    on a grid of 42 x 82 x 2 points... (gamma does not enter for k = 0)."""
    from richmol_amd import rotdens
    from richmol_amd.ensemble import basis_rows, thermal_sample
    bas = rot.linear_basis(0.2029, 40)
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    rows = basis_rows(n, thermal_sample(bas, 50.0, 512, seed=1))
    rng = np.random.default_rng(3)
    rows = rows * np.exp(2j * np.pi * rng.random(size=(512, 1)))
    x, w = np.polynomial.legendre.leggauss(42)
    beta = np.arccos(x)
    na, ng = 82, 2
    alpha = 2 * np.pi * np.arange(na) / na
    gamma = 2 * np.pi * np.arange(ng) / ng
    rho = rotdens.density(bas, rows, alpha, beta, gamma).cpu().numpy()
    # d(cos beta) = sin(beta) d(beta): the Gauss weights replace sin(beta) d(beta)
    integral = np.einsum("abg,b->", rho / np.sin(beta)[None, :, None], w) * (2 * np.pi / na) * (2 * np.pi / ng)
    assert abs(integral - 512.0) < 1e-9
    assert rho.min() >= 0.0
