"""The product's host rotational densities (richmol_amd.rotdens) against the
oracle's restatement of rotdens.py (pinned in tests/test_oracle_rotdens.py),
and the primitive-coefficient form the device density evaluates."""
import numpy as np
import pytest

from richmol_amd import rot, rotdens

GRID = (np.array([0.0, 0.9, 2.2, 4.0]), np.array([0.2, 1.1, 2.0, 2.9]), np.array([0.3, 1.7, 5.5]))


def _bases():
    return {"linear": rot.linear_basis(1.3, 4), "asym": rot.asymtop_basis(9.28, 27.88, 14.51, 3)}


@pytest.mark.parametrize("kind", ["linear", "asym"])
def test_psi_and_dens_equal_the_oracle(kind):
    from oracle import rotdens as ord_
    bas = _bases()[kind]
    p = rotdens.psi_grid(bas, *GRID)
    q = ord_.psi_grid(bas, *GRID)
    for J in bas.Jlist:
        for s in bas.symlist[J]:
            assert np.abs(p[J][s] - q[J][s]).max() < 1e-13
    d = rotdens.dens_grid(bas, *GRID)
    e = ord_.dens_grid(bas, *GRID)
    for key, blk in e.items():
        for sk, v in blk.items():
            assert np.abs(d[key][sk] - v).max() < 1e-13
    dd = rotdens.dens_grid(bas, *GRID, diag_only=True, form="full")
    full = rotdens.psi_grid(bas, *GRID, form="full")
    assert np.abs(dd - np.abs(full) ** 2 * np.sin(GRID[1])[None, :, None, None]).max() < 1e-13


@pytest.mark.parametrize("J", [0, 2, 5])
def test_wigner_d_equals_the_oracle(J):
    from oracle import rotdens as ord_
    b = np.linspace(0.0, np.pi, 7)
    assert np.abs(rotdens.wigner_d(J, b) - ord_.wig_d(J, b)).max() < 1e-13


@pytest.mark.parametrize("kind", ["linear", "asym"])
def test_primitive_coefficients_reproduce_the_wavepacket(kind):
    """psi_w = sum_J sum_mk a^J_mk e^{i m alpha} d^J_mk(beta) e^{i k gamma}
    (the form rm_rotdens evaluates) equals sum_n c_wn psi_n."""
    bas = _bases()[kind]
    full = rotdens.psi_grid(bas, *GRID, form="full")
    rng = np.random.default_rng(5)
    rows = rng.normal(size=(3, full.shape[-1])) + 1j * rng.normal(size=(3, full.shape[-1]))
    a = rotdens.primitive_coefficients(bas, rows)
    jmax = max(bas.Jlist)
    mu = np.arange(-jmax, jmax + 1)
    al, be, ga = GRID
    dt = np.zeros((jmax + 1, len(be), 2 * jmax + 1, 2 * jmax + 1))
    for J in range(jmax + 1):
        dt[J, :, jmax - J:jmax + J + 1, jmax - J:jmax + J + 1] = np.moveaxis(rotdens.wigner_d(J, be), 2, 0)
    T = np.einsum("wjmk,jbmk->wbmk", a, dt)
    psi = np.einsum("am,wbmk,gk->wabg", np.exp(1j * np.outer(al, mu)), T, np.exp(1j * np.outer(ga, mu)))
    want = np.einsum("abgn,wn->wabg", full, rows)
    assert np.abs(psi - want).max() < 1e-12 * np.abs(want).max()
