"""Long-domain kernels of rm_propagate: the cooperative kernel (one
wavepacket at a time on every workgroup of the device, grid barriers between
the Lanczos reductions) and the block kernel (one workgroup per wavepacket).
They take the domains longer than one wave holds (RM_SUBSPACE_MAX_ROWS): the
cooperative kernel up to 8 such wavepackets per call, the block kernel more;
forced with kernel 'coop' / 'block', any domain.  Checked against the FP64
oracle on the same inputs: 1e-10 absolute in every amplitude (entries
outside the domain exactly zero) and identical adaptive Krylov dimensions
per step.
"""
import numpy as np
import pytest

from _fixtures import gaussian_pulse_field, linear_rotor
from test_gpu_subspace import TOL_PSI, _basis_columns, _device_run, _oracle_run

pytestmark = pytest.mark.gpu


def _xz_columns(n, ncols, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(ncols, n)) + 1j * rng.normal(size=(ncols, n))
    return X / np.linalg.norm(X, axis=1)[:, None]


@pytest.mark.parametrize("fixed", [0, 12])
def test_long_domain_xz_field(device, fixed):
    """J<=40 (N=1681) with an X+Z field: m is mixed, the two J-parity
    subspaces hold 861 and 820 states and dense columns span both (1681 rows,
    over a hundred workgroups).  'auto' routes them to the cooperative kernel."""
    h0, H, cos2, bas = linear_rotor(40)
    n = 1681
    X = _xz_columns(n, 2, 31)
    amps = np.linspace(4e8, 1.2e9, 5)
    fields = [[0.5 * a, 0.0, a] for a in amps]
    psi_d, k_d, _ = _device_run(h0, H, fields, X, kernel="auto", fixed=fixed)
    psi_o, k_o = _oracle_run(h0, H, fields, X, fixed=fixed or None)
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    if not fixed:
        assert np.array_equal(k_d, k_o)


def test_long_domain_one_parity(device):
    """A column inside one parity subspace (861 rows: a forced cooperative run
    on a domain that is a single component) next to a two-component column."""
    h0, H, cos2, bas = linear_rotor(40)
    n = 1681
    X = np.zeros((2, n), dtype=np.complex128)
    X[0, 0] = 1.0                       # |0,0>: even J only
    X[1, 3] = 0.8
    X[1, 7] = 0.6j
    fields = [[0.4e9, 0.0, a] for a in np.linspace(6e8, 1.4e9, 4)]
    psi_d, k_d, _ = _device_run(h0, H, fields, X, kernel="coop")
    psi_o, k_o = _oracle_run(h0, H, fields, X)
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    assert np.all(psi_d[psi_o == 0] == 0)
    assert np.array_equal(k_d, k_o)


@pytest.mark.parametrize("kernel", ["coop", "block"])
@pytest.mark.parametrize("fixed", [0, 12])
def test_forced_thermal_columns_jmax40(device, fixed, kernel):
    """The headline basis with basis-state columns (domains of <= 21 rows, a
    single workgroup each) and one-state domains that exhaust the fixed
    Krylov space and restart (tdse.py:459-465), all forced through the
    cooperative kernel."""
    h0, H, cos2, bas = linear_rotor(40)
    n = 1681
    rng = np.random.default_rng(21)
    idx = list(rng.choice(1500, size=4, replace=False)) + [1600, 1680]
    X = _basis_columns(n, idx)
    t = np.arange(28, 33) * 10.0 + 5.0
    fields = [[0.0, 0.0, gaussian_pulse_field(tt)] for tt in t]
    psi_d, k_d, _ = _device_run(h0, H, fields, X, kernel=kernel, fixed=fixed)
    psi_o, k_o = _oracle_run(h0, H, fields, X, fixed=fixed or None)
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    assert np.all(psi_d[psi_o == 0] == 0)
    if not fixed:
        assert np.array_equal(k_d, k_o)


@pytest.mark.parametrize("kernel", ["coop", "block"])
@pytest.mark.parametrize("fixed", [0, 6])
def test_forced_breakdown(device, fixed, kernel):
    """A diagonal operator: every state is its own subspace, beta == 0 and
    the ones-vector restart on one-row and sixteen-row domains."""
    from richmol_amd import rot
    bas = rot.linear_basis(0.2029, 3)
    h0 = rot.hamiltonian(bas)
    D = rot.hamiltonian(rot.linear_basis(7.3, 3))
    n = 16
    X = np.zeros((3, n), dtype=np.complex128)
    X[0, 1] = 1.0
    X[1, 5] = 1.0
    rng = np.random.default_rng(8)
    X[2] = rng.normal(size=n) + 1j * rng.normal(size=n)
    X[2] /= np.linalg.norm(X[2])
    fields = [[0.0, 0.0, 1.0]] * 3
    psi_d, k_d, _ = _device_run(h0, D, fields, X, kernel=kernel, fixed=fixed, dt=100.0)
    psi_o, k_o = _oracle_run(h0, D, fields, X, fixed=fixed or None, dt=100.0)
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    if not fixed:
        assert np.array_equal(k_d, k_o)


@pytest.mark.parametrize("kernel", ["coop", "block"])
def test_forced_unnormalised_and_zero_field_steps(device, kernel):
    """Rows of norm != 1 and zero-field steps (phases only, Krylov dimension
    0) through the long-domain kernels."""
    h0, H, cos2, bas = linear_rotor(6)
    n = 49
    rng = np.random.default_rng(5)
    X = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    X *= (np.array([0.2, 0.55, 1.7]) / np.linalg.norm(X, axis=1))[:, None]
    fields = [[0.0, 0.0, 0.0]] * 2 + [[0.0, 0.0, a] for a in np.linspace(8e8, 1.5e9, 4)] + [[0.0, 0.0, 0.0]]
    psi_d, k_d, _ = _device_run(h0, H, fields, X, kernel=kernel)
    psi_o, k_o = _oracle_run(h0, H, fields, X)
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    assert np.array_equal(k_d, k_o)
    assert np.all(k_d[:2] == 0) and np.all(k_d[-1] == 0)


@pytest.mark.parametrize("kernel", ["coop", "block"])
def test_long_kernels_agree_with_wave_kernel(device, kernel):
    """Forced long-domain and one-wave subspace kernels on 24 columns agree
    to rounding (different summation orders)."""
    h0, H, cos2, bas = linear_rotor(20)
    n = 441
    rng = np.random.default_rng(3)
    X = _basis_columns(n, rng.choice(n, size=24, replace=False))
    t = np.arange(20, 32) * 10.0 + 5.0
    fields = [[0.0, 0.0, gaussian_pulse_field(tt)] for tt in t]
    a, ka, _ = _device_run(h0, H, fields, X, kernel="subspace", fixed=12)
    b, kb, _ = _device_run(h0, H, fields, X, kernel=kernel, fixed=12)
    assert np.abs(a - b).max() <= 1e-12


def test_coop_fp32_long_domain(device):
    """FP32-matvec mode on the cooperative kernel (complex64 step operator and
    matvec input, FP32 accumulation, FP64 elsewhere): amplitudes within 2e-6
    and <cos^2 theta> within 1e-6 of the FP64 oracle."""
    from oracle.cartens import OracleCarTens
    h0, H, cos2, bas = linear_rotor(40)
    n = 1681
    X = _xz_columns(n, 2, 33)
    fields = [[0.5 * a, 0.0, a] for a in np.linspace(4e8, 1.2e9, 6)]
    psi_d, _, _ = _device_run(h0, H, fields, X, kernel="auto", fixed=12, precision="fp32")
    psi_o, _ = _oracle_run(h0, H, fields, X, fixed=12)
    C = OracleCarTens.from_attrs(cos2).tomat(form="full", cart="0")
    c2 = lambda P: np.real(np.einsum("bi,bi->b", P.conj(), (C @ P.T).T))
    assert np.abs(psi_d - psi_o).max() <= 2e-6
    assert np.abs(c2(psi_d) - c2(psi_o)).max() <= 1e-6
    assert np.abs(np.linalg.norm(psi_d, axis=1) - 1.0).max() <= 1e-6


@pytest.mark.parametrize("fixed", [0, 12])
def test_many_long_domains_block_kernel(device, fixed):
    """Ten dense columns on J<=40 with an X+Z field (every domain 1681 rows):
    more than the cooperative kernel takes, so 'auto' runs them on the block
    kernel, one workgroup each."""
    h0, H, cos2, bas = linear_rotor(40)
    n = 1681
    X = _xz_columns(n, 10, 35)
    fields = [[0.5 * a, 0.0, a] for a in np.linspace(5e8, 1.1e9, 3)]
    psi_d, k_d, _ = _device_run(h0, H, fields, X, kernel="auto", fixed=fixed)
    psi_o, k_o = _oracle_run(h0, H, fields, X, fixed=fixed or None)
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    if not fixed:
        assert np.array_equal(k_d, k_o)


def test_block_fp32_and_two_slots(device):
    """Block kernel in FP32-matvec mode with two field slots (X and Z
    products: the step operator array path) against FP64."""
    from oracle.cartens import OracleCarTens
    h0, H, cos2, bas = linear_rotor(30)
    n = 961
    X = _xz_columns(n, 4, 37)
    fields = [[0.5 * a, 0.0, a] for a in np.linspace(4e8, 1.2e9, 5)]
    psi_d, _, _ = _device_run(h0, H, fields, X, kernel="block", fixed=12, precision="fp32")
    psi_o, _ = _oracle_run(h0, H, fields, X, fixed=12)
    C = OracleCarTens.from_attrs(cos2).tomat(form="full", cart="0")
    c2 = lambda P: np.real(np.einsum("bi,bi->b", P.conj(), (C @ P.T).T))
    assert np.abs(psi_d - psi_o).max() <= 2e-6
    assert np.abs(c2(psi_d) - c2(psi_o)).max() <= 1e-6
