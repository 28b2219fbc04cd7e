"""Invariant-subspace kernel (rm_propagate, one wave per wavepacket on the
union of the operator's connected components its initial state touches)
against the FP64 oracle on the same inputs.

Tolerances: 1e-10 absolute in every amplitude of the full wavepacket (the
entries outside the domain must come back exactly zero where the oracle's are)
and identical Krylov dimensions per step; the FP32-matvec precision mode is
held to the error bounds stated in its test.
"""
import numpy as np
import pytest

from _fixtures import gaussian_pulse_field, linear_rotor
from oracle.cartens import OracleCarTens
from oracle.tdse import OracleTDSE

pytestmark = pytest.mark.gpu

TOL_PSI = 1e-10


def _oracle_run(h0, H, fields, X, *, fixed=None, tol=1e-15, dt=10.0, split=True):
    Oh0 = OracleCarTens.from_attrs(h0)
    OH = OracleCarTens.from_attrs(H)
    nst = len(fields)
    to = OracleTDSE(t_start=0.0, t_end=dt * nst, dt=dt, t_units="fs")
    to.time_grid()
    psi = X.copy()
    ks = []
    for F in fields:
        OH.field(F)
        stats = []
        if split:
            psi, _ = to.update(OH, psi, H0=Oh0, tol=tol, fixed_m=fixed, stats=stats)
        else:
            psi, _ = to.update(Oh0 + OH, psi, tol=tol, fixed_m=fixed, stats=stats)
        ks.append(stats if stats else [0] * len(X))
    return psi, np.array(ks)


def _device_run(h0, H, fields, X, *, kernel="subspace", fixed=0, tol=1e-15, dt=10.0,
                precision="fp64"):
    import torch
    from richmol_amd import ensemble
    from richmol_amd.tdse import TDSE
    f = TDSE(t_end=dt * len(fields), dt=dt, t_units="fs").exp_fac
    e = ensemble.Ensemble(H, h0, f, X, krylov_fixed=fixed, tol=tol)
    e.set_kernel(kernel)
    table = e.prepare(fields)
    e.set_precision(precision)
    kused = torch.zeros((len(fields), len(X)), dtype=torch.int32, device=e.psi.device)
    e.propagate(None, table=table, kused=kused)
    e.check()
    return e.psi.cpu().numpy(), kused.cpu().numpy(), e


def _basis_columns(n, idx):
    X = np.zeros((len(idx), n), dtype=np.complex128)
    X[np.arange(len(idx)), idx] = 1.0
    return X


def test_subspaces_match_connected_components(device):
    """The plan's subspace count and size equal scipy's connected components
    of the flattened operator's pattern (Z field: (m, J parity) blocks)."""
    from scipy.sparse.csgraph import connected_components
    from richmol_amd import plan as _plan
    h0, H, cos2, bas = linear_rotor(10)
    H.field([0.0, 0.0, 1e9])
    p = _plan.plan_for(H)
    arr = _plan.flatten_operator(H, p.active)
    from scipy.sparse import csr_matrix
    A = csr_matrix((np.ones(arr["nnz"]), arr["col_idx"], arr["row_ptr"]), shape=(arr["n"], arr["n"]))
    nc, lab = connected_components(A, directed=False)
    count, mx = p.subspaces()
    assert count == nc
    assert mx == np.bincount(lab).max()
    assert count == 2 * 21 - 2   # m = -10..10, two J parities each, |m| = 10 has one


@pytest.mark.parametrize("fixed", [0, 12])
def test_thermal_columns_jmax40(device, fixed):
    """Headline basis (J<=40, N=1681), basis-state columns spread over m, a
    pulse window around the peak: every column's domain is one (m, parity)
    block of <= 21 states.  The last three columns start in one-state blocks
    (|J=40, m=-40>, |40, 40>, |39, -39>), so the fixed 12-dimensional Krylov
    space exhausts the domain and restarts (tdse.py:459-465) on it."""
    h0, H, cos2, bas = linear_rotor(40)
    n = 1681
    rng = np.random.default_rng(21)
    idx = list(rng.choice(1500, size=6, replace=False)) + [1600, 1680, 1521]
    X = _basis_columns(n, idx)
    t = np.arange(28, 36) * 10.0 + 5.0
    fields = [[0.0, 0.0, gaussian_pulse_field(tt)] for tt in t]
    psi_d, k_d, _ = _device_run(h0, H, fields, X, fixed=fixed)
    psi_o, k_o = _oracle_run(h0, H, fields, X, fixed=fixed or None)
    assert np.isfinite(psi_d).all()
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    assert np.all(psi_d[psi_o == 0] == 0)
    if not fixed:
        assert np.array_equal(k_d, k_o)


@pytest.mark.parametrize("precision", ["fp64"])
def test_dense_columns_full_domain(device, precision):
    """Random dense columns touch every component: the domain is the whole
    basis (N=169 <= 256 rows, 4 rows per lane)."""
    h0, H, cos2, bas = linear_rotor(12)
    n = 169
    rng = np.random.default_rng(22)
    X = rng.normal(size=(5, n)) + 1j * rng.normal(size=(5, n))
    X /= np.linalg.norm(X, axis=1)[:, None]
    fields = [[0.0, 0.0, a] for a in [3e8, 9e8, 1.4e9, 0.0, 1.1e9]]
    psi_d, k_d, _ = _device_run(h0, H, fields, X, precision=precision)
    psi_o, k_o = _oracle_run(h0, H, fields, X)
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    assert np.array_equal(k_d, k_o)


def test_xz_field_mixed_domains(device):
    """X+Z field couples m: the components are the two J parities; columns are
    a mix of basis states, a two-component superposition and zero rows
    (which stay zero, with Krylov dimension 0)."""
    h0, H, cos2, bas = linear_rotor(8)
    n = 81
    X = np.zeros((4, n), dtype=np.complex128)
    X[0, 0] = 1.0
    X[1, 5] = 0.6
    X[1, 10] = 0.8j
    X[3, 40] = 1.0
    amps = np.linspace(2e8, 1.6e9, 6)
    fields = [[0.6 * a, 0.0, a] for a in amps]
    psi_d, k_d, _ = _device_run(h0, H, fields, X)
    psi_o, k_o = _oracle_run(h0, H, fields, X)
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    assert np.all(psi_d[2] == 0)
    assert np.array_equal(k_d[:, [0, 1, 3]], k_o[:, [0, 1, 3]])


@pytest.mark.parametrize("fixed", [0, 6])
def test_exact_breakdown_single_state_domains(device, fixed):
    """A diagonal operator makes every state its own subspace: beta_1 == 0,
    the restart vector is built on a one-row domain, and the result and the
    Krylov dimensions still equal the reference's full-vector restart (fixed
    dimension: the two-wavepackets-per-wave kernel, one half breaking down
    while the other does not)."""
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
    psi_d, k_d, _ = _device_run(h0, D, fields, X, dt=100.0, fixed=fixed)
    psi_o, k_o = _oracle_run(h0, D, fields, X, dt=100.0, fixed=fixed or None)
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    if not fixed:
        assert np.array_equal(k_d, k_o)


@pytest.mark.parametrize("fixed", [0, 12])
def test_unnormalised_and_zero_field_steps(device, fixed):
    """Rows of norm != 1 (unnormalised first Krylov vector) and zero-field
    steps (Lanczos skipped, phases only, Krylov dimension 0)."""
    h0, H, cos2, bas = linear_rotor(6)
    n = 49
    rng = np.random.default_rng(5)
    X = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    X *= (np.array([0.2, 0.55, 1.7]) / np.linalg.norm(X, axis=1))[:, None]
    fields = [[0.0, 0.0, 0.0]] * 2 + [[0.0, 0.0, a] for a in np.linspace(8e8, 1.5e9, 4)] + [[0.0, 0.0, 0.0]]
    psi_d, k_d, _ = _device_run(h0, H, fields, X, fixed=fixed)
    psi_o, k_o = _oracle_run(h0, H, fields, X, fixed=fixed or None)
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    if not fixed:
        assert np.array_equal(k_d, k_o)
    else:
        assert np.all(k_d[2:-1] == fixed)
    assert np.all(k_d[:2] == 0) and np.all(k_d[-1] == 0)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_pair_kernel_result_independent_of_partner(device, precision):
    """Fixed Krylov dimension <= 16 runs two wavepackets per wave: nine
    thermal-like columns (an odd count: one empty half) propagated together,
    in another order, and one at a time give the same bits per column, and
    the FP64 run is within 1e-10 of the oracle."""
    h0, H, cos2, bas = linear_rotor(20)
    n = 441
    rng = np.random.default_rng(19)
    X = _basis_columns(n, rng.choice(120, size=9, replace=False))
    t = np.arange(25, 31) * 10.0 + 5.0
    fields = [[0.0, 0.0, gaussian_pulse_field(tt)] for tt in t]
    a, _, _ = _device_run(h0, H, fields, X, fixed=12, precision=precision)
    order = rng.permutation(9)
    b, _, _ = _device_run(h0, H, fields, X[order], fixed=12, precision=precision)
    assert np.array_equal(a[order], b)
    for c in (0, 4, 8):
        one, _, _ = _device_run(h0, H, fields, X[c:c + 1], fixed=12, precision=precision)
        assert np.array_equal(one[0], a[c])
    if precision == "fp64":
        psi_o, _ = _oracle_run(h0, H, fields, X, fixed=12)
        assert np.abs(a - psi_o).max() <= TOL_PSI


def test_pair_kernel_operator_complex_on_some_domains(device):
    """A Z-field operator made complex and non-Hermitian on the m = 0 chains
    only (H + 0.1j H restricted to m = 0): a wave pairs a domain that takes
    the real forms (real coefficients, real alpha, the real exponential) with
    one that takes the complex ones, decided per half.  Every column's bits
    equal its run alone (and in another order), and the oracle's sum of the
    two tensors to 1e-10."""
    import copy
    from scipy.sparse import csr_matrix
    h0, H, cos2, bas = linear_rotor(12)
    H = copy.deepcopy(H)
    # the m = 0 -> m = 0 entries of every M matrix (row and column index m + J)
    Hm = copy.deepcopy(H)
    for (J1, J2), a1 in Hm.mmat.items():
        for a2 in a1.values():
            for a3 in a2.values():
                for cart, m in list(a3.items()):
                    c = m.tocoo()
                    keep = (c.row == J1) & (c.col == J2)
                    a3[cart] = csr_matrix((c.data[keep], (c.row[keep], c.col[keep])), shape=m.shape)
    Hm = Hm * 0.1j
    H.field([0.0, 0.0, 1.0])
    Hm.field([0.0, 0.0, 1.0])
    Hc = H + Hm
    n = 169
    jm = [(0, 0), (1, 1), (2, 0), (3, -1), (4, 0), (2, 2), (6, 0), (5, -3)]
    X = _basis_columns(n, [J * J + J + m for J, m in jm])
    t = np.arange(25, 31) * 10.0 + 5.0
    fields = [[0.0, 0.0, gaussian_pulse_field(tt)] for tt in t]
    a, _, _ = _device_run(h0, Hc, fields, X, fixed=12)
    order = np.array([1, 0, 3, 2, 5, 4, 7, 6])
    b, _, _ = _device_run(h0, Hc, fields, X[order], fixed=12)
    assert np.array_equal(a[order], b)
    for c in (0, 1, 5):
        one, _, _ = _device_run(h0, Hc, fields, X[c:c + 1], fixed=12)
        assert np.array_equal(one[0], a[c])
    Oh0 = OracleCarTens.from_attrs(h0)
    OA, OB = OracleCarTens.from_attrs(H), OracleCarTens.from_attrs(Hm)
    to = OracleTDSE(t_start=0.0, t_end=10.0 * len(fields), dt=10.0, t_units="fs")
    to.time_grid()
    psi = X.copy()
    for F in fields:
        OA.field(F)
        OB.field(F)
        psi, _ = to.update(OA + OB, psi, H0=Oh0, fixed_m=12)
    assert np.abs(a - psi).max() <= TOL_PSI
    # the m = 0 columns did feel the complex part
    assert np.abs(np.abs(a[0]) - np.abs(_device_run(h0, H, fields, X[:1], fixed=12)[0][0])).max() > 1e-6


def test_auto_equals_subspace_and_ell_agrees(device):
    """'auto' takes the subspace kernel for rm_propagate (bitwise the same
    run) and agrees with the full-vector ELL kernel to rounding."""
    h0, H, cos2, bas = linear_rotor(20)
    n = 441
    rng = np.random.default_rng(3)
    X = _basis_columns(n, rng.choice(n, size=40, replace=False))
    t = np.arange(20, 40) * 10.0 + 5.0
    fields = [[0.0, 0.0, gaussian_pulse_field(tt)] for tt in t]
    a, ka, _ = _device_run(h0, H, fields, X, kernel="subspace", fixed=12)
    b, kb, _ = _device_run(h0, H, fields, X, kernel="auto", fixed=12)
    c, kc, _ = _device_run(h0, H, fields, X, kernel="ell", fixed=12)
    assert np.array_equal(a, b)
    assert np.abs(a - c).max() <= 1e-12


def test_subspace_refuses_thresholds_long_domains_go_cooperative(device):
    """Forced 'subspace' raises with MF thresholds; domains longer than a
    wave holds are advanced by the cooperative kernel (same result as the
    oracle)."""
    import torch
    from richmol_amd import ensemble
    from richmol_amd._lib import RichmolAmdError
    from richmol_amd.tdse import TDSE
    h0, H, cos2, bas = linear_rotor(40)
    n = 1681
    rng = np.random.default_rng(9)
    X = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    X /= np.linalg.norm(X, axis=1)[:, None]
    fields = [[0.3e9, 0.0, 1e9]] * 2     # X+Z: one parity block holds ~840 states, domain = all 1681
    f = TDSE(t_end=20.0, dt=10.0, t_units="fs").exp_fac
    e = ensemble.Ensemble(H, h0, f, X)
    e.set_kernel("subspace")
    table = e.prepare(fields, thresh=1e-3)
    with pytest.raises(RichmolAmdError):
        e.propagate(None, thresh=1e-3, table=table)
    psi_d, k_d, _ = _device_run(h0, H, fields, X, kernel="subspace")
    psi_o, k_o = _oracle_run(h0, H, fields, X)
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    assert np.array_equal(k_d, k_o)


def test_fp32_thermal_ensemble_cos2(device):
    """FP32-matvec mode on the headline workload's shape (J<=40, Boltzmann
    columns at 1 K, fixed m=12) through the pulse peak: every column's
    <cos^2 theta> within 1e-6 of the FP64 oracle after each block of steps,
    and the ensemble average too."""
    from oracle.cartens import OracleCarTens
    from richmol_amd import ensemble as _ens
    h0, H, cos2, bas = linear_rotor(40)
    n = 1681
    idx = _ens.thermal_sample(bas, 1.0, 48, seed=11)
    X = _basis_columns(n, idx)
    C = OracleCarTens.from_attrs(cos2).tomat(form="full", cart="0")
    c2 = lambda P: np.real(np.einsum("bi,bi->b", P.conj(), (C @ P.T).T)) + 1 / 3
    psi_d, psi_o = X.copy(), X.copy()
    for blk in range(3):
        t = (np.arange(8) + 26 + 8 * blk) * 10.0 + 5.0
        fields = [[0.0, 0.0, gaussian_pulse_field(tt)] for tt in t]
        psi_d, _, _ = _device_run(h0, H, fields, psi_d, kernel="auto", fixed=12, precision="fp32")
        psi_o, _ = _oracle_run(h0, H, fields, psi_o, fixed=12)
        err = np.abs(c2(psi_d) - c2(psi_o))
        assert err.max() <= 1e-6, (blk, err.max())
        assert abs(c2(psi_d).mean() - c2(psi_o).mean()) <= 1e-6


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_pair_kernel_unbanded_domains(device, precision):
    """A dipole plus a polarizability on a linear rotor under a Z field: each m
    block (J = |m|..20, at most 21 rows) couples J to J +- 1 and J +- 2, so its
    domains fit half a wave but are not banded -- the two-per-wave kernel with
    the LDS gather; FP64 against the oracle (1e-10), FP32 partner independent."""
    import torch
    from richmol_amd import ensemble, rot, units
    from richmol_amd.tdse import TDSE
    from oracle.cartens import OracleCarTens
    from oracle.tdse import OracleTDSE
    h0, A, cos2, bas = linear_rotor(20)
    D = rot.tensor(np.array([0.0, 0.0, 0.3]), bas) * (-units.AUdip_x_Vm_to_invcm())
    n = 441
    rng = np.random.default_rng(41)
    X = _basis_columns(n, rng.choice(200, size=7, replace=False))
    t = np.arange(25, 31) * 10.0 + 5.0
    pul = np.array([[0.0, 0.0, gaussian_pulse_field(tt)] for tt in t])
    stat = np.tile([[0.0, 0.0, 2e7]], (len(t), 1))
    fields = [stat, pul]
    f = TDSE(t_end=10.0 * len(t), dt=10.0, t_units="fs").exp_fac

    def dev(Xr):
        D.field(list(stat[0]))
        A.field(list(pul[0]))
        e = ensemble.Ensemble(D + A, h0, f, Xr, krylov_fixed=12)
        e.set_precision(precision)
        e.propagate(None, table=e.prepare(fields, per_term=True))
        e.check()
        return e.psi.cpu().numpy()

    a = dev(X)
    one = dev(X[2:3])
    assert np.array_equal(one[0], a[2])
    if precision == "fp64":
        OL = [OracleCarTens.from_attrs(D), OracleCarTens.from_attrs(A)]
        Oh0 = OracleCarTens.from_attrs(h0)
        to = OracleTDSE(t_start=0.0, t_end=10.0 * len(t), dt=10.0, t_units="fs")
        to.time_grid()
        psi = X.copy()
        for i in range(len(t)):
            for O, F in zip(OL, fields):
                O.field(list(F[i]))
            psi, _ = to.update(OL[0] + OL[1], psi, H0=Oh0, fixed_m=12)
        assert np.abs(a - psi).max() <= TOL_PSI
