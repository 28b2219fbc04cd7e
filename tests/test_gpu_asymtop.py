"""Asymmetric top (water-like rotor of the asymmetric-top configuration:
static Z field on the dipole plus an X pulse on the polarizability) through
rm_propagate against the FP64 oracle: every amplitude to 1e-10 and the
adaptive Krylov dimension of every step equal to the oracle's.
"""
import numpy as np
import pytest

from _fixtures import water, water_fields
from oracle.cartens import OracleCarTens
from oracle.tdse import OracleTDSE

pytestmark = pytest.mark.gpu

TOL_PSI = 1e-10
DT = 5.0


def _summed(leaves, fields, i):
    for T, F in zip(leaves, fields):
        T.field(list(F[i]))
    return leaves[0] + leaves[1] + leaves[2]


def _oracle(h0, leaves, fields, X, fixed=None):
    Oh0 = OracleCarTens.from_attrs(h0)
    OL = [OracleCarTens.from_attrs(T) for T in leaves]
    nst = len(fields[0])
    to = OracleTDSE(t_start=0.0, t_end=DT * nst, dt=DT, t_units="fs")
    to.time_grid()
    psi, ks = X.copy(), []
    for i in range(nst):
        for O, F in zip(OL, fields):
            O.field(list(F[i]))
        stats = []
        psi, _ = to.update(OL[0] + OL[1] + OL[2], psi, H0=Oh0, fixed_m=fixed, stats=stats)
        ks.append(stats if stats else [0] * len(X))
    return psi, np.array(ks)


def _device(h0, leaves, fields, X, kernel="auto", fixed=0):
    import torch
    from richmol_amd import ensemble
    from richmol_amd.tdse import TDSE
    f = TDSE(t_end=DT * len(fields[0]), dt=DT, t_units="fs").exp_fac
    H = _summed(leaves, fields, 0)
    e = ensemble.Ensemble(H, h0, f, X, krylov_fixed=fixed)
    e.set_kernel(kernel)
    table = e.prepare(fields, per_term=True)
    kused = torch.zeros((len(fields[0]), len(X)), dtype=torch.int32, device=e.psi.device)
    e.propagate(None, table=table, kused=kused)
    e.check()
    return e.psi.cpu().numpy(), kused.cpu().numpy(), e


@pytest.mark.parametrize("kernel", ["auto", "ell", "global"])
def test_water_j6_adaptive(device, kernel):
    """J <= 6 (N = 455), ground state plus a superposition, 24 steps across
    the pulse rise; adaptive Lanczos."""
    h0, leaves, bas = water(6)
    n = sum(len(bas.mlist[J]) * sum(c.shape[1] for _, c in bas.kfun[J].values()) for J in bas.Jlist)
    X = np.zeros((2, n), dtype=np.complex128)
    X[0, 0] = 1.0
    rng = np.random.default_rng(4)
    X[1] = rng.normal(size=n) + 1j * rng.normal(size=n)
    X[1] /= np.linalg.norm(X[1])
    t = (np.arange(24) + 60.5) * DT
    fields = water_fields(t, fwhm_fs=60.0)
    psi_d, k_d, _ = _device(h0, leaves, fields, X, kernel=kernel)
    psi_o, k_o = _oracle(h0, leaves, fields, X)
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    assert np.array_equal(k_d, k_o)


@pytest.mark.parametrize("kernel", ["auto", "block"])
def test_water_j20_configuration(device, kernel):
    """The configuration's basis: J <= 20 with all m (N = 12 341, 1.5 M stored
    entries, four invariant subspaces of ~3 100 states), ground state, a few
    steps at the pulse peak: 'auto' takes the cooperative kernel (one
    wavepacket on a long domain), 'block' one workgroup reading its matvec
    input from memory (a domain longer than the block kernel stages in LDS)."""
    h0, leaves, bas = water(20)
    n = 12341
    X = np.zeros((1, n), dtype=np.complex128)
    X[0, 0] = 1.0
    t = (np.arange(4) + 79.5) * DT
    fields = water_fields(t)
    psi_d, k_d, e = _device(h0, leaves, fields, X, kernel=kernel)
    assert e.plan.n == n
    psi_o, k_o = _oracle(h0, leaves, fields, X)
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    assert np.array_equal(k_d, k_o)


@pytest.mark.parametrize("fixed", [0, 12])
@pytest.mark.parametrize("jmax", [4, 6])
def test_water_wide_operator_classes(device, jmax, fixed):
    """Invariant subspaces of 40-45 (J <= 4) and 112-119 (J <= 6) states with
    operator rows of up to 22 and 41 entries: the one- and two-rows-per-lane
    subspace classes that read the operator at every matvec.  One basis state
    per column (each column's domain is one subspace), adaptive and fixed
    Krylov dimension, against the oracle."""
    h0, leaves, bas = water(jmax)
    n = sum(len(bas.mlist[J]) * sum(c.shape[1] for _, c in bas.kfun[J].values()) for J in bas.Jlist)
    rng = np.random.default_rng(jmax)
    cols = np.sort(rng.choice(n, size=6, replace=False))
    X = np.zeros((len(cols), n), dtype=np.complex128)
    X[np.arange(len(cols)), cols] = 1.0
    t = (np.arange(10) + 75.5) * DT
    fields = water_fields(t, fwhm_fs=60.0)
    psi_d, k_d, _ = _device(h0, leaves, fields, X, fixed=fixed)
    psi_o, k_o = _oracle(h0, leaves, fields, X, fixed=fixed or None)
    assert np.abs(psi_d - psi_o).max() <= TOL_PSI
    if not fixed:
        assert np.array_equal(k_d, k_o)
