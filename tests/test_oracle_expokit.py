"""Pins of the oracle's Expokit restatement (oracle/expokit.py: ZGPADM
expokit.f:763-933, ZHEXPV expokit.f:2783-3170) by what the mathematics
fixes -- no reference test exercises propag='external', and the Fortran
extension is not built here:

* ZGPADM against closed forms (diagonal, nilpotent, rotation generator) and
  scipy's expm over sizes and time scales that need 0 to 10 squarings;
* the step-size rounding of expokit.f:2964-2965 on hand-derived values;
* ZHEXPV for a Hermitian A against exp(tA) v (both signs of t, a step count
  above one, a first step forced too long so that it is rejected, the happy
  breakdown on an invariant subspace smaller than m), and for A = iH as the
  reference calls it (tdse.py:383-391);
* the error exits (m >= n, more than mxstep steps);
* the oracle's TDSE.update(propag='external') on the OCS test basis against
  the internal Lanczos propagator and against dense expm.
"""
import numpy as np
import pytest
from scipy.linalg import expm

from oracle.expokit import ExpokitError, _round_step, zgpadm, zhexpv


def _herm(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return (A + A.conj().T) / 2, rng.normal(size=n) + 1j * rng.normal(size=n)


def test_zgpadm_closed_forms():
    d = np.array([0.3, -1.2, 2.0 + 0.5j])
    E, _ = zgpadm(np.diag(d), 1.7)
    assert np.abs(E - np.diag(np.exp(1.7 * d))).max() < 1e-14 * np.abs(E).max()
    E, ns = zgpadm(np.array([[0.0, 1.0], [0.0, 0.0]]), -3.0)
    assert np.abs(E - np.array([[1.0, -3.0], [0.0, 1.0]])).max() < 1e-14
    t = 2.5
    E, _ = zgpadm(np.array([[0.0, -1.0], [1.0, 0.0]]), t)
    R = np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])
    assert np.abs(E - R).max() < 1e-14


@pytest.mark.parametrize("m", [1, 3, 8, 14])
@pytest.mark.parametrize("t", [1e-3, -0.3, 2.0, 17.0])
def test_zgpadm_vs_expm(m, t):
    rng = np.random.default_rng(m)
    H = rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))
    E, ns = zgpadm(H, t)
    ref = expm(t * H)
    assert np.abs(E - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    # the scaling rule: ||t H / 2^ns||_inf < 1/2 (expokit.f:833-846)
    assert abs(t) * np.abs(H).sum(axis=1).max() / 2 ** ns < 0.5


def test_step_rounding_hand_values():
    # p1 = 10^(nint(log10(t) - sqrt(0.1)) - 1); t -> aint(t/p1 + 0.55) p1: two
    # significant digits for mantissas below 10^(1/2 + sqrt(0.1)) ~ 6.55, one
    # above (87 -> 90, 7.31e-4 -> 7e-4)
    for t, want in ((0.0123456, 0.012), (0.0155, 0.016), (3.4, 3.4), (0.5, 0.5), (87.0, 90.0),
                    (0.000731, 0.0007), (6.2, 6.2)):
        assert _round_step(t) == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("t", [-0.05, 0.02, -0.4])
def test_zhexpv_hermitian(t):
    A, v = _herm(120, 1)
    anorm = np.abs(A).sum(axis=0).max()
    st = {}
    w = zhexpv(v, anorm, 12, t, lambda x: A @ x, tol=1e-12, stats=st)
    ref = expm(t * A) @ v
    assert np.abs(w - ref).max() <= 1e-10 * np.abs(ref).max()
    assert st["nstep"] >= 2 and st["t_now"] == pytest.approx(t)


def test_zhexpv_rejected_first_step():
    """An anorm far below the true norm makes the first step too long: it is
    rejected and shortened, and the result stays within the tolerance."""
    A, v = _herm(80, 2)
    st = {}
    w = zhexpv(v, 0.01 * np.abs(A).sum(axis=0).max(), 12, -0.2, lambda x: A @ x, tol=1e-12, stats=st)
    assert st["nreject"] >= 1
    ref = expm(-0.2 * A) @ v
    assert np.abs(w - ref).max() <= 1e-10 * np.abs(ref).max()


def test_zhexpv_happy_breakdown():
    """v in a 3-dimensional invariant subspace: the recurrence breaks down at
    j = 3 < m and the result is exp(tA) v."""
    rng = np.random.default_rng(3)
    n = 40
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    lam = np.linspace(-2.0, 3.0, n)
    A = (Q * lam) @ Q.conj().T
    v = Q[:, :3] @ np.array([1.0, 0.5j, -0.25])
    st = {}
    w = zhexpv(v, np.abs(A).sum(axis=0).max(), 12, -0.7, lambda x: A @ x, tol=1e-12, stats=st)
    assert st["ibrkflag"] == 1 and st["mbrkdwn"] == 3
    assert np.abs(w - expm(-0.7 * A) @ v).max() < 1e-12


def test_zhexpv_anti_hermitian_as_called():
    """The reference's call: A = iH, t = Im(f) < 0 -- exp(t iH) v."""
    H, v = _herm(100, 4)
    st = {}
    w = zhexpv(v, np.abs(H).sum(axis=0).max(), 12, -0.05, lambda x: 1j * (H @ x), tol=1e-15, stats=st)
    ref = expm(-0.05j * H) @ v
    assert np.abs(w - ref).max() <= 1e-12 * np.abs(ref).max()
    assert abs(np.linalg.norm(w) - np.linalg.norm(v)) <= 1e-12 * np.linalg.norm(v)


def test_zhexpv_error_exits():
    A, v = _herm(10, 5)
    with pytest.raises(ExpokitError, match="bad input"):
        zhexpv(v, 1.0, 10, 0.1, lambda x: A @ x)
    with pytest.raises(ExpokitError, match="maximum number of steps"):
        zhexpv(v, np.abs(A).sum(axis=0).max(), 2, 1e4, lambda x: 1j * (A @ x), tol=1e-14)


def test_tdse_update_external_ocs():
    """OracleTDSE.update(propag='external') on the OCS test basis (16 states,
    a polarizability field step and the split-operator phases) equals the
    internal Lanczos propagator and dense expm within 1e-10."""
    from _fixtures import ocs_fixture
    from oracle import units as ounits
    from oracle.cartens import OracleCarTens
    from oracle.tdse import OracleTDSE
    h0, a, d = ocs_fixture()
    Oh0 = OracleCarTens.from_attrs(h0)
    OH = OracleCarTens.from_attrs(a) * (-0.5 * ounits.AUpol_x_Vm_to_invcm())
    OH.field([0.0, 0.0, 3e9])
    n = 16
    rng = np.random.default_rng(6)
    X = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    X /= np.linalg.norm(X, axis=1)[:, None]
    out = {}
    for prop in ("internal", "external"):
        td = OracleTDSE(t_end=1.0, dt=0.01, t_units="ps")
        td.time_grid()
        out[prop], _ = td.update(OH, X, H0=Oh0, propag=prop)
    f = td.exp_fac
    H = OH.tomat(form="full").toarray()
    P = np.exp(f / 2 * Oh0.tomat(form="full", cart="0").diagonal())
    ref = (P[None, :] * ((expm(f * H) @ (P[None, :] * X).T).T))
    assert np.abs(out["external"] - ref).max() <= 1e-10
    assert np.abs(out["external"] - out["internal"]).max() <= 1e-10
