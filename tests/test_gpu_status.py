"""Device checks of the call contract around rm_propagate and of the oracle
pins on the device:

* domain reuse: calls on the same rows reuse their domains; rows rewritten
  between calls are caught on the device (RM_STALE_DOMAIN, rows untouched),
  and Ensemble.propagate repeats such a call with fresh domains;
* the cooperative kernel's timeout path, forced from the host (the abort
  flag raised before the launch, so every bounded wait gives up at once and
  every kernel exits normally): status bit, Ensemble.check, and the return
  code when the caller passes no status;
* the full-vector kernels refuse a plan set to FP32;
* the thresholded field contraction on the device equals the hand-derived
  closed form of tests/test_oracle_pins.py;
* init_state rows (Boltzmann-weighted eigenvectors, diagonal and eigh
  branches) propagated on the device equal the oracle's.
"""
import numpy as np
import pytest

from _fixtures import gaussian_pulse_field, linear_rotor
from test_oracle_pins import FIELD, MF_THRESH, THRESH, hand_tensor

pytestmark = pytest.mark.gpu


def _oracle_steps(h0, H, X, fields, krylov=12, dt=10.0):
    from oracle.cartens import OracleCarTens
    from oracle.tdse import OracleTDSE
    Oh0, OH = OracleCarTens.from_attrs(h0), OracleCarTens.from_attrs(H)
    td = OracleTDSE(t_end=dt * (len(fields) + 1), dt=dt, t_units="fs")
    td.time_grid()
    for F in fields:
        OH.field(F)
        X, _ = td.update(OH, X, H0=Oh0, fixed_m=krylov)
    return X


def _fields(n, t0=300.0):
    return [[0.0, 0.0, gaussian_pulse_field(t0 + 10.0 * i)] for i in range(n)]


def test_domain_reuse_and_stale_rows(device):
    import torch
    from richmol_amd import _lib
    from richmol_amd.ensemble import Ensemble, basis_rows
    from richmol_amd.tdse import TDSE
    h0, H, _, bas = linear_rotor(8)
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    f = TDSE(t_end=100.0, dt=10.0, t_units="fs").exp_fac
    X = basis_rows(n, [0, 5, 17, 40, 80])
    fields = _fields(6)
    ens = Ensemble(H, h0, f, X, krylov_fixed=12)
    ens.propagate(fields[:3])
    ens.propagate(fields[3:])                      # domains reused
    want = _oracle_steps(h0, H, X, fields)
    assert np.abs(ens.psi.cpu().numpy() - want).max() <= 1e-10
    # rows rewritten behind the ensemble's back: |1> and |2> lie in other
    # subspaces than the kept domains of rows 0 and 1
    Y = basis_rows(n, [1, 2, 17, 40, 80])
    ens.psi.copy_(torch.from_numpy(Y).to(device))
    ens.propagate(fields[:2], check=False)
    st = ens.status.cpu().numpy()
    assert np.all(st & _lib.RM_STALE_DOMAIN)
    assert np.array_equal(ens.psi.cpu().numpy(), Y)          # left as it was
    with pytest.raises(RuntimeError, match="changed"):
        ens.check()
    # with the default check the call is repeated with fresh domains
    ens.propagate(fields[:2])
    assert not np.any(ens.status.cpu().numpy())
    assert np.abs(ens.psi.cpu().numpy() - _oracle_steps(h0, H, Y, fields[:2])).max() <= 1e-10
    # load() forgets the kept domains: no stale bit
    ens.load(X)
    ens.propagate(fields, check=False)
    assert not np.any(ens.status.cpu().numpy())
    assert np.abs(ens.psi.cpu().numpy() - want).max() <= 1e-10


def test_coop_timeout_path_forced(device):
    """Cooperative launches with the abort flag raised beforehand: the status
    bit, the Python errors, and a normal run afterwards."""
    import torch
    from richmol_amd import _lib
    from richmol_amd.ensemble import Ensemble, basis_rows
    from richmol_amd.tdse import TDSE
    h0, H, _, bas = linear_rotor(6)
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    f = TDSE(t_end=100.0, dt=10.0, t_units="fs").exp_fac
    X = basis_rows(n, [0, 3])
    fields = [[3e8, 0.0, F[2]] for F in _fields(3)]
    ens = Ensemble(H, h0, f, X, krylov_fixed=8)
    ens.set_kernel("coop")
    table = ens.prepare(fields)
    print("coop abort hook: J<=6, 2 rows, forced cooperative kernel", flush=True)
    ens.plan.set_coop_abort(True)
    try:
        ens.propagate(None, table=table, check=False)
        torch.cuda.synchronize()
        assert np.all(ens.status.cpu().numpy() & _lib.RM_DEVICE_TIMEOUT)
        with pytest.raises(RuntimeError, match="timed out"):
            ens.check()
        # no status array: the call itself reports the timeout (fresh rows:
        # the aborted run left garbage in them)
        ens.load(X)
        with pytest.raises(RuntimeError, match="bounded wait"):
            ens.plan.propagate(ens.psi, ens.out, table, f, phase=ens.phase, krylov_fixed=8, status=None)
    finally:
        ens.plan.set_coop_abort(False)
    ens.load(X)
    ens.propagate(None, table=table)
    assert np.abs(ens.psi.cpu().numpy() - _oracle_steps(h0, H, X, fields, krylov=8)).max() <= 1e-10


def test_full_vector_kernels_refuse_fp32(device):
    import torch
    from richmol_amd import _lib
    from richmol_amd.ensemble import Ensemble, basis_rows
    from richmol_amd.tdse import TDSE
    h0, H, _, bas = linear_rotor(4)
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    f = TDSE(t_end=100.0, dt=10.0, t_units="fs").exp_fac
    ens = Ensemble(H, h0, f, basis_rows(n, [0]), krylov_fixed=8)
    ens.set_precision("fp32")
    with pytest.raises(_lib.RichmolAmdError, match="FP64 only"):
        ens.step([0.0, 0.0, 1e9])
    with pytest.raises(_lib.RichmolAmdError, match="FP64 only"):
        ens.propagate([[0.0, 0.0, 1e9]], thresh=1e-3)
    # the subspace path runs FP32 as before
    ens.propagate([[0.0, 0.0, 1e9]])
    assert torch.isfinite(ens.psi).all()


def test_threshold_contraction_closed_form(device):
    """The device contraction with field-product and MF thresholds equals the
    hand-derived MF (K = 2): columns of H are H e_j."""
    import torch
    from richmol_amd.field import CarTens
    from richmol_amd.plan import DevicePlan
    T = hand_tensor(CarTens, kval=2.0)
    T.field(FIELD, thresh=THRESH)
    p = DevicePlan(T)
    p.set_field(T)
    x = torch.eye(3, dtype=torch.complex128, device=device)
    y = torch.empty_like(x)
    p.matvec(x, y)
    # row b of y = H e_b = column b of H
    got = y.cpu().numpy().T
    assert np.allclose(got, 2.0 * MF_THRESH, rtol=0, atol=1e-14)


@pytest.mark.parametrize("branch", ["diagonal", "eigh"])
def test_init_state_rows_on_device(device, branch):
    """Boltzmann rows of init_state (temp 30 K, J <= 6) from the product and
    from the oracle, each propagated four steps (device vs oracle)."""
    import torch
    from oracle.cartens import OracleCarTens
    from oracle.tdse import OracleTDSE
    from richmol_amd.tdse import TDSE
    h0, H, _, bas = linear_rotor(6)
    td = TDSE(t_end=100.0, dt=10.0, t_units="fs")
    to = OracleTDSE(t_end=100.0, dt=10.0, t_units="fs")
    td.time_grid()
    to.time_grid()
    if branch == "diagonal":
        rows, orows = td.init_state(h0, temp=30.0), to.init_state(OracleCarTens.from_attrs(h0), temp=30.0)
    else:
        # a static field on the polarizability makes H0 + V non-diagonal
        Hs = linear_rotor(6)[1]
        Hs.field([0.0, 0.0, 5e9])
        Os = OracleCarTens.from_attrs(Hs)
        Os.field([0.0, 0.0, 5e9])
        rows = td.init_state(Hs + h0, temp=30.0)
        orows = to.init_state(Os + OracleCarTens.from_attrs(h0), temp=30.0)
    assert rows.shape == orows.shape and rows.shape[0] > 3
    assert np.abs(rows - orows).max() <= 1e-12
    fields = _fields(4)
    psi = torch.from_numpy(rows).to(device)
    for F in fields:
        H.field(F)
        psi, _ = td.update(H, psi, H0=h0)
    want = _oracle_steps(h0, H, orows, fields, krylov=None)
    assert np.abs(psi.cpu().numpy() - want).max() <= 1e-10
