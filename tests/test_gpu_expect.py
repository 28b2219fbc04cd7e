"""Observables read inside the propagation launch (rm_propagate_expect): the
loop of examples/ocs_alignment.py:89-103 -- TDSE.update, then
<psi_b|cos^2 theta|psi_b> of every vector and their sum -- with the host out
of it, against the oracle after every step.

* the linear-rotor configuration (J <= 30, all m, |0,0>, Z pulse, 1000 steps)
  in one launch, fixed Krylov dimension 12 (two wavepackets per wave, banded
  form) and the adaptive rule (one wave per wavepacket): <cos^2 theta>(t)
  within 1e-10 of the oracle fixture at every step;
* every launch form of the subspace path (one wave with 1, 2 and 4 rows per
  lane, the cooperative and the block kernel) and the full-vector kernels
  (field thresholds), on X+Z fields that mix m, with field-free steps among
  them: per step and per wavepacket within 1e-10 of the oracle;
* FP32 matvecs on a 64-state thermal ensemble at J <= 40 across the pulse
  peak: the ensemble <cos^2 theta>(t) within 1e-6 of the FP64 run, which is
  itself within 1e-10 of the oracle;
* the sums equal the per-wavepacket values, zero wavepackets give zeros, and
  the retried call after rows were rewritten (stale domains) gives the right
  values.
"""
import os

import numpy as np
import pytest

from _fixtures import GOLDEN, gaussian_pulse_field, linear_rotor

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _oracle(h0, H, C, X, fields, krylov=12, dt=10.0, thresh=None):
    """Oracle run: final rows and <psi_b|C|psi_b> after every step."""
    from oracle.cartens import OracleCarTens, flatten, unflatten
    from oracle.tdse import OracleTDSE
    Oh0, OH, OC = (OracleCarTens.from_attrs(T) for T in (h0, H, C))
    OC.field([0.0, 0.0, 1.0])
    td = OracleTDSE(t_end=dt * (len(fields) + 1), dt=dt, t_units="fs")
    td.time_grid()
    ex = np.zeros((len(fields), X.shape[0]), dtype=np.complex128)
    for s, F in enumerate(fields):
        OH.field(F, thresh=thresh)
        X, _ = td.update(OH, X, H0=Oh0, fixed_m=krylov)
        for b in range(X.shape[0]):
            ex[s, b] = np.vdot(X[b], flatten(OC, OC.vec(unflatten(OC, X[b]))))
    return X, ex


def _tdse_f():
    from richmol_amd.tdse import TDSE
    return TDSE(t_end=100.0, dt=10.0, t_units="fs").exp_fac


@pytest.mark.parametrize("mode", ["fixed", "adaptive"])
def test_config0_one_launch(device, mode):
    import torch
    from richmol_amd.ensemble import Ensemble
    from richmol_amd.tdse import TDSE
    gold = np.load(os.path.join(GOLDEN, "config0_oracle.npz"))
    nsteps = 1000
    h0, H, cos2, bas = linear_rotor(30)
    td = TDSE(t_end=10.0 * (nsteps + 1), dt=10.0, t_units="fs")
    tc = td.time_grid()
    fields = [[0.0, 0.0, gaussian_pulse_field(t)] for t in tc[:nsteps]]
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    psi0 = np.zeros((1, n), dtype=np.complex128)
    psi0[0, 0] = 1.0
    ens = Ensemble(H, h0, td.exp_fac, psi0, krylov_fixed=12 if mode == "fixed" else 0)
    cplan = Ensemble.observable_plan(cos2)
    print(f"config 0 one launch with <cos^2>, {mode}", flush=True)
    ex, exs = ens.propagate(fields, obs=cplan, expect=True, expect_sum=True)
    c2 = ex.cpu().numpy()[:, 0]
    d = np.abs(c2.real + 1.0 / 3.0 - gold[f"cos2_{mode}"]).max()
    print(f"linear J<=30, {mode}: max over 1000 steps |d<cos^2>| = {d:.2e}")
    assert d <= TOL
    assert np.abs(c2.imag).max() < 1e-14
    assert np.array_equal(exs.cpu().numpy(), c2)
    out = ens.psi.cpu().numpy()[0]
    assert np.abs(out[gold["idx"]] - gold[f"psi_{mode}"][-1]).max() <= TOL


def _xz_fields(nsteps, zero_at=(3, 4)):
    out = []
    for i in range(nsteps):
        fz = gaussian_pulse_field(300.0 + 10.0 * i, intensity=2e12)
        F = [0.6 * fz, 0.0, fz]
        if i in zero_at:
            F = [0.0, 0.0, 0.0]   # a field-free step: phases only
        out.append(F)
    return out


# (kernel, jmax, rows): domains of the linear rotor under X+Z fields are the
# J-parity halves (about N/2 rows, ELL width > 4): jmax 8 -> 1 row per lane
# (class 8), 12 -> 2 rows (class 9), 20 -> 4 rows (class 3); 'coop' and
# 'block' force the long-domain kernels
CASES = [("auto", 8, 12), ("auto", 12, 12), ("auto", 20, 12), ("auto", 12, 0),
         ("coop", 8, 12), ("block", 8, 12), ("block", 8, 0)]


@pytest.mark.parametrize("kernel,jmax,krylov", CASES)
def test_every_launch_form(device, kernel, jmax, krylov):
    from richmol_amd.ensemble import Ensemble, basis_rows
    h0, H, cos2, bas = linear_rotor(jmax)
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    rng = np.random.default_rng(jmax)
    X = basis_rows(n, [0, 1, 4, n - 1])
    X[1] = rng.normal(size=n) + 1j * rng.normal(size=n)   # both parities: one long domain
    X[1] /= np.linalg.norm(X[1])
    X = np.vstack([X, np.zeros((1, n))])                  # a zero row: zeros out
    fields = _xz_fields(8)
    ens = Ensemble(H, h0, _tdse_f(), X, krylov_fixed=krylov)
    ens.set_kernel("subspace" if kernel == "auto" else kernel)
    cplan = Ensemble.observable_plan(cos2)
    print(f"launch form {kernel}, J<={jmax}, krylov {krylov}", flush=True)
    ex, exs = ens.propagate(fields, obs=cplan, expect=True, expect_sum=True)
    got = ex.cpu().numpy()
    want_psi, want = _oracle(h0, H, cos2, X, fields, krylov=krylov or None)
    assert np.abs(ens.psi.cpu().numpy() - want_psi).max() <= TOL
    assert np.abs(got - want).max() <= TOL
    assert not np.any(got[:, -1])
    assert np.abs(exs.cpu().numpy() - got.sum(axis=1)).max() <= 1e-14


def test_full_vector_kernels_thresholds(device):
    """MF thresholds send the call to the full-vector kernels: one observable
    launch after every step."""
    from richmol_amd.ensemble import Ensemble, basis_rows
    h0, H, cos2, bas = linear_rotor(10)
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    X = basis_rows(n, [0, 2, 7])
    fields = _xz_fields(6)
    thresh = 1e3
    ens = Ensemble(H, h0, _tdse_f(), X, krylov_fixed=12)
    cplan = Ensemble.observable_plan(cos2)
    ex, exs = ens.propagate(fields, thresh=thresh, obs=cplan, expect=True, expect_sum=True)
    want_psi, want = _oracle(h0, H, cos2, X, fields, thresh=thresh)
    assert np.abs(ens.psi.cpu().numpy() - want_psi).max() <= TOL
    assert np.abs(ex.cpu().numpy() - want).max() <= TOL
    assert np.abs(exs.cpu().numpy() - want.sum(axis=1)).max() <= 1e-13


def test_sum_only_and_stale_retry(device):
    """Only the sums asked for (the plan's own per-wavepacket buffer); rows
    rewritten between calls: the repeated call's values are right."""
    import torch
    from richmol_amd.ensemble import Ensemble, basis_rows
    h0, H, cos2, bas = linear_rotor(8)
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    X = basis_rows(n, [0, 5, 17, 40])
    fields = [[0.0, 0.0, gaussian_pulse_field(300.0 + 10.0 * i)] for i in range(6)]
    ens = Ensemble(H, h0, _tdse_f(), X, krylov_fixed=12)
    cplan = Ensemble.observable_plan(cos2)
    _, s1 = ens.propagate(fields[:3], obs=cplan, expect_sum=True)
    _, s2 = ens.propagate(fields[3:], obs=cplan, expect_sum=True)
    _, want = _oracle(h0, H, cos2, X, fields)
    got = np.concatenate([s1.cpu().numpy(), s2.cpu().numpy()])
    assert np.abs(got - want.sum(axis=1)).max() <= TOL
    Y = basis_rows(n, [1, 2, 17, 40])
    ens.psi.copy_(torch.from_numpy(Y).to(device))
    ex, _ = ens.propagate(fields[:2], obs=cplan)        # stale, repeated with fresh domains
    _, want = _oracle(h0, H, cos2, Y, fields[:2])
    assert np.abs(ex.cpu().numpy() - want).max() <= TOL


def test_fp32_thermal_window(device):
    """Ensemble <cos^2 theta>(t) with FP32 matvecs within 1e-6 of FP64 over
    the 76 steps across the pulse peak (J <= 40, 64 states at 1 K); the FP64
    run's per-wavepacket values on 4 of them within 1e-10 of the oracle."""
    from richmol_amd import pulses
    from richmol_amd.ensemble import Ensemble, basis_rows, thermal_sample
    h0, H, cos2, bas = linear_rotor(40)
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    X = basis_rows(n, thermal_sample(bas, 1.0, 64, seed=3)) / 8.0
    fields = [[0.0, 0.0, pulses.envelope(5.0 + 10.0 * i, 300.0, 5e11)] for i in range(76)]
    cplan = Ensemble.observable_plan(cos2)
    sums = {}
    for prec in ("fp64", "fp32"):
        ens = Ensemble(H, h0, _tdse_f(), X, krylov_fixed=12)
        ens.set_precision(prec)
        ex, s = ens.propagate(fields, obs=cplan, expect=True, expect_sum=True)
        sums[prec] = s.cpu().numpy().real
        if prec == "fp64":
            _, want = _oracle(h0, H, cos2, X[:4], fields)
            assert np.abs(ex.cpu().numpy()[:, :4] - want).max() <= TOL
    d = np.abs(sums["fp32"] - sums["fp64"]).max()
    print(f"FP32 ensemble <cos^2>(t) over 76 steps: max |d| = {d:.2e}")
    assert 0.0 < d <= 1e-6


def test_tdse_run_equals_update_loop(device):
    """TDSE.run (one device call) against the reference-style loop of
    update() and <cos^2 theta> per step, and its time and call counter."""
    from richmol_amd.tdse import TDSE
    from oracle.cartens import OracleCarTens
    h0, H, cos2, bas = linear_rotor(10)
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    vecs0 = np.zeros((2, n), dtype=np.complex128)
    vecs0[0, 0] = 1.0
    vecs0[1, 5] = 1.0
    pulse = lambda t: [0.0, 0.0, gaussian_pulse_field(t)]
    ta = TDSE(t_start=200.0, t_end=400.0, dt=10.0, t_units="fs")
    ta.time_grid()
    vecs_run, t_run, ex = ta.run(H, vecs0, pulse, H0=h0, obs=cos2, nsteps=12, krylov_fixed=12)
    from oracle.cartens import flatten, unflatten
    OC = OracleCarTens.from_attrs(cos2)
    OC.field([0.0, 0.0, 1.0])
    cexp = lambda v: np.vdot(v, flatten(OC, OC.vec(unflatten(OC, v))))
    tb = TDSE(t_start=200.0, t_end=400.0, dt=10.0, t_units="fs")
    vecs = vecs0
    want = []
    for t in tb.time_grid()[:12]:
        H.field(pulse(t))
        vecs, t_ = tb.update(H, vecs, H0=h0, krylov_fixed=12)
        want.append(sum(cexp(v) for v in vecs))
    assert t_run == t_ == 320.0
    assert np.abs(vecs_run - vecs).max() <= TOL
    assert np.abs(ex - np.array(want)).max() <= TOL
    # the counter moved on: the next update returns the 13th step's time
    _, t13 = ta.update(H, vecs_run, H0=h0, krylov_fixed=12)
    assert t13 == 330.0
