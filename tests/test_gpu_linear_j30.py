"""The linear-rotor configuration on the device: J <= 30 with all m (N = 961),
psi0 = |J=0, m=0>, cycle-averaged Gaussian Z pulse (FWHM 300 fs, 5e11 W/cm^2,
centred at 375 fs), dt = 10 fs, 1000 steps, split operator, FP64 -- through
the production rm_propagate path, with a fixed Krylov dimension 12 and with
the reference's adaptive rule, against the oracle's trajectories
(tests/golden/config0_oracle.npz, written by make_config0_fixture.py; its
first steps are re-derived from the live oracle in test_golden_config0.py).

Bar: every amplitude and <cos^2 theta> within 1e-10 of the oracle after every
step of the run, amplitudes outside the reachable states exactly zero, and
(adaptive) the Krylov dimension of every step equal to the oracle's.
"""
import os

import numpy as np
import pytest

from _fixtures import GOLDEN, gaussian_pulse_field, linear_rotor

pytestmark = pytest.mark.gpu
NSTEPS = 1000
TOL = 1e-10


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLDEN, "config0_oracle.npz"))


@pytest.fixture(scope="module")
def system(device):
    from richmol_amd.plan import plan_for
    from richmol_amd.tdse import TDSE
    h0, H, cos2, bas = linear_rotor(30)
    td = TDSE(t_end=10.0 * (NSTEPS + 1), dt=10.0, t_units="fs")
    tc = td.time_grid()
    fields = [[0.0, 0.0, gaussian_pulse_field(t)] for t in tc[:NSTEPS]]
    cos2.field([0, 0, 1])
    cplan = plan_for(cos2)
    cplan.set_field(cos2)
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    psi0 = np.zeros((1, n), dtype=np.complex128)
    psi0[0, 0] = 1.0
    return h0, H, cplan, fields, psi0, td.exp_fac


@pytest.mark.parametrize("mode", ["fixed", "adaptive"])
def test_one_launch_1000_steps(system, gold, mode):
    """All 1000 steps in one rm_propagate call: final wavepacket and the
    Krylov dimension of every step against the oracle."""
    import torch
    from richmol_amd.ensemble import Ensemble
    h0, H, cplan, fields, psi0, f = system
    ens = Ensemble(H, h0, f, psi0, krylov_fixed=12 if mode == "fixed" else 0)
    kused = torch.zeros((NSTEPS, 1), dtype=torch.int32, device=ens.psi.device)
    ens.propagate(fields, kused=kused)
    out = ens.psi.cpu().numpy()[0]
    idx = gold["idx"]
    assert np.abs(out[idx] - gold[f"psi_{mode}"][-1]).max() <= TOL
    assert not np.any(np.delete(out, idx))
    assert np.array_equal(kused.cpu().numpy()[:, 0], gold[f"kused_{mode}"])
    assert abs(np.linalg.norm(out) - 1.0) < 1e-12
    print(f"linear J<=30, {mode}: 1000 steps in one launch, |dpsi| = "
          f"{np.abs(out[idx] - gold[f'psi_{mode}'][-1]).max():.2e}")


@pytest.mark.parametrize("mode", ["fixed", "adaptive"])
def test_every_step(system, gold, mode):
    """One step per rm_propagate call (the drop-in time loop), the wavepacket
    and <cos^2 theta> compared after every step."""
    import torch
    from richmol_amd.ensemble import Ensemble
    h0, H, cplan, fields, psi0, f = system
    ens = Ensemble(H, h0, f, psi0, krylov_fixed=12 if mode == "fixed" else 0)
    table = ens.prepare(fields)
    idx = gold["idx"]
    psi_dev = torch.zeros((NSTEPS, len(idx)), dtype=torch.complex128, device=ens.psi.device)
    c2_dev = torch.zeros(NSTEPS, dtype=torch.complex128, device=ens.psi.device)
    kused = torch.zeros((1, 1), dtype=torch.int32, device=ens.psi.device)
    kall = torch.zeros(NSTEPS, dtype=torch.int32, device=ens.psi.device)
    didx = torch.from_numpy(idx).to(ens.psi.device)
    for i in range(NSTEPS):
        ens.propagate(None, table=table[i:i + 1], kused=kused, check=(i % 100 == 0))
        psi_dev[i] = ens.psi[0, didx]
        c2_dev[i] = ens.expect(cplan)[0]
        kall[i] = kused[0, 0]
    ens.check()
    out = ens.psi.cpu().numpy()[0]
    assert not np.any(np.delete(out, idx))
    dpsi = np.abs(psi_dev.cpu().numpy() - gold[f"psi_{mode}"]).max()
    dc2 = np.abs(c2_dev.real.cpu().numpy() + 1.0 / 3.0 - gold[f"cos2_{mode}"]).max()
    print(f"linear J<=30, {mode}: every step, max |dpsi| = {dpsi:.2e}, max |dcos2| = {dc2:.2e}")
    assert dpsi <= TOL
    assert dc2 <= TOL
    assert np.array_equal(kall.cpu().numpy(), gold[f"kused_{mode}"])
    # the alignment peaks after the pulse and the field-free revival structure follows
    assert gold[f"cos2_{mode}"].max() > 0.5
