"""Launch choices of rm_propagate's subspace path (rm_plan_set_tuning) and
the autotuner over them: every choice gives the default's result (summation
order aside) and the oracle's, including a cooperative grid capped below the
domain's row groups (the memory form looping over them); the tuner times the
candidates that apply to the rows' domains and its choice propagates like
the default."""
import numpy as np
import pytest

from _fixtures import linear_rotor
from test_gpu_subspace import TOL_PSI, _device_run, _oracle_run

pytestmark = pytest.mark.gpu


def _run_tuned(h0, H, fields, X, tuning, fixed=0):
    import torch
    from richmol_amd import ensemble
    from richmol_amd.tdse import TDSE
    f = TDSE(t_end=10.0 * len(fields), dt=10.0, t_units="fs").exp_fac
    e = ensemble.Ensemble(H, h0, f, X, krylov_fixed=fixed)
    e.set_tuning(**tuning)
    table = e.prepare(fields)
    kused = torch.zeros((len(fields), len(X)), dtype=torch.int32, device=e.psi.device)
    e.propagate(None, table=table, kused=kused)
    out = e.psi.cpu().numpy(), kused.cpu().numpy()
    e.set_tuning()          # the plan is shared: back to the defaults
    return out


@pytest.mark.parametrize("tuning", [{"wave_max_rows": 1024}, {"wave_max_rows": 1024, "coop_min_entries": 0},
                                    {"wave_max_rows": 16}])
def test_mid_domain_choices_equal_oracle(device, tuning):
    """An asymmetric top on its m = 0 basis (J <= 12, Z field: one domain of
    ~50 rows for the ground state, one of ~40 for an excited state): the
    one-wave kernels up to 1024 rows, the long forms from 17 rows on, and the
    default -- each against the oracle, adaptive."""
    from richmol_amd import rot, units
    bas = rot.asymtop_basis(9.28, 27.88, 14.51, 12, m_filter=lambda J, m: m == 0)
    h0 = rot.hamiltonian(bas)
    H = rot.tensor(np.diag([9.14, 9.87, 9.45]), bas) * (-0.5 * units.AUpol_x_Vm_to_invcm())
    n = sum(sum(c.shape[1] for _, c in bas.kfun[J].values()) for J in bas.Jlist)
    X = np.zeros((2, n), dtype=np.complex128)
    X[0, 0] = 1.0
    X[1, 3] = 1.0
    fields = [[0.0, 0.0, a] for a in np.linspace(2e9, 6e9, 5)]
    base, kb = _run_tuned(h0, H, fields, X, {})
    got, kg = _run_tuned(h0, H, fields, X, tuning)
    psi_o, k_o = _oracle_run(h0, H, fields, X)
    assert np.abs(base - psi_o).max() <= TOL_PSI
    assert np.abs(got - psi_o).max() <= TOL_PSI
    assert np.array_equal(kg, k_o) and np.array_equal(kb, k_o)


@pytest.mark.parametrize("tuning", [{"pair": 0}, {"fork": 0}, {"pair": 0, "fork": 0}])
def test_wave_choices_equal_default(device, tuning):
    """J <= 12 all m, Z field, fixed Krylov dimension 12: pairable domains
    (<= 32 rows) with and without pairing, classes forked or in order."""
    h0, H, cos2, bas = linear_rotor(12)
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    rng = np.random.default_rng(5)
    X = rng.normal(size=(9, n)) + 1j * rng.normal(size=(9, n))
    X /= np.linalg.norm(X, axis=1)[:, None]
    fields = [[0.0, 0.0, a] for a in np.linspace(5e8, 1.5e9, 6)]
    base, kb = _run_tuned(h0, H, fields, X, {}, fixed=12)
    got, kg = _run_tuned(h0, H, fields, X, tuning, fixed=12)
    assert np.abs(got - base).max() <= 1e-12
    psi_o, _ = _oracle_run(h0, H, fields, X, fixed=12)
    assert np.abs(got - psi_o).max() <= TOL_PSI


@pytest.mark.parametrize("tuning", [{"coop_grid": 2}, {"coop_grid": 6}, {"coop_max": 0}, {"coop_max": 1}])
@pytest.mark.parametrize("fixed", [0, 12])
def test_long_domain_choices_equal_default(device, tuning, fixed):
    """J <= 40 with an X+Z field: two dense columns on 1681-row domains (27
    row groups of 64).  A cooperative grid of 2 or 6 workgroups (fewer waves
    than row groups: the memory form loops over them), or the one-workgroup
    form for both (coop_max 0, or 1 below their count) -- each equal to the
    default (cooperative: two domains of ~2e4 entries each) and to the
    oracle, same Krylov dimensions."""
    h0, H, cos2, bas = linear_rotor(40)
    rng = np.random.default_rng(31)
    X = rng.normal(size=(2, 1681)) + 1j * rng.normal(size=(2, 1681))
    X /= np.linalg.norm(X, axis=1)[:, None]
    fields = [[0.5 * a, 0.0, a] for a in np.linspace(4e8, 1.2e9, 4)]
    base, kb = _run_tuned(h0, H, fields, X, {}, fixed=fixed)
    got, kg = _run_tuned(h0, H, fields, X, tuning, fixed=fixed)
    assert np.abs(got - base).max() <= 1e-12
    assert np.array_equal(kg, kb)
    psi_o, k_o = _oracle_run(h0, H, fields, X, fixed=fixed or None)
    assert np.abs(got - psi_o).max() <= TOL_PSI
    if not fixed:
        assert np.array_equal(kg, k_o)


def test_tuning_argument_checks(device):
    from richmol_amd import _lib, plan as _plan
    h0, H, cos2, bas = linear_rotor(3)
    H.field([0.0, 0.0, 1e9])
    p = _plan.plan_for(H)
    for knob, value in ((_lib.TUNE_KNOBS["pair"], 2), (_lib.TUNE_KNOBS["fork"], -1),
                        (_lib.TUNE_KNOBS["coop_grid"], -3), (17, 0)):
        with pytest.raises(_lib.RichmolAmdError, match="rm_plan_set_tuning"):
            _lib.check(p.lib.rm_plan_set_tuning(p.handle, knob, value), "rm_plan_set_tuning")
    with pytest.raises(ValueError, match="unknown tuning choice"):
        p.set_tuning(waves=3)


def test_autotuner_over_production_choices(device):
    """The tuner on 24 Boltzmann-like basis columns at J <= 10 (fixed m 12):
    it times the subspace path with and without pairing, in order, and the
    full-vector ELL configurations; the ensemble then runs its choice and
    equals an untuned run."""
    from richmol_amd import autotune, ensemble
    from richmol_amd.tdse import TDSE
    h0, H, cos2, bas = linear_rotor(10)
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    X = np.zeros((24, n), dtype=np.complex128)
    X[np.arange(24), np.arange(24) * 3 % n] = 1.0
    fields = [[0.0, 0.0, a] for a in np.linspace(6e8, 1.4e9, 6)]
    f = TDSE(t_end=100.0, dt=10.0, t_units="fs").exp_fac
    e = ensemble.Ensemble(H, h0, f, X, krylov_fixed=12)
    (kern, tun), times = autotune.tune(e, fields, steps=6, reps=1, save=False)
    assert "auto" in times and "auto pair=0" in times and any(k.startswith("ell:") for k in times)
    assert e.kernel == "tuned" and e._tuned == (kern, tun)
    e.load(X)
    e.propagate(fields)
    ref, _, _ = _device_run(h0, H, fields, X, kernel="auto", fixed=12)
    assert np.abs(e.psi.cpu().numpy() - ref).max() <= 1e-12
    # the key carries the domain shape: 24 one-m domains of <= 32 rows
    shape = autotune.domain_shape(e.plan, X)
    assert shape[0] == 24 and shape[2] == 0
