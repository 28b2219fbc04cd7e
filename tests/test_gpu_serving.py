"""Serving loop (continuous batching through rm_batch_propagate) on the device.

Consistency bar: a request's final wavepacket is bit-for-bit the same when
it is served alone and when it is served in a mix of other requests (its
launch class, chunking and arithmetic depend only on the request itself);
against the FP64 oracle on the same Hamiltonian and pulse: 1e-10 absolute.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def server():
    from richmol_amd import serving
    return serving.Server(chunk=48, jmax_master=24)


def _reqs(n, seed, jmax=(10, 24), steps=(60, 240)):
    from richmol_amd import serving
    return serving.make_requests(n, seed=seed, jmax=jmax, steps=steps, fwhm=(40.0, 200.0))


def test_result_alone_equals_result_in_the_mix(device, server):
    mix = _reqs(40, seed=5)
    target = mix[7]
    from richmol_amd import serving
    alone = serving.Request(0, 0.0, target.B, target.jmax, target.fwhm, target.intensity, target.nsteps)
    a = server.run([alone])[0]
    done = {r.rid: r for r in server.run(mix)}
    b = done[target.rid]
    assert np.array_equal(a.psi, b.psi)
    assert np.array_equal(a.kused, b.kused)
    assert len(done) == len(mix) and server.used == 0
    for r in done.values():
        assert np.isfinite(r.psi).all()
        assert abs(np.linalg.norm(r.psi) - 1.0) < 1e-12


def test_request_matches_oracle(device, server):
    """One request through the server against the oracle's split-operator
    Lanczos (fixed Krylov dimension 12) on the same operator and pulse."""
    from oracle.cartens import OracleCarTens
    from oracle.tdse import OracleTDSE
    from richmol_amd import rot, serving, units
    req = serving.Request(0, 0.0, B=1.3, jmax=12, fwhm=60.0, intensity=3e11, nsteps=70)
    r = server.run([req])[0]
    bas = rot.linear_basis(req.B, req.jmax)
    h0 = rot.hamiltonian(bas)
    H = rot.tensor(np.diag(serving.LINEAR_ALPHA), bas)
    OH = OracleCarTens.from_attrs(H) * (-0.5 * units.AUpol_x_Vm_to_invcm())
    Oh0 = OracleCarTens.from_attrs(h0)
    td = OracleTDSE(t_end=req.nsteps * req.dt, dt=req.dt, t_units="fs")
    td.time_grid()
    v = np.zeros((1, (req.jmax + 1) ** 2), dtype=np.complex128)
    v[0, 0] = 1.0
    for F in req.fields():
        OH.field(F)
        v, _ = td.update(OH, v, H0=Oh0, fixed_m=12)
    assert np.abs(r.psi - v[0]).max() <= 1e-10


def test_admission_respects_hbm_budget(device):
    """A budget that holds three requests: all requests complete, never more
    than three resident at once, and the budget is released."""
    from richmol_amd import serving
    reqs = _reqs(10, seed=9)
    fp = max(serving.Server.footprint(None, r) for r in reqs)
    srv = serving.Server(chunk=32, jmax_master=24, hbm_budget=3 * fp)
    peak = []
    orig = srv._admit

    def spy(req, now):
        out = orig(req, now)
        peak.append(srv.used)
        return out

    srv._admit = spy
    done = srv.run(reqs)
    assert len(done) == 10 and srv.used == 0
    assert max(peak) <= 3 * fp
