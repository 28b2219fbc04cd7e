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


def _asym_oracle(req, fixed=12, nsteps=None):
    """The oracle's split-operator Lanczos on the general builder's operator
    for an asymmetric-top request (m = 0 block, Z pulse, J = 0 start)."""
    from oracle.cartens import OracleCarTens
    from oracle.tdse import OracleTDSE
    from richmol_amd import rot, serving, units
    a, b, c = req.abc
    bas = rot.asymtop_basis(b, c, a, req.jmax, m_filter=lambda J, m: m == 0)
    h0 = rot.hamiltonian(bas)
    H = rot.tensor(np.diag(serving.ASYM_ALPHA) if req.alpha is None else req.alpha, bas)
    OH = OracleCarTens.from_attrs(H) * (-0.5 * units.AUpol_x_Vm_to_invcm())
    Oh0 = OracleCarTens.from_attrs(h0)
    F = req.fields()[: nsteps or req.nsteps]
    td = OracleTDSE(t_end=len(F) * req.dt, dt=req.dt, t_units="fs")
    td.time_grid()
    v = np.zeros((1, req.n), dtype=np.complex128)
    v[0, 0] = 1.0
    for f in F:
        OH.field(f)
        v, _ = td.update(OH, v, H0=Oh0, fixed_m=fixed)
    return v[0]


def test_asym_request_matches_oracle(device, server):
    """An asymmetric-top request (its own m = 0 operator, built by the
    serving path's direct builder) against the oracle on the general
    builder's operator, FP64: 1e-10."""
    from richmol_amd import serving
    req = serving.Request(0, 0.0, B=0.0, jmax=14, fwhm=60.0, intensity=3e11, nsteps=70, kind="asym",
                          abc=(21.0, 7.5, 3.2))
    r = server.run([req])[0]
    assert np.abs(r.psi - _asym_oracle(req)).max() <= 1e-10
    assert server.used == 0


def test_mixed_stream_asym_alone_equals_mix(device, server):
    """Linear and asymmetric-top requests in one stream: an asymmetric-top
    request's result is bit-for-bit the same alone and in the mix."""
    from richmol_amd import serving
    mix = serving.make_requests(30, seed=11, jmax=(10, 24), steps=(60, 200), fwhm=(40.0, 200.0),
                                asym_fraction=0.5)
    target = next(r for r in mix if r.kind == "asym")
    alone = serving.Request(0, 0.0, target.B, target.jmax, target.fwhm, target.intensity, target.nsteps,
                            kind="asym", abc=target.abc)
    a = server.run([alone])[0]
    done = {r.rid: r for r in server.run(mix)}
    b = done[target.rid]
    assert np.array_equal(a.psi, b.psi) and np.array_equal(a.kused, b.kused)
    assert len(done) == len(mix) and server.used == 0
    assert sum(r.kind == "asym" for r in mix) >= 10
    for r in done.values():
        assert abs(np.linalg.norm(r.psi) - 1.0) < 1e-12


def test_fp32_server_close_to_fp64(device):
    """The serving precision (complex64 matvec, FP64 Krylov): a linear and an
    asymmetric-top request within 1e-6 of the FP64 oracle in every amplitude
    and in <J=0| population."""
    from richmol_amd import serving
    srv = serving.Server(chunk=48, jmax_master=24, precision="fp32")
    req = serving.Request(0, 0.0, B=0.0, jmax=16, fwhm=60.0, intensity=5e11, nsteps=80, kind="asym",
                          abc=(18.0, 6.0, 2.5))
    r = srv.run([req])[0]
    v = _asym_oracle(req)
    assert np.abs(r.psi - v).max() <= 1e-6
    assert abs(abs(r.psi[0]) ** 2 - abs(v[0]) ** 2) <= 1e-6


def test_long_domain_job_block_kernel(device, server):
    """An asymmetric top whose polarizability has off-diagonal elements
    (species mixed) at Jmax 50: the domain of J = 0 (even J of the m = 0
    block) has more rows than a wave holds, so the batch runs it on the
    block kernel; it matches the oracle over 16 steps and is the same alone
    and next to short jobs."""
    from richmol_amd import serving
    alpha = np.array([[12.0, 1.5, -0.8], [1.5, 8.0, 0.6], [-0.8, 0.6, 16.0]])
    kw = dict(B=0.0, jmax=50, fwhm=50.0, intensity=4e11, nsteps=16, kind="asym", abc=(9.0, 4.0, 1.5),
              alpha=alpha)
    req = serving.Request(0, 0.0, **kw)
    m = serving._AsymModel(req)
    assert m.len > 1024
    a = server.run([req])[0]
    assert np.abs(a.psi - _asym_oracle(req)).max() <= 1e-10
    short = serving.make_requests(6, seed=3, jmax=(10, 20), steps=(16, 40), fwhm=(40.0, 100.0))
    done = {r.rid: r for r in server.run(short + [serving.Request(99, 0.0, **kw)])}
    assert np.array_equal(done[99].psi, a.psi)
