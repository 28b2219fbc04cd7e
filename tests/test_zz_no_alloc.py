"""rm_batch_propagate_static (the allocation-free batch call) against
rm_batch_propagate: the same jobs give bit-identical wavepackets, Krylov
dimensions and status, on two calls in a row -- the second one smaller --
with a fixed Krylov dimension (two wavepackets per wave, bases in LDS) and
with the adaptive rule (Gram matrices and Krylov bases in the scratch
array); and the entry point refuses what its fixed array cannot hold."""
import ctypes

import numpy as np
import pytest

from _fixtures import gaussian_pulse_field, linear_rotor

pytestmark = pytest.mark.gpu


def _batch_setup(jmax, rows, nsteps, device, fx=2e8):
    """Plan, device field table and step flags, phases and initial rows of a
    linear rotor under an X+Z field (fx != 0: two subspaces, J parity; every
    job on a wave that reads its operator at each matvec and keeps its Krylov
    basis in the workspace) or a Z field (chains of <= 9 states: two
    wavepackets per wave, bases in LDS)."""
    import torch
    from richmol_amd import _lib
    from richmol_amd.ensemble import Ensemble
    from richmol_amd.tdse import TDSE
    h0, H, _, bas = linear_rotor(jmax)
    f = TDSE(t_end=100.0, dt=10.0, t_units="fs").exp_fac
    fields = [[fx, 0.0, gaussian_pulse_field(300.0 + 10.0 * i)] for i in range(nsteps)]
    ens = Ensemble(H, h0, f, rows)
    table = ens.prepare(fields)
    lib = _lib.load()
    plan = ens.plan
    s = torch.cuda.current_stream().cuda_stream
    sflag = torch.empty(nsteps, dtype=torch.int32, device=device)
    _lib.check(lib.rm_plan_step_flags(plan.handle, nsteps, ctypes.c_void_p(table.data_ptr()),
                                      ctypes.c_void_p(sflag.data_ptr()), ctypes.c_void_p(s)), "step flags")
    lohi = np.zeros(2 * rows.shape[0], dtype=np.int32)
    _lib.check(lib.rm_plan_domains(plan.handle, ctypes.c_void_p(ens.psi.data_ptr()), ens.psi.stride(0),
                                   rows.shape[0], _lib.ptr(lohi, ctypes.c_int32), ctypes.c_void_p(s)), "domains")
    return dict(plan=plan, table=table, sflag=sflag, phase=ens.phase, f=f, lohi=lohi, rows=rows, ens=ens)


def _jobs(st, psi, kused, status, cols, nsteps):
    from richmol_amd.serving import JOB, job_array
    recs = []
    for j, b in enumerate(cols):
        rec = np.zeros(1, dtype=JOB)
        rec["plan"] = st["plan"].handle.value
        rec["psi"] = psi.data_ptr()
        rec["ld"] = psi.stride(0)
        rec["col"] = b
        rec["lo"], rec["len"] = st["lohi"][2 * b], st["lohi"][2 * b + 1] - st["lohi"][2 * b]
        rec["fprod_dev"] = st["table"].data_ptr()
        rec["sflag_dev"] = st["sflag"].data_ptr()
        rec["nsteps"] = nsteps
        rec["f_re"], rec["f_im"] = st["f"].real, st["f"].imag
        rec["phase"] = st["phase"].data_ptr()
        rec["kused"] = kused[j].data_ptr()
        rec["status"] = status[j:j + 1].data_ptr()
        recs.append(rec)
    return job_array(recs)


def _run(lib, st, cols, nsteps, kfix, static, batch_handle):
    import torch
    from richmol_amd import _lib
    psi = torch.from_numpy(st["rows"]).to("cuda")
    kused = torch.zeros((len(cols), nsteps), dtype=torch.int32, device="cuda")
    status = torch.zeros(len(cols), dtype=torch.int32, device="cuda")
    jobs = _jobs(st, psi, kused, status, cols, nsteps)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = (ctypes.c_void_p(jobs.ctypes.data), len(jobs), jobs.dtype.itemsize, kfix, 40, 1e-15, 0, s)
    if static:
        _lib.check(lib.rm_batch_propagate_static(*args), "rm_batch_propagate_static")
    else:
        _lib.check(lib.rm_batch_propagate(batch_handle, *args), "rm_batch_propagate")
    torch.cuda.synchronize()
    return psi.cpu().numpy(), kused.cpu().numpy(), status.cpu().numpy()


@pytest.mark.parametrize("fx", [2e8, 0.0])
@pytest.mark.parametrize("kfix", [12, 0])
def test_static_equals_batch_twice(device, kfix, fx):
    from richmol_amd import _lib
    from richmol_amd.ensemble import basis_rows
    lib = _lib.load()
    h = ctypes.c_void_p()
    _lib.check(lib.rm_batch_create(ctypes.byref(h)), "rm_batch_create")
    try:
        n = 17 * 17
        rng = np.random.default_rng(3)
        rows = basis_rows(n, [0, 1, 5, 12, 40, 99, 150, 200, 260, 288])
        dense = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
        rows = np.vstack([rows, dense / np.linalg.norm(dense, axis=1)[:, None]])
        st = _batch_setup(16, rows, 6, device, fx)
        print(f"static batch: linear rotor J<=16, fx={fx}, {rows.shape[0]} jobs, krylov_fixed={kfix}, domains "
              f"{sorted(set(np.diff(st['lohi'].reshape(-1, 2), axis=1).ravel().tolist()))}", flush=True)
        for cols, nsteps in ((list(range(rows.shape[0])), 6), ([0, 3, 10], 4)):
            a = _run(lib, st, cols, nsteps, kfix, False, h)
            b = _run(lib, st, cols, nsteps, kfix, True, h)
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
            assert not np.any(a[2])
            assert np.all(a[1] > 0)
    finally:
        lib.rm_batch_destroy(h)


def test_static_refuses_what_it_cannot_hold(device):
    import torch
    from richmol_amd import _lib
    lib = _lib.load()
    from richmol_amd.serving import JOB
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    many = np.zeros(_lib_max_jobs() + 1, dtype=JOB)
    rc = lib.rm_batch_propagate_static(ctypes.c_void_p(many.ctypes.data), len(many), many.dtype.itemsize,
                                       12, 40, 1e-15, 0, s)
    assert rc == -1
    assert b"exceed the maximum" in lib.rm_last_error()
    # a domain longer than the one-wave kernels take (J <= 40, X+Z, dense row: 1681 rows)
    n = 41 * 41
    rng = np.random.default_rng(0)
    row = rng.normal(size=(1, n)) + 0j
    st = _batch_setup(40, row / np.linalg.norm(row), 2, device)
    assert st["lohi"][1] - st["lohi"][0] == n
    psi = torch.from_numpy(st["rows"]).to("cuda")
    kused = torch.zeros((1, 2), dtype=torch.int32, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    jobs = _jobs(st, psi, kused, status, [0], 2)
    rc = lib.rm_batch_propagate_static(ctypes.c_void_p(jobs.ctypes.data), 1, jobs.dtype.itemsize, 12, 40,
                                       1e-15, 0, s)
    assert rc == -1 and b"needs the block kernel" in lib.rm_last_error()
    assert np.array_equal(psi.cpu().numpy(), st["rows"])


def _lib_max_jobs():
    import os
    import re
    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                            "richmol_amd.h")).read()
    return int(re.search(r"#define RM_STATIC_MAX_JOBS (\d+)", src).group(1))
