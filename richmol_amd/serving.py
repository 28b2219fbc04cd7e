"""Continuous-batching propagation server.

A request is one field-driven propagation: a rotor Hamiltonian, a pulse, a
number of steps, an initial state.  The server keeps every admitted request's
wavepacket, phases and field table resident on the device and advances all
active requests together, ``chunk`` steps per tick, with one
``rm_batch_propagate`` call (one wave per request on its invariant subspace,
jobs of different operators in the same launch).  Requests join at tick
boundaries as they arrive -- admitted while slots and the HBM budget allow --
and leave when their last step is done; host-side admission of new requests
overlaps the device work of the running tick.

Latency is measured from arrival to the completion of the tick that finished
the request.  ``benchmark`` drives a seeded Poisson stream of requests drawn
as the serving configuration describes (linear rotors: B U(0.05, 30) cm^-1,
Jmax U{10..60}, pulse FWHM U(100, 1000) fs, peak intensity log-U(1e10, 1e12)
W/cm^2, 200-2000 steps of 10 fs, psi0 = |J=0, m=0>) at 0.8 of the capacity
measured on the same stream in a closed-loop run.
"""
import ctypes
import time
from collections import deque

import numpy as np

from . import _lib, rot, units

LINEAR_ALPHA = (26.0, 26.0, 53.6)   # alpha_perp, alpha_perp, alpha_par [au]
DT_FS = 10.0


def _job_dtype():
    f = _lib.RmJob
    names = [n for n, _ in f._fields_]
    fmt = {ctypes.c_void_p: np.uint64, ctypes.c_int64: np.int64, ctypes.c_int32: np.int32,
           ctypes.c_double: np.float64}
    return np.dtype({"names": names, "formats": [fmt[t] for _, t in f._fields_],
                     "offsets": [getattr(f, n).offset for n in names], "itemsize": ctypes.sizeof(f)})


JOB = _job_dtype()


def job_array(recs):
    """Concatenate job records keeping the rm_job layout (numpy's structured
    promotion would otherwise drop the padding and shift every field after
    the first gap)."""
    if not recs:
        return np.zeros(0, dtype=JOB)
    return np.concatenate(recs, dtype=JOB)


def batch_propagate(lib, handle, jobs, krylov_fixed, maxorder, tol, precision, stream):
    """rm_batch_propagate on a JOB array (layout checked here and by the
    library through the record size)."""
    if jobs.dtype != JOB or jobs.dtype.itemsize != ctypes.sizeof(_lib.RmJob) or not jobs.flags["C_CONTIGUOUS"]:
        raise TypeError(f"job array must be C-contiguous with the rm_job layout, got {jobs.dtype}")
    _lib.check(lib.rm_batch_propagate(handle, ctypes.c_void_p(jobs.ctypes.data), len(jobs),
                                      jobs.dtype.itemsize, krylov_fixed, maxorder, tol, precision,
                                      ctypes.c_void_p(stream)), "rm_batch_propagate")


class Request:
    """One propagation job.  ``pulse``: (fwhm_fs, intensity W/cm^2), a Z
    Gaussian envelope centred at 4 fwhm and cut beyond 4 fwhm (as in the
    benchmark workload); ``B``: rotational constant [cm^-1]."""

    __slots__ = ("rid", "arrival", "kind", "B", "jmax", "fwhm", "intensity", "nsteps", "dt",
                 "t_admit", "t_done", "psi", "steps_done", "kused")

    def __init__(self, rid, arrival, B, jmax, fwhm, intensity, nsteps, kind="linear", dt=DT_FS):
        self.rid, self.arrival, self.kind = rid, float(arrival), kind
        self.B, self.jmax, self.fwhm, self.intensity = float(B), int(jmax), float(fwhm), float(intensity)
        self.nsteps, self.dt = int(nsteps), float(dt)
        self.t_admit = self.t_done = None
        self.psi = None
        self.steps_done = 0

    def fields(self):
        """(nsteps, 3) Z field at the step midpoints [V/m]."""
        t = (np.arange(self.nsteps) + 0.5) * self.dt
        e0 = units.intensity_to_field(self.intensity)
        t0 = 4.0 * self.fwhm
        ez = e0 * np.exp(-2 * np.log(2) * (t - t0) ** 2 / self.fwhm ** 2) / np.sqrt(2.0)
        ez[np.abs(t - t0) > 4.0 * self.fwhm] = 0.0
        F = np.zeros((self.nsteps, 3))
        F[:, 2] = ez
        return F


def make_requests(n, seed=0, rate=None, jmax=(10, 60), B=(0.05, 30.0), fwhm=(100.0, 1000.0),
                  intensity=(1e10, 1e12), steps=(200, 2000)):
    """Seeded request stream; Poisson arrivals at ``rate`` per second (all at
    t = 0 when rate is None)."""
    rng = np.random.default_rng(seed)
    arr = np.zeros(n) if rate is None else np.cumsum(rng.exponential(1.0 / rate, size=n))
    out = []
    for i in range(n):
        out.append(Request(i, arr[i], B=rng.uniform(*B), jmax=rng.integers(jmax[0], jmax[1] + 1),
                           fwhm=rng.uniform(*fwhm),
                           intensity=10 ** rng.uniform(np.log10(intensity[0]), np.log10(intensity[1])),
                           nsteps=rng.integers(steps[0], steps[1] + 1)))
    return out


class _Model:
    """Operator, device plan and the |J=0,m=0> domain shared by every
    linear-rotor request with this Jmax (B enters only the phases)."""

    def __init__(self, master, jmax):
        from .plan import DevicePlan, field_table
        import torch
        self.jmax = jmax
        self.H = rot.restrict(master, jmax) * (-0.5 * units.AUpol_x_Vm_to_invcm())
        self.n = (jmax + 1) ** 2
        self.J = np.repeat(np.arange(jmax + 1), 2 * np.arange(jmax + 1) + 1).astype(np.float64)
        active, _, _ = field_table(self.H, np.array([[0.0, 0.0, 1.0]]))
        self.plan = DevicePlan(self.H, active)
        self.nslot = self.plan.nfield
        psi0 = torch.zeros((1, self.n), dtype=torch.complex128, device="cuda")
        psi0[0, 0] = 1.0
        lohi = np.zeros(2, dtype=np.int32)
        _lib.check(self.plan.lib.rm_plan_domains(self.plan.handle, ctypes.c_void_p(psi0.data_ptr()),
                                                 self.n, 1, _lib.ptr(lohi, ctypes.c_int32), None),
                   "rm_plan_domains")
        self.lo, self.len = int(lohi[0]), int(lohi[1] - lohi[0])

    def table(self, F):
        from .plan import field_table
        _, cols, nst = field_table(self.H, F)
        return self.plan.table_array(cols, nst)


class Server:
    """Continuous-batching engine on one device stream.

    hbm_budget: bytes of device memory the resident request state may use
    (admission stops at the first queued request that does not fit);
    max_jobs: concurrent requests; chunk: steps per tick per request;
    krylov_fixed: fixed Krylov dimension (0: adaptive with tol, maxorder).
    """

    def __init__(self, hbm_budget=8 << 30, max_jobs=8192, chunk=64, krylov_fixed=12, maxorder=100,
                 tol=1e-15, precision="fp64", jmax_master=60):
        import torch
        self.torch = torch
        self.lib = _lib.load()
        self.budget, self.max_jobs, self.chunk = int(hbm_budget), int(max_jobs), int(chunk)
        self.kfix, self.maxorder, self.tol = int(krylov_fixed), int(maxorder), float(tol)
        self.precision = _lib.PRECISIONS[precision]
        self.f = -1j * DT_FS * 1e-15 * units.enr_units()["invcm"] / units.HBAR
        self.stream = torch.cuda.current_stream()
        h = ctypes.c_void_p()
        _lib.check(self.lib.rm_batch_create(ctypes.byref(h)), "rm_batch_create")
        self.handle = h
        bas = rot.linear_basis(1.0, jmax_master)
        self.master = rot.tensor(np.diag(LINEAR_ALPHA), bas)
        self.models = {}
        self.used = 0

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self.lib.rm_batch_destroy(h)
            except Exception:
                pass

    def model(self, jmax):
        m = self.models.get(jmax)
        if m is None:
            m = self.models[jmax] = _Model(self.master, jmax)
        return m

    def footprint(self, req):
        """Resident device bytes of a request: wavepacket, phases, field table,
        step flags, Krylov dimensions."""
        n = (req.jmax + 1) ** 2
        return 16 * n * 2 + req.nsteps * (16 * 1 + 4 + 4)

    # -- admission: device state of one request, job record in numpy form
    def _admit(self, req, now):
        torch = self.torch
        m = self.model(req.jmax)
        dev = "cuda"
        E = req.B * m.J * (m.J + 1.0)
        d = torch.from_numpy(E.astype(np.complex128)).to(dev, non_blocking=False)
        phase = torch.empty(m.n, dtype=torch.complex128, device=dev)
        _lib.check(self.lib.rm_phase_init(m.n, ctypes.c_void_p(d.data_ptr()), float(self.f.real),
                                          float(self.f.imag), ctypes.c_void_p(phase.data_ptr()),
                                          ctypes.c_void_p(self.stream.cuda_stream)), "rm_phase_init")
        tab = torch.from_numpy(m.table(req.fields())).to(dev)
        sflag = torch.empty(req.nsteps, dtype=torch.int32, device=dev)
        _lib.check(self.lib.rm_plan_step_flags(m.plan.handle, req.nsteps, ctypes.c_void_p(tab.data_ptr()),
                                               ctypes.c_void_p(sflag.data_ptr()),
                                               ctypes.c_void_p(self.stream.cuda_stream)), "rm_plan_step_flags")
        psi = torch.zeros((1, m.n), dtype=torch.complex128, device=dev)
        psi[0, 0] = 1.0
        kused = torch.zeros(req.nsteps, dtype=torch.int32, device=dev)
        status = torch.zeros(1, dtype=torch.int32, device=dev)
        rec = np.zeros(1, dtype=JOB)
        rec["plan"] = m.plan.handle.value
        rec["psi"] = psi.data_ptr()
        rec["ld"] = m.n
        rec["col"] = 0
        rec["lo"], rec["len"] = m.lo, m.len
        rec["fprod_dev"] = tab.data_ptr()
        rec["sflag_dev"] = sflag.data_ptr()
        rec["f_re"], rec["f_im"] = self.f.real, self.f.imag
        rec["phase"] = phase.data_ptr()
        rec["kused"] = kused.data_ptr()
        rec["status"] = status.data_ptr()
        req.t_admit = now
        req.steps_done = 0
        keep = (d, phase, tab, sflag, psi, kused, status)
        self.used += self.footprint(req)
        return rec, keep, m.nslot

    def run(self, requests, clock=time.perf_counter):
        """Serve ``requests`` (arrival times in seconds from the start);
        returns them with t_admit / t_done set and psi on the host."""
        torch = self.torch
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        queue = deque()
        active, recs, keeps, nslots = [], [], [], []
        jobs = np.zeros(0, dtype=JOB)
        ev = torch.cuda.Event()
        t0 = clock()
        done = []

        def admit(now):
            nonlocal jobs
            changed = False
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            while queue and len(active) < self.max_jobs and self.used + self.footprint(queue[0]) <= self.budget:
                r = queue.popleft()
                rec, keep, ns = self._admit(r, now)
                active.append(r); recs.append(rec); keeps.append(keep); nslots.append(ns)
                changed = True
            if changed:
                jobs = job_array(recs)

        while pending or queue or active:
            now = clock() - t0
            admit(now)
            if not active:
                if pending:
                    wait = pending[0].arrival - (clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 1e-3))
                continue
            # this tick: each active request advances min(chunk, remaining) steps
            steps_done = np.array([r.steps_done for r in active], dtype=np.int64)
            total = np.array([r.nsteps for r in active], dtype=np.int64)
            nst = np.minimum(self.chunk, total - steps_done)
            base_tab = np.array([k[2].data_ptr() for k in keeps], dtype=np.uint64)
            base_flag = np.array([k[3].data_ptr() for k in keeps], dtype=np.uint64)
            base_k = np.array([k[5].data_ptr() for k in keeps], dtype=np.uint64)
            ns = np.array(nslots, dtype=np.uint64)
            jobs["fprod_dev"] = base_tab + steps_done.astype(np.uint64) * ns * np.uint64(16)
            jobs["sflag_dev"] = base_flag + steps_done.astype(np.uint64) * np.uint64(4)
            jobs["kused"] = base_k + steps_done.astype(np.uint64) * np.uint64(4)
            jobs["nsteps"] = nst.astype(np.int32)
            batch_propagate(self.lib, self.handle, jobs, self.kfix, self.maxorder, self.tol, self.precision,
                            self.stream.cuda_stream)
            ev.record(self.stream)
            # overlap: admit the requests that arrived meanwhile (they join the next tick)
            n_before = len(active)
            admit(clock() - t0)
            ev.synchronize()
            t_tick = clock() - t0
            keep_idx = []
            for i in range(n_before):
                r = active[i]
                r.steps_done += int(nst[i])
                if r.steps_done >= r.nsteps:
                    r.t_done = t_tick
                    r.psi = keeps[i][4]
                    r.kused = keeps[i][5]
                    self.used -= self.footprint(r)
                    done.append(r)
                else:
                    keep_idx.append(i)
            keep_idx += list(range(n_before, len(active)))
            if len(keep_idx) != len(active):
                active = [active[i] for i in keep_idx]
                recs = [recs[i] for i in keep_idx]
                keeps = [keeps[i] for i in keep_idx]
                nslots = [nslots[i] for i in keep_idx]
                jobs = job_array(recs)
            elif len(jobs) != len(active):
                jobs = job_array(recs)
        torch.cuda.synchronize()
        for r in done:
            r.psi = r.psi.cpu().numpy()[0]
            r.kused = r.kused.cpu().numpy()
        return done


def summarize(done, wall):
    lat = np.array([r.t_done - r.arrival for r in done])
    steps = sum(r.nsteps for r in done)
    return {"requests": len(done), "wall_s": wall, "requests_per_s": len(done) / wall,
            "wavepacket_steps_per_s": steps / wall, "latency_p50_ms": float(np.percentile(lat, 50) * 1e3),
            "latency_p99_ms": float(np.percentile(lat, 99) * 1e3),
            "latency_max_ms": float(lat.max() * 1e3), "mean_steps": steps / len(done)}


def benchmark(n_requests=200, seed=0, load=0.8, chunk=64, **kw):
    """Closed-loop capacity, then a Poisson stream at ``load`` x capacity."""
    srv = Server(chunk=chunk, **kw)
    reqs = make_requests(n_requests, seed=seed)
    for j in sorted({r.jmax for r in reqs}):   # operator cache warm (shared per Jmax)
        srv.model(j)
    t0 = time.perf_counter()
    done = srv.run(reqs)
    closed = time.perf_counter() - t0
    cap = len(done) / closed
    rate = load * cap
    reqs2 = make_requests(n_requests, seed=seed, rate=rate)
    t0 = time.perf_counter()
    done2 = srv.run(reqs2)
    wall = time.perf_counter() - t0
    rep = summarize(done2, wall)
    rep.update({"offered_rate_per_s": rate, "closed_loop_capacity_per_s": cap, "load": load,
                "chunk_steps": chunk, "krylov": srv.kfix, "models": len(srv.models),
                "stream": "linear rotors, B U(0.05,30) cm^-1, Jmax U{10..60}, FWHM U(100,1000) fs, "
                          "I log-U(1e10,1e12) W/cm^2, 200-2000 steps of 10 fs, psi0=|0,0>, "
                          f"Poisson arrivals at {load} x closed-loop capacity (seeded)"})
    return rep
