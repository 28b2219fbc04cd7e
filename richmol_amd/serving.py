"""Continuous-batching propagation server.

A request is one field-driven propagation: a rotor Hamiltonian, a pulse, a
number of steps, an initial state.  The server keeps every admitted request's
wavepacket, phases and field table resident on the device and advances all
active requests together, ``chunk`` steps per tick, with one
``rm_batch_propagate`` call (one wave -- or, for domains longer than a wave
holds, one workgroup -- per request on its invariant subspace, jobs of
different operators in the same launch).  Requests join at tick boundaries
as they arrive -- admitted while slots and the HBM budget allow -- and leave
when their last step is done; host-side admission of new requests (operator
build included) overlaps the device work of the running tick.

Requests are linear rotors (B, polarizability LINEAR_ALPHA, all m; operators
shared per Jmax) or asymmetric tops (rotational constants A >= B >= C along
z, x, y, polarizability ASYM_ALPHA; the m = 0 block, which a Z pulse never
leaves from |J=0>; operator built per request).  psi0 is the J = 0 state.

Latency is measured from arrival to the completion of the tick that finished
the request.  ``benchmark`` drives a seeded Poisson stream of requests drawn
as the serving configuration describes (linear or asymmetric top 50/50,
rotational constants U(0.05, 30) cm^-1, Jmax U{10..60}, pulse FWHM
U(100, 1000) fs, peak intensity log-U(1e10, 1e12) W/cm^2, 200-2000 steps of
10 fs, complex64 matvec with FP64 Krylov) at 0.8 of the capacity measured on
the same stream in a closed-loop run.
"""
import ctypes
import time
from collections import deque

import numpy as np

from . import _lib, rot, units

LINEAR_ALPHA = (26.0, 26.0, 53.6)   # alpha_perp, alpha_perp, alpha_par [au]
ASYM_ALPHA = (12.0, 8.0, 16.0)      # alpha_xx, alpha_yy, alpha_zz [au]; x = b, y = c, z = a axis
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
    benchmark workload); ``B``: rotational constant of a linear rotor
    [cm^-1]; ``abc``: (A, B, C) of an asymmetric top (kind 'asym'), whose
    polarizability is ``alpha`` (3 x 3, molecular frame; default
    diag(ASYM_ALPHA))."""

    __slots__ = ("rid", "arrival", "kind", "B", "abc", "alpha", "jmax", "fwhm", "intensity", "nsteps", "dt",
                 "t_admit", "t_done", "psi", "steps_done", "kused", "index", "cumfield", "build")

    def __init__(self, rid, arrival, B, jmax, fwhm, intensity, nsteps, kind="linear", dt=DT_FS, abc=None,
                 alpha=None):
        if kind not in ("linear", "asym"):
            raise ValueError(f"unknown request kind '{kind}'")
        if kind == "asym" and (abc is None or len(abc) != 3):
            raise ValueError("an asymmetric-top request needs abc = (A, B, C)")
        self.rid, self.arrival, self.kind = rid, float(arrival), kind
        self.build = None   # future of the operator build (a server with builder processes)
        self.B, self.jmax, self.fwhm, self.intensity = float(B), int(jmax), float(fwhm), float(intensity)
        self.abc = None if abc is None else tuple(float(x) for x in abc)
        # molecular-frame polarizability of an asymmetric top (default diag(ASYM_ALPHA))
        self.alpha = None if alpha is None else np.asarray(alpha, dtype=np.float64)
        self.nsteps, self.dt = int(nsteps), float(dt)
        self.t_admit = self.t_done = None
        self.psi = None
        self.index = None   # basis positions of the propagated states (asymmetric top: its subspace)
        self.steps_done = 0

    @property
    def n(self):
        """States of the request's basis: (Jmax+1)^2 for both kinds (all m of
        a linear rotor; the m = 0 block of an asymmetric top)."""
        return (self.jmax + 1) ** 2

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
                  intensity=(1e10, 1e12), steps=(200, 2000), asym_fraction=0.0):
    """Seeded request stream; Poisson arrivals at ``rate`` per second (all at
    t = 0 when rate is None); a fraction ``asym_fraction`` of asymmetric
    tops, their three constants drawn from ``B`` and sorted A >= B >= C."""
    rng = np.random.default_rng(seed)
    arr = np.zeros(n) if rate is None else np.cumsum(rng.exponential(1.0 / rate, size=n))
    out = []
    for i in range(n):
        r = Request(i, arr[i], B=rng.uniform(*B), jmax=rng.integers(jmax[0], jmax[1] + 1),
                    fwhm=rng.uniform(*fwhm),
                    intensity=10 ** rng.uniform(np.log10(intensity[0]), np.log10(intensity[1])),
                    nsteps=rng.integers(steps[0], steps[1] + 1))
        if asym_fraction > 0:
            asym = rng.random() < asym_fraction
            abc = np.sort(rng.uniform(*B, size=3))[::-1]
            if asym:
                r.kind, r.abc = "asym", tuple(float(x) for x in abc)
        out.append(r)
    return out


def _domain(A, n):
    """Domain [lo, lo + len) of the J = 0 state (row 0) in the plan's
    permuted order: the subspaces are numbered by their first row, so row 0's
    subspace comes first (lo = 0) and len is its size (host, from the
    sparsity pattern; no device call)."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import breadth_first_order
    P = csr_matrix(A)
    P = csr_matrix((np.ones(P.nnz), P.indices, P.indptr), shape=(n, n))
    reach = breadth_first_order(P + P.T, 0, directed=False, return_predecessors=False)
    return 0, int(reach.size)


# Device time of one field step (microseconds, fixed Krylov dimension 12,
# complex64 matvec) by launch form: (fixed, per stored entry of the domain).
# Measured on MI355X with one request at a time (2 000 steps, 800 of them with
# field; the tick overhead of ~10 us per step taken off): linear rotors on the
# two-per-wave kernel ~6 us at any Jmax; asymmetric tops on one wave (domains
# of 21 and 36 rows: 24 and 36 us); on one workgroup (66-496 rows, 1.4e3 to
# 2.1e4 entries: 48, 93, 147, 320 us).
STEP_US = {"pair": (6.0, 0.0), "wave": (15.0, 0.039), "block": (30.0, 0.0142)}


def _domain_stats(A, n):
    """(rows, stored entries, longest row) of the J = 0 state's domain (row
    0's invariant subspace; host, from the sparsity pattern)."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import breadth_first_order
    P = csr_matrix(A)
    P = csr_matrix((np.ones(P.nnz), P.indices, P.indptr), shape=(n, n))
    reach = breadth_first_order(P + P.T, 0, directed=False, return_predecessors=False)
    lens = np.diff(P.indptr)[reach]
    return int(reach.size), int(lens.sum()), int(lens.max())


class _Model:
    """Operator, device plan and the |J=0,m=0> domain shared by every
    linear-rotor request with this Jmax (B enters only the phases)."""

    def __init__(self, master, jmax):
        from .plan import DevicePlan, field_table
        self.jmax = jmax
        self.H = rot.restrict(master, jmax) * (-0.5 * units.AUpol_x_Vm_to_invcm())
        self.n = (jmax + 1) ** 2
        self.J = np.repeat(np.arange(jmax + 1), 2 * np.arange(jmax + 1) + 1).astype(np.float64)
        active, _, _ = field_table(self.H, np.array([[0.0, 0.0, 1.0]]))
        self.plan = DevicePlan(self.H, active)
        self.plan.subspaces()
        self.nslot = self.plan.nfield
        arr = self.plan._arrays
        pattern = (np.ones(arr["nnz"]), arr["col_idx"], arr["row_ptr"])
        from scipy.sparse import csr_matrix
        P = csr_matrix(pattern, shape=(self.n, self.n))
        self.lo, self.len = _domain(P, self.n)
        self.stats = _domain_stats(P, self.n)

    def energies(self, req):
        return req.B * self.J * (self.J + 1.0)

    def table(self, F):
        from .plan import field_table
        _, cols, nst = field_table(self.H, F)
        return self.plan.table_array(cols, nst)


_SPECIES = {}


def _reachable_species(alpha):
    """D2 species of the m = 0 states reachable from |J=0, A> under a Z field
    through the polarizability alpha (a symmetry property, the same at every
    J: found on a J <= 4 basis with generic constants; cached per alpha)."""
    key = np.asarray(alpha, dtype=np.float64).tobytes()
    got = _SPECIES.get(key)
    if got is None:
        bas = rot.asymtop_basis(1.7, 1.3, 2.9, 4, m_filter=lambda J, m: m == 0)
        _, _, sub = rot.zfield_m0_operator(bas, alpha, start=(0, "A"))
        blocks = [(J, sp) for J in bas.Jlist for sp in bas.symlist[J]
                  for _ in range(bas.kfun[J][sp][1].shape[1])]
        got = _SPECIES[key] = frozenset(blocks[i][1] for i in sub)
    return got


def _m0_positions(bas, jmax):
    """Whole-m = 0-basis position (every J <= jmax, all four species, in the
    order J, species, state) of every state of a species-filtered basis."""
    full, off = {}, 0
    for J in range(jmax + 1):
        for sp, ns in rot.species_sizes(J).items():
            full[(J, sp)] = off
            off += ns
    pos = [full[(J, sp)] + np.arange(bas.kfun[J][sp][1].shape[1]) for J in bas.Jlist for sp in bas.symlist[J]]
    return np.concatenate(pos).astype(np.int64)


def _m0_only(J, m):
    return m == 0


def asym_host_build(abc, jmax, alpha=None):
    """The host part of an asymmetric-top request's model -- everything but
    the device plan, so that it can run in a builder process while the
    server's thread admits and launches: H(F) = F_z^2 A on the invariant
    subspace of the J = 0 state inside the m = 0 block
    (rot.zfield_m0_operator from block (0, 'A'); the polarizability
    ASYM_ALPHA unless given, -1/2 alpha F F), the rigid rotor's energies with
    A, B, C along z, x, y, the m = 0 basis position of every subspace state
    (the J = 0 state first) and the domain's extent and statistics.
    Returns plain arrays: ((data, indices, indptr, shape), E, index, lo,
    len, stats)."""
    a, b, c = abc
    alpha = np.diag(ASYM_ALPHA) if alpha is None else np.asarray(alpha, dtype=np.float64)
    # only the species the J = 0 state reaches are diagonalised and coupled
    keep = _reachable_species(alpha)
    bas = rot.asymtop_basis(b, c, a, jmax, m_filter=_m0_only, species=keep)
    A0, E, sub = rot.zfield_m0_operator(bas, alpha, start=(0, "A"))
    index = _m0_positions(bas, jmax)[sub]
    H = (A0 * (-0.5 * units.AUpol_x_Vm_to_invcm())).tocsr()
    H.sort_indices()
    n = H.shape[0]
    lo, ln = _domain(H, n)
    return (H.data, H.indices, H.indptr, H.shape), E, index, lo, ln, _domain_stats(H, n)


class _AsymModel:
    """Operator and device plan of one asymmetric-top request (the host part
    from asym_host_build, given or computed here)."""

    def __init__(self, req, built=None):
        from scipy.sparse import csr_matrix
        from .plan import DevicePlan
        if built is None:
            built = asym_host_build(req.abc, req.jmax, req.alpha)
        (data, indices, indptr, shape), self.E, self.index, self.lo, self.len, self.stats = built
        self.H = csr_matrix((data, indices, indptr), shape=shape)
        self.nfull = req.n
        self.n = self.H.shape[0]
        self.plan = DevicePlan.from_matrix(self.H, "zz")
        self.plan.subspaces()   # the subspace arrays now, during the running tick
        self.nslot = 1

    def energies(self, req):
        return self.E

    def table(self, F):
        tab = np.zeros((len(F), 1), dtype=np.complex128)
        tab[:, 0] = F[:, 2] * F[:, 2]
        return tab


class Server:
    """Continuous-batching engine on one device stream.

    hbm_budget: bytes of device memory the resident request state may use
    (admission stops at the first queued request that does not fit);
    max_jobs: concurrent requests; chunk: field steps per tick of the
    costliest active request -- a request whose field steps cost c runs
    chunk * c_max / c of them per tick, plus the field-free steps (phases
    only) between them, at most chunk_max steps in all, so every request's
    part of a tick takes about as long and cheap requests are not held back
    by expensive ones (a result does not depend on how its steps are split);
    krylov_fixed: fixed Krylov dimension (0: adaptive with tol, maxorder);
    block_rows: jobs on domains longer than this run one workgroup each
    (default: only those longer than a wave holds);
    builders: processes that build asymmetric tops' operators on the host
    from their arrival on (0: built at admission, on the server's thread).
    """

    def __init__(self, hbm_budget=8 << 30, max_jobs=8192, chunk=64, krylov_fixed=12, maxorder=100,
                 tol=1e-15, precision="fp64", jmax_master=60, block_rows=64, chunk_max=4096, builders=0):
        import torch
        self.torch = torch
        self.lib = _lib.load()
        self.budget, self.max_jobs, self.chunk = int(hbm_budget), int(max_jobs), int(chunk)
        self.chunk_max = int(chunk_max)
        self.kfix, self.maxorder, self.tol = int(krylov_fixed), int(maxorder), float(tol)
        self.precision = _lib.PRECISIONS[precision]
        self.f = -1j * DT_FS * 1e-15 * units.enr_units()["invcm"] / units.HBAR
        # ticks run on a stream of their own, so admission's uploads (plan
        # creation, tensors) do not wait for the running tick
        self.stream = torch.cuda.Stream()
        self.t_host = 0.0     # seconds spent admitting (operators, uploads)
        h = ctypes.c_void_p()
        _lib.check(self.lib.rm_batch_create(ctypes.byref(h)), "rm_batch_create")
        self.handle = h
        _lib.track(self)
        if block_rows is not None:
            _lib.check(self.lib.rm_batch_set_block_rows(h, int(block_rows)), "rm_batch_set_block_rows")
        self.block_rows = 1024 if block_rows is None else int(block_rows)
        bas = rot.linear_basis(1.0, jmax_master)
        self.master = rot.tensor(np.diag(LINEAR_ALPHA), bas)
        self.models = {}
        self.used = 0
        self.pool = None
        if builders > 0:
            import multiprocessing
            from concurrent.futures import ProcessPoolExecutor
            # spawned (not forked) workers: they never touch the device
            self.pool = ProcessPoolExecutor(max_workers=int(builders),
                                            mp_context=multiprocessing.get_context("spawn"))

    def close(self):
        """Destroy the device batch object and stop the builders (idempotent)."""
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.rm_batch_destroy(h)
        self.handle = None
        pool = getattr(self, "pool", None)
        if pool is not None:
            pool.shutdown(wait=True, cancel_futures=True)
        self.pool = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def model(self, jmax):
        m = self.models.get(jmax)
        if m is None:
            m = self.models[jmax] = _Model(self.master, jmax)
            m.cost = self.step_cost(m)
        return m

    def step_cost(self, m):
        """Modelled device time of one field step of a model's J = 0 job
        (STEP_US by the launch form rm_batch_propagate picks for its domain)."""
        rows, nnz, width = m.stats
        if rows > self.block_rows:
            form = "block"
        elif width <= 4 and rows <= 32 and 0 < self.kfix <= 16:
            form = "pair"
        else:
            form = "wave"
        a, b = STEP_US[form]
        return a + b * nnz

    def footprint(self, req):
        """Resident device bytes of a request: wavepacket, phases, field table,
        step flags, Krylov dimensions; an asymmetric top also holds its own
        operator plan (about 50 stored entries per state at Jmax 60, ~100
        bytes each over the plan's arrays).  An asymmetric top is counted on
        its whole m = 0 block, an upper bound of the subspace it runs on."""
        n = req.n
        own = 0 if req.kind == "linear" else 100 * 50 * n
        return 16 * n * 2 + req.nsteps * (16 * 1 + 4 + 4) + own

    # -- admission: device state of one request, job record in numpy form
    def _admit(self, req, now):
        t0 = time.perf_counter()
        out = self._admit_one(req, now)
        self.t_host += time.perf_counter() - t0
        return out

    def _admit_one(self, req, now):
        torch = self.torch
        if req.kind == "linear":
            m = self.model(req.jmax)
        else:
            fut = req.build
            m = _AsymModel(req, built=fut.result() if fut is not None else None)
            req.build = None
            m.cost = self.step_cost(m)
        dev = "cuda"
        E = m.energies(req)
        d = torch.from_numpy(E.astype(np.complex128)).to(dev, non_blocking=False)
        phase = torch.empty(m.n, dtype=torch.complex128, device=dev)
        _lib.check(self.lib.rm_phase_init(m.n, ctypes.c_void_p(d.data_ptr()), float(self.f.real),
                                          float(self.f.imag), ctypes.c_void_p(phase.data_ptr()),
                                          ctypes.c_void_p(self.stream.cuda_stream)), "rm_phase_init")
        F = req.fields()
        # field steps up to each step (the ones that run a Lanczos expv)
        req.cumfield = np.concatenate([[0], np.cumsum(np.any(F != 0, axis=1))]).astype(np.int64)
        tab = torch.from_numpy(m.table(F)).to(dev)
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
        keep = (d, phase, tab, sflag, psi, kused, status, m)   # m: an asymmetric top's own plan
        self.used += self.footprint(req)
        return rec, keep, m.nslot

    def run(self, requests, clock=time.perf_counter):
        """Serve ``requests`` (arrival times in seconds from the start);
        returns them with t_admit / t_done set and psi on the host."""
        torch = self.torch
        self.t_host = 0.0
        self.ticks, self.t_device = 0, 0.0
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        queue = deque()
        active, recs, keeps, nslots = [], [], [], []
        jobs = np.zeros(0, dtype=JOB)
        ev = torch.cuda.Event()
        t0 = clock()
        done = []

        def admit(now, busy=None):
            """Admit queued requests that fit; with ``busy`` (the running
            tick's event) at least one and then only while the tick runs, so
            host admission never holds the device idle between ticks."""
            nonlocal jobs
            changed = False
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                if self.pool is not None and r.kind != "linear":
                    # the operator is built while the request waits for a slot
                    r.build = self.pool.submit(asym_host_build, r.abc, r.jmax, r.alpha)
                queue.append(r)
            while queue and len(active) < self.max_jobs and self.used + self.footprint(queue[0]) <= self.budget:
                if busy is not None and changed and busy.query():
                    break
                r = queue.popleft()
                rec, keep, ns = self._admit(r, now)
                active.append(r); recs.append(rec); keeps.append(keep); nslots.append(ns)
                changed = True
            if changed:
                jobs = job_array(recs)

        while pending or queue or active:
            now = clock() - t0
            if not active:
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
            # every request's part of the tick costs about chunk field steps of
            # the costliest one; field-free steps (phases only) are not counted
            cost = np.array([k[7].cost for k in keeps])
            fsteps = np.maximum(1, (self.chunk * cost.max() / cost).astype(np.int64))
            nst = np.empty(len(active), dtype=np.int64)
            for i, r in enumerate(active):
                cf = r.cumfield
                end = np.searchsorted(cf, cf[r.steps_done] + fsteps[i], side="right") - 1
                nst[i] = max(1, min(end - r.steps_done, self.chunk_max))
            nst = np.minimum(nst, total - steps_done)
            base_tab = np.array([k[2].data_ptr() for k in keeps], dtype=np.uint64)
            base_flag = np.array([k[3].data_ptr() for k in keeps], dtype=np.uint64)
            base_k = np.array([k[5].data_ptr() for k in keeps], dtype=np.uint64)
            ns = np.array(nslots, dtype=np.uint64)
            jobs["fprod_dev"] = base_tab + steps_done.astype(np.uint64) * ns * np.uint64(16)
            jobs["sflag_dev"] = base_flag + steps_done.astype(np.uint64) * np.uint64(4)
            jobs["kused"] = base_k + steps_done.astype(np.uint64) * np.uint64(4)
            jobs["nsteps"] = nst.astype(np.int32)
            # admission initialised the new requests' tensors on the default stream
            self.stream.wait_stream(torch.cuda.current_stream())
            ev0.record(self.stream)
            batch_propagate(self.lib, self.handle, jobs, self.kfix, self.maxorder, self.tol, self.precision,
                            self.stream.cuda_stream)
            ev1.record(self.stream)
            ev.record(self.stream)
            # overlap: admit the requests that arrived meanwhile (they join the next tick)
            n_before = len(active)
            admit(clock() - t0, busy=ev)
            ev.synchronize()
            t_tick = clock() - t0
            self.ticks += 1
            self.t_device += ev0.elapsed_time(ev1) * 1e-3
            keep_idx = []
            for i in range(n_before):
                r = active[i]
                r.steps_done += int(nst[i])
                if r.steps_done >= r.nsteps:
                    r.t_done = t_tick
                    r.psi = keeps[i][4]
                    r.kused = keeps[i][5]
                    r.index = getattr(keeps[i][7], "index", None)
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
        self.t_host_run = self.t_host
        for r in done:
            v = r.psi.cpu().numpy()[0]
            if r.index is not None:   # back to the whole m = 0 basis
                full = np.zeros(r.n, dtype=np.complex128)
                full[r.index] = v
                v = full
            r.psi = v
            r.kused = r.kused.cpu().numpy()
        return done


def summarize(done, wall):
    lat = np.array([r.t_done - r.arrival for r in done])
    steps = sum(r.nsteps for r in done)
    return {"requests": len(done), "wall_s": wall, "requests_per_s": len(done) / wall,
            "wavepacket_steps_per_s": steps / wall, "latency_p50_ms": float(np.percentile(lat, 50) * 1e3),
            "latency_p99_ms": float(np.percentile(lat, 99) * 1e3),
            "latency_max_ms": float(lat.max() * 1e3), "mean_steps": steps / len(done)}


def benchmark(n_requests=200, seed=0, load=0.8, chunk=64, asym_fraction=0.5, precision="fp32", builders=8,
              **kw):
    """Closed-loop capacity, then a Poisson stream at ``load`` x capacity;
    asymmetric tops' operators built by ``builders`` host processes."""
    srv = Server(chunk=chunk, precision=precision, builders=builders, **kw)
    reqs = make_requests(n_requests, seed=seed, asym_fraction=asym_fraction)
    for j in sorted({r.jmax for r in reqs if r.kind == "linear"}):   # linear operators (shared per Jmax)
        srv.model(j)
    t0 = time.perf_counter()
    done = srv.run(reqs)
    closed = time.perf_counter() - t0
    cap = len(done) / closed
    rate = load * cap
    reqs2 = make_requests(n_requests, seed=seed, rate=rate, asym_fraction=asym_fraction)
    t0 = time.perf_counter()
    done2 = srv.run(reqs2)
    wall = time.perf_counter() - t0
    rep = summarize(done2, wall)
    rep.update({"admission_host_s": srv.t_host_run, "ticks": srv.ticks, "device_busy_s": srv.t_device,
                "queue_wait_p50_ms": float(np.percentile([r.t_admit - r.arrival for r in done2], 50) * 1e3),
                "queue_wait_p99_ms": float(np.percentile([r.t_admit - r.arrival for r in done2], 99) * 1e3),"offered_rate_per_s": rate, "closed_loop_capacity_per_s": cap, "load": load,
                "chunk_steps": chunk, "krylov": srv.kfix, "models": len(srv.models),
                "precision": precision, "asym_fraction": asym_fraction,
                "requests_asym": sum(r.kind == "asym" for r in done2),
                "stream": f"linear rotors / asymmetric tops ({asym_fraction:.0%} asym), rotational constants "
                          "U(0.05,30) cm^-1, Jmax U{10..60}, FWHM U(100,1000) fs, I log-U(1e10,1e12) W/cm^2, "
                          "200-2000 steps of 10 fs, Z pulse, psi0 = J=0 state, "
                          f"Poisson arrivals at {load} x closed-loop capacity (seeded)",
                "host_builders": builders})
    srv.close()
    return rep
