"""Drop-in time propagator: the interface of richmol.tdse.TDSE on the GPU.

Same constructor, ``time_grid``, ``init_state`` and ``update(H, vecs, **kw)
-> (vecs, t)`` as richmol/tdse.py:27-414.  ``update`` runs the split-operator
step and the Lanczos expv of every wavepacket in one HIP launch
(``rm_step``): the reference's Python loop over rows (tdse.py:379-382) and
per-matvec block loop (field.py:1212-1243) become one workgroup per
wavepacket.

``vecs`` may be a numpy (R, N) complex array (copied to the device and back,
like the reference returns a new array) or a torch complex128 device tensor
(stays on the device; a new tensor is returned).
"""
import contextlib
import functools

import numpy as np
from scipy.sparse import diags, csr_matrix

from . import units
from .field import CarTens


def update_counter(func):
    """tdse.py:11-24: return (vecs, t_2[icall]), counting calls per name."""
    @functools.wraps(func)
    def wrapper(self, *args, **kwargs):
        vecs = func(self, *args, **kwargs)
        name = func.__name__
        if name not in self._ncalls:
            self._ncalls[name] = 0
        icall = self._ncalls[name]
        t = self._time_grid[1][icall]
        self._ncalls[name] += 1
        return vecs, t
    return wrapper


class TDSE:
    """Time propagator (richmol/tdse.py:27-143).

    Kwargs: t_start, t_end, dt, t_units ('fs', 'ps', 'ns', 'au'),
    enr_units ('invcm', 'MHz').
    """

    def __init__(self, **kwargs):
        self._ncalls = dict()
        if "t_start" in kwargs:
            assert type(kwargs["t_start"]) in [int, float], \
                f"`initial time `t_start` has bad type: '{type(kwargs['t_start'])}', (use 'int', 'float')"
            self._t_start = kwargs["t_start"]
        else:
            self._t_start = 0.0
        assert "t_end" in kwargs, "terminal time 't_end' not found in kwargs"
        assert type(kwargs["t_end"]) in [int, float], \
            f"terminal time `t_end` has bad type: '{type(kwargs['t_end'])}', (use 'int', 'float')"
        assert kwargs["t_end"] > self._t_start, \
            f"terminal time `t_end` has bad value: '{kwargs['t_end']}', (must be > '{self._t_start}')"
        self._t_end = kwargs["t_end"]
        assert "dt" in kwargs, "time step `dt` not found in kwargs"
        assert type(kwargs["dt"]) in [int, float], \
            f"time step `dt` has bad type: '{type(kwargs['dt'])}', (use 'int', 'float')"
        assert kwargs["dt"] <= (self._t_end - self._t_start), \
            f"time step `dt` has bad value : '{kwargs['dt']}', (must be <= '{self._t_end - self._t_start}')"
        self._dt = kwargs["dt"]
        t_units = units.T_UNITS
        if "t_units" in kwargs:
            assert type(kwargs["t_units"]) == str, \
                f"time units `t_units` has bad type: '{type(kwargs['t_units'])}', (use 'str')"
            assert kwargs["t_units"] in list(t_units.keys()), \
                f"time units `t_units` has bad value: '{kwargs['t_units']}', (use 'fs', 'ps', 'ns', 'au')"
            self._t_to_s = t_units[kwargs["t_units"]]
        else:
            self._t_to_s = t_units["ps"]
        enr = units.enr_units()
        if "enr_units" in kwargs:
            assert type(kwargs["enr_units"]) == str, \
                f"energy units `enr_units` has bad type: '{type(kwargs['enr_units'])}', (use 'str')"
            assert kwargs["enr_units"].lower() in list(enr.keys()), \
                f"energy units `enr_units` has bad value: '{kwargs['enr_units']}', (use 'invcm', 'MHz')"
            self._enr_to_J = enr[kwargs["enr_units"].lower()]
        else:
            self._enr_to_J = enr["invcm"]

    @property
    def exp_fac(self):
        """Exponent factor f of exp(f H) per step (tdse.py:336-337)."""
        return -1j * self._dt * self._t_to_s * self._enr_to_J / units.HBAR

    def time_grid(self, grid_type="equidistant"):
        """tdse.py:146-174: centres t_c; (t_1, t_2, t_c) kept for update()."""
        assert grid_type.lower() in ["equidistant"], \
            f"time grid type `grid_type` has bad value: '{grid_type}', (use 'equidistant')"
        n = int((self._t_end - self._t_start) / self._dt)
        t1 = np.linspace(self._t_start, self._t_end, num=n, endpoint=False)
        t2 = t1 + self._dt
        tc = t1 + self._dt / 2
        self._time_grid = (t1, t2, tc)
        return tc

    def init_state(self, H, temp=None, thresh=1e-3, zpe=None, sparse=False):
        """Eigenvectors of H, Boltzmann-weighted (tdse.py:177-262).

        Host setup (dense eigh, or the identity when H is a rank-0 diagonal
        Hamiltonian), as in the reference; rows are complex128.
        """
        if temp is not None:
            assert temp >= 0, f"temperature `temp` has negative value: '{temp}'"
        assert thresh >= 0, f"partition function threshold `thresh` has negative or zero value: '{thresh}'"
        is_diag = False
        cart = getattr(H, "cart", None)
        if cart is not None and len(cart) > 0 and cart[0] == "0" and not H.is_sum:
            H.field([0, 0, 1])
            is_diag = True
        try:
            H.mfmat
        except AttributeError:
            raise AttributeError("hamiltonian `H` has inappropriate type (must be Hamiltonian)") from None
        if is_diag:
            enrs = H.tomat(form="full", repres="csr_matrix").diagonal()
            vecs = diags(np.ones(len(enrs)), format="csr")
        else:
            enrs, v = np.linalg.eigh(H.tomat(form="full", repres="dense"))
            vecs = csr_matrix(v)
        if zpe is not None:
            assert zpe <= abs(enrs[0]), \
                f"zero-point energy `zpe` has a value that is larger than the lowest energy: zpe = '{zpe}' > emin = '{enrs[0]}'"
        else:
            zpe = enrs[0]
        enrs = enrs - zpe
        vecs = vecs.transpose().tocsr()
        if temp is None:
            pass
        elif temp == 0:
            vecs = vecs.getrow(0)
        else:
            e = enrs * self._enr_to_J
            beta = 1.0 / (units.BOLTZMANN * temp)
            w = np.exp(-beta * e)
            w /= np.sum(w)
            inds = [i for i in range(len(w)) if (1 - np.sum(w[: i + 1])) > thresh]
            vecs = vecs[inds].multiply(np.sqrt(w[inds])[:, None]).tocsr()
        if not sparse:
            vecs = np.asarray(vecs.toarray())
        return vecs.astype(np.complex128)

    @update_counter
    def update(self, H, vecs, **kwargs):
        """Propagate wavepackets (rows of ``vecs``) by one time step.

        Kwargs as tdse.py:276-286: H0 (diagonal field-free CarTens: split
        operator), matvec_lib ('hip'), propag ('internal' = the reference's
        adaptive Lanczos, here on the device), tol (default 1e-15).
        Device extensions: maxorder (default 100), krylov_fixed (fixed
        Krylov dimension, default 0 = adaptive), stream (torch stream),
        kernel ('subspace': the step runs as a one-step rm_propagate on the
        invariant-subspace kernels, with the wavepackets' domains kept from
        the previous call; a full-vector kernel name of plan.set_kernel
        ('ell', 'ell:<i>', 'resident', 'global'): rm_step; 'auto' (default):
        'subspace' from SUBSPACE_MIN_ROWS rows of vecs on, rm_step with the
        plan's own full-vector choice below -- the host work per call sets the
        time of a step of a few wavepackets, and rm_step's is smaller),
        operator_form ('auto' (default), 'flat' or 'kron'): 'kron' keeps the
        operator factored as richmol stores it (Kronecker-block plan; needs
        every J with all m and the full k space and no field thresholds),
        'auto' takes it for such bases of at least KRON_MIN_STATES states;
        precision ('fp64', or 'fp32' dense products, Kronecker-block form
        only).
        """
        from . import plan as _plan
        import torch

        if "H0" in kwargs:
            assert isinstance(kwargs["H0"], CarTens), \
                f"field-free Hamiltonian `H0` is of bad (sub-)class: '{type(kwargs['H0'])}', " \
                f"(must be (sub-class of) '{CarTens}')"
            H0 = kwargs["H0"]
        else:
            H0 = None
        matvec_lib = kwargs.get("matvec_lib", "hip")
        if matvec_lib != "hip":
            raise ValueError(f"matvec_lib='{matvec_lib}' selects a CPU backend of the reference; "
                             "this package runs matvec_lib='hip' only")
        propag = kwargs.get("propag", "internal")
        assert type(propag) is str, f"propagator 'propag' has bad type: '{type(propag)}', (must be 'str')"
        assert propag in ["external", "internal", "hip"], \
            f"propagator 'propag' has bad value: '{propag}', (use 'external', 'internal')"
        if "tol" in kwargs:
            assert type(kwargs["tol"]) in [int, float], \
                f"tolerance `tol` has bad type: '{type(kwargs['tol'])}', (must be 'int', 'float')"
            assert 0 < kwargs["tol"] <= 1, \
                f"tolerance `tol` has bad value: '{kwargs['tol']}', (must be > 0 and <= 1)"
            tol = kwargs["tol"]
        else:
            tol = 1e-15
        maxorder = int(kwargs.get("maxorder", 100))
        krylov_fixed = int(kwargs.get("krylov_fixed", 0))
        stream = kwargs.get("stream", None)
        operator_form = kwargs.get("operator_form", "auto")
        if propag == "external":
            operator_form = "flat"
        if operator_form not in ("auto", "flat", "kron"):
            raise ValueError(f"operator_form has bad value: '{operator_form}' (use 'auto', 'flat', 'kron')")
        precision = kwargs.get("precision", "fp64")

        f = self.exp_fac
        on_device = isinstance(vecs, torch.Tensor)
        if on_device:
            psi = vecs
            if psi.dtype != torch.complex128 or not psi.is_cuda:
                raise TypeError("device wavepackets must be complex128 tensors on the GPU")
            if psi.dim() != 2 or psi.stride(1) != 1:
                psi = psi.contiguous()
        else:
            arr = np.asarray(vecs)
            if arr.ndim != 2:
                raise ValueError("vecs must be a 2-D array of wavepacket rows")
            psi = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.complex128)).cuda()

        kron = None
        if H is None or (not H.is_sum and "_fprod" not in H.__dict__):
            if H0 is None:
                raise AttributeError("you need to multiply tensor with field before applying it to a vector")
            plan = None
        elif operator_form == "kron" or (operator_form == "auto" and psi.shape[1] >= KRON_MIN_STATES
                                         and _kron_eligible(H)):
            kron = _kron_plan(H)
            plan = None
        else:
            if precision != "fp64":
                raise ValueError("precision applies to the Kronecker-block form; use Ensemble for FP32 matvecs")
            plan = _plan.plan_for(H)
            # plans are shared with ensembles that may have picked another
            # kernel or precision
            kernel = kwargs.get("kernel", "auto")
            sub_step = kernel == "subspace" or (kernel == "auto" and psi.shape[0] >= SUBSPACE_MIN_ROWS)
            want = "auto" if kernel == "subspace" else kernel
            if plan.kernel != want:
                plan.set_kernel(want)
            plan.set_tuning()
            if plan.precision != "fp64":
                plan.set_precision("fp64")

        phase = None
        if H0 is not None:
            if "_exp_fac_H0" not in self.__dict__:
                h0 = H0.tomat(form="full", cart="0")
                assert (h0 - diags(h0.diagonal())).nnz == 0, \
                    "field-free Hamiltonian `H0` is not diagonal -> split-operator approach not implemented"
                d = torch.from_numpy(np.ascontiguousarray(h0.diagonal(), dtype=np.complex128)).cuda()
                ph = torch.empty_like(d)
                _plan_phase(d, f, ph, stream)
                self._exp_fac_H0 = ph
            phase = self._exp_fac_H0
            if phase.numel() != psi.shape[1]:
                raise ValueError("H0 dimension does not match the wavepackets")

        out = torch.empty_like(psi)
        ncols = psi.shape[0]
        if kron is not None:
            kop, dp, perm = kron
            if psi.shape[1] != kop.n:
                raise ValueError("wavepacket length does not match the operator's basis")
            row = kop.slot_row([fp for _, fp, _ in H.terms()])[None, :]
            work = torch.empty_like(psi)
            work[:, perm] = psi
            ph = None
            if phase is not None:
                ph = torch.empty_like(phase)
                ph[perm] = phase
            kused = torch.zeros((1, ncols), dtype=torch.int32, device=psi.device)
            status = torch.zeros(ncols, dtype=torch.int32, device=psi.device)
            dp.propagate(work, row, f, phase=ph, krylov_fixed=krylov_fixed, precision=precision, stream=stream,
                         maxorder=maxorder, tol=tol, kused=kused, status=status)
            out = work[:, perm]
            self.last_krylov_dims = kused[0]
            if krylov_fixed == 0 and bool((status != 0).any()):
                raise ValueError(f"Lanczos reached maximum order of '{maxorder}' without convergence")
        elif plan is None:
            # no field on the operator: only the two H0 phases (tdse.py:375, 377, 394)
            _phase_apply(psi, out, phase, stream)
        elif propag == "external":
            # Expokit's ZHEXPV as tdse.py:383-391 / 404-412 call it: Krylov
            # size 12, t = Im(f), the operator iH; one rm_propagate step
            fprod, thresh = plan.field_arrays(H)
            if np.any(thresh > 0):
                raise ValueError("propag='external' takes no field thresholds on the device")
            table = torch.from_numpy(np.ascontiguousarray(fprod[None, :], dtype=np.complex128)).to(psi.device)
            status = torch.zeros(ncols, dtype=torch.int32, device=psi.device)
            kused = torch.zeros(ncols, dtype=torch.int32, device=psi.device)
            out.copy_(psi)
            work = torch.empty_like(psi)
            plan.set_propagator("external")
            try:
                plan.propagate(out, work, table, f, phase=phase, krylov_fixed=int(kwargs.get("krylov_fixed", 12) or 12),
                               tol=tol, kused=kused, status=status, stream=stream)
            finally:
                plan.set_propagator("internal")
            self.last_krylov_dims = kused
            if bool((status != 0).any()):
                from .ensemble import ExpokitError
                raise ExpokitError("maximum number of steps reached without convergence")
        elif sub_step:
            out = self._step_subspace(plan, H, psi, f, phase, krylov_fixed, maxorder, tol, stream)
        else:
            status = torch.zeros(ncols, dtype=torch.int32, device=psi.device)
            kused = torch.zeros(ncols, dtype=torch.int32, device=psi.device)
            plan.set_field(H, stream)
            plan.step(psi, out, f, phase=phase, krylov_fixed=krylov_fixed, maxorder=maxorder,
                      tol=tol, kused=kused, status=status, stream=stream)
            self.last_krylov_dims = kused
            if krylov_fixed == 0 and bool((status != 0).any()):
                raise ValueError(f"Lanczos reached maximum order of '{maxorder}' without convergence")
        if on_device:
            return out
        return out.cpu().numpy()

    _RING = 8   # field rows in flight between the host and the device

    def _step_subspace(self, plan, H, psi, f, phase, krylov_fixed, maxorder, tol, stream):
        """One step as rm_propagate over a one-row field table.  The rows are
        copied into a buffer kept per plan, so that the next call on them
        finds the same wavepackets and keeps their domains (no host round
        trip for them); the device checks that they still hold, and a step
        whose rows left them (status RM_STALE_DOMAIN, rows untouched) is run
        again with fresh domains.  Rows that are the previous call's output,
        unchanged (the same tensor, its version counter as returned), lie in
        their domains by construction: with a fixed Krylov dimension their
        step then reads no status and the host does not wait for the device.
        Field rows go up through a ring of pinned staging rows, each reused
        after its copy has run."""
        import torch
        from . import _lib
        ncols, n = psi.shape
        st = self.__dict__.setdefault("_sub_state", {})
        key = (id(plan), ncols, n, psi.device)
        s = st.get(key)
        if s is None or s["plan"] is not plan:
            st.clear()     # one buffer set per TDSE: the newest plan's
            dev = psi.device
            nf = max(plan.nfield, 1)
            s = st[key] = {
                "plan": plan,
                "buf": torch.empty((ncols, n), dtype=torch.complex128, device=dev),
                "work": torch.empty((ncols, n), dtype=torch.complex128, device=dev),
                "table": torch.empty((self._RING, nf), dtype=torch.complex128, device=dev),
                "host": torch.empty((self._RING, nf), dtype=torch.complex128).pin_memory(),
                "events": [None] * self._RING,
                "slot": 0,
                "kused": torch.empty((1, ncols), dtype=torch.int32, device=dev),
                "status": torch.empty(ncols, dtype=torch.int32, device=dev),
                "last": None, "last_ver": None,
            }
        trusted = s["last"] is psi and psi._version == s["last_ver"] and plan.domain_owner == id(self)
        fprod, thresh = plan.field_arrays(H)
        i = s["slot"]
        s["slot"] = (i + 1) % self._RING
        if s["events"][i] is not None:
            s["events"][i].synchronize()     # this staging row's previous copy has run
        s["host"][i].copy_(torch.from_numpy(fprod))
        # the caller's stream, or (no switch, no lookup) the current one
        on = torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext()
        table = s["table"][i:i + 1]
        buf = s["buf"]
        with on:
            table.copy_(s["host"][i:i + 1], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record()
            s["events"][i] = ev
            if buf.data_ptr() != psi.data_ptr():
                buf.copy_(psi)
        if plan.domain_owner != id(self):
            plan.set_domain_reuse(True, owner=id(self))
        th = [float(t) if t > 0 else None for t in thresh]
        # no status array on the fast path: the library then reports a
        # cooperative launch's timeout itself (its return code raises)
        fast = trusted and krylov_fixed > 0
        for attempt in range(2):
            plan.propagate(buf, s["work"], table, f, phase=phase, thresh=th, krylov_fixed=krylov_fixed,
                           maxorder=maxorder, tol=tol, kused=s["kused"], status=None if fast else s["status"],
                           stream=stream)
            if fast:
                bad = None
                break
            bad = s["status"].cpu().numpy()
            if attempt == 0 and np.any(bad & _lib.RM_STALE_DOMAIN):
                plan.set_domain_reuse(True, owner=id(self))   # forget the kept domains
                with on:
                    buf.copy_(psi)
                continue
            break
        self.last_krylov_dims = s["kused"][0]
        if bad is not None:
            if krylov_fixed == 0 and np.any(bad & _lib.RM_NOT_CONVERGED):
                raise ValueError(f"Lanczos reached maximum order of '{maxorder}' without convergence")
            if np.any(bad & ~(_lib.RM_NOT_CONVERGED | _lib.RM_STALE_DOMAIN)):
                raise RuntimeError(f"rm_propagate reported status {int(bad.max())}")
        with on:
            out = buf.clone()
        s["last"], s["last_ver"] = out, out._version
        return out

    def run(self, H, vecs, field, H0=None, obs=None, thresh=None, nsteps=None, **kwargs):
        """The reference's time loop (examples/ocs_alignment.py:89-103) as one
        device call::

            for i, t in enumerate(tdse.time_grid()):
                H.field(field(t), thresh=thresh)
                vecs, t_ = tdse.update(H, H0=H0, vecs=vecs)
                expval[i] = sum(vdot(vecs[b], O @ vecs[b]) for b in rows)

        field: a callable t -> [Fx, Fy, Fz] taken at the step centres of the
        time grid (from the step update() would take next), or an
        (nsteps, 3) array.  obs: an observable CarTens O (its matrix for the
        field [0, 0, 1], as tomat(form="full", cart="0") takes it, e.g.
        cos^2 theta) or None.  nsteps: default, the rest of the grid.
        Kwargs as update(): tol, maxorder, krylov_fixed, stream, precision.
        Returns (vecs, t, expval): t the end of the last step (update()'s
        call counter advances by nsteps), expval the complex (nsteps,) sums
        (None without obs).  Runs every step in one rm_propagate_expect call
        on the flattened operator, or (operator_form as in update(), 'auto'
        for large bases) one rm_kplan_propagate_expect call on the
        Kronecker-block form."""
        import torch
        from .ensemble import Ensemble
        if "_time_grid" not in self.__dict__:
            self.time_grid()
        icall = self._ncalls.get("update", 0)
        t2, tc = self._time_grid[1], self._time_grid[2]
        if nsteps is None:
            nsteps = len(tc) - icall
        if nsteps < 1 or icall + nsteps > len(tc):
            raise ValueError(f"{nsteps} steps from step {icall} run past the time grid ({len(tc)} steps)")
        if callable(field):
            fields = np.array([np.asarray(field(t), dtype=np.float64) for t in tc[icall:icall + nsteps]])
        else:
            fields = np.asarray(field, dtype=np.float64)
            if fields.shape != (nsteps, 3):
                raise ValueError(f"field array of shape {fields.shape}, expected ({nsteps}, 3)")
        if H0 is not None:
            h0 = H0.tomat(form="full", cart="0")
            assert (h0 - diags(h0.diagonal())).nnz == 0, \
                "field-free Hamiltonian `H0` is not diagonal -> split-operator approach not implemented"
        on_device = isinstance(vecs, torch.Tensor)
        n = vecs.shape[1]
        form = kwargs.get("operator_form", "auto")
        if form == "kron" or (form == "auto" and n >= KRON_MIN_STATES and _kron_eligible(H)):
            out, expval = self._run_kron(H, vecs, fields, H0, obs, thresh, kwargs)
            self._ncalls["update"] = icall + nsteps
            return (out if on_device else out.cpu().numpy()), t2[icall + nsteps - 1], expval
        ens = Ensemble(H, H0, self.exp_fac, vecs if on_device else np.asarray(vecs),
                       krylov_fixed=int(kwargs.get("krylov_fixed", 0)),
                       maxorder=int(kwargs.get("maxorder", 100)), tol=float(kwargs.get("tol", 1e-15)),
                       stream=kwargs.get("stream", None))
        ens.set_precision(kwargs.get("precision", "fp64"))
        expval = None
        if obs is None:
            ens.propagate(fields, thresh=thresh)
        else:
            _, s = ens.propagate(fields, thresh=thresh, obs=Ensemble.observable_plan(obs), expect_sum=True)
            expval = s.cpu().numpy()
        self._ncalls["update"] = icall + nsteps
        out = ens.psi if on_device else ens.psi.cpu().numpy()
        return out, t2[icall + nsteps - 1], expval

    def _run_kron(self, H, vecs, fields, H0, obs, thresh, kwargs):
        """run() on the Kronecker-block form: per-step slot products from
        H.field(F) on the host, one rm_kplan_propagate_expect call."""
        import torch
        if thresh is not None:
            raise ValueError("field thresholds are not supported by the Kronecker-block form")
        H.field(list(fields[0]))
        kop, dp, perm = _kron_plan(H)
        fp = []
        for F in fields:
            H.field(list(F))
            fp.append(kop.slot_row([q for _, q, _ in H.terms()]))
        fp = np.array(fp)
        src = vecs if isinstance(vecs, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(vecs, dtype=np.complex128)).cuda()
        if src.shape[1] != kop.n:
            raise ValueError("wavepacket length does not match the operator's basis")
        work = torch.empty_like(src)
        work[:, perm] = src
        ph = None
        if H0 is not None:
            d = torch.from_numpy(np.ascontiguousarray(H0.tomat(form="full", cart="0").diagonal(),
                                                      dtype=np.complex128)).cuda()
            phase = torch.empty_like(d)
            _plan_phase(d, self.exp_fac, phase, kwargs.get("stream"))
            ph = torch.empty_like(phase)
            ph[perm] = phase
        krylov_fixed = int(kwargs.get("krylov_fixed", 0))
        maxorder = int(kwargs.get("maxorder", 100))
        status = torch.zeros(src.shape[0], dtype=torch.int32, device=src.device)
        common = dict(phase=ph, krylov_fixed=krylov_fixed, precision=kwargs.get("precision", "fp64"),
                      stream=kwargs.get("stream"), maxorder=maxorder, tol=float(kwargs.get("tol", 1e-15)),
                      status=status)
        expval = None
        if obs is None:
            dp.propagate(work, fp, self.exp_fac, **common)
        else:
            obs.field([0.0, 0.0, 1.0])
            okop, odp, operm = _kron_plan(obs)
            if okop.Jlist != kop.Jlist or not torch.equal(operm, perm):
                raise ValueError("the observable's basis differs from the Hamiltonian's")
            sums = torch.empty(len(fields), dtype=torch.complex128, device=src.device)
            dp.propagate(work, fp, self.exp_fac, obs=odp,
                         obs_fprod=okop.slot_row([q for _, q, _ in obs.terms()]), expect_sum=sums, **common)
            expval = sums.cpu().numpy()
        if krylov_fixed == 0 and bool((status != 0).any()):
            raise ValueError(f"Lanczos reached maximum order of '{maxorder}' without convergence")
        return work[:, perm], expval


# TDSE.update(kernel='auto'): rows of vecs from which a step runs as a one-step
# rm_propagate on the subspace kernels (measured per call on the J <= 40 linear
# rotor: 1 row 76 us vs 53 us through rm_step, 16 rows 79 vs 72 us fixed-m and
# 118 vs 124 us adaptive, 512 rows 75 vs 97 us)
SUBSPACE_MIN_ROWS = 16

# 'auto' operator form: bases at least this large with every J carrying all m
# and the full k space keep the operator factored (a flattened asymmetric top
# at J <= 80 would hold ~3e9 entries)
KRON_MIN_STATES = 150_000
_KRON_CACHE = {}


def _kron_eligible(H):
    leaf = H.terms()[0][0]
    try:
        if list(leaf.Jlist1) != list(leaf.Jlist2):
            return False
        for J in leaf.Jlist1:
            syms = leaf.symlist1[J]
            if any(leaf.dim_m1[J][s] != 2 * J + 1 for s in syms):
                return False
            if sum(leaf.dim_k1[J][s] for s in syms) != 2 * J + 1:
                return False
    except (AttributeError, KeyError):
        return False
    return True


def _kron_plan(H):
    """(KronOperator, DeviceKronPlan, device permutation) for H's leaf
    structure, built once per set of leaves (the field only changes the
    per-step products)."""
    import torch
    from .kron import DeviceKronPlan, KronOperator
    key = tuple(leaf._uid for leaf, _, _ in H.terms())
    hit = _KRON_CACHE.get(key)
    if hit is None:
        kop, _ = KronOperator.from_cartens(H)
        hit = (kop, DeviceKronPlan(kop), torch.from_numpy(kop.perm).cuda())
        # two plans at most: an operator and its observable
        while len(_KRON_CACHE) >= 2:
            _KRON_CACHE.pop(next(iter(_KRON_CACHE)))
        _KRON_CACHE[key] = hit
    else:
        _KRON_CACHE[key] = _KRON_CACHE.pop(key)
    return hit


def _phase_apply(psi, out, phase, stream=None):
    import ctypes
    from . import _lib
    from .plan import _stream_handle
    lib = _lib.load()
    _lib.check(lib.rm_phase_apply(psi.shape[1], ctypes.c_void_p(psi.data_ptr()),
                                  ctypes.c_void_p(out.data_ptr()), psi.stride(0), psi.shape[0],
                                  ctypes.c_void_p(phase.data_ptr()),
                                  ctypes.c_void_p(_stream_handle(stream))), "rm_phase_apply")


def _plan_phase(diag, f, out, stream=None):
    import ctypes
    from . import _lib
    from .plan import _stream_handle
    lib = _lib.load()
    _lib.check(lib.rm_phase_init(diag.numel(), ctypes.c_void_p(diag.data_ptr()), float(f.real),
                                 float(f.imag), ctypes.c_void_p(out.data_ptr()),
                                 ctypes.c_void_p(_stream_handle(stream))), "rm_phase_init")
