"""Batched propagation of many wavepackets that share one Hamiltonian.

A thermal ensemble (Boltzmann-sampled initial states) or a field-parameter
scan at fixed structure is one (B, N) block of rows on the device; every time
step is one field contraction and one ``rm_step`` launch for all B rows
(the reference loops over rows in Python, tdse.py:379-382).
"""
import numpy as np

from . import units


class ExpokitError(ValueError):
    """The external propagator's failures, as richmol/pyexpokit.py:5-14
    names them."""


def thermal_sample(basis, temp, nsamp, seed=0):
    """Indices of nsamp basis states |J, m, k> drawn with probability
    proportional to exp(-E/kT) (all m equally weighted), seeded."""
    E = np.array([e for J in basis.Jlist for s in basis.symlist[J]
                  for m in basis.mlist[J] for e in basis.energy[J][s]])
    kT = units.BOLTZMANN * temp * units.J_to_invcm()
    w = np.exp(-(E - E.min()) / kT)
    p = w / w.sum()
    rng = np.random.default_rng(seed)
    return rng.choice(len(E), size=nsamp, p=p)


def basis_rows(n, idx):
    """Unit wavepackets |idx_b> as a (B, N) complex128 array."""
    X = np.zeros((len(idx), n), dtype=np.complex128)
    X[np.arange(len(idx)), idx] = 1.0
    return X


class Ensemble:
    """(B, N) wavepackets on the device propagated with a shared operator.

    H: CarTens interaction (its field is set per step by ``step(F)``);
    H0: diagonal field-free CarTens (split operator) or None;
    f: exponent factor of the step (TDSE.exp_fac).
    """

    def __init__(self, H, H0, f, psi0, krylov_fixed=0, maxorder=100, tol=1e-15, stream=None):
        import torch
        from . import plan as _plan
        from .tdse import _plan_phase
        self.H = H
        self.f = complex(f)
        self.krylov_fixed = int(krylov_fixed)
        self.maxorder = int(maxorder)
        self.tol = float(tol)
        self.stream = stream
        self.plan = None
        self._active = None
        self.kernel = "auto"
        self.tuning = {}
        self._tuned = None
        self.precision = "fp64"
        dev = torch.device("cuda", torch.cuda.current_device())
        self.psi = torch.as_tensor(np.ascontiguousarray(psi0, dtype=np.complex128)).to(dev)
        self.out = torch.empty_like(self.psi)
        B = self.psi.shape[0]
        self.kused = torch.zeros(B, dtype=torch.int32, device=dev)
        self.status = torch.zeros(B, dtype=torch.int32, device=dev)
        self.phase = None
        if H0 is not None:
            h0 = H0.tomat(form="full", cart="0")
            d = torch.from_numpy(np.ascontiguousarray(h0.diagonal(), dtype=np.complex128)).to(dev)
            self.phase = torch.empty_like(d)
            _plan_phase(d, self.f, self.phase, stream)

    def set_field(self, F, thresh=None):
        """Field of the next step: products on the host, contraction on the
        device; switches plan when the set of active field components changes."""
        from . import plan as _plan
        self.H.field(F, thresh)
        active = _plan.active_carts(self.H)
        if active != self._active:
            self.plan = _plan.plan_for(self.H)
            self._apply_kernel()
            self._active = active
        self.plan.set_field(self.H, self.stream)

    def set_kernel(self, kernel):
        """'auto', 'tuned' (the autotuner's choice for these rows, 'auto' when
        never tuned), 'global', 'resident', 'ell', 'ell:<i>' or 'subspace'."""
        self.kernel = kernel
        self._tuned = None
        if self.plan is not None:
            self._apply_kernel()

    def set_tuning(self, **choices):
        """Launch choices of the subspace path (pair, coop_max, coop_grid,
        fork; see Plan.set_tuning) for this ensemble's calls."""
        self.tuning = dict(choices)
        if self.plan is not None:
            self._apply_kernel()

    def set_precision(self, precision):
        """'fp64' or 'fp32' (FP32 operator and matvec inputs, FP64 elsewhere)."""
        from . import _lib
        if precision not in _lib.PRECISIONS:
            raise ValueError(f"unknown precision '{precision}'")
        self.precision = precision
        if self.plan is not None:
            self._apply_kernel()

    def step(self, F, thresh=None):
        """Set the field of the interaction and advance every row by one step."""
        self.set_field(F, thresh)
        self.launch()

    def prepare(self, fields, thresh=None, per_term=False):
        """Device table of field products for a run of steps (one row per
        step) and the plan that covers every component the run activates.
        per_term: ``fields`` holds one (nsteps, 3) array per summed tensor
        (e.g. a static field on the dipole and a pulse on the polarizability)."""
        import torch
        from . import plan as _plan
        if per_term:
            fields = [None if F is None else np.asarray(F, dtype=np.float64) for F in fields]
        else:
            fields = np.asarray(fields, dtype=np.float64)
        active, cols, nst = _plan.field_table(self.H, fields, thresh, per_term=per_term)
        if active != self._active:
            self.plan = _plan.plan_for(self.H, active)
            self._apply_kernel()
            self._active = active
        tab = self.plan.table_array(cols, nst)
        return torch.from_numpy(tab).to(self.psi.device)

    def _apply_kernel(self):
        """Set the plan's kernel and precision; 'tuned' takes the autotuner's
        cached choice for this shape, or 'auto' when the shape was never tuned.
        Plans are shared between ensembles over the same operator, so this
        runs before every launch."""
        k, tun = self.kernel, getattr(self, "tuning", {})
        if k == "tuned":
            if getattr(self, "_tuned", None) is None:
                from . import autotune
                shape = autotune.domain_shape(self.plan, self.psi.cpu().numpy())
                self._tuned = autotune.lookup(self.plan, self.psi.shape[0], self.krylov_fixed, shape) or ("auto", {})
            k, tun = self._tuned
        if self.plan.kernel != k:
            self.plan.set_kernel(k)
        self.plan.set_tuning(**tun)
        if self.plan.precision != self.precision:
            self.plan.set_precision(self.precision)

    def propagate(self, fields, thresh=None, table=None, kused=None, check=True,
                  obs=None, expect=None, expect_sum=None):
        """Advance every row through len(fields) steps with one device call
        (rm_propagate).  ``table`` may hold a prepared field table.

        obs: an observable's device plan (``observable_plan``); after every
        step <psi_b|O|psi_b> goes to ``expect`` (device (nsteps, B)) and their
        sum over the rows to ``expect_sum`` (device (nsteps,)), both complex128
        -- the observable loop of the reference's examples inside the launch
        (rm_propagate_expect).  Pass True for either to have it allocated;
        the two are returned as (expect, expect_sum).

        The plan keeps the rows' domains from one call to the next (the
        rows never leave them); a call on rows rewritten in between is
        caught on the device and leaves them untouched.  check=True (the
        default) reads the status back: such a call is then repeated with
        fresh domains, and a cooperative timeout or (adaptive rule) maxorder
        raises.  check=False returns without waiting (call check() later)."""
        import torch
        if table is None:
            table = self.prepare(fields, thresh)
        th = None if thresh is None else [thresh]
        self._apply_kernel()
        if self.plan.domain_owner != id(self):
            self.plan.set_domain_reuse(True, owner=id(self))
        if obs is not None:
            nsteps, B, dev = table.shape[0], self.psi.shape[0], self.psi.device
            if expect is True or (expect is None and expect_sum is None):
                expect = torch.empty((nsteps, B), dtype=torch.complex128, device=dev)
            if expect_sum is True:
                expect_sum = torch.empty(nsteps, dtype=torch.complex128, device=dev)
        for attempt in (0, 1):
            self.plan.propagate(self.psi, self.out, table, self.f, phase=self.phase, thresh=th,
                                krylov_fixed=self.krylov_fixed, maxorder=self.maxorder, tol=self.tol,
                                kused=kused, status=self.status, stream=self.stream,
                                obs=obs, expect=expect, expect_sum=expect_sum)
            if not check:
                return (expect, expect_sum) if obs is not None else None
            from . import _lib
            if attempt == 0 and np.any(self.status.cpu().numpy() & _lib.RM_STALE_DOMAIN):
                self.plan.set_domain_reuse(True, owner=id(self))   # forget the kept domains
                continue
            break
        self.check()
        if obs is not None:
            return expect, expect_sum

    def load(self, psi0):
        """Replace the rows (same shape) by psi0 (numpy or device array)."""
        import torch
        src = torch.as_tensor(np.ascontiguousarray(psi0, dtype=np.complex128)) \
            if not isinstance(psi0, torch.Tensor) else psi0
        self.psi.copy_(src.to(self.psi.device))
        if self.kernel == "tuned":
            self._tuned = None   # the rows' domains may differ: look the shape up again
        if self.plan is not None and self.plan.domain_owner == id(self):
            self.plan.set_domain_reuse(True, owner=id(self))

    def launch(self):
        """The Krylov step of all rows with the field already on the device."""
        self._apply_kernel()
        self.plan.step(self.psi, self.out, self.f, phase=self.phase, krylov_fixed=self.krylov_fixed,
                       maxorder=self.maxorder, tol=self.tol, kused=self.kused, status=self.status,
                       stream=self.stream)
        self.psi, self.out = self.out, self.psi

    @staticmethod
    def observable_plan(O, F=(0.0, 0.0, 1.0)):
        """Device plan of an observable CarTens O with its matrix formed for
        field F (default: the "0"/Z component as tomat(cart="0") takes it,
        e.g. cos^2 theta); for Ensemble.propagate(obs=...) and expect()."""
        from . import plan as _plan
        O.field(list(F))
        p = _plan.plan_for(O)
        p.set_field(O)
        return p

    def expect(self, obs_plan, out=None):
        """<psi_b|O|psi_b> for every row (complex, device)."""
        import torch
        if out is None:
            out = torch.empty(self.psi.shape[0], dtype=torch.complex128, device=self.psi.device)
        obs_plan.expect(self.psi, out, self.stream)
        return out

    def check(self):
        """Raise on the status bits of the last call: a cooperative timeout,
        rows rewritten since the domains were kept, or (adaptive rule) a row
        that reached maxorder."""
        from . import _lib
        st = self.status.cpu().numpy()
        if np.any(st & _lib.RM_DEVICE_TIMEOUT):
            raise RuntimeError("a cooperative propagation launch timed out at a grid barrier; "
                               "its wavepackets are invalid")
        if np.any(st & _lib.RM_STALE_DOMAIN):
            raise RuntimeError("the rows changed since the last propagate call kept their domains "
                               "(use Ensemble.load); they were not propagated")
        if bool(np.any(st & _lib.RM_NOT_CONVERGED)):
            if self.plan is not None and getattr(self.plan, "propag", "internal") == "external":
                raise ExpokitError("maximum number of steps reached without convergence")
            if self.krylov_fixed == 0:
                raise ValueError(f"Lanczos reached maximum order of '{self.maxorder}' without convergence")
