"""Device plan: a CarTens operator flattened once and uploaded to the GPU.

The reference applies its operator block by block every matvec
(richmol/field.py:1212-1243) and re-contracts the M tensors with the field
every step (field.py:1115-1142).  Here the block structure is flattened once
into a sparse matrix in the full-form state order (field.py:659-692) whose
entries remember how they are made:

    A_e  = sum_p K_p * MF_{term_p}
    MF_t = thr( sum_q fprod[slot_q] * M_q )

so that per step only the field products (<= 32 scalars) go to the device,
where ``k_field_contract`` and ``k_assemble`` rebuild the values.
"""
import collections

import numpy as np

from . import _lib
from .field import block_offsets

_CACHE = collections.OrderedDict()
_CACHE_MAX = 16


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.RichmolAmdError("no HIP device visible: the propagation runs only on the GPU")
    return torch


def _stream_handle(stream=None):
    if stream is not None:
        return stream.cuda_stream
    torch = _torch()
    try:
        # the current stream of the current device, without torch's Python
        # wrapper around the same lookup
        sid, dev, dtype = torch._C._cuda_getCurrentStream(torch._C._cuda_getDevice())
        return torch.cuda.Stream(stream_id=sid, device_index=dev, device_type=dtype).cuda_stream
    except (AttributeError, TypeError):
        return torch.cuda.current_stream().cuda_stream


def leaf_carts(leaf):
    """Sorted Cartesian components present in a leaf tensor's M matrices
    (kept on the tensor for its structure id: a changed tensor gets a new
    one)."""
    got = leaf.__dict__.get("_carts_of")
    uid = leaf.__dict__.get("_uid")
    if got is not None and got[0] == uid:
        return got[1]
    out = sorted({c for Jp in leaf.mmat for sp in leaf.mmat[Jp]
                  for w in leaf.mmat[Jp][sp] for c in leaf.mmat[Jp][sp][w]})
    leaf.__dict__["_carts_of"] = (uid, out)
    return out


def _basis_signature(T):
    got = T.__dict__.get("_basis_sig")
    key = (T.__dict__.get("_uid"), id(T.Jlist1), id(T.Jlist2), id(T.dim1), id(T.dim2))
    if got is not None and got[0] == key:
        return got[1]
    sig = (tuple(T.Jlist1), tuple(T.Jlist2),
           tuple((J, tuple(T.symlist1[J]), tuple(T.dim1[J][s] for s in T.symlist1[J])) for J in T.Jlist1),
           tuple((J, tuple(T.symlist2[J]), tuple(T.dim2[J][s] for s in T.symlist2[J])) for J in T.Jlist2))
    T.__dict__["_basis_sig"] = (key, sig)
    return sig


def active_carts(T):
    """Per term of T, the Cartesian components whose field product is non-zero
    (after the field's threshold): only they can contribute to MF."""
    out = []
    for leaf, fp, _ in T.terms():
        if fp is None:
            out.append(())
        else:
            out.append(tuple(c for c in leaf_carts(leaf) if fp.get(c, 0) != 0))
    return tuple(out)


def structure_key(T, active=None):
    if active is None:
        active = active_carts(T)
    return (tuple(leaf._uid for leaf, _, _ in T.terms()), _basis_signature(T), active)


def flatten_operator(T, active=None):
    """Host arrays of the plan (see module docstring).  ``active`` (per term,
    the Cartesian components to include; default: those with non-zero field
    products) restricts the sparsity pattern to the entries the field can
    reach, e.g. Delta m = 0 only for a Z-polarised field."""
    if active is None:
        active = active_carts(T)
    off1, n1 = block_offsets(T, 1)
    off2, n2 = block_offsets(T, 2)
    R, C, PT, PK = [], [], [], []
    TF, TV, TL, TCNT = [], [], [], []
    nterm = 0
    slot_base = 0
    slots = []
    terms = T.terms()
    for li, (leaf, _, _) in enumerate(terms):
        carts = [c for c in leaf_carts(leaf) if c in active[li]]
        cidx = {c: slot_base + i for i, c in enumerate(carts)}
        slots.append(carts)
        for Jp in sorted(set(leaf.kmat.keys()) & set(leaf.mmat.keys())):
            J1, J2 = Jp
            for sp in sorted(set(leaf.kmat[Jp].keys()) & set(leaf.mmat[Jp].keys())):
                s1, s2 = sp
                if (J1, s1) not in off1 or (J2, s2) not in off2:
                    continue
                dk1, dk2 = T.dim_k1[J1][s1], T.dim_k2[J2][s2]
                dm2 = T.dim_m2[J2][s2]
                Kb, Mb = leaf.kmat[Jp][sp], leaf.mmat[Jp][sp]
                for w in sorted(set(Kb.keys()) & set(Mb.keys()), key=str):
                    K = Kb[w].tocoo()
                    nzk = K.data != 0
                    kr, kc, kv = K.row[nzk], K.col[nzk], K.data[nzk].astype(np.complex128)
                    if kv.size == 0:
                        continue
                    mr, mc, ms, mv = [], [], [], []
                    for c in sorted(Mb[w].keys()):
                        if c not in cidx:
                            continue
                        M = Mb[w][c].tocoo()
                        nz = M.data != 0
                        mr.append(M.row[nz]); mc.append(M.col[nz])
                        ms.append(np.full(int(nz.sum()), cidx[c], dtype=np.int32))
                        mv.append(M.data[nz].astype(np.complex128))
                    if not mr:
                        continue
                    mr = np.concatenate(mr); mc = np.concatenate(mc)
                    ms = np.concatenate(ms); mv = np.concatenate(mv)
                    if mr.size == 0:
                        continue
                    key = mr.astype(np.int64) * dm2 + mc
                    ukey, inv = np.unique(key, return_inverse=True)
                    nt = ukey.size
                    order = np.argsort(inv, kind="stable")
                    TF.append(ms[order]); TV.append(mv[order])
                    TCNT.append(np.bincount(inv, minlength=nt))
                    TL.append(np.full(nt, li, dtype=np.int32))
                    ur, uc = ukey // dm2, ukey % dm2
                    R.append((off1[(J1, s1)] + ur[:, None] * dk1 + kr[None, :]).ravel())
                    C.append((off2[(J2, s2)] + uc[:, None] * dk2 + kc[None, :]).ravel())
                    PT.append(np.broadcast_to((nterm + np.arange(nt))[:, None], (nt, kv.size)).ravel())
                    PK.append(np.broadcast_to(kv[None, :], (nt, kv.size)).ravel())
                    nterm += nt
        slot_base += len(carts)
    if R:
        R = np.concatenate(R).astype(np.int64); C = np.concatenate(C).astype(np.int64)
        PT = np.concatenate(PT).astype(np.int32); PK = np.concatenate(PK)
        TF = np.concatenate(TF).astype(np.int32); TV = np.concatenate(TV)
        TL = np.concatenate(TL).astype(np.int32); TCNT = np.concatenate(TCNT)
    else:
        R = C = np.zeros(0, np.int64); PT = TF = TL = np.zeros(0, np.int32)
        PK = TV = np.zeros(0, np.complex128); TCNT = np.zeros(0, np.int64)
    order = np.lexsort((C, R))
    R, C, PT, PK = R[order], C[order], PT[order], PK[order]
    if R.size:
        first = np.ones(R.size, dtype=bool)
        first[1:] = (R[1:] != R[:-1]) | (C[1:] != C[:-1])
        ent = np.flatnonzero(first)
    else:
        ent = np.zeros(0, np.int64)
    nnz = ent.size
    pair_ptr = np.append(ent, R.size).astype(np.int64)
    col_idx = C[ent].astype(np.int32)
    row_ptr = np.zeros(n1 + 1, dtype=np.int64)
    np.cumsum(np.bincount(R[ent], minlength=n1), out=row_ptr[1:])
    term_ptr = np.zeros(nterm + 1, dtype=np.int64)
    if nterm:
        np.cumsum(TCNT, out=term_ptr[1:])
    return dict(n=n1, n2=n2, nnz=nnz, row_ptr=row_ptr, col_idx=col_idx, pair_ptr=pair_ptr,
                pair_term=PT, pair_kval=np.ascontiguousarray(PK), nterm=nterm, term_ptr=term_ptr,
                term_field=TF, term_mval=np.ascontiguousarray(TV), term_leaf=TL,
                slots=slots, nfield=slot_base, nleaf=len(terms))


def matrix_arrays(A, cart):
    """Plan arrays (the layout of flatten_operator) of H(F) = p(F) A: one
    term whose M value is 1 in the single field slot ``cart``."""
    from scipy.sparse import csr_matrix
    A = csr_matrix(A, dtype=np.complex128)
    A.sort_indices()
    n = A.shape[0]
    if A.shape[1] != n:
        raise ValueError("the operator matrix must be square")
    nnz = A.nnz
    return dict(n=n, n2=n, nnz=nnz, row_ptr=A.indptr.astype(np.int64), col_idx=A.indices.astype(np.int32),
                pair_ptr=np.arange(nnz + 1, dtype=np.int64), pair_term=np.zeros(nnz, np.int32),
                pair_kval=np.ascontiguousarray(A.data), nterm=1, term_ptr=np.array([0, 1], np.int64),
                term_field=np.zeros(1, np.int32), term_mval=np.ones(1, np.complex128),
                term_leaf=np.zeros(1, np.int32), slots=[[cart]], nfield=1, nleaf=1)


class DevicePlan:
    """rm_plan handle plus the slot layout of the field products.

    A plan covers the field components active when it was built; set_field
    refuses a field that activates any other component (plan_for() then
    returns the plan for the new set)."""

    def __init__(self, T, active=None, arrays=None):
        import ctypes
        self.lib = _lib.load()
        self.kernel = "auto"
        self.precision = "fp64"
        self.propag = "internal"
        if arrays is None:
            self.active = active_carts(T) if active is None else active
            arr = flatten_operator(T, self.active)
        else:
            self.active, arr = active, arrays
        if arr["n"] != arr["n2"]:
            raise ValueError("bra and ket bases differ: the propagator needs a square operator")
        if arr["nfield"] > _lib.RM_MAX_FIELD:
            raise ValueError(f"{arr['nfield']} field slots exceed {_lib.RM_MAX_FIELD}")
        if arr["nleaf"] > _lib.RM_MAX_LEAF:
            raise ValueError(f"{arr['nleaf']} summed tensors exceed {_lib.RM_MAX_LEAF}")
        self.n = arr["n"]
        self.nnz = arr["nnz"]
        self.slots = arr["slots"]
        self.nfield = arr["nfield"]
        self.nleaf = arr["nleaf"]
        self.nterm = arr["nterm"]
        self._arrays = arr
        h = ctypes.c_void_p()
        c_i64, c_i32, c_dbl = ctypes.c_int64, ctypes.c_int32, ctypes.c_double
        _lib.check(self.lib.rm_plan_create(
            self.n, self.nnz, _lib.ptr(arr["row_ptr"], c_i64), _lib.ptr(arr["col_idx"], c_i32),
            _lib.ptr(arr["pair_ptr"], c_i64), _lib.ptr(arr["pair_term"], c_i32),
            _lib.ptr(arr["pair_kval"].view(np.float64), c_dbl), self.nterm,
            _lib.ptr(arr["term_ptr"], c_i64), _lib.ptr(arr["term_field"], c_i32),
            _lib.ptr(arr["term_mval"].view(np.float64), c_dbl), _lib.ptr(arr["term_leaf"], c_i32),
            self.nfield, self.nleaf, ctypes.byref(h)), "rm_plan_create")
        self.handle = h
        self.domain_owner = None
        self.tuning = dict(_lib.TUNE_DEFAULTS)
        _lib.track(self)

    @classmethod
    def from_matrix(cls, A, cart):
        """Plan of H(F) = p(F) A for one field product p (component label
        ``cart``, e.g. 'zz' for F_z^2): a single summed tensor with a single
        term, every entry of the sparse matrix A its K value."""
        return cls(None, active=[[cart]], arrays=matrix_arrays(A, cart))

    def close(self):
        """Destroy the device plan (idempotent)."""
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.rm_plan_destroy(h)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    KERNELS = {"auto": 0, "global": 1, "resident": 2, "ell": 3, "subspace": 4, "coop": 5, "block": 6}
    ELL_CFG0 = 16

    def set_kernel(self, variant):
        """Select the step kernel: 'auto', 'global', 'resident', 'ell',
        'ell:<i>' for ELL launch configuration i (see ell_configs()),
        'subspace' (propagate only: one wave per wavepacket on its invariant
        subspaces, a cooperative grid-wide launch per wavepacket for long
        domains; an error when it does not apply), 'coop' or 'block'
        (propagate only: every wavepacket through the cooperative kernel, or
        through the one-workgroup-per-wavepacket block kernel)."""
        if isinstance(variant, str) and variant.startswith("ell:"):
            code = self.ELL_CFG0 + int(variant[4:])
        else:
            code = self.KERNELS[variant]
        _lib.check(self.lib.rm_plan_set_kernel(self.handle, code), "rm_plan_set_kernel")
        self.kernel = variant

    PROPAGATORS = {"internal": 0, "external": 1}

    def set_propagator(self, propag):
        """'internal' (the Lanczos expv, default) or 'external' (Expokit's
        ZHEXPV as tdse.py:383-391 calls it; krylov_fixed is its Krylov size
        m, 12 in the reference) for propagate()."""
        _lib.check(self.lib.rm_plan_set_propagator(self.handle, self.PROPAGATORS[propag]),
                   "rm_plan_set_propagator")
        self.propag = propag

    def set_precision(self, precision):
        """'fp64' (default) or 'fp32': complex64 bulk matvec in the subspace
        kernel with FP64 Krylov recurrence, norms, exp(fT) and phases."""
        _lib.check(self.lib.rm_plan_set_precision(self.handle, _lib.PRECISIONS[precision]),
                   "rm_plan_set_precision")
        self.precision = precision

    def set_tuning(self, **choices):
        """Launch choices of rm_propagate's subspace path (rm_plan_set_tuning):
        pair, coop_max, coop_grid, fork; a choice not given returns to its
        default."""
        want = dict(_lib.TUNE_DEFAULTS)
        for k, v in choices.items():
            if k not in want:
                raise ValueError(f"unknown tuning choice '{k}' (use {sorted(want)})")
            want[k] = int(v)
        for k, v in want.items():
            if self.tuning.get(k) != v:
                _lib.check(self.lib.rm_plan_set_tuning(self.handle, _lib.TUNE_KNOBS[k], v), "rm_plan_set_tuning")
                self.tuning[k] = v

    def set_domain_reuse(self, on=True, owner=None):
        """Keep rm_propagate's column domains for its next call on the same
        wavepackets (checked on the device); every call forgets the kept
        ones.  ``owner`` records which caller's wavepackets they belong to."""
        _lib.check(self.lib.rm_plan_set_domain_reuse(self.handle, 1 if on else 0),
                   "rm_plan_set_domain_reuse")
        self.domain_owner = owner if on else None

    def set_coop_abort(self, preset):
        """Test hook: cooperative launches start with their abort flag raised."""
        _lib.check(self.lib.rm_plan_set_coop_abort(self.handle, 1 if preset else 0), "rm_plan_set_coop_abort")

    def domain_sizes(self, X):
        """Rows and stored entries of each column's domain (host, from the
        flattened pattern): the union of the connected components that the
        column's non-zero entries touch -- the work of the subspace kernel."""
        from scipy.sparse import csr_matrix
        from scipy.sparse.csgraph import connected_components
        arr = self._arrays
        n = arr["n"]
        A = csr_matrix((np.ones(arr["nnz"]), arr["col_idx"], arr["row_ptr"]), shape=(n, n))
        nc, lab = connected_components(A, directed=False)
        size = np.bincount(lab, minlength=nc)
        nnz = np.bincount(lab, weights=np.diff(arr["row_ptr"]), minlength=nc)
        rows, ents = [], []
        for x in np.asarray(X):
            comps = np.unique(lab[np.nonzero(x)[0]])
            rows.append(int(size[comps].sum()))
            ents.append(int(nnz[comps].sum()))
        return np.array(rows), np.array(ents)

    def subspaces(self):
        """(number of invariant subspaces, rows of the largest)."""
        import ctypes
        c, m = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self.lib.rm_plan_subspaces(self.handle, ctypes.byref(c), ctypes.byref(m)),
                   "rm_plan_subspaces")
        return c.value, m.value

    # ---------------------------------------------------------------- field
    def field_arrays(self, T):
        """Field products and thresholds of T's terms in slot order."""
        fprod = np.zeros(max(self.nfield, 1), dtype=np.complex128)
        thresh = np.full(max(self.nleaf, 1), -1.0)
        base = 0
        terms = T.terms()
        if len(terms) != len(self.slots):
            raise ValueError("operator does not match the plan's terms")
        for li, (leaf, fp, thr) in enumerate(terms):
            carts = self.slots[li]
            if fp is not None:
                for c, v in fp.items():
                    if v != 0 and c not in carts and c in leaf_carts(leaf):
                        raise ValueError(f"field component '{c}' is not in the plan "
                                         f"(built for {self.active[li]})")
                for i, c in enumerate(carts):
                    fprod[base + i] = complex(fp.get(c, 0.0))
            if thr is not None and thr > 0:
                thresh[li] = float(thr)
            base += len(carts)
        return fprod, thresh

    def set_field(self, T, stream=None):
        import ctypes
        fprod, thresh = self.field_arrays(T)
        _lib.check(self.lib.rm_plan_set_field(
            self.handle, _lib.ptr(fprod.view(np.float64), ctypes.c_double),
            _lib.ptr(thresh, ctypes.c_double), ctypes.c_void_p(_stream_handle(stream))),
            "rm_plan_set_field")

    # ------------------------------------------------------------- kernels
    def matvec(self, x, y, stream=None):
        """y[b] = H x[b] for torch complex128 (B, N) device tensors."""
        import ctypes
        _check_cols(x, self.n); _check_cols(y, self.n)
        _lib.check(self.lib.rm_matvec(self.handle, ctypes.c_void_p(x.data_ptr()),
                                      ctypes.c_void_p(y.data_ptr()), x.stride(0), x.shape[0],
                                      ctypes.c_void_p(_stream_handle(stream))), "rm_matvec")

    def step(self, psi_in, psi_out, f, phase=None, krylov_fixed=0, maxorder=100, tol=1e-15,
             kused=None, status=None, stream=None):
        import ctypes
        _check_cols(psi_in, self.n); _check_cols(psi_out, self.n)
        if psi_in.stride(0) != psi_out.stride(0):
            raise ValueError("psi_in and psi_out need the same row stride")
        vp = ctypes.c_void_p
        _lib.check(self.lib.rm_step(
            self.handle, vp(psi_in.data_ptr()), vp(psi_out.data_ptr()), psi_in.stride(0),
            psi_in.shape[0], float(f.real), float(f.imag),
            vp(phase.data_ptr() if phase is not None else None), int(krylov_fixed), int(maxorder),
            float(tol), vp(kused.data_ptr() if kused is not None else None),
            vp(status.data_ptr() if status is not None else None),
            vp(_stream_handle(stream))), "rm_step")

    def timing(self, enable=True):
        """Record HIP events around every Krylov-step launch of this plan."""
        _lib.check(self.lib.rm_plan_timing(self.handle, 1 if enable else 0), "rm_plan_timing")

    def kernel_time(self):
        """(summed launch time in ms, number of launches) since the last call."""
        import ctypes
        ms, nl = ctypes.c_double(), ctypes.c_int64()
        _lib.check(self.lib.rm_plan_kernel_time(self.handle, ctypes.byref(ms), ctypes.byref(nl)),
                   "rm_plan_kernel_time")
        return ms.value, nl.value

    def table_array(self, cols, nsteps):
        """(nsteps, nfield) complex table in slot order from field_table()."""
        tab = np.zeros((nsteps, max(self.nfield, 1)), dtype=np.complex128)
        base = 0
        for li, carts in enumerate(self.slots):
            for c, v in cols[li].items():
                if c not in carts:
                    if np.any(np.asarray(v) != 0):
                        raise ValueError(f"field component '{c}' is not in the plan")
                    continue
                tab[:, base + carts.index(c)] = v
            base += len(carts)
        return tab

    def propagate(self, psi, work, table, f, phase=None, thresh=None, krylov_fixed=0,
                  maxorder=100, tol=1e-15, kused=None, status=None, stream=None,
                  obs=None, expect=None, expect_sum=None):
        """len(table) steps of every row of psi (in place); table: device
        (nsteps, nfield) complex128 tensor of field products.

        obs: a DevicePlan of an observable (its field set), read after every
        step into expect (device (nsteps, ncols) complex128) and/or
        expect_sum (device (nsteps,) complex128, the sum over the rows) --
        rm_propagate_expect."""
        import ctypes
        import torch
        _check_cols(psi, self.n); _check_cols(work, self.n)
        if psi.stride(0) != work.stride(0):
            raise ValueError("psi and work need the same row stride")
        if thresh is None:
            # no thresholds: one array kept per plan (its address passed every call)
            if getattr(self, "_th_none", None) is None:
                self._th_none = np.full(max(self.nleaf, 1), -1.0)
                self._th_none_ptr = _lib.ptr(self._th_none, ctypes.c_double)
            th_ptr = self._th_none_ptr
        else:
            th = np.full(max(self.nleaf, 1), -1.0)
            th[: len(thresh)] = [t if (t is not None and t > 0) else -1.0 for t in thresh]
            th_ptr = _lib.ptr(th, ctypes.c_double)
        vp = ctypes.c_void_p
        args = (self.handle, vp(psi.data_ptr()), vp(work.data_ptr()), psi.stride(0), psi.shape[0],
                table.shape[0], vp(table.data_ptr()), th_ptr,
                float(f.real), float(f.imag), vp(phase.data_ptr() if phase is not None else None),
                int(krylov_fixed), int(maxorder), float(tol),
                vp(kused.data_ptr() if kused is not None else None),
                vp(status.data_ptr() if status is not None else None))
        if obs is None:
            if expect is not None or expect_sum is not None:
                raise ValueError("expect / expect_sum need an observable plan")
            name = "rm_propagate"
            rc = _lib.check(self.lib.rm_propagate(*args, vp(_stream_handle(stream))), name)
        else:
            nsteps, ncols = table.shape[0], psi.shape[0]
            for t, shape in ((expect, (nsteps, ncols)), (expect_sum, (nsteps,))):
                if t is None:
                    continue
                if t.dtype != torch.complex128 or not t.is_cuda or tuple(t.shape) != shape or not t.is_contiguous():
                    raise ValueError(f"expected a contiguous device complex128 tensor of shape {shape}")
            name = "rm_propagate_expect"
            rc = _lib.check(self.lib.rm_propagate_expect(
                *args, obs.handle, vp(expect.data_ptr() if expect is not None else None),
                vp(expect_sum.data_ptr() if expect_sum is not None else None),
                vp(_stream_handle(stream))), name)
        if rc == _lib.RM_DEVICE_TIMEOUT:
            raise RuntimeError(name + ": " + self.lib.rm_last_error().decode(errors="replace"))

    def expect(self, psi, out, stream=None):
        import ctypes
        _check_cols(psi, self.n)
        _lib.check(self.lib.rm_expect(self.handle, ctypes.c_void_p(psi.data_ptr()), psi.stride(0),
                                      psi.shape[0], ctypes.c_void_p(out.data_ptr()),
                                      ctypes.c_void_p(_stream_handle(stream))), "rm_expect")


def _check_cols(t, n):
    import torch
    if t.dtype != torch.complex128:
        raise TypeError(f"expected complex128 wavepackets, got {t.dtype}")
    if not t.is_cuda:
        raise TypeError("wavepackets must be device tensors")
    if t.dim() != 2 or t.shape[1] != n or t.stride(1) != 1 or t.stride(0) < n:
        raise ValueError(f"expected (ncols, {n}) rows with unit inner stride, got {tuple(t.shape)}")


def cart_products(cart, F):
    """Field products of one Cartesian string for a (nsteps, 3) field array
    (field.py:1094-1105): the product of the named components, "0" -> 1."""
    F = np.asarray(F)
    if cart == "0" or cart == "":
        return np.ones(F.shape[0], dtype=np.result_type(F.dtype, np.float64))
    out = F[:, "xyz".index(cart[0])].copy()
    for ch in cart[1:]:
        out = out * F[:, "xyz".index(ch)]
    return out


def field_table(T, fields, thresh=None, per_term=False):
    """Per-step field products for every term of T.

    fields: one (nsteps, 3) array for every term, or with per_term=True a
    list with one entry per term (an array, or None to keep that term's
    present field products).  thresh screens the products as
    CarTens.field(F, thresh) does.  Returns (active, cols, nsteps): per term
    the active components and {cart: products (array or scalar)}.
    """
    terms = T.terms()
    if not per_term:
        fields = [fields] * len(terms)
    if len(fields) != len(terms):
        raise ValueError(f"{len(fields)} field arrays for {len(terms)} terms")
    cols, active, nst = [], [], None
    for (leaf, fp, thr0), F in zip(terms, fields):
        carts = leaf_carts(leaf)
        cdict = {}
        if F is None:
            for c in carts:
                v = 0.0 if fp is None else complex(fp.get(c, 0.0))
                cdict[c] = v
        else:
            F = np.asarray(F)
            nst = F.shape[0] if nst is None else nst
            if F.shape[0] != nst:
                raise ValueError("per-term field arrays differ in length")
            thr = thresh
            for c in carts:
                v = cart_products(c, F)
                if thr is not None:
                    v = np.where(np.abs(v) >= thr, v, 0)
                cdict[c] = v
        cols.append(cdict)
        active.append(tuple(c for c in carts if np.any(np.asarray(cdict[c]) != 0)))
    return tuple(active), cols, nst


def plan_for(T, active=None):
    """Cached DevicePlan for T's structure (leaf tensors, basis, active field
    components; by default the components T's present field activates)."""
    if active is None:
        active = active_carts(T)
    key = structure_key(T, active)
    p = _CACHE.get(key)
    if p is None:
        p = DevicePlan(T, active)
        _CACHE[key] = p
        while len(_CACHE) > _CACHE_MAX:
            _CACHE.popitem(last=False)
    else:
        _CACHE.move_to_end(key)
    return p


def apply(T, X):
    """H X for a numpy (B, N) array through the device (CarTens.vec)."""
    torch = _torch()
    p = plan_for(T)
    x = torch.from_numpy(np.ascontiguousarray(X, dtype=np.complex128)).cuda()
    y = torch.empty_like(x)
    p.set_field(T)
    p.matvec(x, y)
    return y.cpu().numpy()


def ell_configs():
    """[(threads, rows per thread, wavepackets per workgroup, operator in LDS)]."""
    import ctypes
    lib = _lib.load()
    out = []
    i = 0
    while True:
        v = [ctypes.c_int32() for _ in range(4)]
        if lib.rm_ell_config(i, *[ctypes.byref(x) for x in v]) != 0:
            return out
        out.append(tuple(x.value for x in v))
        i += 1
