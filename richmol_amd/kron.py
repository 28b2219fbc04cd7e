"""Kronecker-block operators for large rigid-rotor bases.

richmol applies a tensor as Y[J1] += MF_w . X[J2] . K_w^T per (J1, J2) block
pair and irrep w (field.py:1235-1237), X[J] being the wavepacket's (m, k)
block.  Flattened into one sparse matrix that factorisation costs dk1*dm
entries per row -- about 3e9 stored entries for an asymmetric top at
J <= 80 with all m -- so for large bases the operator is kept factored:

* per J the states are laid out m-major with the four D2 species' k-states
  concatenated (dk = 2J+1 columns), so X[J] is one (2J+1) x (2J+1) matrix;
* per (tensor, J1, J2, w): K_w = C1^H P_w C2 dense over all species (C: the
  species' eigenvector columns over primitive k, P_w the primitive-k matrix
  of labtens.py:482-496), and M_w,cart as its non-zero diagonals m1 - m2 =
  sigma (labtens.py:501-523; |sigma| <= w);
* per step the field products weight the M diagonals into MF_w, and the
  matvec is, per J1, one dense product of the stacked MF_w X[J2] blocks with
  the stacked K_w -- a GEMM over every wavepacket at once.

The matrix elements are those of rot.tensor (same primitives, thresholds and
signs); the 3j symbols come from wigner3j_array (extended-precision
log-factorials, a few 1e-15 relative) instead of exact rationals.
"""
import math

import numpy as np

from . import rot

_NLF = 512
_LF = np.zeros(_NLF + 1, dtype=np.longdouble)
_LF[1:] = np.cumsum(np.log(np.arange(1, _NLF + 1, dtype=np.longdouble)))


def wigner3j_array(j1, j2, j3, m1, m2, m3):
    """Wigner 3j symbols for integer arrays (broadcast), by the Racah sum with
    log-factorials held in extended precision (relative error a few 1e-15
    for the small second angular momenta used here).  Values below 1e-14 in
    magnitude -- left by terms that cancel exactly -- are returned as 0."""
    j1, j2, j3, m1, m2, m3 = np.broadcast_arrays(*[np.asarray(x, dtype=np.int64) for x in (j1, j2, j3, m1, m2, m3)])
    if np.any(j1 + j2 + j3 > _NLF - 1):
        raise ValueError("angular momenta too large for the log-factorial table")
    valid = ((m1 + m2 + m3 == 0) & (j3 >= np.abs(j1 - j2)) & (j3 <= j1 + j2) & (np.abs(m1) <= j1)
             & (np.abs(m2) <= j2) & (np.abs(m3) <= j3))

    def lf(n):
        return _LF[np.clip(n, 0, _NLF)]

    kmin = np.maximum.reduce([np.zeros_like(j1), j2 - j3 - m1, j1 - j3 + m2])
    kmax = np.minimum.reduce([j1 + j2 - j3, j1 - m1, j2 + m2])
    logpre = 0.5 * (lf(j1 + j2 - j3) + lf(j1 - j2 + j3) + lf(-j1 + j2 + j3) - lf(j1 + j2 + j3 + 1)
                    + lf(j1 + m1) + lf(j1 - m1) + lf(j2 + m2) + lf(j2 - m2) + lf(j3 + m3) + lf(j3 - m3))
    s = np.zeros(j1.shape, dtype=np.longdouble)
    nt = int(np.max(np.where(valid, kmax - kmin, 0))) + 1 if valid.any() else 0
    for t in range(nt):
        k = kmin + t
        use = valid & (k <= kmax)
        logden = (lf(k) + lf(j3 - j2 + k + m1) + lf(j3 - j1 + k - m2) + lf(j1 + j2 - j3 - k)
                  + lf(j1 - k - m1) + lf(j2 - k + m2))
        s += np.where(use, np.where(k % 2 == 0, 1.0, -1.0) * np.exp(logpre - logden), 0.0)
    sign = np.where((j1 - j2 - m3) % 2 == 0, 1.0, -1.0)
    v = np.where(valid, sign * s, 0.0).astype(np.float64)
    v[np.abs(v) < 1e-14] = 0.0
    return v


def _kprim(comps_w, w, J1, J2):
    """P_w[k1, k2] = sum_s 3j(J2 w J1; k2 s -k1) UsT_s (-1)^|k1| with
    k1 = k2 + s (rot._k_prim, all primitive k of both J)."""
    P = np.zeros((2 * J1 + 1, 2 * J2 + 1), dtype=np.complex128)
    k2 = np.arange(-J2, J2 + 1)
    for s, u in comps_w.items():
        k1 = k2 + s
        ok = np.abs(k1) <= J1
        t = wigner3j_array(J2, w, J1, k2[ok], s, -k1[ok])
        P[k1[ok] + J1, k2[ok] + J2] += t * u * np.where(np.abs(k1[ok]) % 2 == 0, 1.0, -1.0)
    return P


def _mdiag(ucols_w, w, J1, J2, ncart):
    """Diagonals of M_w,cart (rot._m_block with all m): for every sigma,
    the (ncart, 2J1+1) values at m2 = m1 - sigma (0 where m2 is out of
    range)."""
    fac = math.sqrt((2 * J1 + 1) * (2 * J2 + 1))
    m1 = np.arange(-J1, J1 + 1)
    out = {}
    for s, col in ucols_w.items():
        m2 = m1 - s
        ok = np.abs(m2) <= J2
        t = np.zeros(m1.size)
        t[ok] = wigner3j_array(J2, w, J1, m2[ok], s, -m1[ok])
        base = t * fac * np.where(np.abs(m1) % 2 == 0, 1.0, -1.0)
        d = np.asarray(col, dtype=np.complex128)[:ncart, None] * base[None, :]
        d[np.abs(d) < rot._EPS] = 0
        if np.any(d):
            out[s] = d
    return out


class KronOperator:
    """Factored operator of summed rank-1/2 tensors on a rotor basis with
    every J carrying all m = -J..J.

    blocks: list of (leaf, J1, J2, w, K, mdiag) with K (2J1+1, 2J2+1) complex
    and mdiag {sigma: (ncart, 2J1+1) complex}; carts[leaf]: the Cartesian
    labels of the leaf's M components; layout: Jlist, per-J state offsets
    (internal order: m-major, species-concatenated k) and the permutation
    `perm` with internal index i = perm[full-form index].
    """

    def __init__(self, basis, tensors, scales=None):
        self.Jlist = list(basis.Jlist)
        for J in self.Jlist:
            if list(basis.mlist[J]) != list(range(-J, J + 1)):
                raise ValueError("KronOperator needs every J with all m = -J..J")
        if scales is None:
            scales = [1.0] * len(tensors)
        self.carts, self.ranks, self.blocks = [], [], []
        C = {J: np.concatenate([basis.kfun[J][s][1] for s in basis.symlist[J]], axis=1) for J in self.Jlist}
        for J in self.Jlist:
            if C[J].shape != (2 * J + 1, 2 * J + 1):
                raise ValueError("KronOperator needs the full k space of every J")
        for li, (tens, scale) in enumerate(zip(tensors, scales)):
            rank, comps, ucols = rot.spherical_components(tens)
            carts = rot.CART[rank]
            self.carts.append(list(carts))
            self.ranks.append(rank)
            dJ = max(comps.keys()) if comps else 0
            for J2 in self.Jlist:
                for J1 in self.Jlist:
                    if abs(J1 - J2) > dJ:
                        continue
                    for w in sorted(comps):
                        if abs(J1 - J2) > w:
                            continue
                        K = C[J1].conj().T @ _kprim(comps[w], w, J1, J2) @ C[J2]
                        K[np.abs(K) < rot._EPS] = 0
                        if not np.any(K):
                            continue
                        md = _mdiag(ucols[w], w, J1, J2, len(carts))
                        if not md:
                            continue
                        self.blocks.append((li, J1, J2, w, K * scale, md))
        # internal layout and the permutation from the reference's full-form order
        off, pos = {}, 0
        for J in self.Jlist:
            off[J] = pos
            pos += (2 * J + 1) ** 2
        self.off, self.n = off, pos
        perm = np.empty(pos, dtype=np.int64)
        full = 0
        for J in self.Jlist:
            dm = 2 * J + 1
            kc = 0
            for s in basis.symlist[J]:
                dks = basis.kfun[J][s][1].shape[1]
                im, ik = np.meshgrid(np.arange(dm), np.arange(dks), indexing="ij")
                perm[full:full + dm * dks] = (off[J] + im * (2 * J + 1) + kc + ik).ravel()
                full += dm * dks
                kc += dks
        self.perm = perm
        self.energies = np.empty(pos)
        for J in self.Jlist:
            e = np.concatenate([basis.energy[J][s] for s in basis.symlist[J]])
            self.energies[off[J]:off[J] + (2 * J + 1) ** 2] = np.tile(e, 2 * J + 1)

    @classmethod
    def from_cartens(cls, H):
        """The factored operator of a CarTens leaf or sum as richmol stores it
        (kmat[(J1, J2)][(s1, s2)][w] K blocks per species pair, mmat[...][w][cart]
        M matrices, field.py:58-170): per (leaf, J1, J2, w) the species' K
        blocks assembled over the concatenated k of each J (symlist order) and
        the M matrix as its diagonals m1 - m2 = sigma.  Needs every J with all
        m = -J..J and the full k space (2J + 1 states per m); returns the
        operator and the per-leaf field-product dicts of H's terms."""
        terms = H.terms()
        leaf0 = terms[0][0]
        Jlist = list(leaf0.Jlist1)
        if list(leaf0.Jlist2) != Jlist:
            raise ValueError("the Kronecker-block form needs the same basis on both sides")
        sym = {J: list(leaf0.symlist1[J]) for J in Jlist}
        dkc = {}
        for J in Jlist:
            dms = {leaf0.dim_m1[J][s] for s in sym[J]}
            dkc[J] = np.concatenate([[0], np.cumsum([leaf0.dim_k1[J][s] for s in sym[J]])]).astype(int)
            if dms != {2 * J + 1} or dkc[J][-1] != 2 * J + 1:
                raise ValueError("the Kronecker-block form needs every J with all m and the full k space")
        self = cls.__new__(cls)
        self.Jlist = Jlist
        self.carts, self.ranks, self.blocks = [], [], []
        fprods = []
        for li, (leaf, fprod, thr) in enumerate(terms):
            if fprod is None:
                raise AttributeError("tensor must be multiplied with a field before it is applied")
            if thr is not None:
                raise ValueError("field thresholds (CarTens.field(F, thresh)) are not supported by the "
                                 "Kronecker-block form")
            carts = list(leaf.cart)
            self.carts.append(carts)
            self.ranks.append(leaf.rank)
            fprods.append(fprod)
            for (J1, J2), blk in leaf.kmat.items():
                ws = sorted({w for d in blk.values() for w in d})
                for w in ws:
                    K = np.zeros((2 * J1 + 1, 2 * J2 + 1), dtype=np.complex128)
                    Mw = None
                    for (s1, s2), d in blk.items():
                        if w not in d:
                            continue
                        i1, i2 = sym[J1].index(s1), sym[J2].index(s2)
                        K[dkc[J1][i1]:dkc[J1][i1 + 1], dkc[J2][i2]:dkc[J2][i2 + 1]] = d[w].toarray()
                        if Mw is None:
                            Mw = leaf.mmat[(J1, J2)][(s1, s2)][w]
                    if not np.any(K) or not Mw:
                        continue
                    md = {}
                    for ci, c in enumerate(carts):
                        if c not in Mw:
                            continue
                        M = Mw[c].tocoo()
                        sig = (M.row - J1) - (M.col - J2)
                        if np.any(np.abs(sig) > 2):
                            raise ValueError("M blocks with |m1 - m2| > 2: not a rank <= 2 tensor in m order")
                        for sv in np.unique(sig):
                            d = md.setdefault(int(sv), np.zeros((len(carts), 2 * J1 + 1), dtype=np.complex128))
                            sel = sig == sv
                            d[ci, M.row[sel]] += M.data[sel]
                    md = {sv: d for sv, d in md.items() if np.any(d)}
                    if md:
                        self.blocks.append((li, J1, J2, w, K, md))
        off, pos = {}, 0
        for J in Jlist:
            off[J] = pos
            pos += (2 * J + 1) ** 2
        self.off, self.n = off, pos
        perm = np.empty(pos, dtype=np.int64)
        full = 0
        for J in Jlist:
            dm = 2 * J + 1
            for i, s in enumerate(sym[J]):
                dks = dkc[J][i + 1] - dkc[J][i]
                im, ik = np.meshgrid(np.arange(dm), np.arange(dks), indexing="ij")
                perm[full:full + dm * dks] = (off[J] + im * dm + dkc[J][i] + ik).ravel()
                full += dm * dks
        self.perm = perm
        self.energies = None
        return self, fprods

    def slot_row(self, fprods):
        """Field products of one step in slot order from per-leaf dicts
        (CarTens.field products; a missing label is 0)."""
        return np.array([complex(fp.get(c, 0.0)) for fp, carts in zip(fprods, self.carts) for c in carts],
                        dtype=np.complex128)

    def field_products(self, fields):
        """Per leaf {cart: product of field components} for F (3,) per leaf."""
        out = []
        for li, F in enumerate(fields):
            F = np.asarray(F, dtype=np.float64)
            out.append({c: float(np.prod([F["xyz".index(ch)] for ch in c])) for c in self.carts[li]})
        return out

    def matvec(self, X, fprod):
        """Y = H(F) X on internal-layout rows X (B, n) (host reference)."""
        X = np.asarray(X)
        Y = np.zeros_like(X, dtype=np.complex128)
        B = X.shape[0]
        for li, J1, J2, w, K, md in self.blocks:
            dm1, dm2 = 2 * J1 + 1, 2 * J2 + 1
            x = X[:, self.off[J2]:self.off[J2] + dm2 * dm2].reshape(B, dm2, dm2)
            t = x @ K.T                                      # (B, m2, k1)
            y = np.zeros((B, dm1, 2 * J1 + 1), dtype=np.complex128)
            fp = np.array([fprod[li][c] for c in self.carts[li]])
            for s, d in md.items():
                mf = fp @ d                                  # (2J1+1,) MF values at m2 = m1 - s
                m1 = np.arange(dm1)
                m2 = m1 - J1 - s + J2
                ok = (m2 >= 0) & (m2 < dm2)
                y[:, m1[ok], :] += mf[ok][None, :, None] * t[:, m2[ok], :]
            Y[:, self.off[J1]:self.off[J1] + dm1 * dm1] += y.reshape(B, -1)
        return Y


def slot_products(kop, fields):
    """Field products of every step in slot order: fields[leaf] (nsteps, 3)
    -> (nsteps, nslot) complex, slot = (leaf, Cartesian label) in the order
    of kop.carts (CarTens.field, field.py:1095-1110: the product of the
    field components named by the label)."""
    cols = []
    for li, carts in enumerate(kop.carts):
        F = np.asarray(fields[li], dtype=np.float64).reshape(-1, 3)
        for c in carts:
            v = np.ones(F.shape[0])
            for ch in c:
                v = v * F[:, "xyz".index(ch)]
            cols.append(v)
    return np.ascontiguousarray(np.stack(cols, axis=1).astype(np.complex128))


class DeviceKronPlan:
    """A KronOperator on the device (rm_kplan of include/richmol_amd.h):
    per output J the stacked K blocks in real form for the dense products,
    the M diagonals and the block/J tables.  Wavepackets are rows of a
    (ncols, n) complex128 device tensor in the operator's internal order
    (kop.perm maps the reference's full-form index to it)."""

    def __init__(self, kop):
        import ctypes
        from . import _lib
        self.kop = kop
        self._lib = _lib
        lib = _lib.load()
        Js = np.asarray(kop.Jlist, dtype=np.int32)
        base = np.concatenate([[0], np.cumsum([len(c) for c in kop.carts])]).astype(np.int32)
        self.nslot = int(base[-1])
        J1, J2, slot0, ncart, k_off, sig_off, sig, d_off = [], [], [], [], [0], [0], [], [0]
        Ks, Ds = [], []
        for li, j1, j2, w, K, md in kop.blocks:
            J1.append(j1)
            J2.append(j2)
            slot0.append(base[li])
            ncart.append(len(kop.carts[li]))
            Ks.append(np.ascontiguousarray(K, dtype=np.complex128).ravel())
            k_off.append(k_off[-1] + K.size)
            ss = sorted(md)
            sig.extend(ss)
            sig_off.append(sig_off[-1] + len(ss))
            for s in ss:
                Ds.append(np.ascontiguousarray(md[s], dtype=np.complex128).ravel())
            d_off.append(d_off[-1] + sum(md[s].size for s in ss))
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        i64 = lambda a: np.ascontiguousarray(a, dtype=np.int64)
        K = np.concatenate(Ks) if Ks else np.zeros(1, np.complex128)
        D = np.concatenate(Ds) if Ds else np.zeros(1, np.complex128)
        arrs = [i32(J1), i32(J2), i32(slot0), i32(ncart), i64(k_off), K.view(np.float64), i64(sig_off),
                i32(sig), i64(d_off), D.view(np.float64)]
        self._keep = arrs
        h = ctypes.c_void_p()
        P = _lib.ptr
        _lib.check(lib.rm_kplan_create(
            len(Js), P(Js, ctypes.c_int32), self.nslot, len(J1), P(arrs[0], ctypes.c_int32),
            P(arrs[1], ctypes.c_int32), P(arrs[2], ctypes.c_int32), P(arrs[3], ctypes.c_int32),
            P(arrs[4], ctypes.c_int64), P(arrs[5], ctypes.c_double), P(arrs[6], ctypes.c_int64),
            P(arrs[7], ctypes.c_int32), P(arrs[8], ctypes.c_int64), P(arrs[9], ctypes.c_double),
            ctypes.byref(h)), "rm_kplan_create")
        self._keep = None
        self.handle = h
        _lib.track(self)
        n = ctypes.c_int64()
        fl = ctypes.c_double()
        _lib.check(lib.rm_kplan_info(h, ctypes.byref(n), ctypes.byref(fl)), "rm_kplan_info")
        self.n = int(n.value)
        self.gemm_flops_per_col = float(fl.value)
        self.kernel = "gemm"

    def close(self):
        """Destroy the device plan (idempotent)."""
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self._lib.load().rm_kplan_destroy(h)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    KERNELS = {"fused": 0, "gemm": 1, "mfma": 2}

    def set_kernel(self, kernel):
        """'gemm' (default): the m-mixed input stored, rocBLAS per output J;
        'mfma': the m-mixed input stored, the library's own matrix-core
        product per J chunk; 'fused': one MFMA launch per matvec, the m-mixed
        input built in LDS."""
        self._lib.check(self._lib.load().rm_kplan_set_kernel(self.handle, self.KERNELS[kernel]),
                        "rm_kplan_set_kernel")
        self.kernel = kernel

    def matvec(self, x, y, fprod_row, precision="fp64", stream=None):
        """y = H(F) x (device complex128 (ncols, n) tensors); fprod_row:
        device complex128 (nslot,)."""
        import ctypes
        from .plan import _check_cols, _stream_handle
        _check_cols(x, self.n)
        _check_cols(y, self.n)
        self._lib.check(self._lib.load().rm_kplan_matvec(
            self.handle, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), x.stride(0), x.shape[0],
            ctypes.c_void_p(fprod_row.data_ptr()), self._lib.PRECISIONS[precision],
            ctypes.c_void_p(_stream_handle(stream))), "rm_kplan_matvec")

    def phases(self, f, stream=None):
        """exp(f/2 E) on the device in the internal order (tdse.py:373)."""
        import torch
        from .tdse import _plan_phase
        d = torch.from_numpy(self.kop.energies.astype(np.complex128)).cuda()
        out = torch.empty_like(d)
        _plan_phase(d, complex(f), out, stream)
        return out

    def propagate(self, psi, fprod, f, phase=None, krylov_fixed=12, precision="fp64", stream=None,
                  maxorder=100, tol=1e-15, kused=None, status=None, obs=None, obs_fprod=None,
                  expect=None, expect_sum=None):
        """len(fprod) steps of TDSE.update on the rows of psi (in place);
        fprod: host (nsteps, nslot) complex from slot_products.  krylov_fixed
        0: the reference's adaptive rule with maxorder and tol; kused (device
        int32 (nsteps, ncols)) and status (device int32 (ncols,)) optional.

        obs: a DeviceKronPlan of an observable over the same J list, with
        obs_fprod (host (nslot,) complex, its fixed products); after every
        step <psi_b|O|psi_b> goes to expect (device (nsteps, ncols)) and the
        sums over the rows to expect_sum (device (nsteps,)), complex128
        (rm_kplan_propagate_expect)."""
        import ctypes
        import torch
        from .plan import _check_cols, _stream_handle
        _check_cols(psi, self.n)
        fp = np.ascontiguousarray(fprod, dtype=np.complex128).reshape(-1, max(self.nslot, 1))
        vp = lambda t: ctypes.c_void_p(t.data_ptr() if t is not None else 0)
        args = (self.handle, ctypes.c_void_p(psi.data_ptr()), psi.stride(0), psi.shape[0], fp.shape[0],
                self._lib.ptr(fp.view(np.float64), ctypes.c_double), float(complex(f).real),
                float(complex(f).imag), vp(phase), int(krylov_fixed), int(maxorder), float(tol), vp(kused),
                vp(status), self._lib.PRECISIONS[precision])
        lib = self._lib.load()
        if obs is None:
            self._lib.check(lib.rm_kplan_propagate(*args, ctypes.c_void_p(_stream_handle(stream))),
                            "rm_kplan_propagate")
            return
        nsteps, ncols = fp.shape[0], psi.shape[0]
        for t, shape in ((expect, (nsteps, ncols)), (expect_sum, (nsteps,))):
            if t is not None and (t.dtype != torch.complex128 or not t.is_cuda or tuple(t.shape) != shape
                                  or not t.is_contiguous()):
                raise ValueError(f"expected a contiguous device complex128 tensor of shape {shape}")
        ofp = np.ascontiguousarray(obs_fprod, dtype=np.complex128).reshape(-1)
        if ofp.size != max(obs.nslot, 1) and obs.nslot:
            raise ValueError(f"obs_fprod: {obs.nslot} products expected, got {ofp.size}")
        self._lib.check(lib.rm_kplan_propagate_expect(
            *args, obs.handle, self._lib.ptr(ofp.view(np.float64), ctypes.c_double), vp(expect),
            vp(expect_sum), ctypes.c_void_p(_stream_handle(stream))), "rm_kplan_propagate_expect")
