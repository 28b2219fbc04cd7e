"""Lab-frame Cartesian tensor operator: the data contract of richmol's CarTens.

Mirrors the interface of richmol/field.py:25-1301 for the propagation path:
attributes ``Jlist1/2, symlist1/2, dim1/2, dim_m1/2, dim_k1/2, kmat, mmat,
cart, rank, os`` and the methods ``field``, ``vec``, ``tomat``,
``full_form``, ``mul``, ``add_cartens`` and the ``*``, ``+``, ``-``
operators.

What differs from the reference is where the work happens.  ``field(F)``
records the field products (3**rank host scalars) and does not build the
contracted M matrices; the device plan (``richmol_amd.plan``) contracts them
on the GPU each step.  ``vec`` runs on the GPU.  ``mfmat`` is kept as a host
view, assembled on first access for inspection and ``tomat()``; the
propagation never reads it.
"""
import copy
import itertools
from collections import defaultdict

import numpy as np
import scipy.sparse
from scipy.sparse import csr_matrix, kron

_uid = itertools.count(1)


def _nested():
    return defaultdict(_nested)


def field_products(rank, field, thresh=None):
    """Field products of CarTens.field (field.py:1090-1108): one entry per
    Cartesian string of length ``rank`` plus "0" -> 1, optionally screened."""
    try:
        fx, fy, fz = field[:3]
        fxyz = np.array([fx, fy, fz])
    except (TypeError, IndexError):
        raise IndexError(
            "field variable must be an iterable with three items which represent "
            "field's X, Y, and Z components") from None
    fprod = {"".join("xyz"[c] for c in comb): np.prod(fxyz[list(comb)])
             for comb in itertools.product((0, 1, 2), repeat=rank)}
    fprod["0"] = 1
    if thresh is not None:
        fprod = {k: v for k, v in fprod.items() if abs(v) >= thresh}
    return fprod


class CarTens:
    """Cartesian tensor operator in a block basis of (J, symmetry) subspaces.

    ``kmat[(J1, J2)][(sym1, sym2)][irrep]`` is a CSR dk1 x dk2 matrix,
    ``mmat[(J1, J2)][(sym1, sym2)][irrep][cart]`` a CSR dm1 x dm2 matrix; the
    basis index inside a block is ``im * dk + ik`` (field.py:58-170).
    Leaf tensors own ``kmat``/``mmat``; sums (``add_cartens``) keep the list
    of their leaf terms with the field products each had when summed.
    """

    def __init__(self, filename=None, name=None, **kwargs):
        self._uid = next(_uid)
        if filename is not None:
            raise NotImplementedError(
                "reading CarTens HDF5 files needs h5py, which this package does not use; "
                "build the tensor with richmol_amd.rot or richmol_amd.trove instead")

    # ------------------------------------------------------------------ state
    def _touch(self):
        """New structure id (after copies and in-place value changes)."""
        self._uid = next(_uid)
        self.__dict__.pop("_mfmat_view", None)

    @property
    def is_sum(self):
        return "_terms" in self.__dict__

    def terms(self):
        """[(leaf, fprod dict or None, thresh)] making up this operator."""
        if self.is_sum:
            return list(self._terms)
        return [(self, self.__dict__.get("_fprod"), self.__dict__.get("_fthresh"))]

    def __deepcopy__(self, memo):
        # a deep copy has the same K/M content, so it keeps the structure id
        cls = self.__class__
        new = cls.__new__(cls)
        memo[id(self)] = new
        for k, v in self.__dict__.items():
            if k == "_mfmat_view":
                continue
            setattr(new, k, copy.deepcopy(v, memo))
        return new

    def __getattr__(self, name):
        # merged K view of a sum, built on first access (field.py:1038-1054)
        if name == "kmat" and "_terms" in self.__dict__:
            km = _nested()
            for i, (leaf, _, _) in enumerate(self.__dict__["_terms"]):
                sfx = "_" + str(i + 1) if len(self.__dict__["_terms"]) > 2 else ("_1", "_2")[i]
                for Jp, blk in leaf.kmat.items():
                    for sp, d in blk.items():
                        km[Jp][sp].update({str(w) + sfx: v for w, v in d.items()})
            self.__dict__["kmat"] = km
            return km
        raise AttributeError(f"'{type(self).__name__}' object has no attribute '{name}'")

    # ------------------------------------------------------ field.py:1073-1142
    def field(self, field, thresh=None):
        """Multiply the tensor with a field (X, Y, Z); in place, returns None.

        Only the field products are computed here; the contraction with the
        M tensors runs on the device when the operator is applied.
        """
        if self.is_sum:
            raise AttributeError("a sum of tensors has no M tensor to multiply with a field "
                                 "(field.py:1012-1015 deletes it)")
        self._fprod = field_products(self.rank, field, thresh)
        self._fthresh = thresh
        self.__dict__.pop("_mfmat_view", None)

    @property
    def mfmat(self):
        """Host view of the field-contracted M tensor (field.py:1110-1142)."""
        view = self.__dict__.get("_mfmat_view")
        if view is not None:
            return view
        if self.is_sum:
            view = self._sum_mfmat_view()
        else:
            if "_fprod" not in self.__dict__:
                raise AttributeError("'CarTens' object has no attribute 'mfmat'")
            view = _contract(self.mmat, self._fprod, self._fthresh)
        self.__dict__["_mfmat_view"] = view
        return view

    def _sum_mfmat_view(self):
        mf = _nested()
        terms = self.__dict__["_terms"]
        for i, (leaf, fprod, thr) in enumerate(terms):
            sfx = "_" + str(i + 1) if len(terms) > 2 else ("_1", "_2")[i]
            view = _contract(leaf.mmat, fprod, thr) if fprod is not None else {}
            for Jp, blk in view.items():
                for sp, d in blk.items():
                    mf[Jp][sp].update({str(w) + sfx: v for w, v in d.items()})
        return mf

    # ------------------------------------------------------ field.py:1145-1245
    def vec(self, vec, matvec_lib="hip"):
        """H |vec> for a dict {J: {sym: block}} of numpy blocks, on the GPU."""
        from collections.abc import Mapping
        if not isinstance(vec, Mapping):
            raise TypeError(f"bad argument type for `vec`: '{type(vec)}'") from None
        if matvec_lib != "hip":
            raise ValueError(f"bad argument for `matvec_lib`: '{matvec_lib}' (this package runs 'hip')")
        from . import plan as _plan
        x = flatten(self, vec, side=2)
        y = _plan.apply(self, x[None, :])[0]
        return unflatten(self, y, side=1)

    # ------------------------------------------------ field.py:449-569, 659-692
    def tomat(self, form="block", repres="csr_matrix", thresh=None, cart=None):
        assert form in ("block", "full"), f"`form` unknown: '{form}' (use 'block', 'full')"
        mat = _nested()
        if cart is None:
            mf = self.mfmat
            for Jp in set(mf.keys()) & set(self.kmat.keys()):
                for sp in set(mf[Jp].keys()) & set(self.kmat[Jp].keys()):
                    m, k = mf[Jp][sp], self.kmat[Jp][sp]
                    me = sum(kron(m[w], k[w]) for w in set(m.keys()) & set(k.keys()))
                    if not np.isscalar(me):
                        mat[Jp][sp] = me
        else:
            if cart not in self.cart:
                raise ValueError(f"specified Cartesian component '{cart}' is not contained in "
                                 f"tensor components '{self.cart}'") from None
            for Jp in set(self.mmat.keys()) & set(self.kmat.keys()):
                for sp in set(self.mmat[Jp].keys()) & set(self.kmat[Jp].keys()):
                    m = {w: v[cart] for w, v in self.mmat[Jp][sp].items() if cart in v}
                    k = self.kmat[Jp][sp]
                    me = sum(kron(m[w], k[w]) for w in set(m.keys()) & set(k.keys()))
                    if not np.isscalar(me):
                        mat[Jp][sp] = me
        if thresh is not None and thresh > 0:
            for Jp in mat:
                for sp in mat[Jp]:
                    c = mat[Jp][sp].tocoo()
                    keep = np.abs(c.data) > thresh
                    mat[Jp][sp] = csr_matrix((c.data[keep], (c.row[keep], c.col[keep])), shape=c.shape)
        if form == "full":
            return self.full_form(mat, repres)
        for Jp in mat:
            for sp in mat[Jp]:
                mat[Jp][sp] = (mat[Jp][sp].toarray() if repres == "dense"
                               else getattr(scipy.sparse, repres)(mat[Jp][sp]))
        return mat

    def full_form(self, mat, repres="csr_matrix", thresh=None):
        rows = []
        for J1 in self.Jlist1:
            for s1 in self.symlist1[J1]:
                row = []
                for J2 in self.Jlist2:
                    for s2 in self.symlist2[J2]:
                        if (J1, J2) in mat.keys() and (s1, s2) in mat[(J1, J2)].keys():
                            row.append(mat[(J1, J2)][(s1, s2)])
                        else:
                            row.append(csr_matrix((self.dim1[J1][s1], self.dim2[J2][s2])))
                rows.append(row)
        res = scipy.sparse.bmat(rows)
        return res.toarray() if repres == "dense" else getattr(scipy.sparse, repres)(res)

    # ------------------------------------------------------ field.py:932-948
    def mul(self, arg):
        """In-place multiplication of the K tensor with a scalar."""
        scalar = (int, float, complex, np.integer, np.floating, np.complexfloating)
        if not isinstance(arg, scalar):
            raise TypeError(f"bad argument type for `arg` : '{type(arg)}'") from None
        if self.is_sum:
            self._terms = [(leaf * arg, fp, thr) for leaf, fp, thr in self._terms]
            self.__dict__.pop("kmat", None)
        else:
            # new dict trees: the old ones may be shared with shallow copies
            self.kmat = {Jp: {sp: {w: v * arg for w, v in d.items()} for sp, d in blk.items()}
                         for Jp, blk in self.kmat.items()}
        self._touch()

    # ------------------------------------------------------ field.py:951-1070
    def add_cartens(self, arg):
        if not isinstance(arg, CarTens):
            raise TypeError(f"bad argument type for `arg`: '{type(arg)}'") from None
        for name in ("Jlist1", "Jlist2", "symlist1", "symlist2", "dim1", "dim2",
                     "dim_k1", "dim_k2", "dim_m1", "dim_m2"):
            if getattr(self, name) != getattr(arg, name):
                raise ValueError(f"tensors defined with respect to different basis sets ({name})")
        for t in (self, arg):
            if not t.is_sum and _first_cart(t) == "0":
                t.field([0, 0, 1])
        res = CarTens()
        keep = {k: v for k, v in self.__dict__.items()
                if k not in ("os", "rank", "cart", "mmat", "kmat", "_fprod", "_fthresh",
                             "_mfmat_view", "_terms", "_uid")}
        res.__dict__.update(keep)
        res._terms = []
        for t in (self, arg):
            for leaf, fprod, thr in t.terms():
                if fprod is None:
                    raise AttributeError(
                        "tensor must be multiplied with a field before it is summed "
                        "(field.py:1025-1030 reads its mfmat)")
                res._terms.append((leaf, fprod, thr))
        # kmat / mfmat of the sum: merged host views with renamed irreps, built
        # on first access (field.py:1017-1068)
        return res

    # ---------------------------------------------------- field.py:1248-1301
    def __mul__(self, arg):
        # The reference deep-copies (field.py:1255-1262).  Here the copy shares
        # the K/M trees, which no method mutates in place: mul() builds new
        # trees and field() only replaces the field products.
        scalar = (int, float, complex, np.integer, np.floating, np.complexfloating)
        if isinstance(arg, scalar):
            res = copy.copy(self)
            res.__dict__.pop("_mfmat_view", None)
            res.mul(arg)
        elif isinstance(arg, (np.ndarray, list, tuple)):
            res = copy.copy(self)
            res.__dict__.pop("_mfmat_view", None)
            res.field(arg)
        else:
            raise TypeError(f"unsupported operand type(s) for '*': "
                            f"'{self.__class__.__name__}' and '{type(arg)}'") from None
        return res

    def __add__(self, arg):
        if isinstance(arg, CarTens):
            return self.add_cartens(arg)
        raise TypeError(f"unsupported operand type(s) for '+': "
                        f"'{self.__class__.__name__}' and '{type(arg)}'") from None

    def __sub__(self, arg):
        if isinstance(arg, CarTens):
            return self.add_cartens(arg * (-1))
        raise TypeError(f"unsupported operand type(s) for '-': "
                        f"'{self.__class__.__name__}' and '{type(arg)}'") from None

    __rmul__ = __mul__
    __radd__ = __add__


def _first_cart(t):
    cart = getattr(t, "cart", None)
    if cart is None or len(cart) == 0:
        return None
    return cart[0]


def _contract(mmat, fprod, thresh):
    """Host view of MF = sum_cart fprod[cart] * M[cart] (field.py:1110-1142)."""
    out = _nested()
    if len(fprod) == 0:
        return out
    for Jp, mJ in mmat.items():
        for sp, msym in mJ.items():
            res = dict()
            for w, m in msym.items():
                mats = [fprod[c] * m[c].toarray() for c in sorted(set(m.keys()) & set(fprod.keys()))]
                if not mats:
                    continue
                mat = csr_matrix(sum(mats))
                if thresh is not None and thresh > 0:
                    mat.data[np.abs(mat.data) < thresh] = 0
                    mat.eliminate_zeros()
                if mat.nnz > 0:
                    res[w] = mat
            if res:
                out[Jp][sp] = res
    return out


def block_offsets(T, side):
    """{(J, sym): offset} in the full-form order (field.py:678-686)."""
    Jlist = T.Jlist1 if side == 1 else T.Jlist2
    symlist = T.symlist1 if side == 1 else T.symlist2
    dim = T.dim1 if side == 1 else T.dim2
    off, pos = dict(), 0
    for J in Jlist:
        for s in symlist[J]:
            off[(J, s)] = pos
            pos += dim[J][s]
    return off, pos


def flatten(T, vec, side=2):
    off, n = block_offsets(T, side)
    dim = T.dim1 if side == 1 else T.dim2
    x = np.zeros(n, dtype=np.complex128)
    for (J, s), o in off.items():
        if J in vec and s in vec[J]:
            x[o:o + dim[J][s]] = vec[J][s]
    return x


def unflatten(T, x, side=1):
    off, _ = block_offsets(T, side)
    dim = T.dim1 if side == 1 else T.dim2
    out = dict()
    for (J, s), o in off.items():
        out.setdefault(J, dict())[s] = x[o:o + dim[J][s]]
    return out
