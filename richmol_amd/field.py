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
    f3 = [fxyz[0].item(), fxyz[1].item(), fxyz[2].item()] if fxyz.dtype.kind in "iufc" else list(fxyz)
    fprod = {}
    for comb in itertools.product((0, 1, 2), repeat=rank):
        v = 1
        for c in comb:
            v = v * f3[c]
        fprod["".join("xyz"[c] for c in comb)] = np.float64(v) if isinstance(v, float) else v
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
            from . import h5
            if h5.is_hdf5(filename):
                # read tensor from an HDF5 file (field.py:173-184)
                if name is None:
                    raise ValueError(
                        f"please specify the name of the data group in HDF5 file '{filename}'") from None
                self.read(filename, name=name, **kwargs)

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

    # ------------------------------------------------------ field.py:187-401
    def filter(self, thresh=None, bra=lambda **kw: True, ket=lambda **kw: True):
        """State selection: keeps the bra (ket) states for which ``bra(J=,
        sym=, k=, enr=)`` and ``bra(J=, sym=, m=)`` hold, drops (J, symmetry)
        blocks left empty, and cuts the K and M tensors to what is kept (with
        |element| < ``thresh`` removed).  In place; new K/M trees, so copies
        that share the old ones are not changed."""
        if self.is_sum:
            raise AttributeError("a sum of tensors has no M tensor to filter "
                                 "(field.py:1012-1015 deletes it)")
        for side, flt in ((1, bra), (2, ket)):
            Jl, sl = getattr(self, f"Jlist{side}"), getattr(self, f"symlist{side}")
            qk, qm = getattr(self, f"quanta_k{side}"), getattr(self, f"quanta_m{side}")
            ik = {J: {s: [i for i, (q, e) in enumerate(qk[J][s]) if flt(J=J, sym=s, k=q, enr=e)]
                      for s in sl[J]} for J in Jl}
            im = {J: {s: [i for i, q in enumerate(qm[J][s]) if flt(J=J, sym=s, m=q)]
                      for s in sl[J]} for J in Jl}
            qk = {J: {s: [qk[J][s][i] for i in ik[J][s]] for s in sl[J]} for J in Jl}
            qm = {J: {s: [qm[J][s][i] for i in im[J][s]] for s in sl[J]} for J in Jl}
            # a (J, symmetry) block stays when it keeps k and m states
            keep = {J: [s for s in sl[J] if ik[J][s] and im[J][s]] for J in Jl}
            Jl = sorted(J for J in Jl if keep[J])
            sl = {J: sorted(keep[J]) for J in Jl}
            sub = lambda d: {J: {s: d[J][s] for s in sl[J]} for J in Jl}   # noqa: E731
            ik, im, qk, qm = sub(ik), sub(im), sub(qk), sub(qm)
            setattr(self, f"Jlist{side}", Jl)
            setattr(self, f"symlist{side}", sl)
            setattr(self, f"ind_k{side}", ik)
            setattr(self, f"ind_m{side}", im)
            setattr(self, f"quanta_k{side}", qk)
            setattr(self, f"quanta_m{side}", qm)
            setattr(self, f"dim_k{side}", {J: {s: len(ik[J][s]) for s in sl[J]} for J in Jl})
            setattr(self, f"dim_m{side}", {J: {s: len(im[J][s]) for s in sl[J]} for J in Jl})
            setattr(self, f"dim{side}", {J: {s: len(ik[J][s]) * len(im[J][s]) for s in sl[J]} for J in Jl})

        def cut(mat, rows, cols):
            m = mat.tocsr()[rows, :].tocsc()[:, cols].tocsr()
            if thresh is not None:
                m.data[np.abs(m.data) < thresh] = 0
                m.eliminate_zeros()
            return m

        kmat, mmat = {}, {}
        for Jp in set(self.mmat.keys()) & set(self.kmat.keys()):
            J1, J2 = Jp
            if J1 not in self.Jlist1 or J2 not in self.Jlist2:
                continue
            for sp in set(self.mmat[Jp].keys()) & set(self.kmat[Jp].keys()):
                s1, s2 = sp
                if s1 not in self.symlist1[J1] or s2 not in self.symlist2[J2]:
                    continue
                ik1, ik2 = self.ind_k1[J1][s1], self.ind_k2[J2][s2]
                im1, im2 = self.ind_m1[J1][s1], self.ind_m2[J2][s2]
                kd, md = {}, {}
                for w in set(self.kmat[Jp][sp].keys()) & set(self.mmat[Jp][sp].keys()):
                    k = cut(self.kmat[Jp][sp][w], ik1, ik2)
                    m = {c: cut(v, im1, im2) for c, v in self.mmat[Jp][sp][w].items()}
                    m = {c: v for c, v in m.items() if v.nnz > 0}
                    if k.nnz > 0 and m:
                        kd[w], md[w] = k, m
                if kd:
                    kmat.setdefault(Jp, {})[sp] = kd
                    mmat.setdefault(Jp, {})[sp] = md
        self.kmat, self.mmat = kmat, mmat
        self._touch()

    # ---------------------------------------------------- field.py:1304-1848
    def class_name(self):
        """Type tag of the tensor's HDF5 group: the reference's class path
        (``richmol.field.CarTens`` for this class), so that each package reads
        the other's files (field.py:1845-1849, 1633-1638)."""
        mod = self.__module__.split(".")
        return ".".join(["richmol"] + mod[1:] + [self.__class__.__name__])

    def store(self, filename, name=None, comment=None, replace=False, replace_k=False, replace_m=False,
              store_k=True, store_m=True, thresh=None, store_basis=True):
        """Write the tensor into the HDF5 group ``name`` of ``filename``
        (field.py:1304-1597): the basis attributes (natively, or as JSON
        under ``<name>__json``), and per coupled (J1, J2) and (sym1, sym2)
        the K and M tensors as concatenated CSR arrays with their split
        sizes.  Elements below ``thresh`` are not written (the tensor itself
        is not changed).  ``store_basis`` has no effect: this package's
        tensors carry no symmetric-top basis group."""
        from . import h5
        if self.is_sum:
            raise AttributeError("a sum of tensors has no M tensor to store "
                                 "(field.py:1012-1015 deletes it)")
        if name is None:
            name = _retrieve_name(self)
        with h5.File(filename, "a") as fl:
            if name in fl and replace:
                del fl[name]
            group = fl.require_group(name)
            group.attrs["__class_name__"] = self.class_name()
            doc = group.attrs["__doc__"] if "__doc__" in group.attrs else "Cartesian tensor operator"
            import datetime
            doc += ", store date: " + datetime.datetime.now().strftime("%Y-%m-%d %H:%M:%S")
            if comment is not None:
                doc += ", comment: " + " ".join(comment.split())
            group.attrs["__doc__"] = doc
            exclude = {"mmat", "kmat", "molecule", "basis", "eigvec", "symtop_basis"}
            exclude |= set(getattr(self, "store_exclude", ()))
            for attr, val in self.__dict__.items():
                if attr in exclude or attr.startswith("_"):
                    continue
                try:
                    group.attrs[attr] = val
                except TypeError:
                    group.attrs[attr + "__json"] = h5.json_dumps(val)

            def drop(m):
                m = m.tocsr(copy=True)
                if thresh is not None:
                    m.data[np.abs(m.data) < thresh] = 0
                    m.eliminate_zeros()
                return m

            def put(gsym, what, mats, keys, replace_):
                if replace_:
                    for key in ("_data", "_indices", "_indptr"):
                        if what + key in gsym:
                            del gsym[what + key]
                if any(what + key in gsym for key in ("_data", "_indices", "_indptr")):
                    raise RuntimeError(
                        f"found existing {what[0].upper()}-tensor dataset for `(J1, J2)` = '{(J1, J2)}' and "
                        f"`(sym1, sym2)` = '{(s1, s2)}'  in file '{filename}', use `replace_{what[0]}` = True "
                        f"to replace {what[0].upper()}-tensor datasets") from None
                gsym.create_dataset(what + "_data", np.concatenate([m.data for m in mats]).astype(np.complex128))
                gsym.create_dataset(what + "_indices", np.concatenate([m.indices for m in mats]).astype(np.int32))
                gsym.create_dataset(what + "_indptr", np.concatenate([m.indptr for m in mats]).astype(np.int32))
                gsym.attrs[what + "_nnz"] = np.array([len(m.data) for m in mats], dtype=np.int64)
                gsym.attrs[what + "_nind"] = np.array([len(m.indices) for m in mats], dtype=np.int64)
                gsym.attrs[what + "_nptr"] = np.array([len(m.indptr) for m in mats], dtype=np.int64)
                gsym.attrs[what + "_shape"] = np.array([m.shape for m in mats], dtype=np.int64)
                gsym.attrs[what + ("_irreps" if what == "kmat" else "_irrep_cart")] = keys

            for (J1, J2) in sorted(set(self.mmat.keys()) & set(self.kmat.keys())):
                for (s1, s2) in sorted(set(self.mmat[(J1, J2)].keys()) & set(self.kmat[(J1, J2)].keys())):
                    gsym = None
                    if store_k:
                        kd = {w: drop(m) for w, m in self.kmat[(J1, J2)][(s1, s2)].items()}
                        kd = {w: m for w, m in kd.items() if m.nnz > 0}
                        if kd:
                            gsym = group.require_group(J_group_key(J1, J2)).require_group(sym_group_key(s1, s2))
                            put(gsym, "kmat", list(kd.values()), np.array(list(kd.keys()), dtype=np.int64),
                                replace_k)
                    if store_m:
                        md = {(w, c): drop(m) for w, mw in self.mmat[(J1, J2)][(s1, s2)].items()
                              for c, m in mw.items()}
                        md = {key: m for key, m in md.items() if m.nnz > 0}
                        if md:
                            gsym = gsym or group.require_group(J_group_key(J1, J2)).require_group(
                                sym_group_key(s1, s2))
                            put(gsym, "mmat", list(md.values()),
                                np.array([(str(w), c) for w, c in md.keys()], dtype="S"), replace_m)

    def read(self, filename, name=None, thresh=None, **kwargs):
        """Read the tensor of the HDF5 group ``name`` (field.py:1600-1842):
        attributes first, then the state filters ``bra``/``ket`` of kwargs,
        then the K and M tensors of the kept states (elements below
        ``thresh`` dropped).  J quanta that are whole numbers come back as
        ints; the symmetric-top basis, eigenvector and rotational-density
        groups the reference may also hold are not read (this package builds
        densities from its own basis)."""
        import re
        from . import h5
        if name is None:
            name = _retrieve_name(self)
        with h5.File(filename, "r") as fl:
            if name not in fl:
                raise KeyError(f"file '{filename}' has no dataset with the name '{name}'") from None
            group = fl[name]
            cls = group.attrs["__class_name__"]
            if cls != self.class_name():
                raise TypeError(f"dataset with the name '{name}' in file '{filename}' "
                                f"has different type: '{cls}'") from None
            attrs = {}
            for key, val in group.attrs.items():
                if key.find("__json") == -1:
                    attrs[key] = val
                else:
                    attrs[key.replace("__json", "")] = h5.json_loads(val)
            self.__dict__.update(_normalise_attrs(attrs))
            self.kmat, self.mmat = {}, {}
            self.filter(**kwargs)
            Jre = re.compile(r"^J:(\d+\.\d+),(\d+\.\d+)$")
            sre = re.compile(r"^sym:(\w+),(\w+)$")
            for key in group.keys():
                mj = Jre.match(key)
                if not mj:
                    continue
                J1, J2 = (_jnum(round(float(x), 1)) for x in mj.groups())
                if J1 not in self.Jlist1 or J2 not in self.Jlist2:
                    continue
                gj = group[key]
                for key2 in gj.keys():
                    ms = sre.match(key2)
                    if not ms:
                        continue
                    s1, s2 = ms.groups()
                    if s1 not in self.symlist1[J1] or s2 not in self.symlist2[J2]:
                        continue
                    gs = gj[key2]
                    ik1, ik2 = self.ind_k1[J1][s1], self.ind_k2[J2][s2]
                    im1, im2 = self.ind_m1[J1][s1], self.ind_m2[J2][s2]
                    kd = _read_csr(gs, "kmat")
                    if kd is not None:
                        mats, keys = kd
                        kd = {int(w): _sel(m, ik1, ik2, thresh) for w, m in zip(keys, mats)}
                        kd = {w: m for w, m in kd.items() if m.nnz > 0}
                        if kd:
                            self.kmat.setdefault((J1, J2), {})[(s1, s2)] = kd
                    md = _read_csr(gs, "mmat")
                    if md is not None:
                        mats, keys = md
                        out = {}
                        for (w, c), m in zip(keys, mats):
                            m = _sel(m, im1, im2, thresh)
                            if m.nnz > 0:
                                out.setdefault(int(w), {})[c] = m
                        if out:
                            self.mmat.setdefault((J1, J2), {})[(s1, s2)] = out
        self.filter(thresh=thresh, **kwargs)

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


def filter(obj, bra, ket, thresh=None):
    """Copy of a tensor with the state selection applied (field.py:1859-1883)."""
    if not isinstance(obj, CarTens):
        raise TypeError(f"bad argument type for 'obj': '{type(obj)}'") from None
    res = copy.deepcopy(obj)
    res.filter(bra=bra, ket=ket, thresh=thresh)
    return res


def J_group_key(J1, J2):
    """HDF5 group of the (J1, J2) block (field.py:1886-1890)."""
    return "J:" + str(round(float(J1), 1)) + "," + str(round(float(J2), 1))


def sym_group_key(sym1, sym2):
    """HDF5 group of the (sym1, sym2) block (field.py:1893-1897)."""
    return "sym:" + str(sym1) + "," + str(sym2)


def _retrieve_name(var):
    """Name a variable bound to ``var`` has in the calling frames, outermost
    first (the reference's default group name)."""
    import inspect
    for fi in reversed(inspect.stack()):
        for n, v in fi.frame.f_locals.items():
            if v is var and not n.startswith("_") and n != "self":
                return n
    raise ValueError("no name for the tensor: pass `name`")


def _jnum(J):
    """J quantum as an int when it is a whole number."""
    f = float(J)
    return int(f) if f.is_integer() else round(f, 1)


_JKEYED = ("symlist1", "symlist2", "dim1", "dim2", "dim_k1", "dim_k2", "dim_m1", "dim_m2",
           "ind_k1", "ind_k2", "ind_m1", "ind_m2", "quanta_k1", "quanta_k2", "quanta_m1", "quanta_m2")


def _normalise_attrs(attrs):
    """Attributes as read from a file in the types the tensor code uses:
    J quanta and J keys as ints when whole, lists for arrays of labels,
    (assignment, energy) pairs as tuples."""
    out = dict(attrs)
    for k in ("Jlist1", "Jlist2"):
        if k in out:
            out[k] = [_jnum(J) for J in np.atleast_1d(out[k])]
    for k in _JKEYED:
        if k in out and isinstance(out[k], dict):
            out[k] = {_jnum(J): dict(v) if isinstance(v, dict) else v for J, v in out[k].items()}
    for k in ("quanta_k1", "quanta_k2"):
        if k in out:
            out[k] = {J: {s: [tuple(q) for q in v] for s, v in d.items()} for J, d in out[k].items()}
    if "cart" in out:
        out["cart"] = [str(c) for c in np.atleast_1d(out["cart"])]
    if "os" in out:
        out["os"] = [tuple(int(x) for x in r) for r in np.atleast_2d(out["os"])]
    if "rank" in out:
        out["rank"] = int(out["rank"])
    return out


def _read_csr(gs, what):
    """CSR matrices and their keys of one block (field.py:1745-1835)."""
    if what + "_nnz" not in gs.attrs or what + "_data" not in gs:
        return None
    nnz, nind, nptr = (np.asarray(gs.attrs[what + k]).ravel() for k in ("_nnz", "_nind", "_nptr"))
    shape = np.asarray(gs.attrs[what + "_shape"]).reshape(-1, 2)
    data = np.split(np.asarray(gs[what + "_data"].read()), np.cumsum(nnz))[:-1]
    ind = np.split(np.asarray(gs[what + "_indices"].read()), np.cumsum(nind))[:-1]
    ptr = np.split(np.asarray(gs[what + "_indptr"].read()), np.cumsum(nptr))[:-1]
    mats = [csr_matrix((d.astype(np.complex128), i, p), shape=tuple(int(x) for x in sh))
            for d, i, p, sh in zip(data, ind, ptr, shape)]
    if what == "kmat":
        keys = [int(w) for w in np.atleast_1d(gs.attrs["kmat_irreps"])]
    else:
        raw = np.asarray(gs.attrs["mmat_irrep_cart"]).reshape(-1, 2)
        dec = lambda x: x.decode("utf-8") if isinstance(x, bytes) else str(x)   # noqa: E731
        keys = [(int(dec(w)), dec(c)) for w, c in raw]
    return mats, keys


def _sel(m, rows, cols, thresh):
    m = m[rows, :].tocsc()[:, cols].tocsr()
    if thresh is not None:
        m.data[np.abs(m.data) < thresh] = 0
        m.eliminate_zeros()
    return m


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
