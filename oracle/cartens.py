"""CPU restatement of the reference's Cartesian tensor container (oracle).

Follows richmol/field.py:
  tomat        449-569      full_form  659-692
  mul          932-948      add_cartens 951-1070
  field        1073-1142    vec        1145-1245   (matvec_lib='scipy')
  __mul__/__add__/__sub__   1248-1301

Data contract (field.py:58-170): kmat[(J1,J2)][(s1,s2)][irrep] is a CSR
dk1 x dk2 matrix, mmat[(J1,J2)][(s1,s2)][irrep][cart] a CSR dm1 x dm2 matrix,
basis index inside a (J, sym) block is im*dk + ik (m-major).
"""
import copy
import itertools
from collections import defaultdict

import numpy as np
import scipy.sparse
from scipy.sparse import csr_matrix, kron


def _nested():
    return defaultdict(_nested)


_ATTRS = ("Jlist1", "Jlist2", "symlist1", "symlist2", "dim1", "dim2",
          "dim_m1", "dim_m2", "dim_k1", "dim_k2", "kmat", "mmat", "mfmat",
          "cart", "rank", "os")


class OracleCarTens:
    """Block tensor operator with the reference's attribute layout."""

    @classmethod
    def from_attrs(cls, src):
        """Deep copy of the data-contract attributes of any tensor object."""
        obj = cls()
        for name in _ATTRS:
            if hasattr(src, name):
                val = getattr(src, name)
                if name in ("kmat", "mmat", "mfmat"):
                    val = _copy_blocks(val)
                else:
                    val = copy.deepcopy(val)
                setattr(obj, name, val)
        return obj

    # -- field.py:1073-1142 ------------------------------------------------
    def field(self, field, thresh=None):
        try:
            fx, fy, fz = field[:3]
            fxyz = np.array([fx, fy, fz])
        except (TypeError, IndexError):
            raise IndexError("field variable must be an iterable with three items") from None
        field_prod = {
            "".join("xyz"[c] for c in comb): np.prod(fxyz[list(comb)])
            for comb in itertools.product((0, 1, 2), repeat=self.rank)
        }
        field_prod["0"] = 1
        if thresh is not None:
            field_prod = {k: v for k, v in field_prod.items() if abs(v) >= thresh}
        self.mfmat = _nested()
        if len(field_prod) == 0:
            return
        for (J1, J2), mmat_J in self.mmat.items():
            for (s1, s2), mmat_sym in mmat_J.items():
                res = dict()
                for irrep, mmat in mmat_sym.items():
                    mats = [field_prod[c] * mmat[c].toarray()
                            for c in sorted(set(mmat.keys()) & set(field_prod.keys()))]
                    if len(mats) == 0:
                        continue
                    mat = csr_matrix(mats[0] if len(mats) == 1 else sum(mats))
                    if thresh is not None and thresh > 0:
                        mat.data[abs(mat.data) < thresh] = 0
                        mat.eliminate_zeros()
                    if mat.nnz > 0:
                        res[irrep] = mat
                if len(res) > 0:
                    self.mfmat[(J1, J2)][(s1, s2)] = res

    # -- field.py:1145-1245 ------------------------------------------------
    def vec(self, vec):
        if not hasattr(self, "mfmat"):
            raise AttributeError("you need to multiply tensor with field before applying it to a vector")
        out = dict()
        for (J1, J2) in sorted(set(self.mfmat.keys()) & set(self.kmat.keys())):
            mf_J = self.mfmat[(J1, J2)]
            k_J = self.kmat[(J1, J2)]
            out.setdefault(J1, dict())
            for (s1, s2) in sorted(set(mf_J.keys()) & set(k_J.keys())):
                dm2 = self.dim_m2[J2][s2]
                dk2 = self.dim_k2[J2][s2]
                dim = self.dim1[J1][s1]
                try:
                    xT = np.transpose(vec[J2][s2].reshape(dm2, dk2))
                except KeyError:
                    continue
                res = []
                for irrep in sorted(set(mf_J[(s1, s2)].keys()) & set(k_J[(s1, s2)].keys()),
                                    key=str):
                    tmat = k_J[(s1, s2)][irrep].dot(xT)                       # dk1 x dm2
                    res.append(mf_J[(s1, s2)][irrep].dot(np.transpose(tmat)).reshape(dim))
                if s1 in out[J1]:
                    out[J1][s1] = out[J1][s1] + sum(res)
                else:
                    out[J1][s1] = sum(res)
        return out

    # -- field.py:449-569, 659-692 -----------------------------------------
    def tomat(self, form="block", repres="csr_matrix", thresh=None, cart=None):
        mat = _nested()
        if cart is None:
            src = self.mfmat
            for Jp in set(src.keys()) & set(self.kmat.keys()):
                for sp in set(src[Jp].keys()) & set(self.kmat[Jp].keys()):
                    mm, kk = src[Jp][sp], self.kmat[Jp][sp]
                    me = sum(kron(mm[w], kk[w]) for w in set(mm.keys()) & set(kk.keys()))
                    if not np.isscalar(me):
                        mat[Jp][sp] = me
        else:
            if cart not in self.cart:
                raise ValueError(f"Cartesian component '{cart}' not in {self.cart}")
            for Jp in set(self.mmat.keys()) & set(self.kmat.keys()):
                for sp in set(self.mmat[Jp].keys()) & set(self.kmat[Jp].keys()):
                    mm = {w: v[cart] for w, v in self.mmat[Jp][sp].items() if cart in v}
                    kk = self.kmat[Jp][sp]
                    me = sum(kron(mm[w], kk[w]) for w in set(mm.keys()) & set(kk.keys()))
                    if not np.isscalar(me):
                        mat[Jp][sp] = me
        if thresh is not None and thresh > 0:
            for Jp in mat:
                for sp in mat[Jp]:
                    m = mat[Jp][sp].tocoo()
                    keep = abs(m.data) > thresh
                    mat[Jp][sp] = csr_matrix((m.data[keep], (m.row[keep], m.col[keep])), shape=m.shape)
        if form == "full":
            return self.full_form(mat, repres)
        return mat

    def full_form(self, mat, repres="csr_matrix"):
        rows = []
        for J1 in self.Jlist1:
            for s1 in self.symlist1[J1]:
                row = []
                for J2 in self.Jlist2:
                    for s2 in self.symlist2[J2]:
                        if (J1, J2) in mat and (s1, s2) in mat[(J1, J2)]:
                            row.append(mat[(J1, J2)][(s1, s2)])
                        else:
                            row.append(csr_matrix((self.dim1[J1][s1], self.dim2[J2][s2])))
                rows.append(row)
        res = scipy.sparse.bmat(rows)
        if repres == "dense":
            return res.toarray()
        return getattr(scipy.sparse, repres)(res)

    # -- field.py:932-948 --------------------------------------------------
    def mul(self, arg):
        for Jp in self.kmat:
            for sp in self.kmat[Jp]:
                self.kmat[Jp][sp] = {k: v * arg for k, v in self.kmat[Jp][sp].items()}

    # -- field.py:951-1070 -------------------------------------------------
    def add_cartens(self, arg):
        for t in (self, arg):
            if getattr(t, "cart", [None])[0] == "0":
                t.field([0, 0, 1])
        res = OracleCarTens()
        res.__dict__.update(self.__dict__)
        for a in ("os", "rank", "cart", "mmat"):
            if hasattr(res, a):
                delattr(res, a)
        ir1 = {w for J in self.kmat for s in self.kmat[J] for w in self.kmat[J][s]} | \
              {w for J in self.mfmat for s in self.mfmat[J] for w in self.mfmat[J][s]}
        ir2 = {w for J in arg.kmat for s in arg.kmat[J] for w in arg.kmat[J][s]} | \
              {w for J in arg.mfmat for s in arg.mfmat[J] for w in arg.mfmat[J][s]}
        ren1 = {w: str(w) + "_1" for w in ir1}
        ren2 = {w: str(w) + "_2" for w in ir2}
        res.kmat = _nested()
        res.mfmat = _nested()
        for src, ren, attr in ((self, ren1, "kmat"), (self, ren1, "mfmat"),
                               (arg, ren2, "kmat"), (arg, ren2, "mfmat")):
            for Jp, blk in getattr(src, attr).items():
                for sp, d in blk.items():
                    getattr(res, attr)[Jp][sp].update({ren[w]: v for w, v in d.items()})
        return res

    # -- field.py:1248-1301 ------------------------------------------------
    def __mul__(self, arg):
        if isinstance(arg, (int, float, complex, np.number)):
            res = copy.deepcopy(self)
            res.mul(arg)
        elif isinstance(arg, (np.ndarray, list, tuple)):
            res = copy.deepcopy(self)
            res.field(arg)
        else:
            raise TypeError(f"unsupported operand type for '*': '{type(arg)}'")
        return res

    __rmul__ = __mul__

    def __add__(self, arg):
        return self.add_cartens(arg)

    def __sub__(self, arg):
        return self.add_cartens(arg * (-1))


def _copy_blocks(d):
    out = _nested()
    for Jp, blk in d.items():
        for sp, irr in blk.items():
            out[Jp][sp] = {w: (dict(v) if isinstance(v, dict) else v.copy())
                           for w, v in irr.items()}
    return out


def flatten(H, vec_dict):
    """dict {J: {sym: block}} -> flat vector in Jlist2/symlist2 order (tdse.py:352-360)."""
    return np.concatenate([
        vec_dict[J][s] if (J in vec_dict and s in vec_dict[J])
        else np.zeros(H.dim2[J][s], dtype=np.complex128)
        for J in H.Jlist2 for s in H.symlist2[J]])


def unflatten(H, v):
    """flat vector -> dict in Jlist2/symlist2 order (tdse.py:343-348)."""
    out, ind = dict(), 0
    for J in H.Jlist2:
        out[J] = dict()
        for s in H.symlist2[J]:
            out[J][s] = v[ind: ind + H.dim2[J][s]]
            ind += H.dim2[J][s]
    return out
