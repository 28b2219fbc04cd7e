"""Reader of richmol-format ASCII tensor files (TROVE output, ``.rchm``).

Builds a CarTens from a states file and, optionally, one matrix-element file
per (J1, J2) pair, following the file semantics of richmol/trove.py:
read_states 52-198 (states, Wang-free K basis, rank-0 H0 when no matrix
elements are given) and read_trans 325-666 (M/K tensors, complex flags,
conjugation and transposition rules).  State selection follows
CarTens.filter (richmol/field.py:187-446): ``bra``/``ket`` callables receive
J, sym, m (float) or k (assignment string) and enr, and return True to keep
a state.
"""
import os
import re

import numpy as np
from scipy.sparse import csr_matrix

from .field import CarTens

# (ncart, nirrep) -> irreps (trove.py:357-364); ncart -> rank (trove.py:367)
_IRREPS = {
    (3, 1): [(1, -1), (1, 0), (1, 1)],
    (9, 1): [(2, -2), (2, -1), (2, 0), (2, 1), (2, 2)],
    (9, 2): [(0, 0), (2, -2), (2, -1), (2, 0), (2, 1), (2, 2)],
    (6, 2): [(0, 0), (2, -2), (2, -1), (2, 0), (2, 1), (2, 2)],
    (9, 3): [(0, 0), (1, -1), (1, 0), (1, 1), (2, -2), (2, -1), (2, 0), (2, 1), (2, 2)],
}
_RANKS = {3: 1, 9: 2, 6: 2}


def _keep_all(**kw):
    return True


class CarTensTrove(CarTens):

    def __init__(self, filename, matelem=None, bra=_keep_all, ket=_keep_all, thresh=None):
        super().__init__()
        self._read_states(filename, bra, ket)
        if matelem is None:
            self._h0()
        else:
            self._read_trans(matelem, thresh)

    # ------------------------------------------------------------- states
    def _read_states(self, filename, bra, ket):
        energy, assign, kmap = {}, {}, {}
        with open(filename) as fl:
            for line in fl:
                w = line.split()
                if not w:
                    continue
                try:
                    J = round(float(w[0]), 1)
                    sid = int(w[1])
                    sym = w[2]
                    ndeg = int(w[3])
                    enr = float(w[4])
                    q = " ".join(w[5:])
                except (IndexError, ValueError):
                    raise ValueError(f"error while reading file '{filename}'") from None
                for ideg in range(ndeg):
                    energy.setdefault(J, {}).setdefault(sym, []).append(enr)
                    assign.setdefault(J, {}).setdefault(sym, []).append(q)
                    kmap[(J, sid, ideg + 1)] = (len(energy[J][sym]) - 1, sym)
        if not energy:
            raise Exception(f"zero number of states in file '{filename}'")
        mq = {J: [round(float(m), 1) for m in np.linspace(-J, J, int(2 * J) + 1)] for J in energy}
        self._energy_all, self._kmap = energy, kmap
        self._mq_all = mq
        self._mmap = {(J, m): i for J in mq for i, m in enumerate(mq[J])}
        for side, filt in ((1, bra), (2, ket)):
            ind_k, ind_m = {}, {}
            for J in energy:
                for sym in energy[J]:
                    ik = [i for i, (e, q) in enumerate(zip(energy[J][sym], assign[J][sym]))
                          if filt(J=J, sym=sym, k=q, enr=e)]
                    im = [i for i, m in enumerate(mq[J]) if filt(J=J, sym=sym, m=m)]
                    if ik and im:
                        ind_k.setdefault(J, {})[sym] = ik
                        ind_m.setdefault(J, {})[sym] = im
            Jlist = sorted(ind_k.keys())
            symlist = {J: sorted(ind_k[J].keys()) for J in Jlist}
            setattr(self, f"Jlist{side}", Jlist)
            setattr(self, f"symlist{side}", symlist)
            setattr(self, f"ind_k{side}", ind_k)
            setattr(self, f"ind_m{side}", ind_m)
            setattr(self, f"dim_k{side}", {J: {s: len(ind_k[J][s]) for s in symlist[J]} for J in Jlist})
            setattr(self, f"dim_m{side}", {J: {s: len(ind_m[J][s]) for s in symlist[J]} for J in Jlist})
            setattr(self, f"dim{side}", {J: {s: len(ind_k[J][s]) * len(ind_m[J][s]) for s in symlist[J]}
                                         for J in Jlist})
            setattr(self, f"quanta_m{side}", {J: {s: [mq[J][i] for i in ind_m[J][s]] for s in symlist[J]}
                                              for J in Jlist})
            setattr(self, f"quanta_k{side}", {J: {s: [(assign[J][s][i], energy[J][s][i]) for i in ind_k[J][s]]
                                                  for s in symlist[J]} for J in Jlist})

    def _h0(self):
        """Rank-0 field-free Hamiltonian of the selected states (trove.py:176-194)."""
        self.cart, self.os, self.rank = ["0"], [(0, 0)], 0
        self.kmat, self.mmat = {}, {}
        for J in self.Jlist1:
            if J not in self.Jlist2:
                continue
            for s in self.symlist1[J]:
                if s not in self.symlist2[J]:
                    continue
                e = np.array(self._energy_all[J][s])
                K = np.diag(e)[np.ix_(self.ind_k1[J][s], self.ind_k2[J][s])]
                M = np.eye(len(self._mq_all[J]))[np.ix_(self.ind_m1[J][s], self.ind_m2[J][s])]
                self.kmat.setdefault((J, J), {})[(s, s)] = {0: csr_matrix(K.astype(np.complex128))}
                self.mmat.setdefault((J, J), {})[(s, s)] = {0: {"0": csr_matrix(M.astype(np.complex128))}}

    # ------------------------------------------------------ matrix elements
    def _read_trans(self, template, thresh):
        self.cart, self.kmat, self.mmat = [], {}, {}
        nirrep0 = ncart0 = None
        carts = set()
        for J1_ in self.Jlist1:
            for J2_ in self.Jlist2:
                fname, transp = _find_file(template, J1_, J2_)
                if fname is None:
                    continue
                J1, J2 = (J2_, J1_) if transp else (J1_, J2_)
                head, mdat, kdat = _parse_rchm(fname)
                nirrep, ncart = head
                if nirrep0 is not None and (nirrep0 != nirrep or ncart0 != ncart):
                    raise ValueError(f"'{fname}': irreps/Cartesian components differ from previous files")
                nirrep0, ncart0 = nirrep, ncart
                try:
                    self.os = _IRREPS[(ncart, nirrep)]
                    self.rank = _RANKS[ncart]
                except KeyError:
                    raise ValueError(f"'{fname}': cannot infer tensor irreps from ncart={ncart}, "
                                     f"nirrep={nirrep}") from None
                irreps = sorted({o for o, _ in self.os})
                # M tensor: per (irrep, cart) COO in file orientation
                mrow, mcol, mval = {}, {}, {}
                for cart, m1, m2, vals in mdat:
                    carts.add(cart)
                    i1, i2 = self._mmap[(J1, m1)], self._mmap[(J2, m2)]
                    for ir, v in zip(irreps, vals):
                        if thresh is not None and abs(v) < thresh:
                            continue
                        mrow.setdefault((ir, cart), []).append(i1)
                        mcol.setdefault((ir, cart), []).append(i2)
                        mval.setdefault((ir, cart), []).append(v)
                krow, kcol, kval = {}, {}, {}
                for id1, id2, g1, g2, vals in kdat:
                    s1i, s1 = self._kmap[(J1, id1, g1)]
                    s2i, s2 = self._kmap[(J2, id2, g2)]
                    for ir, v in zip(irreps, vals):
                        if thresh is not None and abs(v) < thresh:
                            continue
                        key = ((s1, s2), ir)
                        krow.setdefault(key, []).append(s1i)
                        kcol.setdefault(key, []).append(s2i)
                        kval.setdefault(key, []).append(v)
                for sp in sorted({k[0] for k in kval}):
                    s1, s2 = sp
                    t1, t2 = (s2, s1) if transp else (s1, s2)
                    try:
                        ik1, ik2 = self.ind_k1[J1_][t1], self.ind_k2[J2_][t2]
                        im1, im2 = self.ind_m1[J1_][t1], self.ind_m2[J2_][t2]
                    except KeyError:
                        continue
                    mshape = (len(self._mq_all[J1_]), len(self._mq_all[J2_]))
                    kshape = (len(self._energy_all[J1_][t1]), len(self._energy_all[J2_][t2]))
                    mm, km = {}, {}
                    for (ir, cart), v in mval.items():
                        if transp:   # file holds <J2_|T|J1_>: transpose, no conjugation (trove.py:604-621)
                            A = csr_matrix((v, (mcol[(ir, cart)], mrow[(ir, cart)])), shape=mshape)
                        else:        # conjugate (trove.py:623-640)
                            A = csr_matrix((np.conj(v), (mrow[(ir, cart)], mcol[(ir, cart)])), shape=mshape)
                        A = A[im1, :].tocsc()[:, im2].tocsr()
                        A.eliminate_zeros()
                        if A.nnz:
                            mm.setdefault(ir, {})[cart] = A
                    for (spk, ir), v in kval.items():
                        if spk != sp:
                            continue
                        if transp:
                            A = csr_matrix((v, (kcol[(spk, ir)], krow[(spk, ir)])), shape=kshape)
                        else:
                            A = csr_matrix((np.conj(v), (krow[(spk, ir)], kcol[(spk, ir)])), shape=kshape)
                        A = A[ik1, :].tocsc()[:, ik2].tocsr()
                        A.eliminate_zeros()
                        if A.nnz and ir in mm:
                            km[ir] = A
                    mm = {ir: v for ir, v in mm.items() if ir in km}
                    if km:
                        self.kmat.setdefault((J1_, J2_), {})[(t1, t2)] = km
                        self.mmat.setdefault((J1_, J2_), {})[(t1, t2)] = mm
        self.cart = sorted(carts)


def _find_file(template, J1, J2):
    """File of the (J1, J2) pair, else of (J2, J1) with transp=True (trove.py:387-414)."""
    for a, b, transp in ((J1, J2, False), (J2, J1, True)):
        name = re.sub(r"\<f1\>", str(round(a, 1)), template)
        name = re.sub(r"\<f2\>", str(round(b, 1)), name)
        name = re.sub(r"\<j1\>", str(int(round(a, 0))), name)
        name = re.sub(r"\<j2\>", str(int(round(b, 0))), name)
        if os.path.exists(name):
            return name, transp
    return None, False


def _parse_rchm(fname):
    """(nirrep, ncart), M records (cart, m1, m2, values), K records
    (id1, id2, ideg1, ideg2, values) of one matrix-element file (trove.py:416-556)."""
    mdat, kdat = [], []
    head = None
    mode = None
    cart = None
    fac_m = fac_k = 1.0
    with open(fname) as fl:
        lines = [ln.rstrip("\n") for ln in fl]
    if not lines or lines[0] != "Start richmol format":
        raise Exception(f"file '{fname}' has bogus header")
    done = False
    for ln in lines[1:]:
        if ln == "End richmol format":
            done = True
            break
        w = ln.split()
        if head is None:
            head = (int(w[1]), int(w[2]))
            continue
        if ln == "M-tensor":
            mode = "m"
            continue
        if "K-tensor" in ln:
            mode = "k"
            # "K-tensor <icmplx>": 0 real, -1 imaginary; without the flag: old format, real
            fac_k = (1j, 1.0)[int(w[1]) + 1] if len(w) > 1 else 1.0
            continue
        if mode == "m" and w[0] == "alpha":
            fac_m = (1j, 1.0)[int(w[2]) + 1]
            cart = w[3].lower()
            continue
        if mode == "m":
            mdat.append((cart, round(float(w[0]), 1), round(float(w[1]), 1),
                         [float(v) * fac_m for v in w[2:]]))
        elif mode == "k":
            kdat.append((int(w[0]), int(w[1]), int(w[2]), int(w[3]), [float(v) * fac_k for v in w[4:]]))
    if not done:
        raise Exception(f"'{fname}' has bogus footer")
    return head, mdat, kdat
