"""Operator builders for rigid rotors: the K/M tensors the propagator consumes.

Produces CarTens objects with the layout richmol's LabTensor produces
(richmol/rot/labtens.py:114-231, tens_proj 411-530, matelem 533-638):

  K_w[bra, ket] = sum_k conj(c_bra,k) c_ket,k' (-1)^|k| sum_s 3j(J2 w J1; k' s -k) UsT[w][s]
  M_w,cart[m1, m2] = sqrt((2J1+1)(2J2+1)) (-1)^|m1| sum_s Ux[cart, (w,s)] 3j(J2 w J1; m2 s -m1)

with (J1, m1, k) the bra and (J2, m2, k') the ket, UsT = Us . tensor the
spherical components of the molecular-frame tensor, Ux the spherical-to-
Cartesian matrix, and c the symmetric-top (Wang) coefficients of the basis.
Linear rotors use the single k = 0 function with c = 1 (J even) or i (J odd),
as richmol/rot/basis.py:729-801 does.  The Cartesian-spherical transformation
table is the one richmol uses (labtens.py:89-112) so operators built here
equal the reference's on the same molecule.
"""
import functools
import math
from fractions import Fraction

import numpy as np
from scipy.sparse import csr_matrix

from .field import CarTens

# Cartesian -> spherical tensor transformation, rows (omega, sigma), columns
# Cartesian components in the order of CART[rank].
_S2 = math.sqrt(2.0)
_US = {
    1: np.array([[_S2 / 2, -_S2 * 1j / 2, 0],
                 [0, 0, 1.0],
                 [-_S2 / 2, -_S2 * 1j / 2, 0]], dtype=np.complex128),
    2: np.array([[-1 / math.sqrt(3), 0, 0, 0, -1 / math.sqrt(3), 0, 0, 0, -1 / math.sqrt(3)],
                 [0, 0, -0.5, 0, 0, 0.5j, 0.5, -0.5j, 0],
                 [0, 1 / _S2 * 1j, 0, -1 / _S2 * 1j, 0, 0, 0, 0, 0],
                 [0, 0, -0.5, 0, 0, -0.5j, 0.5, 0.5j, 0],
                 [0.5, -0.5j, 0, -0.5j, -0.5, 0, 0, 0, 0],
                 [0, 0, 0.5, 0, 0, -0.5j, 0.5, -0.5j, 0],
                 [-1 / math.sqrt(6), 0, 0, 0, -1 / math.sqrt(6), 0, 0, 0, math.sqrt(6) / 3],
                 [0, 0, -0.5, 0, 0, -0.5j, -0.5, -0.5j, 0],
                 [0.5, 0.5j, 0, 0.5j, -0.5, 0, 0, 0, 0]], dtype=np.complex128),
}
_UX = {r: np.linalg.pinv(u) for r, u in _US.items()}
CART = {1: ["x", "y", "z"], 2: ["xx", "xy", "xz", "yx", "yy", "yz", "zx", "zy", "zz"]}
IRREPS = {1: [(1, -1), (1, 0), (1, 1)], 2: [(o, s) for o in range(3) for s in range(-o, o + 1)]}
_EPS = np.finfo(np.complex128).eps


@functools.lru_cache(maxsize=None)
def _fact(n):
    return math.factorial(n)


@functools.lru_cache(maxsize=1 << 20)
def wigner3j(j1, j2, j3, m1, m2, m3):
    """Wigner 3j symbol for integer arguments (Racah formula, exact rational
    arithmetic, one final rounding)."""
    if m1 + m2 + m3 != 0:
        return 0.0
    if j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    kmin = max(0, j2 - j3 - m1, j1 - j3 + m2)
    kmax = min(j1 + j2 - j3, j1 - m1, j2 + m2)
    s = Fraction(0)
    for k in range(kmin, kmax + 1):
        den = (_fact(k) * _fact(j3 - j2 + k + m1) * _fact(j3 - j1 + k - m2) *
               _fact(j1 + j2 - j3 - k) * _fact(j1 - k - m1) * _fact(j2 - k + m2))
        s += Fraction((-1) ** k, den)
    if s == 0:
        return 0.0
    tri = Fraction(_fact(j1 + j2 - j3) * _fact(j1 - j2 + j3) * _fact(-j1 + j2 + j3),
                   _fact(j1 + j2 + j3 + 1))
    pre = tri * (_fact(j1 + m1) * _fact(j1 - m1) * _fact(j2 + m2) * _fact(j2 - m2) *
                 _fact(j3 + m3) * _fact(j3 - m3))
    val = math.sqrt(float(pre * s * s))
    sign = (1 if s > 0 else -1) * (-1) ** ((j1 - j2 - m3) % 2)
    return sign * val


def spherical_components(tens):
    """(rank, {omega: {sigma: UsT}}, {omega: {sigma: Ux column}}) of a
    molecular-frame tensor (labtens.py:237-282 for the Us/Ux pruning)."""
    tens = np.asarray(tens)
    rank = tens.ndim
    if rank not in (1, 2) or any(d != 3 for d in tens.shape):
        raise ValueError(f"input tensor has inappropriate shape: '{tens.shape}'")
    if np.any(np.isnan(tens)):
        raise ValueError("some of elements of input tensor are NaN")
    flat = np.array([tens.item(tuple("xyz".index(c) for c in cart)) for cart in CART[rank]])
    us = _US[rank].copy()
    ux = _UX[rank].copy()
    us[np.abs(us) < _EPS] = 0
    ux[np.abs(ux) < _EPS] = 0
    ust = us @ flat
    comps, ucols = {}, {}
    for i, (o, s) in enumerate(IRREPS[rank]):
        if abs(ust[i]) > _EPS:
            comps.setdefault(o, {})[s] = ust[i]
        ucols.setdefault(o, {})[s] = ux[:, i]
    return rank, comps, ucols


class RotorBasis:
    """Block basis of a rigid rotor: per J one symmetry block (label 'A' for
    linear rotors) with m = -J..J (optionally filtered) and K functions given
    by coefficient columns over primitive k.

    ``kfun[J][sym]`` is (kprim list, coefficient matrix nprim x nstate),
    ``energy[J][sym]`` the field-free energies of the states.
    """

    def __init__(self, jlist, mlist, kfun, energy):
        self.Jlist = list(jlist)
        self.mlist = mlist
        self.kfun = kfun
        self.energy = energy
        self.symlist = {J: list(kfun[J].keys()) for J in self.Jlist}

    def attrs(self, T):
        T.Jlist1 = list(self.Jlist); T.Jlist2 = list(self.Jlist)
        T.symlist1 = {J: list(s) for J, s in self.symlist.items()}
        T.symlist2 = {J: list(s) for J, s in self.symlist.items()}
        T.dim_m1 = {J: {s: len(self.mlist[J]) for s in self.symlist[J]} for J in self.Jlist}
        T.dim_k1 = {J: {s: self.kfun[J][s][1].shape[1] for s in self.symlist[J]} for J in self.Jlist}
        T.dim1 = {J: {s: T.dim_m1[J][s] * T.dim_k1[J][s] for s in self.symlist[J]} for J in self.Jlist}
        T.dim_m2 = {J: dict(v) for J, v in T.dim_m1.items()}
        T.dim_k2 = {J: dict(v) for J, v in T.dim_k1.items()}
        T.dim2 = {J: dict(v) for J, v in T.dim1.items()}
        T.quanta_m1 = {J: {s: list(self.mlist[J]) for s in self.symlist[J]} for J in self.Jlist}
        T.quanta_m2 = T.quanta_m1
        T.quanta_k1 = {J: {s: [(str(i), e) for i, e in enumerate(self.energy[J][s])]
                           for s in self.symlist[J]} for J in self.Jlist}
        T.quanta_k2 = T.quanta_k1


def linear_basis(B, Jmax, Jmin=0, m_filter=None, J_filter=None):
    """Linear rotor basis, energies B J(J+1) (rot/solution.py:239-251)."""
    jl, ml, kf, en = [], {}, {}, {}
    for J in range(Jmin, Jmax + 1):
        if J_filter is not None and not J_filter(J):
            continue
        ms = [m for m in range(-J, J + 1) if m_filter is None or m_filter(J, m)]
        if not ms:
            continue
        jl.append(J)
        ml[J] = ms
        c = 1.0 if J % 2 == 0 else 1j      # Wang function |J,0,tau=J mod 2> (basis.py:790-795)
        kf[J] = {"A": ([0], np.array([[c]], dtype=np.complex128))}
        en[J] = {"A": np.array([B * J * (J + 1)], dtype=np.float64)}
    return RotorBasis(jl, ml, kf, en)


# D2 species of a Wang function |J K ±> with z the quantization axis, from its
# behaviour under C2(z) (K parity) and C2(x) (|J k> -> (-1)^J |J -k>).
_D2 = {(0, 1): "A", (0, -1): "B1", (1, -1): "B2", (1, 1): "B3"}


def rigid_rotor_matrix(J, Bx, By, Bz):
    """<J k'|Bx Jx^2 + By Jy^2 + Bz Jz^2|J k> in the symmetric-top basis
    k = -J..J (molecule-fixed angular momentum, z the quantization axis):
    diagonal (Bx+By)/2 (J(J+1) - k^2) + Bz k^2, and (Bx-By)/4 times the
    Delta k = +-2 ladder elements."""
    ks = np.arange(-J, J + 1, dtype=np.float64)
    jj = J * (J + 1.0)
    H = np.diag((Bx + By) / 2.0 * (jj - ks ** 2) + Bz * ks ** 2)
    if J >= 1:
        k = ks[:-2]
        v = (Bx - By) / 4.0 * np.sqrt((jj - k * (k + 1)) * (jj - (k + 1) * (k + 2)))
        i = np.arange(k.size)
        H[i + 2, i] = v
        H[i, i + 2] = v
    return H


# Watson's centrifugal-distortion terms (rot/solution.py:288-386): the
# reference's attribute names of the constants per reduction.
WATSON_A = ("DeltaJ", "DeltaJK", "DeltaK", "deltaJ", "deltaK", "HJ", "HJK", "HKJ", "HK",
            "phiJ", "phiJK", "phiK")
WATSON_S = ("DeltaJ", "DeltaJK", "DeltaK", "d1", "d2", "HJ", "HJK", "HKJ", "HK", "h1", "h2", "h3")


def _ladder2(J):
    """J+^2 + J-^2 over k = -J..J: <k+2|.|k> = <k|.|k+2> =
    sqrt((J(J+1) - k(k+1)) (J(J+1) - (k+1)(k+2))) -- symmetric in the raising
    and lowering convention of the molecule-fixed J+-, as every Watson term
    uses it."""
    ks = np.arange(-J, J + 1, dtype=np.float64)
    jj = J * (J + 1.0)
    U = np.zeros((2 * J + 1, 2 * J + 1))
    if J >= 1:
        k = ks[:-2]
        i = np.arange(k.size)
        U[i + 2, i] = np.sqrt((jj - k * (k + 1)) * (jj - (k + 1) * (k + 2)))
    return U


def watson_matrix(J, reduction, consts):
    """The distortion terms of Watson's A- or S-reduced Hamiltonian at J over
    k = -J..J (rot/solution.py:288-386), constants by the reference's names
    (missing ones are zero):
      A: -DJ J^4 - DJK J^2 Jz^2 - DK Jz^4 - 1/2 {dJ J^2 + dK Jz^2, J+^2 + J-^2}
         + HJ J^6 + HJK J^4 Jz^2 + HKJ J^2 Jz^4 + HK Jz^6
         + 1/2 {phiJ J^4 + phiJK J^2 Jz^2 + phiK Jz^4, J+^2 + J-^2}
      S: -DJ J^4 - DJK J^2 Jz^2 - DK Jz^4 + d1 J^2 (J+^2 + J-^2) + d2 (J+^4 + J-^4)
         + HJ J^6 + HJK J^4 Jz^2 + HKJ J^2 Jz^4 + HK Jz^6
         + h1 J^4 (J+^2 + J-^2) + h2 J^2 (J+^4 + J-^4) + h3 (J+^6 + J-^6)."""
    names = {"A": WATSON_A, "watson_a": WATSON_A, "S": WATSON_S, "watson_s": WATSON_S}.get(reduction)
    if names is None:
        raise ValueError(f"Hamiltonian '{reduction}' is not available (use 'watson_a' or 'watson_s')")
    unknown = set(consts) - set(names)
    if unknown:
        raise ValueError(f"constants {sorted(unknown)} are not terms of the {reduction} reduction")
    c = {k: float(consts.get(k, 0.0)) for k in names}
    ks = np.arange(-J, J + 1, dtype=np.float64)
    j2 = J * (J + 1.0)
    Z2 = np.diag(ks ** 2)
    U = _ladder2(J)
    P2 = U + U.T
    anti = lambda D, X: 0.5 * (D @ X + X @ D)
    H = np.diag(-c["DeltaJ"] * j2 ** 2 - c["DeltaJK"] * j2 * ks ** 2 - c["DeltaK"] * ks ** 4
                + c["HJ"] * j2 ** 3 + c["HJK"] * j2 ** 2 * ks ** 2 + c["HKJ"] * j2 * ks ** 4 + c["HK"] * ks ** 6)
    if names is WATSON_A:
        H = H - anti(c["deltaJ"] * j2 * np.eye(ks.size) + c["deltaK"] * Z2, P2)
        H = H + anti(c["phiJ"] * j2 ** 2 * np.eye(ks.size) + c["phiJK"] * j2 * Z2 + c["phiK"] * Z2 @ Z2, P2)
    else:
        U2 = U @ U
        P4 = U2 + U2.T
        U3 = U2 @ U
        P6 = U3 + U3.T
        H = H + c["d1"] * j2 * P2 + c["d2"] * P4 + c["h1"] * j2 ** 2 * P2 + c["h2"] * j2 * P4 + c["h3"] * P6
    return H


@functools.lru_cache(maxsize=None)
def _wang(J):
    """Wang functions of J as columns over k = -J..J (K = 0, then K = 1..J
    with the + and - combinations) and their D2 species labels (cached:
    callers read, never modify, the arrays)."""
    Kv = np.concatenate([[0], np.repeat(np.arange(1, J + 1), 2)]).astype(np.int64)
    Sv = np.concatenate([[1], np.tile([1, -1], J)]).astype(np.int64)
    W = np.zeros((2 * J + 1, Kv.size))
    c = np.arange(Kv.size)
    W[J, 0] = 1.0
    W[J + Kv[1:], c[1:]] = 1.0 / math.sqrt(2.0)
    W[J - Kv[1:], c[1:]] = Sv[1:] / math.sqrt(2.0)
    c2x = Sv * (-1) ** J
    labels = np.array([_D2[(int(K) % 2, int(x))] for K, x in zip(Kv, c2x)])
    W.setflags(write=False)
    labels.setflags(write=False)
    return W, labels


def asymtop_basis(Bx, By, Bz, Jmax, Jmin=0, m_filter=None, J_filter=None, species=None,
                  watson=None, distortion=None):
    """Rigid asymmetric-top basis (rot/solution.py:160-391, rot/basis.py:702-801):
    per J the eigenstates of Bx Jx^2 + By Jy^2 + Bz Jz^2 [cm^-1], with the
    rotational constants along the molecule-frame axes of the tensors that
    will be built on it.  The Wang functions split each J into the four D2
    species (the Hamiltonian couples Delta k = 0, +-2 and commutes with
    k -> -k); a species block is diagonalised on its own and its states are
    stored as coefficient columns over primitive k = -J..J, ordered by energy.
    species: the D2 species to keep (all four by default).
    watson: 'watson_a' or 'watson_s' adds Watson's centrifugal-distortion
    terms with the constants of ``distortion`` (a dict by the reference's
    names, rot/solution.py:288-386; see watson_matrix): they couple
    Delta k = 0, +-2, +-4, +-6 and keep the D2 species."""
    jl, ml, kf, en = [], {}, {}, {}
    blocks = []   # (J, species, Wang columns, species block of H)
    for J in range(Jmin, Jmax + 1):
        if J_filter is not None and not J_filter(J):
            continue
        ms = [m for m in range(-J, J + 1) if m_filter is None or m_filter(J, m)]
        if not ms:
            continue
        H = None
        if watson is not None:
            H = rigid_rotor_matrix(J, Bx, By, Bz) + watson_matrix(J, watson, distortion or {})
        any_kept = False
        for sym, Wm, (X, Y, Z) in _species_blocks(J):
            if species is not None and sym not in species:
                continue
            # the rigid rotor's species block from the cached blocks of Jx^2, Jy^2, Jz^2
            Hs = Bx * X + By * Y + Bz * Z if H is None else Wm.T @ H @ Wm
            blocks.append((J, sym, Wm, Hs))
            any_kept = True
        if not any_kept:   # no kept species at this J
            continue
        jl.append(J)
        ml[J] = ms
        kf[J], en[J] = {}, {}
    # every species block diagonalised in one batched call: blocks padded to
    # a common order with decoupled diagonal entries above every block's
    # Gershgorin bound, so each block's eigenpairs come first in its slice
    if blocks:
        nmax = max(b[3].shape[0] for b in blocks)
        big = max(float(np.abs(b[3]).sum(axis=1).max()) for b in blocks) + 1.0
        S = np.zeros((len(blocks), nmax, nmax))
        for i, (_, _, _, Hs) in enumerate(blocks):
            ns = Hs.shape[0]
            S[i, :ns, :ns] = Hs
            pad = np.arange(ns, nmax)
            S[i, pad, pad] = big * (2.0 + pad)
        E, V = np.linalg.eigh(S)
        for i, (J, sym, Wm, Hs) in enumerate(blocks):
            ns = Hs.shape[0]
            kf[J][sym] = (list(range(-J, J + 1)), (Wm @ V[i, :ns, :ns]).astype(np.complex128))
            en[J][sym] = E[i, :ns].copy()
    return RotorBasis(jl, ml, kf, en)


@functools.lru_cache(maxsize=None)
def _species_blocks(J):
    """Per D2 species of J (in the order A, B1, B2, B3, present ones only):
    its Wang columns and the species blocks of Jx^2, Jy^2, Jz^2 (cached, read
    only): the rigid rotor's block is their combination with Bx, By, Bz."""
    W, labels = _wang(J)
    X = rigid_rotor_matrix(J, 1.0, 0.0, 0.0)
    Y = rigid_rotor_matrix(J, 0.0, 1.0, 0.0)
    Z = rigid_rotor_matrix(J, 0.0, 0.0, 1.0)
    out = []
    for sym in ("A", "B1", "B2", "B3"):
        sel = labels == sym
        if not np.any(sel):
            continue
        Wm = W[:, sel]
        mats = tuple(Wm.T @ M @ Wm for M in (X, Y, Z))
        for a in (Wm,) + mats:
            a.setflags(write=False)
        out.append((sym, Wm, mats))
    return tuple(out)


@functools.lru_cache(maxsize=None)
def species_sizes(J):
    """Number of states of each D2 species at J (the Wang-function count;
    cached, do not modify)."""
    _, labels = _wang(J)
    return {sym: int(np.sum(labels == sym)) for sym in ("A", "B1", "B2", "B3")}


def hamiltonian(basis):
    """Field-free Hamiltonian as a rank-0 tensor: K = diag(E), M = identity
    (labtens.py:285-308 with the eigenbasis of H0)."""
    T = CarTens()
    basis.attrs(T)
    T.rank, T.cart, T.os = 0, ["0"], [(0, 0)]
    T.kmat, T.mmat = {}, {}
    for J in basis.Jlist:
        for s in basis.symlist[J]:
            dm = len(basis.mlist[J])
            T.kmat.setdefault((J, J), {})[(s, s)] = {0: csr_matrix(np.diag(basis.energy[J][s]).astype(np.complex128))}
            T.mmat.setdefault((J, J), {})[(s, s)] = {0: {"0": csr_matrix(np.eye(dm, dtype=np.complex128))}}
    return T


def _k_prim(comps_w, w, J1, kp1, J2, kp2):
    """Primitive-k matrix of K_w: sum_s 3j(J2 w J1; k2 s -k1) UsT_s (-1)^|k1|;
    only k1 = k2 + s couples."""
    prim = np.zeros((len(kp1), len(kp2)), dtype=np.complex128)
    index1 = {k: a for a, k in enumerate(kp1)}
    for b, k2 in enumerate(kp2):
        for s, u in comps_w.items():
            a = index1.get(k2 + s)
            if a is None:
                continue
            t = wigner3j(J2, w, J1, k2, s, -(k2 + s))
            if t != 0.0:
                prim[a, b] += t * u * (-1) ** abs(k2 + s)
    return prim


def _k_block(comps_w, w, J1, kf1, J2, kf2, cache=None):
    """K_w[bra J1 states, ket J2 states] (labtens.py:482-496 + overlap_k).  The
    primitive matrix is shared by every species pair of (J1, J2) via cache."""
    kp1, c1 = kf1
    kp2, c2 = kf2
    key = (w, J1, tuple(kp1), J2, tuple(kp2))
    prim = None if cache is None else cache.get(key)
    if prim is None:
        prim = _k_prim(comps_w, w, J1, kp1, J2, kp2)
        if cache is not None:
            cache[key] = prim
    return c1.conj().T @ prim @ c2


def _m_block(ucols_w, w, J1, ms1, J2, ms2, carts):
    """M_w,cart[m1, m2] for every cart (labtens.py:501-523 + overlap_m)."""
    out = {c: np.zeros((len(ms1), len(ms2)), dtype=np.complex128) for c in carts}
    fac = math.sqrt((2 * J1 + 1) * (2 * J2 + 1))
    index2 = {m: b for b, m in enumerate(ms2)}
    for a, m1 in enumerate(ms1):
        for s, col in ucols_w.items():          # only m2 = m1 - s couples (|s| <= w)
            b = index2.get(m1 - s)
            if b is None:
                continue
            t = wigner3j(J2, w, J1, m1 - s, s, -m1)
            if t == 0.0:
                continue
            for i, c in enumerate(carts):
                if col[i] != 0:
                    out[c][a, b] += col[i] * t * fac * (-1) ** abs(m1)
    return out


def restrict(T, Jmax):
    """Copy of a built operator on the sub-basis J <= Jmax (its blocks are the
    same matrices: matrix elements of a tensor between (J1, J2) do not depend
    on the rest of the basis)."""
    from .field import CarTens
    R = CarTens()
    keep = [J for J in T.Jlist1 if J <= Jmax]
    R.Jlist1 = list(keep); R.Jlist2 = list(keep)
    for name in ("symlist1", "symlist2", "dim1", "dim2", "dim_m1", "dim_m2", "dim_k1", "dim_k2",
                 "quanta_m1", "quanta_m2", "quanta_k1", "quanta_k2"):
        if hasattr(T, name):
            setattr(R, name, {J: v for J, v in getattr(T, name).items() if J <= Jmax})
    R.rank, R.cart, R.os = T.rank, list(T.cart), list(getattr(T, "os", []))
    R.kmat = {Jp: v for Jp, v in T.kmat.items() if Jp[0] <= Jmax and Jp[1] <= Jmax}
    R.mmat = {Jp: v for Jp, v in T.mmat.items() if Jp[0] <= Jmax and Jp[1] <= Jmax}
    return R


def _build(basis, rank, carts, comps, ucols):
    T = CarTens()
    basis.attrs(T)
    T.rank, T.cart = rank, list(carts)
    T.os = [(o, s) for o in sorted(comps) for s in sorted(comps[o])]
    T.kmat, T.mmat = {}, {}
    dJ = max(comps.keys()) if comps else 0
    kcache, mcache = {}, {}
    for J2 in basis.Jlist:
        for J1 in basis.Jlist:
            if abs(J1 - J2) > dJ:
                continue
            for s1 in basis.symlist[J1]:
                for s2 in basis.symlist[J2]:
                    kd, md = {}, {}
                    for w in sorted(comps):
                        K = _k_block(comps[w], w, J1, basis.kfun[J1][s1], J2, basis.kfun[J2][s2], kcache)
                        K[np.abs(K) < _EPS] = 0
                        if not np.any(K):
                            continue
                        mkey = (w, J1, J2)
                        if mkey not in mcache:
                            mcache[mkey] = _m_block(ucols[w], w, J1, basis.mlist[J1], J2, basis.mlist[J2], carts)
                        Ms = {c: M.copy() for c, M in mcache[mkey].items()}
                        mm = {}
                        for c, M in Ms.items():
                            M[np.abs(M) < _EPS] = 0
                            if np.any(M):
                                mm[c] = csr_matrix(M)
                        if mm:
                            kd[w] = csr_matrix(K)
                            md[w] = mm
                    if kd:
                        T.kmat.setdefault((J1, J2), {})[(s1, s2)] = kd
                        T.mmat.setdefault((J1, J2), {})[(s1, s2)] = md
    return T


def tensor(tens, basis):
    """Lab-frame operator of a rank-1 or rank-2 molecular-frame tensor (in au)."""
    rank, comps, ucols = spherical_components(tens)
    return _build(basis, rank, CART[rank], comps, ucols)


_KPRIM = {}
_MZ = {}


def zfield_m0_operator(basis, tens, start=None):
    """Operator of a rank-1 or rank-2 molecular-frame tensor on an m = 0
    basis (every J holds m = 0 only) for a field along Z: H(F) = F_z^rank A.
    A is one sparse matrix in the block order of the built operators (J,
    then species, then state) -- the matrix that tensor(tens, basis)
    flattens to under a Z field, up to rounding -- made from the same
    primitives with one dense product per (J1, J2, species of J2):
    K_w = C1^T P_w c2 over every species of J1 at once, scaled by
    M_w,z..z(m=0, m=0).  Entries below 1e-12 of their product's largest
    (rounding left by symmetry-forbidden couplings) are dropped.  The
    primitive-k matrices P_w and the M elements are kept across calls.

    start=None: returns (A, energies) over the whole basis.  start = (J,
    species): only the blocks reachable from that block are built -- the
    invariant subspace its states live in under every Z field -- and
    (A, energies, index) is returned, index[i] being the whole-basis
    position of domain state i (domain states in basis order)."""
    for J in basis.Jlist:
        if list(basis.mlist[J]) != [0]:
            raise ValueError("zfield_m0_operator needs a basis with m = 0 only")
    rank, comps, ucols = spherical_components(tens)
    cart = "z" * rank
    blocks, pos = [], 0
    boff, rows_of = {}, {}
    for J in basis.Jlist:
        r = 0
        for sp in basis.symlist[J]:
            ns = basis.kfun[J][sp][1].shape[1]
            boff[(J, sp)] = pos
            rows_of[(J, sp)] = (r, r + ns)
            blocks.append((J, sp))
            pos += ns
            r += ns
    n = pos
    Cs = {J: np.concatenate([basis.kfun[J][sp][1] for sp in basis.symlist[J]], axis=1) for J in basis.Jlist}
    real = all(np.all(C.imag == 0) for C in Cs.values())
    Cr = {J: (C.real if real else C) for J, C in Cs.items()}
    dJ = max(comps.keys()) if comps else 0
    ckey = tuple((w, tuple(sorted(comps[w].items()))) for w in sorted(comps))

    def mz(w, J1, J2):
        key = (rank, w, J1, J2)
        v = _MZ.get(key)
        if v is None:
            v = _MZ[key] = _m_block(ucols[w], w, J1, [0], J2, [0], CART[rank])[cart][0, 0]
        return v

    def prim(w, J1, J2):
        key = (ckey, w, J1, J2)
        P = _KPRIM.get(key)
        if P is None:
            P = _k_prim(comps[w], w, J1, list(range(-J1, J1 + 1)), J2, list(range(-J2, J2 + 1)))
            if real and np.all(P.imag == 0):
                P = P.real
            P = csr_matrix(P)   # Delta k = sigma: at most five diagonals
            _KPRIM[key] = P
        return P

    def column(b2):
        """Non-zero blocks (b1, K) of the column block b2 = (J2, species)."""
        J2, s2 = b2
        r0, r1 = rows_of[b2]
        c2 = Cr[J2][:, r0:r1]
        out = []
        for J1 in range(J2 - dJ, J2 + dJ + 1):
            if J1 not in Cr:
                continue
            acc = None
            for w in sorted(comps):
                m = mz(w, J1, J2)
                if abs(m) < _EPS:
                    continue
                K = Cr[J1].conj().T @ (prim(w, J1, J2) @ c2)
                K[np.abs(K) < _EPS] = 0
                acc = K * m if acc is None else acc + K * m
            if acc is None:
                continue
            top = np.abs(acc).max()
            if top == 0:
                continue
            acc[np.abs(acc) < 1e-12 * top] = 0
            for s1 in basis.symlist[J1]:
                q0, q1 = rows_of[(J1, s1)]
                sub = acc[q0:q1]
                if np.any(sub):
                    out.append(((J1, s1), sub))
        return out

    ent = {}
    if start is None:
        keep = blocks
        for b2 in blocks:
            for b1, K in column(b2):
                ent[(b1, b2)] = K
    else:
        start = tuple(start)
        if start not in boff:
            raise ValueError(f"start block {start} is not in the basis")
        seen, queue = {start}, [start]
        while queue:
            b2 = queue.pop()
            for b1, K in column(b2):
                ent[(b1, b2)] = K
                if b1 not in seen:
                    seen.add(b1)
                    queue.append(b1)
        keep = [b for b in blocks if b in seen]
    doff, index, dpos = {}, [], 0
    for b in keep:
        ns = rows_of[b][1] - rows_of[b][0]
        doff[b] = dpos
        index.append(boff[b] + np.arange(ns))
        dpos += ns
    index = np.concatenate(index) if index else np.zeros(0, dtype=np.int64)
    R, Ccol, V = [], [], []
    for (b1, b2), K in ent.items():
        r, c = np.nonzero(K)
        R.append(r + doff[b1]); Ccol.append(c + doff[b2]); V.append(K[r, c].astype(np.complex128))
    if R:
        A = csr_matrix((np.concatenate(V), (np.concatenate(R), np.concatenate(Ccol))), shape=(dpos, dpos))
    else:
        A = csr_matrix((dpos, dpos), dtype=np.complex128)
    A.sort_indices()
    energies = np.concatenate([basis.energy[J][sp] for (J, sp) in keep]) if keep else np.zeros(0)
    if start is None:
        return A, energies
    return A, energies, index


def named(name, basis):
    """'cos2theta' (cos^2 theta - 1/3) or 'costheta' as rank-0 tensors
    (labtens.py:362-385)."""
    if name.lower() == "cos2theta":
        comps, ucols = {2: {0: 1.0}}, {2: {0: np.array([2.0 / 3.0])}}
    elif name.lower() == "costheta":
        comps, ucols = {1: {0: 1.0}}, {1: {0: np.array([1.0])}}
    else:
        raise ValueError(f"unknown name for tensor operator: '{name}'")
    return _build(basis, 0, ["0"], comps, ucols)
