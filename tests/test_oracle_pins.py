"""Closed-form pins of the oracle functions that the reference's golden data
leave open: the thresholded field contraction (richmol/field.py:1094-1142),
tomat(thresh=...) (field.py:540-552), __sub__ (field.py:1292-1301), the unit
factors of richmol/convert_units.py, and init_state's temp=None / temp=0 /
temp>0 / zpe branches (richmol/tdse.py:177-262).

Every expected value below is worked out by hand from a small hand-built
tensor, so that a plausible mistake fails a check: '>' for '>=' in either
threshold, thresholding the Cartesian terms before their sum instead of
after, keeping a screened field product, a dropped sqrt on the Boltzmann
weights, an off-by-one in the partition-function cut, A+B for A-B, a factor
h for hbar or a missing 1e2.  The product's host mirror (richmol_amd.field,
richmol_amd.tdse.init_state) is held to the same values.
"""
import numpy as np
import pytest
from scipy.sparse import csr_matrix

from oracle import units as ounits
from oracle.cartens import OracleCarTens, flatten, unflatten
from oracle.tdse import OracleTDSE
from richmol_amd import units as punits
from richmol_amd.field import CarTens
from richmol_amd.tdse import TDSE

# A rank-2 tensor on one (J=1, 'A') block with dm = 3, dk = 1.  M matrices per
# Cartesian component (only these five are non-zero):
M_CART = {
    "xx": [[1.0, 0, 0], [0, 0.3, 0], [0, 0, 0.375]],
    "yy": [[-0.4, 0, 0], [0, 0.1, 0], [0, 0, 0]],
    "xz": [[0, 5.0, 0], [0, 0, 0], [0, 0, 0]],
    "yz": [[0, 0, 0], [2.0, 0, 0], [0, 0, 0]],
    "xy": [[0, 0, 0], [0, 0, 0], [0.125, 0, 0]],
}
# Field (2, 3, 0.5): products xx 4, yy 9, xy = yx 6, xz = zx 1, yz = zy 1.5,
# zz 0.25.  With thresh 1.5 the products xz, zx (1 < 1.5) and zz are dropped
# and yz (exactly 1.5) is kept; then
#   MF[0,0] = 4*1 + 9*(-0.4) = 0.4   dropped after the sum (each term >= 1.5)
#   MF[1,1] = 4*0.3 + 9*0.1  = 2.1   kept (each term < 1.5, the sum is not)
#   MF[0,1] = xz*5                   absent (its field product is screened)
#   MF[1,0] = yz*2           = 3     kept (yz is exactly at the threshold)
#   MF[2,2] = 4*0.375        = 1.5   kept (exactly at the threshold)
#   MF[2,0] = xy*0.125       = 0.75  dropped
FIELD = [2.0, 3.0, 0.5]
THRESH = 1.5
MF_THRESH = np.array([[0, 0, 0], [3.0, 4 * 0.3 + 9 * 0.1, 0], [0, 0, 1.5]])
MF_FULL = np.array([[4 - 9 * 0.4, 5.0, 0], [3.0, 4 * 0.3 + 9 * 0.1, 0], [0.75, 0, 1.5]])


def hand_tensor(cls, kval=2.0, scale=1.0, rank=2, mcart=None):
    T = cls()
    mcart = M_CART if mcart is None else mcart
    for side in ("1", "2"):
        setattr(T, "Jlist" + side, [1])
        setattr(T, "symlist" + side, {1: ["A"]})
        setattr(T, "dim" + side, {1: {"A": 3}})
        setattr(T, "dim_m" + side, {1: {"A": 3}})
        setattr(T, "dim_k" + side, {1: {"A": 1}})
    T.rank = rank
    T.cart = list(mcart)
    T.os = [(0, 0)]
    T.kmat = {(1, 1): {("A", "A"): {0: csr_matrix(np.array([[kval]], dtype=np.complex128))}}}
    T.mmat = {(1, 1): {("A", "A"): {0: {c: csr_matrix(scale * np.array(M, dtype=np.complex128))
                                        for c, M in mcart.items()}}}}
    return T


def _mf_dense(T):
    blk = T.mfmat.get((1, 1), {}).get(("A", "A"), {})
    return blk[0].toarray() if 0 in blk else np.zeros((3, 3))


@pytest.mark.parametrize("cls", [OracleCarTens, CarTens])
def test_field_threshold_closed_form(cls):
    T = hand_tensor(cls)
    T.field(FIELD, thresh=THRESH)
    mf = _mf_dense(T)
    assert np.array_equal(np.abs(mf) > 0, np.abs(MF_THRESH) > 0)
    assert np.allclose(mf, MF_THRESH, rtol=0, atol=1e-15)
    # without a threshold every product and entry stays
    T.field(FIELD)
    assert np.allclose(_mf_dense(T), MF_FULL, rtol=0, atol=1e-15)
    # a threshold above every product: no MF block at all (update then skips the expv)
    T.field(FIELD, thresh=10.0)
    assert len(T.mfmat) == 0


@pytest.mark.parametrize("cls", [OracleCarTens, CarTens])
def test_field_threshold_rank0_product(cls):
    """The '0' string's product is 1: kept at thresh 1 (>=), dropped above."""
    T = hand_tensor(cls, rank=0, mcart={"0": [[7.0, 0, 0], [0, 8.0, 0], [0, 0, 9.0]]})
    T.field([0.0, 0.0, 1.0], thresh=1.0)
    assert np.allclose(_mf_dense(T), np.diag([7.0, 8.0, 9.0]))
    T.field([0.0, 0.0, 1.0], thresh=1.0000001)
    assert len(T.mfmat) == 0


@pytest.mark.parametrize("cls", [OracleCarTens, CarTens])
def test_tomat_threshold_is_strict(cls):
    """tomat keeps |x| > thresh: K = 2 times MF_FULL has the entries
    0.8, 10, 6, 4.2, 1.5, 3.0; at thresh 3.0 the 3.0 entry goes (strict)."""
    T = hand_tensor(cls, kval=2.0)
    T.field(FIELD)
    full = T.tomat(form="full", repres="dense")
    assert np.allclose(full, 2.0 * MF_FULL, rtol=0, atol=1e-15)
    cut = T.tomat(form="full", repres="dense", thresh=3.0)
    want = np.where(np.abs(2.0 * MF_FULL) > 3.0, 2.0 * MF_FULL, 0.0)
    assert np.count_nonzero(want) == 3
    assert np.allclose(cut, want, rtol=0, atol=1e-15)
    # one Cartesian component: M_xx kron K = diag(2, 0.6, 0.75); thresh 0.6 drops 0.6
    cx = T.tomat(form="full", repres="dense", cart="xx", thresh=0.6)
    assert np.allclose(cx, np.diag([2.0, 0.0, 0.75]), rtol=0, atol=1e-15)


@pytest.mark.parametrize("cls", [OracleCarTens, CarTens])
def test_sub_is_dense_difference(cls):
    """(A - B) equals the dense difference of the two operators (not A + B):
    B = 3 x the M tensors of A with K = 5, so A - B = (2 - 15) MF_FULL."""
    A = hand_tensor(cls, kval=2.0)
    B = hand_tensor(cls, kval=5.0, scale=3.0)
    A.field(FIELD)
    B.field(FIELD)
    D = A - B
    want = (2.0 - 15.0) * MF_FULL
    assert np.allclose(D.tomat(form="full", repres="dense"), want, rtol=0, atol=1e-13)
    if cls is OracleCarTens:
        v = np.array([1.0, -2.0j, 0.5])
        got = flatten(D, D.vec(unflatten(D, v)))
        assert np.allclose(got, want @ v, rtol=0, atol=1e-13)
        # the operands are left as they were
        assert np.allclose(B.tomat(form="full", repres="dense"), 15.0 * MF_FULL, rtol=0, atol=1e-13)


def test_unit_factors_from_defining_relations():
    """Unit factors from their definitions with independently tabulated
    constants (scipy.constants): MHz = 1e6 Hz / (c in cm/s); the atomic unit
    of dipole is e a0, of polarizability e^2 a0^2 / E_h; the energy of
    1 au x 1 V/m in cm^-1 is (au)/(h c 100)."""
    import scipy.constants as sc
    c_cm = sc.c * 100.0
    for U in (ounits, punits):
        assert U.MHz_to_invcm() == pytest.approx(1e6 / c_cm, rel=1e-15)
        assert U.MHz_to_invcm() == pytest.approx(3.33564095e-5, rel=1e-8)
        e, a0, Eh = sc.e, sc.physical_constants["Bohr radius"][0], sc.physical_constants["Hartree energy"][0]
        assert U.AU_DIP == pytest.approx(e * a0, rel=1e-9)
        assert U.AU_POL == pytest.approx(e ** 2 * a0 ** 2 / Eh, rel=1e-9)
        assert U.HBAR == pytest.approx(sc.h / (2 * np.pi), rel=1e-15)
        assert U.AUdip_x_Vm_to_invcm() == pytest.approx(e * a0 / (sc.h * c_cm), rel=1e-9)
        assert U.AUdip_x_Vm_to_invcm() == pytest.approx(4.26810e-7, rel=1e-5)
        assert U.AUpol_x_Vm_to_invcm() == pytest.approx(e ** 2 * a0 ** 2 / Eh / (sc.h * c_cm), rel=1e-9)
        # MHz as an energy unit: h x 1e6 Hz = MHz_to_invcm cm^-1 (tdse.py:129-133)
        assert U.enr_units()["mhz"] == pytest.approx(sc.h * 1e6, rel=1e-15)
        assert U.enr_units()["invcm"] == pytest.approx(sc.h * c_cm, rel=1e-15)


def _diag_h0(cls, energies):
    """Rank-0 diagonal Hamiltonian with the given energies (one J block per state)."""
    T = cls()
    Js = list(range(len(energies)))
    for side in ("1", "2"):
        setattr(T, "Jlist" + side, Js)
        setattr(T, "symlist" + side, {J: ["A"] for J in Js})
        for name in ("dim", "dim_m", "dim_k"):
            setattr(T, name + side, {J: {"A": 1} for J in Js})
    T.rank, T.cart, T.os = 0, ["0"], [(0, 0)]
    T.kmat = {(J, J): {("A", "A"): {0: csr_matrix(np.array([[e]], dtype=np.complex128))}}
              for J, e in zip(Js, energies)}
    T.mmat = {(J, J): {("A", "A"): {0: {"0": csr_matrix(np.eye(1, dtype=np.complex128))}}} for J in Js}
    return T


# energies in cm^-1 of five states and kT at 1 K = 0.6950348 cm^-1
ENRS = [5.0, 5.3, 5.9, 6.5, 9.0]
KT_1K = 0.6950348


def _init(cls, H, **kw):
    t = (OracleTDSE if cls is OracleCarTens else TDSE)(t_end=10.0, dt=1.0)
    return t.init_state(H, **kw)


@pytest.mark.parametrize("cls", [OracleCarTens, CarTens])
def test_init_state_branches_diagonal(cls):
    # temp=None: every eigenvector, unweighted (identity rows for a diagonal H)
    v = _init(cls, _diag_h0(cls, ENRS))
    assert v.shape == (5, 5) and np.array_equal(v, np.eye(5))
    # temp=0: the lowest one only
    v = _init(cls, _diag_h0(cls, ENRS), temp=0)
    assert v.shape == (1, 5) and np.array_equal(v, np.eye(5)[:1])
    # temp=1 K: rows sqrt(w_i) e_i, w_i = exp(-(E_i - E_0)/kT)/Z, kept while
    # 1 - sum_{j<=i} w_j > thresh
    w = np.exp(-(np.array(ENRS) - ENRS[0]) / KT_1K)
    w /= w.sum()
    tail = 1.0 - np.cumsum(w)
    for thresh in (1e-3, tail[2] * 1.0001, tail[2] * 0.9999):
        keep = [i for i in range(5) if tail[i] > thresh]
        v = _init(cls, _diag_h0(cls, ENRS), temp=1.0, thresh=thresh)
        want = np.zeros((len(keep), 5))
        want[np.arange(len(keep)), keep] = np.sqrt(w[keep])
        assert v.shape == want.shape
        assert np.allclose(v, want, rtol=1e-6, atol=0)
    assert len([i for i in range(5) if tail[i] > tail[2] * 1.0001]) == 2
    assert len([i for i in range(5) if tail[i] > tail[2] * 0.9999]) == 3


@pytest.mark.parametrize("cls", [OracleCarTens, CarTens])
def test_init_state_zpe(cls):
    # zpe shifts the energies before the weights: zpe = 0 gives the same rows
    # as the default (lowest energy) while exp(-E/kT) does not underflow ...
    a = _init(cls, _diag_h0(cls, ENRS), temp=1.0)
    b = _init(cls, _diag_h0(cls, ENRS), temp=1.0, zpe=0.0)
    assert np.allclose(a, b, rtol=1e-12, atol=0)
    # ... and with energies near 5000 cm^-1 at 1 K it does (exp(-7194) = 0):
    # zero weights, 0/0, no row passes the cut -- only the default zpe works
    high = [e + 4995.0 for e in ENRS]
    assert np.allclose(_init(cls, _diag_h0(cls, high), temp=1.0), a, rtol=1e-9, atol=0)
    with np.errstate(invalid="ignore"):
        assert _init(cls, _diag_h0(cls, high), temp=1.0, zpe=0.0).shape[0] == 0
    # zpe above the lowest energy is refused
    with pytest.raises(AssertionError):
        _init(cls, _diag_h0(cls, ENRS), temp=1.0, zpe=5.5)


def test_init_state_eigh_branch():
    """Non-diagonal H (a sum, so the eigh branch): rows are eigenvectors of H
    (H v = E v, orthonormal), ascending, Boltzmann weights of their energies."""
    h0 = _diag_h0(OracleCarTens, ENRS)
    T = OracleCarTens()
    for attr in ("Jlist1", "Jlist2", "symlist1", "symlist2", "dim1", "dim2", "dim_m1", "dim_m2",
                 "dim_k1", "dim_k2"):
        setattr(T, attr, getattr(h0, attr))
    T.rank, T.cart, T.os = 1, ["z"], [(1, 0)]
    T.kmat = {(0, 1): {("A", "A"): {0: csr_matrix(np.array([[1.0 + 0j]]))}},
              (1, 0): {("A", "A"): {0: csr_matrix(np.array([[1.0 + 0j]]))}},
              (2, 3): {("A", "A"): {0: csr_matrix(np.array([[0.5j]]))}},
              (3, 2): {("A", "A"): {0: csr_matrix(np.array([[-0.5j]]))}}}
    T.mmat = {Jp: {("A", "A"): {0: {"z": csr_matrix(np.eye(1, dtype=np.complex128))}}} for Jp in T.kmat}
    T.field([0.0, 0.0, 0.2])
    Hs = T + h0
    Hd = Hs.tomat(form="full", repres="dense")
    assert np.allclose(Hd, Hd.conj().T)
    assert abs(Hd[0, 1] - 0.2) < 1e-15 and abs(Hd[2, 3] - 0.1j) < 1e-15
    v = _init(OracleCarTens, Hs, temp=1.0, thresh=1e-6)
    E = np.linalg.eigvalsh(Hd)
    w = np.exp(-(E - E[0]) / KT_1K)
    w /= w.sum()
    keep = [i for i in range(5) if 1 - w[: i + 1].sum() > 1e-6]
    assert v.shape[0] == len(keep)
    for r, i in enumerate(keep):
        x = v[r]
        assert abs(np.vdot(x, x).real - w[i]) < 1e-7 * max(w[i], 1e-300) + 1e-300
        u = x / np.linalg.norm(x)
        assert np.allclose(Hd @ u, E[i] * u, atol=1e-12)
    # temp=None: all eigenvectors unweighted and orthonormal
    V = _init(OracleCarTens, T + h0)
    assert np.allclose(V @ V.conj().T, np.eye(5), atol=1e-12)
