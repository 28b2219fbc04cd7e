"""Shared builders for the tests: operators from the committed golden data and
the synthetic linear-rotor workloads."""
import os

import numpy as np
from scipy.sparse import csr_matrix

from richmol_amd import rot, units
from richmol_amd.field import CarTens

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _one_state_basis(T, Js):
    T.Jlist1 = list(Js); T.Jlist2 = list(Js)
    T.symlist1 = {J: ["A"] for J in Js}; T.symlist2 = {J: ["A"] for J in Js}
    for side in (1, 2):
        for name in ("dim", "dim_m", "dim_k"):
            setattr(T, f"{name}{side}", {J: {"A": 1} for J in Js})


def ocs_fixture():
    """(H0, alpha, data) of the reference TDSE test basis (even J<=30, m=0)."""
    d = np.load(os.path.join(GOLDEN, "ocs_alignment.npz"))
    Js = [int(J) for J in d["J"]]
    carts = [str(c) for c in d["carts"]]
    h0 = CarTens()
    _one_state_basis(h0, Js)
    h0.rank, h0.cart, h0.os = 0, ["0"], [(0, 0)]
    h0.kmat = {(J, J): {("A", "A"): {0: csr_matrix(np.array([[e]], dtype=np.complex128))}}
               for J, e in zip(Js, d["energies"])}
    h0.mmat = {(J, J): {("A", "A"): {0: {"0": csr_matrix(np.eye(1, dtype=np.complex128))}}} for J in Js}
    a = CarTens()
    _one_state_basis(a, Js)
    a.rank, a.cart, a.os = 2, carts, [(0, 0)] + [(2, s) for s in range(-2, 3)]
    a.kmat, a.mmat = {}, {}
    for (J1, J2), ir, k, mv in zip(d["blk_J"], d["blk_irrep"], d["blk_k"], d["blk_m"]):
        J1, J2, ir = int(J1), int(J2), int(ir)
        a.kmat.setdefault((J1, J2), {}).setdefault(("A", "A"), {})[ir] = csr_matrix(np.array([[k]]))
        a.mmat.setdefault((J1, J2), {}).setdefault(("A", "A"), {})[ir] = {
            c: csr_matrix(np.array([[m]])) for c, m in zip(carts, mv) if m != 0}
    return h0, a, d


from richmol_amd.molecules import (LARGE_ALPHA, LARGE_B, LARGE_CHIRP, LARGE_E0, LARGE_MU,  # noqa: F401
                                   WATER_ALPHA, WATER_B, WATER_MU, WATER_STATIC, centrifuge_fields,
                                   large_top, large_top_tensors, linear_rotor, water, water_fields)
from richmol_amd import pulses


def gaussian_pulse_field(t_fs, fwhm_fs=300.0, intensity=5e11):
    """Cycle-averaged Z field of a Gaussian pulse: E(t) = E0 exp(-2 ln2 (t-t0)^2/tau^2)/sqrt(2),
    tau = intensity FWHM, t0 = 2.5 tau / 2, E0 from the peak intensity [W/cm^2];
    exactly zero where the envelope factor is below 1e-30."""
    return pulses.envelope(t_fs, fwhm_fs, intensity)
