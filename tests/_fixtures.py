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


def linear_rotor(Jmax, B=0.2029, alpha=(26.0, 26.0, 53.6), m_filter=None):
    """(H0, -1/2 alpha AUpol, cos^2-1/3 operator, basis) of a linear rotor."""
    bas = rot.linear_basis(B, Jmax, m_filter=m_filter)
    h0 = rot.hamiltonian(bas)
    pol = rot.tensor(np.diag(alpha), bas)
    H = pol * (-0.5 * units.AUpol_x_Vm_to_invcm())
    cos2 = rot.named("cos2theta", bas)
    return h0, H, cos2, bas


# Water-like asymmetric top (the asymmetric-top configuration): A, B, C =
# 27.88, 14.51, 9.28 cm^-1, frame x = c, y = a, z = b (the C2 axis, along
# which the dipole lies), dipole (0, 0, -0.7288) au, polarizability
# diag(9.14, 9.87, 9.45) au.
WATER_B = (9.28, 27.88, 14.51)          # (Bx, By, Bz)
WATER_MU = (0.0, 0.0, -0.7288)
WATER_ALPHA = (9.14, 9.87, 9.45)
WATER_STATIC = 5e6                      # 50 kV/cm along Z [V/m]


def water(jmax, m_filter=None):
    """(H0, [dipole, polarizability (static), polarizability (pulse)], basis):
    the interaction -mu.F_s - 1/2 F_s.alpha.F_s - 1/2 <F_p.alpha.F_p> with a
    static Z field F_s and a cycle-averaged X pulse F_p (the cross term of the
    static and the oscillating field averages to zero, so the two act through
    separate tensors)."""
    bas = rot.asymtop_basis(*WATER_B, jmax, m_filter=m_filter)
    h0 = rot.hamiltonian(bas)
    mu = rot.tensor(np.array(WATER_MU), bas) * (-units.AUdip_x_Vm_to_invcm())
    als = rot.tensor(np.diag(WATER_ALPHA), bas) * (-0.5 * units.AUpol_x_Vm_to_invcm())
    alp = rot.tensor(np.diag(WATER_ALPHA), bas) * (-0.5 * units.AUpol_x_Vm_to_invcm())
    return h0, [mu, als, alp], bas


def water_fields(t_fs, fwhm_fs=100.0, intensity=5e12):
    """Per-tensor field arrays at times t_fs: static Z on the dipole and the
    static polarizability term, the X pulse envelope (centred at 4 FWHM, cut
    beyond 4 FWHM) on the pulse polarizability term."""
    t = np.asarray(t_fs, dtype=np.float64)
    e0 = units.intensity_to_field(intensity)
    t0 = 4.0 * fwhm_fs
    ex = e0 * np.exp(-2 * np.log(2) * (t - t0) ** 2 / fwhm_fs ** 2) / np.sqrt(2.0)
    ex[np.abs(t - t0) > 4.0 * fwhm_fs] = 0.0
    stat = np.zeros((len(t), 3))
    stat[:, 2] = WATER_STATIC
    pul = np.zeros((len(t), 3))
    pul[:, 0] = ex
    return [stat, stat, pul]


def gaussian_pulse_field(t_fs, fwhm_fs=300.0, intensity=5e11):
    """Cycle-averaged Z field of a Gaussian pulse: E(t) = E0 exp(-2 ln2 (t-t0)^2/tau^2)/sqrt(2),
    tau = intensity FWHM, t0 = 2.5 tau / 2, E0 from the peak intensity [W/cm^2]."""
    e0 = units.intensity_to_field(intensity)
    t0 = 2.5 * fwhm_fs / 2
    return e0 * np.exp(-2 * np.log(2) * (t_fs - t0) ** 2 / fwhm_fs ** 2) / np.sqrt(2.0)


# Large asymmetric top (the Kronecker-block configuration): H2S-like rotor,
# A, B, C = 10.36, 9.02, 4.73 cm^-1, frame x = c, y = a, z = b (the C2 axis),
# dipole 0.3816 au (0.97 D) along z, and a full polarizability tensor (au;
# the off-diagonal elements make every Cartesian component act).  The field
# is circularly polarised in the XY plane with a linearly chirped frequency,
# 0 -> 10 THz over 100 ps: E(t) = E0 (cos phi, sin phi, 0), phi = pi r t^2,
# r = 1e23 Hz/s, E0 = 1e8 V/m, acting on both the dipole and the
# polarizability (H = -mu.E - 1/2 E.alpha.E).
LARGE_B = (4.73, 10.36, 9.02)           # (Bx, By, Bz)
LARGE_MU = (0.0, 0.0, 0.3816)
LARGE_ALPHA = ((24.0, 0.6, -0.4), (0.6, 25.5, 0.3), (-0.4, 0.3, 24.7))
LARGE_E0 = 1e8
LARGE_CHIRP = 1e23


def large_top_tensors():
    """The two interaction tensors and their energy scales (cm^-1 per field
    product): -mu in au x V/m, -1/2 alpha in au x (V/m)^2."""
    return ([np.array(LARGE_MU), np.array(LARGE_ALPHA)],
            [-units.AUdip_x_Vm_to_invcm(), -0.5 * units.AUpol_x_Vm_to_invcm()])


def large_top(jmax):
    """(KronOperator, basis) of the large asymmetric top with J <= jmax, all m."""
    from richmol_amd.kron import KronOperator
    bas = rot.asymtop_basis(*LARGE_B, jmax)
    tens, scales = large_top_tensors()
    return KronOperator(bas, tens, scales), bas


def centrifuge_fields(t_fs, e0=LARGE_E0, rate=LARGE_CHIRP):
    """Per-tensor (nsteps, 3) field arrays of the chirped circular field at
    times t_fs (the same field on the dipole and the polarizability)."""
    t = np.asarray(t_fs, dtype=np.float64) * 1e-15
    phi = np.pi * rate * t ** 2
    E = np.stack([e0 * np.cos(phi), e0 * np.sin(phi), np.zeros_like(t)], axis=1)
    return [E, E]
