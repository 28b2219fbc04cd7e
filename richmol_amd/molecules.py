"""The molecules and fields of the propagation workloads, as operators.

* ``linear_rotor``: OCS-like linear rotor, B = 0.2029 cm^-1,
  alpha = diag(26.0, 26.0, 53.6) au (alpha_perp = 26.0, Delta alpha = 27.6).
* ``water``: water-like asymmetric top, A, B, C = 27.88, 14.51, 9.28 cm^-1,
  molecule frame x = c, y = a, z = b (the C2 axis, along which the dipole
  lies -- the frame of richmol's water example, examples/h2o_stark_rigrot.py:
  19-37), dipole (0, 0, -0.7288) au, polarizability diag(9.14, 9.87, 9.45) au;
  a static 50 kV/cm Z field plus a cycle-averaged X pulse.
* ``large_top``: H2S-like asymmetric top, A, B, C = 10.36, 9.02, 4.73 cm^-1,
  frame x = c, y = a, z = b, dipole 0.3816 au along z and a full
  polarizability; a circularly polarised XY field chirped 0 -> 10 THz over
  100 ps, E0 = 1e8 V/m, acting on both tensors.

Interaction operators carry the energy factors of richmol's drivers:
-mu.F (AUdip_x_Vm_to_invcm) and -1/2 F.alpha.F (AUpol_x_Vm_to_invcm).
"""
import numpy as np

from . import pulses, rot, units

LINEAR_B = 0.2029
LINEAR_ALPHA = (26.0, 26.0, 53.6)

WATER_B = (9.28, 27.88, 14.51)          # (Bx, By, Bz)
WATER_MU = (0.0, 0.0, -0.7288)
WATER_ALPHA = (9.14, 9.87, 9.45)
WATER_STATIC = 5e6                      # 50 kV/cm along Z [V/m]

LARGE_B = (4.73, 10.36, 9.02)           # (Bx, By, Bz)
LARGE_MU = (0.0, 0.0, 0.3816)
LARGE_ALPHA = ((24.0, 0.6, -0.4), (0.6, 25.5, 0.3), (-0.4, 0.3, 24.7))
LARGE_E0 = 1e8
LARGE_CHIRP = 1e23


def linear_rotor(Jmax, B=LINEAR_B, alpha=LINEAR_ALPHA, m_filter=None):
    """(H0, -1/2 alpha AUpol, cos^2 theta - 1/3 operator, basis)."""
    bas = rot.linear_basis(B, Jmax, m_filter=m_filter)
    h0 = rot.hamiltonian(bas)
    H = rot.tensor(np.diag(alpha), bas) * (-0.5 * units.AUpol_x_Vm_to_invcm())
    cos2 = rot.named("cos2theta", bas)
    return h0, H, cos2, bas


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
    """Per-tensor field arrays of ``water``'s three tensors at times t_fs:
    static Z on the dipole and the static polarizability term, the X pulse
    envelope (centred at 4 FWHM, zero beyond 4 FWHM of the centre) on the
    pulse polarizability term."""
    t = np.asarray(t_fs, dtype=np.float64)
    t0 = 4.0 * fwhm_fs
    ex = pulses.envelope(t, fwhm_fs, intensity, t0_fs=t0, floor=0.0)
    ex[np.abs(t - t0) > 4.0 * fwhm_fs] = 0.0
    stat = pulses.static(len(t), WATER_STATIC)
    return [stat, stat, pulses.polarised(ex, (1.0, 0.0, 0.0))]


def large_top_tensors():
    """The two interaction tensors of ``large_top`` and their energy scales
    (cm^-1 per field product): -mu in au x V/m, -1/2 alpha in au x (V/m)^2."""
    return ([np.array(LARGE_MU), np.array(LARGE_ALPHA)],
            [-units.AUdip_x_Vm_to_invcm(), -0.5 * units.AUpol_x_Vm_to_invcm()])


def large_top(jmax):
    """(KronOperator, basis) of the large asymmetric top, J <= jmax, all m."""
    from .kron import KronOperator
    bas = rot.asymtop_basis(*LARGE_B, jmax)
    tens, scales = large_top_tensors()
    return KronOperator(bas, tens, scales), bas


def centrifuge_fields(t_fs, e0=LARGE_E0, rate=LARGE_CHIRP):
    """Per-tensor (nsteps, 3) field arrays of the chirped circular field at
    times t_fs (the same field on the dipole and the polarizability)."""
    E = pulses.chirped_circular(t_fs, e0, rate)
    return [E, E]
