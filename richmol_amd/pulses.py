"""Field generators for propagation runs: (nsteps, 3) arrays of lab-frame
field vectors [V/m] at the step centres, the input of ``CarTens.field`` /
``Ensemble.prepare`` (richmol takes the field per step from the caller,
examples/ocs_alignment.py:89-93).

Non-resonant pulses act through the polarizability as the cycle average of
E(t)^2: a linearly polarised pulse E0 g(t) cos(w t) then enters as the
envelope field E0 g(t) / sqrt(2) (``envelope``); ``carrier`` gives the
explicit oscillating field instead, for dipole couplings and resonant
pulses.  g(t) = exp(-2 ln2 (t - t0)^2 / tau^2) with tau the intensity FWHM.
"""
import numpy as np

from . import units

C_NM_FS = 299.792458   # speed of light in nm/fs

# Envelope factor below which a pulse is exactly zero.  At 1e-30 of the peak
# field the interaction is ~1e-60 of its peak value; further out the field
# products underflow, and the reference's fixed-order Lanczos then meets two
# exact breakdowns in a row whose Gram-Schmidt restart divides 0/0
# (tdse.py:459-465).  Zero-field steps are phase-only, as in the reference
# when the field leaves no MF block (tdse.py:377).
ENVELOPE_FLOOR = 1e-30


def gaussian_factor(t_fs, fwhm_fs, t0_fs=None, floor=ENVELOPE_FLOOR):
    """g(t) = exp(-2 ln2 (t - t0)^2 / fwhm^2), zero where g < floor;
    t0 defaults to 2.5 fwhm / 2."""
    t = np.asarray(t_fs, dtype=np.float64)
    t0 = 2.5 * fwhm_fs / 2 if t0_fs is None else t0_fs
    g = np.exp(-2.0 * np.log(2.0) * (t - t0) ** 2 / fwhm_fs ** 2)
    return np.where(g >= floor, g, 0.0)[()]   # a scalar for a scalar time


def envelope(t_fs, fwhm_fs, intensity, t0_fs=None, floor=ENVELOPE_FLOOR):
    """Cycle-averaged amplitude [V/m] of a linearly polarised Gaussian pulse
    of peak intensity ``intensity`` [W/cm^2]: E0 g(t) / sqrt(2)."""
    e0 = units.intensity_to_field(intensity)
    return e0 * gaussian_factor(t_fs, fwhm_fs, t0_fs, floor) / np.sqrt(2.0)


def carrier(t_fs, fwhm_fs, intensity, wavelength_nm, t0_fs=None, cep=0.0, floor=ENVELOPE_FLOOR):
    """Explicit oscillating field [V/m]: E0 g(t) cos(w (t - t0) + cep)."""
    t = np.asarray(t_fs, dtype=np.float64)
    t0 = 2.5 * fwhm_fs / 2 if t0_fs is None else t0_fs
    w = 2.0 * np.pi * C_NM_FS / wavelength_nm     # rad/fs
    e0 = units.intensity_to_field(intensity)
    return e0 * gaussian_factor(t, fwhm_fs, t0, floor) * np.cos(w * (t - t0) + cep)


def polarised(amplitude, axis=(0.0, 0.0, 1.0)):
    """(nsteps, 3) field vectors: amplitude(t) along a unit axis."""
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    return np.asarray(amplitude, dtype=np.float64)[:, None] * a[None, :]


def static(nsteps, amplitude, axis=(0.0, 0.0, 1.0)):
    """(nsteps, 3) constant field of the given amplitude [V/m] along axis."""
    return polarised(np.full(nsteps, float(amplitude)), axis)


def chirped_circular(t_fs, e0, rate, phase0=0.0):
    """Circularly polarised field in the XY plane whose frequency grows
    linearly in time ("optical centrifuge"): E0 (cos phi, sin phi, 0) with
    phi = pi rate t^2 + phase0 (rate in Hz/s; instantaneous frequency
    rate * t, e.g. rate = 1e23 Hz/s sweeps 0 -> 10 THz over 100 ps)."""
    t = np.asarray(t_fs, dtype=np.float64) * 1e-15
    phi = np.pi * rate * t ** 2 + phase0
    return np.stack([e0 * np.cos(phi), e0 * np.sin(phi), np.zeros_like(t)], axis=1)
