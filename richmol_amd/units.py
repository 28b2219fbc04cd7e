"""Physical constants and unit factors used to build exponents and operators.

Mirrors the factor functions of richmol/convert_units.py:5-123 and the
time/energy unit tables of richmol/tdse.py:111-143.  The constants are frozen
CODATA 2022 literals (the values scipy 1.15.3 reports) so that every
propagation uses bit-identical factors no matter which scipy is installed.
"""

PLANCK = 6.62607015e-34                # J s (exact)
HBAR = 1.0545718176461565e-34          # J s, h/(2 pi)
C_LIGHT = 299792458.0                  # m/s (exact)
BOLTZMANN = 1.380649e-23               # J/K (exact)
AU_POL = 1.64877727212e-41             # C^2 m^2 / J
AU_DIP = 8.4783536198e-30              # C m
AU_TIME = 2.4188843265864e-17          # s
ELEM_CHARGE = 1.602176634e-19          # C (exact)
BOHR = 5.29177210544e-11               # m
EPS0 = 8.8541878188e-12                # F/m


def _convert(fac, *args):
    """richmol/convert_units.py:113-123"""
    if len(args) == 0:
        return fac
    res = []
    for arg in args:
        try:
            res.append([elem * fac for elem in arg])
        except TypeError:
            res.append(arg * fac)
    return tuple(res)


def J_to_invcm(*args):
    return _convert(1.0 / (PLANCK * 1e2 * C_LIGHT), *args)


def MHz_to_invcm(*args):
    return _convert(1.0 / C_LIGHT * 1e4, *args)


def Debye_to_au(*args):
    return _convert(1e-21 / C_LIGHT / ELEM_CHARGE / BOHR, *args)


def AUpol_x_Vm_to_invcm(*args):
    """polarizability [au] x field^2 [V^2/m^2] -> cm^-1"""
    return _convert(AU_POL / (PLANCK * C_LIGHT) / 1e2, *args)


def AUdip_x_Vm_to_invcm(*args):
    """dipole [au] x field [V/m] -> cm^-1"""
    return _convert(AU_DIP / (PLANCK * C_LIGHT) / 1e2, *args)


def Debye_x_Vm_to_invcm(*args):
    return _convert(AU_DIP / (PLANCK * C_LIGHT) / 1e2 * Debye_to_au(), *args)


T_UNITS = {"fs": 1e-15, "ps": 1e-12, "ns": 1e-9, "au": AU_TIME}


def enr_units():
    return {"invcm": 1.0 / J_to_invcm(), "mhz": MHz_to_invcm() / J_to_invcm()}


def intensity_to_field(intensity_w_cm2):
    """Peak field amplitude E0 [V/m] of a light wave of peak intensity I [W/cm^2]:
    I = c eps0 E0^2 / 2."""
    return (2.0 * intensity_w_cm2 * 1e4 / (C_LIGHT * EPS0)) ** 0.5
