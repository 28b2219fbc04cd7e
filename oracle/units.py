"""Frozen physical constants and unit factors (oracle copy).

Values are CODATA 2022 as shipped by scipy 1.15.3 (``scipy.constants.value``),
frozen as literals so that the oracle and the device path use bit-identical
factors.  The factor definitions follow richmol/convert_units.py:5-110 and the
time/energy unit tables of richmol/tdse.py:111-143.
"""

PLANCK = 6.62607015e-34                # J s (exact)
HBAR = 1.0545718176461565e-34          # J s, h/(2 pi) as scipy returns it
C_LIGHT = 299792458.0                  # m/s (exact)
BOLTZMANN = 1.380649e-23               # J/K (exact)
AU_POL = 1.64877727212e-41             # C^2 m^2 / J (atomic unit of polarizability)
AU_DIP = 8.4783536198e-30              # C m (atomic unit of dipole moment)
AU_TIME = 2.4188843265864e-17          # s


def J_to_invcm():
    """richmol/convert_units.py:11-15"""
    return 1.0 / (PLANCK * 1e2 * C_LIGHT)


def MHz_to_invcm():
    """richmol/convert_units.py:5-8"""
    return 1.0 / C_LIGHT * 1e4


def AUpol_x_Vm_to_invcm():
    """richmol/convert_units.py:102-110"""
    return AU_POL / (PLANCK * C_LIGHT) / 1e2


def AUdip_x_Vm_to_invcm():
    """richmol/convert_units.py:83-91"""
    return AU_DIP / (PLANCK * C_LIGHT) / 1e2


T_UNITS = {"fs": 1e-15, "ps": 1e-12, "ns": 1e-9, "au": AU_TIME}   # tdse.py:111-116


def enr_units():
    """richmol/tdse.py:129-133"""
    return {"invcm": 1.0 / J_to_invcm(), "mhz": MHz_to_invcm() / J_to_invcm()}
