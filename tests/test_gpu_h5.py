"""Device propagation of tensors read from an HDF5 file that richmol wrote
(tests/golden/water.h5: dipole, polarizability and field-free Hamiltonian of
water, J <= 5, stored by the reference's documentation notebook) against the
FP64 oracle on the same tensors: the drop-in loop of that notebook
(filter to one m, init_state, H.field, TDSE.update) and the full basis under a
static Z field on the dipole plus an X pulse on the polarizability.  Every
amplitude of every step to 1e-10.
"""
import os

import numpy as np
import pytest

from oracle.cartens import OracleCarTens
from oracle.tdse import OracleTDSE

pytestmark = pytest.mark.gpu

TOL_PSI = 1e-10
WATER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "water.h5")


def _pulse(t_fs, amp, t0=150.0, fwhm=100.0):
    return amp * np.exp(-4 * np.log(2) * ((t_fs - t0) / fwhm) ** 2)


@pytest.mark.parametrize("m", [0, -2])
def test_notebook_loop_one_m(device, m):
    import torch
    from richmol_amd import units
    from richmol_amd.field import CarTens, filter
    from richmol_amd.tdse import TDSE
    flt = lambda **kw: kw["m"] == m if "m" in kw else True   # noqa: E731
    h0 = filter(CarTens(WATER, name="h0"), bra=flt, ket=flt)
    pol = filter(CarTens(WATER, name="pol"), bra=flt, ket=flt)
    H = pol * (-0.5 * units.AUpol_x_Vm_to_invcm())
    nst, dt = 30, 10.0
    td = TDSE(t_end=nst * dt, dt=dt, t_units="fs")
    to = OracleTDSE(t_end=nst * dt, dt=dt, t_units="fs")
    to.time_grid()
    Oh0, OH = OracleCarTens.from_attrs(h0), OracleCarTens.from_attrs(pol) * (-0.5 * units.AUpol_x_Vm_to_invcm())
    vd = torch.from_numpy(td.init_state(h0, temp=0)).to(device)
    vo = to.init_state(Oh0, temp=0)
    worst = 0.0
    for i, t in enumerate(td.time_grid()):
        F = [0.0, 0.0, _pulse(t, 6e9)]
        H.field(F)
        OH.field(F)
        vd, t_d = td.update(H, vd, H0=h0)
        vo, t_o = to.update(OH, vo, H0=Oh0)
        assert t_d == t_o
        worst = max(worst, np.abs(vd.cpu().numpy() - vo).max())
        if i + 1 == nst:
            break
    assert worst <= TOL_PSI, worst
    # the pulse did something: the ground state is no longer pure
    assert np.abs(vo[0]).max() < 0.9999


def test_full_basis_dc_plus_pulse(device):
    import torch
    from richmol_amd import units
    from richmol_amd.field import CarTens
    from richmol_amd.tdse import TDSE
    h0, dip, pol = (CarTens(WATER, name=n) for n in ("h0", "dip", "pol"))
    D = dip * (-units.AUdip_x_Vm_to_invcm())
    P = pol * (-0.5 * units.AUpol_x_Vm_to_invcm())
    OD = OracleCarTens.from_attrs(dip) * (-units.AUdip_x_Vm_to_invcm())
    OP = OracleCarTens.from_attrs(pol) * (-0.5 * units.AUpol_x_Vm_to_invcm())
    Oh0 = OracleCarTens.from_attrs(h0)
    n = h0.tomat(form="full", cart="0").shape[0]
    rng = np.random.default_rng(11)
    X = np.zeros((3, n), dtype=np.complex128)
    X[0, 0] = 1.0
    X[1:] = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    X[1:] /= np.linalg.norm(X[1:], axis=1, keepdims=True)
    nst, dt = 16, 5.0
    td = TDSE(t_end=nst * dt, dt=dt, t_units="fs")
    to = OracleTDSE(t_end=nst * dt, dt=dt, t_units="fs")
    to.time_grid()
    vd, vo = torch.from_numpy(X.copy()).to(device), X.copy()
    worst = 0.0
    for i, t in enumerate(td.time_grid()):
        Fd, Fp = [0.0, 0.0, 3e7], [_pulse(t, 4e9, t0=40.0, fwhm=40.0), 0.0, 0.0]
        D.field(Fd)
        P.field(Fp)
        OD.field(Fd)
        OP.field(Fp)
        vd, _ = td.update(D + P, vd, H0=h0)
        vo, _ = to.update(OD + OP, vo, H0=Oh0)
        worst = max(worst, np.abs(vd.cpu().numpy() - vo).max())
        if i + 1 == nst:
            break
    assert worst <= TOL_PSI, worst
