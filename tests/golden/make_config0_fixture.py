"""Oracle trajectories of the linear-rotor configuration (J <= 30 with all m,
N = 961; psi0 = |J=0, m=0>; cycle-averaged Gaussian Z pulse, intensity FWHM
300 fs, peak 5e11 W/cm^2, centred at t0 = 2.5 FWHM / 2; dt = 10 fs, 1000
steps; split operator with the field-free phases), written to
config0_oracle.npz: one run with a fixed Krylov dimension 12 and one with the
reference's adaptive rule (tol 1e-15, maxorder 100).

Per step: the wavepacket on the 16 states it can reach (m = 0, even J; every
other amplitude is checked to stay exactly zero), <cos^2 theta>, and the
Krylov dimension used.  The field is evaluated at the step centres, as a
richmol time loop does (tdse.py:146-174, examples/ocs_alignment.py:89-100).

    python tests/golden/make_config0_fixture.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

NSTEPS = 1000


def run(fixed_m, nsteps=NSTEPS):
    from _fixtures import gaussian_pulse_field, linear_rotor
    from oracle.cartens import OracleCarTens, flatten, unflatten
    from oracle.tdse import OracleTDSE
    h0, H, cos2, bas = linear_rotor(30)
    Oh0, OH, OC = (OracleCarTens.from_attrs(T) for T in (h0, H, cos2))
    OC.field([0.0, 0.0, 1.0])
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    idx = reachable(bas)
    td = OracleTDSE(t_end=10.0 * (nsteps + 1), dt=10.0, t_units="fs")
    tc = td.time_grid()
    v = np.zeros((1, n), dtype=np.complex128)
    v[0, 0] = 1.0
    psi = np.zeros((nsteps, len(idx)), dtype=np.complex128)
    c2 = np.zeros(nsteps)
    kused = np.zeros(nsteps, dtype=np.int32)
    for i in range(nsteps):
        OH.field([0.0, 0.0, gaussian_pulse_field(tc[i])])
        st = []
        v, t = td.update(OH, v, H0=Oh0, fixed_m=fixed_m, stats=st)
        rest = np.delete(v[0], idx)
        assert not np.any(rest), "amplitude outside the m = 0, even-J states"
        psi[i] = v[0, idx]
        c2[i] = np.vdot(v[0], flatten(OC, OC.vec(unflatten(OC, v[0])))).real + 1.0 / 3.0
        kused[i] = st[0] if st else 0   # no field: phases only, no Krylov space
    return psi, c2, kused


def reachable(bas):
    """Basis indices (full-form order) of the m = 0, even-J states."""
    out, pos = [], 0
    for J in bas.Jlist:
        for m in bas.mlist[J]:
            if m == 0 and J % 2 == 0:
                out.append(pos)
            pos += 1
    return np.array(out)


def main():
    from _fixtures import linear_rotor
    _, _, _, bas = linear_rotor(30)
    pf, cf, kf = run(12)
    pa, ca, ka = run(None)
    np.savez_compressed(os.path.join(HERE, "config0_oracle.npz"), idx=reachable(bas),
                        psi_fixed=pf, cos2_fixed=cf, kused_fixed=kf,
                        psi_adaptive=pa, cos2_adaptive=ca, kused_adaptive=ka)
    print("cos2 max", cf.max(), "adaptive Krylov dims", np.bincount(ka))


if __name__ == "__main__":
    main()
