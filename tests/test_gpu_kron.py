"""Kronecker-block operator on the device (rm_kplan: per output J one dense
product of the stacked K blocks with the m-mixed input of every wavepacket,
rocBLAS) and its fixed-order Lanczos propagator, against the host
factorisation and the FP64 oracle (tdse.py:365-486 restated) on the large
asymmetric top (dipole + full polarizability, chirped circular XY field) at
small J.  Tolerances: FP64 1e-10 absolute in every amplitude; FP32 products
2e-6 in amplitudes and 1e-6 in <cos^2 theta>."""
import numpy as np
import pytest

from _fixtures import centrifuge_fields, large_top, large_top_tensors

pytestmark = pytest.mark.gpu


def _dev_plan(kop):
    from richmol_amd.kron import DeviceKronPlan
    return DeviceKronPlan(kop)


def _to_int(kop, Xf):
    Xi = np.zeros_like(Xf)
    Xi[:, kop.perm] = Xf
    return Xi


@pytest.mark.parametrize("kernel", ["fused", "gemm", "mfma"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 2e-6)])
def test_matvec_vs_host_factorisation(device, precision, tol, kernel):
    """J <= 12 (2625 states; m blocks of up to 25 rows span two Z tiles), 5
    wavepackets, a general field on both tensors: the device matvec equals
    KronOperator.matvec (relative to the largest output)."""
    import torch
    from richmol_amd.kron import slot_products
    kop, bas = large_top(12)
    dp = _dev_plan(kop)
    dp.set_kernel(kernel)
    rng = np.random.default_rng(11)
    X = rng.normal(size=(5, kop.n)) + 1j * rng.normal(size=(5, kop.n))
    F = [np.array([[0.3e8, -0.8e8, 0.5e8]]), np.array([[0.6e8, 0.2e8, -0.9e8]])]
    fp = slot_products(kop, F)
    ref = kop.matvec(X, kop.field_products([F[0][0], F[1][0]]))
    x = torch.from_numpy(X).to(device)
    y = torch.zeros_like(x)
    dp.matvec(x, y, torch.from_numpy(fp[0]).to(device), precision=precision)
    got = y.cpu().numpy()
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()


def test_matvec_strided_rows(device):
    """Rows with a leading dimension larger than n; the padding is untouched."""
    import torch
    from richmol_amd.kron import slot_products
    kop, bas = large_top(3)
    dp = _dev_plan(kop)
    rng = np.random.default_rng(12)
    n = kop.n
    buf = torch.from_numpy(rng.normal(size=(2, n + 7)) + 1j * rng.normal(size=(2, n + 7))).to(device)
    out = torch.full((2, n + 7), 3.0 + 4.0j, dtype=torch.complex128, device=device)
    F = [np.array([[1e8, 2e7, 0.0]])] * 2
    fp = slot_products(kop, F)
    dp.matvec(buf[:, :n], out[:, :n], torch.from_numpy(fp[0]).to(device))
    ref = kop.matvec(buf[:, :n].cpu().numpy(), kop.field_products([F[0][0], F[1][0]]))
    assert np.abs(out[:, :n].cpu().numpy() - ref).max() <= 1e-13 * np.abs(ref).max()
    assert torch.all(out[:, n:] == 3.0 + 4.0j)


def test_fused_matvec_long_stacks(device):
    """J <= 24 (2 401 states; output blocks of up to 49 x 98 reals, stacks of
    up to nine source blocks, several row and column tiles per J): the fused
    kernel and the stored form's own matrix-core product equal the stored-Z
    rocBLAS form to rounding, and the host factorisation, in FP64; 3
    wavepackets."""
    import torch
    from richmol_amd.kron import slot_products
    kop, bas = large_top(24)
    dp = _dev_plan(kop)
    rng = np.random.default_rng(13)
    X = rng.normal(size=(3, kop.n)) + 1j * rng.normal(size=(3, kop.n))
    F = [np.array([[0.4e8, 0.9e8, -0.3e8]]), np.array([[-0.2e8, 0.5e8, 0.8e8]])]
    fp = torch.from_numpy(slot_products(kop, F)[0]).to(device)
    x = torch.from_numpy(X).to(device)
    out = {}
    for k in ("fused", "gemm", "mfma"):
        dp.set_kernel(k)
        y = torch.zeros_like(x)
        dp.matvec(x, y, fp)
        out[k] = y.cpu().numpy()
    ref = kop.matvec(X, kop.field_products([F[0][0], F[1][0]]))
    scale = np.abs(ref).max()
    assert np.abs(out["fused"] - out["gemm"]).max() <= 1e-13 * scale
    assert np.abs(out["mfma"] - out["gemm"]).max() <= 1e-13 * scale
    assert np.abs(out["fused"] - ref).max() <= 1e-13 * scale
    assert np.abs(out["mfma"] - ref).max() <= 1e-13 * scale


def _oracle_run(bas, fields, Xf, fixed, dt):
    from oracle.cartens import OracleCarTens
    from oracle.tdse import OracleTDSE
    from richmol_amd import rot
    tens, scales = large_top_tensors()
    OL = [OracleCarTens.from_attrs(rot.tensor(t, bas) * s) for t, s in zip(tens, scales)]
    Oh0 = OracleCarTens.from_attrs(rot.hamiltonian(bas))
    nst = len(fields[0])
    to = OracleTDSE(t_start=0.0, t_end=dt * nst, dt=dt, t_units="fs")
    to.time_grid()
    psi = Xf.copy()
    for i in range(nst):
        for O, F in zip(OL, fields):
            O.field(list(F[i]))
        psi, _ = to.update(OL[0] + OL[1], psi, H0=Oh0, fixed_m=fixed)
    return psi, to.exp_fac


def _device_run(kop, fields, Xf, fixed, f, precision="fp64", kernel="gemm"):
    import torch
    from richmol_amd.kron import slot_products
    dp = _dev_plan(kop)
    dp.set_kernel(kernel)
    psi = torch.from_numpy(_to_int(kop, Xf)).cuda()
    ph = dp.phases(f)
    dp.propagate(psi, slot_products(kop, fields), f, phase=ph, krylov_fixed=fixed, precision=precision)
    return psi.cpu().numpy()[:, kop.perm]


def _columns(n, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((4, n), dtype=np.complex128)
    X[0, 0] = 1.0                                  # ground state
    X[1, 5] = 1.0                                  # an excited basis state
    X[2] = rng.normal(size=n) + 1j * rng.normal(size=n)
    X[2] /= np.linalg.norm(X[2])
    X[3] = 0.4 * (rng.normal(size=n) + 1j * rng.normal(size=n)) / np.sqrt(n)   # norm != 1
    return X


@pytest.mark.parametrize("kernel", ["gemm", "fused", "mfma"])
@pytest.mark.parametrize("fixed", [12, 5])
def test_propagate_matches_oracle(device, fixed, kernel):
    """J <= 5 (286 states), four wavepackets (basis states, dense, unnormalised),
    eight 10 fs steps of the chirped circular field starting late in the
    sweep (several THz), plus two zero-field steps (phases only): FP64 within
    1e-10 of the oracle."""
    kop, bas = large_top(5)
    dt = 10.0
    t = 40000.0 + (np.arange(8) + 0.5) * dt
    fields = centrifuge_fields(t, e0=3e9)
    for F in fields:
        F[3] = 0.0
        F[6] = 0.0
    X = _columns(kop.n, 3)
    ref, f = _oracle_run(bas, fields, X, fixed, dt)
    got = _device_run(kop, fields, X, fixed, f, kernel=kernel)
    assert np.abs(got - ref).max() <= 1e-10
    if fixed == 12:   # converged expv: the norms are kept (unitary step)
        assert np.abs(np.linalg.norm(got, axis=1) - np.linalg.norm(X, axis=1)).max() <= 1e-12


@pytest.mark.parametrize("kernel", ["gemm", "fused", "mfma"])
def test_propagate_fp32_products(device, kernel):
    """FP32 dense products (complex64 operands and output; recurrence, norms,
    exp(fT), phases and wavepackets FP64): amplitudes within 2e-6 and
    <cos^2 theta> within 1e-6 of the FP64 oracle over 12 steps at J <= 6."""
    from oracle.cartens import OracleCarTens
    from richmol_amd import rot
    kop, bas = large_top(6)
    dt = 10.0
    t = 20000.0 + (np.arange(12) + 0.5) * dt
    fields = centrifuge_fields(t, e0=3e9)
    X = _columns(kop.n, 5)[:3]
    ref, f = _oracle_run(bas, fields, X, 12, dt)
    got = _device_run(kop, fields, X, 12, f, precision="fp32", kernel=kernel)
    C = OracleCarTens.from_attrs(rot.named("cos2theta", bas)).tomat(form="full", cart="0")
    c2 = lambda P: np.real(np.einsum("bi,bi->b", P.conj(), (C @ P.T).T))
    assert np.abs(got - ref).max() <= 2e-6
    assert np.abs(c2(got) - c2(ref)).max() <= 1e-6


def test_propagate_exact_breakdown(device):
    """Isotropic polarizability only: the operator is a multiple of the
    identity, so a basis-state column breaks down exactly after the first
    Krylov vector (the reference restarts, tdse.py:459-465; here a zero vector
    with zero coefficient); a dense column beside it is unaffected."""
    import torch
    from oracle.cartens import OracleCarTens
    from oracle.tdse import OracleTDSE
    from richmol_amd import rot, units
    from richmol_amd.kron import DeviceKronPlan, KronOperator, slot_products
    from _fixtures import LARGE_B
    bas = rot.asymtop_basis(*LARGE_B, 3)
    a = np.diag([20.0, 20.0, 20.0])
    s = -0.5 * units.AUpol_x_Vm_to_invcm()
    kop = KronOperator(bas, [a], [s])
    dt = 10.0
    E = np.tile([[2e9, 1e9, 0.5e9]], (3, 1))
    X = _columns(kop.n, 7)[:3]
    O = OracleCarTens.from_attrs(rot.tensor(a, bas) * s)
    Oh0 = OracleCarTens.from_attrs(rot.hamiltonian(bas))
    to = OracleTDSE(t_start=0.0, t_end=dt * 3, dt=dt, t_units="fs")
    to.time_grid()
    ref = X.copy()
    for i in range(3):
        O.field(list(E[i]))
        ref, _ = to.update(O, ref, H0=Oh0, fixed_m=6)
    dp = DeviceKronPlan(kop)
    psi = torch.from_numpy(_to_int(kop, X)).cuda()
    dp.propagate(psi, slot_products(kop, [E]), to.exp_fac, phase=dp.phases(to.exp_fac), krylov_fixed=6)
    got = psi.cpu().numpy()[:, kop.perm]
    assert np.abs(got - ref).max() <= 1e-10


def test_propagate_rejects_bad_arguments(device):
    import torch
    from richmol_amd import _lib
    kop, bas = large_top(2)
    dp = _dev_plan(kop)
    psi = torch.zeros((1, kop.n), dtype=torch.complex128, device=device)
    fp = np.zeros((1, dp.nslot), dtype=np.complex128)
    with pytest.raises(_lib.RichmolAmdError):
        dp.propagate(psi, fp, 1j, krylov_fixed=-1)
    with pytest.raises(_lib.RichmolAmdError):
        dp.propagate(psi, fp, 1j, krylov_fixed=_lib.RM_MAX_KRYLOV + 1)
    with pytest.raises(_lib.RichmolAmdError):   # adaptive rule: maxorder in [1, RM_MAX_KRYLOV], tol > 0
        dp.propagate(psi, fp, 1j, krylov_fixed=0, maxorder=0)
    with pytest.raises(_lib.RichmolAmdError):
        dp.propagate(psi, fp, 1j, krylov_fixed=0, tol=0.0)


def test_full_size_large_top_properties(device):
    """The large configuration itself (J <= 80 with all m, N = 708 561), two
    wavepackets (ground state and a dense random row), where the oracle
    cannot run: the device operator is Hermitian (<x|Hy> = conj<y|Hx>), two
    steps keep the norms (1e-12), the fused and stored matvec forms give the
    same wavepackets (1e-11), and FP32 products stay within 1e-6 of FP64."""
    import torch
    from richmol_amd.kron import slot_products
    kop, bas = large_top(80)
    assert kop.n == 708561
    dp = _dev_plan(kop)
    n = kop.n
    rng = np.random.default_rng(17)
    F = centrifuge_fields(np.array([50005.0, 50015.0]))
    fp = slot_products(kop, F)
    fpd = torch.from_numpy(fp[0]).to(device)
    xy = torch.from_numpy(rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))).to(device)
    h = torch.zeros_like(xy)
    dp.matvec(xy, h, fpd)
    a = torch.vdot(xy[0], h[1]).item()
    b = torch.vdot(xy[1], h[0]).item()
    scale = float(torch.linalg.vector_norm(xy[0]) * torch.linalg.vector_norm(h[1]))
    assert abs(a - np.conj(b)) <= 1e-12 * scale
    X = np.zeros((2, n), dtype=np.complex128)
    X[0, int(np.argmin(kop.energies))] = 1.0
    X[1] = rng.normal(size=n) + 1j * rng.normal(size=n)
    X[1] /= np.linalg.norm(X[1])
    from richmol_amd.tdse import TDSE
    f = TDSE(t_end=20.0, dt=10.0, t_units="fs").exp_fac
    ph = dp.phases(f)
    out = {}
    for kern, prec in (("gemm", "fp64"), ("fused", "fp64"), ("gemm", "fp32")):
        dp.set_kernel(kern)
        psi = torch.from_numpy(X).to(device)
        dp.propagate(psi, fp, f, phase=ph, krylov_fixed=12, precision=prec)
        out[(kern, prec)] = psi.cpu().numpy()
    ref = out[("gemm", "fp64")]
    assert np.abs(np.linalg.norm(ref, axis=1) - 1.0).max() <= 1e-12
    assert np.abs(out[("fused", "fp64")] - ref).max() <= 1e-11
    assert np.abs(out[("gemm", "fp32")] - ref).max() <= 1e-6


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 2e-6)])
def test_tdse_update_kronecker_form(device, precision, tol):
    """The drop-in call: richmol's TDSE.update(H, vecs, H0=...) with H the sum
    of two field-multiplied CarTens (dipole, full polarizability) routed to
    the Kronecker-block plan (operator_form='kron'; 'auto' does so from
    150 000 states): rows in richmol's order in and out, four steps against
    the oracle's update on the same tensors."""
    from richmol_amd import rot
    from richmol_amd.tdse import TDSE
    kop, bas = large_top(5)
    tens, scales = large_top_tensors()
    T = [rot.tensor(t, bas) * s for t, s in zip(tens, scales)]
    h0 = rot.hamiltonian(bas)
    dt = 10.0
    t = 40000.0 + (np.arange(4) + 0.5) * dt
    fields = centrifuge_fields(t, e0=3e9)
    X = _columns(kop.n, 23)[:3]
    ref, f = _oracle_run(bas, fields, X, 12, dt)
    td = TDSE(t_start=0.0, t_end=dt * 4, dt=dt, t_units="fs")
    td.time_grid()
    psi = X.copy()
    for i in range(4):
        for Ti, F in zip(T, fields):
            Ti.field(list(F[i]))
        psi, _ = td.update(T[0] + T[1], psi, H0=h0, krylov_fixed=12, operator_form="kron", precision=precision)
    assert np.abs(psi - ref).max() <= tol


def _oracle_run_adaptive(bas, fields, Xf, dt, maxorder=100, tol=1e-15):
    from oracle.cartens import OracleCarTens
    from oracle.tdse import OracleTDSE
    from richmol_amd import rot
    tens, scales = large_top_tensors()
    OL = [OracleCarTens.from_attrs(rot.tensor(t, bas) * s) for t, s in zip(tens, scales)]
    Oh0 = OracleCarTens.from_attrs(rot.hamiltonian(bas))
    nst = len(fields[0])
    to = OracleTDSE(t_start=0.0, t_end=dt * nst, dt=dt, t_units="fs")
    to.time_grid()
    psi, ks = Xf.copy(), []
    for i in range(nst):
        for O, F in zip(OL, fields):
            O.field(list(F[i]))
        stats = []
        if np.any([np.any(F[i]) for F in fields]):
            psi, _ = to.update(OL[0] + OL[1], psi, H0=Oh0, tol=tol, maxorder=maxorder, stats=stats)
        else:
            psi, _ = to.update(OL[0] + OL[1], psi, H0=Oh0, tol=tol, maxorder=maxorder)
        ks.append(stats if stats else [0] * len(Xf))
    return psi, np.array(ks), to.exp_fac


@pytest.mark.parametrize("kernel", ["gemm", "fused"])
def test_propagate_adaptive_matches_oracle(device, kernel):
    """The reference's adaptive rule (tol 1e-15, maxorder 100) on the
    Kronecker-block path: four wavepackets at J <= 5 over six steps of the
    chirped field (one zero-field step): amplitudes within 1e-10 of the oracle
    and every step's Krylov dimension per wavepacket equal to the oracle's."""
    import torch
    from richmol_amd.kron import slot_products
    kop, bas = large_top(5)
    dt = 10.0
    t = 40000.0 + (np.arange(6) + 0.5) * dt
    fields = centrifuge_fields(t, e0=3e9)
    for F in fields:
        F[2] = 0.0
    X = _columns(kop.n, 29)
    ref, k_o, f = _oracle_run_adaptive(bas, fields, X, dt)
    dp = _dev_plan(kop)
    dp.set_kernel(kernel)
    psi = torch.from_numpy(_to_int(kop, X)).cuda()
    kused = torch.zeros((6, 4), dtype=torch.int32, device=device)
    status = torch.ones(4, dtype=torch.int32, device=device)
    dp.propagate(psi, slot_products(kop, fields), f, phase=dp.phases(f), krylov_fixed=0, kused=kused, status=status)
    got = psi.cpu().numpy()[:, kop.perm]
    assert np.abs(got - ref).max() <= 1e-10
    assert np.array_equal(kused.cpu().numpy(), k_o)
    assert int(status.abs().sum()) == 0


def test_propagate_adaptive_maxorder_reached(device):
    """A maxorder too small for tol 1e-15: the steps end at maxorder and
    status flags every wavepacket (the reference raises ValueError), and
    TDSE.update on the Kronecker form raises it."""
    import torch
    from richmol_amd import rot
    from richmol_amd.kron import slot_products
    from richmol_amd.tdse import TDSE
    kop, bas = large_top(4)
    F = centrifuge_fields(np.array([40005.0]), e0=3e9)
    dp = _dev_plan(kop)
    X = _columns(kop.n, 31)
    psi = torch.from_numpy(_to_int(kop, X)).cuda()
    kused = torch.zeros((1, 4), dtype=torch.int32, device=device)
    status = torch.zeros(4, dtype=torch.int32, device=device)
    dp.propagate(psi, slot_products(kop, F), -1e-2j, krylov_fixed=0, maxorder=3, kused=kused, status=status)
    assert np.all(kused.cpu().numpy() == 3)
    assert np.all(status.cpu().numpy() == 1)
    tens, scales = large_top_tensors()
    T = [rot.tensor(t, bas) * s for t, s in zip(tens, scales)]
    for Ti, Fi in zip(T, F):
        Ti.field(list(Fi[0]))
    td = TDSE(t_start=0.0, t_end=10.0, dt=10.0, t_units="fs")
    td.time_grid()
    with pytest.raises(ValueError):
        td.update(T[0] + T[1], X, H0=rot.hamiltonian(bas), operator_form="kron", maxorder=3)


@pytest.mark.parametrize("kernel", ["gemm", "fused"])
@pytest.mark.parametrize("fixed", [12, 0])
def test_propagate_observable_every_step(device, fixed, kernel):
    """rm_kplan_propagate_expect: <cos^2 theta> (molecular z on lab Z, the
    Kronecker-block operator of the field [0, 0, 1]) of every wavepacket
    after every step of the oracle test's run (two field-free steps among
    them), fixed dimension 12 and the adaptive rule, within 1e-10 of the
    oracle's states under rot.named('cos2theta') (cos^2 theta - 1/3) plus
    |psi|^2 / 3; the sums equal the per-wavepacket values; the final rows
    equal those of the call without an observable."""
    import torch
    from oracle.cartens import OracleCarTens
    from oracle.tdse import OracleTDSE
    from richmol_amd import rot
    from richmol_amd.kron import KronOperator, slot_products
    kop, bas = large_top(5)
    dt = 10.0
    t = 40000.0 + (np.arange(8) + 0.5) * dt
    fields = centrifuge_fields(t, e0=3e9)
    for F in fields:
        F[3] = 0.0
        F[6] = 0.0
    X = _columns(kop.n, 3)
    # oracle states after every step
    tens, scales = large_top_tensors()
    OL = [OracleCarTens.from_attrs(rot.tensor(q, bas) * s) for q, s in zip(tens, scales)]
    Oh0 = OracleCarTens.from_attrs(rot.hamiltonian(bas))
    to = OracleTDSE(t_start=0.0, t_end=dt * 8, dt=dt, t_units="fs")
    to.time_grid()
    C = OracleCarTens.from_attrs(rot.named("cos2theta", bas)).tomat(form="full", cart="0")
    psi, want = X.copy(), np.zeros((8, 4), dtype=np.complex128)
    for i in range(8):
        for O, F in zip(OL, fields):
            O.field(list(F[i]))
        psi, _ = to.update(OL[0] + OL[1], psi, H0=Oh0, fixed_m=fixed or None)
        want[i] = np.einsum("bi,bi->b", psi.conj(), (C @ psi.T).T) + np.sum(np.abs(psi) ** 2, axis=1) / 3.0
    dp = _dev_plan(kop)
    dp.set_kernel(kernel)
    kc = KronOperator(bas, [np.diag([0.0, 0.0, 1.0])], [1.0])
    dc = _dev_plan(kc)
    cfp = slot_products(kc, [np.array([[0.0, 0.0, 1.0]])])[0]
    f = to.exp_fac
    ph = dp.phases(f)
    fp = slot_products(kop, fields)
    x = torch.from_numpy(_to_int(kop, X)).cuda()
    x2 = x.clone()
    ex = torch.zeros((8, 4), dtype=torch.complex128, device=device)
    exs = torch.zeros(8, dtype=torch.complex128, device=device)
    dp.propagate(x, fp, f, phase=ph, krylov_fixed=fixed, obs=dc, obs_fprod=cfp, expect=ex, expect_sum=exs)
    dp.propagate(x2, fp, f, phase=ph, krylov_fixed=fixed)
    got = ex.cpu().numpy()
    assert np.abs(got - want).max() <= 1e-10
    assert np.abs(exs.cpu().numpy() - got.sum(axis=1)).max() <= 1e-14
    assert torch.equal(x, x2)
    assert np.abs(x.cpu().numpy()[:, kop.perm] - psi).max() <= 1e-10


@pytest.mark.parametrize("fixed", [12, 0])
def test_tdse_run_kron_equals_flat(device, fixed):
    """TDSE.run on the Kronecker-block form (operator_form='kron') equals the
    flattened form's run -- final rows and the per-step ensemble
    <cos^2 theta - 1/3> -- for the polarizability of the large top at J <= 5
    under a pulsed X+Z field (field-free steps at the start)."""
    from richmol_amd import rot
    from richmol_amd.tdse import TDSE
    kop, bas = large_top(5)
    tens, scales = large_top_tensors()
    H = rot.tensor(tens[1], bas) * scales[1]
    h0 = rot.hamiltonian(bas)
    C = rot.named("cos2theta", bas)
    X = _columns(kop.n, 7)[:3]
    field = lambda t: [0.5e9 * np.exp(-((t - 60.0) / 30.0) ** 2), 0.0,
                       1.5e9 * np.exp(-((t - 60.0) / 30.0) ** 2) * (t > 20.0)]
    out = {}
    for form in ("kron", "flat"):
        td = TDSE(t_start=0.0, t_end=120.0, dt=10.0, t_units="fs")
        td.time_grid()
        out[form] = td.run(H, X, field, H0=h0, obs=C, krylov_fixed=fixed, operator_form=form)
    (vk, tk, ek), (vf, tf, ef) = out["kron"], out["flat"]
    assert tk == tf == 120.0
    assert np.abs(vk - vf).max() <= 1e-10
    assert np.abs(ek - ef).max() <= 1e-10
    assert np.abs(ek).max() > 1e-3
