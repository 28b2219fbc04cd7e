"""Throughput of the field-driven rotational TDSE step on one MI355X.

Workload (the thermal-ensemble configuration the headline metric is quoted
on): linear rigid rotor, B = 0.2029 cm^-1, alpha_perp = 26.0 au,
Delta alpha = 27.6 au, J = 0..40 with all m (N = 1681 states); 512 initial
|J, m> states drawn with Boltzmann weights at T = 1 K (seeded); Gaussian
non-resonant Z pulse (cycle-averaged envelope, intensity FWHM 300 fs, peak
5e11 W/cm^2, centred at t0 = 2.5 FWHM / 2 = 375 fs, zero where the envelope
is below 1e-30 of its peak, i.e. from step 249 on), dt = 10 fs; split
operator with the field-free phases; Lanczos with a fixed Krylov dimension
12; complex128.  Steps 0..248 carry a field, so every timed step runs the
Krylov expv.

``value``: wavepacket-steps per second of the production path (kernel
'auto': for this Z-polarised field every wavepacket stays in its (m, J-parity)
invariant subspace, <= 21 states, and the whole run is one launch with two
wavepackets per wave), inputs resident in HBM.  Beside it: the same workload
through the full-vector ELL kernel (every wavepacket as a 1681-vector), the
FP32-matvec precision mode with its observable errors against FP64 at every
step of a window across the pulse peak, the plain single-wavepacket FP64
path (one TDSE.update call per step), the asymmetric-top, large-top and
serving legs, and the FP64 oracle on the host (one core, and all the cores
this process may use) as CPU baseline.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "wavepacket-steps/s at Jmax=40 (⟨cos²θ⟩ ≤1e-6 vs FP64) + p99 request latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FWHM_FS = 300.0
DT_FS = 10.0
T0_FS = 2.5 * FWHM_FS / 2
# the FP32 accuracy window: steps 0..75, the pulse peak (375 fs) in its middle
ACCURACY_STEPS = int(round(2 * T0_FS / DT_FS)) + 1
# HBM bytes per launch of the headline kernel from rocprofv3 FETCH_SIZE /
# WRITE_SIZE passes scaled by a 1 GiB copy (FETCH x2.00, WRITE x1.00), measured
# at two step counts: bytes = fixed + per_step * steps.  Key: (kernel, jmax,
# wavepackets, Krylov dimension).
TRAFFIC_MODEL = {
    ("pair", 40, 512, 12): (2.515e6, 13.2e3, "profiles/k_sub_pair_linear_j40_b512_m12.md"),
}
MEASURED_TRAFFIC_FULL = {
    ("ell", 40, 512, 12): (351.1e6, "profiles/k_lanczos_ell_linear_j40_b512_m12_pmc.md"),
}


def workload(jmax=40):
    from richmol_amd import molecules
    h0, H, cos2, bas = molecules.linear_rotor(jmax)
    return bas, h0, H, cos2


def pulse(t_fs):
    """Cycle-averaged Z field [V/m] of the workload's pulse at times t_fs."""
    from richmol_amd import pulses
    return pulses.envelope(t_fs, FWHM_FS, 5e11)


def pulse_steps():
    """Number of leading steps whose field is non-zero."""
    f = pulse((np.arange(2000) + 0.5) * DT_FS)
    return int(np.argmax(f == 0.0)) if np.any(f == 0.0) else len(f)


def exp_fac(dt_fs):
    from richmol_amd import units
    return -1j * dt_fs * 1e-15 * units.enr_units()["invcm"] / units.HBAR


def usable_cores():
    """Cores this process may run on (the affinity mask), at most 16 (a GPU
    box's share per GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def _oracle_run(args):
    """One process's share of the CPU baseline: the reference's time loop
    (examples/ocs_alignment.py:89-103) on the oracle -- TDSE.update (row
    loop, block matvec, Lanczos), then <cos^2 theta> of every row -- on its
    rows, one thread."""
    h0, H, C, rows, fields, krylov = args
    from threadpoolctl import threadpool_limits
    from oracle.cartens import OracleCarTens, flatten, unflatten
    from oracle.tdse import OracleTDSE
    Oh0, OH, OC = (OracleCarTens.from_attrs(T) for T in (h0, H, C))
    OC.field([0.0, 0.0, 1.0])
    td = OracleTDSE(t_end=DT_FS * (len(fields) + 1), dt=DT_FS, t_units="fs")
    td.time_grid()
    X = rows.copy()
    c2 = np.zeros(len(fields), dtype=np.complex128)
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        for i, F in enumerate(fields):
            OH.field(F)
            X, _ = td.update(OH, X, H0=Oh0, fixed_m=krylov)
            c2[i] = sum(np.vdot(x, flatten(OC, OC.vec(unflatten(OC, x)))) for x in X)
        el = time.perf_counter() - t0
    return X, el, c2


def cpu_baselines(h0, H, C, psi0, fields, krylov, ncols, ncols_all):
    """The oracle (a port of richmol's CPU path) on the host, before the GPU
    is touched: ncols rows on one core, then ncols_all rows split over every
    usable core (one process per core, as richmol's MPI notebook splits rows)."""
    X1, el1, c2 = _oracle_run((h0, H, C, psi0[:ncols], fields, krylov))
    one = {"value": ncols * len(fields) / el1, "seconds": el1, "rows": ncols, "X": X1, "cos2": c2}
    cores = usable_cores()
    allc = None
    if cores > 1 and ncols_all > 0:
        import multiprocessing as mp
        parts = np.array_split(np.arange(ncols_all), cores)
        jobs = [(h0, H, C, psi0[p], fields, krylov) for p in parts if len(p)]
        ctx = mp.get_context("fork")
        t0 = time.perf_counter()
        with ctx.Pool(len(jobs)) as pool:
            res = pool.map(_oracle_run, jobs)
        wall = time.perf_counter() - t0
        allc = {"value": ncols_all * len(fields) / wall, "seconds": wall, "rows": ncols_all,
                "cores": len(jobs), "max_worker_seconds": max(r[1] for r in res)}
    return one, allc


def timed_run(ens, warm, timed, precision="fp64", obs=None):
    """Warm-up run, then the timed run (no host read inside it):
    (wall s, summed kernel ms, launches, per-step ensemble sums of the
    observable or None).  obs: the observable read after every step inside
    the launch (rm_propagate_expect), as the reference's loop reads it."""
    import torch
    # the field tables and the outputs first, so that the warm-up calls
    # directly precede the timed one; timing on from the warm-up, so the timed
    # call creates no events.  The warm-up steps go in two calls (one if
    # there is a single step): the timed call is then a repeated call of the
    # time loop, as a user's loop makes them (the first repeat had carried
    # ~60 us of one-off host work)
    h = (len(warm) + 1) // 2
    warm_parts = [w for w in (warm[:h], warm[h:]) if len(w)]
    tabs = [ens.prepare(w) for w in warm_parts] + [ens.prepare(timed)]
    it = len(tabs) - 1
    sums = None
    if obs is not None:
        # the caller's output arrays (per wavepacket and summed), allocated
        # before the timed call like the wavepackets
        dev, B = ens.psi.device, ens.psi.shape[0]
        sums = [torch.empty(t.shape[0], dtype=torch.complex128, device=dev) for t in tabs]
        exs = [torch.empty((t.shape[0], B), dtype=torch.complex128, device=dev) for t in tabs]
    kw = lambda i: {} if obs is None else {"obs": obs, "expect": exs[i], "expect_sum": sums[i]}
    ens.set_precision(precision)
    ens.plan.timing(True)
    for i in range(it):
        ens.propagate(None, table=tabs[i], check=False, **kw(i))
        ens.plan.kernel_time()
        ens.check()
    torch.cuda.synchronize()
    # Python's collector is paused during the timed call, as timeit does: a
    # collection pass of the set-up's objects landing in a call of ~0.2 ms
    # would otherwise be timed with it
    gc_on = gc.isenabled()
    gc.disable()
    try:
        t0 = time.perf_counter()
        ens.propagate(None, table=tabs[it], check=False, **kw(it))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    finally:
        if gc_on:
            gc.enable()
    ms, nl = ens.plan.kernel_time()
    ens.plan.timing(False)
    ens.check()
    return wall, ms, nl, (sums[it].cpu().numpy() if sums else None)


def cos2_of(ens, cplan):
    return ens.expect(cplan).real.cpu().numpy() + 1.0 / 3.0


def accuracy_window(H, h0, f, psi0, m, cplan, precisions=("fp64", "fp32")):
    """Every step of the window 0..ACCURACY_STEPS-1 (across the pulse peak)
    for each precision in one launch, <cos^2 theta> of every wavepacket read
    inside it after each step: {precision: (nsteps, ncols) array, final rows}."""
    from richmol_amd import ensemble
    out = {}
    t = (np.arange(ACCURACY_STEPS) + 0.5) * DT_FS
    fields = [[0.0, 0.0, F] for F in pulse(t)]
    for prec in precisions:
        e = ensemble.Ensemble(H, h0, f, psi0, krylov_fixed=m)
        e.set_kernel("subspace")
        e.set_precision(prec)
        ex, _ = e.propagate(fields, obs=cplan)
        out[prec] = (ex.real.cpu().numpy() + 1.0 / 3.0, e.psi.cpu().numpy())
    return out


def asymtop_leg(nsteps, cpu_steps=3):
    """The asymmetric-top configuration: water-like rotor (A, B, C = 27.88,
    14.51, 9.28 cm^-1, D2), J <= 20 with all m (N = 12 341), dipole
    (0, 0, -0.7288) au and polarizability diag(9.14, 9.87, 9.45) au, static
    50 kV/cm Z field plus a cycle-averaged X pulse (FWHM 100 fs, 5e12 W/cm^2,
    centred at 400 fs), psi0 = ground state, dt = 5 fs, adaptive Lanczos,
    FP64: the whole run as one rm_propagate call."""
    import torch
    from richmol_amd import ensemble
    from richmol_amd.molecules import water, water_fields
    dt = 5.0
    t = (np.arange(nsteps) + 0.5) * dt
    h0, leaves, bas = water(20)
    fields = water_fields(t)
    for T, F in zip(leaves, fields):
        T.field(list(F[0]))
    H = leaves[0] + leaves[1] + leaves[2]
    f = -1j * dt * 1e-15 * __import__("richmol_amd").units.enr_units()["invcm"] / \
        __import__("richmol_amd").units.HBAR
    n = 12341
    X = np.zeros((1, n), dtype=np.complex128)
    X[0, 0] = 1.0
    ens = ensemble.Ensemble(H, h0, f, X)
    warm = ens.prepare([F[:10] for F in fields], per_term=True)
    ens.propagate(None, table=warm)
    ens.load(X)
    table = ens.prepare(fields, per_term=True)
    kused = torch.zeros((nsteps, 1), dtype=torch.int32, device=ens.psi.device)
    torch.cuda.synchronize()
    ens.plan.timing(True)
    t0 = time.perf_counter()
    ens.propagate(None, table=table, kused=kused)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms, nl = ens.plan.kernel_time()
    ens.plan.timing(False)
    ens.check()
    rows, ents = ens.plan.domain_sizes(X)
    out = {"n_states": n, "nnz": ens.plan.nnz, "steps": nsteps, "dt_fs": dt,
           "wavepacket_steps_per_s": nsteps / wall, "ms_per_step": wall / nsteps * 1e3,
           "kernel_ms_per_step": kms / nsteps, "launches": nl,
           "domain_rows": int(rows[0]), "domain_entries": int(ents[0]),
           "kernel": "k_long (cooperative form, one launch for the run)" if rows[0] > 1024 else ens.plan.kernel,
           "mean_krylov_dim": float(kused.float().mean()),
           "norm_err": float(abs(np.linalg.norm(ens.psi.cpu().numpy()) - 1)),
           "ground_population_end": float(abs(ens.psi[0, 0].item()) ** 2)}
    if cpu_steps:
        from oracle.cartens import OracleCarTens
        from oracle.tdse import OracleTDSE
        from threadpoolctl import threadpool_limits
        Oh0 = OracleCarTens.from_attrs(h0)
        OL = [OracleCarTens.from_attrs(T) for T in leaves]
        to = OracleTDSE(t_end=dt * (cpu_steps + 1), dt=dt, t_units="fs")
        to.time_grid()
        v = X.copy()
        with threadpool_limits(limits=1):
            t0 = time.perf_counter()
            for i in range(cpu_steps):
                for O, F in zip(OL, fields):
                    O.field(list(F[i]))
                v, _ = to.update(OL[0] + OL[1] + OL[2], v, H0=Oh0)
            el = time.perf_counter() - t0
        out["cpu_oracle_steps_per_s"] = cpu_steps / el
    return out


def kron_leg(nsteps=3, ncols=64, jmax=80, krylov=12, t0_fs=50000.0, kernel="gemm", cpu_matvecs=0,
             accuracy_steps=0):
    """The large asymmetric-top configuration: H2S-like rotor (A, B, C =
    10.36, 9.02, 4.73 cm^-1), J <= jmax with all m (N = 708 561 at J <= 80),
    dipole + full polarizability, circularly polarised XY field chirped
    0 -> 10 THz over 100 ps (E0 = 1e8 V/m), dt = 10 fs, fixed Krylov
    dimension 12, ncols wavepackets (the ncols lowest field-free states) on
    the Kronecker-block operator (rm_kplan_propagate: per output J one dense
    product over all wavepackets).  One warm-up step, then nsteps timed
    steps from t0_fs, in FP64 and with FP32 dense products from the same
    start; <cos^2 theta> (molecular z = C2 axis against lab Z) of both."""
    import torch
    from richmol_amd.kron import DeviceKronPlan, KronOperator, slot_products
    from richmol_amd.molecules import centrifuge_fields, large_top
    t_b = time.perf_counter()
    kop, bas = large_top(jmax)
    dp = DeviceKronPlan(kop)
    dp.set_kernel(kernel)
    kc = KronOperator(bas, [np.diag([0.0, 0.0, 1.0])], [1.0])
    dc = DeviceKronPlan(kc)
    build_s = time.perf_counter() - t_b
    n = kop.n
    f = exp_fac(10.0)
    t = t0_fs + (np.arange(nsteps + 1) + 0.5) * 10.0
    fp = slot_products(kop, centrifuge_fields(t))
    cfp = torch.from_numpy(slot_products(kc, [np.array([[0.0, 0.0, 1.0]])])[0]).cuda()
    order = np.argsort(kop.energies, kind="stable")[:ncols]
    psi0 = torch.zeros((ncols, n), dtype=torch.complex128, device="cuda:0")
    psi0[torch.arange(ncols), torch.from_numpy(order).cuda()] = 1.0
    ph = dp.phases(f)
    y = torch.empty_like(psi0)

    def cos2(psi):
        dc.matvec(psi, y, cfp)
        return torch.real(torch.sum(psi.conj() * y, dim=1)).cpu().numpy()

    out = {"n_states": n, "ncols": ncols, "jmax": jmax, "krylov_fixed": krylov, "dt_fs": 10.0,
           "matvec_form": kernel,
           "blocks": len(kop.blocks), "host_build_s": build_s, "steps_timed": nsteps,
           "t_start_fs": float(t[1] - 5.0), "config_steps": 10000,
           "gemm_flops_per_step": dp.gemm_flops_per_col * ncols * krylov}
    res = {}
    for prec, peak in (("fp64", 78.6), ("fp32", 157.3)):
        psi = psi0.clone()
        dp.propagate(psi, fp[:1], f, phase=ph, krylov_fixed=krylov, precision=prec)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        dp.propagate(psi, fp[1:], f, phase=ph, krylov_fixed=krylov, precision=prec)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t1
        tf = out["gemm_flops_per_step"] * nsteps / wall / 1e12
        res[prec] = (psi, cos2(psi))
        out[prec] = {"wavepacket_steps_per_s": ncols * nsteps / wall, "ms_per_step": wall / nsteps * 1e3,
                     "dense_tflops_effective": tf, "peak_tflops": peak, "frac_of_peak": tf / peak,
                     "norm_err": float((torch.linalg.vector_norm(psi, dim=1) - 1).abs().max())}
        del psi
    if cpu_matvecs:
        # CPU sample: the reference's block loop Y[J1] += MF_w X[J2] K_w^T (field.py:1235-1237)
        # on the same factored operator in numpy, one wavepacket, one thread
        from threadpoolctl import threadpool_limits
        fpr = kop.field_products([F[0] for F in centrifuge_fields(t[:1])])
        x1 = psi0[:1].cpu().numpy()
        with threadpool_limits(limits=1):
            t1 = time.perf_counter()
            for _ in range(cpu_matvecs):
                kop.matvec(x1, fpr)
            tmv = (time.perf_counter() - t1) / cpu_matvecs
        out["cpu_port"] = {"wavepacket_steps_per_s": 1.0 / (krylov * tmv), "matvec_s": tmv, "cores": 1,
                           "sample": f"{cpu_matvecs} host matvec(s) of one wavepacket (numpy block loop on "
                                     f"the factored operator, 1 thread); a step is {krylov} matvecs"}
    if accuracy_steps:
        # FP32 products against FP64 over a longer stretch from the same start:
        # <cos^2 theta> of every wavepacket and the norms every 10 steps
        ta = t0_fs + (np.arange(accuracy_steps) + 0.5) * 10.0
        fpa = slot_products(kop, centrifuge_fields(ta))
        pa = {q: psi0.clone() for q in ("fp64", "fp32")}
        worst_c2, worst_norm, c2_track = 0.0, 0.0, []
        for k0 in range(0, accuracy_steps, 10):
            for q in ("fp64", "fp32"):
                dp.propagate(pa[q], fpa[k0:k0 + 10], f, phase=ph, krylov_fixed=krylov, precision=q)
            c64, c32 = cos2(pa["fp64"]), cos2(pa["fp32"])
            worst_c2 = max(worst_c2, float(np.abs(c32 - c64).max()))
            worst_norm = max(worst_norm, float((torch.linalg.vector_norm(pa["fp32"], dim=1) - 1).abs().max()))
            c2_track.append(float(c64.mean()))
        out["fp32_vs_fp64_long"] = {"steps": accuracy_steps, "max_abs_cos2_any_wavepacket": worst_c2,
                                    "max_norm_err_fp32": worst_norm,
                                    "max_abs_amplitude_end": float((pa["fp32"] - pa["fp64"]).abs().max()),
                                    "cos2_fp64_mean_every_10_steps": c2_track}
        del pa
    out["fp32_vs_fp64"] = {"max_abs_amplitude": float((res["fp32"][0] - res["fp64"][0]).abs().max()),
                           "max_abs_cos2": float(np.abs(res["fp32"][1] - res["fp64"][1]).max()),
                           "cos2_fp64_mean": float(res["fp64"][1].mean())}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--ncols", type=int, default=512)
    ap.add_argument("--jmax", type=int, default=40)
    ap.add_argument("--krylov", type=int, default=12)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-cols", type=int, default=16)
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--cpu-cols-all", type=int, default=256,
                    help="rows of the all-core CPU baseline (split over the usable cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", default="auto",
                    help="headline kernel: auto (subspace when it applies) | subspace | tuned | ell | "
                         "ell:<i> | resident | global")
    ap.add_argument("--full-kernel", default="tuned",
                    help="kernel of the full-vector comparison run (tuned = autotuner cache)")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline run only (no full-vector, precision, single-wavepacket, serving legs)")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests of the serving stream (the serving configuration's 2000)")
    ap.add_argument("--asym-steps", type=int, default=2000,
                    help="steps of the asymmetric-top leg (0: skip)")
    ap.add_argument("--asym-only", action="store_true",
                    help="only the asymmetric-top leg (for profiling its kernel) and exit")
    ap.add_argument("--kron-steps", type=int, default=3,
                    help="timed steps of the large asymmetric-top (Kronecker-block) leg (0: skip)")
    ap.add_argument("--kron-accuracy-steps", type=int, default=60,
                    help="steps of the large-top FP32-vs-FP64 comparison (0: skip)")
    ap.add_argument("--kron-kernel", default="gemm", choices=["gemm", "mfma", "fused"],
                    help="matvec form of the large asymmetric-top leg")
    ap.add_argument("--kron-only", action="store_true",
                    help="only the large asymmetric-top leg (for profiling its kernels) and exit")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the serving leg (for profiling its kernels) and exit")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer 3 times (reference for the HBM "
                         "byte counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        nb = args.calibrate_copy * 1024 * 1024
        x = torch.ones(nb // 8, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": nb, "calibrate_copy_bytes_written": nb,
                          "copies": 3}))
        return
    if args.kron_only:
        print(json.dumps({"large_top_j80": kron_leg(args.kron_steps, kernel=args.kron_kernel,
                                                    accuracy_steps=args.kron_accuracy_steps)}), flush=True)
        return
    if args.serve_only:
        import torch
        from richmol_amd import serving
        torch.cuda.set_device(0)
        print(json.dumps({"serving": serving.benchmark(n_requests=args.serve_requests, seed=args.seed)}),
              flush=True)
        return
    if args.asym_only:
        print(json.dumps({"asymtop_water_j20": asymtop_leg(args.asym_steps, cpu_steps=0)}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(wavepacket columns are independent: a multi-GPU run would shard them)", file=sys.stderr)
        sys.exit(2)
    nfield = pulse_steps()
    if args.warmup + args.steps > nfield:
        print(f"warmup + steps must stay within the pulse (<= {nfield} steps of {DT_FS} fs carry a field)",
              file=sys.stderr)
        sys.exit(2)

    from richmol_amd import ensemble
    bas, h0, H, cos2 = workload(args.jmax)
    n = sum(len(bas.mlist[J]) for J in bas.Jlist)
    idx = ensemble.thermal_sample(bas, 1.0, args.ncols, seed=args.seed)
    psi0 = ensemble.basis_rows(n, idx)
    f = exp_fac(DT_FS)
    m = args.krylov
    # steps around the pulse peak for the CPU baseline (and its device check)
    s0 = int(T0_FS / DT_FS) - args.cpu_steps // 2
    peak = [[0.0, 0.0, F] for F in pulse((s0 + np.arange(args.cpu_steps) + 0.5) * DT_FS)]
    cpu_one = cpu_all = None
    if not args.no_cpu:
        # before the GPU is initialised: the all-core leg forks worker processes
        cpu_one, cpu_all = cpu_baselines(h0, H, cos2, psi0, peak, m, args.cpu_cols, args.cpu_cols_all)
        print(json.dumps({"progress": "cpu baseline", "one_core": cpu_one["value"],
                          "all_cores": cpu_all and cpu_all["value"]}), file=sys.stderr, flush=True)

    import torch
    from richmol_amd.plan import plan_for
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    nst = args.warmup + args.steps
    fields = [[0.0, 0.0, F] for F in pulse((np.arange(nst) + 0.5) * DT_FS)]
    warm, timed = fields[:args.warmup], fields[args.warmup:]
    cos2.field([0, 0, 1])
    cplan = plan_for(cos2)
    cplan.set_field(cos2)

    # ---- headline: production path
    ens = ensemble.Ensemble(H, h0, f, psi0, krylov_fixed=m)
    ens.set_kernel(args.kernel)
    wall, kms, nl, c2_sums = timed_run(ens, warm, timed, obs=cplan)
    value = args.ncols * args.steps / wall
    p = ens.plan
    used = "subspace" if (nl == 1 and args.kernel in ("auto", "subspace")) else p.kernel
    c2_fp64 = cos2_of(ens, cplan)

    # algorithmic FP64 work per wavepacket-step on the rows it spans (d rows,
    # e stored entries): m matvecs of 8 e flops + 7 complex vector operations
    # of 8 d flops per iteration
    pair = False
    if used == "subspace":
        d_rows, d_nnz = p.domain_sizes(psi0)
        flops_launch = float(np.sum(8.0 * m * d_nnz + 8.0 * m * d_rows * 7)) * args.steps
        # fixed dimension <= 16 and domains <= 32 rows: two wavepackets per wave
        pair = m <= 16 and int(np.max(d_rows)) <= 32
        kernel_name = "k_sub_pair" if pair else "k_sub_prop"
    else:
        flops_launch = args.ncols * (8.0 * m * p.nnz + 8.0 * m * n * 7)
        kernel_name = "k_lanczos_ell"
    kern_ms = kms / nl
    tflops = flops_launch / (kern_ms * 1e-3) / 1e12
    tm = TRAFFIC_MODEL.get(("pair" if pair else "other", args.jmax, args.ncols, m))
    traffic = (tm[0] + tm[1] * args.steps) if (tm and used == "subspace") else None

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "wavepacket-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic (Boltzmann-sampled |J,m> initial states, analytic pulse)",
        "config": {"workload": "thermal ensemble: linear rotor J<=40 all m (N=1681), "
                               "512 wavepackets at 1 K, Z pulse 300 fs / 5e11 W/cm^2 centred at 375 fs, "
                               "dt=10 fs, split operator + fixed Krylov m=12, FP64",
                   "jmax": args.jmax, "n_states": n, "wavepackets": args.ncols, "krylov": m,
                   "dt_fs": DT_FS, "first_step": args.warmup, "kernel": args.kernel, "kernel_used": used,
                   "launch": "one rm_propagate call for the timed steps",
                   "parallelism": ("one GPU; two wavepackets per wave on their invariant subspaces" if pair
                                   else "one GPU; one wave per wavepacket on its invariant subspace")
                   if used == "subspace" else "one GPU; wavepackets split over persistent workgroups"},
        "roofline": {"bound": "fp64-valu (latency)", "achieved": tflops, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": tflops / FP64_PEAK_TFLOPS,
                     "traffic": traffic, "precision": "fp64",
                     "kernel": kernel_name, "kernel_ms": kern_ms, "steps_per_launch": args.steps,
                     "flops_per_launch": flops_launch,
                     "traffic_model": ({"fixed_bytes": tm[0], "bytes_per_step": tm[1], "source": tm[2]}
                                       if tm else None),
                     "note": "serial Krylov recurrence of each wavepacket in registers and LDS; "
                             "no MFMA instruction, HBM traffic ~1e-3 of the 8 TB/s roofline"},
        "cos2_ensemble": float(c2_fp64.mean()),
        "cos2_ensemble_timed_steps": [float(c2_sums.real.min() / args.ncols + 1.0 / 3.0),
                                      float(c2_sums.real.max() / args.ncols + 1.0 / 3.0)],
        "observable": "<cos^2 theta> of every wavepacket and their sum read after every step "
                      "inside the timed launch (rm_propagate_expect), as the reference's loop does",
    }
    print(json.dumps({"progress": "headline", "value": value, "kernel_used": used}), file=sys.stderr,
          flush=True)

    if not args.no_extras:
        psi_fp64 = ens.psi.cpu().numpy()
        # ---- full-vector kernel on the same workload
        e2 = ensemble.Ensemble(H, h0, f, psi0, krylov_fixed=m)
        e2.set_kernel(args.full_kernel)
        w2, k2, n2, _ = timed_run(e2, warm, timed, obs=cplan)
        flops_full = args.ncols * (8.0 * m * e2.plan.nnz + 8.0 * m * n * 7)
        k2l = k2 / max(n2, 1)
        meas2 = MEASURED_TRAFFIC_FULL.get(("ell", args.jmax, args.ncols, m))
        result["full_vector"] = {
            "kernel": e2.plan.kernel, "value": args.ncols * args.steps / w2,
            "kernel_ms_per_step": k2l, "tflops": flops_full / (k2l * 1e-3) / 1e12,
            "traffic_per_launch": meas2[0] if meas2 else None,
            "max_abs_diff_vs_headline": float(np.abs(e2.psi.cpu().numpy() - psi_fp64).max())}
        # ---- FP32 matvec precision mode: throughput on the timed steps, and
        # its observable errors against FP64 after every step of a window
        # across the pulse peak (independent of --steps)
        e3 = ensemble.Ensemble(H, h0, f, psi0, krylov_fixed=m)
        e3.set_kernel("subspace")
        w3, k3, n3, _ = timed_run(e3, warm, timed, precision="fp32", obs=cplan)
        acc = accuracy_window(H, h0, f, psi0, m, cplan)
        c64, c32 = acc["fp64"][0], acc["fp32"][0]
        p64, p32 = acc["fp64"][1], acc["fp32"][1]
        result["precision_fp32"] = {
            "value": args.ncols * args.steps / w3,
            "window_steps": ACCURACY_STEPS, "window_fs": [0.0, ACCURACY_STEPS * DT_FS],
            "max_abs_psi_err_end": float(np.abs(p32 - p64).max()),
            "max_cos2_err_per_wavepacket_any_step": float(np.abs(c32 - c64).max()),
            "max_cos2_ensemble_err_any_step": float(np.abs(c32.mean(axis=1) - c64.mean(axis=1)).max()),
            "cos2_ensemble_swing": [float(c64.mean(axis=1).min()), float(c64.mean(axis=1).max())],
            "norm_err_end": float(np.abs(np.linalg.norm(p32, axis=1) - 1).max()),
            "gate": 1e-6,
            "note": "complex64 operator and gathered vector, FP32 accumulation; Krylov recurrence, "
                    "norms, exp(fT), phases and the wavepacket in FP64"}
        # ---- plain single-wavepacket FP64 path: one TDSE.update call per step,
        # the default (one-step rm_propagate on the subspace kernels) and the
        # full-vector rm_step
        from richmol_amd.tdse import TDSE
        nsw = min(60, args.steps)

        def per_step(kernel):
            td = TDSE(t_end=DT_FS * (nst + 1), dt=DT_FS, t_units="fs")
            td.time_grid()
            v = torch.from_numpy(psi0[:1].copy()).to(dev)
            for F in warm:
                H.field(F)
                v, _ = td.update(H, v, H0=h0, krylov_fixed=m, kernel=kernel)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for F in timed[:nsw]:
                H.field(F)
                v, _ = td.update(H, v, H0=h0, krylov_fixed=m, kernel=kernel)
            torch.cuda.synchronize()
            return nsw / (time.perf_counter() - t0), v

        single, v_auto = per_step("auto")
        single_sub, v_sub = per_step("subspace")
        result["single_wavepacket_fp64"] = {
            "value": single, "steps": nsw,
            "path": "TDSE.update per step, one wavepacket, FP64 (host loop; update's default for one "
                    "row: rm_step on the full-vector kernel)",
            "value_subspace": single_sub,
            "path_subspace": "the same loop with kernel='subspace' (one-step rm_propagate on the subspace "
                             "kernels, domains kept between calls)",
            "max_abs_diff_paths": float((v_sub - v_auto).abs().max()),
            "speedup_of_value": value / single}
        print(json.dumps({"progress": "extras"}), file=sys.stderr, flush=True)

        # ---- the asymmetric-top configuration (single wavepacket, N = 12 341)
        if args.asym_steps > 0:
            result["asymtop_water_j20"] = asymtop_leg(args.asym_steps)
            print(json.dumps({"progress": "asymtop"}), file=sys.stderr, flush=True)

        # ---- the large asymmetric-top configuration (64 wavepackets, N = 708 561)
        if args.kron_steps > 0:
            result["large_top_j80"] = kron_leg(args.kron_steps, kernel=args.kron_kernel,
                                               cpu_matvecs=0 if args.no_cpu else 1,
                                               accuracy_steps=args.kron_accuracy_steps)
            print(json.dumps({"progress": "large_top"}), file=sys.stderr, flush=True)

        # ---- serving: continuous batching of independent propagation requests
        if args.serve_requests > 0:
            from richmol_amd import serving
            result["serving"] = serving.benchmark(n_requests=args.serve_requests, seed=args.seed)

    if cpu_one is not None:
        # the same sample on the device, from the same initial rows
        ens2 = ensemble.Ensemble(H, h0, f, psi0[:args.cpu_cols], krylov_fixed=m)
        _, c2s = ens2.propagate(peak, obs=cplan, expect_sum=True)
        dev_err = float(np.abs(ens2.psi.cpu().numpy() - cpu_one["X"]).max())
        c2_err = float(np.abs(c2s.cpu().numpy() - cpu_one["cos2"]).max())
        result["cpu_baseline"] = {
            "value": cpu_one["value"], "unit": "wavepacket-steps/s", "cores": 1, "kind": "port",
            "sample": f"{args.cpu_cols} wavepackets x {args.cpu_steps} steps of the same workload at the "
                      f"pulse peak ({cpu_one['seconds']:.1f} s), FP64 oracle (numpy/scipy restatement of "
                      "richmol's TDSE.update row loop and block matvec, then <cos^2 theta> of every row, as "
                      "examples/ocs_alignment.py:89-103 loops), 1 thread"}
        if cpu_all is not None:
            result["cpu_baseline_all_cores"] = {
                "value": cpu_all["value"], "unit": "wavepacket-steps/s", "cores": cpu_all["cores"],
                "kind": "port",
                "sample": f"{cpu_all['rows']} wavepackets x {args.cpu_steps} steps, rows split over "
                          f"{cpu_all['cores']} processes of 1 thread each ({cpu_all['seconds']:.1f} s wall)"}
        result["host_cores"] = {"affinity": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count(),
                                "used_all_cores_leg": cpu_all["cores"] if cpu_all else 1}
        result["parity_vs_oracle_max_abs"] = dev_err
        result["parity_vs_oracle_cos2_sum_max_abs"] = c2_err
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
