"""Launch-shape autotuner for the Krylov step kernel.

The ELL kernel has 12 launch configurations (threads per workgroup, rows per
thread, wavepackets per workgroup, operator staged in LDS or read from L2;
``plan.ell_configs()``).  Which is fastest depends on the basis size N, the
ELL width, the number of wavepackets and the Krylov mode, so it is measured:
every valid configuration runs a few steps of the caller's own workload
(HIP-event timing of the step launches), and the winner is cached per key
(N, ELL width, wavepackets, Krylov mode, device) in memory and in
``tuned_kernels.json`` next to this file.
"""
import json
import os

import numpy as np

CACHE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_kernels.json")
_MEM = None


def _load():
    global _MEM
    if _MEM is None:
        try:
            with open(CACHE_PATH) as f:
                _MEM = json.load(f)
        except (OSError, ValueError):
            _MEM = {}
    return _MEM


def _device_name():
    import torch
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return getattr(p, "gcnArchName", p.name).split(":")[0]


def key(plan, ncols, krylov_fixed):
    mode = f"m{krylov_fixed}" if krylov_fixed else "adaptive"
    width = int(plan._arrays["row_ptr"].size and np.diff(plan._arrays["row_ptr"]).max(initial=0))
    return f"{_device_name()}|n{plan.n}|w{width}|b{ncols}|{mode}"


def candidates(plan, ncols):
    from .plan import ell_configs
    out = ["auto"]
    for i, (ntr, r, c, opl) in enumerate(ell_configs()):
        if plan.n <= ntr * r and (c == 1 or ncols >= 2):
            out.append(f"ell:{i}")
    if len(out) == 1:
        out.append("global")
    return out


def lookup(plan, ncols, krylov_fixed):
    return _load().get(key(plan, ncols, krylov_fixed))


def tune(ens, fields, steps=6, reps=2, save=True, verbose=False):
    """Time every candidate kernel on ``steps`` steps of ``fields`` for the
    ensemble's own rows (restored afterwards) and select the fastest."""
    import torch
    table = ens.prepare(fields[:steps])
    plan = ens.plan
    psi0 = ens.psi.clone()
    results = {}
    for cand in candidates(plan, ens.psi.shape[0]):
        try:
            plan.set_kernel(cand)
        except Exception:
            continue
        best = None
        for _ in range(reps):
            ens.load(psi0)
            torch.cuda.synchronize()
            plan.timing(True)
            try:
                ens.propagate(None, table=table)
            except Exception:
                best = None
                plan.timing(False)
                break
            ms, nl = plan.kernel_time()
            plan.timing(False)
            t = ms / max(nl, 1)
            best = t if best is None else min(best, t)
        if best is not None:
            results[cand] = best
            if verbose:
                print(f"  {cand:8s} {best * 1e3:9.1f} us/launch", flush=True)
    ens.load(psi0)
    choice = min(results, key=results.get)
    plan.set_kernel(choice)
    ens.kernel = choice
    k = key(plan, ens.psi.shape[0], ens.krylov_fixed)
    _load()[k] = choice
    if save:
        try:
            with open(CACHE_PATH, "w") as f:
                json.dump(_MEM, f, indent=1, sort_keys=True)
        except OSError:
            pass
    return choice, results
