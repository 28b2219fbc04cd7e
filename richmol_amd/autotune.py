"""Launch autotuner for the propagation kernels.

Measured on the caller's own workload (its rows, operator and field steps),
per key -- device, basis size N, operator width, number of wavepackets,
Krylov mode and the shape of the wavepackets' domains:

* the production invariant-subspace path of rm_propagate ('auto'): two
  wavepackets per wave or one (``pair``), the launch classes side by side
  on auxiliary streams or in order (``fork``), and for domains longer than
  a wave holds the cooperative / one-workgroup split (``coop_max``) and the
  cooperative grid (``coop_grid``) -- rm_plan_set_tuning;
* the full-vector ELL kernel's launch configurations (threads per
  workgroup, rows per thread, wavepackets per workgroup, operator staged in
  LDS or read from L2; ``plan.ell_configs()``).

Every candidate runs the same few steps from the same rows (restored after
each); a call is timed with HIP events on the caller's stream around the
whole rm_propagate call (host enqueue included), after one untimed call that
sets up its launch lists.  The fastest is kept in memory and in
``tuned_kernels.json`` next to this file, as {"kernel": ..., "tuning": {...}};
``Ensemble(kernel="tuned")`` applies it.
"""
import json
import os

import numpy as np

CACHE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_kernels.json")
_MEM = None
SUBSPACE_MAX_ROWS = 1024   # RM_SUBSPACE_MAX_ROWS


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


def domain_shape(plan, X):
    """(domains of <= 32 rows, of 33-64, of 65-1024, longer, rows of the
    longest) of the rows X: the part of the key that decides the launch
    classes."""
    rows, _ = plan.domain_sizes(np.asarray(X))
    rows = rows[rows > 0]
    return (int(np.sum(rows <= 32)), int(np.sum((rows > 32) & (rows <= 64))),
            int(np.sum((rows > 64) & (rows <= SUBSPACE_MAX_ROWS))), int(np.sum(rows > SUBSPACE_MAX_ROWS)),
            int(rows.max(initial=0)))


def key(plan, ncols, krylov_fixed, shape=None):
    mode = f"m{krylov_fixed}" if krylov_fixed else "adaptive"
    width = int(plan._arrays["row_ptr"].size and np.diff(plan._arrays["row_ptr"]).max(initial=0))
    k = f"{_device_name()}|n{plan.n}|w{width}|b{ncols}|{mode}"
    if shape is not None:
        k += "|d" + ",".join(str(x) for x in shape)
    return k


def candidates(plan, ncols, krylov_fixed=0, shape=None):
    """[(kernel, tuning)] to time: the subspace path's choices that apply to
    this shape, then the ELL configurations that fit N."""
    from .plan import ell_configs
    out = [("auto", {})]
    if shape is not None:
        npair, nwave, nmid, nlong, lmax = shape
        if npair and 1 <= krylov_fixed <= 16:
            out.append(("auto", {"pair": 0}))
        if (npair > 0) + (nwave > 0) + (nmid + nlong > 0) > 1:
            out.append(("auto", {"fork": 0}))
        if nmid:
            out.append(("auto", {"wave_max_rows": SUBSPACE_MAX_ROWS}))     # one wave each up to 1024 rows
        nbig = nmid + nlong
        if nbig:
            groups = (lmax + 63) // 64
            if nbig <= 8:
                # one workgroup each / cooperative, in turn, whatever the entries
                out.append(("auto", {"coop_max": 0}))
                out.append(("auto", {"coop_min_entries": 0}))
                for g in (16, 32, 64):
                    if g < groups:
                        out.append(("auto", {"coop_grid": g, "coop_min_entries": 0}))
            else:
                out.append(("auto", {"coop_max": nbig, "coop_min_entries": 0}))   # cooperative, in turn
    for i, (ntr, r, c, opl) in enumerate(ell_configs()):
        if plan.n <= ntr * r and (c == 1 or ncols >= 2):
            out.append((f"ell:{i}", {}))
    if len(out) == 1:
        out.append(("global", {}))
    return out


def lookup(plan, ncols, krylov_fixed, shape=None):
    """The cached choice as (kernel, tuning), or None."""
    mem = _load()
    v = mem.get(key(plan, ncols, krylov_fixed, shape)) if shape is not None else None
    if v is None:
        v = mem.get(key(plan, ncols, krylov_fixed))
    if v is None:
        return None
    if isinstance(v, str):
        return v, {}
    return v["kernel"], dict(v.get("tuning", {}))


def tune(ens, fields, steps=6, reps=2, save=True, verbose=False, per_term=False):
    """Time every candidate on ``steps`` steps of ``fields`` for the
    ensemble's own rows (restored afterwards), select the fastest and keep
    it for the ensemble's shape.  per_term: ``fields`` holds one (nsteps, 3)
    array per summed tensor, as Ensemble.prepare takes it.  Returns
    ((kernel, tuning), {label: ms per call})."""
    import torch
    table = ens.prepare([F[:steps] for F in fields] if per_term else fields[:steps], per_term=per_term)
    plan = ens.plan
    psi0 = ens.psi.clone()
    shape = domain_shape(plan, psi0.cpu().numpy())
    stream = ens.stream or torch.cuda.current_stream()
    results = {}
    for kern, tun in candidates(plan, ens.psi.shape[0], ens.krylov_fixed, shape):
        label = kern + "".join(f" {k}={v}" for k, v in sorted(tun.items()))
        try:
            ens.kernel, ens.tuning, ens._tuned = kern, dict(tun), None
            ens._apply_kernel()
        except Exception:
            continue
        best = None
        try:
            for rep in range(reps + 1):
                ens.load(psi0)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                ens.propagate(None, table=table, check=False)
                e1.record(stream)
                ens.check()
                e1.synchronize()
                if rep > 0:   # the first call sets up the launch lists
                    t = e0.elapsed_time(e1)
                    best = t if best is None else min(best, t)
        except Exception:
            best = None
        if best is not None:
            results[label] = (best, kern, tun)
            if verbose:
                print(f"  {label:28s} {best * 1e3:9.1f} us/call", flush=True)
    ens.load(psi0)
    label = min(results, key=lambda x: results[x][0])
    _, kern, tun = results[label]
    _load()[key(plan, ens.psi.shape[0], ens.krylov_fixed, shape)] = {"kernel": kern, "tuning": tun}
    if save:
        try:
            with open(CACHE_PATH, "w") as f:
                json.dump(_MEM, f, indent=1, sort_keys=True)
        except OSError:
            pass
    ens.kernel, ens.tuning = "tuned", {}
    ens._tuned = (kern, dict(tun))
    ens._apply_kernel()
    return (kern, tun), {k: v[0] for k, v in results.items()}
