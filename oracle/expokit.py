"""CPU restatement of Expokit's ZHEXPV and ZGPADM (oracle; test
infrastructure only).

The reference's ``propag='external'`` (richmol/tdse.py:383-391, 404-412)
calls ``zhexpv(vec, onenormest(H.tomat(form='full')), 12, exp_fac.imag,
lambda v: cartensvec(v) * 1j, tol=tol)`` (richmol/pyexpokit.py:55-89), the
f2py wrapper of the Fortran routine ZHEXPV (expokit/expokit.f:2783-3170),
whose small exponentials come from ZGPADM (expokit.f:763-933, ideg = 6).
This module restates both, step for step, in numpy:

* ``zgpadm``: irreducible rational Pade approximation of type (6, 6) with
  scaling and squaring (expokit.f:818-932);
* ``zhexpv``: Krylov projection with step-size control -- the three-term
  recurrence that ZHEXPV runs for a Hermitian A, the `happy breakdown' test
  against the absolute 1e-7, the corrected (m + 2) x (m + 2) matrix, the
  local error estimate, rejection and next-step rules, mxstep = 500,
  mxreject = 0, delta = 1.2, gamma = 0.9 (expokit.f:2882-3170).

The reference hands ZHEXPV the matrix A = iH, which is anti-Hermitian: the
recurrence still satisfies A V = V Hs + (last column), with Hs the
tridiagonal it stores, so w = beta V exp(t Hs) e1 is a polynomial
approximation of exp(tA) v, but its basis is not orthonormal.  The
restatement keeps that behaviour (it runs the same recurrence); for a
Hermitian A it is the textbook method and must equal exp(tA) v within the
tolerance, which is how it is pinned (tests/test_oracle_expokit.py).
"""
import math

import numpy as np
from scipy.linalg import lu_factor, lu_solve

# expokit.f:2882-2887
MXSTEP = 500
MXREJECT = 0
IDEG = 6
DELTA = 1.2
GAMMA = 0.9


class ExpokitError(Exception):
    """pyexpokit.py:5-14"""


def _machine_eps():
    """The loop of expokit.f:2944-2948: eps = |3 (4/3 - 1) - 1|."""
    p1 = 4.0 / 3.0
    while True:
        p2 = p1 - 1.0
        p3 = p2 + p2 + p2
        eps = abs(p3 - 1.0)
        if eps != 0.0:
            return eps


def _round_step(t):
    """expokit.f:2964-2965 (and 3085-3086, 3122-3123): keep ~2 significant
    digits, rounding half up."""
    if not math.isfinite(t):
        return t
    sqr1 = math.sqrt(0.1)
    p1 = 10.0 ** (_nint(math.log10(t) - sqr1) - 1)
    return math.trunc(t / p1 + 0.55) * p1


def _nint(x):
    """Fortran NINT: round half away from zero."""
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def zgpadm(H, t, ideg=IDEG):
    """exp(t H) of a complex square matrix (expokit.f:763-932); returns
    (E, ns).  The norm is the largest row sum of |H| (expokit.f:833-844)."""
    H = np.asarray(H, dtype=np.complex128)
    m = H.shape[0]
    hnorm = abs(t * np.max(np.sum(np.abs(H), axis=1)))
    if hnorm == 0.0:
        raise ExpokitError("null H in input of ZGPADM")
    ns = max(0, int(math.log(hnorm) / math.log(2.0)) + 2)
    scale = t / float(2 ** ns)
    scale2 = scale * scale
    # Pade coefficients (expokit.f:850-855)
    coef = [1.0]
    i, j = ideg + 1, 2 * ideg + 1
    for k in range(1, ideg + 1):
        coef.append(coef[k - 1] * float(i - k) / float(k * (j - k)))
    H2 = scale2 * (H @ H)
    I = np.eye(m, dtype=np.complex128)
    p = coef[ideg - 1] * I
    q = coef[ideg] * I
    # Horner (expokit.f:875-889): odd passes update q, even passes p
    odd = True
    k = ideg - 1
    while True:
        if odd:
            q = q @ H2 + coef[k - 1] * I
        else:
            p = p @ H2 + coef[k - 1] * I
        odd = not odd
        k -= 1
        if k <= 0:
            break
    # (+/-)(I + 2 (q - p)^-1 p) with the odd polynomial times scale H (expokit.f:893-913)
    if odd:
        q = scale * (q @ H)
    else:
        p = scale * (p @ H)
    E = 2.0 * lu_solve(lu_factor(q - p), p) + I
    if ns == 0 and odd:
        return -E, ns
    for _ in range(ns):
        E = E @ E
    return E, ns


def zhexpv(v, anorm, m, t, matvec, tol=0.0, stats=None):
    """w = exp(t A) v by ZHEXPV (expokit.f:2783-3170), A given by matvec.

    Raises ExpokitError for the exit flags pyexpokit.check_exit_flag turns
    into errors (pyexpokit.py:7-14).  stats (a dict), when given, receives
    the accounts of expokit.f:3146-3169 (nmult, nexph, nscale, nstep,
    nreject, ibrkflag, mbrkdwn, t_now, hump, ...).
    """
    v = np.asarray(v, dtype=np.complex128).ravel()
    n = v.shape[0]
    if m >= n or m <= 0:
        raise ExpokitError("bad input arguments")
    k1 = 2
    mh = m + 2
    ibrkflag, mbrkdwn = 0, m
    nmult = nreject = nexph = nscale = 0
    t_out = abs(t)
    tbrkdwn = 0.0
    step_min, step_max = t_out, 0.0
    nstep = 0
    s_error = x_error = 0.0
    t_now = t_new = 0.0
    eps = _machine_eps()
    if tol <= eps:
        tol = math.sqrt(eps)
    rndoff = eps * anorm
    break_tol = 1.0e-7
    sgn = math.copysign(1.0, t)
    w = v.copy()
    beta = np.linalg.norm(w)
    vnorm = beta
    hump = beta
    # the first step size (expokit.f:2960-2965)
    xm = 1.0 / float(m)
    p2 = tol * (((m + 1) / 2.72) ** (m + 1)) * math.sqrt(2.0 * 3.14 * (m + 1))
    t_new = (1.0 / anorm) * (p2 / (4.0 * beta * anorm)) ** xm
    t_new = _round_step(t_new)
    iflag = 0
    avnorm = 0.0
    while t_now < t_out:
        nstep += 1
        t_step = min(t_out - t_now, t_new)
        beta = np.linalg.norm(w)
        V = np.zeros((m + 1, n), dtype=np.complex128)
        Hs = np.zeros((mh, mh), dtype=np.complex128)
        V[0] = (1.0 / beta) * w
        # Lanczos loop (expokit.f:2985-3013); k1 and mbrkdwn keep their
        # values from one step to the next, as in the Fortran
        broke = False
        for j in range(m):
            nmult += 1
            p = matvec(V[j])
            if j > 0:
                p = p - Hs[j - 1, j] * V[j - 1]
            hjj = np.vdot(V[j], p)
            p = p - hjj * V[j]
            hj1j = np.linalg.norm(p)
            Hs[j, j] = hjj
            if hj1j <= break_tol:
                k1 = 0
                ibrkflag = 1
                mbrkdwn = j + 1
                tbrkdwn = t_now
                t_step = t_out - t_now
                broke = True
                break
            Hs[j, j + 1] = hj1j
            Hs[j + 1, j] = hj1j
            V[j + 1] = (1.0 / hj1j) * p
        if not broke:
            nmult += 1
            avnorm = np.linalg.norm(matvec(V[m]))
        # the corrected scheme (expokit.f:3019-3021)
        Hs[m - 1, m] = 0.0
        Hs[m + 1, m] = 1.0
        ireject = 0
        while True:
            nexph += 1
            mx = mbrkdwn + k1
            E, ns = zgpadm(Hs[:mx, :mx], sgn * t_step)
            nscale += ns
            e1 = E[:, 0]
            # local error estimate (expokit.f:3045-3062)
            if k1 == 0:
                err_loc = tol
            else:
                p1 = abs(e1[m]) * beta
                p2 = abs(e1[m + 1]) * beta * avnorm
                if p1 > 10.0 * p2:
                    err_loc = p2
                    xm = 1.0 / float(m)
                elif p1 > p2:
                    err_loc = (p1 * p2) / (p1 - p2)
                    xm = 1.0 / float(m)
                else:
                    err_loc = p1
                    xm = 1.0 / float(m - 1)
            # rejection (expokit.f:3066-3097)
            if k1 != 0 and err_loc > DELTA * t_step * tol and (MXREJECT == 0 or ireject < MXREJECT):
                t_step = GAMMA * t_step * (t_step * tol / err_loc) ** xm
                t_step = _round_step(t_step)
                ireject += 1
                nreject += 1
                if MXREJECT != 0 and ireject > MXREJECT:
                    raise ExpokitError("requested tolerance was too high")
                continue
            break
        # w = beta V exp(t_step Hs) e1 (expokit.f:3101-3105)
        mx = mbrkdwn + max(0, k1 - 1)
        w = beta * (e1[:mx] @ V[:mx])
        beta = np.linalg.norm(w)
        hump = max(hump, beta)
        t_new = GAMMA * t_step * (t_step * tol / err_loc) ** xm
        t_new = _round_step(t_new)
        err_loc = max(err_loc, rndoff)
        t_now = t_now + t_step
        step_min = min(step_min, t_step)
        step_max = max(step_max, t_step)
        s_error += err_loc
        x_error = max(x_error, err_loc)
        if not (MXSTEP == 0 or nstep < MXSTEP):
            iflag = 1
            break
    if stats is not None:
        stats.update(nmult=nmult, nexph=nexph, nscale=nscale, nstep=nstep, nreject=nreject,
                     ibrkflag=ibrkflag, mbrkdwn=mbrkdwn, step_min=step_min, step_max=step_max,
                     x_error=x_error, s_error=s_error, tbrkdwn=tbrkdwn, t_now=sgn * t_now,
                     hump=hump / vnorm, scaled_norm=beta / vnorm)
    if iflag == 1:
        raise ExpokitError("maximum number of steps reached without convergence")
    return w
