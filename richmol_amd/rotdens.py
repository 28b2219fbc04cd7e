"""Rotational densities on grids of Euler angles (richmol/rotdens.py).

``psi_grid`` and ``dens_grid`` keep the reference's interface
(rotdens.py:10-283: wave functions of the basis states, and the densities
conj(psi_i) psi_j sin(beta) of state pairs) for a rigid-rotor basis of
``rot`` (the reference reads the same coefficients from h.symtop_basis):
state index im * nk + ik inside a (J, symmetry) block, as the operators'.

``density`` is the at-scale form on the device: the density
|psi_w(alpha, beta, gamma)|^2 sin(beta) of wavepackets given by their rows
(full-form order of the basis, e.g. TDSE rows after a propagation), for every
wavepacket or summed over them (a Boltzmann ensemble's density), one
``rm_rotdens`` call.  The symmetric-top functions are
sqrt((2J+1)/(8 pi^2)) e^{i m alpha} d^J_mk(beta) e^{i k gamma} (the complex
conjugate of D^J_mk; d^J from the eigenvectors of J_y, rot/wig.py:5-28).
"""
import functools

import numpy as np


@functools.lru_cache(maxsize=None)
def _jy_eigvecs(J):
    """Eigenvectors of J_y = (J+ - J-)/(2i) over m = -J..J (ascending
    eigenvalues -J..J), J+ |m> = sqrt(J(J+1) - m(m+1)) |m+1>."""
    m = np.arange(-J, J, dtype=np.float64)
    up = np.sqrt(J * (J + 1.0) - m * (m + 1.0))          # <m+1|J+|m>
    Jy = np.zeros((2 * J + 1, 2 * J + 1), dtype=np.complex128)
    i = np.arange(2 * J)
    Jy[i + 1, i] = up / 2j
    Jy[i, i + 1] = -up / 2j
    _, v = np.linalg.eigh(Jy)
    v.setflags(write=False)
    return v


def wigner_d(J, beta):
    """d^J_mk(beta) = <J m| e^{-i beta J_y} |J k>, array (2J+1, 2J+1, nbeta)
    over m, k = -J..J (real, Condon-Shortley phases)."""
    beta = np.atleast_1d(np.asarray(beta, dtype=np.float64))
    v = _jy_eigvecs(int(J))
    mu = np.arange(-J, J + 1, dtype=np.float64)
    e = np.exp(-1j * mu[None, :] * beta[:, None])
    return np.einsum("mi,gi,ki->mkg", v, e, np.conj(v)).real


def psi_grid(basis, alpha, beta, gamma, form="block"):
    """Wave functions of the basis states on the grid (rotdens.py:140-283):
    'block': {J: {sym: (na, nb, ng, nm * nk)}}; 'full': (na, nb, ng, N) in
    the basis' full-form order."""
    if form not in ("block", "full"):
        raise ValueError(f"Unknown value of parameter 'form' = {form}")
    alpha, beta, gamma = (np.atleast_1d(np.asarray(x, dtype=np.float64)) for x in (alpha, beta, gamma))
    na, nb, ng = len(alpha), len(beta), len(gamma)
    out = {}
    for J in basis.Jlist:
        d = wigner_d(J, beta)                                     # (m, k, b)
        fac = np.sqrt((2 * J + 1) / (8 * np.pi ** 2))
        ms = np.asarray(basis.mlist[J], dtype=np.int64)
        ea = np.exp(1j * ms[None, :] * alpha[:, None])             # (a, m)
        out[J] = {}
        for sym in basis.symlist[J]:
            kprim, ck = basis.kfun[J][sym]
            kp = np.asarray(kprim, dtype=np.int64)
            eg = np.exp(1j * kp[None, :] * gamma[:, None])         # (g, p)
            dmk = d[(ms + J)[:, None], (kp + J)[None, :], :]        # (m, p, b)
            res = fac * np.einsum("am,mpb,gp,ps->abgms", ea, dmk, eg, ck)
            out[J][sym] = res.reshape(na, nb, ng, -1)
    if form == "full":
        return np.concatenate([out[J][s] for J in basis.Jlist for s in basis.symlist[J]], axis=-1)
    return out


def dens_grid(basis, alpha, beta, gamma, diag_only=False, form="block"):
    """conj(psi_i) psi_j sin(beta) of state pairs (rotdens.py:10-137):
    'block': {(J1, J2): {(sym1, sym2): (na, nb, ng, n1, n2)}} ((na, nb, ng, n)
    with diag_only, equal J and symmetry only); 'full': the same over the
    full-form order."""
    beta = np.atleast_1d(np.asarray(beta, dtype=np.float64))
    psi = psi_grid(basis, alpha, beta, gamma)
    vol = np.sin(beta)
    dens = {}
    for J1, p1 in psi.items():
        for s1, a in p1.items():
            for J2, p2 in psi.items():
                for s2, b in p2.items():
                    if diag_only and (J1 != J2 or s1 != s2):
                        continue
                    if diag_only:
                        d = np.conj(a) * b * vol[None, :, None, None]
                    else:
                        d = np.einsum("abgk,abgl,b->abgkl", np.conj(a), b, vol)
                    dens.setdefault((J1, J2), {})[(s1, s2)] = d
    if form == "full":
        keys = [(J, s) for J in basis.Jlist for s in basis.symlist[J]]
        if diag_only:
            return np.concatenate([dens[(J, J)][(s, s)] for J, s in keys], axis=-1)
        return np.concatenate([np.concatenate([dens[(J1, J2)][(s1, s2)] for J2, s2 in keys], axis=-1)
                               for J1, s1 in keys], axis=-2)
    return dens


def primitive_coefficients(basis, rows):
    """a^J_mk of wavepacket rows (full-form order): (nrows, jmax+1, M, M)
    complex, M = 2 jmax + 1, m and k from -jmax, including
    sqrt((2J+1)/(8 pi^2))."""
    rows = np.atleast_2d(np.asarray(rows, dtype=np.complex128))
    jmax = max(basis.Jlist)
    M = 2 * jmax + 1
    a = np.zeros((rows.shape[0], jmax + 1, M, M), dtype=np.complex128)
    off = 0
    for J in basis.Jlist:
        fac = np.sqrt((2 * J + 1) / (8 * np.pi ** 2))
        ms = np.asarray(basis.mlist[J], dtype=np.int64)
        for sym in basis.symlist[J]:
            kprim, ck = basis.kfun[J][sym]
            kp = np.asarray(kprim, dtype=np.int64)
            nk = ck.shape[1]
            blk = rows[:, off:off + len(ms) * nk].reshape(rows.shape[0], len(ms), nk)
            a[:, J, (ms + jmax)[:, None], (kp + jmax)[None, :]] += fac * np.einsum("wms,ps->wmp", blk, ck)
            off += len(ms) * nk
    if off != rows.shape[1]:
        raise ValueError(f"rows of {rows.shape[1]} states for a basis of {off}")
    return a


def density(basis, rows, alpha, beta, gamma, per_wavepacket=False, stream=None):
    """|psi_w|^2 sin(beta) of wavepacket rows on the grid, on the device
    (rm_rotdens): (na, nb, ng) summed over the rows, or (nrows, na, nb, ng);
    a torch float64 device tensor."""
    import ctypes
    import torch
    from . import _lib
    from .plan import _stream_handle
    if isinstance(rows, torch.Tensor):
        rows = rows.detach().cpu().numpy()
    alpha, beta, gamma = (np.atleast_1d(np.asarray(x, dtype=np.float64)) for x in (alpha, beta, gamma))
    jmax = max(basis.Jlist)
    M = 2 * jmax + 1
    mu = np.arange(-jmax, jmax + 1, dtype=np.float64)
    dev = torch.device("cuda", torch.cuda.current_device())
    a = torch.from_numpy(primitive_coefficients(basis, rows)).to(dev)
    dtab = np.zeros((jmax + 1, len(beta), M, M))
    for J in range(jmax + 1):
        dtab[J, :, jmax - J:jmax + J + 1, jmax - J:jmax + J + 1] = np.moveaxis(wigner_d(J, beta), 2, 0)
    dtab = torch.from_numpy(dtab).to(dev)
    ea = torch.from_numpy(np.exp(1j * alpha[:, None] * mu[None, :])).to(dev)
    eg = torch.from_numpy(np.exp(1j * gamma[:, None] * mu[None, :])).to(dev)
    sb = torch.from_numpy(np.sin(beta)).to(dev)
    nw = a.shape[0]
    shape = (nw, len(alpha), len(beta), len(gamma)) if per_wavepacket else (len(alpha), len(beta), len(gamma))
    rho = torch.empty(shape, dtype=torch.float64, device=dev)
    vp = ctypes.c_void_p
    lib = _lib.load()
    _lib.check(lib.rm_rotdens(jmax, nw, vp(a.data_ptr()), len(alpha), vp(ea.data_ptr()), len(beta),
                              vp(dtab.data_ptr()), vp(sb.data_ptr()), len(gamma), vp(eg.data_ptr()),
                              1 if per_wavepacket else 0, vp(rho.data_ptr()), vp(_stream_handle(stream))),
               "rm_rotdens")
    return rho
