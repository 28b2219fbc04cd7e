"""CPU restatement of the reference's rotational densities (oracle; test
infrastructure only).

Follows richmol/rot/wig.py:5-28 (jy_eig, wig_d) and richmol/rotdens.py:
psi_grid 140-283 (no vibrational quanta: one v) and dens_grid 10-137.  The
basis is given as the product's rot.RotorBasis (kfun[J][sym] = (k list,
coefficients over primitive k), mlist[J] the primitive m of each m state);
the reference reads the same coefficients from h.symtop_basis.
"""
import numpy as np
from scipy.sparse import diags


def jy_eig(j):
    """wig.py:5-21 (trid=True): eigenvectors of the tridiagonal matrix the
    reference builds and hands to eigh (which reads its lower triangle)."""
    x = lambda m: np.sqrt((j + m) * (j - m + 1))
    ld = np.array([x(n) for n in np.linspace(-j, j, int(2 * j + 1))])
    ud = np.array([x(-n) for n in np.linspace(-j, j, int(2 * j + 1))])
    jmat = diags([ld[1:], np.zeros(2 * j + 1), ud[:-1]], [-1, 0, 1]).toarray() * 0.5j
    _, v = np.linalg.eigh(jmat)
    return v


def wig_d(j, beta):
    """wig.py:24-28: d[m, k, b] over m, k = -j..j."""
    v = jy_eig(j)
    e = np.exp(-1j * np.linspace(-j, j, int(2 * j + 1))[None, :] * np.asarray(beta)[:, None])
    return np.einsum("gi,mi,ki->mkg", e, np.conj(v), v)


def psi_grid(basis, alpha, beta, gamma):
    """rotdens.py:224-283 with one vibrational quantum: {J: {sym: (na, nb,
    ng, nm * nk)}} -- state index im * nk + ik (m states, then k states)."""
    alpha, beta, gamma = (np.asarray(x, dtype=np.float64) for x in (alpha, beta, gamma))
    na, nb, ng = len(alpha), len(beta), len(gamma)
    out = {}
    for J in basis.Jlist:
        mu = np.linspace(-int(J), int(J), int(2 * J) + 1)
        wig = jy_eig(int(J))
        fac = np.sqrt((2 * J + 1) / (8 * np.pi ** 2))
        out[J] = {}
        for sym in basis.symlist[J]:
            kprim, ck = basis.kfun[J][sym]
            k = np.array(kprim, dtype=np.int64)
            k_ind = np.where(k[:, None] == mu)[1]
            egamma = np.exp(1j * k[None, :] * gamma[:, None])
            wig_k = np.einsum("pk,pi,gp->gki", ck, wig[k_ind, :], egamma)
            m = np.array(basis.mlist[J], dtype=np.int64)
            m_ind = np.where(m[:, None] == mu)[1]
            cm = np.eye(len(m))
            ealpha = np.exp(1j * m[None, :] * alpha[:, None])
            wig_m = fac * np.einsum("pm,pi,ap->ami", cm, np.conj(wig)[m_ind, :], ealpha)
            ebeta = np.exp(-1j * mu[None, :] * beta[:, None])
            res = np.einsum("ami,bi,gki->abgmk", wig_m, ebeta, wig_k)
            out[J][sym] = res.reshape(na, nb, ng, -1)
    return out


def dens_grid(basis, alpha, beta, gamma, diag_only=False):
    """rotdens.py:99-137: {(J1, J2): {(sym1, sym2): conj(psi1) psi2 sin(beta)}},
    (na, nb, ng, n1, n2), or (na, nb, ng, n) for diag_only."""
    psi = psi_grid(basis, alpha, beta, gamma)
    vol = np.sin(np.asarray(beta, dtype=np.float64))
    dens = {}
    for J1, p1 in psi.items():
        for s1, a in p1.items():
            for J2, p2 in psi.items():
                for s2, b in p2.items():
                    if diag_only and (J1 != J2 or s1 != s2):
                        continue
                    if diag_only:
                        d = np.einsum("abgk,abgk,b->abgk", np.conj(a), b, vol)
                    else:
                        d = np.einsum("abgk,abgl,b->abgkl", np.conj(a), b, vol)
                    dens.setdefault((J1, J2), {})[(s1, s2)] = d
    return dens
