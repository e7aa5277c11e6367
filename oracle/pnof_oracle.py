"""FP64 CPU oracle for the occupation step of a PNOF orbital optimisation.

TEST INFRASTRUCTURE ONLY. Nothing in the product imports this module: it is
used by ``tests/`` as the checker. It restates, in numpy, the reference's
occupation parametrisations, the PNOF5 / PNOF7 / GNOF coefficient matrices
C^J, C^K and their derivatives, the energy and the occupation gradient (closed
and open shell); every function cites the reference lines it follows (paths
relative to the reference root, ``src/pynof/pnof.py`` unless named).

Where the reference's code has quirks, they are restated as they are:
``ocupacion_softmax`` writes the strong occupations and their derivative rows
at index i (not no1 + i) and leaves n[:no1] at zero (pnof.py:790-799).

Pinning (tests/test_pnof_oracle.py):
- occupations: the sum rule of every subspace (strong + weak = 1), the range
  0 < n < 1, and dn/dgamma against central finite differences of n(gamma);
- coefficients: the Hartree-Fock limit (n in {0, 1}) against the closed-form
  HF energy; PNOF5 against the energy of an antisymmetrised product of
  strongly orthogonal geminals (APSG), built independently from the
  two-electron Slater rules; PNOF7 minus PNOF5 against -sum'' Phi_p Phi_q K_pq;
  symmetry;
- derivatives: the occupation gradient against central finite differences
  of the energy, for every PNOF and both parametrisations, closed and open
  shell (MSpin != 0);
- the open-shell energy (pnof.py:924-988) against explicit loops over the
  reference's blocks; it has no independent closed form here (its
  singly-occupied terms are not the ROHF expression), so beyond the loops and
  the gradient consistency its terms follow the code.
The GNOF inter-pair static and dynamic terms (pnof.py:378-396) have no
independent closed form here: they are pinned by the HF limit (where they
vanish), by the single-pair limit, by symmetry and by the finite-difference
consistency of their derivatives, and otherwise follow the code.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

H_CUT = 0.02 * np.sqrt(2.0)  # GNOF hole cut-off, pnof.py:355
TINY = 1e-15                 # derivative guards, pnof.py:229-231, :308, :456


@dataclass
class Pnof:
    """The fields of PyNOF's ``param`` the occupation step reads (parameters.py:46-110)."""
    no1: int
    ndoc: int
    nsoc: int
    ncwo: int
    ipnof: int = 8
    ista: int = 0
    occ_method: str = "Softmax"
    HighSpin: bool = False
    MSpin: int = 0

    @property
    def nbeta(self):
        return self.no1 + self.ndoc

    @property
    def nalpha(self):
        return self.nbeta + self.nsoc

    @property
    def ndns(self):
        return self.ndoc + self.nsoc

    @property
    def nv(self):
        return self.ncwo * self.ndoc  # parameters.py:94

    @property
    def nbf5(self):
        return self.no1 + self.ndoc * (1 + self.ncwo) + self.nsoc  # parameters.py:65, :157-158


def weak_of(p: Pnof, l: int) -> np.ndarray:
    """Weak orbitals coupled to strong orbital no1 + l: no1 + ndns + (ndoc - l - 1) + ndoc t,
    t < ncwo (the slice ll:ul:ndoc of pnof.py:187-190)."""
    w = p.no1 + p.ndns + (p.ndoc - l - 1) + p.ndoc * np.arange(p.ncwo)
    return w[w < p.nbf5]


# ---------------------------------------------------------------- occupations
def occupation_softmax(x, p: Pnof):
    """ocupacion_softmax (pnof.py:770-810): n (nbf5) and dn/dx (nbf5, nv)."""
    n = np.zeros(p.nbf5)
    dn = np.zeros((p.nbf5, p.nv))
    n[p.no1 + p.ndoc:p.no1 + p.ndns] = 1.0 if p.HighSpin else 0.5
    e = np.exp(np.asarray(x, dtype=np.float64))
    for i in range(p.ndoc):
        w = weak_of(p, i)
        cols = (p.ndoc - i - 1) + p.ndoc * np.arange(len(w))
        den = 1.0 + np.sum(e[cols])
        n[i] = 1.0 / den
        nw = e[cols] / den
        n[w] = nw
        dn[i, cols] = -n[i] * nw
        dn[np.ix_(w, cols)] = -np.outer(nw, nw) + np.diag(nw)
    return n, dn


def occupation_trigonometric(gamma, p: Pnof):
    """ocupacion_trigonometric (pnof.py:704-768): n (nbf5) and dn/dgamma (nbf5, nv).
    Strong n = (1 + cos^2 g_i)/2; weak t of subspace i takes the hole h_i times
    prod_{s<t} cos^2 g_s times sin^2 g_t (no sin^2 factor for the last weak)."""
    gamma = np.asarray(gamma, dtype=np.float64)
    n = np.zeros(p.nbf5)
    dn = np.zeros((p.nbf5, p.nv))
    n[:p.no1] = 1.0
    g = gamma[:p.ndoc]
    n[p.no1:p.no1 + p.ndoc] = 0.5 * (1.0 + np.cos(g) ** 2)
    dstrong = -0.5 * np.sin(2.0 * g)
    for i in range(p.ndoc):
        dn[p.no1 + i, i] = dstrong[i]
    n[p.no1 + p.ndoc:p.no1 + p.ndns] = 1.0 if p.HighSpin else 0.5
    for i in range(p.ndoc):
        w = weak_of(p, i)
        m = len(w)
        gcol = p.ndoc + (p.ndoc - i - 1) + p.ndoc * np.arange(max(m - 1, 0))
        gp = gamma[gcol]
        c2, s2, s2g = np.cos(gp) ** 2, np.sin(gp) ** 2, np.sin(2.0 * gp)
        hole = 1.0 - n[p.no1 + i]
        for t in range(m):
            f = 1.0
            for s in range(t):
                f *= c2[s]
            last = s2[t] if t < m - 1 else 1.0
            n[w[t]] = hole * f * last
            dn[w[t], i] = -dstrong[i] * f * last
            # d/dg_j of the factor cos^2 g_j (j < t): -sin 2g_j
            for j in range(t):
                fj = 1.0
                for s in range(t):
                    if s != j:
                        fj *= c2[s]
                dn[w[t], gcol[j]] = -hole * s2g[j] * fj * last
            if t < m - 1:  # d/dg_t of sin^2 g_t
                dn[w[t], gcol[t]] = hole * f * s2g[t]
    return n, dn


def occupation(gamma, p: Pnof):
    """ocupacion (pnof.py:692-701), Softmax and Trigonometric."""
    if p.occ_method == "Softmax":
        return occupation_softmax(gamma, p)
    if p.occ_method == "Trigonometric":
        return occupation_trigonometric(gamma, p)
    raise ValueError(f"occupation method {p.occ_method!r} is not restated")


# ---------------------------------------------------------------- coefficients
def _phi(n, ista):
    """Phi of PNOF7/GNOF (pnof.py:252-257, :364-366): sqrt(n(1-n)), or 2n(1-n) for ista != 0."""
    if ista == 0:
        f = n * (1.0 - n)
        f[f <= 0] = 0.0
        return np.sqrt(f)
    return 2.0 * n * (1.0 - n)


def _intra(cj, ck, n, p):
    """The intra-pair blocks shared by PNOF5/7/GNOF (pnof.py:184-200, :271-287, :402-418):
    C^J = 0 between a strong orbital and its weak orbitals and among those weak
    orbitals; C^K = sqrt(n_s n_w) strong-weak, -sqrt(n_w n_w') weak-weak."""
    for l in range(p.ndoc):
        s = p.no1 + l
        w = weak_of(p, l)
        cj[s, w] = 0.0
        cj[w, s] = 0.0
        cj[np.ix_(w, w)] = 0.0
        sw = np.sqrt(n[s] * n[w])
        ck[s, w] = sw
        ck[w, s] = sw
        ck[np.ix_(w, w)] = -np.sqrt(np.outer(n[w], n[w]))


def cjck5(n, p: Pnof):
    """CJCKD5 (pnof.py:155-202)."""
    cj = 2.0 * np.outer(n, n)
    ck = np.outer(n, n)
    b, a = p.nbeta, p.nalpha
    if p.MSpin == 0 and p.nsoc > 1:
        ck[b:a, b:a] = 2.0 * np.outer(n[b:a], n[b:a])
    _intra(cj, ck, n, p)
    return cj, ck


def cjck7(n, p: Pnof):
    """CJCKD7 (pnof.py:248-289): PNOF5 plus Phi_p Phi_q in C^K between subspaces."""
    fi = _phi(n, p.ista)
    cj = 2.0 * np.outer(n, n)
    ck = np.outer(n, n) + np.outer(fi, fi)
    b, a = p.nbeta, p.nalpha
    if p.MSpin == 0 and p.nsoc > 1:
        ck[b:a, b:a] = 2.0 * np.outer(n[b:a], n[b:a])
    _intra(cj, ck, n, p)
    return cj, ck


def gnof_dynamic(n, p: Pnof):
    """n_d = n F_g with F_g = exp(-((1 - n_strong)/h_cut)^2) on each subspace, 0 elsewhere
    (pnof.py:355-373); returns (n_d, F per subspace)."""
    nd = np.zeros(p.nbf5)
    F = np.zeros(p.ndoc)
    for l in range(p.ndoc):
        s = p.no1 + l
        F[l] = np.exp(-(((1.0 - n[s]) / H_CUT) ** 2))
        nd[s] = n[s] * F[l]
        w = weak_of(p, l)
        nd[w] = n[w] * F[l]
    return nd, F


def cjck8(n, p: Pnof):
    """CJCKD8, GNOF (pnof.py:350-422)."""
    nd, _ = gnof_dynamic(n, p)
    nd12 = np.sqrt(nd)
    fi = _phi(n, 0)
    o0, b, a = p.no1, p.nbeta, p.nalpha
    cj = 2.0 * np.outer(n, n)
    ck = np.outer(n, n)
    ck[o0:b, a:] += np.outer(fi[o0:b], fi[a:])
    ck[a:, o0:b] += np.outer(fi[a:], fi[o0:b])
    ck[a:, a:] += np.outer(fi[a:], fi[a:])
    if p.MSpin == 0 and p.nsoc > 0:
        ck[o0:b, b:a] += 0.25 * fi[o0:b, None]
        ck[b:a, o0:b] += 0.25 * fi[None, o0:b]
        ck[b:a, a:] += 0.5 * fi[None, a:]
        ck[a:, b:a] += 0.5 * fi[a:, None]
    if p.MSpin == 0 and p.nsoc > 1:
        ck[b:a, b:a] = 0.5
    ck[o0:b, a:] += np.outer(nd12[o0:b], nd12[a:]) - np.outer(nd[o0:b], nd[a:])
    ck[a:, o0:b] += np.outer(nd12[a:], nd12[o0:b]) - np.outer(nd[a:], nd[o0:b])
    ck[a:, a:] += -np.outer(nd12[a:], nd12[a:]) - np.outer(nd[a:], nd[a:])
    _intra(cj, ck, n, p)
    return cj, ck


def coefficients(n, p: Pnof):
    """PNOFi_selector (pnof.py:657-667) for PNOF5, PNOF7 and GNOF."""
    n = np.asarray(n, dtype=np.float64)
    if p.ipnof == 5:
        return cjck5(n, p)
    if p.ipnof == 7:
        return cjck7(n, p)
    if p.ipnof == 8:
        return cjck8(n, p)
    raise ValueError(f"ipnof {p.ipnof} is not restated")


# ---------------------------------------------------------------- derivatives
def _d_intra(Dcj, Dck, n, dn, p):
    """Intra-pair blocks of the derivatives (pnof.py:220-243, :324-347, :498-525):
    zero C^J blocks; d sqrt(n_s n_w) taken on the row orbital, guarded by 1e-15."""
    for l in range(p.ndoc):
        s = p.no1 + l
        w = weak_of(p, l)
        Dcj[s, w, :] = 0.0
        Dcj[w, s, :] = 0.0
        Dcj[np.ix_(w, w)] = 0.0
        a = max(n[s], TINY)
        bw = np.maximum(n[w], TINY)
        Dck[s, w, :] = 0.5 / np.sqrt(a) * np.outer(np.sqrt(n[w]), dn[s, :])
        Dck[w, s, :] = 0.5 / np.sqrt(bw)[:, None] * dn[w, :] * np.sqrt(n[s])
        # -1/2 (dn_w / sqrt(b_w)) sqrt(n_w'): rows w, columns w', derivative index k
        Dck[np.ix_(w, w)] = -0.5 * (dn[w, :] / np.sqrt(bw)[:, None])[:, None, :] * np.sqrt(n[w])[None, :, None]


def _dphi(n, dn, p, ista):
    """dPhi/dgamma for orbitals >= no1 (pnof.py:301-309, :460-465)."""
    fi = _phi(n, ista)
    d = np.zeros_like(dn)
    a = np.maximum(fi[p.no1:], TINY)
    if ista == 0:
        d[p.no1:] = (1.0 / (2.0 * a) * (1.0 - 2.0 * n[p.no1:]))[:, None] * dn[p.no1:]
    else:
        d[p.no1:] = (2.0 * (1.0 - 2.0 * n[p.no1:]))[:, None] * dn[p.no1:]
    return fi, d


def der_coefficients(n, dn, p: Pnof):
    """der_PNOFi_selector (pnof.py:670-689): Dcj, Dck (nbf5, nbf5, nv), the derivative of
    C^J[i, j], C^K[i, j] taken through the row orbital i, as the reference forms them."""
    n = np.asarray(n, dtype=np.float64)
    Dcj = 2.0 * n[None, :, None] * dn[:, None, :]
    Dck = 1.0 * n[None, :, None] * dn[:, None, :]
    if p.ipnof == 7:
        fi, dfi = _dphi(n, dn, p, p.ista)
        Dck = Dck + fi[None, :, None] * dfi[:, None, :]
    elif p.ipnof == 8:
        o0, b, a = p.no1, p.nbeta, p.nalpha
        nd, F = gnof_dynamic(n, p)
        dnd = np.zeros_like(dn)
        for l in range(p.ndoc):
            s = p.no1 + l
            w = weak_of(p, l)
            coc = (1.0 - n[s]) / H_CUT
            dnd[s] = F[l] * dn[s] * (1.0 + n[s] * 2.0 * coc / H_CUT)
            dnd[w] = F[l] * (dn[w] + (2.0 * coc / H_CUT) * np.outer(n[w], dn[s]))
        nd12 = np.sqrt(nd)
        dnd12 = 0.5 * dnd / np.maximum(nd12, TINY)[:, None]
        fi, dfi = _dphi(n, dn, p, 0)
        # Phi terms (pnof.py:474-477)
        Dck[o0:b, a:, :] += dfi[o0:b, None, :] * fi[None, a:, None]
        Dck[a:, o0:b, :] += fi[None, o0:b, None] * dfi[a:, None, :]
        Dck[a:, a:, :] += fi[None, a:, None] * dfi[a:, None, :]
        if p.MSpin == 0 and p.nsoc > 0:  # pnof.py:479-482
            Dck[o0:b, b:a, :] += 0.25 * dfi[o0:b, None, :]
            Dck[a:, b:a, :] += 0.5 * dfi[a:, None, :]
        if p.MSpin == 0 and p.nsoc > 1:  # pnof.py:484-485
            Dck[b:a, b:a, :] = 0.0
        # dynamic terms (pnof.py:487-494)
        Dck[o0:b, a:, :] += dnd12[o0:b, None, :] * nd12[None, a:, None] - dnd[o0:b, None, :] * nd[None, a:, None]
        Dck[a:, o0:b, :] += nd12[None, o0:b, None] * dnd12[a:, None, :] - nd[None, o0:b, None] * dnd[a:, None, :]
        Dck[a:, a:, :] += -nd12[None, a:, None] * dnd12[a:, None, :] - nd[None, a:, None] * dnd[a:, None, :]
    elif p.ipnof != 5:
        raise ValueError(f"ipnof {p.ipnof} is not restated")
    _d_intra(Dcj, Dck, n, dn, p)
    return Dcj, Dck


# ---------------------------------------------------------------- energy and gradient
def calce(n, cj12, ck12, J_MO, K_MO, H_core, p: Pnof) -> float:
    """calce (pnof.py:905-990). MSpin = 0: 2 sum n H + sum n_i J_ii over [:nbeta] and
    [nalpha:nbf5] + sum' C^J J - sum' C^K K, diagonals of C^J, C^K removed. MSpin != 0:
    the C^J/C^K sums skip the singly occupied block S = [nbeta, nalpha), whose orbitals
    enter through n-weighted J/K terms instead (pnof.py:924-988)."""
    cj = np.array(cj12, dtype=np.float64, copy=True)
    ck = np.array(ck12, dtype=np.float64, copy=True)
    np.fill_diagonal(cj, 0.0)
    np.fill_diagonal(ck, 0.0)
    b, a, n5 = p.nbeta, p.nalpha, p.nbf5
    dj = np.diagonal(J_MO)
    if p.MSpin == 0:
        E = 2.0 * np.dot(n, H_core)
        E += np.dot(n[:b], dj[:b]) + np.dot(n[a:n5], dj[a:n5])
        E += np.sum(cj * J_MO.T) - np.sum(ck * K_MO.T)
        return float(E)
    keep = np.ones(n5, dtype=bool)
    keep[b:a] = False  # rows and columns outside S
    E = np.dot(n[:b], 2.0 * H_core[:b] + dj[:b]) + np.dot(n[b:a], 2.0 * H_core[b:a])
    E += np.dot(n[a:n5], 2.0 * H_core[a:n5] + dj[a:n5])
    blk = np.outer(keep, keep)
    E += np.sum((cj * J_MO.T)[blk]) - np.sum((ck * K_MO.T)[blk])
    sj = J_MO[b:a, :].sum(axis=0)  # sum_{s in S} J[s, i]
    sk = K_MO[b:a, :].sum(axis=0)
    E += 2.0 * np.dot(n[:b], sj[:b]) + 2.0 * np.dot(n[a:n5], sj[a:n5])
    E += 0.5 * (np.dot(n[b:a], sj[b:a]) - np.dot(n[b:a], dj[b:a]))
    E -= np.dot(n[:b], sk[:b]) + np.dot(n[a:n5], sk[a:n5])
    E -= 0.5 * (np.dot(n[b:a], sk[b:a]) + np.dot(n[b:a], np.diagonal(K_MO)[b:a]))
    return float(E)


def calcocce(gamma, J_MO, K_MO, H_core, p: Pnof) -> float:
    """calcocce (pnof.py:994-1001)."""
    n, _ = occupation(gamma, p)
    cj, ck = coefficients(n, p)
    return calce(n, cj, ck, J_MO, K_MO, H_core, p)


def calcoccg(gamma, J_MO, K_MO, H_core, p: Pnof) -> np.ndarray:
    """calcoccg (pnof.py:1004-1100). MSpin = 0: dn^T (2 H + diag J) + 2 sum' Dcj J - 2 sum' Dck K.
    MSpin != 0: rows [no1, nbeta) and [nalpha, nbf5) only, columns outside S for the
    derivative sums, plus 2 dn sum_S J - dn sum_S K (pnof.py:1027-1098)."""
    n, dn = occupation(gamma, p)
    Dcj, Dck = der_coefficients(n, dn, p)
    idx = np.arange(p.nbf5)
    Dcj[idx, idx, :] = 0.0
    Dck[idx, idx, :] = 0.0
    if p.MSpin == 0:
        grad = dn.T @ (2.0 * H_core + np.diagonal(J_MO))
        grad += 2.0 * np.einsum("ijk,ji->k", Dcj, J_MO)
        grad -= 2.0 * np.einsum("ijk,ji->k", Dck, K_MO)
        return grad
    o0, b, a, n5 = p.no1, p.nbeta, p.nalpha, p.nbf5
    rows = np.zeros(n5, dtype=bool)
    rows[o0:b] = True
    rows[a:n5] = True
    cols = np.ones(n5, dtype=bool)
    cols[b:a] = False
    w = 2.0 * H_core + np.diagonal(J_MO)
    grad = dn[rows].T @ w[rows]
    mask = np.outer(rows, cols)[:, :, None]
    grad += 2.0 * np.einsum("ijk,ji->k", Dcj * mask, J_MO)
    grad -= 2.0 * np.einsum("ijk,ji->k", Dck * mask, K_MO)
    sj = J_MO[b:a, :].sum(axis=0)
    sk = K_MO[b:a, :].sum(axis=0)
    grad += 2.0 * dn[rows].T @ sj[rows]
    grad -= dn[rows].T @ sk[rows]
    return grad
