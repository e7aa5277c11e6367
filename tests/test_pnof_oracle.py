"""Pins the occupation-step oracle (oracle/pnof_oracle.py) without the reference:
sum rules and finite differences of the occupations, the Hartree-Fock limit, the
APSG energy from two-electron Slater rules (PNOF5), the published PNOF7 inter-pair
term, and finite differences of the energy for the occupation gradient."""
import numpy as np
import pytest

from oracle import jk_oracle as JO
from oracle import pnof_oracle as O


def mo_integrals(nbf5, naux, seed):
    """J_MO, K_MO, H_core and the 4-index (pq|rs) of a random symmetric B[p, q, P]."""
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((nbf5, nbf5, naux)) * 0.3
    B = 0.5 * (B + B.transpose(1, 0, 2))
    eri = np.einsum("pqP,rsP->pqrs", B, B)
    J = np.einsum("ppP,qqP->pq", B, B)
    K = np.einsum("pqP,pqP->pq", B, B)
    h = rng.standard_normal((nbf5, nbf5))
    h = 0.5 * (h + h.T)
    return J, K, np.diag(h).copy(), eri


def random_n(p, seed):
    """Valid occupations: n = 1 below no1, each subspace sums to 1, singly 0.5."""
    rng = np.random.default_rng(seed)
    n = np.zeros(p.nbf5)
    n[:p.no1] = 1.0
    n[p.nbeta:p.nalpha] = 0.5
    for l in range(p.ndoc):
        w = O.weak_of(p, l)
        e = np.exp(rng.normal(-2.0, 0.7, len(w) + 1))
        e[0] = np.exp(rng.uniform(0.5, 2.0))
        e /= e.sum()
        n[p.no1 + l] = e[0]
        n[w] = e[1:]
    return n


CONFIGS = [(0, 1, 0, 3), (0, 2, 0, 2), (0, 3, 0, 3), (1, 2, 0, 2), (0, 2, 2, 2), (2, 3, 1, 2), (0, 2, 0, 1)]


@pytest.mark.parametrize("no1,ndoc,nsoc,ncwo", CONFIGS)
@pytest.mark.parametrize("method", ["Softmax", "Trigonometric"])
def test_occupation_sum_rules_and_derivative(no1, ndoc, nsoc, ncwo, method):
    p = O.Pnof(no1, ndoc, nsoc, ncwo, occ_method=method)
    rng = np.random.default_rng(10 + no1 + 3 * ndoc + 7 * ncwo)
    g = rng.uniform(-1.0, 1.5, p.nv)
    n, dn = O.occupation(g, p)
    assert n.shape == (p.nbf5,) and dn.shape == (p.nbf5, p.nv)
    assert np.all(n[p.nbeta:p.nalpha] == 0.5)
    for l in range(ndoc):
        # strong orbital: index l for Softmax (the reference's own indexing), no1 + l for Trigonometric
        s = l if method == "Softmax" else no1 + l
        assert abs(n[s] + n[O.weak_of(p, l)].sum() - 1.0) < 1e-14
    if method == "Trigonometric":
        assert np.all(n[:no1] == 1.0)
        assert np.all((n[no1:] > 0) & (n[no1:] < 1))
    eps = 1e-6
    fd = np.zeros_like(dn)
    for k in range(p.nv):
        gp, gm = g.copy(), g.copy()
        gp[k] += eps
        gm[k] -= eps
        fd[:, k] = (O.occupation(gp, p)[0] - O.occupation(gm, p)[0]) / (2 * eps)
    assert np.max(np.abs(fd - dn)) < 1e-8


@pytest.mark.parametrize("ipnof,ista", [(5, 0), (7, 0), (7, 1), (8, 0)])
@pytest.mark.parametrize("no1,ndoc,ncwo", [(0, 2, 2), (2, 3, 2), (1, 1, 3)])
def test_hartree_fock_limit(ipnof, ista, no1, ndoc, ncwo):
    """n in {0, 1}: every PNOF gives 2 sum_i h_i + sum_ij (2 J_ij - K_ij) over the occupied."""
    p = O.Pnof(no1, ndoc, 0, ncwo, ipnof=ipnof, ista=ista)
    J, K, hc, _ = mo_integrals(p.nbf5, 9, 20 + ipnof)
    n = np.zeros(p.nbf5)
    n[:p.nbeta] = 1.0
    cj, ck = O.coefficients(n, p)
    occ = slice(0, p.nbeta)
    e_hf = 2.0 * hc[occ].sum() + np.sum(2.0 * J[occ, occ] - K[occ, occ])
    assert abs(O.calce(n, cj, ck, J, K, hc, p) - e_hf) < 1e-12 * max(1.0, abs(e_hf))


def apsg_energy(n, p, hc, eri):
    """Energy of an antisymmetrised product of strongly orthogonal geminals with
    coefficients +sqrt(n) (strong, inactive) and -sqrt(n) (weak), from the two-electron
    Slater rules: <p a p b|H|q a q b> = delta_pq (2 h_p + (pp|pp)) + (1 - delta_pq)(pq|pq);
    between geminals, the 1-RDM product n_p n_q (2 (pp|qq) - (pq|qp))."""
    gem = [[i] for i in range(p.no1)]
    sign = {i: 1.0 for i in range(p.no1)}
    for l in range(p.ndoc):
        s = p.no1 + l
        w = list(O.weak_of(p, l))
        gem.append([s] + w)
        sign[s] = 1.0
        for x in w:
            sign[x] = -1.0
    E = 0.0
    for g in gem:
        c = np.array([sign[i] * np.sqrt(n[i]) for i in g])
        M = np.empty((len(g), len(g)))
        for a, pa in enumerate(g):
            for b, qb in enumerate(g):
                M[a, b] = 2.0 * hc[pa] + eri[pa, pa, pa, pa] if a == b else eri[pa, qb, pa, qb]
        E += c @ M @ c
    for gi, g in enumerate(gem):
        for fi, f in enumerate(gem):
            if gi != fi:
                for a in g:
                    for b in f:
                        E += n[a] * n[b] * (2.0 * eri[a, a, b, b] - eri[a, b, b, a])
    return E


@pytest.mark.parametrize("no1,ndoc,ncwo", [(0, 1, 3), (0, 2, 2), (2, 3, 2), (1, 2, 1)])
def test_pnof5_is_apsg(no1, ndoc, ncwo):
    p = O.Pnof(no1, ndoc, 0, ncwo, ipnof=5)
    J, K, hc, eri = mo_integrals(p.nbf5, 11, 30 + ndoc)
    n = random_n(p, 31 + ncwo)
    cj, ck = O.coefficients(n, p)
    e = O.calce(n, cj, ck, J, K, hc, p)
    assert abs(e - apsg_energy(n, p, hc, eri)) < 1e-12 * max(1.0, abs(e))
    # the same through the closed-shell energy of the JK oracle
    assert abs(e - JO.calce_closed_shell(n, cj, ck, J, K, hc, p.nbeta, p.nalpha, p.nbf5)) < 1e-13 * max(1.0, abs(e))


@pytest.mark.parametrize("ista", [0, 1])
@pytest.mark.parametrize("no1,ndoc,ncwo", [(0, 2, 2), (1, 3, 2)])
def test_pnof7_minus_pnof5(ista, no1, ndoc, ncwo):
    """E(PNOF7) - E(PNOF5) = -sum_{p, q in different subspaces} Phi_p Phi_q K_pq."""
    p5 = O.Pnof(no1, ndoc, 0, ncwo, ipnof=5)
    p7 = O.Pnof(no1, ndoc, 0, ncwo, ipnof=7, ista=ista)
    J, K, hc, _ = mo_integrals(p5.nbf5, 7, 40 + ndoc)
    n = random_n(p5, 41)
    e5 = O.calce(n, *O.coefficients(n, p5), J, K, hc, p5)
    e7 = O.calce(n, *O.coefficients(n, p7), J, K, hc, p7)
    phi = np.sqrt(n * (1 - n)) if ista == 0 else 2 * n * (1 - n)
    sub = -np.ones(p5.nbf5, dtype=int)  # inactive orbitals: a subspace each
    sub[:no1] = 100 + np.arange(no1)
    for l in range(ndoc):
        sub[no1 + l] = l
        sub[O.weak_of(p5, l)] = l
    diff = sub[:, None] != sub[None, :]
    assert abs((e7 - e5) + np.sum(np.outer(phi, phi) * K * diff)) < 1e-12


@pytest.mark.parametrize("ipnof", [5, 7, 8])
@pytest.mark.parametrize("no1,ndoc,nsoc,ncwo", CONFIGS)
def test_coefficients_symmetric(ipnof, no1, ndoc, nsoc, ncwo):
    p = O.Pnof(no1, ndoc, nsoc, ncwo, ipnof=ipnof)
    n = random_n(p, 50)
    cj, ck = O.coefficients(n, p)
    assert np.array_equal(cj, cj.T)
    # GNOF's open-shell blocks add 0.25 Phi_o both ways but 0.5 Phi_v only in one
    # orientation each, which is symmetric as a whole
    assert np.max(np.abs(ck - ck.T)) < 1e-15


def fd_grad(g, J, K, hc, p, eps=1e-5):
    out = np.zeros(p.nv)
    for k in range(p.nv):
        gp, gm = g.copy(), g.copy()
        gp[k] += eps
        gm[k] -= eps
        out[k] = (O.calcocce(gp, J, K, hc, p) - O.calcocce(gm, J, K, hc, p)) / (2 * eps)
    return out


# The reference's softmax writes the strong occupations at index i < ndoc, so for no1 > 0
# an inactive orbital's n moves with x while the Phi derivative skips every orbital below
# no1 (pnof.py:301, :460): there its PNOF7/GNOF gradient is not the derivative of its
# energy. The oracle restates that as it is; test_softmax_no1_quirk pins it, and those
# cases are left out of the finite-difference check.
GRAD_CASES = [(ip, st, m, c) for ip, st in [(5, 0), (7, 0), (7, 1), (8, 0)] for m in ["Softmax", "Trigonometric"]
              for c in CONFIGS if not (m == "Softmax" and c[0] > 0 and ip != 5)]


@pytest.mark.parametrize("ipnof,ista,method,cfg", GRAD_CASES)
def test_gradient_matches_energy_differences(ipnof, ista, method, cfg):
    no1, ndoc, nsoc, ncwo = cfg
    p = O.Pnof(no1, ndoc, nsoc, ncwo, ipnof=ipnof, ista=ista, occ_method=method)
    J, K, hc, _ = mo_integrals(p.nbf5, 8, 60 + ipnof + ndoc)
    rng = np.random.default_rng(61 + ncwo)
    g = rng.uniform(-0.8, 1.2, p.nv)
    grad = O.calcoccg(g, J, K, hc, p)
    fd = fd_grad(g, J, K, hc, p)
    assert np.max(np.abs(grad - fd)) < 1e-7 * max(1.0, np.max(np.abs(fd)))


def test_softmax_no1_quirk():
    """Softmax with no1 > 0: strong occupations land at indices 0..ndoc-1, n[no1 + l] of
    the later strong orbitals stays 0, and the PNOF7 gradient misses exactly the Phi
    derivative of the orbitals below no1 (pnof.py:301 loops from no1)."""
    p = O.Pnof(1, 2, 0, 2, ipnof=7, occ_method="Softmax")
    g = np.random.default_rng(80).uniform(-0.8, 1.2, p.nv)
    n, dn = O.occupation(g, p)
    assert n[0] > 0 and n[p.no1 + p.ndoc - 1] == 0.0 and np.any(dn[0] != 0)
    J, K, hc, _ = mo_integrals(p.nbf5, 8, 81)
    fi = np.sqrt(n * (1 - n))
    dfi0 = (1.0 - 2.0 * n[0]) / (2.0 * fi[0]) * dn[0]
    sub = np.full(p.nbf5, -1)
    for l in range(p.ndoc):
        sub[p.no1 + l] = l
        sub[O.weak_of(p, l)] = l
    # the missing term: 2 sum_j (d Phi_0) Phi_j K_j0 over j outside orbital 0's subspace
    other = np.array([j for j in range(p.nbf5) if j != 0 and (sub[j] != sub[0] or sub[0] < 0)])
    missing = -2.0 * dfi0 * np.sum(fi[other] * K[other, 0])
    fd = fd_grad(g, J, K, hc, p)
    assert np.max(np.abs(O.calcoccg(g, J, K, hc, p) + missing - fd)) < 1e-7


def test_planted_mistakes_fail(monkeypatch):
    """A flipped weak-weak sign fails the APSG check; a GNOF Phi derivative with a
    dropped factor 1/2 fails the finite-difference check."""
    p = O.Pnof(0, 2, 0, 2, ipnof=5)
    J, K, hc, eri = mo_integrals(p.nbf5, 11, 70)
    n = random_n(p, 71)
    good = O._intra

    def flipped(cj, ck, n_, p_):
        good(cj, ck, n_, p_)
        for l in range(p_.ndoc):
            w = O.weak_of(p_, l)
            ck[np.ix_(w, w)] *= -1.0

    monkeypatch.setattr(O, "_intra", flipped)
    e = O.calce(n, *O.coefficients(n, p), J, K, hc, p)
    assert abs(e - apsg_energy(n, p, hc, eri)) > 1e-6
    monkeypatch.setattr(O, "_intra", good)

    p8 = O.Pnof(0, 2, 0, 2, ipnof=8)
    J, K, hc, _ = mo_integrals(p8.nbf5, 8, 72)
    g = np.random.default_rng(73).uniform(-0.8, 1.2, p8.nv)
    assert np.max(np.abs(O.calcoccg(g, J, K, hc, p8) - fd_grad(g, J, K, hc, p8))) < 1e-7
    real = O._dphi

    def doubled(n_, dn_, p_, ista):
        fi, d = real(n_, dn_, p_, ista)
        return fi, 2.0 * d

    monkeypatch.setattr(O, "_dphi", doubled)
    assert np.max(np.abs(O.calcoccg(g, J, K, hc, p8) - fd_grad(g, J, K, hc, p8))) > 1e-6


def calce_mspin_loops(n, cj, ck, J, K, hc, p):
    """calce with MSpin != 0 by explicit loops over the reference's blocks (pnof.py:924-988)."""
    b, a, n5 = p.nbeta, p.nalpha, p.nbf5
    E = 0.0
    for i in range(n5):
        E += n[i] * 2.0 * hc[i]
        if i < b or i >= a:
            E += n[i] * J[i, i]
    for i in range(n5):
        for j in range(n5):
            if i != j and not (b <= i < a) and not (b <= j < a):
                E += cj[i, j] * J[j, i] - ck[i, j] * K[j, i]
    for i in list(range(b)) + list(range(a, n5)):
        for s in range(b, a):
            E += 2.0 * n[i] * J[s, i] - n[i] * K[s, i]
    for i in range(b, a):
        for s in range(b, a):
            E += 0.5 * n[i] * J[s, i] - 0.5 * n[i] * K[s, i]
        E += -0.5 * n[i] * J[i, i] - 0.5 * n[i] * K[i, i]
    return E


@pytest.mark.parametrize("ipnof", [5, 7, 8])
@pytest.mark.parametrize("no1,ndoc,nsoc,ncwo", [(0, 2, 1, 2), (1, 2, 2, 2), (0, 3, 3, 1)])
def test_open_shell_energy_loops(ipnof, no1, ndoc, nsoc, ncwo):
    p = O.Pnof(no1, ndoc, nsoc, ncwo, ipnof=ipnof, MSpin=nsoc)
    J, K, hc, _ = mo_integrals(p.nbf5, 8, 90 + ipnof)
    n = random_n(p, 91)
    cj, ck = O.coefficients(n, p)
    E = O.calce(n, cj, ck, J, K, hc, p)
    assert abs(E - calce_mspin_loops(n, cj, ck, J, K, hc, p)) < 1e-12 * max(1.0, abs(E))
    # the singly occupied block's override (MSpin == 0 only, pnof.py:179-180) is absent:
    # PNOF5 keeps n_s n_s' there (PNOF7's n n + Phi Phi equals 2 n n at n = 1/2)
    p0 = O.Pnof(no1, ndoc, nsoc, ncwo, ipnof=ipnof, MSpin=0)
    cj0, ck0 = O.coefficients(n, p0)
    s = slice(p.nbeta, p.nalpha)
    if nsoc > 1 and ipnof == 5:
        assert np.allclose(ck[s, s], np.outer(n[s], n[s])) and np.allclose(ck0[s, s], 2 * np.outer(n[s], n[s]))


OPEN_CASES = [(ip, st, m, c) for ip, st in [(5, 0), (7, 0), (7, 1), (8, 0)] for m in ["Softmax", "Trigonometric"]
              for c in [(0, 2, 1, 2), (0, 2, 2, 2), (2, 3, 1, 2), (0, 3, 3, 1)]
              if not (m == "Softmax" and c[0] > 0)]  # softmax with no1 > 0: see test_softmax_no1_quirk


@pytest.mark.parametrize("ipnof,ista,method,cfg", OPEN_CASES)
def test_open_shell_gradient_matches_energy_differences(ipnof, ista, method, cfg):
    no1, ndoc, nsoc, ncwo = cfg
    p = O.Pnof(no1, ndoc, nsoc, ncwo, ipnof=ipnof, ista=ista, occ_method=method, MSpin=nsoc)
    J, K, hc, _ = mo_integrals(p.nbf5, 8, 95 + ipnof + ndoc)
    g = np.random.default_rng(96 + ncwo).uniform(-0.8, 1.2, p.nv)
    grad = O.calcoccg(g, J, K, hc, p)
    fd = fd_grad(g, J, K, hc, p)
    assert np.max(np.abs(grad - fd)) < 1e-7 * max(1.0, np.max(np.abs(fd)))
