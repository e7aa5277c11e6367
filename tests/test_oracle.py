"""Pins the FP64 oracle (oracle/jk_oracle.py) without the reference importable.

Each check is independent of the oracle's own formula: explicit index loops,
the 4-index route I = b b^T through the reference's JKj_Full/JKH_MO_Full
formulas (integrals.py:197-205, :332-345), closed forms at permutation C and
naux = 1, C-independent sums, and invariants. Planted mistakes in a copy of the
oracle must fail these checks.
"""
import numpy as np
import pytest

from oracle import jk_oracle as O
from oracle import pnof_oracle as P
from pynof_amd import synth


def tiny(nbf=5, naux=7, nbf5=3, seed=11, lam=0.3):
    b = synth.make_b(nbf, naux, seed, lam)
    C = synth.make_C(nbf, seed + 1)
    H = synth.make_H(nbf, seed + 2)
    return b, C, H, nbf5


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


@pytest.mark.parametrize("nbf,naux,nbf5", [(5, 7, 3), (4, 3, 4), (6, 8, 1), (3, 1, 2)])
def test_loops_match(nbf, naux, nbf5):
    b, C, H, _ = tiny(nbf, naux, nbf5)
    J, K = O.jkj_ri(C, b, nbf5)
    Jl, Kl = O.jkj_loops(C, b, nbf5)
    assert rel(J, Jl) < 1e-13 and rel(K, Kl) < 1e-13
    JM, KM, Hc = O.jkh_mo_ri(C, H, b, nbf5)
    JMl, KMl, Hcl = O.jkh_mo_loops(C, H, b, nbf5)
    assert rel(JM, JMl) < 1e-13 and rel(KM, KMl) < 1e-13 and rel(Hc, Hcl) < 1e-13
    Jf, Kf, JMf, KMf, Hf = O.jk_build(C, H, b, nbf5)
    for a, r in ((Jf, Jl), (Kf, Kl), (JMf, JMl), (KMf, KMl), (Hf, Hcl)):
        assert rel(a, r) < 1e-13


@pytest.mark.parametrize("nbf,naux,nbf5", [(6, 9, 4), (8, 5, 8)])
def test_four_index_route(nbf, naux, nbf5):
    b, C, H, _ = tiny(nbf, naux, nbf5, seed=21)
    I4 = O.four_index(b)
    J, K = O.jkj_ri(C, b, nbf5)
    Jf, Kf = O.jkj_full(C, I4, nbf5)
    assert rel(J, Jf) < 1e-13 and rel(K, Kf) < 1e-13
    JM, KM, Hc = O.jkh_mo_ri(C, H, b, nbf5)
    JMf, KMf, Hf = O.jkh_mo_full(C, H, I4, nbf5)
    assert rel(JM, JMf) < 1e-13 and rel(KM, KMf) < 1e-13 and rel(Hc, Hf) < 1e-13


def test_permutation_closed_form():
    nbf, naux = 6, 5
    b, _, H, _ = tiny(nbf, naux, nbf, seed=31)
    perm = np.roll(np.arange(nbf), 2)  # column q of C is the unit vector e_perm[q]
    C = np.zeros((nbf, nbf))
    C[perm, np.arange(nbf)] = 1.0
    J, K, JM, KM, Hc = O.jk_build(C, H, b, nbf)
    for q in range(nbf):
        a = perm[q]
        assert np.allclose(J[q], np.einsum("l,mnl->mn", b[a, a], b), rtol=0, atol=1e-14)
        assert np.allclose(K[q], b[a] @ b[a].T, rtol=0, atol=1e-14)
        assert abs(Hc[q] - H[a, a]) < 1e-15
        for s in range(nbf):
            c = perm[s]
            assert abs(KM[q, s] - np.sum(b[a, c] ** 2)) < 1e-14
            assert abs(JM[q, s] - np.dot(b[a, a], b[c, c])) < 1e-14


def test_rank_one_naux():
    nbf, nbf5 = 5, 4
    b, C, H, _ = tiny(nbf, 1, nbf5, seed=41)
    v = b[:, :, 0]
    J, K, JM, KM, Hc = O.jk_build(C, H, b, nbf5)
    for q in range(nbf5):
        c = C[:, q]
        assert np.allclose(J[q], (c @ v @ c) * v, rtol=1e-13, atol=1e-15)
        w = v @ c
        assert np.allclose(K[q], np.outer(w, w), rtol=1e-13, atol=1e-15)
        for s in range(nbf5):
            e = C[:, s]
            assert abs(JM[q, s] - (c @ v @ c) * (e @ v @ e)) < 1e-14
            assert abs(KM[q, s] - (c @ v @ e) ** 2) < 1e-14


def test_orbital_invariant_sums():
    nbf, naux = 7, 6
    b, C, H, _ = tiny(nbf, naux, nbf, seed=51)
    J, K, *_ = O.jk_build(C, H, b, nbf)
    sumK = sum(b[l] @ b[l].T for l in range(nbf))
    tr = np.einsum("mml->l", b)
    sumJ = np.einsum("l,mnl->mn", tr, b)
    assert rel(J.sum(0), sumJ) < 1e-12
    assert rel(K.sum(0), sumK) < 1e-12


def test_invariants_water():
    w = synth.workload(0)
    b, C, H, n = synth.make_inputs(w)
    J, K, JM, KM, Hc = O.jk_build(C, H, b, w.nbf5)
    assert np.allclose(J, J.transpose(0, 2, 1), rtol=0, atol=1e-13)
    assert np.allclose(K, K.transpose(0, 2, 1), rtol=0, atol=1e-13)
    assert np.allclose(JM, JM.T, atol=1e-13) and np.allclose(KM, KM.T, atol=1e-13)
    assert np.allclose(np.diag(JM), np.diag(KM), rtol=1e-13)
    assert np.min(np.linalg.eigvalsh(JM)) > -1e-10 * np.max(np.abs(JM))
    Cq = C[:, : w.nbf5]
    JMc = np.einsum("qmn,mp,np->qp", J, Cq, Cq)
    KMc = np.einsum("qmn,mp,np->qp", K, Cq, Cq)
    assert rel(JMc, JM) < 1e-12 and rel(KMc, KM) < 1e-12


def test_energy_forms():
    w = synth.workload(0)
    b, C, H, n = synth.make_inputs(w)
    assert abs(n.sum() - w.ndoc) < 1e-12 and np.all(n > 0) and np.all(n <= 1)
    _, _, JM, KM, Hc = O.jk_build(C, H, b, w.nbf5)
    e = O.energy(n, JM, KM, Hc)
    brute = 2 * sum(n[i] * Hc[i] for i in range(len(n)))
    brute += sum(n[p] * n[q] * (2 * JM[p, q] - KM[p, q]) for p in range(len(n)) for q in range(len(n)))
    assert abs(e - brute) < 1e-11 * abs(brute)
    cj, ck = 2 * np.outer(n, n), np.outer(n, n)
    ec = O.calce_closed_shell(n, cj, ck, JM, KM, Hc, w.ndoc, w.ndoc, w.nbf5)
    assert abs(ec - (e + np.sum((n - n * n) * np.diag(JM)))) < 1e-11 * abs(e)


# ---------------------------------------------------------------- planted mistakes
def _mutant_transposed_ht(C, b, nbf5):
    Cq = C[:, :nbf5]
    X = np.tensordot(C.T[:, :nbf5], b, axes=([0], [0]))  # half-transform with a transposed C
    d = np.einsum("nq,qnl->ql", Cq, X)
    J = np.tensordot(d, b, axes=([1], [2]))
    K = np.einsum("qml,qnl->qmn", X, X)
    return J, K


def _mutant_kmo_offdiag(C, H, b, nbf5):
    JM, KM, Hc = O.jkh_mo_ri(C, H, b, nbf5)
    Cq = C[:, :nbf5]
    bp = np.einsum("nq,pnl->pql", Cq, np.tensordot(Cq, b, axes=([0], [0])))
    JM = np.einsum("pql,qpl->pq", bp, bp)  # exchange-like instead of diagonal products
    return JM, KM, Hc


def _mutant_dropped_sqrt2(b):
    return b * np.sqrt(2.0)


def test_planted_mistakes_fail():
    b, C, H, nbf5 = tiny(5, 6, 4, seed=61)
    Jl, Kl = O.jkj_loops(C, b, nbf5)
    Jm, Km = _mutant_transposed_ht(C, b, nbf5)
    assert rel(Jm, Jl) > 1e-3 or rel(Km, Kl) > 1e-3
    JMl, KMl, _ = O.jkh_mo_loops(C, H, b, nbf5)
    JMm, _, _ = _mutant_kmo_offdiag(C, H, b, nbf5)
    assert rel(JMm, JMl) > 1e-3
    # a wrongly scaled b breaks the 4-index identity against the unscaled I
    I4 = O.four_index(b)
    Jw, _ = O.jkj_ri(C, _mutant_dropped_sqrt2(b), nbf5)
    Jf, _ = O.jkj_full(C, I4, nbf5)
    assert rel(Jw, Jf) > 0.5


# ---- orbital-gradient MO integrals (Integrals_MO_RI_CPU, integrals.py:703-707)

@pytest.mark.parametrize("nbf,naux,nbf5", [(5, 7, 3), (4, 3, 4), (6, 2, 1)])
def test_mo_integrals_loops(nbf, naux, nbf5):
    b, C, H, _ = tiny(nbf, naux, nbf5, seed=71)
    Hm, bm = O.integrals_mo_ri(C, H, b, nbf5)
    Hl, bl = O.integrals_mo_loops(C, H, b, nbf5)
    assert Hm.shape == (nbf, nbf5) and bm.shape == (nbf, nbf5, naux)
    assert rel(Hm, Hl) < 1e-13 and rel(bm, bl) < 1e-13


def test_mo_integrals_four_index_route():
    nbf, naux, nbf5 = 6, 8, 4
    b, C, H, _ = tiny(nbf, naux, nbf5, seed=72)
    Hm, bm = O.integrals_mo_ri(C, H, b, nbf5)
    Hf, I_MO = O.integrals_mo_full(C, H, O.four_index(b), nbf5)
    assert rel(Hm, Hf) < 1e-13
    # I_MO[p,q,r,t] = sum_P b_MO[p,q,P] b_MO[r,t,P] for r, t < nbf5
    assert rel(np.einsum("pqP,rtP->pqrt", bm, bm[:nbf5]), I_MO) < 1e-12


def test_mo_integrals_permutation_closed_form():
    nbf, naux, nbf5 = 6, 5, 4
    b, _, H, _ = tiny(nbf, naux, nbf5, seed=73)
    perm = np.roll(np.arange(nbf), 3)  # column p of C is e_perm[p]
    C = np.zeros((nbf, nbf))
    C[perm, np.arange(nbf)] = 1.0
    Hm, bm = O.integrals_mo_ri(C, H, b, nbf5)
    for p in range(nbf):
        for q in range(nbf5):
            assert np.array_equal(bm[p, q], b[perm[p], perm[q]])
            assert Hm[p, q] == H[perm[p], perm[q]]


def test_mo_integrals_reproduce_build_outputs():
    """The nbf5 x nbf5 block of b_MO is Y of the JK build: J_MO = sum_P b_pp b_qq,
    K_MO = sum_P b_pq^2, H_core = diag H_mat; that block is symmetric in (p, q)."""
    b, C, H, nbf5 = tiny(7, 9, 5, seed=74)
    Hm, bm = O.integrals_mo_ri(C, H, b, nbf5)
    JM, KM, Hc = O.jkh_mo_loops(C, H, b, nbf5)
    blk = bm[:nbf5]
    dg = np.einsum("ppP->pP", blk)
    assert rel(dg @ dg.T, JM) < 1e-13
    assert rel(np.einsum("pqP,pqP->pq", blk, blk), KM) < 1e-13
    assert rel(np.diag(Hm[:nbf5]), Hc) < 1e-13
    assert rel(blk, blk.transpose(1, 0, 2)) < 1e-13


def test_mo_integrals_planted_mistakes_fail():
    b, C, H, nbf5 = tiny(5, 6, 3, seed=75)
    Hl, bl = O.integrals_mo_loops(C, H, b, nbf5)
    Cq = C[:, :nbf5]
    # C instead of C^T on the first index
    wrong = np.einsum("pm,nq,mnl->pql", C, Cq, b)
    assert rel(wrong, bl) > 1e-3
    # C transposed on the right of H_mat: C[j, n] instead of C[n, j]
    wrong_h = np.einsum("mi,mn,jn->ij", C, H, C[:nbf5, :])
    assert rel(wrong_h, Hl) > 1e-3


# ---- orbital-gradient Lagrangian (computeLagrange2 with RI, utils.py:262-355)

def lag_inputs(nbf=6, naux=7, nbf5=4, seed=81):
    b, C, H, _ = tiny(nbf, naux, nbf5, seed=seed)
    rng = np.random.default_rng(seed + 3)
    n = rng.uniform(0.05, 0.95, nbf5)
    cj = rng.standard_normal((nbf5, nbf5))
    ck = rng.standard_normal((nbf5, nbf5))
    return b, C, H, n, cj + cj.T, ck + ck.T


def fd_gradient(n, cj, ck, C, H, b, nbf5, nbeta, nalpha, eps=1e-5):
    def E(Cx):
        JM, KM, Hc = O.jkh_mo_ri(Cx, H, b, nbf5)
        return O.calce_closed_shell(n, cj, ck, JM, KM, Hc, nbeta, nalpha, nbf5)
    nv = sum(C.shape[0] - i - 1 for i in range(nbf5))
    g = np.zeros(nv)
    for m in range(nv):
        y = np.zeros(nv)
        y[m] = eps
        g[m] = (E(O.rotate_orbital(y, C, nbf5)) - E(O.rotate_orbital(-y, C, nbf5))) / (2 * eps)
    return g


@pytest.mark.parametrize("nbeta,nalpha", [(2, 2), (1, 3), (4, 4)])
def test_lagrangian_loops_and_energy_gradient(nbeta, nalpha):
    """elag matches explicit loops, and 4 (elag - elag^T) is the derivative of the
    closed-shell energy (pnof.py:905-922) under C -> C exp(kappa) (utils.py:550-565):
    an independent check of every term and sign."""
    b, C, H, n, cj, ck = lag_inputs()
    nbf5 = 4
    el, Hm = O.lagrange2_ri(n, cj, ck, C, H, b, nbf5, nbeta, nalpha)
    ell, _ = O.lagrange2_loops(n, cj, ck, C, H, b, nbf5, nbeta, nalpha)
    assert rel(el, ell) < 1e-13
    assert np.all(el[:, nbf5:] == 0.0)
    g = O.orbital_gradient(el, nbf5)
    fd = fd_gradient(n, cj, ck, C, H, b, nbf5, nbeta, nalpha)
    assert np.max(np.abs(g - fd)) < 1e-7 * max(1.0, np.max(np.abs(g)))


def test_lagrangian_does_not_mutate_and_open_shell_is_zero():
    b, C, H, n, cj, ck = lag_inputs(seed=85)
    cj0, ck0 = cj.copy(), ck.copy()
    el, Hm = O.lagrange2_ri(n, cj, ck, C, H, b, 4, 2, 2, MSpin=1)
    assert np.array_equal(cj, cj0) and np.array_equal(ck, ck0)
    assert np.all(el == 0.0) and Hm.shape == (6, 4)


def test_lagrangian_planted_mistakes_fail():
    b, C, H, n, cj, ck = lag_inputs(seed=86)
    nbf5, nbeta, nalpha = 4, 2, 2
    fd = fd_gradient(n, cj, ck, C, H, b, nbf5, nbeta, nalpha)
    Hm, B = O.integrals_mo_ri(C, H, b, nbf5)
    el, _ = O.lagrange2_ri(n, cj, ck, C, H, b, nbf5, nbeta, nalpha)
    # the exchange term with the wrong sign
    cjz, ckz = cj.copy(), ck.copy()
    np.fill_diagonal(cjz, 0.0)
    np.fill_diagonal(ckz, 0.0)
    xterm = np.einsum("bq,aqk,bqk->ab", ckz, B[:, :nbf5], B[:nbf5, :nbf5])
    wrong = el.copy()
    wrong[:, :nbf5] += 2.0 * xterm
    assert np.max(np.abs(O.orbital_gradient(wrong, nbf5) - fd)) > 1e-3
    # the dJ_pp term dropped
    djj = np.einsum("b,abk,bbk->ab", n[:nbeta], B[:, :nbeta], B[:nbeta, :nbeta])
    wrong2 = el.copy()
    wrong2[:, :nbeta] -= djj
    assert np.max(np.abs(O.orbital_gradient(wrong2, nbf5) - fd)) > 1e-3


# ---- generalised Fock and Lagrange multipliers (computeF_RC_CPU, computeLagrange)

@pytest.mark.parametrize("nbeta,nalpha,no1", [(2, 2, 0), (1, 3, 0), (3, 3, 2)])
def test_fock_loops(nbeta, nalpha, no1):
    b, C, H, n, cj, ck = lag_inputs(seed=90)
    nbf5 = 4
    J, K = O.jkj_ri(C, b, nbf5)
    F = O.fock_rc(J, K, n, H, cj, ck, nbeta, nalpha, no1)
    Fl = O.fock_loops(J, K, n, H, cj, ck, nbeta, nalpha, no1)
    assert F.shape == (nbf5, 6, 6) and rel(F, Fl) < 1e-13


@pytest.mark.parametrize("nbeta,nalpha", [(2, 2), (1, 3)])
def test_fock_lagrangian_equals_mo_lagrangian(nbeta, nalpha):
    """The Lagrangian from the Fock matrices (utils.py:249-259) equals the one from the
    MO integrals (utils.py:262-355), which the energy's finite-difference gradient pins:
    C_a^T J[j] C_b = sum_P d[j,P] b_MO[a,b,P] and C_a^T K[j] C_b = sum_P b_MO[a,j,P] b_MO[b,j,P]."""
    b, C, H, n, cj, ck = lag_inputs(seed=91)
    nbf5 = 4
    J, K = O.jkj_ri(C, b, nbf5)
    F = O.fock_rc(J, K, n, H, cj, ck, nbeta, nalpha)
    el = O.lagrange_fock(F, C, C.shape[0])
    el2, _ = O.lagrange2_ri(n, cj, ck, C, H, b, nbf5, nbeta, nalpha)
    assert rel(el, el2) < 1e-12
    # noptorb < nbf keeps only the first rows
    el3 = O.lagrange_fock(F, C, 3)
    assert np.array_equal(el3[:3], el[:3]) and np.all(el3[3:] == 0.0)


def test_fock_planted_mistake_fails():
    b, C, H, n, cj, ck = lag_inputs(seed=92)
    nbf5 = 4
    J, K = O.jkj_ri(C, b, nbf5)
    F = O.fock_rc(J, K, n, H, cj, ck, 2, 2)
    el2, _ = O.lagrange2_ri(n, cj, ck, C, H, b, nbf5, 2, 2)
    # C^K applied to J instead of K
    wrong = F + np.einsum("ij,jmn->imn", ck - np.diag(np.diag(ck)), K - J)
    assert rel(O.lagrange_fock(wrong, C, 6), el2) > 1e-3


# ---- restricted open-shell Fock (computeF_RO_CPU, utils.py:96-171)

def _ro_energy(C, H, b, n, cj, ck, p):
    """The open-shell energy as a function of C: calce (pnof.py:924-988) with its singly
    occupied block replaced by the textbook ROHF one. For parallel-spin singles of n = 1/2,
    every pair contributes J_is - K_is once: sum_{i < s in S} (J_is - K_is). pnof.py:976-988
    instead carries half of that pair sum and a self-exchange -n_i K_ii; the Fock route of
    utils.py:155-169 follows the textbook block."""
    JM, KM, Hc = O.jkh_mo_ri(C, H, b, p.nbf5)
    e = P.calce(n, cj, ck, JM, KM, Hc, p)
    S = slice(p.nbeta, p.nalpha)
    ns, js, ks = n[S], JM[S, S], KM[S, S]
    ref_block = 0.5 * (ns @ js.sum(axis=0) - ns @ np.diagonal(js)) - 0.5 * (ns @ ks.sum(axis=0) + ns @ np.diagonal(ks))
    m = js - ks
    np.fill_diagonal(m, 0.0)
    return e - ref_block + 0.5 * m.sum()


def _ro_check(F, C, H, b, n, cj, ck, p, eps=1e-5):
    """max |dE/dC - 4 F_i C_i| / max |dE/dC| over the first nbf5 columns (central differences)."""
    nbf, n5 = C.shape[0], p.nbf5
    G = np.einsum("imn,ni->mi", F, C[:, :n5])
    dE = np.zeros((nbf, n5))
    for m in range(nbf):
        for i in range(n5):
            Cp, Cm = C.copy(), C.copy()
            Cp[m, i] += eps
            Cm[m, i] -= eps
            dE[m, i] = (_ro_energy(Cp, H, b, n, cj, ck, p) - _ro_energy(Cm, H, b, n, cj, ck, p)) / (2 * eps)
    return np.max(np.abs(dE - 4.0 * G)) / np.max(np.abs(dE))


def _ro_inputs(p, seed, pnof_coefficients):
    b, C, H, _ = tiny(10, 11, p.nbf5, seed=seed)
    rng = np.random.default_rng(seed + 5)
    if pnof_coefficients:  # the functional's own coefficients (core orbitals at n = 1)
        n, _ = P.occupation(rng.standard_normal(p.nv), p)
        cj, ck = P.coefficients(n, p)
    else:  # any symmetric coefficients, singles at n = 1/2
        n = rng.uniform(0.05, 0.95, p.nbf5)
        n[p.nbeta:p.nalpha] = 0.5
        cj, ck = rng.standard_normal((2, p.nbf5, p.nbf5))
        cj, ck = cj + cj.T, ck + ck.T
    return b, C, H, n, cj, ck


RO_LAYOUTS = [  # (no1, ndoc, nsoc, ncwo, pnof coefficients)
    (0, 2, 2, 1, False), (0, 2, 1, 2, False), (1, 1, 3, 1, False), (0, 3, 2, 1, True),
    (2, 2, 1, 2, True), (3, 1, 2, 2, True)]


@pytest.mark.parametrize("no1,ndoc,nsoc,ncwo,pc", RO_LAYOUTS)
def test_fock_ro_is_the_energy_derivative(no1, ndoc, nsoc, ncwo, pc):
    """G = F_i C_i (utils.py:251) is a quarter of the C-gradient of the open-shell energy
    above: every coefficient of every row (D, S and W) is pinned by central differences."""
    p = P.Pnof(no1=no1, ndoc=ndoc, nsoc=nsoc, ncwo=ncwo, ipnof=8, MSpin=nsoc, occ_method="Trigonometric")
    b, C, H, n, cj, ck = _ro_inputs(p, 120 + no1 + 7 * nsoc, pc)
    J, K = O.jkj_ri(C, b, p.nbf5)
    F = O.fock_ro(J, K, n, H, cj, ck, p.nbeta, p.nalpha, no1)
    assert _ro_check(F, C, H, b, n, cj, ck, p) < 1e-7


@pytest.mark.parametrize("no1,ndoc,nsoc,ncwo,pc", RO_LAYOUTS)
def test_fock_ro_loops(no1, ndoc, nsoc, ncwo, pc):
    """The coefficient of J[j], K[j] per (i, j): this also pins each single's own SUMij term,
    which removes (J[s] - K[s]) / 2 from F[s] (utils.py:159) and which the gradient cannot
    see, since (J[s] - K[s]) C_s = 0."""
    p = P.Pnof(no1=no1, ndoc=ndoc, nsoc=nsoc, ncwo=ncwo, ipnof=8, MSpin=nsoc, occ_method="Trigonometric")
    b, C, H, n, cj, ck = _ro_inputs(p, 140 + no1, pc)
    J, K = O.jkj_ri(C, b, p.nbf5)
    F = O.fock_ro(J, K, n, H, cj, ck, p.nbeta, p.nalpha, no1)
    assert rel(F, O.fock_ro_loops(J, K, n, H, cj, ck, p.nbeta, p.nalpha, no1)) < 1e-13


def test_fock_ro_without_singles_is_fock_rc():
    b, C, H, n, cj, ck = lag_inputs(seed=93)
    J, K = O.jkj_ri(C, b, 4)
    for no1 in (0, 2):
        assert rel(O.fock_ro(J, K, n, H, cj, ck, 3, 3, no1), O.fock_rc(J, K, n, H, cj, ck, 3, 3, no1)) < 1e-14


def test_fock_ro_planted_mistakes_fail():
    p = P.Pnof(no1=0, ndoc=2, nsoc=2, ncwo=1, ipnof=8, MSpin=2, occ_method="Trigonometric")
    b, C, H, n, cj, ck = _ro_inputs(p, 130, False)
    J, K = O.jkj_ri(C, b, p.nbf5)
    F = O.fock_ro(J, K, n, H, cj, ck, p.nbeta, p.nalpha)
    S = slice(p.nbeta, p.nalpha)
    # PRODWROij with K at full weight
    dw = np.r_[0:p.nbeta, p.nalpha:p.nbf5]
    w1 = F.copy()
    w1[S] -= 0.5 * np.einsum("j,jmn->mn", n[dw], K[dw])
    # SUMij on the doubly occupied rows with K at full weight
    w2 = F.copy()
    w2[:p.nbeta] -= 0.5 * np.einsum("i,jmn->imn", n[:p.nbeta], K[S])
    # the closed-shell form (C^J/C^K over the singles too, no SUMij/PRODWROij)
    w3 = O.fock_rc(J, K, n, H, cj, ck, p.nbeta, p.nalpha)
    for w in (w1, w2, w3):
        assert _ro_check(w, C, H, b, n, cj, ck, p) > 1e-3


# ---------------------------------------------------------------- b producer (integrals.py:35-49)
def _three_center(nbf, naux, seed):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((nbf, nbf, naux))
    return 0.5 * (m + m.transpose(1, 0, 2))


def test_b_from_3c_closed_forms():
    mnk = _three_center(5, 7, 900)
    assert np.allclose(O.b_from_3c(mnk, np.eye(7)), mnk, rtol=0, atol=1e-14)
    d = np.linspace(0.3, 4.0, 7)
    assert np.allclose(O.b_from_3c(mnk, np.diag(d)), mnk / np.sqrt(d), rtol=1e-14, atol=1e-14)


def test_b_from_3c_reproduces_ri_integrals():
    """sum_P b[mn,P] b[ls,P] = sum_PQ (mn|P) [(P|Q)^-1]_PQ (Q|ls), by a linear solve (no eigenpairs)."""
    nbf, naux = 6, 11
    mnk = _three_center(nbf, naux, 901)
    rng = np.random.default_rng(902)
    A = rng.standard_normal((naux, naux))
    metric = A @ A.T + naux * np.eye(naux)
    b = O.b_from_3c(mnk, metric)
    lhs = np.einsum("mnP,lsP->mnls", b, b)
    x = np.linalg.solve(metric, mnk.reshape(-1, naux).T)  # (naux, nbf^2)
    rhs = (mnk.reshape(-1, naux) @ x).reshape(nbf, nbf, nbf, nbf)
    assert np.max(np.abs(lhs - rhs)) < 1e-12 * np.max(np.abs(rhs))
    assert np.array_equal(b, b.transpose(1, 0, 2)) or np.max(np.abs(b - b.transpose(1, 0, 2))) < 1e-14
    bad = np.einsum("mnk,kl->mnl", mnk, O.metric_power(metric, +0.5))  # planted: wrong power
    assert np.max(np.abs(np.einsum("mnP,lsP->mnls", bad, bad) - rhs)) > 1e-3


def test_b_from_3c_drops_tiny_eigenvalues():
    """A metric built as V diag(a) V^T with one eigenvalue below 1e-14 max|a|: that direction
    is dropped, b = mnk V diag(a'^-1/2) V^T with a'^-1/2 = 0 there."""
    naux = 8
    rng = np.random.default_rng(903)
    V, _ = np.linalg.qr(rng.standard_normal((naux, naux)))
    a = np.linspace(1.0, 5.0, naux)
    a[2] = 1e-16 * 5.0
    metric = (V * a) @ V.T
    f = 1.0 / np.sqrt(a)
    f[2] = 0.0
    mnk = _three_center(4, naux, 904)
    expect = np.einsum("mnk,kl->mnl", mnk, (V * f) @ V.T)
    got = O.b_from_3c(mnk, metric)
    assert np.max(np.abs(got - expect)) < 1e-12 * np.max(np.abs(expect))
