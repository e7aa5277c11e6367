"""FP64 CPU oracle for the density-fitted JK build.

TEST INFRASTRUCTURE ONLY. Nothing in the product imports this module: it is
used by ``tests/``, by ``__graft_entry__.smoke()`` as the checker and by
``bench.py``'s CPU-baseline leg. It is a restatement, in numpy, of the
reference's RI contractions, written from the mathematics; every function
cites the reference lines it follows (paths relative to the reference root).

Pinning (see tests/test_oracle.py): every function below is checked against
(1) explicit index loops on tiny inputs, (2) the independent 4-index route
I = b b^T through the reference's ``JKj_Full``/``JKH_MO_Full`` formulas,
(3) closed forms at C = identity and naux = 1, (4) invariants (symmetry,
diag K_MO = diag J_MO, J_MO PSD, contraction of J^n/K^n with D_p reproducing
J_MO/K_MO, orbital invariance of sum_q J^q and sum_q K^q, the MO integrals'
nbf5 x nbf5 block reproducing J_MO, K_MO and H_core), and (5) the
committed golden fixtures of the water-sized workload. The reference ships
no tests or stored outputs, so no reference-held vector exists to pin against.

The b producer (``metric_power``, ``b_from_3c``; integrals.py:35-49) is pinned by
closed forms (identity and diagonal metrics), by the RI identity through a
linear solve, and by a metric with a known dropped eigenvector. The eigenvalue
cut restates psi4's Matrix::power (a dependency the reference does not pin);
with no psi4 output available, the cut rule at its margin is parity unpinned.
"""
from __future__ import annotations

import numpy as np


def jkj_ri(C: np.ndarray, b_mnl: np.ndarray, nbf5: int):
    """J^n, K^n with the reference's call sequence (integrals.py:208-218).

    b_qnl[q,nu,P] = sum_mu C[mu,q] b[mu,nu,P]          (tensordot, :211)
    d[q,P]        = sum_nu C[nu,q] b_qnl[q,nu,P]        (einsum, :212)
    J[q,mu,nu]    = sum_P d[q,P] b[mu,nu,P]             (tensordot, :214)
    K[q,mu,nu]    = sum_P b_qnl[q,mu,P] b_qnl[q,nu,P]   (einsum, :216)
    """
    Cq = C[:, 0:nbf5]
    b_qnl = np.tensordot(Cq, b_mnl, axes=([0], [0]))
    b_qql = np.einsum("nq,qnl->ql", Cq, b_qnl, optimize=True)
    J = np.tensordot(b_qql, b_mnl, axes=([1], [2]))
    K = np.einsum("qml,qnl->qmn", b_qnl, b_qnl, optimize=True)
    return J, K


def jkh_mo_ri(C: np.ndarray, H: np.ndarray, b_mnl: np.ndarray, nbf5: int):
    """J_MO, K_MO, H_core with the reference's call sequence (integrals.py:348-362).

    D[i,mu,nu]   = C[mu,i] C[nu,i]                           (:351)
    b_pql[p,q,P] = sum_nu C[nu,q] sum_mu C[mu,p] b[mu,nu,P]  (:353-354)
    J_MO[p,q]    = sum_P b_ppP b_qqP                         (:356)
    K_MO[p,q]    = sum_P b_pqP^2                             (:358)
    H_core[i]    = sum_{mu nu} D[i,mu,nu] H[mu,nu]           (:360)
    """
    Cq = C[:, 0:nbf5]
    D = np.einsum("mi,ni->imn", Cq, Cq, optimize=True)
    b_pnl = np.tensordot(Cq, b_mnl, axes=([0], [0]))
    b_pql = np.einsum("nq,pnl->pql", Cq, b_pnl, optimize=True)
    J_MO = np.einsum("ppl,qql->pq", b_pql, b_pql, optimize=True)
    K_MO = np.einsum("pql,pql->pq", b_pql, b_pql, optimize=True)
    H_core = np.tensordot(D, H, axes=([1, 2], [0, 1]))
    return J_MO, K_MO, H_core


def jk_build(C, H, b_mnl, nbf5):
    """One fused build: the outputs of one computeJKj plus one computeJKH_MO call,
    with the half-transform done once and H_core as diag(C^T H C) (no D tensor).
    Same mathematics as ``jkj_ri`` + ``jkh_mo_ri``; used where the reference's
    duplicate work would only cost time."""
    Cq = np.ascontiguousarray(C[:, 0:nbf5])
    nbf, _, naux = b_mnl.shape
    X = (Cq.T @ b_mnl.reshape(nbf, nbf * naux)).reshape(nbf5, nbf, naux)
    Y = np.einsum("nq,pnl->pql", Cq, X, optimize=True)
    d = np.ascontiguousarray(np.einsum("ppl->pl", Y))
    J = (d @ b_mnl.reshape(nbf * nbf, naux).T).reshape(nbf5, nbf, nbf)
    K = np.matmul(X, X.transpose(0, 2, 1))
    J_MO = d @ d.T
    K_MO = np.einsum("pql,pql->pq", Y, Y, optimize=True)
    H_core = np.einsum("mi,mi->i", Cq, H @ Cq)
    return J, K, J_MO, K_MO, H_core


def jkj_full(C: np.ndarray, I: np.ndarray, nbf5: int):
    """4-index route (integrals.py:197-205): J = D:I over (2,3), K = D:I over (1,3)."""
    Cq = C[:, 0:nbf5]
    D = np.einsum("mi,ni->imn", Cq, Cq, optimize=True)
    J = np.tensordot(D, I, axes=([1, 2], [2, 3]))
    K = np.tensordot(D, I, axes=([1, 2], [1, 3]))
    return J, K


def jkh_mo_full(C: np.ndarray, H: np.ndarray, I: np.ndarray, nbf5: int):
    """4-index route for the MO quantities (integrals.py:332-345)."""
    Cq = C[:, 0:nbf5]
    D = np.einsum("mi,ni->imn", Cq, Cq, optimize=True)
    J = np.tensordot(D, I, axes=([1, 2], [2, 3]))
    J_MO = np.tensordot(J, D, axes=((1, 2), (1, 2)))
    K = np.tensordot(D, I, axes=([1, 2], [1, 3]))
    K_MO = np.tensordot(K, D, axes=([1, 2], [1, 2]))
    H_core = np.tensordot(D, H, axes=([1, 2], [0, 1]))
    return J_MO, K_MO, H_core


def four_index(b_mnl: np.ndarray) -> np.ndarray:
    """I[mu,nu,lam,sig] = sum_P b[mu,nu,P] b[lam,sig,P] (the RI factorisation, integrals.py:30-49)."""
    nbf, _, naux = b_mnl.shape
    B = b_mnl.reshape(nbf * nbf, naux)
    return (B @ B.T).reshape(nbf, nbf, nbf, nbf)


def jkj_loops(C, b_mnl, nbf5):
    """Explicit loops with the accumulation order of JKj_RI_jit (integrals.py:282-301).
    Pure Python: tiny inputs only."""
    nbf, _, naux = b_mnl.shape
    J = np.zeros((nbf5, nbf, nbf))
    K = np.zeros((nbf5, nbf, nbf))
    for q in range(nbf5):
        bq = np.zeros((nbf, naux))
        dq = np.zeros(naux)
        for l in range(naux):
            for n in range(nbf):
                s = 0.0
                for m in range(nbf):
                    s += C[m, q] * b_mnl[m, n, l]
                bq[n, l] = s
                dq[l] += C[n, q] * bq[n, l]
            for n in range(nbf):
                for m in range(nbf):
                    K[q, m, n] += bq[m, l] * bq[n, l]
                    J[q, m, n] += dq[l] * b_mnl[m, n, l]
    return J, K


def jkh_mo_loops(C, H, b_mnl, nbf5):
    """Explicit loops following JKH_MO_RI_jit (integrals.py:457-484), full q range."""
    nbf, _, naux = b_mnl.shape
    bpq = np.zeros((nbf5, nbf5, naux))
    for p in range(nbf5):
        for n in range(nbf):
            for l in range(naux):
                s = 0.0
                for m in range(nbf):
                    s += C[m, p] * b_mnl[m, n, l]
                for q in range(nbf5):
                    bpq[p, q, l] += C[n, q] * s
    H_core = np.zeros(nbf5)
    J_MO = np.zeros((nbf5, nbf5))
    K_MO = np.zeros((nbf5, nbf5))
    for p in range(nbf5):
        for m in range(nbf):
            for n in range(nbf):
                H_core[p] += C[m, p] * H[m, n] * C[n, p]
        for q in range(nbf5):
            for l in range(naux):
                J_MO[p, q] += bpq[p, p, l] * bpq[q, q, l]
                K_MO[p, q] += bpq[p, q, l] * bpq[p, q, l]
    return J_MO, K_MO, H_core


def integrals_mo_ri(C: np.ndarray, H: np.ndarray, b_mnl: np.ndarray, nbf5: int):
    """(H_mat, b_MO) of the orbital gradient, Integrals_MO_RI_CPU (integrals.py:703-707),
    reached through JKH_MO_tmp (:665-682) from computeLagrange2 (utils.py:262-265):

    H_mat[i,j]   = sum_mn C[m,i] H[m,n] C[n,j]         i < nbf, j < nbf5   (:705)
    b_MO[p,q,P]  = sum_mn C[m,p] C[n,q] b[m,n,P]       p < nbf, q < nbf5   (:706)
    """
    Cq = C[:, 0:nbf5]
    H_mat = np.einsum("mi,mn,nj->ij", C, H, Cq, optimize=True)
    b_MO = np.einsum("mp,nq,mnl->pql", C, Cq, b_mnl, optimize=True)
    return H_mat, b_MO


def integrals_mo_loops(C, H, b_mnl, nbf5):
    """The same (H_mat, b_MO) by explicit loops over every index (tiny inputs only):
    an independent route that pins integrals_mo_ri."""
    nbf, naux = b_mnl.shape[0], b_mnl.shape[2]
    H_mat = np.zeros((nbf, nbf5))
    b_MO = np.zeros((nbf, nbf5, naux))
    for i in range(nbf):
        for j in range(nbf5):
            s = 0.0
            for m in range(nbf):
                for n in range(nbf):
                    s += C[m, i] * H[m, n] * C[n, j]
            H_mat[i, j] = s
            for P in range(naux):
                t = 0.0
                for m in range(nbf):
                    for n in range(nbf):
                        t += C[m, i] * C[n, j] * b_mnl[m, n, P]
                b_MO[i, j, P] = t
    return H_mat, b_MO


def integrals_mo_full(C, H, I, nbf5):
    """Integrals_MO_Full_CPU (integrals.py:685-690): H_mat and the 4-index MO integrals
    I_MO[p,q,r,t] = sum C[m,p] C[n,q] I[m,n,s,l] C[s,r] C[l,t] (q, r, t < nbf5). With
    I = b b^T this equals sum_P b_MO[p,q,P] b_MO[r,t,P]."""
    Cq = C[:, 0:nbf5]
    H_mat = np.einsum("mi,mn,nj->ij", C, H, Cq, optimize=True)
    I_MO = np.einsum("mp,nq,mnsl,sr,lt->pqrt", C, Cq, I, Cq, Cq, optimize=True)
    return H_mat, I_MO


def lagrange2_ri(n, cj12, ck12, C, H, b_mnl, nbf5, nbeta, nalpha, MSpin=0):
    """(elag, Hmat) of computeLagrange2 with p.RI on the CPU (utils.py:262-270, :328-355).

    elag[:, b] (b < nbf5; zero for b >= nbf5), with cj, ck = cj12, ck12 with zeroed
    diagonals (the reference zeroes the caller's arrays in place, :269-270):
      n_b Hmat[a,b]                                    (:334)
      + n_b sum_k b_MO[a,b,k] b_MO[b,b,k]   for b < nbeta and nalpha <= b < nbf5 (:337-346)
      + sum_k b_MO[a,b,k] sum_q cj[b,q] b_MO[q,q,k]    (:349-350)
      - sum_q ck[b,q] sum_k b_MO[a,q,k] b_MO[b,q,k]    (:353-355)
    For MSpin != 0 the reference's RI CPU branch leaves elag at zero."""
    Hmat, b_MO = integrals_mo_ri(C, H, b_mnl, nbf5)
    cj = np.array(cj12, dtype=np.float64, copy=True)
    ck = np.array(ck12, dtype=np.float64, copy=True)
    np.fill_diagonal(cj, 0.0)
    np.fill_diagonal(ck, 0.0)
    nbf = C.shape[0]
    elag = np.zeros((nbf, nbf))
    if MSpin == 0:
        elag[:, :nbf5] += np.einsum("b,ab->ab", n, Hmat[:, :nbf5], optimize=True)
        elag[:, :nbeta] += np.einsum("b,abk,bbk->ab", n[:nbeta], b_MO[:, :nbeta, :], b_MO[:nbeta, :nbeta, :],
                                     optimize=True)
        elag[:, nalpha:nbf5] += np.einsum("b,abk,bbk->ab", n[nalpha:nbf5], b_MO[:, nalpha:nbf5, :],
                                          b_MO[nalpha:nbf5, nalpha:nbf5, :], optimize=True)
        tmp = np.einsum("bq,qqk->bk", cj, b_MO[:nbf5, :nbf5, :], optimize=True)
        elag[:, :nbf5] += np.einsum("abk,bk->ab", b_MO[:, :nbf5, :], tmp, optimize=True)
        elag[:, :nbf5] -= np.einsum("bq,aqk,bqk->ab", ck, b_MO[:, :nbf5, :], b_MO[:nbf5, :nbf5, :], optimize=True)
    return elag, Hmat


def lagrange2_loops(n, cj12, ck12, C, H, b_mnl, nbf5, nbeta, nalpha):
    """The MSpin = 0 elag of lagrange2_ri by explicit loops (tiny inputs only), from
    b_MO and Hmat of integrals_mo_loops: an independent route."""
    Hmat, B = integrals_mo_loops(C, H, b_mnl, nbf5)
    nbf, naux = b_mnl.shape[0], b_mnl.shape[2]
    elag = np.zeros((nbf, nbf))
    for a in range(nbf):
        for bb in range(nbf5):
            v = n[bb] * Hmat[a, bb]
            if bb < nbeta or nalpha <= bb:
                v += n[bb] * sum(B[a, bb, k] * B[bb, bb, k] for k in range(naux))
            for q in range(nbf5):
                if q == bb:
                    continue
                v += cj12[bb, q] * sum(B[a, bb, k] * B[q, q, k] for k in range(naux))
                v -= ck12[bb, q] * sum(B[a, q, k] * B[bb, q, k] for k in range(naux))
            elag[a, bb] = v
    return elag, Hmat


def fock_rc(J, K, n, H, cj12, ck12, nbeta, nalpha, no1=0):
    """Generalised Fock matrices of the closed-shell case, computeF_RC_CPU
    (utils.py:25-54): F[i] = n_i H + w_i n_i J[i] + sum_j cj'[i,j] J[j] - sum_j ck'[i,j] K[j]
    with w_i = 1 for ini <= i < nbeta or nalpha <= i < nbf5, ini = no1 if no1 > 1 else 0,
    and cj', ck' the coefficients with the diagonal of their [ini:, ini:] block zeroed (the
    reference zeroes it in the caller's arrays, :45, :50; this copy leaves them alone)."""
    nbf5 = J.shape[0]
    ini = no1 if no1 > 1 else 0
    cj = np.array(cj12, dtype=np.float64, copy=True)
    ck = np.array(ck12, dtype=np.float64, copy=True)
    np.fill_diagonal(cj[ini:, ini:], 0)
    np.fill_diagonal(ck[ini:, ini:], 0)
    F = np.einsum("i,mn->imn", n, H, optimize=True)
    F[ini:nbeta] += np.einsum("i,imn->imn", n[ini:nbeta], J[ini:nbeta], optimize=True)
    F[nalpha:nbf5] += np.einsum("i,imn->imn", n[nalpha:nbf5], J[nalpha:nbf5], optimize=True)
    F += np.einsum("ij,jmn->imn", cj, J, optimize=True)
    F -= np.einsum("ij,jmn->imn", ck, K, optimize=True)
    return F


def fock_ro(J, K, n, H, cj12, ck12, nbeta, nalpha, no1=0):
    """Generalised Fock matrices of the restricted open-shell case, computeF_RO_CPU
    (utils.py:96-171), with D = [0, nbeta), S = [nbeta, nalpha), W = [nalpha, nbf5):
    n_i H on D and W, H / 2 on S (:104-106); n_i J[i] on [ini, nbeta) and W (:109-114);
    C^J', C^K' (diagonal of the [ini:, ini:] block zeroed, :121, :138) over the D/W x D/W
    blocks only (:122-152); SUMij (:155-161): n_i (J[s] - K[s] / 2) summed over s in S on
    D and W rows, and on S rows (J[s] - K[s]) / 2 over s in S less the row's own term;
    PRODWROij (:164-169): sum_j n_j (J[j] - K[j] / 2) over j in D and W on S rows. The
    caller's arrays are left alone."""
    nbf5 = J.shape[0]
    ini = no1 if no1 > 1 else 0
    cj = np.array(cj12, dtype=np.float64, copy=True)
    ck = np.array(ck12, dtype=np.float64, copy=True)
    np.fill_diagonal(cj[ini:, ini:], 0)
    np.fill_diagonal(ck[ini:, ini:], 0)
    b, a = nbeta, nalpha
    F = np.zeros((nbf5,) + H.shape)
    F[:b] += np.einsum("i,mn->imn", n[:b], H, optimize=True)
    F[b:a] += 0.5 * H
    F[a:] += np.einsum("i,mn->imn", n[a:], H, optimize=True)
    F[ini:b] += np.einsum("i,imn->imn", n[ini:b], J[ini:b], optimize=True)
    F[a:] += np.einsum("i,imn->imn", n[a:], J[a:], optimize=True)
    for r in (slice(0, b), slice(a, nbf5)):
        for c in (slice(0, b), slice(a, nbf5)):
            F[r] += np.einsum("ij,jmn->imn", cj[r, c], J[c], optimize=True)
            F[r] -= np.einsum("ij,jmn->imn", ck[r, c], K[c], optimize=True)
    js = J[b:a] - 0.5 * K[b:a]
    F[:b] += np.einsum("i,jmn->imn", n[:b], js, optimize=True)
    F[a:] += np.einsum("i,jmn->imn", n[a:], js, optimize=True)
    F[b:a] += 0.5 * np.sum(J[b:a] - K[b:a], axis=0) - 0.5 * (J[b:a] - K[b:a])
    dw = np.r_[0:b, a:nbf5]
    F[b:a] += np.einsum("j,jmn->mn", n[dw], J[dw] - 0.5 * K[dw], optimize=True)
    return F


def fock_ro_loops(J, K, n, H, cj12, ck12, nbeta, nalpha, no1=0):
    """fock_ro by explicit loops over i, j (tiny inputs only), one coefficient of H, J[j] and
    K[j] per (i, j) from the orbital classes of i and j."""
    nbf5 = J.shape[0]
    ini = no1 if no1 > 1 else 0

    def single(k):
        return nbeta <= k < nalpha

    F = np.zeros((nbf5,) + H.shape)
    for i in range(nbf5):
        F[i] = (0.5 if single(i) else n[i]) * H
        for j in range(nbf5):
            a = c = 0.0  # coefficients of J[j] and K[j]
            if single(i) and single(j):
                if j != i:
                    a, c = 0.5, -0.5
            elif single(i):
                a, c = n[j], -0.5 * n[j]
            elif single(j):
                a, c = n[i], -0.5 * n[i]
            else:
                if not (i == j and i >= ini):
                    a, c = cj12[i, j], -ck12[i, j]
                if i == j and i >= ini:
                    a += n[i]
            F[i] += a * J[j] + c * K[j]
    return F


def lagrange_fock(F, C, noptorb):
    """Lagrange multipliers from the Fock matrices, computeLagrange (utils.py:249-259):
    G[m, i] = sum_n F[i, m, n] C[n, i]; elag[i, j] = sum_m C[m, i] G[m, j] for i < noptorb,
    j < nbf5; zero elsewhere."""
    nbf5, nbf = F.shape[0], F.shape[1]
    G = np.einsum("imn,ni->mi", F, C[:, :nbf5], optimize=True)
    elag = np.zeros((nbf, nbf))
    elag[:noptorb, :nbf5] = np.einsum("mi,mj->ij", C[:, :noptorb], G, optimize=True)[:noptorb, :nbf5]
    return elag


def fock_loops(J, K, n, H, cj12, ck12, nbeta, nalpha, no1=0):
    """fock_rc by explicit loops over i, j (tiny inputs only)."""
    nbf5, nbf = J.shape[0], J.shape[1]
    ini = no1 if no1 > 1 else 0
    F = np.zeros((nbf5, nbf, nbf))
    for i in range(nbf5):
        F[i] = n[i] * H
        if ini <= i < nbeta or nalpha <= i < nbf5:
            F[i] += n[i] * J[i]
        for j in range(nbf5):
            if i == j and i >= ini:
                continue
            F[i] += cj12[i, j] * J[j] - ck12[i, j] * K[j]
    return F


def orbital_gradient(elag, nbf5):
    """calcorbg's gradient vector (pnof.py:1114-1131): 4 (elag - elag^T)[i, j] for
    i < nbf5, j > i, in rotate_orbital's variable order (utils.py:550-565)."""
    g = 4.0 * (elag - elag.T)
    nbf = elag.shape[0]
    return np.array([g[i, j] for i in range(nbf5) for j in range(i + 1, nbf)])


def rotate_orbital(y, C, nbf5):
    """C exp(kappa(y)), kappa[i, j] = y, kappa[j, i] = -y for i < nbf5, j > i
    (utils.py:550-565)."""
    from scipy.linalg import expm

    nbf = C.shape[0]
    kap = np.zeros((nbf, nbf))
    m = 0
    for i in range(nbf5):
        for j in range(i + 1, nbf):
            kap[i, j] = y[m]
            kap[j, i] = -y[m]
            m += 1
    return C @ expm(kap)


def metric_power(metric, alpha=-0.5, cutoff=1e-14):
    """psi4 Matrix::power(alpha, cutoff) as integrals.py:42 uses it (psi4 is a dependency the
    reference does not pin): eigenpairs of the symmetric metric; for alpha < 0, eigenvalues
    with |a| < cutoff * max|a| are dropped; the rest are raised to alpha, non-finite results
    dropped; V diag(a') V^T."""
    a, V = np.linalg.eigh(np.asarray(metric, dtype=np.float64))
    amax = max(abs(a[0]), abs(a[-1]))
    f = np.zeros_like(a)
    for i, x in enumerate(a):
        if alpha < 0.0 and abs(x) < cutoff * amax:
            continue
        with np.errstate(all="ignore"):
            v = x ** alpha
        f[i] = v if np.isfinite(v) else 0.0
    return (V * f) @ V.T


def b_from_3c(mnk, metric, cutoff=1e-14):
    """The reference's b producer (integrals.py:35-49): b[m, n, l] = sum_k (mn|k) [(P|Q)^-1/2][k, l]."""
    W = metric_power(metric, -0.5, cutoff)
    return np.einsum("mnk,kl->mnl", mnk, W, optimize=True)


def energy(n, J_MO, K_MO, H_core):
    """E_1 + E_2 with E_1 = 2 sum n H_core and E_2 = sum_pq n_p n_q (2J_pq - K_pq)."""
    e1 = 2.0 * float(np.dot(n, H_core))
    e2 = float(n @ (2.0 * J_MO - K_MO) @ n)
    return e1 + e2


def calce_closed_shell(n, cj12, ck12, J_MO, K_MO, H_core, nbeta, nalpha, nbf5):
    """Closed-shell (MSpin = 0) energy of pnof.py:905-922, for given C^J, C^K."""
    cj = np.array(cj12, dtype=np.float64, copy=True)
    ck = np.array(ck12, dtype=np.float64, copy=True)
    E = 2.0 * np.einsum("i,i", n, H_core)
    dj = np.diagonal(J_MO)
    E += np.einsum("i,i", n[:nbeta], dj[:nbeta])
    E += np.einsum("i,i", n[nalpha:nbf5], dj[nalpha:nbf5])
    np.fill_diagonal(cj, 0.0)
    E += np.einsum("ij,ji->", cj, J_MO)
    np.fill_diagonal(ck, 0.0)
    E -= np.einsum("ij,ji->", ck, K_MO)
    return float(E)
