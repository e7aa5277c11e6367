"""Seeded synthetic inputs for the density-fitted JK build.

The reference obtains b[mu,nu,P] = sum_Q (mu nu|Q) J^-1/2[Q,P] from psi4
(integrals.py:35-49); psi4 is not part of this project, so every workload uses
a seeded synthetic b with the same layout (nbf, nbf, naux), aux index last, and
the same symmetry in (mu, nu).

Recipes (identical for every caller, so a seed fixes the inputs):

* b: G ~ N(0,1) of shape (nbf, nbf, naux) from ``numpy.random.default_rng``;
  b = (G + G^T_{mu nu}) / sqrt(2); b *= 0.3 * exp(-lam * |mu - nu|).
* C: the orthogonal Q of ``numpy.linalg.qr`` of a seeded N(0,1) (nbf, nbf).
* n: ndoc subspaces, one strong orbital each plus weak orbitals assigned in
  PyNOF's strided order (pnof.py:785-803 with no1 = 0, nsoc = 0); a seeded
  softmax inside each subspace, so each subspace sums to 1 and sum(n) = ndoc.
* H: seeded symmetric N(0,1)/sqrt(nbf).

``device_b_packed`` draws the same distribution directly on the GPU in the
packed lower-triangle layout used on the device (off-diagonal pairs N(0,1),
diagonal pairs N(0,2), times the same decay), for sizes whose host copy
would not fit (nbf = 1000, naux = 3600 is 28.8 GB in FP64).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass(frozen=True)
class Workload:
    name: str
    nbf: int
    naux: int
    nbf5: int
    ndoc: int
    lam: float
    seed_b: int
    seed_c: int
    seed_n: int
    seed_h: int


def default_ndoc(nbf5: int) -> int:
    """Pair count for a synthetic workload: water's 5 pairs per 24 orbitals."""
    return max(1, (5 * nbf5 + 12) // 24)


def ceil8(x: float) -> int:
    return int(math.ceil(x / 8.0) * 8)


def workload(k: int) -> Workload:
    """The named single-instance workloads (k = 0, 1, 2)."""
    if k == 0:
        return Workload("water nbf=24 naux=116 nbf5=24", 24, 116, 24, 5, 0.05, 1000, 2000, 3000, 7000)
    if k == 1:
        return Workload("mid nbf=280 naux=1100 nbf5=280", 280, 1100, 280, default_ndoc(280), 0.05,
                        1001, 2001, 3001, 7001)
    if k == 2:
        return Workload("large nbf=1000 naux=3600 nbf5=300", 1000, 3600, 300, default_ndoc(300), 0.02,
                        1002, 2002, 3002, 7002)
    raise ValueError(f"no single-instance workload {k}")


def batch_workloads(count: int = 64, seed: int = 4000, lo: int = 40, hi: int = 400) -> list[Workload]:
    """Uneven batch: nbf ~ U[lo, hi] (seeded), naux = ceil8(3.6 nbf), nbf5 = nbf."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, size=count)
    out = []
    for i, nbf in enumerate(sizes.tolist()):
        naux = ceil8(3.6 * nbf)
        out.append(Workload(f"batch[{i}] nbf={nbf}", nbf, naux, nbf, default_ndoc(nbf), 0.05,
                            4000 + i, 14000 + i, 24000 + i, 34000 + i))
    return out


def make_b(nbf: int, naux: int, seed: int, lam: float = 0.05) -> np.ndarray:
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((nbf, nbf, naux))
    b = g + g.transpose(1, 0, 2)
    b *= 1.0 / math.sqrt(2.0)
    b *= decay(nbf, lam)[:, :, None]
    return np.ascontiguousarray(b)


def decay(nbf: int, lam: float) -> np.ndarray:
    idx = np.arange(nbf)
    return 0.3 * np.exp(-lam * np.abs(idx[:, None] - idx[None, :]).astype(np.float64))


def make_C(nbf: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((nbf, nbf)))
    return np.ascontiguousarray(q)


def subspace_of(nbf5: int, ndoc: int) -> np.ndarray:
    """Subspace index of each orbital: strong orbital i < ndoc owns weak orbitals
    ndoc + (ndoc - i - 1) + k*ndoc (PyNOF's strided order, pnof.py:785-803);
    leftover orbitals continue the same stride."""
    sub = np.empty(nbf5, dtype=np.int64)
    sub[:ndoc] = np.arange(ndoc)
    w = np.arange(ndoc, nbf5)
    sub[ndoc:] = ndoc - 1 - ((w - ndoc) % ndoc)
    return sub


def make_n(nbf5: int, ndoc: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = rng.normal(-3.0, 1.0, size=nbf5)
    x[:ndoc] = 0.0
    e = np.exp(x)
    sub = subspace_of(nbf5, ndoc)
    den = np.zeros(ndoc)
    np.add.at(den, sub, e)
    return e / den[sub]


def make_H(nbf: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((nbf, nbf))
    return np.ascontiguousarray((a + a.T) / (2.0 * math.sqrt(nbf)))


def make_inputs(w: Workload):
    """(b, C, H, n) on the host for workload ``w``."""
    b = make_b(w.nbf, w.naux, w.seed_b, w.lam)
    C = make_C(w.nbf, w.seed_c)
    H = make_H(w.nbf, w.seed_h)
    n = make_n(w.nbf5, w.ndoc, w.seed_n)
    return b, C, H, n


def rotation(nbf: int, seed: int, norm: float = 1e-2) -> np.ndarray:
    """Orthogonal exp(kappa) for a seeded antisymmetric kappa with ||kappa||_F = norm
    (the orbital update of utils.py:550-565)."""
    from scipy.linalg import expm

    rng = np.random.default_rng(seed)
    a = rng.standard_normal((nbf, nbf))
    k = a - a.T
    k *= norm / np.linalg.norm(k)
    return expm(k)


def energy_terms(n: np.ndarray, J_MO: np.ndarray, K_MO: np.ndarray, H_core: np.ndarray):
    """E_1 = 2 sum_p n_p H_core[p]; E_2 = sum_pq n_p n_q (2 J_MO[p,q] - K_MO[p,q]), in FP64."""
    n = np.asarray(n, dtype=np.float64)
    e1 = 2.0 * float(np.dot(n, H_core))
    e2 = float(n @ (2.0 * J_MO - K_MO) @ n)
    return e1, e2


def device_b_packed(nbf: int, naux: int, seed: int, lam: float, naux_ld: int, device="cuda"):
    """Packed lower triangle of a synthetic b, drawn on the device.

    Row t = mu*(mu+1)/2 + nu (mu >= nu) holds b[mu, nu, 0:naux] followed by
    zeros up to ``naux_ld``. Same distribution as ``make_b`` (the sum of two
    independent N(0,1) over sqrt(2) is N(0,1) off the diagonal and sqrt(2) N(0,1)
    on it), not the same numbers.
    """
    import torch

    ntri = nbf * (nbf + 1) // 2
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    out = torch.zeros((ntri, naux_ld), dtype=torch.float64, device=device)
    mu = torch.arange(nbf, device=device)
    chunk = max(1, (1 << 27) // max(1, naux))  # rows per draw, ~1 GiB of FP64
    rows_mu = torch.repeat_interleave(mu, mu + 1)
    rows_nu = torch.cat([torch.arange(m + 1, device=device) for m in range(nbf)]) if nbf <= 64 else None
    if rows_nu is None:
        t = torch.arange(ntri, device=device)
        rows_nu = t - rows_mu * (rows_mu + 1) // 2
    for t0 in range(0, ntri, chunk):
        t1 = min(ntri, t0 + chunk)
        g = torch.randn((t1 - t0, naux), dtype=torch.float64, device=device, generator=gen)
        m = rows_mu[t0:t1]
        v = rows_nu[t0:t1]
        s = 0.3 * torch.exp(-lam * (m - v).abs().to(torch.float64))
        s = torch.where(m == v, s * math.sqrt(2.0), s)
        out[t0:t1, :naux] = g * s[:, None]
    return out
