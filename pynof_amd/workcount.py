"""Algorithmic work of one JK build: the numerators of the roofline fractions.

Flop counts credit each useful multiply-add once (2 flops), use the symmetry
the outputs have (J^n, K^n over mu >= nu; b_pq over q <= p), and use the
unpadded sizes. Split-precision products are not credited: a BF16x3 product
counts as one.
"""
from __future__ import annotations


def stage_flops(nbf: int, naux: int, nbf5: int) -> dict:
    pairs = nbf * (nbf + 1)  # 2 * number of mu >= nu pairs
    return {
        "half_transform": 2.0 * nbf5 * nbf * nbf * naux,
        "quarter_transform": float(nbf5 * (nbf5 + 1)) * nbf * naux + 2.0 * nbf5 * nbf5 * naux,
        "jmo_syrk": float(nbf5 * (nbf5 + 1)) * naux,
        "coulomb_jn": float(nbf5) * naux * pairs,
        "exchange_kn": float(nbf5) * pairs * naux,
        "hcore": 2.0 * nbf * nbf * nbf5 + 2.0 * nbf * nbf5,
    }


def build_flops(nbf: int, naux: int, nbf5: int) -> float:
    """F_sym: the useful flops of one build (HT once, symmetric outputs)."""
    return (2.0 * nbf5 * nbf * nbf * naux + 2.0 * nbf5 * nbf * naux + 2.0 * nbf5 * naux * nbf * (nbf + 1)
            + float(nbf5 * (nbf5 + 1)) * nbf * naux + 4.0 * nbf5 * nbf5 * naux
            + 2.0 * nbf * nbf * nbf5 + 2.0 * nbf * nbf5)


def build_flops_dense(nbf: int, naux: int, nbf5: int) -> float:
    """The same build without using any output symmetry."""
    return (2.0 * nbf5 * nbf * nbf * naux + 2.0 * nbf5 * nbf * naux + 4.0 * nbf5 * naux * nbf * nbf
            + 2.0 * nbf5 * nbf5 * nbf * naux + 4.0 * nbf5 * nbf5 * naux + 2.0 * nbf * nbf * nbf5
            + 2.0 * nbf * nbf5)


def stage_min_bytes(nbf: int, naux: int, nbf5: int, elem: int = 8) -> dict:
    """Minimum HBM bytes per stage: operands read once, results written once."""
    ntri = nbf * (nbf + 1) // 2
    x = nbf5 * nbf * naux
    return {
        "half_transform": elem * (ntri * naux + nbf * nbf5 + x),
        "quarter_transform": elem * (x + nbf * nbf5 + nbf5 * naux),
        "coulomb_jn": elem * (ntri * naux + nbf5 * naux + nbf5 * nbf * nbf),
        "exchange_kn": elem * (x + nbf5 * nbf * nbf),
    }


def build_min_bytes(nbf: int, naux: int, nbf5: int, elem: int = 8) -> float:
    """B_min: b read once (dense, stored precision) + J, K written once + small operands."""
    return float(elem * (nbf * nbf * naux + 2 * nbf5 * nbf * nbf + 2 * nbf * nbf + 2 * nbf5 * nbf5 + nbf5))
