"""CPU oracle of the RI JK build: test infrastructure, never part of the product.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg import this package, and only as the checker or the timed CPU baseline.
"""
