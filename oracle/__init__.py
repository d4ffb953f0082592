"""CPU oracle for the ProteinSim2015 hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package, and only as the checker.  The product
(``proteinsim2015_amd``) never imports it.
"""
from .oracle import Oracle, ForceParams, GBCoeffs, load_oracle  # noqa: F401
from .reference import Reference, reference_available  # noqa: F401
