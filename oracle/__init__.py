"""CPU parity oracle for the hsolve chanmode-4/5 step -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker / baseline.
The product library (libhsolve_hip) never loads it.

``Oracle`` wraps ``libhsolve_oracle.so`` (``hsolve_oracle.c``), a restatement
of the reference CPU code:

* interpreter  -- genesis/src/hines/hines_chip.c:78-602 (do_chip_hh4_update)
* elimination  -- genesis/src/hines/hines_solve.c:140-296
* phase order  -- genesis/src/hines/hines.c:222-270

``TangentLinear`` wraps ``hsolve_tl.c``: the same step (tabchannel subset)
carrying the derivative of every value along one direction of the initial
vm[], chip[] and ravals[]; the CPU reference for adjoint gradients.

Pinning (what checks each function; see tests/test_oracle_*.py):

* interpreter + elimination: the reference's own CPU goldens for
  hh1952_ap_verify.g (cluster_bringup/accel_regression_golden.txt:4,
  experiments/data/campaign_divergence.csv:2-8), the closed-form passive
  recursion, steady state at rest, and hand-computed single-op cases.
* elimination: brute-force dense solves of the same matrix on random tiny
  trees (numpy.linalg.solve), Crank-Nicolson identity V = 2x - V_old.
* tangent-linear model: central finite differences of the oracle itself
  (tests/test_oracle_tl.py); the closed-form derivative of the passive
  isopotential recursion.
* parity unpinned (no reference test reaches them): GHK / NERNST / MGBLOCK /
  conc ops (libm exp/log), SYN3 (urandom), SPIKE_OP and SYN2 event
  arithmetic.
"""
from .oracle import Oracle, TangentLinear, oracle_lib  # noqa: F401
