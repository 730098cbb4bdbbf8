"""CPU parity oracle for the hsolve chanmode-4/5 step -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker / baseline.
The product library (libhsolve_hip) never loads it.

``Oracle`` wraps ``libhsolve_oracle.so`` (``hsolve_oracle.c``), a restatement
of the reference CPU code:

* interpreter  -- genesis/src/hines/hines_chip.c:78-602 (do_chip_hh4_update)
* Ca_concen    -- genesis/src/hines/hines_conc_chip.c:196-211
                  (do_h4_conc_chip_update, CACONC_OP / LCACONC_OP only)
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
* SPIKE_OP and synaptic events: refractory spacing, arrival step and the
  SYN2 impulse response in closed form, the Mg block against Jahr-Stevens
  (tests/test_oracle_events.py).
* concentrations: the Ca_concen recursion under constant influx in closed
  form, the Nernst equilibrium, the GHK current and its slope against the
  textbook GHK equation (both branches), concentration-gated channels
  settling on the interpolated table (incl. the row cache and the clamps),
  DOMAIN_OP and table currents (tests/test_oracle_conc.py).
* 2-D tables (IPOL2V; NEWCONC2 + IPOL2C) settling on the bilinearly
  interpolated table, table GHK and SAME_GHK (tests/test_oracle_conc.py).
* parity unpinned (no reference test or closed form reaches them here):
  SYN3 (urandom), ELMMPUMP.  The device path refuses both.
"""
from .oracle import Oracle, TangentLinear, oracle_lib  # noqa: F401
