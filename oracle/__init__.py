"""CPU oracle for the split-step elasticity sweeps -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker (or the CPU baseline).  The shipped
library (le2d_amd) never loads it.

`Oracle` wraps liboracle_{d,f}.so (le2d_ref.c, a clean-room restatement of
the reference's simple/le_core.c AOS path).  `Reference` wraps
_ref/libsimple_{d,f}.so, the reference's own simple/ sources compiled by
`make ref` where they lie; it exists only where that build ran.
"""
from .loader import Oracle, Reference, EDGE_CLAMP, EDGE_RING, build_oracle  # noqa: F401
