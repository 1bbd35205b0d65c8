"""CPU oracle for the GraphGP hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import this package, and only as the checker or the timed CPU baseline.  The product
package ``graphgp_amd`` never imports it.
"""
from .oracle import *  # noqa: F401,F403
