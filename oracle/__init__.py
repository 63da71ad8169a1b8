"""CPU oracle of the VSGM enumeration path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker / the reported CPU baseline; the
product (vsgm_amd/) never does. See vsgm_oracle.h for what pins each
function.
"""
from .checker import (  # noqa: F401
    automorphisms,
    count_brute,
    triangles,
    four_cliques,
    four_cliques_ranked,
    diamonds,
    four_cycles,
    count_plan,
    parse_plan,
    build,
)
