"""FP64 CPU oracle of the profile-curves spec -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package; the product (tool_godot_curvenet_amd) never imports it.
"""
from .oracle import *  # noqa: F401,F403
