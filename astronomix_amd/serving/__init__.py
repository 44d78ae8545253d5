"""Serving: continuous batching of independent FV MHD simulations.

``ServingEngine`` admits requests as they arrive into per-grid pools of
slots (``amx_fv_pool_*``), bounded by an HBM budget, and steps every pool
that has running slots each iteration -- the pools on their own HIP streams,
so their launches overlap on the device. Finished simulations are copied
out and their slots reused at once. ``run_one_at_a_time`` serves the same
requests one after another through ``time_integration``'s single-simulation
path, the comparison point for throughput and latency.

No counterpart in the reference (one simulation per ``time_integration``
call, ``time_stepping/time_integration.py:61-236``).
"""

from astronomix_amd.serving.requests import Request, mixed_requests, initial_state, request_config
from astronomix_amd.serving.engine import ServingEngine, run_one_at_a_time, latency_summary

__all__ = ["Request", "mixed_requests", "initial_state", "request_config", "ServingEngine", "run_one_at_a_time",
           "latency_summary"]
