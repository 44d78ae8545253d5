"""The HIP path against the reference's own golden output
(tests/mhd_tests/final_state_orszag_tang.npz of the reference: 2D Orszag-Tang,
2 pi box, N = 200, t_end = 3.0, C_cfl = 0.4, produced by the reference in
float64). The run goes through the public time_integration and the C ABI in
float32 on a grid that the kernels' tiles do not divide.

The golden predates the current reference source: the float64 and float32
oracles both sit 0.7e-3 .. 1.5e-3 relative L1 from it (1057 steps, 2968
fixed-point iterations, in both precisions), so the bar against the golden is
the oracle's own, 2.5e-3; the GPU must also track the float32 oracle
(relative L1 below 2e-4 after the 1057 steps)."""

import numpy as np
import pytest

from tests.test_oracle import GOLDEN, orszag_tang_reference_ic, rel_l1

pytestmark = pytest.mark.gpu


def test_gpu_against_reference_golden(gpu, oracle):
    import torch
    import astronomix_amd as amx
    from astronomix_amd.option_classes import simulation_config as sc
    from astronomix_amd.option_classes.simulation_params import SimulationParams

    print("[gpu] Orszag-Tang 200^2 to t = 3 against the reference golden", flush=True)
    g = np.load(GOLDEN)["final_state"]
    s, box = orszag_tang_reference_ic()
    s32 = s.astype(np.float32)
    cfg = amx.SimulationConfig(dimensionality=2, mhd=True, box_size=box, num_cells=200,
                               boundary_settings=sc.ALL_PERIODIC, return_snapshots=False)
    cfg = amx.finalize_config(cfg, s32.shape)
    params = SimulationParams(C_cfl=0.4, t_end=3.0, gamma=5.0 / 3.0)
    rv = amx.get_registered_variables(cfg)
    out = amx.time_integration(torch.from_numpy(s32).cuda(), cfg, params, rv)
    torch.cuda.synchronize()
    out = out.cpu().numpy().astype(np.float64)

    err = rel_l1(out, g)
    assert err[:6].max() < 2.5e-3, err
    assert np.all(out[6] == 0.0)
    assert abs(out[0].mean() - 25.0 / 9.0) < 1e-5           # mass, to float32 sums
    assert np.abs(out[0] - out[0][::-1, ::-1]).max() < 1e-3  # OT's point symmetry

    oc = oracle.make_config(s32.shape, 2, box_size=box, C_cfl=0.4, t_end=3.0, gamma=5.0 / 3.0, tol=1e-5)
    ref, steps, t, iters = oracle.time_integration(s32, oc)
    err_o = rel_l1(out, ref.astype(np.float64))
    print(f"[gpu] rel L1 vs golden {err[:6].max():.3e}, vs float32 oracle {err_o.max():.3e}", flush=True)
    # 7.8e-5 measured on MI355X after the 1057 steps (the kernels are deterministic)
    assert err_o.max() < 2e-4, err_o
