"""Turbulent forcing between steps (``_physics_modules/_turbulent_forcing/
_turbulent_forcing.py:14-104``, applied in the loop at
``time_stepping/time_integration.py:694-704``).

Each step: the solenoidal forcing spectrum from the native kernel
(``amx_forcing_spectrum``: seeded Philox noise times sqrt(P(k)), projected),
its inverse FFT (rocFFT through ``torch.fft``), then ``amx_forcing_apply``,
which injects ``energy_injection_rate * dt`` of kinetic energy into the
velocity rows. The noise is this library's own seeded stream: the reference
draws from JAX's keys, so runs agree with it in distribution, not sample by
sample. The hook runs on the step loop's stream with dt still on the device.
"""

from __future__ import annotations

import ctypes

from astronomix_amd import _native


class TurbulentForcing:
    def __init__(self, shape, box_size: float, grid_spacing: float, energy_injection_rate: float,
                 seed: int = 42, rho_row: int = 0, vel_row: int = 1, num_vars: int = 8, device=None):
        import torch
        self.lib = _native.load_library()
        self.nx, self.ny, self.nz = (int(v) for v in shape)
        self.box_size = float(box_size)
        self.dx = float(grid_spacing)
        self.edot = float(energy_injection_rate)
        self.seed = int(seed)
        self.rho_row, self.vel_row, self.num_vars = int(rho_row), int(vel_row), int(num_vars)
        dev = device or torch.device("cuda")
        self.cw = torch.empty((3, self.nx, self.ny, self.nz), dtype=torch.complex64, device=dev)
        self.work = torch.empty(int(self.lib.amx_forcing_workspace_bytes()), dtype=torch.uint8, device=dev)
        self.amp = torch.zeros((), dtype=torch.float32, device=dev)
        self.error = None
        self._hook = _native.STEP_HOOK(self._on_step)

    def field(self, step: int, stream=None):
        """Real forcing field of a step, (3, nx, ny, nz) float32 on the GPU."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        _native.check(self.lib.amx_forcing_spectrum(ctypes.c_void_p(self.cw.data_ptr()), self.nx, self.ny, self.nz,
                                                    self.box_size, self.seed, int(step),
                                                    ctypes.c_void_p(s.cuda_stream)), "amx_forcing_spectrum")
        with torch.cuda.stream(s):
            return torch.fft.ifftn(self.cw, dim=(1, 2, 3)).real.contiguous()

    def apply_field(self, state_ptr: int, w, d_dt_ptr: int, stream=None):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        _native.check(self.lib.amx_forcing_apply(ctypes.c_void_p(state_ptr), self.num_vars, self.nx, self.ny,
                                                 self.nz, self.rho_row, self.vel_row, ctypes.c_void_p(w.data_ptr()),
                                                 ctypes.c_void_p(d_dt_ptr), self.edot, self.dx,
                                                 ctypes.c_void_p(self.work.data_ptr()),
                                                 ctypes.c_void_p(self.amp.data_ptr()), ctypes.c_void_p(s.cuda_stream)),
                      "amx_forcing_apply")

    def _on_step(self, user, state_ptr, d_dt_ptr, step, stream_ptr):
        import torch
        try:
            s = torch.cuda.ExternalStream(stream_ptr) if stream_ptr else torch.cuda.current_stream()
            w = self.field(step, s)
            self.apply_field(state_ptr, w, d_dt_ptr, s)
            return 0
        except BaseException as e:  # noqa: BLE001 -- re-raised after the loop returns
            self.error = e
            return 1

    def install(self, ctx, kind: str = "fv"):
        fn = ctx.lib.amx_fv_set_step_hook if kind == "fv" else ctx.lib.amx_fd_set_step_hook
        _native.check(fn(ctx.handle, ctypes.cast(self._hook, ctypes.c_void_p), None), "set_step_hook")

    @staticmethod
    def uninstall(ctx, kind: str = "fv"):
        fn = ctx.lib.amx_fv_set_step_hook if kind == "fv" else ctx.lib.amx_fd_set_step_hook
        _native.check(fn(ctx.handle, None, None), "set_step_hook")
