"""Drag sensitivity of a channel obstacle given as a solid-fraction field, by
the hand-written discrete adjoint (lbm_pbb_* of include/lbm_d2q9.h).

    case = cylinder_case()                       # 1024x256, disk D=40 + noisy box
    eng = PbbEngine(case, storage="f32")
    s = torch.tensor(case["s"], device="cuda", requires_grad=True)
    umax = torch.tensor(case["u_max"], dtype=torch.float64, requires_grad=True)
    loss = DragLoss.apply(s, umax, eng)          # time-averaged drag
    loss.backward()                              # s.grad, umax.grad from the adjoint

The loss is the mean over the last `window` iterations of the momentum the
obstacle takes from the fluid per iteration (lattice units); ``cd_scale``
turns it into a drag coefficient 2 D / (rho U^2 L).
"""
import ctypes
import math

import numpy as np

from lbm_amd import _lib


def cylinder_case(nx=1024, ny=256, diameter=40.0, cx=200.0, cy=128.0, box=60, noise=0.2,
                  seed=0, u_max=0.05, Re=100.0, iterations=2000, window=1000, every=100,
                  sigma=200.0):
    """Channel with a disk of solid fraction 1 and seeded U[0, noise] solid
    fraction in a box x box square around it (outside the disk); parabolic
    inlet of peak u_max ramped with 1 - exp(-it^2 / (2 sigma^2)); Re on the
    mean inlet velocity 2/3 u_max and the diameter; TRT with Lambda = 1/4."""
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    s = np.zeros((nx, ny))
    inbox = (np.abs(ii - cx) <= box / 2) & (np.abs(jj - cy) <= box / 2)
    rng = np.random.default_rng(seed)
    s[inbox] = rng.uniform(0.0, noise, int(inbox.sum()))
    s[(ii - cx) ** 2 + (jj - cy) ** 2 <= (diameter / 2) ** 2] = 1.0
    u_avg = 2.0 * u_max / 3.0
    nu = u_avg * diameter / Re
    tau = 0.5 + 3.0 * nu
    tau_m = 0.25 / (tau - 0.5) + 0.5
    y = np.arange(ny, dtype=np.float64)
    prof = np.zeros((2, ny))
    prof[0] = 4.0 * y * (ny - 1 - y) / (ny - 1) ** 2
    ramp = np.array([1.0 - math.exp(-it ** 2 / (2.0 * sigma ** 2))
                     for it in range(iterations + 1)])
    return dict(nx=nx, ny=ny, s=s, u_max=u_max, om_p=1.0 / tau, om_m=1.0 / tau_m, tau=tau,
                profile=prof, ramp=ramp, iterations=iterations, window=window, every=every,
                diameter=diameter, u_avg=u_avg, rho_out=1.0)


def cd_scale(case):
    """Factor turning the drag loss into a drag coefficient 2 D / (rho U^2 L)."""
    return 2.0 / (1.0 * case["u_avg"] ** 2 * case["diameter"])


class PbbEngine:
    def __init__(self, case, storage="f32"):
        self.L = _lib.load()
        self.case = case
        self.storage = storage
        self.nx, self.ny, self.T = case["nx"], case["ny"], case["iterations"]
        self._prof = np.ascontiguousarray(case["profile"], dtype=np.float64)
        self._ramp = np.ascontiguousarray(case["ramp"], dtype=np.float64)
        d = _lib.PbbDesc()
        d.nx, d.ny = self.nx, self.ny
        d.storage = _lib.STORE[storage]
        d.om_p, d.om_m = case["om_p"], case["om_m"]
        d.rho_out = case.get("rho_out", 1.0)
        d.inlet_profile_host = self._prof.ctypes.data_as(_lib.c_dp)
        d.ramp_host = self._ramp.ctypes.data_as(_lib.c_dp)
        d.iterations = self.T
        d.window = case["window"]
        d.checkpoint_every = case["every"]
        h = _lib.c_vp()
        _lib.check(self.L.lbm_pbb_create(ctypes.byref(d), ctypes.byref(h)), "lbm_pbb_create")
        self.h = h
        self._s = None
        self.generation = 0          # forwards run so far; a backward pairs with the last

    def close(self):
        if getattr(self, "h", None):
            self.L.lbm_pbb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _dtype(self):
        import torch
        return torch.float64 if self.storage == "f64" else torch.float32

    def forward(self, s, u_max):
        """s: CUDA tensor [nx, ny]; returns the loss (float)."""
        import torch
        if not isinstance(s, torch.Tensor) or not s.is_cuda:
            raise TypeError("s must be a CUDA tensor")
        if tuple(s.shape) != (self.nx, self.ny):
            raise ValueError(f"s has shape {tuple(s.shape)}, expected {(self.nx, self.ny)}")
        self._s = s.detach().to(self._dtype()).contiguous()
        self.generation += 1
        out = np.zeros(1)
        _lib.check(self.L.lbm_pbb_forward(self.h, self._s.data_ptr(), float(u_max),
                                          out.ctypes.data_as(_lib.c_vp), _lib.stream_ptr()),
                   "lbm_pbb_forward")
        return float(out[0])

    def backward(self):
        """dL/ds (float64 CUDA tensor [nx, ny]) and dL/du_max of the last forward."""
        import torch
        g = torch.empty((self.nx, self.ny), dtype=torch.float64, device="cuda")
        gu = np.zeros(1)
        _lib.check(self.L.lbm_pbb_backward(self.h, g.data_ptr(), gu.ctypes.data_as(_lib.c_vp),
                                           _lib.stream_ptr()), "lbm_pbb_backward")
        return g, float(gu[0])

    def drag_history(self):
        out = np.zeros(self.T + 1)
        _lib.check(self.L.lbm_pbb_drag_history(self.h, out.ctypes.data_as(_lib.c_vp)))
        return out

    def final_state(self):
        import torch
        t = torch.empty((9, self.nx, self.ny), dtype=self._dtype(), device="cuda")
        _lib.check(self.L.lbm_pbb_final_state(self.h, t.data_ptr(), _lib.stream_ptr()))
        return t


def _make_function():
    import torch

    class DragLoss(torch.autograd.Function):
        """loss = time-averaged drag; gradients from the adjoint."""

        @staticmethod
        def forward(ctx, s, u_max, engine):
            loss = engine.forward(s, float(u_max))
            ctx.engine = engine
            ctx.generation = engine.generation
            ctx.s_dtype = s.dtype
            return torch.tensor(loss, dtype=torch.float64, device=s.device)

        @staticmethod
        def backward(ctx, grad):
            # the engine keeps the checkpoints of its last forward only
            if ctx.engine.generation != ctx.generation:
                raise RuntimeError("DragLoss.backward: the engine has run another forward "
                                   "since this loss was computed; recompute the loss first")
            gs, gu = ctx.engine.backward()
            g = grad.to(torch.float64)
            return (g * gs).to(ctx.s_dtype), g * gu, None

    return DragLoss


try:
    DragLoss = _make_function()
except ImportError:  # torch missing: the engine API above still works
    DragLoss = None
