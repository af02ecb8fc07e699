"""Learned warm start of the DFSPH pressure solves.

A per-particle MLP (13 -> 64 -> 64 -> 64 -> 2, SiLU, bf16 weights and activations with float32 accumulation on the
MFMA units) predicts the initial kappa of the divergence and the density solve from features of the particle's state
at the start of the step (neighbour count, density deviation, velocity divergence, the previous step's kappa of both
solves and eight neighbourhood moments, computed by the solver's feature kernel). It replaces solve_init's warm-start
rule (pressure.py:254, 261: warm x the previous kappa, clamped); the solves still iterate until the same per-world
tolerance holds, so a prediction can only change the iteration count, never the accepted error.

Training data are (features -> converged kappa) pairs the plain solver logs while it steps; the targets are
asinh(kappa / ks), ks = |g| h. Forward, backward, the deterministic gradient reduction and the Adam update are the
library's own HIP kernels (dfsph_mlp_*); torch only holds the buffers and draws the minibatch indices.
"""

from __future__ import annotations

import ctypes
import math
import os

import numpy as np
import torch

from . import _abi

N_FEATURES = 13
FEAT_STRIDE = 16
# the network trained by ``python -m alloy_amd.sph.train_warm_start`` (see that module for the data it learns from)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warm_start_mlp.npy")


def _vp(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _lib():
    return _abi.load_library()


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class WarmStartMLP:
    """The network's float32 master parameters on the device plus its Adam state."""

    def __init__(self, device="cuda", seed: int = 0):
        self.lib = _lib()
        self.n_params = int(self.lib.dfsph_mlp_param_count())
        self.device = torch.device(device)
        self.params = torch.tensor(self.init_params(seed), device=self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.t = 0

    @staticmethod
    def init_params(seed: int = 0) -> np.ndarray:
        """Uniform +-sqrt(3 / fan_in) weights, zero biases; padded input columns and output rows zero."""
        from_fan = lambda rng, out, fan: rng.uniform(-1, 1, (out, fan)) * math.sqrt(3.0 / fan)  # noqa: E731
        rng = np.random.default_rng(seed)
        H = 64
        W1 = np.zeros((H, 16))
        W1[:, :N_FEATURES] = from_fan(rng, H, N_FEATURES)
        W2, W3 = from_fan(rng, H, H), from_fan(rng, H, H)
        W4 = np.zeros((16, H))
        W4[:2] = from_fan(rng, 2, H) * 0.1
        parts = [W1, np.zeros(H), W2, np.zeros(H), W3, np.zeros(H), W4, np.zeros(16)]
        return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)

    def _scratch(self, n: int) -> torch.Tensor:
        """Per-wave gradient partials for n samples, sized by the library."""
        return torch.empty(int(self.lib.dfsph_mlp_grad_scratch_floats(int(n))), dtype=torch.float32,
                           device=self.device)

    def forward(self, X: torch.Tensor) -> torch.Tensor:
        """Y [n, 2] = MLP(X [n, 16])."""
        X = X.contiguous()
        Y = torch.empty((X.shape[0], 2), dtype=torch.float32, device=self.device)
        rc = self.lib.dfsph_mlp_forward(_vp(self.params), _vp(X), int(X.shape[0]), _vp(Y), _stream())
        if rc != 0:
            raise RuntimeError(f"dfsph_mlp_forward failed ({rc})")
        return Y

    def grad(self, X: torch.Tensor, T: torch.Tensor, n_total: int | None = None) -> tuple[torch.Tensor, float]:
        """(gradient [n_params], squared-error sum) of sum((y - T)^2) / (2 n_total)."""
        n = int(X.shape[0])
        scratch = self._scratch(n)
        g = torch.empty(self.n_params + 1, dtype=torch.float32, device=self.device)
        rc = self.lib.dfsph_mlp_grad(_vp(self.params), _vp(X.contiguous()), _vp(T.contiguous()), n, int(n_total or n),
                                     _vp(scratch), scratch.numel(), _vp(g), _stream())
        if rc != 0:
            raise RuntimeError(f"dfsph_mlp_grad failed ({rc})")
        return g[: self.n_params], float(g[self.n_params].item())

    def train(self, X: torch.Tensor, T: torch.Tensor, steps: int = 2000, batch: int = 4096, lr: float = 2e-3,
              seed: int = 0, log_every: int = 0) -> list[float]:
        """Adam on random minibatches; returns the minibatch mean squared errors."""
        gen = torch.Generator(device=self.device).manual_seed(seed)
        n = int(X.shape[0])
        scratch = self._scratch(batch)
        g = torch.empty(self.n_params + 1, dtype=torch.float32, device=self.device)
        losses = []
        st = _stream()
        for k in range(steps):
            idx = torch.randint(0, n, (batch,), device=self.device, generator=gen)
            xb, tb = X.index_select(0, idx).contiguous(), T.index_select(0, idx).contiguous()
            rc = self.lib.dfsph_mlp_grad(_vp(self.params), _vp(xb), _vp(tb), batch, batch, _vp(scratch),
                                         scratch.numel(), _vp(g), st)
            if rc != 0:
                raise RuntimeError(f"dfsph_mlp_grad failed ({rc})")
            self.t += 1
            cur_lr = lr * 0.5 * (1.0 + math.cos(math.pi * k / max(steps, 1)))  # cosine schedule
            rc = self.lib.dfsph_mlp_adam(_vp(self.params), _vp(g), _vp(self.m), _vp(self.v), cur_lr, 0.9, 0.999,
                                         1e-8, self.t, st)
            if rc != 0:
                raise RuntimeError(f"dfsph_mlp_adam failed ({rc})")
            if log_every and (k % log_every == 0 or k == steps - 1):
                losses.append(float(g[self.n_params].item()) / (2 * batch))
        return losses

    def save(self, path: str) -> None:
        """The float32 master parameters as a .npy file (numpy's format, no pickling)."""
        np.save(path, self.params.detach().cpu().numpy().astype(np.float32), allow_pickle=False)

    @classmethod
    def load(cls, path: str, device="cuda") -> "WarmStartMLP":
        net = cls(device)
        p = np.load(path, allow_pickle=False)
        if p.shape != (net.n_params,) or p.dtype != np.float32:
            raise ValueError(f"{path}: expected float32[{net.n_params}] parameters, got {p.dtype}{list(p.shape)}")
        net.params.copy_(torch.from_numpy(p).to(net.device))
        return net

    def attach(self, solver) -> None:
        """Use this network as ``solver``'s warm start (the parameter tensor must stay alive)."""
        _abi.check(solver.lib.dfsph_set_warm_start_model(solver._h, _vp(self.params), None), solver.lib)
        solver._warm_model = self.params

    @staticmethod
    def detach(solver) -> None:
        _abi.check(solver.lib.dfsph_set_warm_start_model(solver._h, None, None), solver.lib)
        solver._warm_model = None


def collect(solver, state, steps: int, dt: float, per_step: int | None = None, seed: int = 0,
            warmup: int = 0) -> tuple[torch.Tensor, torch.Tensor]:
    """Step ``solver`` and log (features -> converged kappa targets) of its liquid particles: ``per_step`` random
    liquid particles per step (all when None). Returns X [m, 16], T [m, 2] on the device."""
    lib = _lib()
    n = solver.model.particle_count
    F = torch.zeros((n, FEAT_STRIDE), dtype=torch.float32, device=solver.device)
    Tg = torch.zeros((n, 2), dtype=torch.float32, device=solver.device)
    m = solver.model
    fluid = (((m.particle_flags & 1) != 0) & (m.sph.phase >= 1)).nonzero().reshape(-1)
    gen = torch.Generator(device=solver.device).manual_seed(seed)
    xs, ts = [], []
    for _ in range(warmup):
        solver.step(state, state, None, None, dt)
    _abi.check(lib.dfsph_set_logging(solver._h, _vp(F), _vp(Tg)), lib)
    try:
        for _ in range(steps):
            solver.step(state, state, None, None, dt)
            sel = fluid
            if per_step is not None and per_step < len(fluid):
                sel = fluid[torch.randperm(len(fluid), device=solver.device, generator=gen)[:per_step]]
            xs.append(F.index_select(0, sel).clone())
            ts.append(Tg.index_select(0, sel).clone())
    finally:
        _abi.check(lib.dfsph_set_logging(solver._h, None, None), lib)
    return torch.cat(xs), torch.cat(ts)
