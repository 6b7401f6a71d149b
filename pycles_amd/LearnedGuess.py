"""Learned first guess for the saturation adjustment of ThermodynamicsSA.

The reference's eos_c (thermodynamics_sa.h:17-66) starts its secant from the
unsaturated temperature and iterates to |dT| < 1e-3 K.  In the FP32 form a
small network (3 -> 64 -> 64 -> 1, tanh) maps normalised (s, qt, p0) to a
temperature guess for the saturated cells, one or two Newton corrections
follow, and an FP64 residual check sends any cell that misses the tolerance
back to the FP64 secant.

Everything runs in the library's HIP kernels: the training pairs come from
the FP64 saturation adjustment on the GPU, training uses the library's
forward/backward and Adam kernels, and inference (bf16 matrix cores, FP32
accumulate) is fused into the thermodynamics kernel.  torch only allocates.

    guess = SaturationGuess(Th.CC).fit()      # generate pairs, train, pack
    Th.mlp = guess                             # ThermodynamicsSA uses it (FP32)
"""
from __future__ import annotations

import numpy as np
import torch

from . import _device as D
from ._device import _capi


class SaturationGuess:
    def __init__(self, CC, n_pairs: int = 1 << 22, batch: int = 4096, epochs: int = 40,
                 lr_max: float = 3e-3, lr_min: float = 1e-5, seed: int = 20240611, max_iter: int = 0,
                 saturated_only: bool = True):
        if n_pairs & (n_pairs - 1):
            raise ValueError("n_pairs must be a power of two")
        self.CC = CC
        self.n = int(n_pairs)
        self.batch = int(batch)
        self.epochs = int(epochs)
        self.lr_max = float(lr_max)
        self.lr_min = float(lr_min)
        self.seed = int(seed)
        self.max_iter = int(max_iter)
        # the guess is only used on the saturated branch, so by default the
        # pairs are the sampling box conditioned on saturation
        self.saturated_only = bool(saturated_only)
        self.dev = D.require_gpu()
        npar = _capi.mlp_param_count()
        self.params = torch.zeros(npar, dtype=torch.float32, device=self.dev)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.norm = torch.zeros(_capi.mlp_norm_bytes(), dtype=torch.uint8, device=self.dev)
        self.weights = torch.zeros(_capi.mlp_weights_bytes(), dtype=torch.uint8, device=self.dev)
        self.workspace = torch.zeros(_capi.mlp_workspace_bytes(self.batch), dtype=torch.uint8, device=self.dev)
        self.data = None
        self.steps = 0
        self.loss_history: list[float] = []
        self._tab = D.to_device(np.asarray(CC.pv, dtype=np.float64))
        t = CC.table
        self._lt = _capi.Lookup(self._tab.data_ptr(), len(CC.pv), t.x_min, t.dx, t.dxi)

    # -- data ----------------------------------------------------------------
    def pairs(self, n: int, seed: int, saturated_only: bool | None = None) -> torch.Tensor:
        """[n][4] = (s, qt, p0, T) from the FP64 saturation adjustment."""
        sat = self.saturated_only if saturated_only is None else saturated_only
        out = torch.empty((n, 4), dtype=torch.float64, device=self.dev)
        _capi.sa_training_pairs(self._lt, n, seed, self.max_iter, out.data_ptr(), D.stream(), 1 if sat else 0)
        return out

    def generate(self):
        self.data = self.pairs(self.n, self.seed)
        _capi.mlp_normalisation(self.data.data_ptr(), self.n, self.norm.data_ptr(), self.workspace.data_ptr(),
                                D.stream())
        _capi.mlp_init(self.params.data_ptr(), self.m.data_ptr(), self.v.data_ptr(), self.seed, D.stream())
        self.steps = 0
        return self

    # -- training ------------------------------------------------------------
    def train(self, epochs: int | None = None, total_epochs: int | None = None):
        """Run epochs of Adam; the cosine schedule spans total_epochs."""
        if self.data is None:
            self.generate()
        epochs = self.epochs if epochs is None else int(epochs)
        total = (self.epochs if total_epochs is None else int(total_epochs)) * (self.n // self.batch)
        losses = torch.zeros(epochs, dtype=torch.float64, device=self.dev)
        _capi.mlp_train(self.params.data_ptr(), self.m.data_ptr(), self.v.data_ptr(), self.norm.data_ptr(),
                        self.data.data_ptr(), self.n, self.batch, epochs, self.lr_max, self.lr_min,
                        self.seed, self.steps, total, losses.data_ptr(), self.workspace.data_ptr(), D.stream())
        self.steps += epochs * (self.n // self.batch)
        self.loss_history += [float(x) / self.n for x in losses.cpu().numpy()]
        self.pack()
        return self

    def fit(self):
        return self.generate().train()

    def pack(self):
        _capi.mlp_pack(self.params.data_ptr(), self.norm.data_ptr(), self.weights.data_ptr(), D.stream())
        return self

    def handle_ptr(self) -> int:
        """Device pointer of the packed block (pycles_sa_options.mlp)."""
        return self.weights.data_ptr()

    # -- evaluation ----------------------------------------------------------
    def predict(self, data: torch.Tensor, packed: bool = True) -> torch.Tensor:
        n = data.shape[0]
        out = torch.empty(n, dtype=torch.float32, device=self.dev)
        if packed:
            _capi.mlp_infer(self.weights.data_ptr(), data.data_ptr(), n, out.data_ptr(), D.stream())
        else:
            _capi.mlp_forward(self.params.data_ptr(), self.norm.data_ptr(), data.data_ptr(), n, out.data_ptr(),
                              D.stream())
        return out

    def gradient(self, data: torch.Tensor):
        """Mean-loss gradient over data (rows a multiple of 128) and the loss sum."""
        g = torch.empty_like(self.params)
        loss = torch.zeros(1, dtype=torch.float64, device=self.dev)
        ws = torch.zeros(_capi.mlp_workspace_bytes(data.shape[0]), dtype=torch.uint8, device=self.dev)
        _capi.mlp_grad(self.params.data_ptr(), self.norm.data_ptr(), data.data_ptr(), data.shape[0],
                       g.data_ptr(), loss.data_ptr(), ws.data_ptr(), D.stream())
        return g, loss

    def normalisation(self) -> dict:
        """(mu, isd) of (s, qt, p0) and (mu_T, sd_T) as stored on the device."""
        v = self.norm.view(torch.float32).cpu().numpy()
        return {"mu": v[0:3].copy(), "isd": v[3:6].copy(), "mu_T": float(v[6]), "sd_T": float(v[7])}

    def state_dict(self) -> dict:
        return {"params": self.params.cpu().numpy(), "norm": self.norm.cpu().numpy()}

    def load_state_dict(self, st: dict):
        self.params.copy_(torch.from_numpy(np.asarray(st["params"], dtype=np.float32)))
        self.norm.copy_(torch.from_numpy(np.asarray(st["norm"], dtype=np.uint8)))
        return self.pack()
