"""mV view of the membrane state.

The engine holds the membrane potential relative to rest, ``u = v - v_rest`` (float32): same
model, with every rounding near threshold 8x finer than at |v| ~ 45 mV.  Callers of the
reference read and write ``v`` in mV (flyzero tests/test_gpu.py:43 ``gb.v[:, :40] = -40.0``);
``MembraneView`` keeps that working: reads return ``u + v_rest``, writes store
``u = f32(f32(value) - f32(v_rest))``.
"""

from __future__ import annotations

import numpy as np
import torch


class MembraneView:
    def __init__(self, u: torch.Tensor, v_rest: float):
        self._u = u
        self._rest = np.float32(v_rest)

    def __getitem__(self, k) -> torch.Tensor:
        return self._u[k] + float(self._rest)

    def __setitem__(self, k, value):
        if isinstance(value, torch.Tensor):
            val = value.to(device=self._u.device, dtype=torch.float32) - torch.tensor(
                float(self._rest), dtype=torch.float32, device=self._u.device)
        else:
            val = torch.from_numpy(np.atleast_1d(np.asarray(value, np.float32) - self._rest)).to(self._u.device)
            if np.ndim(value) == 0:
                val = val.reshape(())
        self._u[k] = val

    def tensor(self) -> torch.Tensor:
        return self._u + float(self._rest)

    def fill_(self, value):
        self._u.fill_(float(np.float32(value) - self._rest))
        return self

    def cpu(self):
        return self.tensor().cpu()

    def numpy(self):
        return self.cpu().numpy()

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a if dtype is None else a.astype(dtype)

    @property
    def shape(self):
        return self._u.shape

    @property
    def dtype(self):
        return self._u.dtype

    @property
    def device(self):
        return self._u.device

    def __len__(self):
        return len(self._u)

    def __repr__(self):
        return f"MembraneView(mV, shape={tuple(self._u.shape)})"
