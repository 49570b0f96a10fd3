"""Device array handle over a PyTorch-ROCm tensor with the wp.array surface callers use.

Callers of the reference read and write solver buffers through `.numpy()` (synchronising
copy), `.assign(values)`, `.zero_()`, `.fill_(value)` and slice views `a[i:i+1]`
(solver.py:290-301, paper_freefall.py:78-79).  `DArray` gives the same surface over a torch
tensor that owns the device memory; `ptr` is the raw device address handed to the C-ABI.
A write hook lets the solver react to host writes (new cut codes invalidate the field
table, as the reference's registration would on its next material step).
"""

import numpy as np
import torch


class DArray:
    def __init__(self, tensor, on_write=None):
        self.tensor = tensor
        self._on_write = on_write

    # -- wp.array surface
    def numpy(self):
        return self.tensor.detach().cpu().numpy().copy()

    def assign(self, values):
        src = torch.as_tensor(np.asarray(values), dtype=self.tensor.dtype)
        if src.numel() != self.tensor.numel():
            raise ValueError(f"assign: {src.numel()} values for an array of {self.tensor.numel()}")
        self.tensor.copy_(src.reshape(self.tensor.shape).to(self.tensor.device))
        self._written()

    def zero_(self):
        self.tensor.zero_()
        self._written()

    def fill_(self, value):
        self.tensor.fill_(value)
        self._written()

    def __getitem__(self, index):
        return DArray(self.tensor[index], self._on_write)

    def __len__(self):
        return self.tensor.shape[0]

    @property
    def shape(self):
        return tuple(self.tensor.shape)

    @property
    def dtype(self):
        return self.tensor.dtype

    @property
    def ptr(self):
        return self.tensor.data_ptr()

    def _written(self):
        if self._on_write is not None:
            self._on_write()

    def __repr__(self):
        return f"DArray(shape={self.shape}, dtype={self.dtype})"
