"""Dense ReLU network on the device, trained and evaluated by the MFMA kernels of libbm_hip.

`Mlp` holds the parameters, Adam moments, activations and the packed bfloat16 inference
weights as PyTorch-ROCm tensors and calls `bm_mlp_*` on torch's current stream:
- `train_step(x, t)`: forward, mean squared error, backward and one Adam step (float32);
- `fit(X, Y, ...)`: epochs of shuffled minibatches, each step one replay of a HIP graph;
- `infer(x)`: one fused bfloat16 launch (float32 accumulation) with the standardisation of
  the inputs and outputs applied inside the kernel.

The network has no counterpart in the reference (see `surrogate.py`).
"""

import ctypes

import numpy as np
import torch

from . import _lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _fptr(t):
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), _lib.c_f32p) if t is not None else None


class Mlp:
    """inputs -> hidden (ReLU) -> hidden (ReLU) -> outputs.

    capacity: largest batch of `forward`, `backward` and `train_step` (activation buffers).
    """

    def __init__(self, inputs=96, hidden=256, outputs=6, capacity=4096, device="cuda:0", seed=0, lr=1.0e-3,
                 weight_decay=0.0):
        if not torch.cuda.is_available():
            raise ValueError("CUDA device required")
        self._lib = _lib.load()
        self.inputs, self.hidden, self.outputs, self.capacity = int(inputs), int(hidden), int(outputs), int(capacity)
        self.device = torch.device(device)
        count, packed, work = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self._lib.bm_mlp_sizes(self.inputs, self.hidden, self.outputs, ctypes.byref(count),
                                          ctypes.byref(packed), ctypes.byref(work)), "bm_mlp_sizes")
        dev, f32 = self.device, torch.float32
        self.params = torch.tensor(self.initial_params(seed), dtype=f32, device=dev)
        assert self.params.numel() == count.value
        z = lambda *shape: torch.zeros(*shape, dtype=f32, device=dev)  # noqa: E731
        self.grads, self.adam_m, self.adam_v = z(count.value), z(count.value), z(count.value)
        self.h1, self.h2, self.out = z(capacity, hidden), z(capacity, hidden), z(capacity, outputs)
        self.d_h1, self.d_h2, self.d_out = z(capacity, hidden), z(capacity, hidden), z(capacity, outputs)
        self.loss = z(1)
        self.step = torch.zeros(1, dtype=torch.int32, device=dev)
        # Learning rate and decoupled weight decay of the AdamW step, read on the device.
        self.lr = torch.tensor([float(lr), float(weight_decay)], dtype=f32, device=dev)
        self.packed = torch.zeros(packed.value, dtype=torch.int16, device=dev)
        self.work = z(work.value)
        self.x_shift, self.x_scale = z(inputs), torch.ones(inputs, dtype=f32, device=dev)
        self.y_shift, self.y_scale = z(outputs), torch.ones(outputs, dtype=f32, device=dev)
        self._packed_current = False
        m = _lib.Mlp(inputs=self.inputs, hidden=self.hidden, outputs=self.outputs)
        for name in ("params", "grads", "adam_m", "adam_v", "h1", "h2", "out", "d_out", "d_h2", "d_h1", "loss", "lr",
                     "x_shift", "x_scale", "y_shift", "y_scale", "work"):
            setattr(m, name, _fptr(getattr(self, name)))
        m.step = ctypes.cast(_ptr(self.step), _lib.c_i32p)
        m.packed = ctypes.cast(_ptr(self.packed), ctypes.POINTER(ctypes.c_uint16))
        m.work_len = self.work.numel()
        self._m = m

    def initial_params(self, seed):
        """He-uniform weights (bound sqrt(6 / fan_in)), zero biases, in the augmented layout."""
        rng = np.random.default_rng(seed)
        parts = []
        for rows, cols in ((self.inputs, self.hidden), (self.hidden, self.hidden), (self.hidden, self.outputs)):
            w = np.zeros((rows + 1, cols), dtype=np.float32)
            w[:rows] = rng.uniform(-1.0, 1.0, size=(rows, cols)) * np.sqrt(6.0 / rows)
            parts.append(w.ravel())
        return np.concatenate(parts)

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _batch(self, x, width):
        if x.device != self.device or x.dtype != torch.float32 or not x.is_contiguous():
            raise ValueError("Inputs must be contiguous float32 tensors on the network's device")
        if x.dim() != 2 or x.shape[1] != width:
            raise ValueError(f"Expected a [batch][{width}] tensor, got {tuple(x.shape)}")
        return x.shape[0]

    # ------------------------------------------------------------------ training (float32)
    def forward(self, x):
        """Network output for standardised inputs x [batch][inputs]; a view of `out`."""
        b = self._batch(x, self.inputs)
        if b > self.capacity:
            raise ValueError(f"Batch {b} exceeds the capacity {self.capacity}")
        _lib.check(self._lib.bm_mlp_forward(ctypes.byref(self._m), _ptr(x), b, self._stream()), "bm_mlp_forward")
        return self.out[:b]

    def backward(self, x, target):
        """Loss (device scalar) and gradients in `grads`; parameters unchanged."""
        b = self._batch(x, self.inputs)
        if self._batch(target, self.outputs) != b or b > self.capacity or b < 1:
            raise ValueError("Bad batch")
        _lib.check(self._lib.bm_mlp_backward(ctypes.byref(self._m), _ptr(x), _ptr(target), b, self._stream()),
                   "bm_mlp_backward")
        return self.loss

    def train_step(self, x, target):
        b = self._batch(x, self.inputs)
        if self._batch(target, self.outputs) != b or b > self.capacity or b < 1:
            raise ValueError("Bad batch")
        _lib.check(self._lib.bm_mlp_train_step(ctypes.byref(self._m), _ptr(x), _ptr(target), b, self._stream()),
                   "bm_mlp_train_step")
        self._packed_current = False
        return self.loss

    def standardize(self, X, Y, x_floor=None, y_floor=None):
        """Set the input/output standardisation from a data set (mean 0, std 1 per column).

        x_floor, y_floor: smallest standard deviation per column (a column that barely varies
        in the data must not be amplified beyond its physical resolution)."""
        x_std = X.std(0).clamp_min(1.0e-6)
        y_std = Y.std(0).clamp_min(1.0e-9)
        if x_floor is not None:
            x_std = torch.maximum(x_std, torch.as_tensor(x_floor, dtype=torch.float32, device=self.device))
        if y_floor is not None:
            y_std = torch.maximum(y_std, torch.as_tensor(y_floor, dtype=torch.float32, device=self.device))
        self.x_shift.copy_(X.mean(0))
        self.x_scale.copy_(1.0 / x_std)
        self.y_shift.copy_(Y.mean(0))
        self.y_scale.copy_(y_std)
        self._packed_current = False

    def loss_on(self, X, Y):
        """Mean squared error on a data set in the standardised space (float32 forward)."""
        total, n = 0.0, self._batch(X, self.inputs)
        for a in range(0, n, self.capacity):
            xs = ((X[a:a + self.capacity] - self.x_shift) * self.x_scale).contiguous()
            out = self.forward(xs)
            total += float(((out - (Y[a:a + self.capacity] - self.y_shift) / self.y_scale) ** 2).sum())
        return total / max(n * self.outputs, 1)

    def fit(self, X, Y, epochs=10, batch=1024, lr=None, seed=0, graph=True, log=None, lr_end=None,
            weight_decay=None, validation=None):
        """Minibatch AdamW on the standardised data set (X, Y device tensors).

        Each step gathers its shuffled rows into static buffers and replays a captured graph
        of the training step.  lr_end: cosine decay of the learning rate from lr to lr_end over
        the epochs.  validation: (Xv, Yv); the parameters of the epoch with the lowest
        validation loss are kept.  Returns the mean training loss of every epoch (and sets
        `validation_history`).
        """
        n = self._batch(X, self.inputs)
        if self._batch(Y, self.outputs) != n:
            raise ValueError("X and Y differ in length")
        batch = min(int(batch), self.capacity, n)
        lr0 = float(self.lr[0]) if lr is None else float(lr)
        if weight_decay is not None:
            self.lr[1] = float(weight_decay)
        Xs = (X - self.x_shift) * self.x_scale
        Ys = (Y - self.y_shift) / self.y_scale
        dev = self.device
        xb = torch.empty(batch, self.inputs, dtype=torch.float32, device=dev)
        yb = torch.empty(batch, self.outputs, dtype=torch.float32, device=dev)
        idx = torch.arange(batch, dtype=torch.int64, device=dev)
        total = torch.zeros(1, dtype=torch.float32, device=dev)

        def step():
            torch.index_select(Xs, 0, idx, out=xb)
            torch.index_select(Ys, 0, idx, out=yb)
            self.train_step(xb, yb)
            total.add_(self.loss)

        step_graph = None
        if graph:
            # One graph per call (it reads this call's data set): gather, forward, backward, AdamW
            # and the loss sum.  Warm-up on a side stream first, then restore the optimiser state.
            saved = [t.clone() for t in (self.params, self.adam_m, self.adam_v, self.step)]
            side = torch.cuda.Stream(dev)
            side.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(side):
                step()
            torch.cuda.current_stream(dev).wait_stream(side)
            step_graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(step_graph):
                step()
            for t, v in zip((self.params, self.adam_m, self.adam_v, self.step), saved):
                t.copy_(v)
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        history = []
        self.validation_history = []
        best = None
        steps = n // batch
        for epoch in range(epochs):
            rate = lr0 if lr_end is None else lr_end + 0.5 * (lr0 - lr_end) * (1 + np.cos(np.pi * epoch / epochs))
            self.lr[0] = rate
            perm = torch.randperm(n, device=dev, generator=gen)
            total.zero_()
            for s in range(steps):
                idx.copy_(perm[s * batch:(s + 1) * batch])
                if step_graph is not None:
                    step_graph.replay()
                else:
                    step()
            history.append(float(total[0]) / max(steps, 1))
            if validation is not None:
                self.validation_history.append(self.loss_on(*validation))
                if best is None or self.validation_history[-1] < best[0]:
                    best = (self.validation_history[-1], self.params.clone())
            if log is not None:
                log(epoch, history[-1])
        if best is not None:
            self.params.copy_(best[1])
        self._packed_current = False
        return history

    # ------------------------------------------------------------------ inference (bf16)
    def pack(self):
        _lib.check(self._lib.bm_mlp_pack(ctypes.byref(self._m), self._stream()), "bm_mlp_pack")
        self._packed_current = True

    def infer(self, x, out=None):
        """Outputs for raw (unstandardised) inputs x [batch][inputs], bfloat16 weights."""
        b = self._batch(x, self.inputs)
        if not self._packed_current:
            self.pack()
        if out is None:
            out = torch.empty(b, self.outputs, dtype=torch.float32, device=self.device)
        _lib.check(self._lib.bm_mlp_infer(ctypes.byref(self._m), _ptr(x), _ptr(out), b, self._stream()),
                   "bm_mlp_infer")
        return out

    # ------------------------------------------------------------------ persistence
    def state_dict(self):
        return {k: getattr(self, k).detach().cpu().numpy() for k in ("params", "x_shift", "x_scale", "y_shift",
                                                                      "y_scale")}

    def load_state_dict(self, state):
        for k in ("params", "x_shift", "x_scale", "y_shift", "y_scale"):
            value = torch.as_tensor(np.asarray(state[k], dtype=np.float32))
            if value.numel() != getattr(self, k).numel():
                raise ValueError(f"{k}: {value.numel()} values for {getattr(self, k).numel()}")
            getattr(self, k).copy_(value.reshape(getattr(self, k).shape))
        self._packed_current = False

    def save(self, path):
        np.savez(path, inputs=self.inputs, hidden=self.hidden, outputs=self.outputs, **self.state_dict())

    @classmethod
    def load(cls, path, device="cuda:0", capacity=4096):
        with np.load(path) as z:
            net = cls(int(z["inputs"]), int(z["hidden"]), int(z["outputs"]), capacity=capacity, device=device)
            net.load_state_dict({k: z[k] for k in ("params", "x_shift", "x_scale", "y_shift", "y_scale")})
        return net
