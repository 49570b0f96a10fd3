"""The no-allocation piece-label entry point (bm_piece_labels_static, scratch in a static
device array) against the allocating one (bm_piece_labels): labels and piece counts equal bit
for bit, on two calls in a row, the second on a smaller solver."""

import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from blademaster_hip import Config, Solver, _lib  # noqa: E402


def _solver(size, seed, cuts):
    """A block split by `cuts` planar cuts (x, then y): 2**cuts pieces."""
    s = Solver(Config(size=size, grid_spacing=0.006, dt=0.0002, cache_topology=True, sampling_seed=seed))
    rng = np.random.default_rng(seed)
    x = s.reference
    codes = np.zeros(s.n, np.int64)
    for axis in range(cuts):
        codes += 3**axis * np.where(x[:, axis] < rng.uniform(-0.005, 0.005), 1, 2)
    s.codes.assign(codes)
    return s


def _static_labels(s, fields=16):
    lib = _lib.load()
    dev = s.device.torch
    bits = (np.asarray(s.weight_numpy) > 0).astype(np.uint8) << np.arange(8, dtype=np.uint8)
    mask = torch.tensor(bits.sum(axis=1).astype(np.uint8), device=dev)
    labels = torch.full((s.n,), -1, dtype=torch.int32, device=dev)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    errors = torch.full((1,), 7, dtype=torch.int32, device=dev)  # the call clears it
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    code = lib.bm_piece_labels_static(ctypes.byref(s._params), ctypes.byref(s._state), fields,
                                      ctypes.c_void_p(mask.data_ptr()), ctypes.c_void_p(labels.data_ptr()),
                                      ctypes.c_void_p(count.data_ptr()), ctypes.c_void_p(errors.data_ptr()), stream)
    _lib.check(code, "bm_piece_labels_static")
    torch.cuda.synchronize(dev)
    return labels.cpu().numpy().astype(np.int64), int(count.item()), int(errors.item())


def test_static_scratch_labels_equal_the_allocating_entry_point():
    large = _solver((0.06, 0.036, 0.024), 1, cuts=2)
    small = _solver((0.036, 0.024, 0.018), 2, cuts=1)
    assert small.n < large.n
    for s in (large, small):
        print(f"static labels: n={s.n} nodes={s.node_count}", flush=True)
        labels, count, errors = _static_labels(s)
        want, sizes = s.piece_labels()
        assert errors == 0
        assert count == len(sizes) == 2 ** (2 if s is large else 1)
        np.testing.assert_array_equal(labels, want)
