"""Single operators of the substep on caller-provided arrays, for the callers and tests that
launch the reference's kernels one at a time (`material.couple`, `material.g2p`,
`contact.resist_cut`; tests/test_mechanics.py:80-105, tests/test_particle_transport.py:10-45,
tests/test_cut_resistance.py:15-42).  Each takes numpy arrays or CUDA tensors, runs one C-ABI
call on the current stream, and returns numpy results.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .solver import CUT_ADVANCE, CUT_ON


def _dev(a, dtype):
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=dtype).contiguous()
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def couple(keys, heads, links, mass, state, node_count, result=None):
    """Compatible-field exchange (Eq. 8) of `state` over the fields listed per node
    (material.py:155-209).  result: the array the reference passes in (fields with a negative
    key or no mass keep its values); zeros by default."""
    keys = _dev(keys, torch.int64)
    count = keys.numel()
    out = _dev(np.zeros((count, 3)) if result is None else result, torch.float32).reshape(count, 3)
    args = [keys, _dev(heads, torch.int32), _dev(links, torch.int32), _dev(mass, torch.float32),
            _dev(state, torch.float32).reshape(count, 3), out]
    _lib.check(_lib.load().bm_couple(*[_p(a) for a in args], count, int(node_count), _stream()), "bm_couple")
    return out.cpu().numpy()


def g2p(x, v, slots, weights, old_v, new_v, dt, flip):
    """Grid-to-particle transfer (material.py:343-384); returns the new (x, v)."""
    x, v = _dev(x, torch.float32).reshape(-1, 3), _dev(v, torch.float32).reshape(-1, 3)
    n = x.shape[0]
    old_v, new_v = _dev(old_v, torch.float32).reshape(-1, 3), _dev(new_v, torch.float32).reshape(-1, 3)
    if old_v.shape != new_v.shape:
        raise ValueError("old_v and new_v must have the same shape")
    args = [x, v, _dev(slots, torch.int32).reshape(n, 8), _dev(weights, torch.float32).reshape(n, 8), old_v, new_v]
    code = _lib.load().bm_g2p(*[_p(a) for a in args], float(dt), float(flip), n, old_v.shape[0], _stream())
    _lib.check(code, "bm_g2p")
    return x.cpu().numpy(), v.cpu().numpy()


def resist_cut(j, x, v, origin, tool_velocity, orientation_xyzw, tool_angular, half, radius, spacing, damage,
               threshold, cutting="relative"):
    """Cutting resistance at the knife edge (contact.py:182-232) for each particle (x, v,
    damage)[i] under a knife moving at tool_velocity[i]; cutting: "relative" (paper Eq. 5),
    "advance" (the edge's own speed), or 0 for no cutting.  Returns (impulses, new damage,
    work)."""
    x = _dev(x, torch.float32).reshape(-1, 3)
    count = x.shape[0]
    damage = _dev(damage, torch.float32).reshape(count).clone()
    out = torch.zeros(count, 3, dtype=torch.float32, device="cuda")
    work = torch.zeros(1, dtype=torch.float32, device="cuda")
    params = np.concatenate([np.ravel(j), np.ravel(origin), np.ravel(orientation_xyzw), np.ravel(tool_angular),
                             np.ravel(half), [radius, spacing, threshold]]).astype(np.float32)
    if params.size != 19:
        raise ValueError("j, origin, tool_angular and half take 3 values, orientation_xyzw 4")
    if cutting in (0, None, False):
        bits = 0
    elif cutting in ("relative", "advance"):
        bits = CUT_ON | (CUT_ADVANCE if cutting == "advance" else 0)
    else:
        raise ValueError("cutting must be 'relative', 'advance' or 0")
    args = [x, _dev(v, torch.float32).reshape(count, 3), _dev(tool_velocity, torch.float32).reshape(count, 3), damage,
            _dev(params, torch.float32)]
    code = _lib.load().bm_probe_resist_cut(*[_p(a) for a in args], bits, _p(out), _p(work), count, _stream())
    _lib.check(code, "bm_probe_resist_cut")
    return out.cpu().numpy(), damage.cpu().numpy(), float(work.cpu().numpy()[0])
