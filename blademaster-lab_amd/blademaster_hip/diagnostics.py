"""Kinetic, gravitational and elastic energy of the material state (diagnostics.py:7-48).

One wave-reduced kernel over the particles; the columns and units are the reference's.
"""

import ctypes
import numpy as np
import torch


class Metrics:
    columns = ["kinetic_J", "gravity_J", "elastic_diagnostic_J", "max_speed_m_s", "max_z_m", "min_det_F"]

    def __init__(self, solver):
        self.s = solver
        self.array = torch.zeros(6, dtype=torch.float32, device=solver.device.torch)

    def read(self):
        self.array.copy_(torch.tensor([0, 0, 0, 0, -1.0e20, 1.0e20], dtype=torch.float32))
        self.s.launch("bm_metrics", ctypes.c_void_p(self.array.data_ptr()))
        return self.array.cpu().numpy().copy()
