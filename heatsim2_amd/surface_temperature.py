"""Surface temperature of an insulating z-min face from the first two layers.

insulating_z_min_surface_temperature (surface_temperature.py:4-36 upstream)
averages two extrapolations to z = 0 from the cell centres at dz/2 and
3dz/2: the quadratic with zero slope at the insulated face,
T0 - (T1 - T0)/8, and the linear one, T0 - (T1 - T0)/2, giving
T0 - (5/16)(T1 - T0).

Computed on the GPU by the HIP library (hs2_surface_temperature).  A device
tensor gives a device tensor; a numpy array is copied to the device and the
result copied back.
"""
import numpy as np


def insulating_z_min_surface_temperature(T, dz):
    import torch
    from heatsim2_amd import _lib
    if isinstance(T, torch.Tensor):
        return _lib.surface_temperature_device(T, dz)
    dev = _lib.device()
    out = _lib.surface_temperature_device(torch.from_numpy(np.ascontiguousarray(T, np.float64)).to(dev), dz)
    return out.cpu().numpy()
