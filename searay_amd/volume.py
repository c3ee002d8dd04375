'''Minimal stand-in for the reference's analytic-density volume as seen by the UPPE propagator.

The propagator only needs ``vol_dict['object'].GetRelDensity(xp)`` (uppe_kernel.py:70), which for SeaRay's
``AnalyticDensity`` volumes evaluates the input dictionary's 'density lambda' at the ray positions
(volume.py:669-674). A reference volume object can be passed instead; this class exists so that UPPE runs can
be set up without the ray-optics geometry.
'''
import numpy as np


class AnalyticBox:
    def __init__(self, name):
        self.name = name
        self.vol_dict = None

    def Initialize(self, input_dict):
        self.vol_dict = input_dict

    def GetRelDensity(self, xp):
        x = xp[..., 1]
        y = xp[..., 2]
        z = xp[..., 3]
        r2 = np.einsum('...i,...i', xp[..., 1:3], xp[..., 1:3])
        return self.vol_dict['density lambda'](x, y, z, r2)
