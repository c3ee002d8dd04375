'''Minimal stand-in for the reference's analytic-density volume as seen by the UPPE propagator, and the UPPE
branch of its Propagate / Report.

The propagator only needs ``vol_dict['object'].GetRelDensity(xp)`` (uppe_kernel.py:70), which for SeaRay's
``AnalyticDensity`` volumes evaluates the input dictionary's 'density lambda' at the ray positions
(volume.py:669-674). A reference volume object can be passed instead; this class exists so that UPPE runs can
be set up without the ray-optics geometry.

``AnalyticBox.Propagate`` stores the five arrays ``track`` returns as the reference's volume does
(volume.py:606-613), and ``Report`` writes them under the reference's file names (volume.py:635-641):
``<basename>_<name>_uppe_wave.npy``, ``_uppe_j``, ``_uppe_chi``, ``_uppe_plasma`` and ``_uppe_plot_ext``, so the
reference's plotter and viewer read them unchanged. ``write_sim_json`` writes the ``<basename>_sim.json``,
``_sources.json`` and ``_diagnostics.json`` dumps of rays.py:117-123.
'''
import json

import numpy as np

UPPE_OUTPUTS = ('uppe_wave', 'uppe_j', 'uppe_chi', 'uppe_plasma', 'uppe_plot_ext')


class AnalyticBox:
    def __init__(self, name):
        self.name = name
        self.vol_dict = None
        self.uppe_wave = self.uppe_j = self.uppe_chi = self.uppe_plasma = self.dom4d = None

    def Initialize(self, input_dict):
        self.vol_dict = input_dict

    def Propagate(self, xp, eikonal, vg, orb=None, dens=None):
        '''The UPPE branch of nonuniform_volume.Propagate (volume.py:606-613): run track and keep its outputs.
        Ray transport through the volume is not part of this path.'''
        from . import uppe_kernel
        if self.vol_dict.get('propagator', 'uppe') != 'uppe':
            raise NotImplementedError('only the UPPE propagator is provided')
        ret = uppe_kernel.track(None, xp, eikonal, vg, self.vol_dict)
        self.uppe_wave, self.uppe_j, self.uppe_chi, self.uppe_plasma, self.dom4d = ret
        return ret

    def Report(self, basename, mks_length=None):
        '''Write the UPPE arrays under the reference's names (volume.py:635-641).'''
        print('    Write UPPE wave data...')
        arrays = (self.uppe_wave, self.uppe_j, self.uppe_chi, self.uppe_plasma, self.dom4d)
        for suffix, arr in zip(UPPE_OUTPUTS, arrays):
            np.save(basename + '_' + self.name + '_' + suffix, arr)

    def GetRelDensity(self, xp):
        x = xp[..., 1]
        y = xp[..., 2]
        z = xp[..., 3]
        r2 = np.einsum('...i,...i', xp[..., 1:3], xp[..., 1:3])
        return self.vol_dict['density lambda'](x, y, z, r2)


def write_sim_json(basename, sim, sources, diagnostics):
    """The run-description dumps of rays.py:117-123 (JSON of the three input lists/dicts)."""
    for suffix, obj in (('_sim.json', sim), ('_sources.json', sources), ('_diagnostics.json', diagnostics)):
        with open(basename + suffix, 'w') as f:
            json.dump(obj, f)
