'''Minimal stand-in for the reference's analytic-density volume as seen by the wave propagators, and the UPPE and
paraxial branches of its Propagate / Report.

The propagator only needs ``vol_dict['object'].GetRelDensity(xp)`` (uppe_kernel.py:70), which for SeaRay's
``AnalyticDensity`` volumes evaluates the input dictionary's 'density lambda' at the ray positions
(volume.py:669-674). A reference volume object can be passed instead; this class exists so that UPPE runs can
be set up without the ray-optics geometry.

``AnalyticBox.Propagate`` runs the propagator the dictionary names ('uppe' or 'paraxial') and stores the five
arrays ``track`` returns as the reference's volume does (volume.py:598-613), and ``Report`` writes them under
the reference's file names (volume.py:627-641): ``<basename>_<name>_uppe_wave.npy``, ``_uppe_j``, ``_uppe_chi``,
``_uppe_plasma`` and ``_uppe_plot_ext`` (``_paraxial_...`` for the envelope), so the reference's plotter and
viewer read them unchanged. ``write_sim_json`` writes the ``<basename>_sim.json``,
``_sources.json`` and ``_diagnostics.json`` dumps of rays.py:117-123.
'''
import json

import numpy as np

UPPE_OUTPUTS = ('uppe_wave', 'uppe_j', 'uppe_chi', 'uppe_plasma', 'uppe_plot_ext')
PARAXIAL_OUTPUTS = ('paraxial_wave', 'paraxial_j', 'paraxial_chi', 'paraxial_plasma', 'paraxial_plot_ext')


class AnalyticBox:
    def __init__(self, name):
        self.name = name
        self.vol_dict = None
        self.uppe_wave = self.uppe_j = self.uppe_chi = self.uppe_plasma = self.dom4d = None
        self.paraxial_wave = self.paraxial_j = self.paraxial_chi = self.paraxial_plasma = None

    def Initialize(self, input_dict):
        self.vol_dict = input_dict

    def Propagate(self, xp, eikonal, vg, orb=None, dens=None):
        '''The wave branches of nonuniform_volume.Propagate (volume.py:598-613): run the propagator's track and keep
        its outputs. Ray transport through the volume is not part of this path.'''
        from . import paraxial_kernel, uppe_kernel
        prop = self.vol_dict.get('propagator', 'uppe')
        if prop == 'paraxial':
            ret = paraxial_kernel.track(None, xp, eikonal, vg, self.vol_dict)
            self.paraxial_wave, self.paraxial_j, self.paraxial_chi, self.paraxial_plasma, self.dom4d = ret
        elif prop == 'uppe':
            ret = uppe_kernel.track(None, xp, eikonal, vg, self.vol_dict)
            self.uppe_wave, self.uppe_j, self.uppe_chi, self.uppe_plasma, self.dom4d = ret
        else:
            raise NotImplementedError("only the 'uppe' and 'paraxial' propagators are provided")
        return ret

    def Report(self, basename, mks_length=None):
        '''Write the wave arrays under the reference's names (volume.py:627-641).'''
        if (self.vol_dict or {}).get('propagator', 'uppe') == 'paraxial':
            print('    Write paraxial wave data...')
            names = PARAXIAL_OUTPUTS
            arrays = (self.paraxial_wave, self.paraxial_j, self.paraxial_chi, self.paraxial_plasma, self.dom4d)
        else:
            print('    Write UPPE wave data...')
            names = UPPE_OUTPUTS
            arrays = (self.uppe_wave, self.uppe_j, self.uppe_chi, self.uppe_plasma, self.dom4d)
        for suffix, arr in zip(names, arrays):
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
