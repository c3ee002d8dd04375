'''Writes tests/golden/reference_rates_chi.npz: input/output vectors of the reference's own rate models, series fit
and air susceptibilities, for the functions that no closed form or stored value pins.

Run only in the build container, where the reference is readable (it never travels with this repository):
    python tests/golden/make_reference_golden.py [/path/to/reference]

It imports the reference's pure numpy/scipy modules (modules/photoionization_base.py, modules/ionization.py,
modules/dispersion.py; none of them imports pyopencl) and records:
  * rates in atomic units of ADK, PPT_Tunneling, PPT, PMPB and StitchedPPT (photoionization_base.py:271-426), cycle
    averaged and static, for O2 (12.07 eV, Z = 0.53) and Ar (15.76 eV, Z = 1) at 800 nm, 80 terms, over fields that
    span Keldysh parameters from about 0.05 to 50 (the PPT/PMPB Fourier-series region included);
  * the quintic coefficients Clog that ionization.Ionizer fits with scipy.optimize.curve_fit (ionization.py:105-109)
    for the series-rate models (PPT, PMPB), and the coefficients Ionizer takes for StitchedPPT;
  * chi(w) of SimpleAir, DryAir and HumidAir (dispersion.py:371-429) on the frequency grids of the workloads, and
    SimpleAir with the UV opacity band of examples/uppe/air_fil.py:36, plus the group velocity of SimpleAir at w = 1.
Only numbers are stored: the inputs and the reference's outputs.
'''
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, 'reference_rates_chi.npz')

E_AU = np.logspace(-3, 0, 48)   # field moduli in atomic units


def main(ref):
    # inputs come from this repository's workload definitions; outputs from the reference
    sys.path.insert(0, ROOT)
    from searay_amd import configs, input_tools
    MKS_LENGTH = configs.MKS_LENGTH
    helper = input_tools.InputHelper(MKS_LENGTH)
    grids = {}
    for name, vol in (('w513', configs.radial_kerr_air()), ('w2049', configs.radial_o2_ionization()),
                      ('w1025', configs.batched_radial(n=1)[0][0])):
        band, nw = vol['frequency band'], vol['wave grid'][0]
        grids[name] = (nw, band[1] - band[0])
    sys.path.insert(0, os.path.join(ref, 'modules'))
    import photoionization_base as izbase
    import ionization
    import dispersion
    out = {'E_au': E_AU}
    au = izbase.AtomicUnits()
    w0_au = au.ta * 299792458.0 / MKS_LENGTH  # w = 1 in simulation units (800 nm), in atomic units
    out['w0_au'] = np.array(w0_au)
    gases = {'o2': (12.07, 0.53), 'ar': (15.76, 1.0)}
    models = {'adk': izbase.ADK, 'ppt_tunneling': izbase.PPT_Tunneling, 'ppt': izbase.PPT, 'pmpb': izbase.PMPB,
              'stitched': izbase.StitchedPPT}
    for gname, (ev, z) in gases.items():
        uion = au.energy_to_au(ev)
        out['uion_au_' + gname] = np.array(uion)
        for mname, cls in models.items():
            for avg in (True, False):
                rg = cls(avg, uion, z, lstar=0, l=0, m=0, w0=w0_au, terms=80)
                out['rate_%s_%s_%s' % (mname, gname, 'avg' if avg else 'static')] = np.asarray(rg.Rate(E_AU.copy()),
                                                                                            dtype=float)
    # the series fits Ionizer makes for the workloads' series-rate models (static rates, sim-unit w0 = 1)
    for gname, (ev, z) in gases.items():
        uion_sim = helper.dnum('%r eV' % ev)
        out['uion_sim_' + gname] = np.array(uion_sim)
        for mname, cls in (('ppt', ionization.PPT), ('pmpb', ionization.PMPB), ('stitched', ionization.StitchedPPT)):
            rg = cls(MKS_LENGTH, False, uion_sim, z, lstar=0, l=0, m=0, w0=1.0, terms=80)
            ion = ionization.Ionizer(rg)
            # (C1, C2, C3): a nonzero C1 selects the closed form, as for StitchedPPT, whose constructor is
            # PPT_Tunneling's; a zero C1 the quintic fit
            out['coeff_%s_%s' % (mname, gname)] = np.array([ion.C1, ion.C2, ion.C3], dtype=float)
            if hasattr(ion, 'Clog'):
                out['clog_%s_%s' % (mname, gname)] = np.asarray(ion.Clog, dtype=float)
    for gname, (nw, band) in grids.items():
        w = np.arange(nw) * (band / nw)
        out[gname] = w
        out['chi_simpleair_' + gname] = np.asarray(dispersion.SimpleAir(MKS_LENGTH).chi(w), dtype=complex)
        out['chi_dryair_' + gname] = np.asarray(dispersion.DryAir(MKS_LENGTH).chi(w), dtype=complex)
        out['chi_humidair_' + gname] = np.asarray(dispersion.HumidAir(MKS_LENGTH, 0.5).chi(w), dtype=complex)
        air = dispersion.SimpleAir(MKS_LENGTH)
        cm, um = 100 * MKS_LENGTH, 1e6 * MKS_LENGTH  # the example's unit names: x / cm is x centimetres
        air.add_opacity_region(0.1 / cm, 0.01 / um, 0.3 / um)
        out['chi_simpleair_uv_' + gname] = np.asarray(air.chi(w), dtype=complex)
    out['vg_simpleair_w1'] = np.asarray(dispersion.SimpleAir(MKS_LENGTH).GroupVelocityMagnitude(1.0), dtype=float)
    np.savez_compressed(OUT, **out)
    print('wrote %s (%d arrays, %d bytes)' % (OUT, len(out), os.path.getsize(OUT)))


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
