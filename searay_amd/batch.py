'''Batched radial UPPE: many independent radial instances stepped together on one GPU.

The reference steps one grid per ``propagator`` call (uppe_kernel.py:260-331); a parameter scan over pulse
energies, gases and grids runs its instances one after another. ``Batch`` keeps that per-instance API -- every
instance is a ``vol_dict`` of the reference's format, set up by ``uppe_kernel.RunSetup`` exactly as ``track``
sets it up, with its own Hankel basis, medium, ionizer and field -- and advances all instances through
``uppe_batch_advance``: members of equal (Nw, Nr, Nk, precision) share every kernel launch of a step, and the
launch shapes of each such run are chosen by timing candidates on the run (autotune).

    b = Batch([vol_0, vol_1, ...], precision='fp32')
    b.set_fields()                 # each instance's 'incoming wave'
    steps = b.advance()            # the first diagnostic interval of every instance
    q = b.members[i].get_state()   # or b.members[i].finish(zf - zi)
'''
import ctypes as ct

import numpy as np

from . import _abi
from . import _native
from . import uppe_kernel


class BatchError(RuntimeError):
    pass


class Batch:
    def __init__(self, vols, precision='fp32', device=0, autotune=True, graph_steps=8):
        if len(vols) < 1:
            raise ValueError('a batch needs at least one instance')
        self.lib = _native.library()
        self.setups = [uppe_kernel.RunSetup(v) for v in vols]
        prec = _abi.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        n = len(self.setups)
        descs = (_abi.uppe_desc * n)()
        for i, rs in enumerate(self.setups):
            if rs.geometry != _abi.UPPE_RADIAL or rs.N[2] != 1:
                raise ValueError('instance %d: batches take radial grids with one azimuthal node' % i)
            T = rs.field_tool.GetTransverseTool()
            descs[i] = _abi.uppe_desc(_abi.UPPE_RADIAL, rs.N[0], rs.N[1], 1, T.Nk, prec, device, _abi.MODEL_UPPE)
        b = ct.c_void_p()
        rc = self.lib.uppe_batch_create(descs, n, ct.byref(b))
        if rc != _abi.UPPE_OK:
            raise BatchError('uppe_batch_create failed (status %d): %s'
                             % (rc, (self.lib.uppe_batch_last_error(b) or b'').decode()))
        self.b = b
        self._check(self.lib.uppe_batch_set_autotune(self.b, int(bool(autotune))), 'set_autotune')
        self.set_graph_steps(graph_steps)
        self.members = []
        for i, rs in enumerate(self.setups):
            h = ct.c_void_p()
            self._check(self.lib.uppe_batch_member(self.b, i, ct.byref(h)), 'member')
            m = _native.BatchMember(h, (rs.N[0], rs.N[1], 1))
            rs.make_stepper(lambda *args, m=m, **kw: m)  # Hankel basis, medium of the first subcycle, rotator
            self.members.append(m)

    def set_graph_steps(self, steps):
        '''Steps of every run that one HIP graph replays (0: launch kernel by kernel).'''
        self._check(self.lib.uppe_batch_set_graph(self.b, int(steps)), 'set_graph')

    def _check(self, rc, what):
        if rc != _abi.UPPE_OK:
            msg = self.lib.uppe_batch_last_error(self.b)
            raise BatchError('uppe_batch_%s failed (status %d): %s' % (what, rc, msg.decode() if msg else ''))

    def set_fields(self, fields=None):
        '''Load every member's initial field (default: each instance's 'incoming wave').'''
        for i, (rs, m) in enumerate(zip(self.setups, self.members)):
            m.set_field(rs.stepper_field() if fields is None else fields[i])

    def controls(self, max_steps=0, plane=0, dphi=0.4, rel_amin=1e-2):
        '''The step control of every member for diagnostic interval `plane`, as propagator builds it.'''
        c = (_abi.uppe_step_control * len(self.setups))()
        for i, rs in enumerate(self.setups):
            zi, zf, L = rs.plane(plane)
            c[i] = _abi.uppe_step_control(float(zi), float(L), 0, int(max_steps), float(rs.dzmin), float(dphi),
                                          float(rel_amin), int(rs.subcycles), 0, None, None, 0, 0)
        return c

    def advance(self, max_steps=0, plane=0):
        '''Step every member through its interval `plane` (or max_steps steps); returns the step counts. The members
        keep the medium of their first subcycle, so an instance whose density changes along z is refused.'''
        for i, rs in enumerate(self.setups):
            if not rs.medium_is_uniform(plane):
                raise ValueError('instance %d: the density varies along z in interval %d; a batch loads the medium '
                                 'once (step such an instance with track)' % (i, plane))
        c = self.controls(max_steps, plane)
        steps = (ct.c_int * len(self.setups))()
        self._check(self.lib.uppe_batch_advance(self.b, c, steps), 'advance')
        return np.array(steps[:], dtype=int)

    def plan(self, i):
        '''Launch shapes of member i's run in the last advance.'''
        out = (ct.c_int * 10)()
        self._check(self.lib.uppe_batch_plan(self.b, int(i), out), 'plan')
        v = list(out)
        return dict(k2r=dict(tiles=v[0], waves=v[1], row_tiles=v[2]), r2k=dict(tiles=v[3], waves=v[4], row_tiles=v[5]),
                    source_samples_per_thread=v[6], run_length=v[7], first_slot=v[8], group=v[9])

    def synchronize(self):
        for m in self.members:
            m.synchronize()

    def close(self):
        if getattr(self, 'b', None):
            for m in self.members:
                m.close()
            self.lib.uppe_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
