'''The C ABI: the HIP library loads and exports every entry point include/searay_uppe.h declares, the oracle
exports the same set with the orc_ prefix, and the ctypes mirror matches the header's struct layouts.
No compute call is made here (no GPU needed); the argument checks that run before any device call are
exercised: bad descriptors and null handles.'''
import ctypes as ct
import os
import re
import subprocess

import pytest

from searay_amd import _abi, _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'searay_uppe.h')


def header_functions():
    text = open(HEADER).read()
    return re.findall(r'^\s*(?:int|const char\s*\*)\s*(uppe_\w+)\s*\(', text, re.M)


def test_header_lists_entry_points():
    names = header_functions()
    for ep in _abi.ENTRY_POINTS + _abi.HIP_ONLY_ENTRY_POINTS:
        assert 'uppe_' + ep in names
    assert len(names) == len(set(names))


def test_hip_library_exports_every_symbol():
    lib = _native.library()
    assert os.path.basename(_native.LIB_PATH) == 'libsearay_uppe.so'
    for name in header_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(['nm', '-D', '--defined-only', _native.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r'\bT (uppe_\w+)', out))
    assert set(header_functions()) <= exported


def test_oracle_exports_same_entry_points(oracle_mod):
    lib = oracle_mod.library()
    for ep in _abi.ENTRY_POINTS:
        assert hasattr(lib, 'orc_uppe_' + ep)


def test_struct_layouts_match_header():
    '''Compile a probe against the header and compare sizeof/offsetof with the ctypes mirror.'''
    probe = r'''
#include <stdio.h>
#include <stddef.h>
#include "searay_uppe.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(uppe_desc), sizeof(uppe_ionization), sizeof(uppe_medium),
         sizeof(uppe_step_control), sizeof(uppe_info));
  printf("%zu %zu %zu %zu\n", offsetof(uppe_medium, ion), offsetof(uppe_step_control, dz_schedule),
         offsetof(uppe_step_control, record_cap), offsetof(uppe_ionization, c));
  printf("%zu %zu %zu\n", sizeof(uppe_rotator), offsetof(uppe_rotator, wdt), offsetof(uppe_rotator, hbar));
  return 0;
}
'''
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, 'p.c')
        exe = os.path.join(d, 'p')
        open(src, 'w').write(probe)
        subprocess.run(['gcc', '-I', os.path.join(ROOT, 'include'), src, '-o', exe], check=True)
        lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split('\n')
    sizes = [int(v) for v in lines[0].split()]
    offs = [int(v) for v in lines[1].split()]
    assert sizes == [ct.sizeof(_abi.uppe_desc), ct.sizeof(_abi.uppe_ionization), ct.sizeof(_abi.uppe_medium),
                     ct.sizeof(_abi.uppe_step_control), ct.sizeof(_abi.uppe_info)]
    assert offs == [_abi.uppe_medium.ion.offset, _abi.uppe_step_control.dz_schedule.offset,
                    _abi.uppe_step_control.record_cap.offset, _abi.uppe_ionization.c.offset]
    rot = [int(v) for v in lines[2].split()]
    assert rot == [ct.sizeof(_abi.uppe_rotator), _abi.uppe_rotator.wdt.offset, _abi.uppe_rotator.hbar.offset]


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_native, 'LIB_PATH', str(tmp_path / 'nope.so'))
    monkeypatch.setattr(_native, '_lib', None)
    with pytest.raises(ImportError):
        _native.library()


def _desc(**kw):
    d = dict(geometry=_abi.UPPE_RADIAL, nw=65, n1=16, n2=1, nk=16, precision=_abi.PREC_FP32, device=0,
             model=_abi.MODEL_UPPE)
    d.update(kw)
    return _abi.uppe_desc(**d)


@pytest.mark.parametrize('bad', [dict(nw=64), dict(nw=2), dict(n1=12), dict(n2=3), dict(nk=0), dict(nk=17),
                                 dict(precision=4), dict(precision=-1), dict(geometry=2), dict(model=2),
                                 dict(model=_abi.MODEL_PARAXIAL, nw=65)])
def test_create_rejects_bad_descriptors_before_any_device_call(bad):
    '''The descriptor rules of the oracle's create (2^n + 1 frequencies for the UPPE, 2^n for the envelope, powers
    of two for the cells, 1 <= nk <= n1, a known precision, geometry and model) hold in the HIP library, and are
    checked before it touches the HIP runtime: no GPU is needed to see UPPE_ERR_ARG.'''
    lib = _native.library()
    h = ct.c_void_p(1)
    assert lib.uppe_create(ct.byref(_desc(**bad)), ct.byref(h)) == _abi.UPPE_ERR_ARG
    assert not h.value
    assert lib.uppe_create(None, ct.byref(h)) == _abi.UPPE_ERR_ARG
    assert lib.uppe_create(ct.byref(_desc()), None) == _abi.UPPE_ERR_ARG


def test_null_handles_are_rejected():
    '''Every handle entry point returns UPPE_ERR_ARG for a null handle (destroy of a null handle is a no-op, as
    free is), without a device call.'''
    lib = _native.library()
    buf = (ct.c_double * 8)()
    sc = _abi.uppe_step_control(length=1.0, n_sub=1)
    z = ct.c_double(0.0)
    calls = {
        'set_hankel': (None, buf, buf), 'set_medium': (None, None), 'set_paraxial_medium': (None, None),
        'set_rotator': (None, None), 'set_field': (None, buf), 'advance': (None, ct.byref(sc), None),
        'advance_static': (None, ct.byref(sc), None), 'advance_checkpointed': (None, ct.byref(sc), 10, None),
        'adjoint': (None, z, buf, buf), 'finish': (None, z, None, None, None, None), 'get_state': (None, buf),
        'get_info': (None, None), 'profile_enable': (None, 1), 'profile_read': (None, 0, None, None),
        'synchronize': (None,),
    }
    for name, args in calls.items():
        assert getattr(lib, 'uppe_' + name)(*args) == _abi.UPPE_ERR_ARG, name
    assert lib.uppe_destroy(None) == _abi.UPPE_OK
    lib.uppe_last_error.restype = ct.c_char_p
    assert lib.uppe_last_error(None) == b'null handle'
    b = ct.c_void_p(1)
    assert lib.uppe_batch_create(None, 1, ct.byref(b)) == _abi.UPPE_ERR_ARG
    assert lib.uppe_batch_create(ct.byref(_desc()), 0, ct.byref(b)) == _abi.UPPE_ERR_ARG
    for name, args in {'batch_member': (None, 0, ct.byref(b)), 'batch_advance': (None, ct.byref(sc), None),
                       'batch_set_autotune': (None, 1), 'batch_set_graph': (None, 8),
                       'batch_plan': (None, 0, buf)}.items():
        assert getattr(lib, 'uppe_' + name)(*args) == _abi.UPPE_ERR_ARG, name
    assert lib.uppe_batch_destroy(None) == _abi.UPPE_OK
    assert lib.uppe_batch_last_error(None) == b'null batch'
