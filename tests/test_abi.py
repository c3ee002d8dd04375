'''The C ABI: the HIP library loads and exports every entry point include/searay_uppe.h declares, the oracle
exports the same set with the orc_ prefix, and the ctypes mirror matches the header's struct layouts.
No compute call is made here (no GPU needed).'''
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
