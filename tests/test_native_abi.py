"""The C ABI: libpmfdtd loads (no GPU needed), exports every function the
header declares, and the ctypes mirror has the header's struct layout."""

import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from prismo_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "pmfdtd.h"
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(pm_\w+)\s*\(", h.read_text(),
                                re.M))
    return sorted(names)


def test_header_declares_what_binding_binds():
    assert set(declared_functions()) == (set(N.SIGNATURES) | set(N.BATCH_SIGNATURES)
                                         | set(N.MLP_SIGNATURES) | set(N.GEOM_SIGNATURES))


def test_library_loads_and_exports_all_symbols():
    lib = N.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.pm_abi_version() == N.ABI_VERSION


def test_errors_without_device_are_reported():
    lib = N.load()
    h = C.c_void_p()
    rc = lib.pm_create(10_000, None, C.byref(h))
    assert rc != 0 and not h.value


PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "pmfdtd.h"
#include "pmbatch.h"
#include "pmgeom.h"
#define P(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  %FIELDS%
  return 0;
}
"""


STRUCTS = (N.pm_plan, N.pm_source, N.pm_sample_box, N.pm_dft_group, N.pm_ade_box,
           N.pm_batch_line, N.pm_batch_design, N.pm_batch_plan, N.pm_shape)


def test_struct_layout_matches_header(tmp_path):
    lines = []
    for cls in STRUCTS:
        lines.append(f'printf("{cls.__name__} %zu\\n", sizeof({cls.__name__}));')
        for fname, _ in cls._fields_:
            lines.append(f"P({cls.__name__}, {fname})")
    src = tmp_path / "probe.c"
    src.write_text(PROBE.replace("%FIELDS%", "\n  ".join(lines)))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", f"-I{HEADER.parent}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(l.rsplit(" ", 1) for l in out if l.strip())
    for cls in STRUCTS:
        assert int(got[cls.__name__]) == C.sizeof(cls), cls.__name__
        for fname, _ in cls._fields_:
            assert int(got[f"{cls.__name__}.{fname}"]) == getattr(cls, fname).offset, (cls, fname)
