"""The ctypes mirrors of the public structs in include/toposzp_hip.h have the C layout:
a small C program built with gcc against the header prints each struct's size and
every member's offset, which must equal ctypes' sizeof / offsets (a mismatch would
make the library read or write the wrong bytes of a caller's struct)."""
import ctypes
import os
import shutil
import subprocess
import tempfile

import pytest

import toposzp_hip as tz

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# C type name -> ctypes class; member names are the ctypes field names, which follow
# the header's member names
STRUCTS = {
    "tszp_topo_info": tz.TopoInfo,
    "tszp_restore_info": tz.RestoreInfo,
    "tszp_batch_field": tz.BatchField,
    "tszp_topo_batch_field": tz.TopoBatchField,
    "tszp_ticket_info": tz.TicketInfo,
    "tszp_server_stats_t": tz.ServerStats,
    "tszp_quality": tz.Quality,
}


def c_layouts():
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "toposzp_hip.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("  return 0;\n}")
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "layout.c")
        exe = os.path.join(tmp, "layout")
        open(src, "w").write("\n".join(lines) + "\n")
        subprocess.run([gcc, "-std=c11", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        cname, what, val = line.split()
        got[(cname, what)] = int(val)
    return got


def test_ctypes_structs_match_the_header():
    got = c_layouts()
    for cname, cls in STRUCTS.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
