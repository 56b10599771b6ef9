"""The ctypes mirrors of the C-ABI structs have the header's layout: a C
program built from include/oe/oe.h prints sizeof and offsetof of every
field (by name), compared with the ctypes structures."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from openems_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = {
    "Excitation": "oe_excitation",
    "EnergySample": "oe_energy_sample",
    "BatchDesc": "oe_batch_desc",
    "BatchResultC": "oe_batch_result",
    "Waveform": "oe_waveform",
    "DftConfig": "oe_dft_config",
    "FarField": "oe_far_field",
    "SweepConfig": "oe_sweep_config",
    "SweepEntry": "oe_sweep_entry",
    "ExtrapDims": "oe_extrap_dims",
    "CpmlConfig": "oe_cpml_config",
    "HostCallbackConfig": "oe_host_callback_config",
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_structs_match_header(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "oe/oe.h"', "int main(void) {"]
    expect = []
    for py, cname in STRUCTS.items():
        cls = getattr(_ffi, py)
        lines.append(f'    printf("%zu\\n", sizeof({cname}));')
        expect.append((f"sizeof {cname}", C.sizeof(cls)))
        for fname, _ in cls._fields_:
            lines.append(f'    printf("%zu\\n", offsetof({cname}, {fname}));')
            expect.append((f"{cname}.{fname}", getattr(cls, fname).offset))
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert len(got) == len(expect)
    bad = [(name, want, have) for (name, want), have in zip(expect, got) if want != have]
    assert not bad, bad
