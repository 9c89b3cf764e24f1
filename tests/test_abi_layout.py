"""The ctypes mirrors in mrilab_amd/abi.py against the C header.

A C program compiled with gcc from include/mrilab_amd.h prints sizeof and
offsetof of every field that the ctypes structures name; each must equal
what ctypes computes, so a field added to one side only fails here rather
than as a shifted struct on the GPU.
"""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from mrilab_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PAIRS = [
    ("mrs_vobj", abi.VObjC), ("mrs_vmag", abi.VMagC), ("mrs_vcoi", abi.VCoiC), ("mrs_vctl", abi.VCtlC),
    ("mrs_vseq", abi.VSeqC), ("mrs_vvar", abi.VVarC), ("mrs_vsig", abi.VSigC),
    ("mrs_callbacks", abi.CallbacksC), ("mrs_ext_state", abi.ExtStateC), ("mrs_options", abi.OptionsC), ("mrs_plan_info", abi.PlanInfoC),
    ("mrs_batch_info", abi.BatchInfoC), ("mrs_excitation", abi.ExcitationC),
    ("mrs_adjoint_info", abi.AdjointInfoC), ("mrs_recon_desc", abi.ReconDescC), ("mrs_vimg", abi.VImgC),
    ("mrs_sar_obj", abi.SarObjC), ("mrs_sar_sig", abi.SarSigC), ("mrs_seq_param", abi.SeqParamC),
    ("mrs_seq_lines", abi.SeqLinesC),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_layout_matches_header(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "mrilab_amd.h"', "int main(void) {"]
    expect = []
    for cname, py in PAIRS:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        expect.append((cname, "sizeof", C.sizeof(py)))
        for f, _ in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {f}));')
            expect.append((cname, f, getattr(py, f).offset))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert len(got) == len(expect)
    bad = [(c, f, want, int(g)) for (c, f, want), g in zip(expect, got) if int(g) != want]
    assert not bad, f"(struct, field, ctypes, C) mismatches: {bad}"
