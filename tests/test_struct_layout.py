"""Binary layout of CHAM_tile_t and CHAM_desc_t.

A caller compiled against Chameleon's own header reads desc->mt,
desc->tiles, desc->mat ... at fixed offsets, so the records of
include/chameleon/struct.h must place every field where the reference's
does.  The expected layout is a ctypes restatement of the reference field
list (include/chameleon/struct.h:49-56 for CHAM_tile_t and :152-243 for
CHAM_desc_t, default build without CHAMELEON_KERNELS_TRACE), written here
independently of this repository's header; the actual layout is what gcc
makes of that header (offsetof / sizeof printed by a small C program).
"""
import ctypes
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

FP = ctypes.c_void_p
I32, I64, U32, I8, SZ = ctypes.c_int, ctypes.c_int64, ctypes.c_uint, ctypes.c_int8, ctypes.c_size_t


class RefTile(ctypes.Structure):
    # reference struct.h:49-56: void *mat; int rank, m, n, ld; int8_t format; cham_flttype_t flttype
    _fields_ = [("mat", FP), ("rank", I32), ("m", I32), ("n", I32), ("ld", I32), ("format", I8),
                ("flttype", I32)]


class RefDesc(ctypes.Structure):
    # reference struct.h:152-243, in declaration order
    _fields_ = [("name", FP), ("get_blktile", FP), ("get_blkaddr", FP), ("get_blkldd", FP),
                ("get_blkdim", FP), ("get_rankof", FP), ("get_rankof_init", FP), ("get_rankof_init_arg", FP),
                ("tiles", FP), ("mat", FP), ("A21", SZ), ("A12", SZ), ("A22", SZ), ("styp", I32),
                ("dtyp", I32), ("mb", I32), ("nb", I32), ("bsiz", I32), ("i", I32), ("j", I32), ("m", I32),
                ("n", I32), ("lm", I32), ("ln", I32), ("mt", I32), ("nt", I32), ("lmt", I32), ("lnt", I32),
                ("data_dist", FP), ("dist_it", I32), ("dist_jt", I32), ("dist_mstride", I32),
                ("dist_nstride", I32), ("llm", I32), ("lln", I32), ("llm1", I32), ("lln1", I32),
                ("llmt", I32), ("llnt", I32), ("id", I32), ("occurences", I32), ("sync", I32),
                ("use_mat", I32), ("alloc_mat", I32), ("register_mat", I32), ("myrank", I32), ("ooc", I32),
                ("runtime_view", I32), ("mpitag", I64), ("mpitag_size", I64), ("mpitag_tile_span", I64),
                ("mpitag_tree_span", I64), ("mpitag_ld", I32), ("mpitag_bits", U32), ("schedopt", FP)]


def _probe(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    tile = [f for f, _ in RefTile._fields_]
    desc = [f for f, _ in RefDesc._fields_ if f != "mpitag_bits"]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "chameleon/struct.h"', 'int main(void) {',
             '  printf("tile.sizeof %zu\\n", sizeof(CHAM_tile_t));',
             '  printf("desc.sizeof %zu\\n", sizeof(CHAM_desc_t));']
    lines += [f'  printf("tile.{f} %zu\\n", offsetof(CHAM_tile_t, {f}));' for f in tile]
    lines += [f'  printf("desc.{f} %zu\\n", offsetof(CHAM_desc_t, {f}));' for f in desc]
    # the three one-bit tag flags share the word after mpitag_ld
    lines += ['  CHAM_desc_t d = {0}; d.mpitag_owner = 1; d.mpitag_use = 1; d.mpitag_dist = 1;',
              '  const unsigned char *b = (const unsigned char *)&d; size_t w = 0;',
              '  for (size_t k = 0; k < sizeof d; ++k) if (b[k]) { w = k; break; }',
              '  printf("desc.mpitag_bits %zu\\n", w);', '  return 0; }']
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run([gcc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (l.split() for l in out.splitlines())}


def test_descriptor_and_tile_records_match_reference_offsets(tmp_path):
    got = _probe(tmp_path)
    assert got["tile.sizeof"] == ctypes.sizeof(RefTile)
    for f, _ in RefTile._fields_:
        assert got[f"tile.{f}"] == getattr(RefTile, f).offset, f
    assert got["desc.sizeof"] == ctypes.sizeof(RefDesc)
    for f, _ in RefDesc._fields_:
        assert got[f"desc.{f}"] == getattr(RefDesc, f).offset, f
