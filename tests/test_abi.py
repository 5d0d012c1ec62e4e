"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
entry point include/pybella_amd.h declares, the ctypes table binds exactly
those, the header's structs match the ctypes layouts, and the product fails
loudly (no CPU fallback) when no GPU is present.  No compute call is made."""
import ctypes as C
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "pybella_amd.h"


def header_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(pb_[a-z0-9_]+)\s*\(", text)))


def header_struct_fields(name):
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    m = re.search(r"typedef\s+struct\s+" + name + r"\s*\{(.*?)\}\s*" + name + r"\s*;", text, flags=re.S)
    assert m, name
    fields = []
    for decl in m.group(1).split(";"):
        decl = decl.strip()
        if decl:   # "double a, b[3]" -> a, b
            for part in decl.split(","):
                fields.append(re.findall(r"([A-Za-z_]\w*)\s*(?:\[\s*\d+\s*\])?\s*$", part.strip())[0])
    return fields


@pytest.fixture(scope="module")
def lib():
    from pybella_amd import _abi
    if not _abi.LIB_PATH.exists():
        pytest.fail(f"{_abi.LIB_PATH} is not built (run __graft_entry__.build())")
    return _abi.load()


def test_every_header_symbol_is_exported(lib):
    names = header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    from pybella_amd import _abi
    assert sorted(_abi.SIGNATURES) == header_functions()


def test_struct_layouts_match_header():
    from pybella_amd import _abi
    for cname, py in (("pb_grid2d", _abi.Grid2D), ("pb_phys", _abi.Phys),
                      ("pb_solve_opts", _abi.SolveOpts), ("pb_solve_info", _abi.SolveInfo)):
        assert header_struct_fields(cname) == [f for f, _ in py._fields_], cname


def test_version_without_gpu(lib):
    assert lib.pb_version() > 0


def test_no_cpu_fallback_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from pybella_amd import _abi
    with pytest.raises(_abi.HipError):
        _abi.Context()


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


@pytest.mark.parametrize("zc", [12, 16, 24])
def test_march_blocks_cover_planes_once(lib, zc):
    """The 3D march's block -> (x tile, y tile, z chunk) mapping, run on the
    host from the library: for 257 x 129 x 129 nodes in zc-plane chunks
    (grid z = ceil(129 / zc)) it is a bijection onto the tiles, every chunk
    starts inside the grid, and the chunks of each (x, y) tile cover planes
    0..128 exactly once."""
    nx, ny, nz = 257, 129, 129
    gx, gy, gz = -(-nx // 64), -(-ny // 8), -(-nz // zc)
    nb = gx * gy * gz
    out = (C.c_int * (3 * nb))()
    assert lib.pb_op3d_march_map(gx, gy, gz, out) == 0
    tiles = np.array(out[:]).reshape(nb, 3)
    assert len({tuple(t) for t in tiles}) == nb
    assert tiles[:, 0].max() == gx - 1 and tiles[:, 1].max() == gy - 1 and tiles[:, 2].max() == gz - 1
    assert np.all(tiles[:, 2] * zc < nz)
    cover = np.zeros((gx, gy, nz), dtype=int)
    for bx, by, bz in tiles:
        kb = bz * zc
        cover[bx, by, kb:min(kb + zc, nz)] += 1
    assert np.all(cover == 1)
