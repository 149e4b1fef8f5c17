"""C ABI checks that need no GPU: the library loads, exports every function
include/rp_abi.h declares, and its struct layouts match the Python mirror."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from rusty_particles_amd import abi, native, scenes
from oracle import read_stl_numpy

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rp_abi.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rp_[a-z_0-9]+)\s*\(", text)))


def test_header_and_mirror_agree():
    assert header_functions() == sorted(abi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = native.lib()
    for name in header_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", native._LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (rp_[a-z_0-9]+)", out))
    assert set(header_functions()) <= exported


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "lay.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "rp_abi.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(rp_params), offsetof(rp_params, particle),'
        ' offsetof(rp_params, gravity), offsetof(rp_params, skin), sizeof(rp_material), sizeof(rp_stats),'
        ' offsetof(rp_stats, excessive_overlap)); return 0;}\n')
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [C.sizeof(abi.RpParams), abi.RpParams.particle.offset, abi.RpParams.gravity.offset,
            abi.RpParams.skin.offset, C.sizeof(abi.RpMaterial), C.sizeof(abi.RpStats),
            abi.RpStats.excessive_overlap.offset]
    assert got == want


def test_params_default_matches_reference_defaults():
    """Simulation::new defaults (src/simulation.rs:62-91) via the C function."""
    lib = native.lib()
    p = abi.RpParams()
    lo = (C.c_double * 3)(-1, -2, -3)
    hi = (C.c_double * 3)(1, 2, 3)
    lib.rp_params_default(C.byref(p), 1e-3, lo, hi)
    q = abi.default_params(1e-3, (-1, -2, -3), (1, 2, 3))
    assert bytes(p) == bytes(q)
    assert (p.particle.youngs_modulus, p.particle.poissons_ratio, p.particle.density) == (1e7, 0.3, 2500.0)
    assert (p.wall.youngs_modulus, p.wall.density) == (1e9, 7800.0)
    assert p.normal_model == abi.RP_NORMAL_HERTZ and p.tangential_model == abi.RP_TANGENTIAL_MINDLIN
    assert list(p.gravity) == [0.0, -9.81, 0.0] and p.ref_cell_size == 1.0


def test_engine_create_without_device_fails_cleanly():
    if native.device_count() > 0:
        pytest.skip("a HIP device is present")
    p = abi.default_params(1e-3, (-1, -1, -1), (1, 1, 1))
    with pytest.raises(native.EngineError) as ei:
        native.Engine(p)
    assert ei.value.code == abi.RP_ERR_DEVICE
    assert "no HIP device" in native.lib().rp_last_error().decode()


def test_invalid_params_rejected():
    lib = native.lib()
    p = abi.default_params(-1.0, (-1, -1, -1), (1, 1, 1))
    h = C.c_void_p()
    assert lib.rp_engine_create(C.byref(p), C.byref(h)) == abi.RP_ERR_INVALID
    p = abi.default_params(1e-3, (-1, -1, -1), (1, 1, 1))
    p.normal_model = 9
    assert lib.rp_engine_create(C.byref(p), C.byref(h)) == abi.RP_ERR_INVALID


@pytest.mark.parametrize("fmt", ["binary", "ascii"])
def test_stl_reader_matches_independent_parser(tmp_path, fmt):
    v9, n3 = scenes.hopper_mesh(0.008, 30.0, 0.075, nseg=16, nring=4)
    path = str(tmp_path / f"h.{fmt}.stl")
    (scenes.write_stl_binary if fmt == "binary" else scenes.write_stl_ascii)(path, v9, n3)
    a9, a3 = native.read_stl(path)
    b9, b3 = read_stl_numpy(path)
    assert a9.shape == (128, 9) and np.array_equal(a9, b9) and np.array_equal(a3, b3)
    assert np.array_equal(a9, v9.astype(np.float32))


def test_stl_reader_errors(tmp_path):
    with pytest.raises(OSError):
        native.read_stl(str(tmp_path / "missing.stl"))
    bad = tmp_path / "trunc.stl"
    bad.write_bytes(b"\0" * 80 + (5).to_bytes(4, "little") + b"\0" * 60)
    with pytest.raises(OSError):
        native.read_stl(str(bad))
