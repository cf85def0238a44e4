"""The C ABI: libarcsim loads and exports every function include/*.h
declares; the host types have the reference's layout; the Python mirror
declares the reference's ctypes signatures.  No device calls."""

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "arc-demo_amd" / "arc" / "libarcsim.so"
HEADERS = sorted((ROOT / "include").glob("*.h"))

# the 13 functions of the reference's csrc/arc_sim.h:25-56
REFERENCE_API = ["arc_sim_create", "arc_sim_destroy", "arc_sim_set_threads", "arc_sim_num_worlds",
                 "arc_sim_num_joints", "arc_sim_num_legs", "arc_sim_robot_info", "arc_sim_reset",
                 "arc_sim_step", "arc_sim_observe", "arc_sim_heights", "arc_sim_terrain_at", "arc_sim_world"]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(arc\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not LIB.exists():
        pytest.skip("libarcsim.so not built")
    return ctypes.CDLL(str(LIB))


def test_headers_declare_reference_api():
    names = declared_functions()
    for f in REFERENCE_API:
        assert f in names


def test_every_declared_symbol_exported(lib):
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_no_device_counts_without_gpu(lib):
    lib.arcx_device_count.restype = ctypes.c_int
    assert lib.arcx_device_count() >= 0
    lib.arc_sim_num_joints.restype = ctypes.c_int
    lib.arc_sim_num_legs.restype = ctypes.c_int
    assert lib.arc_sim_num_joints() == 12 and lib.arc_sim_num_legs() == 4


def test_world_layout(tmp_path):
    """sizeof(arc_world) and field offsets equal the reference's (the renderer
    reads them): compiled with the system C compiler."""
    src = tmp_path / "l.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "arc_sim.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(arc_world), offsetof(arc_world, bodies),'
                   ' offsetof(arc_world, terrain), sizeof(arc_body), sizeof(arc_joint));return 0;}\n')
    exe = tmp_path / "l"
    subprocess.run(["gcc", "-std=c89", "-pedantic", "-Wall", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert int(out[0]) == 62960
    assert int(out[3]) == 368 and int(out[4]) == 296


def test_headers_compile_as_cxx(tmp_path):
    src = tmp_path / "c.cpp"
    src.write_text('#include "arc_gpu.h"\nint main(){return arc_sim_num_joints() == 12 ? 0 : 1;}\n')
    subprocess.run(["g++", "-fsyntax-only", f"-I{ROOT / 'include'}", str(src)], check=True)


def test_sim_raises_without_gpu():
    import sys
    sys.path.insert(0, str(ROOT / "arc-demo_amd"))
    from arc import native
    if native.sim_lib().arcx_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(native.SimError):
        native.Sim(4)
