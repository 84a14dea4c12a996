"""The C ABI library loads and exports every entry point include/gdem.h declares
(no device calls: runs without a GPU)."""
import ctypes
import os
import re

from pacific_amd import gdem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gdem_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(gdem.LIB_PATH)
    names = declared("gdem.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert sorted(gdem.EXPORTED) == declared("gdem.h")


def test_version_string():
    lib = gdem.load_library()
    assert b"gfx950" in lib.gdem_version()


def test_create_rejects_bad_config_without_device_work():
    # a periodic direction with fewer than 3 cells is rejected before any HIP call
    lib = gdem.load_library()
    cfg = gdem._Config()
    for d in range(3):
        cfg.size[d] = 1.0
        cfg.periodic[d] = 1
    cfg.cell_size = 0.6
    h = ctypes.c_void_p()
    assert lib.gdem_create(ctypes.byref(cfg), ctypes.byref(h)) == gdem.ERR_ARG
    assert lib.gdem_create(None, ctypes.byref(h)) == gdem.ERR_ARG


def test_struct_layouts_match_the_header(tmp_path):
    """ctypes mirrors of gdem_config and gdem_batch_system have the C layout
    (sizes and field offsets from a program compiled against the header)."""
    import subprocess
    src = tmp_path / "layout.c"
    src.write_text(
        "#include <stdio.h>\n#include <stddef.h>\n#include \"gdem.h\"\n"
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(gdem_config),"
        " offsetof(gdem_config, cell_size), offsetof(gdem_config, device),"
        " sizeof(gdem_batch_system), offsetof(gdem_batch_system, kind),"
        " offsetof(gdem_batch_system, pp), offsetof(gdem_batch_system, wall_crust));return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)])
    got = [int(v) for v in subprocess.check_output([str(exe)]).split()]
    B = gdem._BatchSystem
    assert got == [ctypes.sizeof(gdem._Config), gdem._Config.cell_size.offset,
                   gdem._Config.device.offset, ctypes.sizeof(B), B.kind.offset, B.pp.offset,
                   B.wall_crust.offset]
