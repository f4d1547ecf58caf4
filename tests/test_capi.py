"""The C ABI (CPU-only checks, no compute calls): libabd_hip.so loads, exports
every entry point include/abd_hip.h declares, and the ctypes mirrors of AbdParams
/ AbdScene match the C layout compiled from the header."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "abd_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(abd_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from abd_hip import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 28
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every binding the Python mirror sets up exists in the header too
    L = _lib.load()
    assert L.abd_version() > 0
    assert _lib.error_string(_lib.ABD_ERR_UNSORTED).startswith("candidate list")


def test_struct_layouts_match_header(tmp_path):
    from abd_hip import _lib
    from abd_hip.const_params import AbdParamsC
    src = tmp_path / "layout.c"
    fields_p = [f for f, _ in AbdParamsC._fields_]
    fields_s = [f for f, _ in _lib.AbdScene._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void) {',
             'printf("%zu %zu\\n", sizeof(AbdParams), sizeof(AbdScene));']
    lines += [f'printf("%zu\\n", offsetof(AbdParams, {f}));' for f in fields_p]
    lines += [f'printf("%zu\\n", offsetof(AbdScene, {f}));' for f in fields_s]
    lines += ['return 0; }']
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    vals = [int(v) for v in out]
    assert vals[0] == ctypes.sizeof(AbdParamsC)
    assert vals[1] == ctypes.sizeof(_lib.AbdScene)
    for k, f in enumerate(fields_p):
        assert vals[2 + k] == getattr(AbdParamsC, f).offset, f
    for k, f in enumerate(fields_s):
        assert vals[2 + len(fields_p) + k] == getattr(_lib.AbdScene, f).offset, f


def test_oracle_params_mirror_matches_header():
    from oracle import oracle as O
    from abd_hip.const_params import AbdParamsC
    assert [f for f, _ in O.OrcParams._fields_] == [f for f, _ in AbdParamsC._fields_]
    assert ctypes.sizeof(O.OrcParams) == ctypes.sizeof(AbdParamsC)


def test_product_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "abd-warp_amd", "abd_hip")
    for fn in os.listdir(pkg):
        if fn.endswith(".py"):
            text = open(os.path.join(pkg, fn)).read()
            assert not re.search(r"^\s*(from|import)\s+oracle", text, re.M), fn


def test_sim_state_and_tuning_layouts_match_header(tmp_path):
    from abd_hip import _lib
    structs = [("AbdSimState", _lib.AbdSimState), ("AbdTuning", _lib.AbdTuning)]
    src = tmp_path / "layout2.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void) {']
    for cname, py in structs:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        lines += [f'printf("%zu\\n", offsetof({cname}, {f}));' for f, _ in py._fields_]
    lines += ['return 0; }']
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout2"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    k = 0
    for cname, py in structs:
        assert vals[k] == ctypes.sizeof(py), cname
        k += 1
        for f, _ in py._fields_:
            assert vals[k] == getattr(py, f).offset, (cname, f)
            k += 1


def test_tuning_round_trip_and_bounds():
    from abd_hip import _lib
    L = _lib.load()
    t0 = _lib.AbdTuning()
    assert L.abd_get_tuning(ctypes.byref(t0)) == 0
    assert (t0.cull_chunk, t0.cull_grid, t0.contact_bound_mib, t0.pair_grid, t0.toi_grid) == (4096, 16384, 0, 0, 0)
    bad = _lib.AbdTuning(3000, 16384)   # not a power of two
    assert L.abd_set_tuning(ctypes.byref(bad)) != 0
    assert L.abd_set_tuning(ctypes.byref(_lib.AbdTuning(4096, 16384, -1))) != 0
    assert L.abd_set_tuning(ctypes.byref(_lib.AbdTuning(4096, 16384, 0, 32, 0))) != 0   # pair_grid below 64
    assert L.abd_set_tuning(ctypes.byref(_lib.AbdTuning(4096, 16384, 0, 0, 70000))) != 0  # toi_grid above 65536
    good = _lib.AbdTuning(8192, 4096)
    assert L.abd_set_tuning(ctypes.byref(good)) == 0
    t1 = _lib.AbdTuning()
    L.abd_get_tuning(ctypes.byref(t1))
    assert (t1.cull_chunk, t1.cull_grid) == (8192, 4096)
    assert L.abd_set_tuning(ctypes.byref(t0)) == 0


def test_contact_params_reject_unknown_psd():
    from abd_hip.const_params import ContactParams
    for psd in (0, 1, 2, 3):
        assert ContactParams(psd=psd).to_c().psd == psd
    for psd in (-1, 4):
        with pytest.raises(ValueError):
            ContactParams(psd=psd).to_c()
