"""The C-ABI library loads and exports every symbol include/musica_gpu.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "musica_gpu.h")).read()
    return sorted(set(re.findall(r"\b(mgpu_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_binding_list():
    from musica2_newgrid_amd import _lib
    assert sorted(_lib.EXPORTS) == declared_symbols()


def test_library_exports_all_symbols():
    from musica2_newgrid_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_struct_layout_matches_header(tmp_path):
    """Every field offset of the ctypes mirrors equals the C compiler's for
    include/musica_gpu.h."""
    import subprocess
    from musica2_newgrid_amd._lib import MgpuBlock, MgpuCase
    assert ctypes.sizeof(MgpuBlock) == 4 * (3 + 24)
    # 8 int32, one pointer, 8 crk, deltat and 9 gas/filter doubles, 2 int32, t_wall and p_ref,
    # is_wall2 and its reserved twin
    assert ctypes.sizeof(MgpuCase) == 32 + 8 + 8 * (8 + 1 + 9) + 8 + 8 + 8 + 8
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{ROOT}/include/musica_gpu.h"', "int main(void) {"]
    for st, cname in ((MgpuCase, "mgpu_case"), (MgpuBlock, "mgpu_block")):
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in st._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    for st, cname in ((MgpuCase, "mgpu_case"), (MgpuBlock, "mgpu_block")):
        assert int(got[cname]) == ctypes.sizeof(st)
        for f, _ in st._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(st, f).offset, (cname, f)
