"""The C-ABI library loads and exports exactly what include/*.h declare
(hcs.h: the engine; hcs_mujoco.h: the MuJoCo host glue), and the ctypes
binding covers all of it (no compute calls: CPU only)."""
import os
import re
import subprocess

from mujoco_contact_surfaces_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            # declarations (a return type before the name), not the callback members of a struct
            names |= set(re.findall(r"^(?:int|void\*|const char\*)\s+(hcs_[a-z0-9_]+)\s*\(", txt, flags=re.M))
    return names


def test_library_exports_every_declared_entry_point():
    lib = A.lib()
    declared = header_functions()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", A.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (hcs_[a-z0-9_]+)", out))
    assert declared == exported


def test_binding_signatures_cover_the_header():
    assert set(A.SIGNATURES) == header_functions()


def test_struct_layouts_match_the_header():
    import ctypes as C
    assert C.sizeof(A.PairIn) == 16 + 36 * 8
    assert C.sizeof(A.PairOut) == 12 * 8 + 4 * 8
    assert C.sizeof(A.Props) == 5 * 8
    assert C.sizeof(A.TaxelGrid) == 16 + 8 + 12 * 8 + 3 * 8 + 8


def test_generators_run_on_the_host():
    from mujoco_contact_surfaces_amd import scenes as S
    V, T, p = S.soft_mesh(2, 2, 0, 0.05, 0.05, 0.05, 1e5)
    assert T.shape == (320 * 4, 4) and V.shape == (1 + 162 * 2, 3)
    V, F = S.rigid_mesh(1, 64, 0.08)
    assert F.shape == (8192, 3)
    V, F = S.rigid_mesh(0, 1, 0.1, 0.1, 0.1)
    assert F.shape == (12, 3)


def test_invalid_arguments_are_rejected_without_a_device():
    lib = A.lib()
    import ctypes as C
    assert lib.hcs_gen_soft_size(-1, 2, None, None) == A.HCS_ERR_ARG
    assert lib.hcs_destroy(None) == A.HCS_ERR_ARG
    g = A.TaxelGrid()
    g.resolution = 0.0
    cx, cy = C.c_int32(), C.c_int32()
    assert lib.hcs_taxel_dims(C.byref(g), C.byref(cx), C.byref(cy)) == A.HCS_ERR_ARG
    g.resolution, g.half_size[0], g.half_size[1] = 0.005, 0.04, 0.04
    assert lib.hcs_taxel_dims(C.byref(g), C.byref(cx), C.byref(cy)) == 0
    assert (cx.value, cy.value) == (16, 16)
