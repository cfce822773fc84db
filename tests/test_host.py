"""Host-side checks that run without a GPU: the tree/connectivity mirror, the
C-ABI library's exported symbols and its host-only eigensolver."""

import ctypes
import os
import re

import numpy as np
import pytest

from srmg import configs, solver
from srmg.tree import (boundaryconditions, connectivity, parentcoord, refined_tree,
                       star_stencil, subblocks, subcoord)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_mesh_index_helpers_match_reference_definitions():
    # src/mesh.jl:48-61: parent = div(c-1, 2)+1, sub = rem(c-1, 2)+1
    assert parentcoord((1, 2, 5)) == (1, 1, 3)
    assert subcoord((1, 2, 5)) == (1, 2, 1)
    for c in [(1, 1), (3, 4), (7, 2)]:
        for ch in subblocks(c):
            assert parentcoord(ch) == c
    # src/stencil.jl:16-33
    assert star_stencil(2) == [(-1, 0), (1, 0), (0, -1), (0, 1)]


def test_boundaryconditions_table_layout():
    # src/boundary.jl:28-38: Julia column-major 3x3, a[2+face...]
    a = boundaryconditions(((1, -1), (-1, -1))).reshape((3, 3), order="F")
    assert a[0, 1] == 1 and a[2, 1] == -1 and a[1, 0] == -1 and a[1, 2] == -1


@pytest.mark.parametrize("maker", [lambda: configs.make_cfg1(0), lambda: configs.make_cfg2(1)])
def test_trees_are_properly_nested(maker):
    p = maker()
    tree, conn = p.tree, p.conn
    D = tree.D
    for li, layer in enumerate(tree):
        nb = len(layer)
        # every face of every block is exactly one of neighbour / boundary / refinement
        faces = {}
        for rec in conn.neighbor[li]:
            faces[(int(rec[1]), tuple(-rec[2:2 + D]))] = "n"
        for rec in conn.boundary[li]:
            faces[(int(rec[0]), tuple(rec[1:1 + D]))] = "b"
        for rec in conn.refboundary[li]:
            faces[(int(rec[0]), tuple(rec[2:2 + D]))] = "r"
        assert len(faces) == nb * 2 * D
        if li > 0:
            assert len(conn.child[li]) == nb
    # children come in complete sibling groups
    for li in range(1, len(tree)):
        parents = {}
        for rec in conn.child[li]:
            parents.setdefault(int(rec[1]), []).append(int(rec[0]))
        assert all(len(v) == 2 ** D for v in parents.values())


def test_config_sizes():
    p1 = configs.make_cfg1(0)
    assert 15_000 <= p1.n_stored <= 25_000
    p2 = configs.make_cfg2(0)
    assert 1.5e6 <= p2.n_stored <= 2.5e6 and len(p2.tree) == 6


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "srmg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(srmg_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(solver.LIB_PATH)
    names = _header_functions()
    assert "srmg_vcycle" in names and len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("n,radial,alow,ahigh", [(8, False, -1, -1), (16, False, 1, -1),
                                                  (64, True, 1, -1), (128, False, -1, 1),
                                                  (32, True, 1, 1)])
def test_axis_decomposition_reconstructs_the_1d_operator(n, radial, alow, ahigh):
    from assemble import axis_matrix
    lam, P, Q = solver.axis_decomp(n, radial, alow, ahigh)
    T = axis_matrix(n, radial, alow, ahigh).toarray()
    np.testing.assert_allclose(P @ np.diag(lam) @ Q, T, atol=1e-12)
    np.testing.assert_allclose(P @ Q, np.eye(n), atol=1e-12)
    np.testing.assert_allclose(np.sort(lam), np.sort(np.linalg.eigvals(T).real), atol=1e-11)


def test_missing_extension_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(solver, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(solver, "_lib", None)
    with pytest.raises(solver.SrmgError):
        solver.lib()


@pytest.mark.parametrize("cname,pyname", [("srmg_op", "_Op"), ("srmg_solve_opts", "_SolveOpts")])
def test_ctypes_records_match_the_c_header_layout(cname, pyname, tmp_path):
    """Size and field offsets of the ctypes mirrors equal what a C compiler
    lays out for include/srmg.h."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    py = getattr(solver, pyname)
    fields = [f[0] for f in py._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'    printf("{f} %zu\\n", offsetof({cname}, {f}));\n' for f in fields)
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "srmg.h"\nint main(void) {\n'
                   f'    printf("sizeof %zu\\n", sizeof({cname}));\n{body}    return 0;\n}}\n')
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                       text=True).stdout.splitlines())
    assert int(out["sizeof"]) == ctypes.sizeof(py)
    for f in fields:
        assert int(out[f]) == getattr(py, f).offset, f


@pytest.mark.parametrize("name,arg", [("srmg_set_group_order", 1), ("srmg_set_axis_gemm", 2),
                                      ("srmg_set_group_tiles", 1)])
def test_tuning_switches_refuse_a_null_context(name, arg):
    """The switches report a null context through the status code and
    srmg_last_error without touching a device."""
    lib = ctypes.CDLL(solver.LIB_PATH)
    fn = getattr(lib, name)
    fn.argtypes, fn.restype = [ctypes.c_void_p, ctypes.c_int], ctypes.c_int
    lib.srmg_last_error.restype = ctypes.c_char_p
    assert fn(None, arg) != 0
    assert lib.srmg_last_error() == b"null context"
