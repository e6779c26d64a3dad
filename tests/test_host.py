"""CPU tests of the host side: the C ABI library, the Python boundary and the input schema.
No GPU compute is issued here (the calls used return before touching the device)."""

import ctypes
import os
import re
import subprocess
import sys

import pytest

from tests.conftest import ROOT, design_path

HEADER = os.path.join(ROOT, "include", "tmx.h")
LIB = os.path.join(ROOT, "topomax_amd", "libtmx.so")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tmx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build libtmx.so first (python -m topomax_amd.build)"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (tmx_[a-z0-9_]+)$", out, flags=re.M))
    declared = header_functions()
    assert declared, "no declarations parsed"
    missing = [f for f in declared if f not in exported]
    assert not missing, f"declared but not exported: {missing}"


def test_ctypes_binding_covers_header():
    from topomax_amd import _lib
    assert sorted(_lib.SIGNATURES) == header_functions()
    lib = _lib.load()
    assert lib.tmx_abi_version() == 2


def test_desc_layout_matches_c_struct():
    # field offsets of tmx_problem_desc as the C compiler sees them
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "tmx.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(tmx_problem_desc), offsetof(tmx_problem_desc, fixed_sides),
         offsetof(tmx_problem_desc, tractions), offsetof(tmx_problem_desc, body_force),
         offsetof(tmx_problem_desc, precision), sizeof(tmx_solve_stats));
  return 0;
}
'''
    import tempfile
    from topomax_amd import _lib
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                               check=True).stdout.split()]
    D = _lib.ProblemDesc
    assert vals == [ctypes.sizeof(D), D.fixed_sides.offset, D.tractions.offset,
                    D.body_force.offset, D.precision.offset, ctypes.sizeof(_lib.SolveStats)]


def test_error_mapping_without_gpu():
    from topomax_amd import _lib
    lib = _lib.load()
    # invalid descriptor is rejected before any device work
    d = _lib.ProblemDesc()
    d.nx, d.ny = 0, 10
    h = ctypes.c_void_p()
    rc = lib.tmx_problem_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == _lib.TMX_ERR_INVALID and not h.value
    with pytest.raises(ValueError, match="nx and ny"):
        _lib.check(rc)
    d.nx, d.ny, d.width, d.height = 10, 10, 1.0, 1.0
    d.young_modulus, d.poisson_ratio, d.penalty_minimum = 1.0, 0.6, 1e-6
    with pytest.raises(ValueError, match="poisson"):
        _lib.check(lib.tmx_problem_create(ctypes.byref(d), ctypes.byref(h)))
    d.poisson_ratio = 0.3
    d.n_tractions = 1
    d.tractions[0].side = 7
    with pytest.raises(ValueError, match="Malformed side"):
        _lib.check(lib.tmx_problem_create(ctypes.byref(d), ctypes.byref(h)))
    with pytest.raises(ValueError):
        _lib.check(lib.tmx_set_penalization(None, 3.0))
    with pytest.raises(_lib.NotConvergedError):
        _lib.check(_lib.TMX_ERR_NOT_CONVERGED)
    with pytest.raises(RuntimeError):
        _lib.check(_lib.TMX_ERR_BREAKDOWN)
    # batch argument checks happen before any device work
    b = ctypes.c_void_p()
    with pytest.raises(ValueError, match="batch"):
        _lib.check(lib.tmx_batch_create(None, 2, ctypes.byref(b)))
    nulls = (ctypes.c_void_p * 2)(None, None)
    with pytest.raises(ValueError, match="null problem"):
        _lib.check(lib.tmx_batch_create(nulls, 2, ctypes.byref(b)))
    assert not b.value
    with pytest.raises(ValueError):
        _lib.check(lib.tmx_batch_set_workers(None, 4))


def test_parse_reference_designs():
    from topomax_amd import Side, parse_design
    dom, par = parse_design(design_path("triangle"))
    assert (dom.width, dom.height, dom.volume_fraction) == (1.0, 1.0, 0.5)
    assert par.fixed_sides == [Side.BOTTOM] and par.tractions is None
    assert par.body_force.region.center == (0.5, 0.5) and par.body_force.value == (0.0, -10.0)
    dom, par = parse_design(design_path("bridge"))
    assert par.fixed_sides == [Side.LEFT, Side.RIGHT] and par.body_force is None
    assert par.tractions[0].side == Side.TOP and par.tractions[0].value == (0.0, -2000.0)


def test_broken_design_raises():
    """tests/test_design_parser.py of the reference: a 3-tuple value raises ValueError."""
    from topomax_amd import parse_design
    with pytest.raises(ValueError):
        parse_design(design_path("broken_design"))


def test_penalizer_reference_semantics():
    from topomax_amd import ElasticPenalizer
    pen = ElasticPenalizer()
    with pytest.raises(ValueError, match="penalization"):
        pen(0.5)
    pen.set_penalization(3.0)
    assert abs(pen(0.5) - (1e-6 + 0.125 * (1 - 1e-6))) < 1e-16
    assert abs(pen.derivative(0.5) - 3 * 0.25 * (1 - 1e-6)) < 1e-16


def test_mesh_from_N_matches_reference_rule():
    from topomax_amd import StructuredMesh
    m = StructuredMesh.from_N(3.0, 1.0, 20)
    assert (m.nx, m.ny) == (60, 20)
    m = StructuredMesh.from_N(10.0, 5.0, 10)   # N per unit length = int(10 / 5) = 2
    assert (m.nx, m.ny) == (20, 10)


def test_bench_refuses_multi_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert p.returncode == 2 and "one GPU" in p.stderr


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "topomax_amd")
    for name in os.listdir(pkg):
        if name.endswith(".py"):
            src = open(os.path.join(pkg, name)).read()
            assert "oracle" not in src.replace("no CPU fallback", ""), name


def test_reference_record_pickle_layout(tmp_path):
    """The EMD output records are written as the reference's src.utils dataclasses without
    importing the reference, and the import table is restored afterwards."""
    import pickletools
    import sys
    from topomax_amd.emd import _dump_reference_record, _penalty_formatter
    before = {k: sys.modules.get(k) for k in ("src", "src.utils")}
    path = tmp_path / "r.dat"
    _dump_reference_record(str(path), "SolverResult",
                           {"exit_condition": "x", "min_objective": 1.0, "objectives": [2.0, 1.0],
                            "iterations": 2, "min_index": 1, "times": [0.1, 0.2]})
    assert {k: sys.modules.get(k) for k in ("src", "src.utils")} == before
    ops = [(op.name, arg) for op, arg, _ in pickletools.genops(path.read_bytes())]
    names = [a for _, a in ops if isinstance(a, str)]
    assert names[:2] == ["src.utils", "SolverResult"]
    assert any(n == "STACK_GLOBAL" for n, _ in ops) and any(n == "NEWOBJ" for n, _ in ops)
    fmt = _penalty_formatter([1.0, 2.5, 10.25])
    assert [fmt(p) for p in (1.0, 2.5, 10.25)] == ["01.00", "02.50", "10.25"]


def test_option_constants_match_header():
    """Every TMX_OPT_* of include/tmx.h has the same value in the ctypes binding (as the C
    compiler evaluates the enum), and a name in ElasticityProblem.set_option."""
    import inspect
    import tempfile
    from topomax_amd import _lib
    from topomax_amd.elasticity_problem import ElasticityProblem
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(TMX_OPT_[A-Z0-9_]+)\s*=", text)))
    assert len(names) >= 20
    body = "".join(f'  printf("%d\\n", (int){n});\n' for n in names)
    src = '#include <stdio.h>\n#include "tmx.h"\nint main(void) {\n' + body + "  return 0;\n}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "o.c")
        exe = os.path.join(d, "o")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                               check=True).stdout.split()]
    for n, v in zip(names, vals):
        assert getattr(_lib, n) == v, n
    src_py = inspect.getsource(ElasticityProblem.set_option)
    for n in names:
        assert f"_lib.{n}" in src_py, f"{n} has no set_option name"


def test_design_record_N_is_per_unit_length(tmp_path):
    """save_function stores mesh_to_N(mesh), the elements per unit length
    (FEM_src/utils.py:29-37, 47-69), and load_function rebuilds int(w N) x int(h N) cells:
    for the 12 x 2 bridge at solver N = 40 that is N = 20 and the solver's own 240 x 40 mesh."""
    import pickle
    import numpy as np
    from topomax_amd.consistency import design_mesh, load_design
    from topomax_amd.emd import design_record, mesh_parameter
    n_unit, full_N = mesh_parameter(40, 12.0, 2.0)
    assert (n_unit, full_N) == (20, 40)
    mesh = design_mesh(n_unit, (12.0, 2.0))
    assert (mesh.nx, mesh.ny) == (int(12.0 * n_unit), int(2.0 * n_unit)) == (240, 40)
    vec = np.linspace(0.0, 1.0, mesh.n_vertices)
    path = tmp_path / "bridge_rho.dat"
    with open(path, "wb") as fh:
        pickle.dump(design_record(vec, n_unit, 12.0, 2.0), fh)
    N, size, got = load_design(str(path))
    assert N == 20 and size == (12.0, 2.0) and np.array_equal(got, vec)


def test_design_record_dof_order_round_trip(tmp_path):
    """With a dof order the vector is written as the reference writes it (no marker) and
    load_design maps it back to vertex order; without the order such a file is refused."""
    import pickle
    import numpy as np
    from topomax_amd.consistency import load_design
    from topomax_amd.emd import design_record
    rng = np.random.default_rng(3)
    vec = rng.uniform(size=121)
    perm = rng.permutation(121)
    rec = design_record(vec, 10, 1.0, 1.0, dof_order=perm)
    assert "order" not in rec and np.array_equal(rec["vector"], vec[perm])
    path = tmp_path / "d_rho.dat"
    with open(path, "wb") as fh:
        pickle.dump(rec, fh)
    with pytest.raises(ValueError, match="dof order"):
        load_design(str(path))
    _, _, back = load_design(str(path), dof_order=perm)
    assert np.array_equal(back, vec)
    with pytest.raises(ValueError, match="permutation"):
        design_record(vec, 10, 1.0, 1.0, dof_order=np.zeros(121, dtype=int))


def test_load_design_refuses_foreign_globals(tmp_path):
    """A design file that names anything but numpy's array reconstructors is refused before
    the named object is imported or called."""
    import pickle
    from topomax_amd.consistency import load_design
    marker = tmp_path / "ran"

    class Payload:
        def __reduce__(self):
            return (open, (str(marker), "w"))

    path = tmp_path / "evil_rho.dat"
    with open(path, "wb") as fh:
        pickle.dump({"N": 10, "domain_size": (1.0, 1.0), "problem": "design",
                     "vector": Payload()}, fh)
    with pytest.raises(pickle.UnpicklingError, match="io.open|builtins.open"):
        load_design(str(path))
    assert not marker.exists()


def test_batch_report_layout_and_modes():
    """tmx_batch_report and the TMX_BATCH_* modes as the C compiler sees them equal the ctypes
    binding's."""
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "tmx.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %d %d\n", sizeof(tmx_batch_report),
         offsetof(tmx_batch_report, solo_designs), offsetof(tmx_batch_report, launches),
         offsetof(tmx_batch_report, iterations), offsetof(tmx_batch_report, record_ms),
         sizeof(((tmx_batch_report*)0)->iterations), TMX_BATCH_MERGED, TMX_BATCH_ALONE);
  return 0;
}
'''
    import tempfile
    from topomax_amd import _lib
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                               check=True).stdout.split()]
    R = _lib.BatchReport
    assert vals == [ctypes.sizeof(R), R.solo_designs.offset, R.launches.offset,
                    R.iterations.offset, R.record_ms.offset, ctypes.sizeof(ctypes.c_int * 3),
                    _lib.TMX_BATCH_MERGED, _lib.TMX_BATCH_ALONE]
