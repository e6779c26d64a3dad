"""Extract the reference's elasticity golden values into a JSON fixture.

The reference stores its only elasticity known answer as two pickles:
``tests/test_data/FEM/triangle/data/correct_data.dat`` (objective after the EMD
run) and ``correct_rho.dat`` (the design, 121 nodal values in dolfin CG1 dof
order), asserted by ``tests/test_elasticity_solver.py:54-55``.

Nothing here unpickles those files.  ``pickletools.genops`` only tokenises the
opcode stream (it imports and calls nothing), and the numbers are read straight
from the BINFLOAT / BININT1 / BINBYTES operands.  The rho vector is the single
968-byte little-endian float64 payload of the ndarray.

Usage (in a checkout that has the reference mounted)::

    python tests/golden/make_triangle_golden.py /root/reference triangle_golden.json
"""

import json
import os
import pickletools
import struct
import sys


def _ops(path):
    with open(path, "rb") as fh:
        data = fh.read()
    return [(op.name, arg) for op, arg, _ in pickletools.genops(data)]


def _scalar_after_key(ops, key):
    for i, (name, arg) in enumerate(ops):
        if name.endswith("UNICODE") and arg == key:
            for name2, arg2 in ops[i + 1:]:
                if name2 in ("BINFLOAT", "BININT1", "BININT", "BININT2"):
                    return arg2
    raise KeyError(key)


def extract(reference_root):
    folder = os.path.join(reference_root, "tests", "test_data", "FEM", "triangle", "data")
    data_ops = _ops(os.path.join(folder, "correct_data.dat"))
    rho_ops = _ops(os.path.join(folder, "correct_rho.dat"))

    objective = _scalar_after_key(data_ops, "objective")
    iteration = _scalar_after_key(data_ops, "iteration")
    penalty = _scalar_after_key(data_ops, "penalty")
    n = _scalar_after_key(rho_ops, "N")

    payloads = [arg for name, arg in rho_ops if name in ("BINBYTES", "SHORT_BINBYTES", "BINBYTES8")
                and isinstance(arg, (bytes, bytearray)) and len(arg) > 8]
    if len(payloads) != 1:
        raise RuntimeError(f"expected one ndarray payload, found {len(payloads)}")
    raw = payloads[0]
    count = len(raw) // 8
    values = list(struct.unpack("<%dd" % count, raw))
    return {
        "source": "tests/test_data/FEM/triangle/data/correct_{data,rho}.dat "
                  "(asserted at tests/test_elasticity_solver.py:54-55)",
        "design": "designs/triangle.json",
        "N": n,
        "objective": objective,
        "iteration": iteration,
        "penalty": penalty,
        "rho_dof_order": values,
        "rho_sorted": sorted(values),
    }


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = sys.argv[2] if len(sys.argv) > 2 else os.path.join(os.path.dirname(__file__),
                                                             "triangle_golden.json")
    golden = extract(ref)
    with open(out, "w") as fh:
        json.dump(golden, fh, indent=1)
        fh.write("\n")
    print(f"wrote {out}: objective={golden['objective']!r}, {len(golden['rho_dof_order'])} rho values")
