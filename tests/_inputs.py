"""Driver input sets written from the parameter fixture
tests/golden/reference_inputs.json (test helper).

``write_input_set(case, directory)`` writes the three files that
filament_amd.driver.prepare reads -- the input file with the reference's
variable names (src/Input.jl:20-200), an initial-condition file and a medium
file -- and returns the input file's path."""
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "reference_inputs.json")


def cases():
    return {k: v for k, v in json.load(open(FIXTURE)).items() if k != "about"}


def _assign(name, value):
    return f"{name} = {value!r}\n"


def write_input_set(case, directory):
    c = cases()[case]
    os.makedirs(directory, exist_ok=True)
    geometry = c["geometry"]
    m = c["model"]
    lines = [_assign("prefix", "results/"), _assign("file_initial_condition", "initial.py"),
             _assign("file_medium", "medium.py"), _assign("geometry", geometry)]
    for group in ("units", "grid", "guards"):
        lines += [_assign(k, v) for k, v in c[group].items()]
    lines += [_assign(k, m[k]) for k in ("zmax", "NONLINEARITY", "PLASMA", "Istop", "z") if k in m]
    lines += [_assign(k, m[k]) for k in ("KPARAXIAL", "QPARAXIAL") if k in m]
    lines.append(_assign("dz_initial", m["zmax"] / m["steps"]))
    lines.append(_assign("dzphimax", math.pi * m["dzphimax_over_pi"]))
    lines.append(f"ALG = {m.get('ALG', 'RK3')}()\n")
    with open(os.path.join(directory, "input.py"), "w") as f:
        f.writelines(lines)

    q = c["pulse"]
    if geometry == "RT":
        ic = f"gaussian_ic({q['a0']!r}, {q['tau0']!r}, {q['I0']!r}, {q['lam0']!r})"
    elif geometry == "T":
        ic = f"gaussian_ic_t({q['tau0']!r}, {q['I0']!r}, {q['lam0']!r})"
    elif geometry == "R":
        ic = f"gaussian_ic_r({q['a0']!r}, {q['I0']!r})"
    elif geometry == "XY":
        ic = f"gaussian_ic_xy({q['a0']!r}, {q['a0']!r}, {q['I0']!r})"
    else:
        ic = f"gaussian_ic_xyt({q['a0']!r}, {q['a0']!r}, {q['tau0']!r}, {q['I0']!r}, {q['lam0']!r})"
    with open(os.path.join(directory, "initial.py"), "w") as f:
        f.write("from filament_amd.workloads import (gaussian_ic, gaussian_ic_r, gaussian_ic_t, gaussian_ic_xy,\n"
                "                                    gaussian_ic_xyt)\n")
        f.write(_assign("lam0", q["lam0"]))
        f.write(f"initial_condition = {ic}\n")

    md = c["medium"]
    with open(os.path.join(directory, "medium.py"), "w") as f:
        f.write(f"from filament_amd.physics import {md['permittivity']} as permittivity\n")
        f.write("from filament_amd.physics import unit_permeability as permeability\n")
        f.write(_assign("n2", md["n2"]))
    return os.path.join(directory, "input.py")
