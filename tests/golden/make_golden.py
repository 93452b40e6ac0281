"""Writes the golden fixtures of tests/golden/ from the reference's own CPU solver.

Run in a container where the reference headers are present (oracle/_ref built by
`make -C oracle ref`):  python tests/golden/make_golden.py
Every fixture holds seeded inputs and the reference's outputs for them; the tests
check the CPU restatement (and through it the HIP path) against these files
where the reference build is not available.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import subprocess  # noqa: E402

from mechanisms import TS1_JSON, all_rate_types, all_types_inputs, ts1_inputs, ts1_mechanism  # noqa: E402
from micm_amd import mechanism_io, synthetic  # noqa: E402
from oracle import Reference, reference_available  # noqa: E402

def export_ts1():
    """tests/golden/ts1_mechanism.json from the reference's TS1 benchmark header
    (oracle/_ref/ts1_export, built by `make -C oracle ts1`), with the compiled orders
    and patterns of this project's mechanism compiler attached."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ts1_export")
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ts1"], check=True)
    import json
    d = json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)
    desc = mechanism_io.from_dict(d)
    mechanism_io.save(desc, TS1_JSON, include_compiled=True, initial=d["initial"])


CASES = {
    # name: (mechanism, inputs(desc, n), tableau, dt, cells, solves)
    "readme": (synthetic.readme_mechanism, lambda d, n: synthetic.readme_inputs(n, seed=3), 3, 500.0, 24, 10),
    "all_types": (all_rate_types, all_types_inputs, 4, 60.0, 16, 3),
    "60_200": (synthetic.mechanism_60_200, synthetic.inputs_60_200, 4, 300.0, 12, 1),
    "210_560": (synthetic.mechanism_210_560, synthetic.inputs_210_560, 4, 900.0, 4, 1),
    # the reference's TS1 benchmark mechanism, its 3-stage solver and 30 s step
    "ts1": (ts1_mechanism, ts1_inputs, 3, 30.0, 6, 3),
}
ALPHA = 1.0 / (37.5 * 0.5728200000000000)


def make(name):
    mech_fn, inputs, tableau, dt, n, solves = CASES[name]
    desc = mech_fn()
    ref = Reference(desc, tableau)
    cells = inputs(desc, n)
    order = ref.species_order()
    y = cells.variables[:, order]
    pat = ref.patterns()
    k = ref.rate_constants(cells.conditions, cells.custom)
    f = ref.forcing(k, y)
    jac = ref.jacobian(k, y)
    a, lower, upper = ref.lu(jac, ALPHA)
    x = ref.lu_solve(lower, upper, f)
    y_cell, st_cell, stats_cell, ft_cell = ref.solve(cells.conditions, cells.custom, y, dt, solves)
    out = dict(species_order=order, jac_rows=pat["jacobian"][0], jac_cols=pat["jacobian"][1],
               lower_rows=pat["lower"][0], lower_cols=pat["lower"][1], upper_rows=pat["upper"][0],
               upper_cols=pat["upper"][1], conditions=cells.conditions, custom=cells.custom, y=y, k=k, forcing=f,
               jacobian=jac, alpha=np.array([ALPHA]), alpha_jacobian=a, lower=lower, upper=upper, solve=x,
               dt=np.array([dt]), n_solves=np.array([solves]), tableau=np.array([tableau]),
               y_per_cell=y_cell, state_per_cell=st_cell, stats_per_cell=stats_cell, final_time_per_cell=ft_cell)
    if name == "readme":
        # config-1 grouping: one State holding 1000 cells (lockstep) next to 1000 one-cell States
        big = synthetic.readme_inputs(1000, seed=1)
        yb = big.variables[:, order]
        out["lockstep_conditions"] = big.conditions
        out["lockstep_y0"] = yb
        out["lockstep_y"], _, out["lockstep_stats"], _ = ref.solve(big.conditions, big.custom, yb, 500.0, 10,
                                                                   lockstep=True)
        out["percell_y"], _, out["percell_stats"], _ = ref.solve(big.conditions, big.custom, yb, 500.0, 10)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(name, {kk: v.shape for kk, v in out.items() if v.size > 1})


def make_tableaus():
    """tests/golden/tableaus.npz: every field of the reference's five RosenbrockSolverParameters
    presets (ref_tableau in oracle/ref_driver.cpp)."""
    out = {}
    for preset in (2, 3, 4, 41, 61):
        for key, value in Reference.tableau(preset).items():
            out[f"{preset}_{key}"] = value
    np.savez_compressed(os.path.join(HERE, "tableaus.npz"), **out)
    print("tableaus", sorted({k.split("_", 1)[1] for k in out}))


if __name__ == "__main__":
    if not reference_available():
        sys.exit("oracle/_ref/libmicm_ref.so is not built (make -C oracle ref)")
    make_tableaus()
    names = sys.argv[1:] or list(CASES)
    if "ts1" in names:
        export_ts1()
    for name in names:
        make(name)
