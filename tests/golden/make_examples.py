"""Copies the reference's example inputs into test fixtures.

For each data/examples/<name>/ of the reference this writes
tests/golden/examples/<name>.csv (the initial_conditions.csv bytes as they
are) and one params.json with each example's simulation_parameters from its
config.json (g_constant, time_step, num_steps, softening_factor, gravity,
integrator).  Run once here, where the reference is mounted; the GPU tests
read only the fixtures.

    python tests/golden/make_examples.py [/root/reference]
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref="/root/reference"):
    src = os.path.join(ref, "data", "examples")
    out = os.path.join(HERE, "examples")
    os.makedirs(out, exist_ok=True)
    params = {}
    for name in sorted(os.listdir(src)):
        d = os.path.join(src, name)
        with open(os.path.join(d, "initial_conditions.csv"), "rb") as f:
            data = f.read()
        with open(os.path.join(out, name + ".csv"), "wb") as f:
            f.write(data)
        with open(os.path.join(d, "config.json")) as f:
            sp = json.load(f)["simulation_parameters"]
        params[name] = {k: sp[k] for k in ("g_constant", "time_step", "num_steps",
                                           "softening_factor", "gravity",
                                           "integrator")}
    with open(os.path.join(out, "params.json"), "w") as f:
        json.dump(params, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main(*sys.argv[1:])
