"""Transcribe the reference's own golden test outputs into JSON fixtures.

Reads, as text, the committed expected-output files of hpsint's test suite
(tests/*.output) and writes the numbers they hold, together with the setup
that produced them (transcribed from the test sources named below), to
tests/golden/*.json.  Run it in a container where the reference is mounted:

    python tests/golden/make_fixtures.py /root/reference

The fixtures are data (inputs and expected outputs); nothing of the
reference's code is kept.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def blocks_from_output(path):
    """Vector blocks printed by residual_tester.h:66-77 (deal.II Vector::print)."""
    text = open(path).read()
    out = []
    for m in re.finditer(r"Vector data:\n([^\n]*)\n", text):
        out.append([float(x) for x in m.group(1).split()])
    return out


def key_values(path):
    vals = {}
    for line in open(path):
        m = re.match(r"\s*(\S+) = (\S+)\s*$", line)
        if m and m.group(1) not in vals:
            vals[m.group(1)] = float(m.group(2))
    return vals


def main(ref):
    T = os.path.join(ref, "tests")
    # tests/include/pf-applications/tests/sintering_model.h:53-99 (2x1 cells on
    # [0,2]x[0,1], Q1, QGauss(2), A=B=1, kappa_c=kappa_p=0.1,
    # ProviderAbstract(1e-1, 1e-8, 1e1, 1e0, 1e1), BDF1 dt=1e-5, old = current)
    # and S/initial_values_debug.h:36-63 (nodal values c, mu, eta0, eta1).
    setup = {
        "dim": 2, "degree": 1, "n_cells": [2, 1], "h": [1.0, 1.0],
        "A": 1.0, "B": 1.0, "kappa_c": 0.1, "kappa_p": 0.1,
        "Mvol": 0.1, "Mvap": 1e-8, "Msurf": 10.0, "Mgb": 1.0, "L": 10.0,
        "dt": 1e-5, "scheme": "BDF1", "history": "old solution = current",
        "nodal_values": {
            "0,0": [1.0, 0.1, 1.0, 0.0], "1,0": [0.8, 0.2, 0.4, 0.4],
            "2,0": [1.0, 0.3, 0.0, 1.0], "0,1": [1.0, 0.4, 1.0, 0.0],
            "1,1": [0.8, 0.5, 0.4, 0.4], "2,1": [1.0, 0.6, 0.0, 1.0],
        },
        # deal.II's DoF numbering of this 2-cell mesh: dof -> (x, y)
        "dof_points": [[0, 0], [1, 0], [0, 1], [1, 1], [2, 0], [2, 1]],
        "digits": 4,
    }
    fx = {"setup": setup}
    for kind in ("scalar", "tensorial"):
        name = "residual_generic_%s_no_rbm.mpirun=1.output" % kind
        fx[kind] = blocks_from_output(os.path.join(T, name))
    json.dump(fx, open(os.path.join(HERE, "residual_generic_2cell.json"), "w"),
              indent=1)

    # tests/free_energy_sintering.cc:29-36: A=2, B=3, c=0.95, etas
    fe = {"A": 2.0, "B": 3.0, "c": 0.95, "etas": [0.3, 0.15, 0.47, 0.0],
          "expected": key_values(os.path.join(T, "free_energy_sintering.output")),
          "digits": 6}
    json.dump(fe, open(os.path.join(HERE, "free_energy_sintering.json"), "w"),
              indent=1)

    # tests/power_sum.cc: {3,4,2,-5} for p=1..4 and {3,4} for p=2,3
    ps = {"general": {"v": [3.0, 4.0, 2.0, -5.0], "p": [1, 2, 3, 4]},
          "small": {"v": [3.0, 4.0], "p": [2, 3]}, "expected": {}}
    section = None
    for line in open(os.path.join(T, "power_sum.output")):
        if line.startswith("Generalized"):
            section = "general"
        elif line.startswith("Optimized"):
            section = "small"
        m = re.match(r"p=(\d): (\S+)", line)
        if m:
            ps["expected"].setdefault(section, {})[m.group(1)] = float(m.group(2))
    json.dump(ps, open(os.path.join(HERE, "power_sum.json"), "w"), indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
