"""Writes packings.json: the seeded particle packings of the benchmark
workloads and their grain -> order-parameter assignments (DSatur greedy
colouring of the interaction graph), as produced by hpsint_amd.packing.

    python tests/golden/make_packings.py

The fixture pins the assignment bit-exactly (tests/test_packing.py): a change
of the generator or of the colouring shows up as a test failure instead of a
silent change of the benchmark workload."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from hpsint_amd import packing  # noqa: E402


def dump(p):
    return {"centers": [[float(x).hex() for x in c] for c in p.centers],
            "radii": [float(r).hex() for r in p.radii],
            "colors": [int(c) for c in p.colors],
            "n_order_params": int(p.n_order_params),
            "lower": [float(x).hex() for x in p.lower],
            "upper": [float(x).hex() for x in p.upper]}


def main():
    out = {"two_particles_2d": dump(packing.two_particles_2d()),
           "cfg1_8_spheres": dump(packing.packing_cfg1(seed=0)),
           "cfg2_27_spheres": dump(packing.packing_cfg2(seed=0)),
           "cfg3_64_spheres": dump(packing.packing_cfg3(seed=0))}
    for seed in range(5):
        out["study_100_spheres_seed%d" % seed] = dump(packing.packing_cfg4(seed=seed))
    with open(os.path.join(HERE, "packings.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
