"""Seeded packings and their grain -> order-parameter assignments
(hpsint_amd/packing.py) against the committed fixture (bit-exact), plus the
invariant of an assignment: no two interacting particles share an order
parameter."""
import json
import os

import numpy as np
import pytest

from hpsint_amd import packing

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "packings.json")))

GEN = {"two_particles_2d": packing.two_particles_2d,
       "cfg1_8_spheres": lambda: packing.packing_cfg1(seed=0),
       "cfg2_27_spheres": lambda: packing.packing_cfg2(seed=0),
       "cfg3_64_spheres": lambda: packing.packing_cfg3(seed=0)}
for _s in range(5):
    GEN["study_100_spheres_seed%d" % _s] = (lambda s: lambda: packing.packing_cfg4(seed=s))(_s)

BUFFER = {"two_particles_2d": 0.0, "cfg1_8_spheres": 0.75, "cfg2_27_spheres": 1.25,
          "cfg3_64_spheres": 2.075}
BUFFER.update({"study_100_spheres_seed%d" % s: 2.0 for s in range(5)})


@pytest.mark.parametrize("name", sorted(GEN))
def test_packing_matches_fixture_bit_exactly(name):
    p = GEN[name]()
    fx = FIX[name]
    assert [[float.fromhex(x) for x in c] for c in fx["centers"]] == \
        [[float(x) for x in c] for c in p.centers]
    assert [float.fromhex(r) for r in fx["radii"]] == [float(r) for r in p.radii]
    assert fx["colors"] == [int(c) for c in p.colors]
    assert fx["n_order_params"] == p.n_order_params
    assert [float.fromhex(x) for x in fx["lower"]] == [float(x) for x in p.lower]
    assert [float.fromhex(x) for x in fx["upper"]] == [float(x) for x in p.upper]


@pytest.mark.parametrize("name", sorted(GEN))
def test_interacting_particles_get_distinct_order_parameters(name):
    p = GEN[name]()
    adj = packing.interaction_graph(p, radius_buffer_ratio=BUFFER[name])
    for i, nbrs in enumerate(adj):
        for j in nbrs:
            assert p.colors[i] != p.colors[j], (i, j)


def test_config_order_parameter_counts():
    assert packing.packing_cfg1(seed=0).n_order_params == 4
    assert packing.packing_cfg2(seed=0).n_order_params == 8
    assert packing.packing_cfg3(seed=0).n_order_params == 10
    assert packing.two_particles_2d().n_order_params == 2
    # the predictor study builds one operator with 12 order parameters
    assert max(packing.packing_cfg4(seed=s).n_order_params for s in range(5)) <= 12
