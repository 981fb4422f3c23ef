"""The CPU oracle against the reference's own golden outputs.

Fixtures are transcribed by tests/golden/make_fixtures.py from
tests/residual_generic_{scalar,tensorial}_no_rbm.mpirun=1.output,
tests/free_energy_sintering.output and tests/power_sum.output.
"""
import json
import os

import numpy as np
import pytest

import pyoracle as po


def _sig(x, digits):
    return float("%.*e" % (digits - 1, x))


def two_cell_params(setup, tensorial):
    dt = setup["dt"]
    return po.make_params(
        2, 1, setup["n_cells"], setup["h"], 4, tensorial=tensorial,
        A=setup["A"], B=setup["B"], kappa_c=setup["kappa_c"],
        kappa_p=setup["kappa_p"], Mvol=setup["Mvol"], Mvap=setup["Mvap"],
        Msurf=setup["Msurf"], Mgb=setup["Mgb"], L=setup["L"],
        w0=1.0 / dt, w_hist=(-1.0 / dt,))


def two_cell_state(setup):
    nx = setup["n_cells"][0] + 1
    u = np.zeros((6, 4))
    for key, vals in setup["nodal_values"].items():
        x, y = map(int, key.split(","))
        u[x + nx * y] = vals
    # deal.II dof -> our lexicographic node
    perm = [p[0] + nx * p[1] for p in setup["dof_points"]]
    return u, perm


@pytest.mark.parametrize("kind", ["scalar", "tensorial"])
def test_residual_golden_2cell(golden_dir, kind):
    fx = json.load(open(os.path.join(golden_dir, "residual_generic_2cell.json")))
    setup = fx["setup"]
    p = two_cell_params(setup, kind == "tensorial")
    u, perm = two_cell_state(setup)
    r = po.residual(p, 2, u, [u]).reshape(6, 4)
    expected = fx[kind]
    assert len(expected) == 4
    for b in range(4):
        got = [_sig(r[perm[i], b], setup["digits"]) for i in range(6)]
        np.testing.assert_allclose(got, expected[b], rtol=0, atol=0,
                                   err_msg="block %d" % b)


def test_free_energy_known_answers(golden_dir):
    fx = json.load(open(os.path.join(golden_dir, "free_energy_sintering.json")))
    e = po.free_energy_eval(fx["A"], fx["B"], fx["c"], fx["etas"])
    exp = fx["expected"]
    got = {"f": e["f"], "df_dc": e["df_dc"], "d2f_dc2": e["d2f_dc2"]}
    etas = fx["etas"]
    for i in range(4):
        got["df_detai_%d" % i] = e["df_detai"][i]
        got["d2f_dcdetai_%d" % i] = e["d2f_dcdetai"][i]
        got["d2f_detai2_%d" % i] = e["d2f_detai2"][i]
        for j in range(i + 1, 4):
            got["d2f_detaidetaj_%d_%d" % (i, j)] = po.d2f_detaidetaj(
                fx["B"], etas[i], etas[j])
    assert set(got) == set(exp)
    for k, v in exp.items():
        assert float("%.6g" % got[k]) == v, k


def test_power_sum_known_answers(golden_dir):
    fx = json.load(open(os.path.join(golden_dir, "power_sum.json")))
    for sec in ("general", "small"):
        v = fx[sec]["v"]
        for p in fx[sec]["p"]:
            assert po.power_sum(v, p) == fx["expected"][sec][str(p)]
