"""Restart points and field output of the time loop.

save_restart / load_restart keep what the reference's driver writes at a
restart point (S/driver.h:2535-2610): the solution or its full history, the
step sizes of the time integrator and the time.  The file is a safetensors
archive (tensors and a string-to-string header, no executable content); the
mesh and the number of components are stored with it and checked on load,
in place of the reference's triangulation file.

write_vtk writes the nodal fields of a state as a legacy-VTK structured
grid (the Cartesian mesh): c, mu, eta_i and the two indicator fields of the
reference's output, bnds = sum_i eta_i^2 and gb = sum_{i>j} eta_i eta_j
(S/operator_sintering_base.h:171-183), evaluated at the nodes.
"""
import json

import numpy as np
import torch
from safetensors.numpy import load_file, save_file


def _mesh_meta(mesh):
    return {"n_cells": list(mesh.n_cells), "lower": [float(x) for x in mesh.lower],
            "upper": [float(x) for x in mesh.upper]}


def save_restart(path, stepper, full_history=True):
    """solution (and, with full_history, the committed old solutions the BDF
    scheme needs), the integrator's step sizes, the time and the step count"""
    op = stepper.op
    vecs = {"solution": stepper.u.detach().cpu().numpy()}
    hist = stepper.history if full_history else stepper.history[:1]
    for i, h in enumerate(hist):
        vecs["history_%d" % i] = h.detach().cpu().numpy()
    vecs["dt"] = np.asarray(stepper.td.dt, dtype=np.float64)
    meta = {"t": repr(float(stepper.t)), "dt_next": repr(float(stepper.dt)),
            "order": str(stepper.order), "n_history": str(len(hist)),
            "n_steps": str(len(stepper.records)), "n_components": str(op.n_components()),
            "fe_degree": str(op.fe_degree), "mesh": json.dumps(_mesh_meta(op.mesh))}
    save_file(vecs, path, metadata=meta)


def load_restart(path, stepper):
    """restores a state written by save_restart into a stepper whose operator
    has the same mesh, finite element and number of components"""
    from safetensors import safe_open
    with safe_open(path, framework="numpy") as f:
        meta = f.metadata()
    op = stepper.op
    if (json.loads(meta["mesh"]) != _mesh_meta(op.mesh) or
            int(meta["n_components"]) != op.n_components() or
            int(meta["fe_degree"]) != op.fe_degree):
        raise ValueError("restart file %s was written for another mesh or operator" % path)
    if int(meta["order"]) != stepper.order:
        raise ValueError("restart file %s holds a BDF%s state" % (path, meta["order"]))
    vecs = load_file(path)
    dev = op.device
    stepper.u = torch.from_numpy(vecs["solution"]).to(dev)
    stepper.history = [torch.from_numpy(vecs["history_%d" % i]).to(dev)
                       for i in range(int(meta["n_history"]))]
    dt_saved = [float(x) for x in vecs["dt"]]
    if len(stepper.history) >= stepper.order:
        stepper.td.dt = dt_saved
    else:
        # too few old solutions for the saved scheme: start again from the
        # lowest order with the current step (S/driver.h:370-377)
        stepper.td.dt = [dt_saved[0]] + [0.0] * (stepper.order - 1)
    stepper.td.update_weights()
    stepper.t = float(meta["t"])
    stepper.dt = float(meta["dt_next"])
    return int(meta["n_steps"])


def write_vtk(path, mesh, u, n_comp, fe_degree=1, title="hpsint_amd state"):
    """legacy-VTK STRUCTURED_POINTS file with the nodal fields of u
    (node-interleaved, n_comp components per node)"""
    if isinstance(u, torch.Tensor):
        u = u.detach().cpu().numpy()
    u = np.asarray(u, dtype=np.float64).reshape(-1, n_comp)
    dim = mesh.dim
    nn = [fe_degree * n + 1 for n in mesh.n_cells] + [1] * (3 - dim)
    if u.shape[0] != nn[0] * nn[1] * nn[2]:
        raise ValueError("vector does not match the mesh's nodes")
    h = [(mesh.upper[d] - mesh.lower[d]) / (fe_degree * mesh.n_cells[d]) for d in range(dim)]
    h += [1.0] * (3 - dim)
    origin = list(mesh.lower) + [0.0] * (3 - dim)
    etas = u[:, 2:]
    bnds = (etas * etas).sum(axis=1)
    gb = 0.5 * (etas.sum(axis=1) ** 2 - bnds)
    fields = [("c", u[:, 0]), ("mu", u[:, 1])]
    fields += [("eta%d" % i, etas[:, i]) for i in range(etas.shape[1])]
    fields += [("bnds", bnds), ("gb", gb)]
    with open(path, "wb") as f:
        head = ("# vtk DataFile Version 3.0\n%s\nBINARY\nDATASET STRUCTURED_POINTS\n"
                "DIMENSIONS %d %d %d\nORIGIN %.17g %.17g %.17g\nSPACING %.17g %.17g %.17g\n"
                "POINT_DATA %d\n" % (title, nn[0], nn[1], nn[2], *origin, *h, u.shape[0]))
        f.write(head.encode())
        for name, x in fields:
            f.write(("SCALARS %s double 1\nLOOKUP_TABLE default\n" % name).encode())
            f.write(np.ascontiguousarray(x, dtype=">f8").tobytes())
            f.write(b"\n")


def read_vtk(path):
    """reads a file written by write_vtk: (dimensions, origin, spacing,
    {field name: values})"""
    data = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        s = data[pos:end].decode()
        pos = end + 1
        return s

    assert line().startswith("# vtk DataFile")
    line()
    assert line() == "BINARY"
    assert line() == "DATASET STRUCTURED_POINTS"
    dims = [int(x) for x in line().split()[1:]]
    origin = [float(x) for x in line().split()[1:]]
    spacing = [float(x) for x in line().split()[1:]]
    n = int(line().split()[1])
    fields = {}
    while pos < len(data):
        head = line()
        if not head:
            continue
        name = head.split()[1]
        line()  # LOOKUP_TABLE
        fields[name] = np.frombuffer(data[pos:pos + 8 * n], dtype=">f8").astype(np.float64)
        pos += 8 * n + 1
    return dims, origin, spacing, fields
