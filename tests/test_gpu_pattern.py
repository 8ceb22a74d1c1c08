"""Device pattern build (tm_pattern_dev_*, tm_element_slots_dev) against the
host pattern, itself pinned bit-exact to SciPy's tocsr (tests/test_host.py):
row pointers, columns and P_E identical on 3D and 2D meshes, with renumbered
nodes (so the atomics fill the incidence lists in a different order) and an
unused node (empty row)."""
import numpy as np
import pytest
import torch

from tensormesh_amd.assemble.topology import P1Topology, device_element_slots, device_pattern
from tensormesh_amd.mesh import kuhn_cube_arrays, structured_rectangle_arrays

pytestmark = pytest.mark.gpu


def _cases():
    pts, conn, _ = kuhn_cube_arrays(5)
    yield "kuhn5", pts, conn, "tetra"
    perm = np.random.default_rng(1).permutation(len(pts))
    yield "kuhn5-renumbered", pts[np.argsort(perm)], perm[conn], "tetra"
    p2, c2, _ = structured_rectangle_arrays(9, jitter=0.2, seed=3)
    yield "tri9", p2, c2, "triangle"
    # one extra node used by no element: an empty row
    yield "tri9-isolated", np.vstack([p2, [[2.0, 2.0]]]), c2, "triangle"


@pytest.mark.parametrize("name,pts,conn,et", list(_cases()), ids=[c[0] for c in _cases()])
def test_device_pattern_matches_host(cuda, name, pts, conn, et):
    print(f"case pattern {name}", flush=True)
    topo = P1Topology(pts, conn, len(pts), et)
    rp, ci, work = device_pattern(torch.from_numpy(conn).to(cuda), len(pts))
    assert np.array_equal(rp.cpu().numpy(), topo.row_ptr)
    assert np.array_equal(ci.cpu().numpy(), topo.col_idx)
    slots = device_element_slots(torch.from_numpy(conn).to(cuda), rp, ci, work, len(pts))
    assert np.array_equal(slots.cpu().numpy(), topo.element_slots().ravel())


def test_device_pattern_rejects_bad_ids(cuda):
    print("case pattern bad ids", flush=True)
    pts, conn, _ = kuhn_cube_arrays(2)
    bad = conn.copy()
    bad[3, 2] = len(pts) + 5
    with pytest.raises(Exception, match="outside"):
        device_pattern(torch.from_numpy(bad).to(cuda), len(pts))


def test_from_mesh_on_the_gpu_builds_no_host_pattern(cuda, monkeypatch):
    """ElementAssembler.from_mesh on a device mesh builds the pattern and P_E on
    the GPU (the host pattern entry points are not called) and assembles the
    same matrix as the host-built topology, bit for bit in the pattern."""
    import tensormesh_amd as tm
    from tensormesh_amd import _lib
    from tensormesh_amd.assemble.topology import clear_topology_cache
    print("case from_mesh device pattern", flush=True)
    pts, conn, bnd = kuhn_cube_arrays(7)
    kappa = torch.tensor(np.exp(np.random.default_rng(2).standard_normal(len(pts))), dtype=torch.float32)
    host_mesh = tm.Mesh(torch.from_numpy(pts), {"tetra": torch.from_numpy(conn)},
                        {"is_boundary": torch.from_numpy(bnd)})
    K_host = tm.KappaLaplaceElementAssembler.from_mesh(host_mesh).to(cuda)(point_data={"kappa": kappa.to(cuda)})
    pe_host = torch.from_numpy(P1Topology(pts, conn, len(pts), "tetra").element_slots())
    clear_topology_cache()
    L = _lib.lib()

    def refuse(*args):
        raise AssertionError("host pattern build called for a device mesh")

    monkeypatch.setattr(L, "tm_pattern_rowptr", refuse)
    monkeypatch.setattr(L, "tm_pattern_colidx", refuse)
    monkeypatch.setattr(L, "tm_element_slots", refuse)
    dev_mesh = host_mesh.to(cuda)
    asm = tm.KappaLaplaceElementAssembler.from_mesh(dev_mesh)
    assert asm.topology.pattern_device == cuda
    K = asm(point_data={"kappa": kappa.to(cuda)})
    assert torch.equal(K.row, K_host.row) and torch.equal(K.col, K_host.col)
    vref = K_host.values.double()
    assert ((K.values.double() - vref).abs().max() / vref.abs().max()).item() <= 1e-6
    # the generic path's P_E comes from the device too
    assert torch.equal(asm.projector.cpu(), pe_host)
