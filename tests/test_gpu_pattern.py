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
