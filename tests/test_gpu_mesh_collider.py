"""GPU mesh-collider voxelization (fx_voxelize_mesh) against the oracle's
restatement of the reference's ray-parity voxelizer
(collision/__init__.py:263-273, :304-367): identical occupancy on closed,
transformed, curved and open meshes, union with analytic shapes, and a
solver substep against a mesh collider."""

import numpy as np
import pytest

from helpers import box_mesh, rotation, uv_sphere

pytestmark = pytest.mark.gpu


def _cases():
    R = rotation((1, 2, 3), 0.7)
    bv, bt = box_mesh((-0.3, -0.2, -0.25), (0.3, 0.2, 0.25))
    M = np.eye(4)
    M[:3, :3] = R @ np.diag([1.5, 0.8, 1.2])  # rotated and non-uniformly scaled
    M[:3, 3] = (0.1, -0.05, 1.0)
    sv, st = uv_sphere((0.0, 0.0, 0.0), 0.55)
    S = np.eye(4)
    S[:3, 3] = (0.05, -0.1, 0.9)
    quad = np.array([[-0.6, -0.6, 1.0], [0.6, -0.6, 1.0], [0.6, 0.6, 1.0], [-0.6, 0.6, 1.0]])
    qt = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    two = (np.concatenate([quad, quad + [0, 0, 0.5]]), np.concatenate([qt, qt + 4]))
    return {"box": (*box_mesh((-0.33, -0.21, 0.37), (0.41, 0.52, 1.13)), None),
            "transformed_box": (bv, bt, M), "sphere": (sv, st, S), "open_quads": (*two, None)}


@pytest.mark.parametrize("name", ["box", "transformed_box", "sphere", "open_quads"])
@pytest.mark.parametrize("res", [24, 57])
def test_mesh_voxelization_matches_oracle(fx, oracle, name, res):
    from flowx_amd.collision import ColliderGrid, MeshCollider

    v, t, M = _cases()[name]
    m = MeshCollider(v, t, M)
    grid = ColliderGrid((-1.0, -1.0, 0.0), (1.0, 1.0, 2.0), res)
    grid.rebuild([("mesh", m)])
    got = grid.occupancy()
    ref = oracle.voxelize_mesh(grid.dims, grid.origin, grid.voxel, m.verts_world, m.tris_np, m.bounds_lo,
                               m.bounds_hi, m.ray_eps)
    print(f"{name} res {res}: {int(ref.sum())} occupied, {int((got != ref).sum())} differ")
    assert ref.sum() > 0 and np.array_equal(got, ref)


def test_mesh_union_with_analytic_shapes(fx, oracle):
    from flowx_amd.collision import ColliderGrid, MeshCollider, occupancy_cpu

    v, t, M = _cases()["sphere"]
    m = MeshCollider(v, t, M)
    grid = ColliderGrid((-1.0, -1.0, 0.0), (1.0, 1.0, 2.0), 40)
    shapes = [("box", (-0.9, -0.9, 0.0), (-0.5, 0.9, 0.3)), ("mesh", m)]
    grid.rebuild(shapes)
    ref = oracle.voxelize_mesh(grid.dims, grid.origin, grid.voxel, m.verts_world, m.tris_np, m.bounds_lo,
                               m.bounds_hi, m.ray_eps).astype(bool)
    ref = ref | occupancy_cpu(grid.origin, grid.voxel, grid.dims, shapes[:1]).reshape(ref.shape).astype(bool)
    assert np.array_equal(grid.occupancy().astype(bool), ref)


def test_substep_against_a_mesh_collider_matches_the_analytic_box(fx):
    """A closed box mesh and the analytic box voxelize identically, so a run
    against either collider is the same run."""
    from flowx_amd.collision import MeshCollider

    cfg_a = fx.dam_break_config()
    lo, hi = cfg_a.extra["colliders"][0][1], cfg_a.extra["colliders"][0][2]
    v, t = box_mesh(lo, hi)
    cfg_b = fx.dam_break_config()
    cfg_b.extra["colliders"] = [("mesh", MeshCollider(v, t))]
    runs = []
    for cfg in (cfg_a, cfg_b):
        s = fx.Solver(cfg, keep_scratch=False)
        s.reset()
        s.step(1 / 24.0)
        s.step(1 / 24.0)
        runs.append((s.collider.occupancy(), s.read_positions()))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
