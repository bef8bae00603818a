"""Mesh ingestion of the Hydra scene bridge (shiro_hip_triangulate,
shiro_hip_smooth_normals), host code of the library.

The reference functions live in src/hydra/SceneBridge.cpp, which needs USD
and cannot be built here; the oracle restates them (oracle.cpp, cited
lines) and is pinned by closed forms: face counts, the fan structure, an
octahedron whose vertex normals are exactly the axes, a flat grid whose
normals are exactly +z, isolated and degenerate-only vertices getting
(0, 1, 0), and a geodesic sphere's normals approaching the radial
direction.  The library must equal the oracle bit for bit.
"""
import numpy as np
import pytest

from oracle import bindings as ob
from shiro_amd import hip
from shiro_amd.hip import HipError
from shiro_amd.workloads import geodesic_sphere


def _random_polygons(rng, faces=400, verts=300):
    counts = rng.integers(0, 9, faces).astype(np.int32)
    idx = rng.integers(0, verts, int(counts.sum())).astype(np.int32)
    return counts, idx


def test_triangulate_matches_oracle_and_fan_structure():
    rng = np.random.default_rng(1)
    counts, idx = _random_polygons(rng)
    got = hip.triangulate(counts, idx)
    ref = ob.triangulate(counts, idx)
    assert np.array_equal(got, ref)
    assert len(got) == 3 * int(np.maximum(counts - 2, 0).sum())
    # Each face's triangles are (root, v_k, v_k+1), k = 1..c-2.
    off, t = 0, 0
    for c in counts:
        face = idx[off:off + c]
        for k in range(1, c - 1):
            assert tuple(got[t:t + 3]) == (face[0], face[k], face[k + 1])
            t += 3
        off += c
    assert t == len(got)


def test_triangulate_rejects_malformed_faces():
    with pytest.raises(HipError):
        hip.triangulate([3, -1], [0, 1, 2])
    with pytest.raises(HipError):
        hip.triangulate([4], [0, 1, 2])
    assert ob.triangulate([4], [0, 1, 2]) is None
    assert len(hip.triangulate([], [])) == 0
    assert len(hip.triangulate([2, 1, 0], [0, 1, 2])) == 0


def test_smooth_normals_match_oracle_bit_for_bit():
    rng = np.random.default_rng(2)
    pos = rng.normal(size=(500, 3)).astype(np.float32)
    idx = rng.integers(0, 500, 3000).astype(np.uint32)
    idx[:6] = [0, 0, 1, 2, 3, 2]  # degenerate: repeated vertex
    idx[6:9] = [5, 600, 7]        # out-of-range index: skipped
    pos[10] = pos[11] = pos[12] = 0.5
    idx[9:12] = [10, 11, 12]      # zero-area triangle
    got = hip.smooth_normals(pos, idx)
    ref = ob.smooth_normals(pos, idx)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_octahedron_normals_are_the_axes():
    pos = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
    faces = []
    for x in (0, 1):
        for y in (2, 3):
            for z in (4, 5):
                tri = [x, y, z]
                # Outward winding: flip when the octant has an odd number of negative axes.
                if (x + y + z - 6) % 2 == 1:
                    tri = [x, z, y]
                faces += tri
    n = hip.smooth_normals(pos, np.array(faces, np.uint32))
    assert np.array_equal(n, pos)


def test_flat_grid_normals_are_up_and_isolated_vertices_default():
    xs, ys = np.meshgrid(np.linspace(-2, 3, 7), np.linspace(-1, 1, 5))
    pos = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)], 1).astype(np.float32)
    pos = np.vstack([pos, [[9, 9, 9]]]).astype(np.float32)  # isolated vertex
    idx = []
    for j in range(4):
        for i in range(6):
            a, b, c, d = j * 7 + i, j * 7 + i + 1, (j + 1) * 7 + i + 1, (j + 1) * 7 + i
            idx += [a, b, c, a, c, d]
    n = hip.smooth_normals(pos, np.array(idx, np.uint32))
    assert np.array_equal(n[:-1], np.tile(np.float32([0, 0, 1]), (len(pos) - 1, 1)))
    assert np.array_equal(n[-1], np.float32([0, 1, 0]))


def test_sphere_normals_approach_radial():
    p, i = geodesic_sphere(16)
    n = hip.smooth_normals(p, i)
    radial = p / np.linalg.norm(p, axis=1, keepdims=True)
    assert np.all(np.abs(np.linalg.norm(n, axis=1) - 1.0) < 1e-6)
    assert np.max(np.linalg.norm(n - radial, axis=1)) < 2e-2


def test_polygon_mesh_like_the_scene_bridge():
    # A unit quad and a pentagon: fan triangles, smooth normals when none
    # are given, the given ones normalised otherwise.
    pts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [2, 0, 0], [3, 0, 0], [3, 1, 0], [2.5, 2, 0],
                    [2, 1, 0]], np.float32)
    counts = [4, 5]
    idx = [0, 1, 2, 3, 4, 5, 6, 7, 8]
    m = hip.polygon_mesh(pts, counts, idx, material_index=2)
    assert m.materialIndex == 2
    assert list(m.indices) == [0, 1, 2, 0, 2, 3, 4, 5, 6, 4, 6, 7, 4, 7, 8]
    assert np.array_equal(m.normals, np.tile(np.float32([0, 0, 1]), (9, 1)))
    given = np.tile(np.float32([0, 0, 2]), (9, 1))
    m2 = hip.polygon_mesh(pts, counts, idx, normals=given)
    assert np.array_equal(m2.normals, np.tile(np.float32([0, 0, 1]), (9, 1)))
    assert hip.polygon_mesh(pts[:2], [2], [0, 1]) is None
