"""Host side of the mesh boundary (psum_amd.mesh): generators and the STL
reader, checked without a GPU.

  * axis_aligned_cube: 12 triangles covering the six faces (each face's two
    triangles have equal area summing to the face area; total area of the
    box surface), vertices on the box;
  * axis_aligned_plane: 2 triangles in the fixed plane, none when no
    coordinate is fixed;
  * revolve_zr_curve: 2 * segments * (points - 1) triangles; every vertex at
    the radius of its curve point; the lateral area converges to the
    frustum formula;
  * load_stl: the binary reference fixture has the triangle count its
    header declares; an ASCII file written here reads back exactly.
"""
import os

import numpy as np

from psum_amd import mesh as Mh

HERE = os.path.dirname(os.path.abspath(__file__))


def _areas(tri):
    t = np.asarray(tri).reshape(-1, 3, 3)
    return 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)


def test_cube_and_plane():
    c = Mh.axis_aligned_cube(-1, -2, -3, 1, 2, 3)
    assert c.shape == (12, 9)
    assert abs(_areas(c).sum() - 2 * (4 * 6 + 2 * 6 + 2 * 4)) < 1e-12
    v = c.reshape(-1, 3)
    assert np.all(np.isin(v[:, 0], [-1, 1])) and np.all(np.isin(v[:, 1], [-2, 2])) and np.all(np.isin(v[:, 2], [-3, 3]))
    p = Mh.axis_aligned_plane(-5, -5, -4.0, 5, 5, -4.0)
    assert p.shape == (2, 9) and np.all(p.reshape(-1, 3)[:, 2] == -4.0) and abs(_areas(p).sum() - 100) < 1e-12
    assert Mh.axis_aligned_plane(0, 0, 0, 1, 1, 1).shape == (0, 9)


def test_revolve():
    pts = [(0.0, 1.0), (1.0, 2.0), (3.0, 2.0)]
    for seg in (16, 256):
        t = Mh.revolve_zr_curve(pts, seg)
        assert t.shape == (2 * seg * 2, 9)
        v = t.reshape(-1, 3)
        r = np.hypot(v[:, 0], v[:, 1])
        assert np.all(np.isclose(r, 1.0) | np.isclose(r, 2.0))
    exact = np.pi * (1 + 2) * np.hypot(1, 1) + 2 * np.pi * 2 * 2
    assert abs(_areas(t).sum() / exact - 1) < 1e-3


def test_stl_binary_fixture_and_ascii_round_trip(tmp_path):
    tri = Mh.load_stl(os.path.join(HERE, "golden", "race.stl"))
    assert tri.shape == (1952, 9) and np.all(np.isfinite(tri))
    path = tmp_path / "two.stl"
    t = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 0, 1.5, 1, 0, 1.5, 0, 1, 1.5]], np.float64)
    with open(path, "w") as f:
        f.write("solid two\n")
        for row in t:
            f.write("  facet normal 0 0 1\n    outer loop\n")
            for k in range(3):
                f.write(f"      vertex {row[3 * k]} {row[3 * k + 1]} {row[3 * k + 2]}\n")
            f.write("    endloop\n  endfacet\n")
        f.write("endsolid two\n")
    assert np.array_equal(Mh.load_stl(str(path)), t)


def test_rescale_about_centroid():
    c = Mh.axis_aligned_cube(0, 0, 0, 2, 2, 2)
    s = Mh.rescale(c, 0.5)
    assert np.allclose(s.reshape(-1, 3).min(axis=0), 0.5) and np.allclose(s.reshape(-1, 3).max(axis=0), 1.5)
