"""Pins of the oracle's triangle-mesh boundary (oracle/mesh.c).

The reference's own checks, restated:
  * a reported segment/triangle hit puts p1 + lambda (p2 - p1) inside the
    triangle by barycentric coordinates
    (test/particle_boundary/test_line_seg_tri_intersect.cpp), and a direct
    3x3 solve of p1 + t d = a + u e1 + v e2 decides the same hits;
  * every cell that a dense sampling of a segment touches is among the
    traversed cells, in order, and no traversed cell is missed by the
    segment (test_trace_to_idxs.cpp: "fail count should be ZERO");
  * box/triangle overlap is true whenever a sampled point of the triangle
    lies inside the box (test_box_tri_overlap.cpp), and agrees with a
    separating-axis test away from touching configurations;
plus: the partition lists exactly the cells of each triangle's range that
overlap it; detect returns the nearest hit among the triangles of the
first traversed cell with a hit, reports a hit exactly when some triangle
is hit (brute force), and that hit is the nearest in nearly every case;
specular reflection off a plane mirrors the normal component;
diffuse re-emission keeps |v| with a cosine law (E[cos] = 2/3); Maxwell
re-emission has E[v_n^2] = E[v_t^2] = 2kT/m; absorption invalidates; the
seed advances by the action's draws.  Planted mistakes (specular without
the velocity flip, traversal of the start cell only) fail these checks.
"""
import numpy as np
import pytest

import oracle_lib as O


def _rng(seed):
    return np.random.default_rng(seed)


def _solve_hit(tri, p1, p2):
    """independent decision: (hit, t, margin) of p1 + t d = a + u e1 + v e2"""
    a, b, c = tri[:3], tri[3:6], tri[6:]
    d = p2 - p1
    M = np.column_stack([d, -(b - a), -(c - a)])
    if abs(np.linalg.det(M)) < 1e-12:
        return None
    t, u, v = np.linalg.solve(M, a - p1)
    margin = min(t, 1 - t, u, v, 1 - u - v)
    return (t > 0 and t <= 1 and u >= 0 and v >= 0 and u + v <= 1), t, margin


def test_seg_tri_against_barycentric_and_solve():
    rng = _rng(1)
    hits = 0
    for _ in range(10000):
        tri = rng.uniform(-1, 1, 9)
        p1, p2 = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        h, lam = O.seg_tri(tri, p1, p2)
        ref = _solve_hit(tri, p1, p2)
        if h:
            hits += 1
            q = p1 + lam * (p2 - p1)
            v0, v1, v2 = tri[6:] - tri[:3], tri[3:6] - tri[:3], q - tri[:3]
            d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
            d20, d21 = v2 @ v0, v2 @ v1
            den = d00 * d11 - d01 * d01
            vv = (d11 * d20 - d01 * d21) / den
            ww = (d00 * d21 - d01 * d20) / den
            assert 0 < lam <= 1 and vv >= -1e-9 and ww >= -1e-9 and vv + ww <= 1 + 1e-9
        if ref is not None and abs(ref[2]) > 1e-9:
            assert h == ref[0]
            if h:
                assert abs(lam - ref[1]) < 1e-9
    assert hits > 500


def _cells_sampled(I, J, K, a, b, m=4000):
    t = np.linspace(0, 1, m + 1)[1:]
    pts = a[None, :] + t[:, None] * (b - a)[None, :]
    ins = np.all(pts >= 0, axis=1) & (pts[:, 0] < I) & (pts[:, 1] < J) & (pts[:, 2] < K)
    f = np.floor(pts[ins]).astype(np.int64)
    idx = f[:, 0] * J * K + f[:, 1] * K + f[:, 2]
    keep = np.ones(len(idx), bool)
    keep[1:] = idx[1:] != idx[:-1]
    return [int(x) for x in idx[keep]]


def _seg_hits_cell(I, J, K, a, b, idx, eps=1e-9):
    i, r = divmod(idx, J * K)
    j, k = divmod(r, K)
    lo, hi = np.array([i, j, k], float), np.array([i + 1, j + 1, k + 1], float)
    d = b - a
    t0, t1 = 0.0, 1.0
    for c in range(3):
        if abs(d[c]) < 1e-15:
            if a[c] < lo[c] - eps or a[c] > hi[c] + eps:
                return False
            continue
        ta, tb = (lo[c] - a[c]) / d[c], (hi[c] - a[c]) / d[c]
        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    return t0 <= t1 + eps


@pytest.mark.parametrize("plant", [0, 2])
def test_trace_covers_sampled_cells(plant):
    rng = _rng(2)
    I, J, K = 4, 3, 2
    bad = 0
    for _ in range(800):
        a, b = rng.uniform(-2, 6, 3), rng.uniform(-2, 6, 3)
        cells = O.trace_cells(I, J, K, a, b)
        if plant == 2:
            cells = cells[:1]
        samp = _cells_sampled(I, J, K, a, b)
        if not set(samp) <= set(cells):
            bad += 1
            continue
        # order along the segment and no spurious cell
        pos = [cells.index(c) for c in samp]
        assert pos == sorted(pos)
        for c in cells:
            assert 0 <= c < I * J * K and _seg_hits_cell(I, J, K, a, b, c)
    assert (bad == 0) == (plant == 0)


def _sat_overlap(tri, box):
    """separating-axis triangle/box test; returns (overlap, margin)"""
    v = tri.reshape(3, 3)
    c = (box[:3] + box[3:]) / 2
    h = (box[3:] - box[:3]) / 2
    v = v - c
    e = [v[1] - v[0], v[2] - v[1], v[0] - v[2]]
    axes = [np.eye(3)[i] for i in range(3)]
    axes.append(np.cross(e[0], e[1]))
    for i in range(3):
        for j in range(3):
            axes.append(np.cross(np.eye(3)[i], e[j]))
    margin = np.inf
    for ax in axes:
        nrm = np.linalg.norm(ax)
        if nrm < 1e-12:
            continue
        ax = ax / nrm
        p = v @ ax
        r = h @ np.abs(ax)
        sep = max(p.min() - r, -r - p.max())
        margin = min(margin, -sep)
        if sep > 0:
            return False, sep
    return True, margin


def test_box_tri_overlap_against_sampling_and_sat():
    rng = _rng(3)
    box = np.array([0, 0, 0, 1, 1, 1.0])
    agree = 0
    for _ in range(4000):
        tri = rng.uniform(-2, 3, 9)
        got = O.box_tri_overlap(tri, box)
        w = rng.dirichlet([1, 1, 1], 3000)
        pts = w @ tri.reshape(3, 3)
        if np.any(np.all((pts > 0) & (pts < 1), axis=1)):
            assert got
        ov, margin = _sat_overlap(tri, box)
        if abs(margin) > 1e-9:
            assert got == ov
            agree += 1
    assert agree > 3900


def _nozzle():
    from psum_amd import mesh as Mh
    return Mh.revolve_zr_curve([(-4, 1.0), (-3, 0.9), (-2, 0.6), (-1, 0.6), (0, 1.0), (1, 1.4), (2, 1.8),
                                (3, 2.0), (4, 2.0)], 24)


def _scene():
    from psum_amd import mesh as Mh
    return [(Mh.axis_aligned_cube(-5, -5, -5, 5, 5, 10), O.MAT["absorb"]),
            (_nozzle(), O.MAT["specular"]),
            (Mh.axis_aligned_plane(-5, -5, -4.0, 5, 5, -4.0), O.MAT["diffuse"]),
            (Mh.axis_aligned_plane(-5, 4.0, -5, 5, 4.0, 10), O.MAT["maxwell_reflect"])]


def _flatten(meshes):
    tri = np.concatenate([np.asarray(t).reshape(-1, 9) for t, _ in meshes])
    ids = np.concatenate([np.full(len(np.asarray(t).reshape(-1, 9)), i, np.int32) for t, i in meshes])
    return tri, ids


def test_partition_lists_exactly_the_overlapping_cells():
    tri, ids = _flatten(_scene())
    I, J, K = 12, 10, 14
    m = O.OracleMesh(tri, ids, I, J, K)
    span = m.box[3:] - m.box[:3]
    d = span / np.array([I, J, K])
    lists = {}
    for c in range(I * J * K):
        for t in m.cell_tri[m.cell_start[c]:m.cell_start[c + 1]]:
            lists.setdefault(int(t), set()).add(c)
    assert np.all(m.kept == 1)
    for t in range(0, len(tri), 13):
        v = tri[t].reshape(3, 3)
        for c in range(I * J * K):
            i, r = divmod(c, J * K)
            j, k = divmod(r, K)
            lo = m.box[:3] + np.array([i, j, k]) * d
            ov, margin = _sat_overlap(tri[t], np.concatenate([lo, lo + d]))
            if abs(margin) > 1e-9 * span.max():
                assert (c in lists.get(t, set())) == ov, (t, c)
        # normals: unit, along (p2-p1) x (p3-p1)
        n = np.cross(v[1] - v[0], v[2] - v[0])
        np.testing.assert_allclose(m.tri12[t, 9:], n / np.linalg.norm(n), rtol=0, atol=1e-15)


def _mt_all(tri, p1, p2):
    """segment parameter of every triangle hit (inf where missed), numpy"""
    a, b, c = tri[:, :3], tri[:, 3:6], tri[:, 6:]
    e1, e2, d = b - a, c - a, p2 - p1
    h = np.cross(d, e2)
    det = np.einsum("ij,ij->i", e1, h)
    ok = np.abs(det) >= 1e-20
    f = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
    o = p1 - a
    u = f * np.einsum("ij,ij->i", o, h)
    q = np.cross(o, e1)
    v = f * (q @ d)
    lam = f * np.einsum("ij,ij->i", e2, q)
    hit = ok & (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (lam > 0) & (lam <= 1)
    return np.where(hit, lam, np.inf)


def test_detect_first_cell_rule_and_completeness():
    """detect returns the nearest hit among the triangles listed in the first
    traversed cell that has one (so it may pass over a nearer hit in a later
    cell whose triangle is not listed earlier); it reports a hit exactly when
    some triangle is hit, and the nearest one in nearly every case"""
    tri, ids = _flatten(_scene())
    I, J, K = 16, 16, 20
    m = O.OracleMesh(tri, ids, I, J, K)
    rd = np.array([I, J, K]) / (m.box[3:] - m.box[:3])
    rng = _rng(4)
    found = nearest = 0
    for _ in range(1500):
        p1 = rng.uniform([-4.9, -4.9, -4.9], [4.9, 4.9, 9.9])
        p2 = p1 + rng.normal(0, 2.0, 3)
        h, k, nrm, mid = m.detect(p1, p2)
        lam = _mt_all(tri, p1, p2)
        assert h == bool(np.isfinite(lam).any())
        # the first-cell rule from the traversal and the partition lists
        want = None
        for c in O.trace_cells(I, J, K, (p1 - m.box[:3]) * rd, (p2 - m.box[:3]) * rd):
            lst = m.cell_tri[m.cell_start[c]:m.cell_start[c + 1]]
            if len(lst) and np.isfinite(lam[lst]).any():
                j = lst[np.argmin(lam[lst])]
                want = (lam[j], ids[j], m.tri12[j, 9:])
                break
        assert (want is not None) == h
        if h:
            found += 1
            assert abs(k - want[0]) < 1e-12 and mid == want[1] and np.array_equal(nrm, want[2])
            nearest += abs(k - lam.min()) < 1e-12
    assert found > 300 and nearest > 0.97 * found


def test_fast_mode_tests_end_cells_only():
    tri, ids = _flatten(_scene())
    m = O.OracleMesh(tri, ids, 16, 16, 20)
    rng = _rng(5)
    for _ in range(500):
        p1 = rng.uniform([-4, -4, -4], [4, 4, 9])
        p2 = p1 + rng.normal(0, 0.3, 3)
        hf = m.detect(p1, p2, fast=True)
        hr = m.detect(p1, p2)
        if hf[0]:
            assert hr[0]
        # a short segment stays within its end cells: both modes agree
        g1 = (p1 - m.box[:3]) * np.array([16, 16, 20]) / (m.box[3:] - m.box[:3])
        g2 = (p2 - m.box[:3]) * np.array([16, 16, 20]) / (m.box[3:] - m.box[:3])
        if np.sum(np.abs(np.floor(g1) - np.floor(g2))) <= 1:      # same or face-adjacent cells
            assert hf[0] == hr[0] and (not hf[0] or hf[1] == hr[1])


@pytest.mark.parametrize("plant", [0, 1])
def test_specular_mirrors(plant):
    n = np.array([0, 0, 1.0])
    hit = np.array([0.3, -0.2, 0.0])
    pos = np.array([0.5, 0.1, -0.4])
    vel = np.array([1.0, 2.0, -3.0])
    O.mesh_plant(plant)
    try:
        p, v, s = O.material_action(O.MAT["specular"], pos, vel, 77, hit, n)
    finally:
        O.mesh_plant(0)
    ok = np.allclose(v, [1.0, 2.0, 3.0], rtol=0, atol=0) and s == 77
    ok &= np.allclose(p, hit + np.array([0.2, 0.3, 0.4]) * 0.001, rtol=0, atol=1e-16)
    assert ok == (plant == 0)


def test_diffuse_and_maxwell_laws():
    rng = _rng(6)
    n = np.array([0.0, 0.6, 0.8])
    hit = np.zeros(3)
    cos, vn2, vt2 = [], [], []
    T, m = 300.0, 6.67e-27
    for s in rng.integers(0, 2**32, 4000, dtype=np.uint64):
        vel = rng.normal(0, 1, 3) * 10
        p, v, s2 = O.material_action(O.MAT["diffuse"], hit - 0.1 * n, vel, int(s), hit, n)
        assert abs(np.linalg.norm(v) / np.linalg.norm(vel) - 1) < 1e-14
        cos.append(v @ n / np.linalg.norm(v))
        assert s2 == O.lcg_draws(int(s), 3)[1]
        p, v, s3 = O.material_action(O.MAT["maxwell_reflect"], hit - 0.1 * n, vel, int(s), hit, n, T, m)
        assert s3 == O.lcg_draws(int(s), 4)[1]
        vn = v @ n
        vn2.append(vn ** 2)
        vt2.append(v @ v - vn ** 2)
    cos = np.array(cos)
    assert cos.min() >= 0 and abs(cos.mean() - 2 / 3) < 5 * np.sqrt(1 / 18 / len(cos))
    s2 = 2 * 1.3805e-23 * T / m
    assert abs(np.mean(vn2) / s2 - 1) < 0.08 and abs(np.mean(vt2) / s2 - 1) < 0.08
    p, v, s4 = O.material_action(O.MAT["absorb"], np.ones(3), np.ones(3), 5, hit, n)
    assert np.isnan(p[0]) and s4 == 5


def test_router_moves_particles_consistently():
    """free flight through the scene with the router: absorbed particles
    stay invalid, survivors stay inside the outer cube, speeds are kept
    by specular and diffuse walls"""
    tri, ids = _flatten(_scene()[:3])
    m = O.OracleMesh(tri, ids, 20, 20, 30)
    rng = _rng(7)
    N = 2000
    pos = [rng.uniform(-0.5, 0.5, N), rng.uniform(-0.5, 0.5, N), rng.uniform(-3.9, -3.8, N)]
    vel = [rng.normal(0, 1, N) for _ in range(3)]
    seed = rng.integers(0, 2**32, N, dtype=np.uint64).astype(np.uint32)
    speed0 = np.sqrt(sum(v ** 2 for v in vel))
    dt = 0.1
    total = 0
    for _ in range(150):
        old = [a.copy() for a in pos]
        pos = [p + v * dt for p, v in zip(pos, vel)]
        r = m.boundary(old, pos, vel, seed)
        pos, vel, seed = r["pos"], r["vel"], r["seed"]
        total += r["hits"]
    alive = np.isfinite(pos[0])
    assert total > 200 and 0 < alive.sum() < N
    for c, (lo, hi) in enumerate([(-5, 5), (-5, 5), (-5, 10)]):
        assert np.all(pos[c][alive] > lo) and np.all(pos[c][alive] < hi)
    np.testing.assert_allclose(np.sqrt(sum(v ** 2 for v in vel))[alive], speed0[alive], rtol=1e-12)


# ---------------------------------------------------------------- distance field

def _lower_bound(occ, i, j, k):
    """exact lower bound on the distance (cell units) from cell (i, j, k) to
    any occupied cell: per axis, cells d apart share no point closer than
    max(|d| - 1, 0) (test_udf.cpp: compute_precise_lower_bound)"""
    o = np.argwhere(occ)
    if len(o) == 0:
        return np.inf
    d = np.maximum(np.abs(o - np.array([i, j, k])) - 1, 0).astype(np.float64)
    return np.sqrt((d ** 2).sum(axis=1)).min()


def test_udf_reference_cases():
    """test_udf.cpp's basic cases: an empty grid gets min(I, J, K) - 1
    everywhere, a full grid zeros, an occupied cell zero"""
    assert np.all(O.compute_udf(np.zeros((10, 10, 10), np.uint8)) == 9.0)
    assert np.all(O.compute_udf(np.zeros((7, 12, 9), np.uint8)) == 6.0)
    assert np.all(O.compute_udf(np.ones((10, 10, 10), np.uint8)) == 0.0)
    occ = np.zeros((10, 10, 10), np.uint8)
    occ[5, 5, 5] = 1
    u = O.compute_udf(occ).reshape(occ.shape)
    assert u[5, 5, 5] == 0.0 and np.all(u[4:7, 4:7, 4:7] == 0.0) and u[0, 0, 0] > 0


def test_udf_line_closed_form():
    """one occupied end cell of a 1-cell-wide row: the field is the exact
    distance max(i - 1, 0) once the passes have raised the capped cells
    (the neighbourhood bound alone stops at udf_k - 1 = 5)"""
    for shape, axis in (((14, 1, 1), 0), ((1, 14, 1), 1), ((1, 1, 14), 2)):
        occ = np.zeros(shape, np.uint8)
        occ[(0, 0, 0)] = 1
        u = O.compute_udf(occ)
        assert np.array_equal(u, np.maximum(np.arange(14) - 1, 0).astype(np.float32)), axis
        # a single pass leaves the cap one higher only at the far cells
        u1 = O.compute_udf(occ, max_loop=0)
        assert np.array_equal(u1, np.minimum(np.maximum(np.arange(14) - 1, 0), [5] * 7 + [6] * 7).astype(np.float32))


def test_udf_random_conservative():
    """test_udf.cpp's random case: 100 cells of a 20^3 grid occupied by the
    reference generator from seed 12345, then 1000 sampled cells; the field
    never exceeds the exact lower bound there (and the neighbourhood bound
    equals it wherever it is below the cap)"""
    I = J = K = 20
    r, _ = O.lcg_draws(12345, 3 * 100 + 3 * 1000)
    occ = np.zeros((I, J, K), np.uint8)
    for n in range(100):
        occ[int(r[3 * n] * I), int(r[3 * n + 1] * J), int(r[3 * n + 2] * K)] = 1
    u = O.compute_udf(occ).reshape(occ.shape)
    u0 = O.compute_udf(occ, max_loop=0).reshape(occ.shape)
    checked = 0
    for n in range(1000):
        i, j, k = (int(r[300 + 3 * n + c] * d) for c, d in enumerate((I, J, K)))
        if occ[i, j, k]:
            continue
        lb = _lower_bound(occ, i, j, k)
        assert u[i, j, k] - lb <= 1e-6, (i, j, k, u[i, j, k], lb)
        assert u[i, j, k] >= min(lb, 5.0) - 1e-6
        if lb < 5.0 and u0[i, j, k] == u[i, j, k]:
            assert abs(u0[i, j, k] - np.float32(lb)) <= 1e-6
        checked += 1
    assert checked > 900


def test_udf_skip_keeps_the_hits():
    """detect with the field's skip against detect without it: on a scene
    whose occupied cells are far apart the skipped cells are empty, so the
    same hits are found (the skip can change a result only where the field
    overestimates, next to occupied cells)"""
    cube = np.array([[-5, -5, -5, 5, -5, -5, 5, 5, -5], [-5, -5, -5, 5, 5, -5, -5, 5, -5],
                     [-5, -5, 5, 5, -5, 5, 5, 5, 5], [-5, -5, 5, 5, 5, 5, -5, 5, 5]], np.float64)
    small = np.array([[-0.5, -0.5, 0.0, 0.5, -0.5, 0.0, 0.0, 0.5, 0.0]])
    tri = np.concatenate([cube, small])
    ids = np.array([0, 0, 0, 0, 1], np.int32)
    plain = O.OracleMesh(tri, ids, 30, 30, 30)
    skip = O.OracleMesh(tri, ids, 30, 30, 30)
    skip.set_udf(True)
    assert skip.udf.max() > 1.0
    rng = _rng(31)
    hits = skipped_same = 0
    for _ in range(3000):
        p1 = rng.uniform(-4.5, 4.5, 3)
        p2 = p1 + rng.normal(0, 3.0, 3)
        a = plain.detect(p1, p2)
        b = skip.detect(p1, p2)
        assert a[0] == b[0]
        if a[0]:
            hits += 1
            assert a[3] == b[3] and np.array_equal(a[2], b[2])
            skipped_same += a[1] == b[1]
    assert hits > 300 and skipped_same == hits
