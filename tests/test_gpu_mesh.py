"""Triangle-mesh particle boundary on the GPU (psum_mesh_* / psum_hip_mesh_*)
against the oracle.

Bar: the partition (cell lists, normals) is identical; detect is
bit-identical (k, normal, id) in both traversal modes, and with the
distance-field skip of the full traversal -- the geometry is
FP64 adds, multiplies and divides, uncontracted on both sides; the router
is bit-identical for specular and absorbing walls and for every seed, and
within 1e-12 (relative to the speed / to the position scale) where the
diffuse and Maxwell materials call sin, cos and log, which the device
library may round differently from the host's by an ulp.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))


def _scene():
    from psum_amd import mesh as Mh
    noz = Mh.revolve_zr_curve([(-4, 1.0), (-3, 0.9), (-2, 0.6), (-1, 0.6), (0, 1.0), (1, 1.4), (2, 1.8),
                               (3, 2.0), (4, 2.0)], 24)
    return [(Mh.axis_aligned_cube(-5, -5, -5, 5, 5, 10), Mh.ABSORB), (noz, Mh.SPECULAR),
            (Mh.axis_aligned_plane(-5, -5, -4.0, 5, 5, -4.0), Mh.DIFFUSE),
            (Mh.axis_aligned_plane(-5, 4.0, -5, 5, 4.0, 10), Mh.MAXWELL_REFLECT)]


def _flatten(meshes):
    tri = np.concatenate([np.asarray(t).reshape(-1, 9) for t, _ in meshes])
    ids = np.concatenate([np.full(len(np.asarray(t).reshape(-1, 9)), i, np.int32) for t, i in meshes])
    return tri, ids


def _cuda(arrs):
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


@pytest.mark.parametrize("shape", [(12, 10, 14), (7, 9, 5)])
def test_partition_matches_oracle(shape):
    from psum_amd import mesh as Mh
    meshes = _scene()
    tri, ids = _flatten(meshes)
    tm = Mh.TriangleMesh(meshes, *shape)
    om = O.OracleMesh(tri, ids, *shape)
    assert np.array_equal(tm.box, om.box)
    cs, ct, t12 = tm.partition()
    assert np.array_equal(cs, om.cell_start) and np.array_equal(ct, om.cell_tri)
    assert np.array_equal(t12, om.tri12)
    assert tm.kept == len(tri) and tm.max_id == 3
    tm.close()


@pytest.mark.parametrize("mode", ["restrict", "fast", "udf"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_detect_bit_exact(mode, dtype):
    from psum_amd import mesh as Mh
    meshes = _scene()
    tri, ids = _flatten(meshes)
    shape = (16, 16, 20)
    fast = mode == "fast"
    tm = Mh.TriangleMesh(meshes, *shape)
    tm.set_fast(fast)
    om = O.OracleMesh(tri, ids, *shape)
    if mode == "udf":
        tm.set_udf(True)
        om.set_udf(True)
    rng = np.random.default_rng(11)
    n = 20000
    p1 = rng.uniform([-4.9, -4.9, -4.9], [4.9, 4.9, 9.9], (n, 3)).astype(dtype)
    p2 = (p1 + rng.normal(0, 1.5, (n, 3))).astype(dtype)
    p2[::997] = p1[::997]                                  # zero-length segments
    p1[5::1013, 1] = np.nan                                # non-finite ends are not traced
    print(f"mesh detect mode={mode} dtype={np.dtype(dtype).name} n={n}", flush=True)
    k, nrm, idt = tm.detect(_cuda(p1.T), _cuda(p2.T))
    k, nrm, idt = k.cpu().numpy(), nrm.cpu().numpy(), idt.cpu().numpy()
    hits = 0
    for i in range(n):
        a, b = p1[i].astype(np.float64), p2[i].astype(np.float64)
        if not (np.all(np.isfinite(a)) and np.all(np.isfinite(b))):
            assert idt[i] == -1
            continue
        h, kk, nn, mid = om.detect(a, b, fast=fast)
        assert (idt[i] >= 0) == h, i
        if h:
            hits += 1
            assert k[i] == kk and idt[i] == mid and np.array_equal(nrm[i], nn), i
    assert hits > 2000
    tm.close()


def test_detect_stl_mesh():
    from psum_amd import mesh as Mh
    tri = Mh.load_stl(os.path.join(HERE, "golden", "race.stl"))
    ids = (np.arange(len(tri)) % 4).astype(np.int32)
    shape = (24, 40, 16)
    tm = Mh.TriangleMesh([(tri[ids == j], j) for j in range(4)], *shape)
    order = np.concatenate([np.nonzero(ids == j)[0] for j in range(4)])
    om = O.OracleMesh(tri[order], ids[order], *shape)
    rng = np.random.default_rng(12)
    n = 20000
    lo, hi = tri.reshape(-1, 3).min(axis=0), tri.reshape(-1, 3).max(axis=0)
    p1 = rng.uniform(lo - 0.1, hi + 0.1, (n, 3))
    p2 = p1 + rng.normal(0, 0.1, (n, 3))
    k, nrm, idt = tm.detect(_cuda(p1.T), _cuda(p2.T))
    k, nrm, idt = k.cpu().numpy(), nrm.cpu().numpy(), idt.cpu().numpy()
    hits = 0
    for i in range(n):
        h, kk, nn, mid = om.detect(p1[i], p2[i])
        assert (idt[i] >= 0) == h
        if h:
            hits += 1
            assert k[i] == kk and idt[i] == mid and np.array_equal(nrm[i], nn)
    assert hits > 500
    tm.close()


@pytest.mark.parametrize("udf", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_router_matches_oracle(dtype, udf):
    """flight through the scene: every step moves x += v dt and lets the
    router act on GPU and oracle copies; compared after each step"""
    from psum_amd import mesh as Mh
    meshes = _scene()
    tri, ids = _flatten(meshes)
    shape = (20, 20, 30)
    router = Mh.BoundaryRouter(temperature=300.0, mass=6.67e-27).set(meshes, *shape)
    om = O.OracleMesh(tri, ids, *shape)
    if udf:
        router.trigger.set_udf(True)
        om.set_udf(True)
    print(f"mesh router dtype={np.dtype(dtype).name} udf={udf}", flush=True)
    rng = np.random.default_rng(13)
    n = 6000
    pos = [rng.uniform(-0.6, 0.6, n), rng.uniform(-0.6, 0.6, n), rng.uniform(-3.9, 3.0, n)]
    vel = [rng.normal(0, 1.0, n) for _ in range(3)]
    pos = [a.astype(dtype) for a in pos]
    vel = [a.astype(dtype) for a in vel]
    seed = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    gp, gv = _cuda(pos), _cuda(vel)
    gs = torch.from_numpy(seed.view(np.int32).copy()).cuda()
    dt = dtype(0.1)
    total = 0
    kinds = np.zeros(4, np.int64)
    for step in range(40):
        old_h = [a.copy() for a in pos]
        pos = [(p + v * dt).astype(dtype) for p, v in zip(pos, vel)]
        ref = om.boundary(old_h, pos, vel, seed, temperature=300.0, mass=6.67e-27)
        gold = _cuda(old_h)
        for c in range(3):
            gp[c].copy_(torch.from_numpy(pos[c]))
        hits, hid = router.deal(gold, gp, gv, gs, hit_id=True)
        assert hits == ref["hits"], step
        assert np.array_equal(hid.cpu().numpy(), ref["hit_id"])
        assert np.array_equal(gs.cpu().numpy().view(np.uint32), ref["seed"])
        speed = np.sqrt(sum(x.astype(np.float64) ** 2 for x in ref["vel"]))
        for c in range(3):
            a, b = gp[c].cpu().numpy(), ref["pos"][c]
            assert np.array_equal(np.isnan(a), np.isnan(b))
            m = ~np.isnan(b)
            tol = 1e-12 * (1 + np.abs(b[m])) if dtype == np.float64 else np.spacing(np.abs(b[m]).astype(np.float32))
            assert np.all(np.abs(a[m] - b[m]) <= tol), (step, c)
            va, vb = gv[c].cpu().numpy(), ref["vel"][c]
            tolv = 1e-12 * speed if dtype == np.float64 else np.spacing(np.abs(vb).astype(np.float32))
            assert np.all(np.abs(va - vb) <= tolv), (step, c)
        # continue both from the oracle's state so that ulp differences of
        # the diffuse / Maxwell actions do not grow over the steps
        pos, vel, seed = ref["pos"], ref["vel"], ref["seed"]
        for c in range(3):
            gv[c].copy_(torch.from_numpy(vel[c]))
        gs.copy_(torch.from_numpy(seed.view(np.int32)))
        total += hits
        kinds += np.bincount(ref["hit_id"][ref["hit_id"] >= 0], minlength=4)
    assert total > 500 and np.all(kinds > 0), kinds
    router.trigger.close()


def test_device_view_describes_the_partition():
    """psum_mesh_get_view: dims, box, cells per length, the traversal mode
    and the device arrays are the partition's; the distance field appears
    in the view only while it is enabled"""
    import ctypes as C
    from psum_amd import mesh as Mh
    from psum_amd._lib import PsumMeshView, lib
    meshes = _scene()
    shape = (12, 10, 14)
    tm = Mh.TriangleMesh(meshes, *shape)
    v = PsumMeshView()
    assert lib().psum_mesh_get_view(tm.h, C.byref(v)) == 0
    assert (v.I, v.J, v.K) == shape and v.fast == 0 and not v.udf
    assert np.array_equal(np.array(v.box[:]), tm.box)
    assert np.allclose(np.array(v.rd[:]), np.array(shape) / (tm.box[3:] - tm.box[:3]), rtol=0, atol=0)
    assert v.cell_start and v.cell_tri and v.tri12 and v.ids
    tm.set_fast(True)
    assert lib().psum_mesh_get_view(tm.h, C.byref(v)) == 0 and v.fast == 1
    tm.set_udf(True)
    assert lib().psum_mesh_get_view(tm.h, C.byref(v)) == 0 and v.udf
    tm.set_udf(False)
    assert lib().psum_mesh_get_view(tm.h, C.byref(v)) == 0 and not v.udf
    tm.close()
