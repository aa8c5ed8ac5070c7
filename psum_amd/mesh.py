"""Triangle-mesh particle boundary on the GPU (psum_mesh_* / psum_hip_mesh_*).

Host side of the reference's particle_boundary module:

  * mesh generators returning (n, 9) float64 triangle arrays (p1, p2, p3)
    -- axis_aligned_cube, axis_aligned_plane, revolve_zr_curve,
    extrude_xy_line_to_width, rescale (mesh_generator.hpp:16-337) -- and an
    STL reader (stl_loader.hpp:14-270);
  * ``TriangleMesh``: the triangle_mesh_trigger built from meshes with ids
    (``set(meshes, I, J, K)``), whose ``detect`` traces segments;
  * ``BoundaryRouter``: boundary_router with basic_material_set -- ids are
    materials (DIFFUSE, SPECULAR, MAXWELL_REFLECT, ABSORB) and ``deal``
    applies them to particles whose segment old -> new position hits the
    mesh.
"""
import ctypes as C
import struct

import numpy as np
import torch

from ._lib import PSUM_F32, PSUM_F64, PsumError, check, lib

DIFFUSE, SPECULAR, MAXWELL_REFLECT, ABSORB = 0, 1, 2, 3
MATERIALS = {"diffuse": DIFFUSE, "specular": SPECULAR, "maxwell_reflect": MAXWELL_REFLECT, "absorb": ABSORB}


# ------------------------------------------------------------ generators

def axis_aligned_cube(x0, y0, z0, x1, y1, z1):
    """12 triangles, two per face, faces in the reference's order
    (mesh_generator.hpp:16-86)."""
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x0, y1, z0], [x1, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x0, y1, z1], [x1, y1, z1]], np.float64)
    faces = [(0, 1, 3, 2), (4, 5, 7, 6), (0, 1, 5, 4), (2, 3, 7, 6), (0, 2, 6, 4), (1, 3, 7, 5)]
    out = []
    for a, b, c, d in faces:
        out.append(np.concatenate([v[a], v[b], v[c]]))
        out.append(np.concatenate([v[a], v[c], v[d]]))
    return np.array(out)


def axis_aligned_plane(x0, y0, z0, x1, y1, z1):
    """Two triangles spanning the rectangle of one fixed coordinate
    (mesh_generator.hpp:134-215); empty when no coordinate is fixed."""
    if x0 == x1:
        v = [[x0, y0, z0], [x0, y1, z0], [x0, y1, z1], [x0, y0, z1]]
    elif y0 == y1:
        v = [[x0, y0, z0], [x1, y0, z0], [x1, y0, z1], [x0, y0, z1]]
    elif z0 == z1:
        v = [[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0]]
    else:
        return np.zeros((0, 9))
    v = np.array(v, np.float64)
    return np.array([np.concatenate([v[0], v[1], v[2]]), np.concatenate([v[0], v[2], v[3]])])


def revolve_zr_curve(zr_points, azimuthal_segments):
    """Surface of revolution of a (z, r) polyline about the z axis
    (mesh_generator.hpp:217-300)."""
    zr = [(float(z), float(r)) for z, r in zr_points]
    if len(zr) < 2:
        return np.zeros((0, 9))
    dth = 2.0 * np.pi / azimuthal_segments
    out = []
    for i in range(azimuthal_segments):
        c0, s0 = np.cos(i * dth), np.sin(i * dth)
        c1, s1 = np.cos((i + 1) * dth), np.sin((i + 1) * dth)
        for j in range(len(zr) - 1):
            z0, r0 = zr[j]
            z1, r1 = zr[j + 1]
            p00 = [r0 * c0, r0 * s0, z0]
            p01 = [r0 * c1, r0 * s1, z0]
            p10 = [r1 * c0, r1 * s0, z1]
            p11 = [r1 * c1, r1 * s1, z1]
            out.append(p00 + p10 + p11)
            out.append(p00 + p11 + p01)
    return np.array(out, np.float64)


def extrude_xy_line_to_width(xy_points, z_width):
    """A polyline in the xy plane extruded over z in [-w/2, w/2]
    (mesh_generator.hpp:88-132)."""
    out = []
    za, zb = -z_width / 2.0, z_width / 2.0
    for (x0, y0), (x1, y1) in zip(xy_points[:-1], xy_points[1:]):
        out.append([x0, y0, za, x1, y1, za, x1, y1, zb])
        out.append([x0, y0, za, x1, y1, zb, x0, y0, zb])
        out.append([x0, y0, za, x1, y1, za, x0, y0, zb])
        out.append([x1, y1, za, x1, y1, zb, x0, y0, zb])
    return np.array(out, np.float64).reshape(-1, 9)


def rescale(tri, scale):
    """Scale about the mean of the triangle centroids (rescale_triangles)."""
    tri = np.array(tri, np.float64).reshape(-1, 9)
    cen = tri.reshape(-1, 3, 3).mean(axis=1).mean(axis=0)
    pts = tri.reshape(-1, 3, 3)
    return ((pts - cen) * scale + cen).reshape(-1, 9)


def load_stl(path):
    """Triangles of an ASCII or binary STL file (stl_loader.hpp): ASCII when
    the file starts with "solid" and its first 1000 bytes mention a facet or
    vertex; the facet normals in the file are ignored."""
    with open(path, "rb") as f:
        data = f.read()
    head = data[:1000].decode("latin-1")
    if data[:5] == b"solid" and ("facet" in head or "vertex" in head):
        verts = []
        for line in data.decode("latin-1").splitlines():
            t = line.split()
            if len(t) == 4 and t[0] == "vertex":
                verts.append([float(t[1]), float(t[2]), float(t[3])])
        if len(verts) % 3:
            raise PsumError(f"load_stl: {path}: vertex count not a multiple of 3")
        return np.array(verts, np.float64).reshape(-1, 9)
    if len(data) < 84:
        raise PsumError(f"load_stl: {path}: too short for a binary STL")
    n = struct.unpack_from("<I", data, 80)[0]
    if len(data) < 84 + 50 * n:
        raise PsumError(f"load_stl: {path}: truncated ({n} triangles declared)")
    rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", 9), ("a", "<u2")]), count=n, offset=84)
    return rec["v"].astype(np.float64)


# ------------------------------------------------------------ trigger / router

def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptrs3(ts):
    arr = (C.c_void_p * 3)()
    for c, t in enumerate(ts):
        if not (t.is_cuda and t.is_contiguous()):
            raise PsumError("coordinate arrays must be contiguous CUDA tensors")
        arr[c] = t.data_ptr()
    return arr


class TriangleMesh:
    """triangle_mesh_trigger: ``meshes`` is a list of (triangles (n, 9), id)
    pairs (triangle_mesh_with_id); I, J, K the partition; ``box`` the
    partitioned box (x0, y0, z0, x1, y1, z1) or None for the reference's
    widened bounding box."""

    def __init__(self, meshes, I, J, K, box=None):
        tris, ids = [], []
        for tri, mid in meshes:
            tri = np.asarray(tri, np.float64).reshape(-1, 9)
            tris.append(tri)
            ids.append(np.full(len(tri), int(mid), np.int32))
        self.tri = np.ascontiguousarray(np.concatenate(tris))
        self.ids = np.ascontiguousarray(np.concatenate(ids))
        if len(self.tri) == 0:
            raise PsumError("TriangleMesh: no triangles")
        b = None
        if box is not None:
            b = np.ascontiguousarray(box, np.float64)
        self.h = lib().psum_mesh_create(self.tri.ctypes.data_as(C.c_void_p), self.ids.ctypes.data_as(C.c_void_p),
                                        len(self.tri), I, J, K,
                                        b.ctypes.data_as(C.c_void_p) if b is not None else None)
        if not self.h:
            raise PsumError("psum_mesh_create: " + lib().psum_hip_last_error().decode(errors="replace"))
        self.shape = (I, J, K)
        info = np.zeros(4, np.int64)
        self.box = np.zeros(6)
        check(lib().psum_mesh_info(self.h, info.ctypes.data_as(C.c_void_p), self.box.ctypes.data_as(C.c_void_p)),
              "psum_mesh_info")
        self.kept, self.entries, self.cells, self.max_id = (int(x) for x in info)

    def set_fast(self, fast=True):
        check(lib().psum_mesh_set_fast(self.h, int(bool(fast))), "psum_mesh_set_fast")

    def set_udf(self, enabled=True):
        """set_acceleration_UDF: the full traversal skips empty space by the
        cells' distance field (computed on the host at the first enable)"""
        check(lib().psum_mesh_set_udf(self.h, int(bool(enabled))), "psum_mesh_set_udf")

    def partition(self):
        """(cell_start, cell_tri, tri12) host copies"""
        cs = np.zeros(self.cells + 1, np.uint32)
        ct = np.zeros(max(self.entries, 1), np.uint32)
        t12 = np.zeros((len(self.tri), 12))
        check(lib().psum_mesh_partition(self.h, cs.ctypes.data_as(C.c_void_p), ct.ctypes.data_as(C.c_void_p),
                                        t12.ctypes.data_as(C.c_void_p)), "psum_mesh_partition")
        return cs, ct[:self.entries], t12

    def detect(self, p1, p2):
        """Trace segments p1 -> p2 (3 CUDA tensors each).  Returns (k,
        normal (n, 3), id) tensors; id -1 where nothing is hit."""
        n = p1[0].numel()
        dt = PSUM_F64 if p1[0].dtype == torch.float64 else PSUM_F32
        dev = p1[0].device
        k = torch.zeros(n, dtype=torch.float64, device=dev)
        nrm = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        idt = torch.empty(n, dtype=torch.int32, device=dev)
        check(lib().psum_hip_mesh_detect(self.h, dt, _ptrs3(p1), _ptrs3(p2), n, C.c_void_p(k.data_ptr()),
                                         C.c_void_p(nrm.data_ptr()), C.c_void_p(idt.data_ptr()), _stream()),
              "psum_hip_mesh_detect")
        return k, nrm, idt

    def close(self):
        if getattr(self, "h", None):
            lib().psum_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class BoundaryRouter:
    """boundary_router<basic_material_set>: ``set(meshes, I, J, K)`` with
    material ids; ``deal(old, pos, vel, seed)`` for all particles."""

    def __init__(self, temperature=273.15, mass=9.11e-31):
        self.temperature = temperature     # basic_material_set defaults
        self.mass = mass
        self.trigger = None

    def set(self, meshes, I, J, K, box=None):
        meshes = [(t, MATERIALS[m] if isinstance(m, str) else int(m)) for t, m in meshes]
        self.trigger = TriangleMesh(meshes, I, J, K, box)
        if self.trigger.max_id > ABSORB:
            raise PsumError("BoundaryRouter: mesh ids must be materials 0..3")
        return self

    def deal(self, old, pos, vel, seed, hit_id=False):
        """Applies the materials to particles whose segment old -> pos hits
        the mesh (3 CUDA tensors each of one dtype; seed an int32 tensor of
        uint32 bit patterns).  Returns (hits, hit_id tensor or None)."""
        if self.trigger is None:
            raise PsumError("BoundaryRouter.deal before set")
        n = pos[0].numel()
        dt = PSUM_F64 if pos[0].dtype == torch.float64 else PSUM_F32
        dev = pos[0].device
        hits = torch.zeros(1, dtype=torch.int64, device=dev)
        hid = torch.empty(n, dtype=torch.int32, device=dev) if hit_id else None
        check(lib().psum_hip_mesh_boundary(self.trigger.h, dt, _ptrs3(old), _ptrs3(pos), _ptrs3(vel),
                                           C.c_void_p(seed.data_ptr()), n, self.temperature, self.mass,
                                           C.c_void_p(hid.data_ptr()) if hid is not None else None,
                                           C.c_void_p(hits.data_ptr()), _stream()),
              "psum_hip_mesh_boundary")
        return int(hits.item()), hid
