"""Mesh-to-particle-lattice importer (mesh_lattice.h:1-281, SimulationHost::importMesh in
simulation_host.cpp:371-490): a closed triangle mesh becomes sphere particles on a cubic lattice,
joined by joints, exactly as the reference builds it.

- interior: per (y, z) lattice row, the surface's x-crossings are sorted and cells between
  crossing pairs are inside (row-parity fill, rows nudged by spacing * 0.0137 off mesh edges);
- surface: every cell a triangle touches (barycentric sampling at <= half-spacing steps, at most
  64 per edge) is added, so thin features still give a connected chain;
- particles: one sphere (radius spacing / 2) per occupied cell, created in z, y, x order;
- joints: to the 13 forward neighbours (faces, edges, corners), skipping a diagonal when an
  occupied cell on a partial step of it already connects the pair; anchors meet halfway.
  Stiffness: automatic (K = 7000 * particle mass, angular 0.2 K), given, or infinite (rigid).

All arithmetic is FP32 in the reference's operation order, so the created bodies and joints are
the reference's bit for bit.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

F = np.float32
FLT_MAX = np.finfo(np.float32).max


@dataclass
class TriMesh:
    verts: np.ndarray          # (n, 3) float32
    tris: np.ndarray           # (3 m,) uint32, three vertex indices per triangle


@dataclass
class LatticeParams:
    """MeshLatticeParams (mesh_lattice.h:33-40)."""
    spacing: float = 0.25          # lattice cell; particle radius = spacing / 2
    density: float = 12.0
    friction: float = 0.6
    stiffness_lin: float = -1.0    # < 0: automatic; inf: rigid lattice
    stiffness_ang: float = -1.0    # < 0: automatic (0.2 * linear)


def torus_mesh(major_radius, minor_radius, major_segments: int, minor_segments: int) -> TriMesh:
    """makeTorusMesh (mesh_lattice.h:42-72): ring around z, two triangles per quad."""
    tau = F(6.28318530718)
    R, r0 = F(major_radius), F(minor_radius)
    verts = []
    for i in range(major_segments):
        a = tau * F(i) / F(major_segments)
        ca, sa = F(math.cos(float(a))), F(math.sin(float(a)))
        for j in range(minor_segments):
            b = tau * F(j) / F(minor_segments)
            r = R + r0 * F(math.cos(float(b)))
            verts.append((ca * r, sa * r, r0 * F(math.sin(float(b)))))
    at = lambda i, j: (i % major_segments) * minor_segments + (j % minor_segments)
    tris = []
    for i in range(major_segments):
        for j in range(minor_segments):
            tris += [at(i, j), at(i + 1, j), at(i + 1, j + 1), at(i, j), at(i + 1, j + 1), at(i, j + 1)]
    return TriMesh(np.asarray(verts, F).reshape(-1, 3), np.asarray(tris, np.uint32))


def read_obj(text: str) -> TriMesh:
    """Wavefront OBJ (v / f lines; polygons fanned into triangles; v/vt/vn and negative indices)."""
    verts, tris = [], []
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "v" and len(parts) >= 4:
            verts.append([F(float(x)) for x in parts[1:4]])
        elif parts[0] == "f" and len(parts) >= 4:
            idx = []
            for p in parts[1:]:
                k = int(p.split("/")[0])
                idx.append(k - 1 if k > 0 else len(verts) + k)
            for t in range(1, len(idx) - 1):
                tris += [idx[0], idx[t], idx[t + 1]]
    tri = np.asarray(tris, np.int64)
    if len(tri) and (tri.min() < 0 or tri.max() >= len(verts)):
        raise ValueError("triangle index out of range")
    return TriMesh(np.asarray(verts, F).reshape(-1, 3), tri.astype(np.uint32))


@dataclass
class LatticeBuild:
    particles: list            # created body indices (builder order)
    centers: np.ndarray        # (k, 3) float32 build-time centres


def _grid(mesh: TriMesh, position, scale, spacing):
    v = mesh.verts.astype(F) * F(scale) + np.asarray(position, F)
    mn, mx = v.min(0), v.max(0)
    dims = [int((mx[k] - mn[k]) / spacing) + 1 for k in range(3)]
    return v, mn, dims


def _fill_rows(v, tris, mn, dims, spacing, inside):
    nx, ny, nz = dims
    p0, p1, p2 = v[tris[0::3]], v[tris[1::3]], v[tris[2::3]]
    e1y, e1z = p1[:, 1] - p0[:, 1], p1[:, 2] - p0[:, 2]
    e2y, e2z = p2[:, 1] - p0[:, 1], p2[:, 2] - p0[:, 2]
    d = e1y * e2z - e2y * e1z
    ok = np.abs(d) >= F(1e-12)
    e1x, e2x = p1[:, 0] - p0[:, 0], p2[:, 0] - p0[:, 0]
    row_eps = spacing * F(0.0137)
    count = 0
    with np.errstate(divide="ignore", invalid="ignore"):
        for iz in range(nz):
            cz = mn[2] + (F(iz) + F(0.5)) * spacing + row_eps
            for iy in range(ny):
                cy = mn[1] + (F(iy) + F(0.5)) * spacing + row_eps
                u = ((cy - p0[:, 1]) * e2z - e2y * (cz - p0[:, 2])) / d
                w = (e1y * (cz - p0[:, 2]) - (cy - p0[:, 1]) * e1z) / d
                hit = ok & ~(u < 0) & ~(w < 0) & ~(u + w > 1)
                xs = np.sort(p0[hit, 0] + u[hit] * e1x[hit] + w[hit] * e2x[hit])
                if len(xs) < 2:
                    continue
                for k in range(0, len(xs) - 1, 2):
                    x0 = int(np.ceil((xs[k] - mn[0]) / spacing - F(0.5)))
                    x1 = int(np.floor((xs[k + 1] - mn[0]) / spacing - F(0.5)))
                    for ix in range(max(x0, 0), min(x1, nx - 1) + 1):
                        inside[iz, iy, ix] = True
                        count += 1
    return count


def _mark_surface(v, tris, mn, dims, spacing, inside):
    nx, ny, nz = dims
    hi = (F(nx - 1), F(ny - 1), F(nz - 1))
    count = 0
    for t in range(0, len(tris) - 2, 3):
        p0, p1, p2 = v[tris[t]], v[tris[t + 1]], v[tris[t + 2]]
        e1, e2, e3 = p1 - p0, p2 - p0, p2 - p1
        dd = lambda e: e[0] * e[0] + e[1] * e[1] + e[2] * e[2]
        max_edge = np.sqrt(max(max(dd(e1), dd(e2)), dd(e3)))
        steps = min(max(int(np.ceil(max_edge / (spacing * F(0.5)))), 1), 64)
        for a in range(steps + 1):
            ta = F(a) / F(steps)
            for b in range(steps - a + 1):
                p = p0 + e1 * ta + e2 * (F(b) / F(steps))
                c = [int(max(F(0), min(hi[k], (p[k] - mn[k]) / spacing))) for k in range(3)]
                if not inside[c[2], c[1], c[0]]:
                    inside[c[2], c[1], c[0]] = True
                    count += 1
    return count


# the 13 forward neighbour directions: faces, edges, corners (mesh_lattice.h:237-240)
FORWARD = ((1, 0, 0), (0, 1, 0), (0, 0, 1),
           (1, 1, 0), (1, -1, 0), (1, 0, 1), (1, 0, -1), (0, 1, 1), (0, 1, -1),
           (1, 1, 1), (1, 1, -1), (1, -1, 1), (1, -1, -1))


def build_lattice(sb, mesh: TriMesh, position, scale=1.0, params: LatticeParams | None = None):
    """buildMeshLattice (mesh_lattice.h:85-281) into a scene builder; None when no cell is occupied."""
    params = params or LatticeParams()
    if len(mesh.verts) == 0 or len(mesh.tris) < 3:
        return None
    spacing = max(F(params.spacing), F(1e-3))
    tris = np.asarray(mesh.tris, np.int64)
    v, mn, dims = _grid(mesh, position, scale, spacing)
    nx, ny, nz = dims
    if nx * ny * nz > (1 << 22):
        return None
    inside = np.zeros((nz, ny, nx), bool)
    count = _fill_rows(v, tris, mn, dims, spacing, inside)
    count += _mark_surface(v, tris, mn, dims, spacing, inside)
    if count == 0:
        return None

    radius = spacing * F(0.5)
    rigid = math.isinf(params.stiffness_lin)
    klin, kang = F(params.stiffness_lin), F(params.stiffness_ang)
    if not rigid and klin < 0:
        mass = F(4.0) / F(3.0) * F(3.14159265) * radius * radius * radius * F(params.density)
        klin = F(7000.0) * mass
    if not rigid and kang < 0:
        kang = F(0.2) * klin

    ids = -np.ones((nz, ny, nx), np.int64)
    particles, centers = [], []
    for iz in range(nz):
        for iy in range(ny):
            for ix in range(nx):
                if not inside[iz, iy, ix]:
                    continue
                c = (mn[0] + (F(ix) + F(0.5)) * spacing, mn[1] + (F(iy) + F(0.5)) * spacing,
                     mn[2] + (F(iz) + F(0.5)) * spacing)
                ids[iz, iy, ix] = sb.sphere(radius, F(params.density), F(params.friction), c)
                particles.append(int(ids[iz, iy, ix]))
                centers.append(c)

    def occupied(x, y, z):
        return 0 <= x < nx and 0 <= y < ny and 0 <= z < nz and ids[z, y, x] >= 0

    half = spacing * F(0.5)
    k_lin = math.inf if rigid else klin
    k_ang = math.inf if rigid else kang
    for iz in range(nz):
        for iy in range(ny):
            for ix in range(nx):
                if ids[iz, iy, ix] < 0:
                    continue
                for d in FORWARD:
                    if not occupied(ix + d[0], iy + d[1], iz + d[2]):
                        continue
                    redundant = False
                    for bits in range(1, 7):
                        s = tuple(d[k] if bits & (1 << k) else 0 for k in range(3))
                        if s == (0, 0, 0) or s == d:
                            continue
                        if occupied(ix + s[0], iy + s[1], iz + s[2]):
                            redundant = True
                            break
                    if redundant:
                        continue
                    ra = (F(d[0]) * half, F(d[1]) * half, F(d[2]) * half)
                    sb.joint(int(ids[iz, iy, ix]), int(ids[iz + d[2], iy + d[1], ix + d[0]]), ra,
                             (-ra[0], -ra[1], -ra[2]), k_lin, k_ang)
    return LatticeBuild(particles, np.asarray(centers, F).reshape(-1, 3))


def skin_bindings(mesh: TriMesh, position, scale, build: LatticeBuild, spacing):
    """Per-vertex skinning (simulation_host.cpp:420-487): up to 4 nearest particles within
    2.5 spacings (else the single nearest), inverse-distance weights, offsets in build frame.
    Returns per vertex a list of (particle index into build.particles, weight, offset xyz)."""
    spacing = F(spacing)
    reach = spacing * F(2.5)
    reach_sq = reach * reach
    world = mesh.verts.astype(F) * F(scale) + np.asarray(position, F)
    out = []
    for wv in world:
        dvec = wv - build.centers
        dist = dvec[:, 0] * dvec[:, 0] + dvec[:, 1] * dvec[:, 1] + dvec[:, 2] * dvec[:, 2]
        best_i, best_d = [-1] * 4, [reach_sq] * 4
        for p, ds in enumerate(dist):
            for k in range(4):
                if ds < best_d[k]:
                    best_i[k + 1:] = best_i[k:3]
                    best_d[k + 1:] = best_d[k:3]
                    best_i[k], best_d[k] = p, ds
                    break
        if best_i[0] < 0 and len(dist):
            best_i[0] = int(np.argmin(dist))
            best_d[0] = dist[best_i[0]]
        wts = [F(1.0) / (np.sqrt(best_d[k]) + spacing * F(0.05)) if best_i[k] >= 0 else F(0) for k in range(4)]
        total = F(0)
        for k in range(4):
            if best_i[k] >= 0:
                total = total + wts[k]
        out.append([(best_i[k], float(wts[k] / total), *(wv - build.centers[best_i[k]]).tolist())
                    for k in range(4) if best_i[k] >= 0])
    return out
