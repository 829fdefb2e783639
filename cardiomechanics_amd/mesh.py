"""Meshes for the mechanics path: generators, Tetgen I/O and LAT files.

Data layout (SI units, metres):
  coords      (N, 3) float64  node coordinates
  elements    (E, npe) int32  0-based node ids; T10 midside order
                              4=(0,1) 5=(1,2) 6=(0,2) 7=(0,3) 8=(1,3) 9=(2,3)
                              (CBElementSolidT10.cpp:151-165, ConvertT4toT10.cpp:262-290)
  bases       (E, nb, 3, 3)   fibre frame per point, row-major with columns f, s, n;
                              nb = 5 for T10 (centroid, then the 4 quadrature points,
                              CBModelLoaderTetgen.cpp:200-338), nb = 1 for T4
  bc          (N,) uint8      bitmask 1=x 2=y 4=z (CBModelLoaderTetgen.cpp:59-155)
  material    (E,) int32      material index
  activation  (E,) float64    activation time per element (s)
  surfaces    (S, nps) int32  pressure/cavity surface elements (T3 or T6),
                              T6 midside order 3=(0,1) 4=(1,2) 5=(2,0)
  surface_id  (S,) int32      surface index (cavity id)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Quadrature points of the T10 element (CBElementSolidT10.cpp:82-86)
_QA = (5 + 3 * math.sqrt(5)) / 20
_QB = (5 - math.sqrt(5)) / 20
T10_QP_BARY = np.array([[_QA, _QB, _QB, _QB], [_QB, _QA, _QB, _QB],
                        [_QB, _QB, _QA, _QB], [_QB, _QB, _QB, _QA]])
# (corner a, corner b) of the midside nodes 4..9
T10_EDGES = np.array([[0, 1], [1, 2], [0, 2], [0, 3], [1, 3], [2, 3]], dtype=np.int64)


@dataclass
class Mesh:
    coords: np.ndarray
    elements: np.ndarray
    bases: np.ndarray
    bc: np.ndarray
    material: np.ndarray
    activation: np.ndarray
    surfaces: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int32))
    surface_id: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    name: str = "mesh"

    @property
    def n_nodes(self) -> int:
        return int(self.coords.shape[0])

    @property
    def n_elements(self) -> int:
        return int(self.elements.shape[0])

    @property
    def npe(self) -> int:
        return int(self.elements.shape[1])

    @property
    def n_dofs(self) -> int:
        return 3 * self.n_nodes


# ---------------------------------------------------------------------------
# tetrahedra helpers
# ---------------------------------------------------------------------------

def signed_volume6(coords: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """6 x signed volume, the orientation test of CBElementSolidT10::CheckNodeSorting (:586-596)."""
    p = coords[tets[:, :4]]
    d1 = p[:, 1] - p[:, 0]
    d2 = p[:, 2] - p[:, 0]
    d3 = p[:, 3] - p[:, 0]
    return np.einsum("ij,ij->i", d1, np.cross(d2, d3))


def orient_t4(coords: np.ndarray, tets: np.ndarray) -> np.ndarray:
    tets = tets.copy()
    neg = signed_volume6(coords, tets) < 0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
    return tets


def check_node_sorting_t10(elements: np.ndarray, coords: np.ndarray, bases: np.ndarray | None = None):
    """CBElementSolidT10::CheckNodeSorting (:573-618): a negatively oriented element
    swaps corners 1<->2, midside 4<->6 and 8<->9, and quadrature bases 2<->3."""
    el = elements.copy()
    neg = signed_volume6(coords, el) < 0
    for a, b in ((1, 2), (6, 4), (8, 9)):
        el[neg, a], el[neg, b] = el[neg, b].copy(), el[neg, a].copy()
    if bases is not None:
        bases = bases.copy()
        bases[neg, 2], bases[neg, 3] = bases[neg, 3].copy(), bases[neg, 2].copy()
    return el, bases


def t4_to_t10(coords: np.ndarray, tets: np.ndarray):
    """Add straight-edge midside nodes (ConvertT4toT10.cpp:255-406 ordering)."""
    n = coords.shape[0]
    e = tets[:, T10_EDGES]                       # (E, 6, 2)
    lo = np.minimum(e[..., 0], e[..., 1]).astype(np.int64)
    hi = np.maximum(e[..., 0], e[..., 1]).astype(np.int64)
    key = (lo * n + hi).ravel()
    uniq, inv = np.unique(key, return_inverse=True)
    a = uniq // n
    b = uniq % n
    mid = 0.5 * (coords[a] + coords[b])
    new_coords = np.concatenate([coords, mid], axis=0)
    mids = (n + inv).reshape(-1, 6)
    el = np.concatenate([tets, mids], axis=1).astype(np.int32)
    return new_coords, el, (a, b)


def boundary_faces(tets: np.ndarray):
    """Faces used by exactly one tetrahedron, oriented with the outward normal
    of the solid (right-handed (v1-v0)x(v2-v0) points away from the 4th vertex)."""
    t = tets[:, :4].astype(np.int64)
    faces = np.concatenate([t[:, [1, 2, 3]], t[:, [0, 3, 2]], t[:, [0, 1, 3]], t[:, [0, 2, 1]]], axis=0)
    srt = np.sort(faces, axis=1)
    n = int(t.max()) + 1
    key = (srt[:, 0] * n + srt[:, 1]) * n + srt[:, 2]
    uniq, idx, cnt = np.unique(key, return_index=True, return_counts=True)
    return faces[idx[cnt == 1]]


def orient_faces_outward(coords, faces, tets):
    """Make (v1-v0)x(v2-v0) point away from the owning tetrahedron's opposite vertex."""
    t = tets[:, :4].astype(np.int64)
    n = int(coords.shape[0])
    srt_t = np.sort(np.concatenate([t[:, [1, 2, 3]], t[:, [0, 2, 3]], t[:, [0, 1, 3]], t[:, [0, 1, 2]]]), axis=1)
    opp = np.concatenate([t[:, 0], t[:, 1], t[:, 2], t[:, 3]])
    key_t = (srt_t[:, 0] * n + srt_t[:, 1]) * n + srt_t[:, 2]
    order = np.argsort(key_t)
    srt_f = np.sort(faces, axis=1)
    key_f = (srt_f[:, 0] * n + srt_f[:, 1]) * n + srt_f[:, 2]
    pos = np.searchsorted(key_t[order], key_f)
    o = opp[order[pos]]
    p = coords[faces]
    nrm = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    inward = np.einsum("ij,ij->i", nrm, coords[o] - p[:, 0]) > 0
    faces = faces.copy()
    faces[inward, 1], faces[inward, 2] = faces[inward, 2].copy(), faces[inward, 1].copy()
    return faces


def t3_to_t6(faces: np.ndarray, edge_ab, n_vertices: int) -> np.ndarray:
    """Midside nodes of surface triangles, order 3=(0,1) 4=(1,2) 5=(2,0)
    (the order CBElementSurfaceT6.cpp:104-107 assumes for its sub-triangles)."""
    a, b = edge_ab
    key_edges = a.astype(np.int64) * n_vertices + b.astype(np.int64)
    order = np.argsort(key_edges)
    out = [faces]
    for i, j in ((0, 1), (1, 2), (2, 0)):
        lo = np.minimum(faces[:, i], faces[:, j]).astype(np.int64)
        hi = np.maximum(faces[:, i], faces[:, j]).astype(np.int64)
        k = lo * n_vertices + hi
        pos = np.searchsorted(key_edges[order], k)
        out.append((n_vertices + order[pos])[:, None])
    return np.concatenate(out, axis=1).astype(np.int32)


def t3_subdivide(faces: np.ndarray, edge_ab, n_vertices: int) -> np.ndarray:
    """Split each straight quadratic face into 4 T3 triangles on its corner and
    midside nodes, keeping the parent's orientation -- how the reference's T10
    meshes carry their pressure surfaces (examples/benchmark2015/tetgen/ellipsoid.sur
    triangles use one corner and two midside nodes, or three midside nodes)."""
    t6 = t3_to_t6(faces, edge_ab, n_vertices)
    c0, c1, c2, m01, m12, m20 = (t6[:, k] for k in range(6))
    tris = np.stack([np.stack([c0, m01, m20], 1), np.stack([m01, c1, m12], 1),
                     np.stack([m20, m12, c2], 1), np.stack([m01, m12, m20], 1)], axis=1)
    return tris.reshape(-1, 3).astype(np.int32)


def quadrature_points(coords: np.ndarray, el: np.ndarray) -> np.ndarray:
    """Physical positions of (centroid, QP1..QP4) for straight-edged T10 -> (E, 5, 3)."""
    corners = coords[el[:, :4]]                              # (E, 4, 3)
    bary = np.concatenate([np.full((1, 4), 0.25), T10_QP_BARY], axis=0)  # (5, 4)
    return np.einsum("pk,ekc->epc", bary, corners)


def frames_from_vectors(f: np.ndarray, s: np.ndarray) -> np.ndarray:
    """Matrix3 with columns f, s, n = f x s (right handed, orthonormal)."""
    f = f / np.linalg.norm(f, axis=-1, keepdims=True)
    s = s - np.sum(s * f, axis=-1, keepdims=True) * f
    s = s / np.linalg.norm(s, axis=-1, keepdims=True)
    n = np.cross(f, s)
    return np.stack([f, s, n], axis=-1)


# ---------------------------------------------------------------------------
# structured hex grids split into tetrahedra
# ---------------------------------------------------------------------------
_KUHN = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def kuhn_tets(corner_index):
    """Six tetrahedra of each cell, all sharing the main diagonal.  corner_index(d)
    maps an offset tuple (d0, d1, d2) in {0,1}^3 to node ids of every cell."""
    tets = []
    for perm in _KUHN:
        d = [0, 0, 0]
        verts = [corner_index(tuple(d))]
        for ax in perm:
            d[ax] = 1
            verts.append(corner_index(tuple(d)))
        tets.append(np.stack(verts, axis=-1))
    return np.concatenate(tets, axis=0)


def cube_mesh(n: int = 6, length: float = 0.01, order: int = 10, surface_type: int = 3,
              fibre_deg=(-60.0, 60.0)) -> Mesh:
    """Unit-cube benchmark: n^3 hexes x 6 tets (1296 T10 for n=6), face x=0 clamped
    (bc=7), pressure surface x=L (outward normal +x), fibre angle linear in x.
    Pressure faces are T3 sub-triangles of the quadratic faces (surface_type=3)
    or T6 faces (surface_type=6)."""
    m = n + 1
    g = np.arange(m) * (length / n)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    idx = np.arange(m ** 3).reshape(m, m, m)
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    tets = kuhn_tets(lambda d: idx[ii + d[0], jj + d[1], kk + d[2]])
    tets = orient_t4(coords, tets)
    faces = orient_faces_outward(coords, boundary_faces(tets), tets)
    on_x1 = np.all(np.isclose(coords[faces][..., 0], length), axis=1)
    surf = faces[on_x1]
    nv = coords.shape[0]
    if order == 10:
        coords, el, edges = t4_to_t10(coords, tets)
        surf = t3_to_t6(surf, edges, nv) if surface_type == 6 else t3_subdivide(surf, edges, nv)
    else:
        el = tets.astype(np.int32)
    bc = np.where(np.isclose(coords[:, 0], 0.0), 7, 0).astype(np.uint8)

    def frame_at(p):
        a = np.deg2rad(fibre_deg[0] + (fibre_deg[1] - fibre_deg[0]) * p[..., 0] / length)
        f = np.stack([np.zeros_like(a), np.cos(a), np.sin(a)], axis=-1)
        s = np.broadcast_to(np.array([1.0, 0.0, 0.0]), f.shape)
        return frames_from_vectors(f, s)

    if order == 10:
        pts = quadrature_points(coords, el)
    else:
        pts = coords[el].mean(axis=1)[:, None, :]
    bases = frame_at(pts)
    E = el.shape[0]
    return Mesh(coords=coords, elements=el.astype(np.int32), bases=np.ascontiguousarray(bases),
                bc=bc, material=np.zeros(E, np.int32), activation=np.zeros(E),
                surfaces=surf.astype(np.int32), surface_id=np.ones(surf.shape[0], np.int32),
                name=f"cube{n}_T{order}")


# ---------------------------------------------------------------------------
# truncated prolate-ellipsoid left ventricle
# ---------------------------------------------------------------------------
LV_ENDO = (7e-3, 17e-3)      # short / long semi-axis of the endocardium
LV_EPI = (10e-3, 20e-3)      # short / long semi-axis of the epicardium
LV_BASE_Z = 5e-3             # truncation plane in the ellipsoid frame


def lv_ring_counts(ns: int, nv: int) -> np.ndarray:
    """Vertices of rings 0..ns of the LV surface grid: the apex (one vertex), then 6j
    on ring j (rings near the pole are nearly circles of radius proportional to j, so
    6j keeps their triangles close to equilateral) up to nv."""
    return np.array([1] + [min(nv, 6 * j) for j in range(1, ns + 1)], np.int64)


def lv_tet_count(nt: int, ns: int, nv: int) -> int:
    """Tetrahedra of lv_mesh(dims=(nt, ns, nv)): 3 per prism, one prism per surface
    triangle and layer; the strip between rings of n and m vertices holds n + m
    triangles, the apex fan n_1."""
    n = lv_ring_counts(ns, nv)
    tris = int(n[1] + np.sum(n[1:-1] + n[2:]))
    return 3 * nt * tris


def lv_dims_for(n_target: int):
    """(nt, ns, nv) giving about n_target tetrahedra: layer and ring counts from the
    wall's aspect (thickness 3 mm, meridian ~30 mm, circumference ~55 mm), then the
    largest ring size nv that comes closest to n_target."""
    c = (n_target / 6.0 / (1.0 * 10.0 * 18.3)) ** (1.0 / 3.0)
    nt = max(2, int(round(1.0 * c)))
    ns = max(4, int(round(10.0 * c)))
    nv0 = max(8, int(round(n_target / 6.0 / (nt * ns))))
    best = min(range(8, 4 * nv0 + 8), key=lambda nv: abs(lv_tet_count(nt, ns, nv) - n_target))
    return nt, ns, best


def ring_strip_triangles(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Triangulate the closed strip between two rings of vertex ids a and b, both
    ordered by angle and starting at angle 0 (vertex k of a ring of n at 2 pi k / n):
    advance along whichever ring's next vertex comes first.  A one-vertex ring a
    gives the fan around it."""
    na, nb = len(a), len(b)
    if na == 1:
        k = np.arange(nb)
        return np.stack([np.full(nb, a[0]), b[k], b[(k + 1) % nb]], axis=-1)
    tris = []
    i = j = 0
    while i < na or j < nb:
        if j >= nb or (i < na and (i + 1) * nb <= (j + 1) * na):
            tris.append((a[i % na], a[(i + 1) % na], b[j % nb]))
            i += 1
        else:
            tris.append((a[i % na], b[(j + 1) % nb], b[j % nb]))
            j += 1
    return np.array(tris, np.int64)


def split_prisms(prisms: np.ndarray) -> np.ndarray:
    """Split triangular prisms (bottom v0 v1 v2, top v3 v4 v5, v3 above v0, ...) into
    3 tetrahedra each so that neighbouring prisms conform: every quadrilateral face is
    cut by the diagonal through its smallest global vertex id (Dompierre, Labbe,
    Vallet, Camarero, "How to subdivide pyramids, prisms and hexahedra into
    tetrahedra", IMR 1999).  The prism is first turned (and flipped top to bottom) so
    that its smallest vertex is v0; the faces at v0 then take their diagonals from v0
    and the third face's diagonal decides between the two splits.  Orientation is left
    to orient_t4."""
    P = np.asarray(prisms, np.int64).copy()
    amin = np.argmin(P, axis=1)
    top = amin >= 3
    P[top] = P[top][:, [3, 4, 5, 0, 1, 2]]
    amin = np.where(top, amin - 3, amin)
    for r, perm in ((1, [1, 2, 0, 4, 5, 3]), (2, [2, 0, 1, 5, 3, 4])):
        P[amin == r] = P[amin == r][:, perm]
    v0, v1, v2, v3, v4, v5 = P.T
    cut15 = np.minimum(v1, v5) < np.minimum(v2, v4)
    a = np.stack([np.stack([v0, v1, v2, v5], -1), np.stack([v0, v1, v5, v4], -1),
                  np.stack([v0, v4, v5, v3], -1)], axis=1)
    b = np.stack([np.stack([v0, v1, v2, v4], -1), np.stack([v0, v4, v2, v5], -1),
                  np.stack([v0, v4, v5, v3], -1)], axis=1)
    return np.where(cut15[:, None, None], a, b).reshape(-1, 4)


def tet_quality(coords: np.ndarray, tets: np.ndarray) -> np.ndarray:
    """6 sqrt(2) |V| / l_rms^3 of each tetrahedron (1 for the regular one, 0 for a flat
    one), l_rms the root mean square of its six edge lengths."""
    p = coords[tets[:, :4]]
    vol = np.abs(signed_volume6(coords, tets)) / 6.0
    l2 = sum(np.sum((p[:, a] - p[:, b]) ** 2, axis=1) for a, b in
             ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))) / 6.0
    return 6.0 * np.sqrt(2.0) * vol / l2 ** 1.5


def lv_mesh(n_target: int = 200_000, dims=None, order: int = 10, surface_type: int = 3,
            fibre_deg=(60.0, -60.0), activation=(0.0, 0.08), endo_normal: str = "cavity") -> Mesh:
    """Idealised truncated prolate-ellipsoid LV (Land et al. 2015 benchmark shape).

    Wall surfaces of transmural parameter t: r_s(t) = 7+3t mm, r_l(t) = 17+3t mm,
    u(s) from -pi (apex) to -acos(5 mm / r_l(t)), so every layer ends on the plane
    z = 5 mm.  Coordinates are shifted by -5 mm so the base lies in z = 0; the cavity
    volume taken about the origin (CBCirculationCavity.cpp:77-94) is then the volume
    closed by the base plane.  dims = (nt, ns, nv): nt layers through the wall; on
    each surface, rings j = 0..ns at s = j/ns, ring j holding min(6j, nv) vertices
    (a polar cap with near-equilateral triangles, where a uniform ring size would make
    needle-shaped elements at the apex).  The surface triangles (ring strips) are
    extruded through the wall into prisms, each split into 3 tetrahedra by the
    smallest-vertex rule, which keeps the mesh conforming.  Base nodes are clamped
    (bc=7).  Surface 1 = endocardium, oriented like the reference's ellipsoid.sur:
    normal out of the wall, into the cavity.  The residual carries the pressure forces
    of CBElementSurfaceT3::CalcForcesDueToPressure (-p A N n), so a negative pressure
    amplitude inflates (as in the reference's benchmark Problem2.xml) and the cavity
    volume sum_i x_0.(a x b)/6 is negative.  Fibres: helix angle linear in t from
    fibre_deg[0] (endocardium) to fibre_deg[1] (epicardium), sheet = transmural.
    Activation rises linearly apex -> base over `activation` seconds.
    endo_normal="wall" turns the endocardium the other way (normal out of the blood
    volume, into the wall): positive pressure inflates and the cavity volume is
    positive, the orientation a circulation-coupled cavity needs.
    """
    if endo_normal not in ("cavity", "wall"):
        raise ValueError("endo_normal must be 'cavity' or 'wall'")
    nt, ns, nv = dims if dims is not None else lv_dims_for(n_target)
    if nt < 1 or ns < 1 or nv < 6:
        raise ValueError("lv_mesh needs nt >= 1, ns >= 1, nv >= 6")
    t = np.linspace(0.0, 1.0, nt + 1)
    rs = LV_ENDO[0] + (LV_EPI[0] - LV_ENDO[0]) * t
    rl = LV_ENDO[1] + (LV_EPI[1] - LV_ENDO[1]) * t
    uend = -np.arccos(LV_BASE_Z / rl)
    counts = lv_ring_counts(ns, nv)
    first = np.concatenate([[0], np.cumsum(counts)])
    nsurf = int(first[-1])
    s_v = np.concatenate([np.full(c, j / ns) for j, c in enumerate(counts)])
    ang_v = np.concatenate([np.arange(c) * (2.0 * np.pi / c) for c in counts])
    tris = np.concatenate([ring_strip_triangles(np.arange(first[j], first[j + 1]),
                                                np.arange(first[j + 1], first[j + 2])) for j in range(ns)])
    # vertex (layer i, surface vertex k) has id i * nsurf + k
    U = -np.pi + s_v[None, :] * (uend[:, None] + np.pi)
    RS, RL, V = rs[:, None], rl[:, None], ang_v[None, :]
    coords = np.stack([RS * np.sin(U) * np.cos(V), RS * np.sin(U) * np.sin(V),
                       np.broadcast_to(RL * np.cos(U), U.shape)], axis=-1).reshape(-1, 3)
    coords[:, 2] -= LV_BASE_Z
    tcoord_v = np.repeat(t, nsurf)                              # transmural parameter of vertices
    off = np.repeat(np.arange(nt, dtype=np.int64) * nsurf, len(tris))[:, None]
    tri = np.tile(tris, (nt, 1))
    tets = orient_t4(coords, split_prisms(np.concatenate([tri + off, tri + off + nsurf], axis=1)))
    faces = orient_faces_outward(coords, boundary_faces(tets), tets)
    is_endo = np.all(tcoord_v[faces] == 0.0, axis=1)
    endo = faces[is_endo]                                      # normal out of the wall, into the cavity
    if endo_normal == "wall":
        endo = endo[:, [0, 2, 1]]
    nvtx = coords.shape[0]
    if order == 10:
        coords, el, edges = t4_to_t10(coords, tets)
        endo = t3_to_t6(endo, edges, nvtx) if surface_type == 6 else t3_subdivide(endo, edges, nvtx)
    else:
        el = tets.astype(np.int32)
    bc = np.where(np.abs(coords[:, 2]) < 1e-12, 7, 0).astype(np.uint8)

    if order == 10:
        pts = quadrature_points(coords, el)
    else:
        pts = coords[el].mean(axis=1)[:, None, :]
    bases = lv_frames(pts, fibre_deg)
    cz = coords[el[:, :4]].mean(axis=1)[:, 2]
    z_apex = -LV_EPI[1] - LV_BASE_Z
    act = activation[0] + (activation[1] - activation[0]) * (cz - z_apex) / (0.0 - z_apex)
    E = el.shape[0]
    return Mesh(coords=coords, elements=el.astype(np.int32), bases=np.ascontiguousarray(bases), bc=bc,
                material=np.zeros(E, np.int32), activation=act, surfaces=endo.astype(np.int32),
                surface_id=np.ones(endo.shape[0], np.int32), name=f"lv{E}_T{order}")


def lv_transmural(p: np.ndarray) -> np.ndarray:
    """Transmural parameter t in [0,1] of points p (shifted frame): solve
    (r/rs(t))^2 + (z/rl(t))^2 = 1 by bisection."""
    r2 = p[..., 0] ** 2 + p[..., 1] ** 2
    z = p[..., 2] + LV_BASE_Z
    lo = np.full(r2.shape, -0.5)
    hi = np.full(r2.shape, 1.5)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        rs = LV_ENDO[0] + (LV_EPI[0] - LV_ENDO[0]) * mid
        rl = LV_ENDO[1] + (LV_EPI[1] - LV_ENDO[1]) * mid
        g = r2 / rs ** 2 + z ** 2 / rl ** 2 - 1.0
        outside = g > 0          # point outside the surface of parameter mid -> t larger
        lo = np.where(outside, mid, lo)
        hi = np.where(outside, hi, mid)
    return 0.5 * (lo + hi)


def lv_frames(p: np.ndarray, fibre_deg=(60.0, -60.0)) -> np.ndarray:
    t = np.clip(lv_transmural(p), 0.0, 1.0)
    rs = LV_ENDO[0] + (LV_EPI[0] - LV_ENDO[0]) * t
    rl = LV_ENDO[1] + (LV_EPI[1] - LV_ENDO[1]) * t
    z = p[..., 2] + LV_BASE_Z
    et = np.stack([p[..., 0] / rs ** 2, p[..., 1] / rs ** 2, z / rl ** 2], axis=-1)
    et /= np.linalg.norm(et, axis=-1, keepdims=True)
    r = np.sqrt(p[..., 0] ** 2 + p[..., 1] ** 2)
    ev = np.stack([-p[..., 1] / r, p[..., 0] / r, np.zeros_like(r)], axis=-1)
    ev = ev - np.sum(ev * et, axis=-1, keepdims=True) * et
    ev /= np.linalg.norm(ev, axis=-1, keepdims=True)
    el = np.cross(et, ev)
    a = np.deg2rad(fibre_deg[0] + (fibre_deg[1] - fibre_deg[0]) * t)[..., None]
    f = np.cos(a) * ev + np.sin(a) * el
    return frames_from_vectors(f, et)


# ---------------------------------------------------------------------------
# Tetgen files (CBModelLoaderTetgen.cpp:59-547) and LAT files (CBFileManager.cpp:120-194)
# ---------------------------------------------------------------------------

def write_tetgen(prefix: str, mesh: Mesh, unit: float = 1.0) -> None:
    N, E = mesh.n_nodes, mesh.n_elements
    with open(prefix + ".node", "w") as fh:
        fh.write(f"{N} 3 1 0\n")
        for i in range(N):
            x, y, z = mesh.coords[i] / unit
            fh.write(f"{i + 1} {x:.17g} {y:.17g} {z:.17g} {int(mesh.bc[i])}\n")
    with open(prefix + ".ele", "w") as fh:
        fh.write(f"{E} {mesh.npe} 1\n")
        for e in range(E):
            nodes = " ".join(str(int(k) + 1) for k in mesh.elements[e])
            fh.write(f"{e + 1} {nodes} {int(mesh.material[e])}\n")
    nb = mesh.bases.shape[1]
    with open(prefix + ".bases", "w") as fh:
        fh.write(f"{E} {nb}\n")
        for e in range(E):
            vals = []
            for q in range(nb):
                B = mesh.bases[e, q]
                vals += list(B[:, 0]) + list(B[:, 1]) + list(B[:, 2])
            fh.write(f"{e + 1} " + " ".join(f"{v:.17g}" for v in vals) + "\n")
    S = mesh.surfaces.shape[0]
    with open(prefix + ".sur", "w") as fh:
        fh.write(f"{S} {mesh.surfaces.shape[1]} 2\n")
        for s in range(S):
            nodes = " ".join(str(int(k) + 1) for k in mesh.surfaces[s])
            fh.write(f"{s + 1} {nodes} 0 {int(mesh.surface_id[s])}\n")


def read_tetgen(prefix: str, unit: float = 1.0, surface_ids=None) -> Mesh:
    def body(path):
        with open(path) as fh:
            lines = [ln.split("#")[0].split() for ln in fh]
        lines = [ln for ln in lines if ln]
        return lines[0], lines[1:]

    hdr, rows = body(prefix + ".node")
    N = int(hdr[0])
    has_attr = int(hdr[2]) > 0
    arr = np.array([[float(v) for v in r[:5 if has_attr else 4]] for r in rows[:N]])
    if not np.array_equal(arr[:, 0].astype(np.int64), np.arange(1, N + 1)):
        raise ValueError("node ids must be 1-based and consecutive")
    coords = arr[:, 1:4] * unit
    bc = arr[:, 4].astype(np.uint8) if has_attr else np.zeros(N, np.uint8)
    hdr, rows = body(prefix + ".ele")
    E, npe = int(hdr[0]), int(hdr[1])
    ea = np.array([[int(float(v)) for v in r[:npe + 2]] for r in rows[:E]])
    elements = (ea[:, 1:npe + 1] - 1).astype(np.int32)
    material = ea[:, npe + 1].astype(np.int32) if ea.shape[1] > npe + 1 else np.zeros(E, np.int32)
    nb = 5 if npe == 10 else 1
    bases = np.tile(np.eye(3), (E, nb, 1, 1))
    try:
        hdr, rows = body(prefix + ".bases")
        nq = int(hdr[1])
        vals = np.array([[float(v) for v in r[1:]] for r in rows[:E]])
        if vals.shape[1] == 9 * nq:
            b = vals.reshape(E, nq, 3, 3).transpose(0, 1, 3, 2)    # stored f, s, n -> columns
            bases = np.broadcast_to(b, (E, nb, 3, 3)).copy() if nq == 1 else b[:, :nb]
    except FileNotFoundError:
        pass
    surfaces = np.zeros((0, 3), np.int32)
    sid = np.zeros(0, np.int32)
    try:
        hdr, rows = body(prefix + ".sur")
        S, nps = int(hdr[0]), int(hdr[1])
        sa = np.array([[int(float(v)) for v in r[:nps + 3]] for r in rows[:S]])
        surfaces = (sa[:, 1:nps + 1] - 1).astype(np.int32)
        sid = sa[:, nps + 2].astype(np.int32)
        if surface_ids is not None:
            keep = np.isin(sid, list(surface_ids))
            surfaces, sid = surfaces[keep], sid[keep]
    except FileNotFoundError:
        pass
    return Mesh(coords=coords, elements=elements, bases=bases, bc=bc, material=material,
                activation=np.zeros(E), surfaces=surfaces, surface_id=sid, name=prefix)


def write_lat(path: str, activation: np.ndarray) -> None:
    with open(path, "w") as fh:
        fh.write("id time\n")
        for e, t in enumerate(activation):
            fh.write(f"{e} {t:.17g}\n")


def read_lat(path: str, n_elements: int) -> np.ndarray:
    """First line skipped; 'id time' rows; NaN -> 0; looked up by row position
    (CBFileManager.cpp:120-194, CBSolver.cpp:207-217)."""
    with open(path) as fh:
        rows = [ln.split() for ln in fh.readlines()[1:] if ln.strip()]
    t = np.array([float(r[1]) for r in rows[:n_elements]])
    return np.nan_to_num(t, nan=0.0)


# ---------------------------------------------------------------------------
# derived per-element data: reference shape-function derivatives
# ---------------------------------------------------------------------------

def t10_shape_derivatives(coords: np.ndarray, elements: np.ndarray):
    """dN/dX at the 4 QPs and the centroid (E, 5, 10, 3) and the centroid detJ (E,),
    CBElementSolidT10::CalcShapeFunctionDerivatives (:94-264), vectorised."""
    X = coords[elements]                     # (E, 10, 3)
    pts = np.concatenate([T10_QP_BARY, np.full((1, 4), 0.25)], axis=0)
    out = np.empty((elements.shape[0], 5, 10, 3))
    detJ = None
    for q, (l1, l2, l3, l4) in enumerate(pts):
        c = [X[:, k, :] for k in range(10)]
        J1 = 4.0 * (c[0] * (l1 - 0.25) + c[4] * l2 + c[6] * l3 + c[7] * l4)
        J2 = 4.0 * (c[4] * l1 + c[1] * (l2 - 0.25) + c[5] * l3 + c[8] * l4)
        J3 = 4.0 * (c[6] * l1 + c[5] * l2 + c[2] * (l3 - 0.25) + c[9] * l4)
        J4 = 4.0 * (c[7] * l1 + c[8] * l2 + c[9] * l3 + c[3] * (l4 - 0.25))
        Jx = [J1[:, 0], J2[:, 0], J3[:, 0], J4[:, 0]]
        Jy = [J1[:, 1], J2[:, 1], J3[:, 1], J4[:, 1]]
        Jz = [J1[:, 2], J2[:, 2], J3[:, 2], J4[:, 2]]

        def dd(v, i, j):
            return v[i - 1] - v[j - 1]
        d = (dd(Jx, 2, 1) * (dd(Jy, 2, 3) * dd(Jz, 3, 4) - dd(Jy, 3, 4) * dd(Jz, 2, 3))
             + dd(Jx, 3, 2) * (dd(Jy, 3, 4) * dd(Jz, 1, 2) - dd(Jy, 1, 2) * dd(Jz, 3, 4))
             + dd(Jx, 4, 3) * (dd(Jy, 1, 2) * dd(Jz, 2, 3) - dd(Jy, 2, 3) * dd(Jz, 1, 2)))
        a = [dd(Jy, 4, 2) * dd(Jz, 3, 2) - dd(Jy, 3, 2) * dd(Jz, 4, 2),
             dd(Jy, 3, 1) * dd(Jz, 4, 3) - dd(Jy, 3, 4) * dd(Jz, 1, 3),
             dd(Jy, 2, 4) * dd(Jz, 1, 4) - dd(Jy, 1, 4) * dd(Jz, 2, 4),
             dd(Jy, 1, 3) * dd(Jz, 2, 1) - dd(Jy, 1, 2) * dd(Jz, 3, 1)]
        b = [dd(Jx, 3, 2) * dd(Jz, 4, 2) - dd(Jx, 4, 2) * dd(Jz, 3, 2),
             dd(Jx, 4, 3) * dd(Jz, 3, 1) - dd(Jx, 1, 3) * dd(Jz, 3, 4),
             dd(Jx, 1, 4) * dd(Jz, 2, 4) - dd(Jx, 2, 4) * dd(Jz, 1, 4),
             dd(Jx, 2, 1) * dd(Jz, 1, 3) - dd(Jx, 3, 1) * dd(Jz, 1, 2)]
        cc = [dd(Jx, 4, 2) * dd(Jy, 3, 2) - dd(Jx, 3, 2) * dd(Jy, 4, 2),
              dd(Jx, 3, 1) * dd(Jy, 4, 3) - dd(Jx, 3, 4) * dd(Jy, 1, 3),
              dd(Jx, 2, 4) * dd(Jy, 1, 4) - dd(Jx, 1, 4) * dd(Jy, 2, 4),
              dd(Jx, 1, 3) * dd(Jy, 2, 1) - dd(Jx, 1, 2) * dd(Jy, 3, 1)]
        fac = 4.0 / d
        for comp, v in enumerate((a, b, cc)):
            out[:, q, 0, comp] = fac * (l1 - 0.25) * v[0]
            out[:, q, 1, comp] = fac * (l2 - 0.25) * v[1]
            out[:, q, 2, comp] = fac * (l3 - 0.25) * v[2]
            out[:, q, 3, comp] = fac * (l4 - 0.25) * v[3]
            out[:, q, 4, comp] = fac * (l1 * v[1] + l2 * v[0])
            out[:, q, 5, comp] = fac * (l2 * v[2] + l3 * v[1])
            out[:, q, 6, comp] = fac * (l3 * v[0] + l1 * v[2])
            out[:, q, 7, comp] = fac * (l1 * v[3] + l4 * v[0])
            out[:, q, 8, comp] = fac * (l2 * v[3] + l4 * v[1])
            out[:, q, 9, comp] = fac * (l3 * v[3] + l4 * v[2])
        detJ = d
    return out, detJ


def t4_shape_derivatives(coords: np.ndarray, elements: np.ndarray):
    """CBElementSolidT4::CalcShapeFunctionDerivatives (:73-121) -> ((E, 4, 3), detJ)."""
    X = coords[elements[:, :4]]
    x = X[..., 0]
    y = X[..., 1]
    z = X[..., 2]
    z43, z42, z41 = z[:, 3] - z[:, 2], z[:, 3] - z[:, 1], z[:, 3] - z[:, 0]
    z32, z31, z21 = z[:, 2] - z[:, 1], z[:, 2] - z[:, 0], z[:, 1] - z[:, 0]
    y43, y42, y41 = y[:, 3] - y[:, 2], y[:, 3] - y[:, 1], y[:, 3] - y[:, 0]
    y32, y31, y21 = y[:, 2] - y[:, 1], y[:, 2] - y[:, 0], y[:, 1] - y[:, 0]
    x41, x31, x21 = x[:, 3] - x[:, 0], x[:, 2] - x[:, 0], x[:, 1] - x[:, 0]
    d = x21 * (y31 * z41 - y41 * z31) + y21 * (x41 * z31 - x31 * z41) + z21 * (x31 * y41 - x41 * y31)
    g = np.empty((X.shape[0], 4, 3))
    g[:, 0, 0] = (y[:, 1] * z43 - y[:, 2] * z42 + y[:, 3] * z32) / d
    g[:, 1, 0] = (-y[:, 0] * z43 + y[:, 2] * z41 - y[:, 3] * z31) / d
    g[:, 2, 0] = (y[:, 0] * z42 - y[:, 1] * z41 + y[:, 3] * z21) / d
    g[:, 3, 0] = (-y[:, 0] * z32 + y[:, 1] * z31 - y[:, 2] * z21) / d
    g[:, 0, 1] = (-x[:, 1] * z43 + x[:, 2] * z42 - x[:, 3] * z32) / d
    g[:, 1, 1] = (x[:, 0] * z43 - x[:, 2] * z41 + x[:, 3] * z31) / d
    g[:, 2, 1] = (-x[:, 0] * z42 + x[:, 1] * z41 - x[:, 3] * z21) / d
    g[:, 3, 1] = (x[:, 0] * z32 - x[:, 1] * z31 + x[:, 2] * z21) / d
    g[:, 0, 2] = (x[:, 1] * y43 - x[:, 2] * y42 + x[:, 3] * y32) / d
    g[:, 1, 2] = (-x[:, 0] * y43 + x[:, 2] * y41 - x[:, 3] * y31) / d
    g[:, 2, 2] = (x[:, 0] * y42 - x[:, 1] * y41 + x[:, 3] * y21) / d
    g[:, 3, 2] = (-x[:, 0] * y32 + x[:, 1] * y31 - x[:, 2] * y21) / d
    return g, d


def smooth_displacement(coords: np.ndarray, length: float, amplitude: float = 0.02, seed: int = 1234):
    """u_i = A L a_i sin(2 pi k_i.X / L + phi_i), the seeded deformed state used for
    element-throughput measurements (all from numpy default_rng(seed))."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1.0, 1.0, 3)
    k = rng.normal(size=(3, 3))
    k /= np.linalg.norm(k, axis=1, keepdims=True)
    phi = rng.uniform(0.0, 2 * np.pi, 3)
    arg = 2 * np.pi * (coords @ k.T) / length + phi
    return amplitude * length * a * np.sin(arg)
