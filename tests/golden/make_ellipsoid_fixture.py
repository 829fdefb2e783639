"""Writes tests/golden/ellipsoid_subset.npz from the reference's benchmark mesh
examples/benchmark2015/tetgen/ellipsoid.{node,ele,bases,sur} (Land et al. 2015
problem 2 geometry, read with Tetgen.Unit = 1e-3 as in Problem2.xml).

The fixture is input data only: the first 400 T10 elements with a face on the
endocardium (after the reference's CheckNodeSorting), their nodes renumbered compactly, their 5 fibre bases as stored
in the file (6 significant digits, so not exactly orthonormal), the nodes' BC
masks, and the endocardial (surface 1) T3 triangles that are faces of the subset's
elements.  Also stores the full mesh's counts as a known-answer check of the reader.
Run in a container that has the reference checkout:
    python tests/golden/make_ellipsoid_fixture.py /root/reference
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from cardiomechanics_amd import mesh as M  # noqa: E402


def main(ref_root):
    prefix = os.path.join(ref_root, "examples/benchmark2015/tetgen/ellipsoid")
    mesh = M.read_tetgen(prefix, unit=1e-3)
    el, bases = M.check_node_sorting_t10(mesh.elements, mesh.coords, mesh.bases)
    n_el = 400
    # elements carrying an endocardial (surface 1) triangle, in file order; the
    # file's T3 triangles are sub-triangles of the quadratic faces (corner/midside nodes)
    endo_all = mesh.surfaces[mesh.surface_id == 1]
    owner = {}
    node_sets = [set(r) for r in el.tolist()]
    inc = {}
    for e, r in enumerate(el.tolist()):
        for nd in r:
            inc.setdefault(nd, []).append(e)
    for t in endo_all.tolist():
        for e in inc.get(t[0], []):
            if t[1] in node_sets[e] and t[2] in node_sets[e]:
                owner.setdefault(e, []).append(t)
                break
    chosen = sorted(owner)[:n_el]
    sub = el[chosen]
    bases = bases[chosen]
    used = np.unique(sub)
    remap = -np.ones(mesh.n_nodes, np.int64)
    remap[used] = np.arange(used.size)
    surf = np.array([t for e in chosen for t in owner[e]], dtype=np.int64)
    keep = np.ones(surf.shape[0], bool)
    out = dict(coords=mesh.coords[used], elements=remap[sub].astype(np.int32), bases=bases,
               bc=mesh.bc[used], surfaces=remap[surf[keep]].astype(np.int32),
               full_counts=np.array([mesh.n_nodes, mesh.n_elements, int((mesh.bc == 7).sum()),
                                     int((mesh.surface_id == 1).sum())]))
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ellipsoid_subset.npz")
    np.savez_compressed(dst, **out)
    print(dst, {k: v.shape for k, v in out.items()}, out["full_counts"])


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
