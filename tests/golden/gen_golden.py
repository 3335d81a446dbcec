"""Writes the golden fixtures under tests/golden/.

1. reference_fixtures.json: inputs and expected outputs transcribed from the
   reference's own tests (each entry cites its file:line).  They are data:
   graphs, field values and asserted results.
2. ref_index_*.npz: vev / ve indices of small meshes computed by the
   REFERENCE's own PathIndexBuilder (src/path_indices.cpp), compiled from the
   reference sources by oracle/Makefile into oracle/_ref/ref_index.  Needs the
   reference checkout; the committed .npz files are what tests read.
3. tet-bunny/: the reference's tetgen bunny mesh (apps/data/tet-bunny, data
   files copied unchanged) with its vev indices (ref_index_bunny_*.npz).

Run:  python tests/golden/gen_golden.py   (after `make -C oracle all ref`)
"""
import json
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

FIXTURES = {
    # ---- test/path_indices-test.cpp ------------------------------------------------
    "pathindex_link_ve": {
        "cite": "test/path_indices-test.cpp:17-56 (createBox(5,1,1), VERIFY_INDEX ve)",
        "nverts": 5, "sets": [[[0, 1], [1, 2], [2, 3], [3, 4]]], "kind": "ve",
        "expected": [[0], [0, 1], [1, 2], [2, 3], [3]],
    },
    "pathindex_link_ve_isolated": {
        "cite": "test/path_indices-test.cpp:70-77 (G = {(v0,v4)}, ve with neighbourless vertices)",
        "nverts": 5, "sets": [[[0, 4]]], "kind": "ve",
        "expected": [[0], [], [], [], [0]],
    },
    "pathindex_exist_and_vev": {
        "cite": "test/path_indices-test.cpp:153-174 (createBox(3,1,1), vev)",
        "nverts": 3, "sets": [[[0, 1], [1, 2]]], "kind": "vev",
        "expected": [[0, 1], [0, 1, 2], [1, 2]],
    },
    "pathindex_or_vev_vfv": {
        "cite": "test/path_indices-test.cpp:113-150 + path_indices-tests.h:36-45 "
                "(createTestGraph0, vev or vfv)",
        "nverts": 4, "sets": [[[0, 1], [1, 2]], [[2, 3], [0, 3], [0, 1]]], "kind": "vev",
        "expected": [[0, 1, 3], [0, 1, 2], [1, 2, 3], [0, 2, 3]],
    },
    "pathindex_alias": {
        "cite": "test/path_indices-test.cpp:315-339 (E = {(v0,v1),(v1,v0)})",
        "nverts": 2, "sets": [[[0, 1], [1, 0]]], "kind": "vev",
        "expected": [[0, 1], [0, 1]],
    },
    "pathindex_hypergraph": {
        "cite": "test/path_indices-test.cpp:341-378 (E cardinality 3, F empty)",
        "nverts": 4, "sets": [[[0, 1, 2], [1, 2, 3]]], "kind": "vev",
        "expected": [[0, 1, 2], [0, 1, 2, 3], [0, 1, 2, 3], [1, 2, 3]],
    },
    # ---- test/program-cg-test.cpp + test/input/program/cg.sim -------------------------
    "program_cg": {
        "cite": "test/program-cg-test.cpp:10-44; test/input/program/cg.sim:1-60",
        "b": [1.0, 2.0, 3.0], "springs": [[0, 1], [1, 2]], "a": [4.0, 5.0],
        "tol": 1e-6, "maxiters": 5,
        "expected_c": [0.95883777239709455653, 1.98789346246973352983,
                       3.05326876513317202466],
        "abs_tol": 1e-13,
    },
    # ---- test/program-isprings.cpp:91-175 + test/input/program/isprings.sim ------------
    "program_isprings": {
        "cite": "test/program-isprings.cpp:91-175; test/input/program/isprings.sim:19-103",
        "l0": 1.0, "stiffness": 100.0, "density": 1000.0, "radius": 0.01,
        "pi": 3.14159265358979, "zfloor": 0.4,
        "x": [[0.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 2.0]],
        "springs": [[0, 1], [1, 2]], "steps": 10,
        "h": 1e-2, "viscous": 1e-1, "grav": [0.0, 0.0, -9.81], "tol": 1e-5, "maxiters": 25,
        "expected_v": [[0.0, 0.0, 0.0], [0.0, 0.0, -0.54482994491692449],
                       [0.0, 0.0, -0.8480293605755741]],
        "expected_x": [[0.0, 0.0, 0.0], [0.0, 0.0, 0.96180188527075183],
                       [0.0, 0.0, 1.9500008677096612]],
        "abs_tol": 1e-13,
    },
    # ---- test/system-gemv-tests.cpp:854-903 + input/system/gemv_blocked.sim -----------
    "gemv_blocked": {
        "cite": "test/system-gemv-tests.cpp:854-903; test/input/system/gemv_blocked.sim",
        "b": [[1.0, 2.0], [3.0, 4.0], [5.0, 6.0]], "springs": [[0, 1], [1, 2]],
        "a": [[1.0, 2.0, 3.0, 4.0], [5.0, 6.0, 7.0, 8.0]],
        "expected_c": [[16.0, 36.0], [116.0, 172.0], [100.0, 136.0]],
    },
    # ---- test/assembly-tests.cpp:264-291 + input/assembly/matrix_vv.sim ---------------
    "assembly_matrix_vv": {
        "cite": "test/assembly-tests.cpp:264-291; test/input/assembly/matrix_vv.sim",
        "a": [1, 1, 1], "springs": [[0, 1], [1, 2]], "expected_b": [2, 4, 2],
    },
    # ---- test/system-vector-tests.cpp:51-124 ------------------------------------------
    "vector_dot": {"cite": "test/system-vector-tests.cpp:51-66",
                   "x": [1.0, 2.0, 3.0], "expected": 14.0},
    "vector_dot_blocked": {"cite": "test/system-vector-tests.cpp:68-84",
                           "x": [1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0],
                           "expected": 285.0},
    "vector_dot_intrinsic": {"cite": "test/system-vector-tests.cpp:86-124",
                             "x": [float(i) for i in range(1, 12)], "expected": 506.0},
}


def reference_indices(tmp):
    """vev indices of small meshes from the reference's own builder."""
    import pyoracle as O
    from simit_amd import meshes

    out = {}
    npts, ep = meshes.create_box(4, 3, 5)
    out["box435_springs"] = (npts, [ep])
    x, tets = meshes.tet5_grid(3)
    out["tet5_n3"] = (x.shape[0], [tets])
    x, hexes = meshes.hex_grid(2)
    out["hex_n2"] = (x.shape[0], [hexes])
    rng = np.random.default_rng(7)
    # a ragged hypergraph with repeated vertices inside elements and isolated vertices
    tri = rng.integers(0, 30, size=(25, 3)).astype(np.int32)
    tri[3] = [5, 5, 9]
    out["ragged_tri"] = (40, [tri])
    for name, (nv, sets) in out.items():
        rp, ci = O.ref_index_vev(nv, sets, tmp)
        arrays = {"nverts": np.array(nv), "rowptr": rp, "colidx": ci}
        for s, e in enumerate(sets):
            arrays[f"set{s}"] = e
        np.savez_compressed(HERE / f"ref_index_{name}.npz", **arrays)
        print(name, nv, "rows", len(ci), "nnz")
    # the reference's createBox edge order
    ep = O.ref_create_box(4, 3, 5, tmp)
    np.savez_compressed(HERE / "ref_create_box_435.npz", endpoints=ep)
    # the reference's tetgen bunny (apps/data/tet-bunny: data files, copied
    # as they are) and its vev indices over the tets and over the edges
    src = Path("/root/reference/apps/data/tet-bunny")
    dst = HERE / "tet-bunny"
    dst.mkdir(exist_ok=True)
    for f in ("bunny.1.node", "bunny.1.ele", "bunny.1.edge"):
        (dst / f).write_bytes((src / f).read_bytes())
    m = meshes.load_tetgen(dst / "bunny.1.node", dst / "bunny.1.ele", dst / "bunny.1.edge")
    nv = m["x"].shape[0]
    for name, sets in (("bunny_tets", [m["elements"]]), ("bunny_edges", [m["edges"]])):
        rp, ci = O.ref_index_vev(nv, sets, tmp)
        arrays = {"nverts": np.array(nv), "rowptr": rp, "colidx": ci, "set0": sets[0]}
        np.savez_compressed(HERE / f"ref_index_{name}.npz", **arrays)
        print(name, nv, "rows", len(ci), "nnz")


if __name__ == "__main__":
    (HERE / "reference_fixtures.json").write_text(json.dumps(FIXTURES, indent=1))
    print("wrote reference_fixtures.json")
    import pyoracle as O
    if O.ref_index_available():
        with tempfile.TemporaryDirectory() as tmp:
            reference_indices(tmp)
    else:
        print("oracle/_ref/ref_index not built: skipped the reference-builder fixtures")
