"""Device-built vev index: bit-exact against the reference's fixtures, the
reference's own builder (golden .npz) and the oracle."""
import numpy as np
import pytest
import torch

from conftest import ref_index_files, to_dev
from simit_amd import meshes

pytestmark = pytest.mark.gpu


def dev_sets(sets):
    out = []
    for s in sets:
        a = np.ascontiguousarray(s, dtype=np.int32)
        out.append(to_dev(a if a.ndim == 2 else a.reshape(len(a), -1)))
    return out


def build(ctx, nv, sets, force_diag=False):
    idx = ctx.build_index(nv, dev_sets(sets), force_diag=force_diag)
    return idx, idx.to_host()


def rows(rp, ci):
    return [list(ci[rp[i]:rp[i + 1]]) for i in range(len(rp) - 1)]


def test_reference_fixtures(ctx, fixtures):
    for name, fx in fixtures.items():
        if not name.startswith("pathindex"):
            continue
        print("fixture", name, flush=True)
        sets = [np.array(s, dtype=np.int32) for s in fx["sets"]]
        idx, (rp, ci) = build(ctx, fx["nverts"], sets)
        if fx["kind"] == "ve":
            vp, ve = idx.ve_to_host(0)
            assert rows(vp, ve) == fx["expected"], name
        else:
            assert rows(rp, ci) == fx["expected"], name


@pytest.mark.parametrize("path", ref_index_files(), ids=lambda p: p.stem)
def test_matches_reference_builder(ctx, path):
    z = np.load(path)
    sets = [z[k] for k in sorted(z.files) if k.startswith("set")]
    _, (rp, ci) = build(ctx, int(z["nverts"]), sets)
    np.testing.assert_array_equal(rp, z["rowptr"])
    np.testing.assert_array_equal(ci, z["colidx"])


@pytest.mark.parametrize("kind,n", [("box", 12), ("tet", 10), ("hex", 9), ("tet", 32)])
def test_matches_oracle_on_grids(ctx, oracle, kind, n):
    if kind == "box":
        nv, ep = meshes.create_box(n, n + 1, n + 2)
    elif kind == "tet":
        x, ep = meshes.tet5_grid(n)
        nv = x.shape[0]
    else:
        x, ep = meshes.hex_grid(n)
        nv = x.shape[0]
    print(kind, n, nv, ep.shape, flush=True)
    idx, (rp, ci) = build(ctx, nv, [ep])
    orp, oci = oracle.index_vev(nv, [ep])
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(ci, oci)
    # incidence lists: ascending element ids, same as Link::ve
    vp, ve = idx.ve_to_host(0)
    ovp, ove = oracle.index_ve(nv, ep)
    np.testing.assert_array_equal(vp, ovp)
    np.testing.assert_array_equal(ve, ove)


@pytest.mark.parametrize("card", [1, 2, 3, 4, 8])
def test_ragged_random_hypergraphs(ctx, oracle, card):
    rng = np.random.default_rng(100 + card)
    nv = 300
    ep = rng.integers(0, nv - 20, size=(500, card)).astype(np.int32)  # 20 isolated vertices
    ep[::7, 0] = ep[::7, -1]  # repeated vertices inside elements
    for fd in (False, True):
        _, (rp, ci) = build(ctx, nv, [ep], force_diag=fd)
        orp, oci = oracle.index_vev(nv, [ep], force_diag=fd)
        np.testing.assert_array_equal(rp, orp)
        np.testing.assert_array_equal(ci, oci)


def test_union_of_edge_sets(ctx, oracle):
    rng = np.random.default_rng(5)
    nv = 200
    sets = [rng.integers(0, nv, size=(300, 2)).astype(np.int32),
            rng.integers(0, nv, size=(100, 4)).astype(np.int32),
            np.zeros((0, 3), np.int32)]
    _, (rp, ci) = build(ctx, nv, sets)
    orp, oci = oracle.index_vev(nv, sets)
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(ci, oci)


def test_wide_rows_take_the_workgroup_path(ctx, oracle):
    """A hub vertex in 400 springs (> 64 neighbours per row and > 64 incidences)."""
    nv = 500
    hub = np.stack([np.zeros(400, np.int32), np.arange(1, 401, dtype=np.int32)], axis=1)
    rng = np.random.default_rng(9)
    rest = rng.integers(1, nv, size=(600, 2)).astype(np.int32)
    ep = np.concatenate([rest[:300], hub, rest[300:]])
    idx, (rp, ci) = build(ctx, nv, [ep])
    orp, oci = oracle.index_vev(nv, [ep])
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(ci, oci)
    vp, ve = idx.ve_to_host(0)
    ovp, ove = oracle.index_ve(nv, ep)
    np.testing.assert_array_equal(ve, ove)


def test_empty_and_bad_inputs(ctx):
    import simit_amd
    _, (rp, ci) = build(ctx, 5, [np.zeros((0, 2), np.int32)])
    assert list(rp) == [0] * 6 and len(ci) == 0
    with pytest.raises(simit_amd.SimitError) as e:
        build(ctx, 3, [np.array([[0, 7]], np.int32)])
    assert e.value.code == simit_amd.ERR_INVALID


def test_locs_match_loc_scan(ctx, oracle):
    x, ep = meshes.tet5_grid(4)
    nv = x.shape[0]
    idx, (rp, ci) = build(ctx, nv, [ep])
    out = torch.empty(ep.shape[0] * 16, dtype=torch.int32, device="cuda")
    idx.locs(0, out)
    got = out.cpu().numpy().reshape(-1, 4, 4)
    for e in range(0, ep.shape[0], 17):
        for a in range(4):
            for b in range(4):
                assert got[e, a, b] == oracle.loc(rp, ci, int(ep[e, a]), int(ep[e, b]))


def test_full_size_tet_index_properties(ctx):
    """96^3 x 5 tets (the 2.7M-DOF mesh): size-independent invariants."""
    n = 96
    x, ep = meshes.tet5_grid(n)
    nv = x.shape[0]
    idx = ctx.build_index(nv, [to_dev(ep)])
    rp, ci = idx.to_host()
    axis, face = 3 * n * (n + 1) ** 2, 3 * n * n * (n + 1)
    assert int(rp[-1]) == nv + 2 * (axis + face) == 11_695_969
    starts = rp[:-1]
    # sorted, unique, diagonal present in every row
    d = np.diff(ci.astype(np.int64))
    row_start = np.zeros(len(ci), bool)
    row_start[starts] = True
    assert (d[~row_start[1:]] > 0).all()
    diag = np.repeat(np.arange(nv), np.diff(rp))
    assert ((ci == diag).sum()) == nv
    # symmetric pattern: (i,j) present iff (j,i)
    keys = diag.astype(np.int64) * nv + ci
    tkeys = np.sort(ci.astype(np.int64) * nv + diag)
    np.testing.assert_array_equal(np.sort(keys), tkeys)
