"""Per-frame input formats: the Houdini ``.geo`` reader/writer and the per-tet
activation sources (src/muscle/util/geo_io.py, src/muscle/util/activation.py).

The reference's own .geo tests (tests/test_geo_io.py, tests/test_fascia_geo_io.py)
run on asset files that are git-lfs pointers in the reference tree, so these tests
write files of the same layout and check the properties those tests check (dtypes,
shapes, index ranges, groups without duplicates, the cached read equal to the full
parse) plus round trips against the arrays that were written."""
import json
import os
import time
from pathlib import Path

import numpy as np
import pytest

from rlmusclev2_amd import geo
from rlmusclev2_amd.activation import MappedActivationSource, RampActivationSource


def _tet_asset(rng, nv=300, nt=500):
    pos = rng.standard_normal((nv, 3)).astype(np.float32)
    tets = np.stack([rng.choice(nv, 4, replace=False) for _ in range(nt)]).astype(np.int32)
    return pos, tets


def test_restpos_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    pos, tets = _tet_asset(rng)
    glue = rng.random(len(pos)) < 0.3
    mid = rng.integers(0, 3, len(tets))
    f = tmp_path / "restpos.geo"
    geo.write_geo(f, pos, tets, point_attrs={"gluetoanimation": glue, "materialW": rng.random((len(pos), 3)),
                                             "tissuetomuscledistance": np.full(len(pos), 0.01)},
                  prim_attrs={"fiberstrength": rng.random(len(tets))},
                  prim_strings={"muscle_id": (mid, ["L_a", "L_b", "R_c"])})
    g = geo.GeoFile(str(f))
    assert g.positions.dtype == np.float32 and g.positions.shape == (len(pos), 3)
    np.testing.assert_array_equal(g.positions, pos)
    assert g.tet_indices.dtype == np.int32 and g.tet_indices.shape == (len(tets), 4)
    np.testing.assert_array_equal(g.tet_indices, tets)
    assert g.gluetoanimation.dtype == np.bool_
    np.testing.assert_array_equal(g.gluetoanimation, glue)
    assert g.point_attr("materialW").shape == (len(pos), 3)
    d = g.point_attr("tissuetomuscledistance")
    assert d.shape == (len(pos),) and np.allclose(d, 0.01, atol=1e-6)
    assert g.has_prim_attr("muscle_id") and not g.has_prim_attr("nope")
    np.testing.assert_array_equal(g.prim_attr("muscle_id"), mid)
    assert g.prim_attr("muscle_id").dtype == np.int32
    assert g.prim_attr_strings("muscle_id") == ["L_a", "L_b", "R_c"]
    assert g.prim_attr("fiberstrength").shape == (len(tets),)
    with pytest.raises(KeyError):
        _ = g.pts
    with pytest.raises(KeyError):
        g.point_attr("nope")


def test_missing_glue_is_all_false(tmp_path):
    pos, tets = _tet_asset(np.random.default_rng(1), 40, 30)
    f = tmp_path / "a.geo"
    geo.write_geo(f, pos, tets)
    g = geo.GeoFile(str(f))
    assert g.gluetoanimation.dtype == np.bool_ and g.gluetoanimation.shape == (40,)
    assert not g.gluetoanimation.any()


def test_cons_frame_and_cache(tmp_path):
    rng = np.random.default_rng(2)
    nc = 257
    pts = rng.choice(1000, nc, replace=False).astype(np.int32)
    tgt = rng.standard_normal((nc, 3)).astype(np.float32)
    f = tmp_path / "cons_45.geo"
    geo.write_geo(f, tgt, point_attrs={"pts": pts, "target_pos": tgt})
    g = geo.GeoFile(str(f))
    assert g.pts.dtype == np.int32 and g.target_pos.dtype == np.float32
    assert g.target_pos.shape == (g.pts.shape[0], 3)
    np.testing.assert_array_equal(g.pts, pts)
    fast = geo.read_cons_frame(str(tmp_path), 45)
    np.testing.assert_array_equal(fast, g.target_pos)
    cache = tmp_path / "_npy_cache" / "cons_45_target_pos.npy"
    assert cache.exists()
    assert not list((tmp_path / "_npy_cache").glob("*.tmp*"))
    # the cached copy is what is read while it is not older than the source
    np.save(cache, tgt + 1.0)
    np.testing.assert_array_equal(geo.read_cons_frame(str(tmp_path), 45), tgt + 1.0)
    # an edited source re-parses and rewrites the cache
    tgt2 = tgt * 2.0
    geo.write_geo(f, tgt2, point_attrs={"pts": pts, "target_pos": tgt2})
    later = time.time() + 5.0
    os.utime(f, (later, later))
    np.testing.assert_array_equal(geo.read_cons_frame(str(tmp_path), 45), tgt2)
    np.testing.assert_array_equal(np.load(cache), tgt2)


def test_point_groups_rle_and_ordered(tmp_path):
    rng = np.random.default_rng(3)
    n = 500
    pos = rng.standard_normal((n, 3)).astype(np.float32)
    inner = np.sort(rng.choice(n, 120, replace=False))
    edge = np.r_[0:17, 480:500]                          # runs touching both ends
    rigid = rng.choice(n, 40, replace=False)              # ordered lists keep their order
    for rle in (True, False):
        f = tmp_path / f"g{int(rle)}.geo"
        geo.write_geo(f, pos, groups={"tissue_inner_points": inner, "edge": edge}, rle_groups=rle)
        g = geo.GeoFile(str(f))
        for name, want in (("tissue_inner_points", inner), ("edge", edge)):
            got = g.point_group(name)
            assert got.dtype == np.int32 and got.ndim == 1
            np.testing.assert_array_equal(got, want)
            assert len(np.unique(got)) == len(got)
    f = tmp_path / "ordered.geo"
    geo.write_geo(f, pos, groups={"rigid": rigid}, rle_groups=False)
    np.testing.assert_array_equal(geo.GeoFile(str(f)).point_group("rigid"), rigid)
    assert geo.GeoFile(str(f)).has_point_group("rigid")
    with pytest.raises(KeyError):
        geo.GeoFile(str(f)).point_group("absent")


def test_bool_rle_decoder_by_hand():
    """Run-length list decoded by hand: 3 off, 2 on, 1 off, 4 on."""
    got = geo._bool_rle(["boolRLE", [3, False, 2, True, 1, False, 4, True]])
    np.testing.assert_array_equal(got, [3, 4, 6, 7, 8, 9])
    assert geo._bool_rle([]).size == 0
    assert geo._bool_rle(["other", [1, True]]).size == 0


def test_mixed_tet_and_polygon_blocks(tmp_path):
    """A file with a Tetrahedron_run followed by polygons keeps the tets only
    (the fascia asset layout, tests/test_fascia_geo_io.py)."""
    rng = np.random.default_rng(4)
    pos, tets = _tet_asset(rng, 60, 25)
    tris = rng.integers(0, 60, (9, 3)).astype(np.int32)
    f = tmp_path / "fascia.geo"
    geo.write_geo(f, pos, tets)
    root = json.loads(f.read_text())
    d = dict(zip(root[0::2], root[1::2]))
    d["topology"][1][1] += [int(i) for i in tris.reshape(-1)]
    d["primitives"].append([["type", "Polygon_run"], ["startvertex", tets.size, "nprimitives", 9,
                                                      "nvertices", [3] * 9]])
    d["primitivecount"] = 34
    f.write_text(json.dumps([x for kv in d.items() for x in kv]))
    g = geo.GeoFile(str(f))
    assert g.tet_indices.shape == (25, 4)
    np.testing.assert_array_equal(g.tet_indices, tets)


def test_triangle_mesh_and_bone_source(tmp_path):
    """Surface triangles without a tet block come back as (ntri, 3); the bone
    source concatenates them and reads positions per frame, cached and scaled."""
    rng = np.random.default_rng(5)
    nv = 50
    tris = rng.integers(0, nv, (70, 3)).astype(np.int32)
    frames = {f: rng.standard_normal((nv, 3)).astype(np.float32) for f in (1, 2, 3)}
    for fr, p in frames.items():
        geo.write_geo(tmp_path / f"bone_{fr}.geo", p, tris)
    g = geo.GeoFile(str(tmp_path / "bone_1.geo"))
    np.testing.assert_array_equal(g.tet_indices, tris)
    src = geo.GeoBoneSource(str(tmp_path), start_frame=1, scale=0.01)
    assert src.n_verts == nv
    np.testing.assert_array_equal(src.tris, tris.reshape(-1))
    for fr, p in frames.items():
        got = src.positions(fr)
        assert got.dtype == np.float32 and got.flags.c_contiguous
        np.testing.assert_allclose(got, p * np.float32(0.01), rtol=0, atol=0)
    assert (tmp_path / "_npy_cache" / "bone_2_P.npy").exists()


def test_geo_writer_rewrites_in_place(tmp_path):
    rng = np.random.default_rng(6)
    pos, tets = _tet_asset(rng, 80, 60)
    mid = rng.integers(0, 2, 60)
    tmpl = tmp_path / "tmpl.geo"
    geo.write_geo(tmpl, pos, tets, point_attrs={"gluetoanimation": np.zeros(80, bool)},
                  prim_strings={"muscle_id": (mid, ["a", "b"])}, groups={"g": np.arange(10)})
    w = geo.GeoWriter(str(tmpl))
    for k in range(2):
        p2 = (pos + k + 1).astype(np.float32)
        col = (rng.random(80) < 0.5).astype(np.int32)
        dep = rng.random(80).astype(np.float32)
        out = tmp_path / f"out{k}.geo"
        w.write(str(out), p2, is_collide=col, pen_depth=dep)
        g = geo.GeoFile(str(out))
        np.testing.assert_array_equal(g.positions, p2)
        np.testing.assert_array_equal(g.point_attr("is_collide"), col)
        np.testing.assert_array_equal(g.point_attr("pen_depth"), dep)
        np.testing.assert_array_equal(g.tet_indices, tets)
        np.testing.assert_array_equal(g.prim_attr("muscle_id"), mid)
        np.testing.assert_array_equal(g.point_group("g"), np.arange(10))
        # the attribute is overwritten, not appended, on the second write
        names = [e[0][5] for e in dict(zip(json.loads(out.read_text())[0::2],
                                           json.loads(out.read_text())[1::2]))["attributes"][1]]
        assert names.count("is_collide") == 1 and names.count("pen_depth") == 1
    with pytest.raises(ValueError):
        w.write(str(tmp_path / "bad.geo"), pos[:10])


# ---------------------------------------------------------------- activation
def test_ramp_activation():
    mask = np.array([0.0, 1.0, 0.5], np.float32)
    src = RampActivationSource(mask, 10, 20)
    np.testing.assert_array_equal(src.per_tet(5), 0 * mask)
    np.testing.assert_allclose(src.per_tet(15), 0.5 * mask)
    np.testing.assert_array_equal(src.per_tet(20), mask)
    np.testing.assert_array_equal(src.per_tet(99), mask)
    assert src.per_tet(12).dtype == np.float32
    one = RampActivationSource(mask, 7, 7)            # a zero span counts as one frame
    np.testing.assert_array_equal(one.per_tet(8), mask)


def _mapping(tmp_path, entries):
    p = tmp_path / "muscle_mapping.json"
    p.write_text(json.dumps({"mappings": entries}))
    return p


@pytest.mark.parametrize("agg", ["mean", "max"])
def test_mapped_activation_matches_per_tet_loop(tmp_path, agg):
    rng = np.random.default_rng(7)
    start, end = 3, 12
    nf = end - start + 1
    strings = ["L_biceps", "L_triceps", "R_soleus", "R_passive"]
    entries = {"Bicep_Long/curve": {"reference": "name", "value": "L_biceps"},
               "Bicep_Short/curve": {"reference": "name", "value": "L_biceps"},
               "Bicep_Short1/curve": {"value": "L_biceps"},
               "Triceps/curve": {"value": "L_triceps"},
               "Soleus/curve": {"value": "R_soleus"},
               "Empty/curve": {},
               "NoValue/curve": {"value": ""},
               "Ghost/curve": {"value": "L_notInAsset"},
               "MissingTrack/curve": {"value": "R_soleus"}}
    tracks = {k: rng.random(nf).astype(np.float32) for k in entries
              if k not in ("Empty/curve", "NoValue/curve", "MissingTrack/curve")}
    mid = rng.integers(0, len(strings), 400)
    src = MappedActivationSource(_mapping(tmp_path, entries), tracks, mid, strings, start, end, aggregate=agg)
    assert src.n_unmapped == 2 and src.n_missing == 1 and src.n_read == 6
    assert sorted(src.driven) == ["L_biceps", "L_triceps", "R_soleus"]
    assert src.absent == ["L_notInAsset"]
    by_muscle = {}
    for k, m in entries.items():
        if m.get("value") and k in tracks:
            by_muscle.setdefault(m["value"], []).append(tracks[k])
    for frame in (start - 5, start, 7, end, end + 3):
        fi = min(max(frame - start, 0), nf - 1)
        got = src.per_tet(frame)
        assert got.dtype == np.float32 and got.shape == (400,)
        for t in range(400):
            ls = by_muscle.get(strings[mid[t]])
            if ls is None:
                want = 0.0
            else:
                vals = [float(x[fi]) for x in ls]
                want = max(vals) if agg == "max" else sum(vals) / len(vals)
            assert got[t] == pytest.approx(want, abs=1e-7)


def test_mapped_activation_errors(tmp_path):
    p = _mapping(tmp_path, {"A/curve": {"value": "m"}})
    with pytest.raises(FileNotFoundError):
        MappedActivationSource(tmp_path / "none.json", {}, [0], ["m"], 0, 3)
    with pytest.raises(ValueError):
        MappedActivationSource(p, {"A/curve": np.zeros(3)}, [0], ["m"], 0, 3)   # 4 frames expected
    with pytest.raises(ValueError):
        MappedActivationSource(p, {}, [0], ["m"], 0, 3, aggregate="sum")
    with pytest.raises(ImportError):
        MappedActivationSource.from_usd("walk.usd", p, [0], ["m"], 0, 3)


REF_MAPPING = Path("/root/reference/data/humanV2/muscle_mapping.json")


@pytest.mark.skipif(not REF_MAPPING.exists(), reason="the humanV2 muscle mapping is not present")
def test_reference_mapping_schema(tmp_path):
    """The reference's humanV2 muscle_mapping.json: 284 curve entries, 6 authored
    empty, many curves per muscle; every mapped muscle gets the mean of its curves."""
    m = json.loads(REF_MAPPING.read_text())["mappings"]
    names = sorted({v["value"] for v in m.values() if isinstance(v, dict) and v.get("value")})
    tracks = {k: np.full(2, float(i % 7) / 7, np.float32) for i, k in enumerate(m)}
    src = MappedActivationSource(REF_MAPPING, tracks, np.arange(len(names)), names, 1, 2)
    assert src.n_unmapped == 6 and src.n_read == 284 - 6 and src.n_missing == 0
    assert sorted(src.driven) == names and not src.absent
    got = src.per_tet(1)
    for i, name in enumerate(names):
        vals = [tracks[k][0] for k, v in m.items() if isinstance(v, dict) and v.get("value") == name]
        assert got[i] == pytest.approx(np.mean(vals), abs=1e-7)
