"""Houdini JSON ``.geo`` files: the per-frame input format of the humanV2 assets.

``GeoFile`` reads what the solver consumes, with the names and return types of the
reference reader (src/muscle/util/geo_io.py:42-216): point positions ``P``, tet (or
triangle) vertex indices from the topology, ``gluetoanimation``, the constraint
``pts`` and ``target_pos`` of ``cons_<frame>.geo``, any named point attribute,
point groups (ordered lists or ``boolRLE`` runs) and primitive attributes (numeric,
or string tables such as ``muscle_id``).  ``GeoWriter`` rewrites ``P`` (and the
optional ``is_collide`` / ``pen_depth`` point attributes) of a template file in
place (geo_io.py:266-313).  ``read_geo_attr_cached`` keeps one extracted attribute
next to the source as ``_npy_cache/<stem>_<attr>.npy``, re-parsing when the source
is newer, written atomically (geo_io.py:316-338).

The layout, as these functions read and write it: the file is a flat pair list
``[key0, value0, key1, value1, ...]``; ``attributes`` holds ``pointattributes`` and
``primitiveattributes``, each a list of ``[metadata, data]`` pair lists; a numeric
attribute's payload is the sixth element of its ``values`` pair list (the tuples of a
vector attribute, or a one-row ``arrays`` list of a scalar one); ``topology`` ->
``pointref`` -> ``indices`` is the flat vertex list of the primitives, and a
``Tetrahedron_run`` primitive block gives the tet count.

``write_geo`` writes a new file in that layout (tests and asset conversion).
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np


def _pairs(lst) -> dict:
    return dict(zip(lst[0::2], lst[1::2]))


def _bool_rle(unordered) -> np.ndarray:
    """['boolRLE', [count, flag, count, flag, ...]] -> the selected indices"""
    if not unordered or unordered[0] != "boolRLE":
        return np.zeros(0, np.int32)
    runs = unordered[1]
    out, pos = [], 0
    for k in range(0, len(runs), 2):
        n = int(runs[k])
        if bool(runs[k + 1]):
            out.append(np.arange(pos, pos + n, dtype=np.int32))
        pos += n
    return np.concatenate(out) if out else np.zeros(0, np.int32)


def _payload(data: dict):
    vals = data["values"]
    return vals[5]


class GeoFile:
    """The fields of a ``.geo`` the solver reads (geo_io.py:42-216)."""

    def __init__(self, path):
        with open(path, "r") as fp:
            root = _pairs(json.load(fp))
        attrs = _pairs(root.get("attributes", []))
        self._attrs: dict[str, np.ndarray] = {}
        self._positions = self._glue = self._pts = self._target = None
        for entry in attrs.get("pointattributes", []):
            meta, data = _pairs(entry[0]), _pairs(entry[1])
            name = meta.get("name")
            if "values" not in data:
                continue
            raw = _payload(data)
            if name == "P":
                self._positions = np.asarray(raw, np.float32)
            elif name == "target_pos":
                self._target = np.asarray(raw, np.float32)
            elif name == "gluetoanimation":
                self._glue = np.asarray(raw[0], np.bool_)
            elif name == "pts":
                self._pts = np.asarray(raw[0], np.int32)
            arr = np.asarray(raw)
            if arr.ndim == 2 and arr.shape[0] == 1:        # a scalar attribute, stored as one row
                arr = arr[0]
            self._attrs[name] = arr
        self._prim: dict[str, np.ndarray] = {}
        self._prim_strings: dict[str, list] = {}
        for entry in attrs.get("primitiveattributes", []):
            meta, data = _pairs(entry[0]), _pairs(entry[1])
            name = meta.get("name")
            if "strings" in data and "indices" in data:
                arr = _pairs(data["indices"]).get("arrays")
                if arr is not None:
                    self._prim[name] = np.asarray(arr).ravel().astype(np.int32)
                    self._prim_strings[name] = list(data["strings"])
            elif "values" in data:
                arr = np.asarray(_payload(data))
                if arr.ndim == 2 and arr.shape[0] == 1:
                    arr = arr[0]
                self._prim[name] = arr
        topo = _pairs(root.get("topology", []))
        idx = np.asarray(_pairs(topo.get("pointref", [])).get("indices", []), np.int32)
        ntet = 0
        for blk in root.get("primitives", []):
            if _pairs(blk[0]).get("type") == "Tetrahedron_run":
                ntet = int(_pairs(blk[1])["nprimitives"])
                break
        if ntet > 0:
            self._tets = idx[:4 * ntet].reshape(ntet, 4)
        else:
            nprim = int(root.get("primitivecount", 0))
            self._tets = idx.reshape(nprim, idx.size // nprim) if nprim > 0 and idx.size else idx.reshape(0, 0)
        self._groups: dict[str, np.ndarray] = {}
        for grp in root.get("pointgroups", []):
            meta = _pairs(grp[0])
            sel = _pairs(_pairs(grp[1])["selection"])
            if "ordered" in sel:
                self._groups[meta["name"]] = np.asarray(sel["ordered"], np.int32)
            elif "unordered" in sel:
                self._groups[meta["name"]] = _bool_rle(sel["unordered"])

    @property
    def positions(self) -> np.ndarray:
        if self._positions is None:
            raise KeyError("attribute 'P' not present in this .geo")
        return self._positions

    @property
    def tet_indices(self) -> np.ndarray:
        return self._tets

    @property
    def gluetoanimation(self) -> np.ndarray:
        if self._glue is None:       # absent: all False (geo_io.py:165-171)
            self._glue = np.zeros(0 if self._positions is None else len(self._positions), np.bool_)
        return self._glue

    @property
    def pts(self) -> np.ndarray:
        if self._pts is None:
            raise KeyError("attribute 'pts' not present in this .geo")
        return self._pts

    @property
    def target_pos(self) -> np.ndarray:
        if self._target is None:
            raise KeyError("attribute 'target_pos' not present in this .geo")
        return self._target

    def point_attr(self, name: str) -> np.ndarray:
        if name not in self._attrs:
            raise KeyError(f"point attribute '{name}' not in {sorted(self._attrs)}")
        return self._attrs[name]

    def has_point_attr(self, name: str) -> bool:
        return name in self._attrs

    def point_group(self, name: str) -> np.ndarray:
        if name not in self._groups:
            raise KeyError(f"point group '{name}' not in {sorted(self._groups)}")
        return self._groups[name]

    def has_point_group(self, name: str) -> bool:
        return name in self._groups

    def has_prim_attr(self, name: str) -> bool:
        return name in self._prim

    def prim_attr(self, name: str) -> np.ndarray:
        if name not in self._prim:
            raise KeyError(f"prim attribute '{name}' not in {sorted(self._prim)}")
        return self._prim[name]

    def prim_attr_strings(self, name: str) -> list:
        return self._prim_strings.get(name, [])


# ---------------------------------------------------------------- writing
def _attr_entry(name, arr):
    arr = np.asarray(arr)
    if arr.dtype.kind in "iub":
        storage, conv = "int32", int
    else:
        storage, conv = "fpreal32", float
    meta = ["scope", "public", "type", "numeric", "name", name, "options", {}]
    if arr.ndim == 2:
        size = int(arr.shape[1])
        vals = ["size", size, "storage", storage, "tuples", [[conv(x) for x in row] for row in arr]]
    else:
        size = 1
        vals = ["size", 1, "storage", storage, "arrays", [[conv(x) for x in arr]]]
    data = ["size", size, "storage", storage, "defaults", ["size", 1, "storage", "fpreal64", "values", [0]],
            "values", vals]
    return [meta, data]


def _string_entry(name, index, strings):
    meta = ["scope", "public", "type", "string", "name", name, "options", {}]
    data = ["size", 1, "storage", "int32", "strings", list(strings),
            "indices", ["size", 1, "storage", "int32", "arrays", [[int(i) for i in index]]]]
    return [meta, data]


def _rle(indices, n):
    sel = np.zeros(n, bool)
    sel[np.asarray(indices, np.int64)] = True
    runs, k = [], 0
    while k < n:
        j = k
        while j < n and sel[j] == sel[k]:
            j += 1
        runs += [j - k, bool(sel[k])]
        k = j
    return ["boolRLE", runs]


def write_geo(path, positions, prims=None, point_attrs=None, prim_attrs=None, prim_strings=None,
              groups=None, rle_groups=True):
    """A new ``.geo`` in the layout GeoFile reads: P plus point attributes, the
    primitives (n, 4) as a Tetrahedron_run or (n, k) polygons, numeric primitive
    attributes, string primitive attributes {name: (index per prim, strings)} and
    point groups (boolRLE runs or ordered lists)."""
    pos = np.asarray(positions, np.float32)
    n = len(pos)
    pattrs = [_attr_entry("P", pos)]
    for name, arr in (point_attrs or {}).items():
        pattrs.append(_attr_entry(name, arr))
    prims = np.zeros((0, 4), np.int32) if prims is None else np.asarray(prims, np.int32)
    qattrs = [_attr_entry(name, arr) for name, arr in (prim_attrs or {}).items()]
    qattrs += [_string_entry(name, idx, strs) for name, (idx, strs) in (prim_strings or {}).items()]
    if prims.shape[1] == 4:
        blocks = [[["type", "Tetrahedron_run"], ["startvertex", 0, "nprimitives", int(len(prims))]]]
    else:
        blocks = [[["type", "Polygon_run"], ["startvertex", 0, "nprimitives", int(len(prims)),
                                            "nvertices", [int(prims.shape[1])] * int(len(prims))]]]
    pgroups = []
    for name, idx in (groups or {}).items():
        sel = (["unordered", _rle(idx, n)] if rle_groups
               else ["ordered", [int(i) for i in np.asarray(idx)]])
        pgroups.append([["name", name], ["selection", sel]])
    root = ["fileversion", "19.5", "pointcount", n, "vertexcount", int(prims.size), "primitivecount",
            int(len(prims)), "topology", ["pointref", ["indices", [int(i) for i in prims.reshape(-1)]]],
            "attributes", ["pointattributes", pattrs, "primitiveattributes", qattrs],
            "primitives", blocks, "pointgroups", pgroups]
    with open(path, "w") as fp:
        json.dump(root, fp)


class GeoWriter:
    """In-place rewrite of a template's ``P`` (+ optional int ``is_collide`` and float
    ``pen_depth`` point attributes, created when missing); everything else of the
    template is kept (geo_io.py:266-313)."""

    def __init__(self, template_path):
        with open(template_path, "r") as fp:
            self._root = json.load(fp)
        root = _pairs(self._root)
        self._npoints = int(root["pointcount"])
        self._pattrs = _pairs(root["attributes"])["pointattributes"]
        if self._find("P") is None:
            raise KeyError("template .geo has no 'P' attribute")

    def _find(self, name):
        for e in self._pattrs:
            if _pairs(e[0]).get("name") == name:
                return e
        return None

    def _put(self, name, arr):
        entry = self._find(name)
        if entry is None:
            self._pattrs.append(_attr_entry(name, arr))
            return
        vals = _pairs(entry[1])["values"]
        key = "tuples" if np.asarray(arr).ndim == 2 else "arrays"
        data = (np.asarray(arr).tolist() if key == "tuples" else [np.asarray(arr).tolist()])
        for i in range(0, len(vals), 2):
            if vals[i] in ("tuples", "arrays"):
                vals[i], vals[i + 1] = key, data
                return
        vals += [key, data]

    def write(self, out_path, positions, is_collide=None, pen_depth=None):
        pos = np.asarray(positions, np.float32)
        if pos.shape[0] != self._npoints:
            raise ValueError(f"positions has {pos.shape[0]} points, template has {self._npoints}")
        self._put("P", pos)
        if is_collide is not None:
            self._put("is_collide", np.asarray(is_collide).astype(np.int32))
        if pen_depth is not None:
            self._put("pen_depth", np.asarray(pen_depth).astype(np.float32))
        with open(out_path, "w") as fp:
            json.dump(self._root, fp)


def read_geo_attr_cached(geo_path, attr: str) -> np.ndarray:
    """One point attribute (``P`` or ``target_pos``), cached as ``.npy`` next to the
    source and re-parsed when the source is newer (geo_io.py:316-338)."""
    geo = Path(geo_path)
    cache = geo.parent / "_npy_cache" / f"{geo.stem}_{attr}.npy"
    if cache.exists() and cache.stat().st_mtime >= geo.stat().st_mtime:
        return np.load(cache)
    g = GeoFile(str(geo))
    arr = g.positions if attr == "P" else getattr(g, attr)
    cache.parent.mkdir(exist_ok=True)
    tmp = cache.with_suffix(".tmp.npy")          # np.save keeps a .npy suffix as given
    np.save(tmp, arr)
    os.replace(tmp, cache)
    return arr


def read_cons_frame(geo_dir, frame: int) -> np.ndarray:
    """target_pos (NCONS, 3) of ``cons_<frame>.geo``, cached"""
    return read_geo_attr_cached(str(Path(geo_dir) / f"cons_{frame}.geo"), "target_pos")


class GeoBoneSource:
    """A directory of ``bone_<frame>.geo`` files (bone_source.py:95-126): triangles
    from the first frame's topology as authored, positions per frame (cached)."""

    def __init__(self, bone_dir, start_frame: int = 1, scale: float = 1.0):
        self.bone_dir = Path(bone_dir)
        self.start_frame = int(start_frame)
        self._scale = float(scale)
        g0 = GeoFile(self._path(self.start_frame))
        self.tris = np.ascontiguousarray(np.asarray(g0.tet_indices, np.int32).reshape(-1))
        self.n_verts = int(np.asarray(g0.positions).shape[0])

    def _path(self, frame: int) -> str:
        return str(self.bone_dir / f"bone_{frame}.geo")

    def positions(self, frame: int) -> np.ndarray:
        p = np.ascontiguousarray(np.asarray(read_geo_attr_cached(self._path(frame), "P"), np.float32))
        return p if self._scale == 1.0 else (p * np.float32(self._scale)).astype(np.float32)
