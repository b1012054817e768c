"""Per-tet muscle activation sources (src/muscle/util/activation.py:1-144).

A source exposes ``per_tet(frame) -> (NT,) float32`` for ``set_fiber_activation``.

* ``RampActivationSource``: a fixed per-tet mask ramped 0 -> 1 over the frame range
  (activation.py:28-43).
* ``MappedActivationSource``: per-curve activation tracks mapped onto muscles by
  ``muscle_mapping.json`` (``{"mappings": {"<Name>/curve": {"value": "<muscle_id>"}}}``,
  many curves to one muscle, aggregated by mean or max, entries without a value
  skipped, muscles without a curve passive) and scattered to the tets by their
  ``muscle_id`` (UsdActivationSource, activation.py:46-144).  The reference reads the
  tracks from a USD curve asset (``primvars:muscle:activation`` on
  ``/Human/Muscles/<Name>/curve``); ``from_usd`` needs the ``pxr`` module, which is not
  installed here, so tracks can also be given directly as ``{key: (n_frames,) array}``.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np


class RampActivationSource:
    def __init__(self, mask, start_frame, end_frame):
        self._mask = np.asarray(mask, dtype=np.float32)
        self.start_frame = int(start_frame)
        self.end_frame = int(end_frame)
        self._span = max(1, self.end_frame - self.start_frame)

    def per_tet(self, frame: int) -> np.ndarray:
        a = min(1.0, max(0.0, (int(frame) - self.start_frame) / self._span))
        return (a * self._mask).astype(np.float32)


class MappedActivationSource:
    """tracks: {"<Name>/curve": per-frame values for frames start..end}"""

    def __init__(self, mapping_path, tracks: dict, muscle_id_np, muscle_id_strings, start_frame, end_frame,
                 aggregate: str = "mean"):
        self.start_frame = int(start_frame)
        self.end_frame = int(end_frame)
        self._n_frames = self.end_frame - self.start_frame + 1
        self._muscle_id = np.asarray(muscle_id_np, np.int64)
        self.aggregate = str(aggregate).lower()
        if self.aggregate not in ("mean", "max"):
            raise ValueError(f"aggregate={aggregate!r}: 'mean' or 'max'")
        if not Path(mapping_path).is_file():
            raise FileNotFoundError(f"muscle mapping not found: {mapping_path}")
        mapping = json.loads(Path(mapping_path).read_text()).get("mappings", {})
        lines: dict[str, list] = {}
        self.n_unmapped = self.n_missing = self.n_read = 0
        for key, m in mapping.items():
            target = m.get("value") if isinstance(m, dict) else None
            if not target:
                self.n_unmapped += 1
                continue
            if key not in tracks:
                self.n_missing += 1
                continue
            arr = np.asarray(tracks[key], np.float32).reshape(-1)
            if arr.shape[0] != self._n_frames:
                raise ValueError(f"track {key!r} has {arr.shape[0]} frames, expected {self._n_frames}")
            lines.setdefault(target, []).append(arr)
            self.n_read += 1
        index = {s: i for i, s in enumerate(muscle_id_strings)}
        self._table = np.zeros((len(muscle_id_strings), self._n_frames), np.float32)
        self.driven, self.absent = [], []
        for name, ls in lines.items():
            st = np.stack(ls, 0)
            merged = st.max(0) if self.aggregate == "max" else st.mean(0)
            i = index.get(name)
            if i is None:
                self.absent.append(name)
                continue
            self._table[i] = merged.astype(np.float32)
            self.driven.append(name)

    @classmethod
    def from_usd(cls, usd_path, mapping_path, muscle_id_np, muscle_id_strings, start_frame, end_frame,
                 muscles_root="/Human/Muscles", aggregate="mean", activation_attr="primvars:muscle:activation"):
        try:
            from pxr import Usd
        except ImportError as e:
            raise ImportError("reading activation curves from USD needs the pxr module (OpenUSD), which is not "
                              "installed; pass the tracks as arrays to MappedActivationSource") from e
        stage = Usd.Stage.Open(str(usd_path))
        if stage is None:
            raise FileNotFoundError(f"cannot open activation USD: {usd_path}")
        mapping = json.loads(Path(mapping_path).read_text()).get("mappings", {})
        tracks = {}
        for key in mapping:
            prim = stage.GetPrimAtPath(f"{muscles_root}/{key}")
            attr = prim.GetAttribute(activation_attr) if prim and prim.IsValid() else None
            if attr is None or not attr.HasValue():
                continue
            vals = [attr.Get(float(f)) for f in range(int(start_frame), int(end_frame) + 1)]
            tracks[key] = np.asarray([0.0 if v is None else float(v) for v in vals], np.float32)
        return cls(mapping_path, tracks, muscle_id_np, muscle_id_strings, start_frame, end_frame, aggregate)

    def per_tet(self, frame: int) -> np.ndarray:
        fi = max(0, min(self._n_frames - 1, int(frame) - self.start_frame))
        return self._table[:, fi][self._muscle_id].astype(np.float32)
