"""JSON scene config, same schema as the reference.

``load_config`` follows src/muscle/mg/mg.py:1437-1443 (json -> dict, relative
asset paths made absolute against cwd); ``find_constraint`` follows
mg.py:71-82 (constraints[] looked up by ``name``, entries with
``enable: false`` ignored).  A ``synthetic`` block (new) names a generated
workload for asset-free runs.
"""
from __future__ import annotations

import json
from pathlib import Path
from types import SimpleNamespace


def load_config(config_path: str) -> dict:
    with open(config_path) as f:
        cfg = json.load(f)
    for key in ("geo_dir", "cons_dir", "rest_path", "fascia_path"):
        if key in cfg and not Path(cfg[key]).is_absolute():
            cfg[key] = str(Path.cwd() / cfg[key])
    return cfg


def find_constraint(cfg: dict, name: str):
    """First enabled constraints[] entry with this name, else None."""
    for c in cfg.get("constraints", []) or []:
        if c.get("name") == name and c.get("enable", True):
            return c
    return None


def find_constraint_by_type(cfg, ctype: str):
    """MuscleSim-style lookup by ``type`` (src/muscle/constraint/builder.py:618-648)."""
    cons = cfg.get("constraints", []) if isinstance(cfg, dict) else getattr(cfg, "constraints", [])
    for c in cons or []:
        if str(c.get("type", "")).lower() == ctype and c.get("enable", True):
            return c
    return None


def as_namespace(cfg: dict) -> SimpleNamespace:
    """SimConfig-like attribute view of a dict (src/muscle/util/util.py:54-74)."""
    return SimpleNamespace(**cfg)
