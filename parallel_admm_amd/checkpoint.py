"""Checkpoint and resume of a backend's state.

The reference has none: `reset` re-seeds x and zeroes v, Ue and Uc
(src/solver/backend_cpu/admm_impl_cpu.cpp:71-88), and its OBJ/PLY export
keeps positions only.  A checkpoint here is an .npz (arrays only, loaded with
allow_pickle=False) holding x and v (n_all x 3), Ue (m x 3), Uc (n_dyn x 3),
the rigid tracks' frame counters, the step index and a hash of the solver
configuration and the scene's sizes.  The values are the device's FP32 state
(exact in FP64), and nothing else is carried from one step to the next (the
contact duals restart at every detection), so a resumed run is bit-identical
to an uninterrupted one.  With a warm start of the contact duals enabled, the
first detection after a resume has no previous window to take duals from.
"""
from __future__ import annotations

import dataclasses
import hashlib
import json

import numpy as np


def config_hash(cfg, sizes: dict) -> str:
    """SHA-256 of the solver configuration and the scene sizes."""
    d = {f.name: getattr(cfg, f.name) for f in dataclasses.fields(cfg)}
    d.update({k: int(sizes[k]) for k in ("n_all", "n_dyn", "m_rows")})
    return hashlib.sha256(json.dumps(d, sort_keys=True, default=float).encode()).hexdigest()


def save(backend, path: str, step_idx: int, cfg) -> None:
    x, v = backend.get_state()
    ue, uc = backend.get_duals()
    np.savez(path, x=x, v=v, ue=ue, uc=uc, tracks=backend.track_frames().astype(np.int32), step=np.int64(step_idx),
             config=np.array(config_hash(cfg, backend.sizes())))


def load(backend, path: str, cfg) -> int:
    """Restores the state saved by `save` into a backend built for the same
    scene and configuration (tracks registered as before); returns the step
    index.  Raises ValueError on a configuration or size mismatch."""
    with np.load(path, allow_pickle=False) as z:
        if str(z["config"]) != config_hash(cfg, backend.sizes()):
            raise ValueError(f"{path}: checkpoint was written for another configuration or scene")
        backend.set_state(z["x"], z["v"])
        backend.set_duals(z["ue"], z["uc"])
        backend.set_track_frames(z["tracks"])
        return int(z["step"])
