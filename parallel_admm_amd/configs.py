"""Solver settings for the benchmark scenes (scene geometry lives in
pa_scene.cpp).  Unspecified keys take ADMMSolverConfig defaults
(src/solver/core/admm_config.h:20-53); w_scale is this build's recorded
choice (the reference scenes use 20..2000).

  cloth32   32x32 cloth over an icosphere: FP64 CPU-runnable case
  tetbox    20^3-cell 5-tet box dropped on a ground quad
  cloth256  256x256 self-contact cloth over a sphere and a cylinder
"""
from __future__ import annotations

from . import AdmmConfig

SCENE_KIND = {"cloth32": 1, "tetbox": 2, "cloth256": 3, "batch": 4}


def config_for(name: str) -> AdmmConfig:
    if name == "cloth32":
        return AdmmConfig(dt=1.0 / 60, g=9.81, admm_max_iter=20, dcd_interval=5, gs_max_iter=10, mu=0.3,
                          kappa=1e6, beta=1e5, thickness=1e-3, damp_kl=1e-4, damp_km=0.0, w_scale=50.0,
                          max_contacts=1 << 16)
    if name == "tetbox":
        return AdmmConfig(dt=1.0 / 100, g=9.81, admm_max_iter=30, dcd_interval=5, gs_max_iter=10, mu=0.5,
                          kappa=1e6, beta=1e5, thickness=2e-3, damp_kl=1e-4, damp_km=0.0, w_scale=50.0,
                          max_contacts=1 << 18)
    if name == "cloth256":
        return AdmmConfig(dt=1.0 / 120, g=9.81, admm_max_iter=50, dcd_interval=5, gs_max_iter=10, mu=0.4,
                          kappa=1e6, beta=1e5, thickness=5e-4, damp_kl=1e-4, damp_km=0.0, w_scale=50.0,
                          max_contacts=1 << 21)
    if name == "batch":
        return AdmmConfig(dt=1.0 / 60, g=9.81, admm_max_iter=30, dcd_interval=5, gs_max_iter=10, mu=0.3,
                          kappa=1e6, beta=1e5, thickness=1e-3, damp_kl=1e-4, damp_km=0.0, w_scale=50.0,
                          max_contacts=1 << 16)
    raise KeyError(name)
