"""Learned warm start of the contact duals: data, training and evaluation.

The reference restarts every contact dual from zero at each detection
(src/solver/contact_solver.cu:243; CPU src/contact/contact_info.h:46).  Here
the backend can seed them (pa_backend_set_warm_start):
  mode 1  a small MLP (10 -> 128 -> 128 -> 128 -> 2, SiLU, BF16 MFMA) maps
          per-contact features to the dual;
  mode 2  the previous window's dual of the same contact (no network);
with plain cold start as the fallback the backend takes when a seed does
worse than zero or the ADMM residual stalls.

Pipeline (all on the GPU, through the C ABI):
  collect_samples  run plain steps of randomised batch-member scenes (side
                   16..96 cloths dropped on a sphere) with the contact log on;
  fit              normalise, then Adam over shuffled batches in the
                   library's own forward/backward kernels;
  evaluate         steps/s and ADMM iterations to the residual tolerance,
                   plain vs learned (and vs mode 2).
`python -m parallel_admm_amd.warm_start train --out PATH` writes a model file
(.npz, arrays only).
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import time

import numpy as np

from . import ADMMBackendHIP, MLP, MLP_PARAMS, Scene

FEATURES = ("gap", "u0_n", "u0_t", "vrel_n", "vrel_t", "mu", "type", "has_prev", "prev_n", "prev_t")
DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warm_start_mlp.npz")


@dataclasses.dataclass
class Model:
    params: np.ndarray             # MLP_PARAMS float32, flat layout of pa_mlp_get_params
    in_shift: np.ndarray           # 10
    in_scale: np.ndarray           # 10
    out_scale: np.ndarray          # 2
    meta: dict = dataclasses.field(default_factory=dict)

    def save(self, path: str) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, params=self.params.astype(np.float32), in_shift=self.in_shift.astype(np.float32),
                 in_scale=self.in_scale.astype(np.float32), out_scale=self.out_scale.astype(np.float32),
                 meta=np.array(json.dumps(self.meta)))

    @classmethod
    def load(cls, path: str = DEFAULT_MODEL) -> "Model":
        with np.load(path, allow_pickle=False) as z:
            m = cls(z["params"].astype(np.float32), z["in_shift"].astype(np.float32), z["in_scale"].astype(np.float32),
                    z["out_scale"].astype(np.float32), json.loads(str(z["meta"])))
        if m.params.size != MLP_PARAMS:
            raise ValueError(f"{path}: expected {MLP_PARAMS} parameters")
        return m


def apply(backend: ADMMBackendHIP, model: Model | None, mode: int = 1, **kw) -> None:
    """Configures the backend's warm start (mode 0 turns it off)."""
    if mode == 1 and model is None:
        raise ValueError("mode 1 needs a model")
    if mode == 1:
        backend.set_warm_start(1, params=model.params, in_shift=model.in_shift, in_scale=model.in_scale,
                               out_scale=model.out_scale, **kw)
    else:
        backend.set_warm_start(mode, **kw)


def batch_member_scenes(count: int, seed: int = 0):
    """Randomised batch members: cloth side ~ U{16..96}, drop ~ U(0.2, 0.6) m,
    sphere radius ~ U(0.1, 0.3) m, mu ~ U(0.1, 0.6).  Yields (params, Scene)."""
    rng = np.random.default_rng(seed)
    for _ in range(count):
        side = int(rng.integers(16, 97))
        drop = float(rng.uniform(0.2, 0.6))
        radius = float(rng.uniform(0.1, 0.3))
        mu = float(rng.uniform(0.1, 0.6))
        yield dict(side=side, drop=drop, radius=radius, mu=mu), Scene.generate(4, side=side, radius=radius, drop=drop)


def _batch_cfg(mu: float):
    from .configs import config_for

    cfg = config_for("batch")
    cfg.mu = mu
    return cfg


def collect_samples(n_samples: int, seed: int = 0, steps: int = 90, max_scenes: int = 200, log=print) -> np.ndarray:
    """Contact-log rows (12 floats: 10 raw features, 2 targets) from plain runs."""
    chunks, total = [], 0
    for k, (par, sc) in enumerate(batch_member_scenes(max_scenes, seed)):
        cfg = _batch_cfg(par["mu"])
        b = ADMMBackendHIP.from_scene(sc, cfg)
        b.set_sync_to_host(False)
        cap = 1 << 22
        b.set_contact_log(cap)
        t0 = time.time()
        for _ in range(steps):
            b.step()
        rows = b.read_contact_log(cap)
        chunks.append(rows)
        total += rows.shape[0]
        log(f"[warm_start] scene {k}: side {par['side']}, drop {par['drop']:.2f}, r {par['radius']:.2f}, mu {par['mu']:.2f}: "
            f"{rows.shape[0]} rows in {time.time() - t0:.1f} s (total {total})")
        if total >= n_samples:
            break
    data = np.concatenate(chunks)[:n_samples]
    return data[np.all(np.isfinite(data), axis=1)]


def normalisation(rows: np.ndarray):
    f = rows[:, :10].astype(np.float64)
    t = rows[:, 10:12].astype(np.float64)
    shift = f.mean(axis=0)
    std = f.std(axis=0)
    scale = np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-6), 0.0)
    out = np.maximum(t.std(axis=0), 1e-6)
    return shift.astype(np.float32), scale.astype(np.float32), out.astype(np.float32)


def fit(rows: np.ndarray, epochs: int = 20, batch: int = 8192, lr: float = 1e-3, seed: int = 0, log=print) -> Model:
    shift, scale, out = normalisation(rows)
    x = (rows[:, :10] - shift) * scale
    t = rows[:, 10:12] / out
    samples = np.concatenate([x, t], axis=1).astype(np.float32)
    net = MLP(seed=seed)
    t0 = time.time()
    losses = net.train(samples, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    dt = time.time() - t0
    log(f"[warm_start] trained {epochs} epochs on {rows.shape[0]} samples in {dt:.1f} s; "
        f"loss {losses[0]:.4f} -> {losses[-1]:.4f} (normalised targets)")
    return Model(net.params(), shift, scale, out, dict(samples=int(rows.shape[0]), epochs=epochs, batch=batch, lr=lr,
                                                        seed=seed, losses=[float(v) for v in losses], train_seconds=dt))


def run(scene: Scene, cfg, steps: int, mode: int = 0, model: Model | None = None, tracks=(), warmup: int = 0):
    """Steps a fresh backend; returns per-step stats and the wall time of the
    timed steps (after `warmup` untimed ones)."""
    b = ADMMBackendHIP.from_scene(scene, cfg)
    b.set_sync_to_host(False)
    for first, count, pos, quat in tracks:
        b.add_rigid_track(first, count, pos, quat)
    if mode:
        apply(b, model, mode)
    for _ in range(warmup):
        b.step()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(b.step())
    wall = time.perf_counter() - t0
    return stats, wall, b


def summary(stats, wall) -> dict:
    it = np.array([s["iterations"] for s in stats], float)
    wasted = np.array([s["ws_wasted_iters"] for s in stats], float)
    res = np.array([max(s["r_primal"], s["r_dual"]) for s in stats], float)
    return dict(steps_per_s=len(stats) / wall, admm_iters_mean=float(it.mean()), admm_iters_incl_wasted=float((it + wasted).mean()),
                converged=int(sum(s["converged"] for s in stats)), final_residual_mean=float(res.mean()),
                fallbacks=int(sum(s["ws_fallback"] for s in stats)), seeded=int(sum(s["ws_seeded"] for s in stats)),
                bad_windows=int(sum(s["ws_bad_windows"] for s in stats)))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--out", required=True)
    tr.add_argument("--samples", type=int, default=2_000_000)
    tr.add_argument("--epochs", type=int, default=20)
    tr.add_argument("--batch", type=int, default=8192)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--seed", type=int, default=0)
    tr.add_argument("--steps", type=int, default=90)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        t0 = time.time()
        rows = collect_samples(a.samples, seed=a.seed, steps=a.steps)
        print(f"[warm_start] {rows.shape[0]} samples in {time.time() - t0:.1f} s", flush=True)
        model = fit(rows, a.epochs, a.batch, a.lr, a.seed)
        model.meta.update(data_seed=a.seed, steps_per_scene=a.steps, source="batch-member scenes (kind 4), plain cold start")
        model.save(a.out)
        print(f"[warm_start] wrote {a.out}", flush=True)


if __name__ == "__main__":
    main()
