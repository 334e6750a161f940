"""Train the learned beam scorer on the GPU.

    python -m synchrowords_amd.train_scorer --n 300 --k 2 --dfas 20000 \
        --out synchrowords_amd/models/scorer_n300_k2.npz

Samples come from the exact search itself: for each seeded random DFA
(seeds from --seed0 on) a cardinality beam proposes candidate sets, --per-depth
of them per beam depth, and the search's forward levels label each with its
forward distance (sw_mlp_samples).  Features are built on the device, and the
network (in_dim -> H -> H -> 1, BF16 on MFMA with FP32 accumulation) trains
with Adam on the device.  The default loss ranks the candidates of one beam
depth by their distance (pairwise, within each group): the beam only ever
compares siblings, and distances beyond the labelled depth are censored, so
a ranking is what the labels support.  Several host threads each drive a
context, so sampling overlaps.  Writes the FP32 master parameters and the
training record.
"""
from __future__ import annotations

import argparse
import json
import threading
import time
from pathlib import Path

import numpy as np

import synchrowords_amd as sw


def collect(n, k, seeds, fdepth, width, bdepth, per_depth, threads, log_every=2000, raw=None):
    xs, ts = [], []
    lock = threading.Lock()
    it = iter(seeds)
    done = [0]
    t0 = time.perf_counter()

    def work():
        ctx = sw.Context(0)
        feat = sw.Mlp(ctx, n, k, hidden=16, seed=0)  # feature builder only
        try:
            while True:
                with lock:
                    s = next(it, None)
                if s is None:
                    return
                d = sw.random_dfa(n, k, s)
                sets, dist = sw.mlp_samples(ctx, n, k, d, fdepth=fdepth, width=width, bdepth=bdepth,
                                            per_depth=per_depth, cap=1 << 14)
                if len(sets) == 0:
                    continue
                x = feat.features(d, sets)
                with lock:
                    xs.append(x)
                    ts.append(dist)
                    if raw is not None:
                        raw.append((s, np.asarray(sets), np.asarray(dist)))
                    done[0] += 1
                    if done[0] % log_every == 0:
                        print(f"[train] sampled {done[0]} automata, {sum(len(t) for t in ts)} samples, "
                              f"{time.perf_counter() - t0:.1f}s", flush=True)
        finally:
            feat.close()
            ctx.close()

    th = [threading.Thread(target=work) for _ in range(threads)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    return np.concatenate(xs), np.concatenate(ts)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=300)
    ap.add_argument("--k", type=int, default=2)
    ap.add_argument("--dfas", type=int, default=20000)
    ap.add_argument("--seed0", type=int, default=100000)
    ap.add_argument("--fdepth", type=int, default=20)
    ap.add_argument("--width", type=int, default=256)
    ap.add_argument("--bdepth", type=int, default=60)
    ap.add_argument("--per-depth", type=int, default=8)
    ap.add_argument("--hidden", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--loss", choices=["rank", "mse"], default="rank")
    ap.add_argument("--out", type=str, required=True)
    ap.add_argument("--dump", type=str, default="", help="also save the raw labelled sets (seed, set, distance)")
    a = ap.parse_args()

    t0 = time.perf_counter()
    raw = [] if a.dump else None
    x, t = collect(a.n, a.k, range(a.seed0, a.seed0 + a.dfas), a.fdepth, a.width, a.bdepth, a.per_depth,
                   a.threads, raw=raw)
    if raw:
        np.savez_compressed(a.dump, seed=np.concatenate([np.full(len(d), s, dtype=np.int64) for s, _, d in raw]),
                            sets=np.concatenate([r for _, r, _ in raw]), dist=np.concatenate([d for _, _, d in raw]))
    t_sample = time.perf_counter() - t0
    print(f"[train] {len(t)} samples from {a.dfas} automata in {t_sample:.1f}s; "
          f"distance mean {t.mean():.2f}, censored {(t > a.fdepth).mean():.3f}", flush=True)
    ctx = sw.Context(0)
    m = sw.Mlp(ctx, a.n, a.k, hidden=a.hidden, seed=7)
    if a.loss == "rank":
        m.set_loss(a.per_depth)
    t1 = time.perf_counter()
    curve = m.train(x, t, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=11)
    t_train = time.perf_counter() - t1
    print(f"[train] loss per epoch {[round(float(c), 4) for c in curve]} in {t_train:.1f}s", flush=True)
    out = Path(a.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    meta = {"n": a.n, "k": a.k, "hidden": a.hidden, "dfas": a.dfas, "loss": a.loss, "seed0": a.seed0, "fdepth": a.fdepth,
            "width": a.width, "bdepth": a.bdepth, "per_depth": a.per_depth, "samples": int(len(t)),
            "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "loss_curve": [float(c) for c in curve],
            "seconds_sampling": t_sample, "seconds_training": t_train}
    np.savez(out, params=m.params(), meta=json.dumps(meta))
    print(json.dumps(meta))
    m.close()
    ctx.close()
    return 0


def load_scorer(ctx: "sw.Context", path: str | Path) -> "sw.Mlp":
    """An Mlp with the parameters saved by this script."""
    z = np.load(path, allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    m = sw.Mlp(ctx, meta["n"], meta["k"], hidden=meta["hidden"], seed=0)
    m.set_params(z["params"])
    return m


if __name__ == "__main__":
    raise SystemExit(main())
