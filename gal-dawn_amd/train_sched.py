"""Train the step scheduler's MLP on per-step mode timings.

1. Record: on each training graph the engine runs batches of 64*W sources
   (W = the graph's lane_words knob, 16 by default: 1024 sources); at
   every step it times PUSH, PULL and BITMAP from the same state (HIP
   events, minimum of `repeats`) and stores the step's 8 features
   (dawn_sched_record).
2. Train: the FP32 MLP 8->32->32->3 regresses log2 of the three step times
   (hand-written HIP forward/backward + Adam, dawn_mlp_train).
3. Save gal-dawn_amd/sched_mlp.json (params, feature normalisation) and the
   recorded table sched_data.npz, and report the schedule time of the
   learned choice vs the fixed alpha/beta rule vs the per-step best on the
   recorded steps.

Training graphs (seeds 10-40): RMAT scales 16-23 (ef 16), Erdos-Renyi, and
grids with 1% long-range edges.  The held-out RMAT scale 24 graph of the
benchmark is not used here.

--single trains the single-source engine's network instead (dawn_bfs; its
own parameter set, dawn_sched_set_mlp_single): the steps of single-source
traversals from `--sources` sampled vertices of each graph, every mode
timed at every step (dawn_sched_record_single), saved to
sched_mlp_single.json and sched_data_single.npz.

  python gal-dawn_amd/train_sched.py [--quick] [--single]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
import __graft_entry__  # noqa: E402

dawn = __graft_entry__.load_package()


def training_graphs(quick: bool):
    if quick:
        return [("rmat", 14, 10), ("rmat", 15, 11), ("er", 1 << 15, 12), ("grid", 128, 13)]
    gs = []
    for i, scale in enumerate(range(16, 24)):
        gs.append(("rmat", scale, 10 + i))
        gs.append(("rmat", scale, 20 + i))
    for i, n in enumerate((1 << 17, 1 << 19, 1 << 21)):
        gs.append(("er", n, 30 + i))
    for i, side in enumerate((256, 512, 1024)):
        gs.append(("grid", side, 35 + i))
    return gs


def make(kind, size, seed):
    if kind == "rmat":
        return dawn.HostGraph.rmat(size, 16, seed)
    if kind == "er":
        return dawn.HostGraph.er(size, 4.0 + (seed % 5) * 3.0, seed)
    return dawn.HostGraph.grid(size, 0.01, seed)


def record(quick: bool, repeats: int, batches: int, single: bool = False, sources: int = 16):
    X, Y, F, meta = [], [], [], []
    for kind, size, seed in training_graphs(quick):
        h = make(kind, size, seed)
        g = dawn.Graph(h)
        if single:
            src = dawn.gen_sources(h.n, 8 * sources, seed)
            src = src[h.degrees()[src] > 0][:sources]  # isolated sources have no steps
        else:
            src = dawn.gen_sources(h.n, 64 * g.knobs()["lane_words"] * batches, seed)
        t0 = time.time()
        feat, times, fixed = g.sched_record(src, repeats=repeats, single=single,
                                            max_rows=16384)
        g.close()
        X.append(feat)
        Y.append(times)
        F.append(fixed)
        meta.append({"kind": kind, "size": size, "seed": seed, "n": h.n, "m": h.m,
                     "rows": int(len(feat)), "record_s": round(time.time() - t0, 2)})
        print(f"record {kind} {size} seed {seed}: n={h.n} m={h.m} rows={len(feat)}", flush=True)
    return np.concatenate(X), np.concatenate(Y), np.concatenate(F), meta


def schedule_cost(times, choice):
    return float(times[np.arange(len(choice)), choice].sum())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--batches", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", default=None)
    ap.add_argument("--data", default=None, help="reuse a recorded sched_data.npz")
    ap.add_argument("--single", action="store_true", help="the single-source engine's network")
    ap.add_argument("--sources", type=int, default=16, help="--single: sources per graph")
    args = ap.parse_args()
    if args.out is None:
        args.out = os.path.join(HERE, "sched_mlp_single.json" if args.single else "sched_mlp.json")

    if args.data:
        d = np.load(args.data, allow_pickle=False)
        X, T, FX = d["features"], d["times_ms"], d["fixed_modes"]
        meta = json.loads(str(d["meta"]))
    else:
        X, T, FX, meta = record(args.quick, args.repeats, args.batches, args.single, args.sources)
        np.savez(os.path.join(os.path.dirname(os.path.abspath(args.out)),
                              "sched_data_single.npz" if args.single else "sched_data.npz"),
                 features=X, times_ms=T, fixed_modes=FX,
                 meta=json.dumps(meta))
    Y = np.log2(np.maximum(T, 1e-4) * 1e3).astype(np.float32)  # log2 microseconds
    mean = X.mean(axis=0)
    std = X.std(axis=0)
    std[std < 1e-6] = 1.0
    Xn = ((X - mean) / std).astype(np.float32)

    p0 = dawn.mlp_init(args.seed)
    params, loss = dawn.mlp_train(p0, Xn, Y, epochs=args.epochs, batch=256, lr=args.lr,
                                  seed=args.seed)
    pred = dawn.mlp_forward(params, Xn)
    learned = pred.argmin(axis=1)
    best = T.argmin(axis=1)
    out = {
        "params": params.tolist(),
        "mean": mean.astype(np.float32).tolist(),
        "inv_std": (1.0 / std).astype(np.float32).tolist(),
        "train": {
            "rows": int(len(X)),
            "final_mse_log2us": loss,
            "graphs": meta,
            "oracle_best_ms": schedule_cost(T, best),
            "learned_ms": schedule_cost(T, learned),
            "fixed_ms": schedule_cost(T, FX),
            "always_push_ms": schedule_cost(T, np.zeros(len(T), int)),
            "always_pull_ms": schedule_cost(T, np.ones(len(T), int)),
            "choice_accuracy": float((learned == best).mean()),
            "epochs": args.epochs,
            "lr": args.lr,
        },
    }
    with open(args.out, "w") as f:
        json.dump(out, f)
    print(json.dumps(out["train"] | {"graphs": len(meta)}), flush=True)


if __name__ == "__main__":
    main()
