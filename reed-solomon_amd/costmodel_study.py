"""Cost-model study on one GPU: exhaustive autotune of sampled shapes, MLP
training on the GPU, regret on held-out shapes.

    python reed-solomon_amd/costmodel_study.py [--shapes 3000] [--test 500]
        [--top-n 3] [--steps 3000] [--out results.json]

Prints progress lines while measuring and one JSON summary at the end:
mean / p90 / max regret of the model-guided choice (top-n measured), of the
model's top-1 pick (no measurement), of the codec's own GPU-calibrated rules
and of the reference's routing (both without a model), all against the
exhaustive optimum of each held-out shape.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--shapes", type=int, default=3000)
    ap.add_argument("--test", type=int, default=500)
    ap.add_argument("--top-n", type=int, default=3)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0xC057)
    ap.add_argument("--cap-mib", type=int, default=32)
    ap.add_argument("--out", default=None)
    ap.add_argument("--detail", default=None,
                    help="JSON lines, one per shape: every timed candidate, the best, the "
                         "heuristic and (held-out shapes) the model's choice")
    args = ap.parse_args()

    import __graft_entry__

    trs = __graft_entry__.load()
    from reed_solomon_amd import autotune as A

    shapes = A.sample_shapes(args.shapes, args.seed)
    timings = []
    t0 = time.time()
    for i, s in enumerate(shapes):
        timings.append(A.exhaustive(s, cap_bytes=args.cap_mib << 20))
        if (i + 1) % 250 == 0:
            print(f"measured {i + 1}/{len(shapes)} shapes in {time.time() - t0:.1f} s", flush=True)
    ntrain = len(shapes) - args.test
    Xtr, Ttr = A.dataset(shapes[:ntrain], timings[:ntrain])
    model = A.CostModel(seed=args.seed)
    t1 = time.time()
    loss = model.fit(Xtr, Ttr, steps=args.steps, lr=args.lr)
    train_s = time.time() - t1
    tuner = A.Tuner(model, top_n=args.top_n)
    reg_model, reg_top1, reg_heur, reg_rule, measured_counts = [], [], [], [], []
    detail = open(args.detail, "w") if args.detail else None
    for idx, (s, meas) in enumerate(zip(shapes, timings)):
        if detail is None:
            break
        best = min(meas, key=meas.get)
        rec = {"field": s.field, "k": s.k, "m": s.m, "length": s.length, "held_out": idx >= ntrain,
               "best": A.VARIANTS[best], "best_ms": meas[best],
               "heuristic": A.VARIANTS[A.heuristic(s)], "rule": A.VARIANTS[A.rule(s)],
               "times_ms": {A.VARIANTS[v]: t for v, t in meas.items()}}
        if idx >= ntrain:
            rec["model"] = A.VARIANTS[tuner.choose(s, meas)]
        detail.write(json.dumps(rec) + "\n")
    if detail is not None:
        detail.close()
    for s, meas in zip(shapes[ntrain:], timings[ntrain:]):
        reg_model.append(A.regret(meas, tuner.choose(s, meas)))
        reg_top1.append(A.regret(meas, A.Tuner(model, top_n=1).choose(s, meas)))
        reg_heur.append(A.regret(meas, A.heuristic(s)))
        reg_rule.append(A.regret(meas, A.rule(s)))
        measured_counts.append(len(meas))

    def stats(r):
        r = np.array(r)
        return {"mean": round(float(r.mean()), 4), "p90": round(float(np.percentile(r, 90)), 4),
                "max": round(float(r.max()), 4), "within_10pct": round(float((r <= 0.10).mean()), 4)}

    summary = {
        "shapes": len(shapes), "train_shapes": ntrain, "test_shapes": args.test,
        "train_samples": int(len(Ttr)), "measure_seconds": round(t1 - t0, 1),
        "train_seconds": round(train_s, 2), "train_loss": round(float(loss), 5),
        "top_n": args.top_n,
        "candidates_per_test_shape": round(float(np.mean(measured_counts)), 2),
        "regret_model_top_n": stats(reg_model),
        "regret_model_top1": stats(reg_top1),
        "regret_heuristic_no_model": stats(reg_heur),
        "regret_codec_rules_no_model": stats(reg_rule),
    }
    print(json.dumps(summary), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(summary, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
