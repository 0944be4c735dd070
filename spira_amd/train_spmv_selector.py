"""Train the single-matrix SpMV kernel selector on one GPU.

    python -m spira_amd.train_spmv_selector --train 2400 --heldout 500 --out <dir>

Generates the seeded corpus (workloads.spmv_corpus_matrix: uniform, banded,
stencil, power-law, block-diagonal and mostly-local matrices up to 2^23
rows), times every FP64 variant on each with the exhaustive autotuner,
trains the MLP on the device on the training seeds and scores it on the
held-out seeds as the fraction of the best variant's throughput, next to the
row-length rule used without a trained model.  Writes
spmv_selector_model.json (weights + scores), spmv_selector_weights.inc (the
weights as the library's include) and spmv_selector_timings.npz to --out.
"""
import argparse
import json
import os
import sys
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2400)
    ap.add_argument("--heldout", type=int, default=500)
    ap.add_argument("--seed", type=int, default=100000)
    ap.add_argument("--heldout-seed", type=int, default=900000)
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--max-rows", type=int, default=1 << 23)
    ap.add_argument("--timings", default="", help="reuse a timings .npz instead of timing the corpus")
    ap.add_argument("--out", default=".")
    args = ap.parse_args()

    import torch

    import spira_amd
    from spira_amd import spmv_selector as S

    def log(m):
        print(m, flush=True)

    stream = torch.cuda.Stream()
    dev = spira_amd.Device(0, stream=stream.cuda_stream)
    t0 = time.time()
    if args.timings:
        d = np.load(args.timings)
        Ft, Tt, Fh, Th = d["Ft"], d["Tt"], d["Fh"], d["Th"]
        meta_h = json.loads(str(d["meta_h"]))
    else:
        with torch.cuda.stream(stream):
            Ft, Tt, _ = S.collect(dev, range(args.seed, args.seed + args.train), args.reps, log, args.max_rows)
            log(f"train corpus timed ({time.time() - t0:.1f}s)")
            Fh, Th, meta_h = S.collect(dev, range(args.heldout_seed, args.heldout_seed + args.heldout), args.reps,
                                       log, args.max_rows)
            log(f"held-out corpus timed ({time.time() - t0:.1f}s)")
        os.makedirs(args.out, exist_ok=True)
        np.savez_compressed(os.path.join(args.out, "spmv_selector_timings.npz"), Ft=Ft, Tt=Tt, Fh=Fh, Th=Th,
                            meta_h=json.dumps(meta_h))
    sel = S.SpmvSelector(dev, seed=11)
    with torch.cuda.stream(stream):
        hist = sel.fit(Ft, Tt, epochs=args.epochs, log=log)
        choice = sel.predict(Fh, np.isfinite(Th))
        tr_choice = sel.predict(Ft, np.isfinite(Tt))
    frac = S.fraction_of_best(Th, choice)
    rule = np.array([m["rule"] for m in meta_h])
    frac_rule = S.fraction_of_best(Th, rule)
    best_h = np.nanmin(Th, axis=1)
    fams = sorted({m["family"] for m in meta_h})
    per_family = {fam: {"selector": float(frac[[m["family"] == fam for m in meta_h]].mean()),
                        "rule": float(frac_rule[[m["family"] == fam for m in meta_h]].mean())} for fam in fams}
    scores = {
        "heldout_mean_fraction_of_best": {"selector": float(frac.mean()), "rule": float(frac_rule.mean())},
        "heldout_total_time_fraction": {"selector": float(best_h.sum() / Th[np.arange(len(choice)), choice].sum()),
                                        "rule": float(best_h.sum() / Th[np.arange(len(rule)), rule].sum())},
        "heldout_min_fraction_of_best": float(frac.min()),
        "heldout_fraction_below_0.9": float((frac < 0.9).mean()),
        "heldout_per_family": per_family,
        "train_mean_fraction_of_best": float(S.fraction_of_best(Tt, tr_choice).mean()),
        "loss_history": hist,
        "corpus": {"train": int(len(Ft)), "heldout": int(len(Fh)), "seed": args.seed,
                   "heldout_seed": args.heldout_seed, "reps": args.reps, "max_rows": args.max_rows},
    }
    log(json.dumps({k: v for k, v in scores.items() if k != "loss_history"}))
    os.makedirs(args.out, exist_ok=True)
    sel.save(os.path.join(args.out, "spmv_selector_model.json"), **scores)
    sel.export_inc(os.path.join(args.out, "spmv_selector_weights.inc"))
    log(f"done in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    sys.exit(main())
