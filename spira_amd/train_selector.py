"""Train the batched-SpMV schedule selector on one GPU.

    python -m spira_amd.train_selector --train 2000 --heldout 500 --out <dir>

Builds the seeded corpus (batched-workload generator: rows log-uniform in
[256, 65536], nnz/row a 50/50 mix of uniform 4..64 and clipped Zipf(1.6)),
times every schedule of every matrix at k = 1, 4, 16, 32 with the exhaustive
autotuner, trains the MLP on the device (hand-written HIP forward/backward/Adam)
on the training seeds, and scores it on the held-out seeds as the fraction of
the best achievable throughput, beside the built-in heuristic.  Writes
selector_model.json (weights + scores) and selector_timings.npz to --out.
"""
import argparse
import json
import os
import sys
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--heldout", type=int, default=500)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--heldout-seed", type=int, default=900000)
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--max-rows", type=int, default=65536)
    ap.add_argument("--out", default=".")
    args = ap.parse_args()

    import torch

    import spira_amd
    from spira_amd import selector as S

    def log(m):
        print(m, flush=True)

    stream = torch.cuda.Stream()
    dev = spira_amd.Device(0, stream=stream.cuda_stream)
    t0 = time.time()
    with torch.cuda.stream(stream):
        train_m, train_mean = S.build_corpus(dev, args.train, args.seed, max_rows=args.max_rows)
        log(f"train corpus: {len(train_m)} matrices, {sum(A.nnz for A in train_m)} nnz ({time.time() - t0:.1f}s)")
        Ft, Tt, Kt, Mt = S.collect(dev, train_m, stream, args.reps, log)
        log(f"train timings done ({time.time() - t0:.1f}s)")
        for A in train_m:
            A.close()
        held_m, held_mean = S.build_corpus(dev, args.heldout, args.heldout_seed, max_rows=args.max_rows)
        Fh, Th, Kh, Mh = S.collect(dev, held_m, stream, args.reps, log)
        log(f"held-out timings done ({time.time() - t0:.1f}s)")

    sel = S.Selector(dev, seed=7)
    hist = sel.fit(Ft, Tt, epochs=args.epochs, log=log)
    choice = sel.predict(Fh, Kh)
    frac_sel = S.fraction_of_best(Th, choice)
    heur = np.array([S.heuristic_schedule(held_mean[i], k) for i, k in zip(Mh, Kh)])
    frac_heur = S.fraction_of_best(Th, heur)
    tr_choice = sel.predict(Ft, Kt)
    frac_train = S.fraction_of_best(Tt, tr_choice)
    per_k = {}
    for k in S.K_LIST:
        m = Kh == k
        per_k[str(k)] = {"selector": float(frac_sel[m].mean()), "heuristic": float(frac_heur[m].mean())}
    # time-weighted: total best time / total chosen time
    tw_sel = float(np.nanmin(Th, 1).sum() / Th[np.arange(len(choice)), choice].sum())
    tw_heur = float(np.nanmin(Th, 1).sum() / Th[np.arange(len(heur)), heur].sum())
    scores = {
        "heldout_mean_fraction_of_best": {"selector": float(frac_sel.mean()), "heuristic": float(frac_heur.mean())},
        "heldout_time_weighted_fraction_of_best": {"selector": tw_sel, "heuristic": tw_heur},
        "heldout_per_k": per_k,
        "train_mean_fraction_of_best": float(frac_train.mean()),
        "loss_history": hist,
        "corpus": {"train": args.train, "heldout": args.heldout, "seed": args.seed,
                   "heldout_seed": args.heldout_seed, "reps": args.reps, "max_rows": args.max_rows},
    }
    log(json.dumps({k: v for k, v in scores.items() if k != "loss_history"}))
    os.makedirs(args.out, exist_ok=True)
    sel.save(os.path.join(args.out, "selector_model.json"), **scores)
    np.savez_compressed(os.path.join(args.out, "selector_timings.npz"), Ft=Ft, Tt=Tt, Kt=Kt, Fh=Fh, Th=Th, Kh=Kh,
                        held_mean=held_mean, Mh=Mh)
    log(f"done in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    sys.exit(main())
