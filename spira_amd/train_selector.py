"""Train the batched-SpMV schedule selector on one GPU.

    python -m spira_amd.train_selector --train 2000 --heldout 500 --out <dir>

Builds the seeded corpus (batched-workload generator: rows log-uniform in
[256, 65536], nnz/row a 50/50 mix of uniform 4..64 and clipped Zipf(1.6)),
runs the exhaustive autotuner on it (every schedule of every matrix at k = 1,
4, 16, 32, costed inside the batch by device timestamps per work item),
trains the MLP on the device (hand-written HIP forward/backward/Adam) on the
training seeds, and scores it on the held-out seeds: per matrix as the
fraction of the best achievable cost, and on the whole held-out batch as
timed launches with the selector's, the built-in and the exhaustive-best
schedules.  Writes selector_model.json (weights + scores) and
selector_timings.npz to --out.
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
    ap.add_argument("--reps", type=int, default=2)
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
        train_m = S.build_corpus(dev, args.train, args.seed, max_rows=args.max_rows)
        log(f"train corpus: {len(train_m)} matrices, {sum(A.nnz for A in train_m)} nnz ({time.time() - t0:.1f}s)")
        Ft, Tt, Kt, Mt, nct = S.collect(dev, train_m, args.reps, log)
        log(f"train costs done ({time.time() - t0:.1f}s)")
        for A in train_m:
            A.close()
        held_m = S.build_corpus(dev, args.heldout, args.heldout_seed, max_rows=args.max_rows)
        Fh, Th, Kh, Mh, nch = S.collect(dev, held_m, args.reps, log)
        log(f"held-out costs done ({time.time() - t0:.1f}s)")

    sel = S.Selector(dev, seed=7)
    hist = sel.fit(Ft, Tt, epochs=args.epochs, log=log)
    choice = sel.predict(Fh, Kh, nch[Mh])
    frac_sel = S.fraction_of_best(Th, choice)
    heur = np.array([S.heuristic_schedule(nch[i], k) for i, k in zip(Mh, Kh)])
    frac_heur = S.fraction_of_best(Th, heur)
    tr_choice = sel.predict(Ft, Kt, nct[Mt])
    frac_train = S.fraction_of_best(Tt, tr_choice)
    per_k = {}
    for k in S.K_LIST:
        m = Kh == k
        per_k[str(k)] = {"selector": float(frac_sel[m].mean()), "heuristic": float(frac_heur[m].mean())}

    # whole held-out batch, timed launches: selector vs built-in vs exhaustive-best
    launches = {}
    with torch.cuda.stream(stream):
        B = spira_amd.Batch(dev, held_m, "f32")
        X = torch.randn((B.n_cols * 32,), dtype=torch.float32, device="cuda")
        Y = torch.empty((B.n_rows * 32,), dtype=torch.float32, device="cuda")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for k in S.K_LIST:
            m = Kh == k
            rec = {}
            for name, idx in (("selector", choice[m]), ("heuristic", heur[m]),
                              ("exhaustive", np.nanargmin(Th[m], axis=1))):
                S.apply(B, k, idx, nch)
                B.spmm(X, Y, sync=True)
                e0.record(stream)
                for _ in range(5):
                    B.spmm(X, Y, sync=False)
                e1.record(stream)
                e1.synchronize()
                rec[name + "_ms"] = e0.elapsed_time(e1) / 5
            rec["selector_fraction_of_exhaustive"] = rec["exhaustive_ms"] / rec["selector_ms"]
            rec["heuristic_fraction_of_exhaustive"] = rec["exhaustive_ms"] / rec["heuristic_ms"]
            launches[str(k)] = rec
            log(f"held-out batch k={k}: {rec}")
        B.close()

    scores = {
        "heldout_mean_fraction_of_best": {"selector": float(frac_sel.mean()), "heuristic": float(frac_heur.mean())},
        "heldout_per_k": per_k,
        "heldout_batch_launches": launches,
        "train_mean_fraction_of_best": float(frac_train.mean()),
        "loss_history": hist,
        "corpus": {"train": args.train, "heldout": args.heldout, "seed": args.seed,
                   "heldout_seed": args.heldout_seed, "reps": args.reps, "max_rows": args.max_rows},
    }
    log(json.dumps({k: v for k, v in scores.items() if k != "loss_history"}))
    os.makedirs(args.out, exist_ok=True)
    sel.save(os.path.join(args.out, "selector_model.json"), **scores)
    np.savez_compressed(os.path.join(args.out, "selector_timings.npz"), Ft=Ft, Tt=Tt, Kt=Kt, Mt=Mt, nct=nct,
                        Fh=Fh, Th=Th, Kh=Kh, Mh=Mh, nch=nch)
    log(f"done in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    sys.exit(main())
