"""Train the learned warm start on the method's own FIRE trajectories and compare it with
plain FIRE (GPU; everything through the C ABI of libpcp_hip.so).

Workload: packings of N=1024 polygons with 16 vertices, bidisperse 1.0/1.4 with +-5%
radial noise, phi ~ U[0.82, 0.88], centres at least (R_i + R_j)/2 apart, squared
lens-area energy, rigid polygons (SHAKE every step), FIRE to max|F| <= 1e-10 with the
reference driver's FIRE parameters except dtMax (default 100: the squared lens energy's
stiffness scales as ell^2, so its natural time step is ~1/ell, far above 0.1).

The defaults reproduce pycudapolygon_amd/weights/warmstart_n1024_n16.npz bit for bit
(every step is deterministic):

  python -m pycudapolygon_amd.train_warmstart --out weights.npz

Steps: (1) FIRE on `batches` batches of `packings` packings, states captured at a few
step counts; (2) targets = displacement of each captured state to its minimized state;
(3) Adam on the GPU (last batch held out for validation); (4) A/B: plain FIRE and
warm-started FIRE on `eval-packings` new packings, both batched in one context each.
Prints one JSON line per phase and saves the weights (params, input/output scales).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from pycudapolygon_amd import warmstart as W   # noqa: E402
else:
    from . import warmstart as W


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--batches", type=int, default=157)
    ap.add_argument("--packings", type=int, default=64)
    ap.add_argument("--N", type=int, default=1024)
    ap.add_argument("--n", type=int, default=16)
    ap.add_argument("--captures", default="0,10,25,50,100")
    ap.add_argument("--max-steps", type=int, default=3000)
    ap.add_argument("--dt-max", type=float, default=100.0)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=2e-4)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--out", default="warmstart_weights.npz")
    ap.add_argument("--eval-packings", type=int, default=32)
    ap.add_argument("--eval-seed", type=int, default=777777)
    ap.add_argument("--applications", type=int, default=1)
    ap.add_argument("--eval-dt0", default="0.001",
                    help="comma-separated initial FIRE time steps of the A/B comparison (both arms)")
    ap.add_argument("--mode", choices=["fresh", "compression"], default="fresh",
                    help="fresh: FIRE trajectories from new packings (the defaults above); compression: "
                         "states just after each packing-fraction step of a ramp, labelled with the state "
                         "FIRE reaches from them")
    ap.add_argument("--replicas", type=int, default=16)
    ap.add_argument("--phi-steps", type=int, default=40)
    args = ap.parse_args(argv)

    log = lambda m: print(m, flush=True)
    caps = tuple(int(c) for c in args.captures.split(","))
    t0 = time.perf_counter()
    if args.mode == "compression":
        ds, runs, inScale = W.generate_compression(args.replicas, args.N, args.n, args.seed, steps=args.phi_steps,
                                                   maxSteps=args.max_steps, dtMax=args.dt_max, log=log)
    else:
        ds, runs, inScale = W.generate(args.batches, args.packings, args.seed, capture_steps=caps,
                                       maxSteps=args.max_steps, log=log, N=args.N, n=args.n, dtMax=args.dt_max)
    outScale = W.label_all(ds, runs)
    tgen = time.perf_counter() - t0
    conv = np.concatenate([r[2] for r in runs])
    print(json.dumps(dict(phase="data", trajectories=int(len(conv)), converged=int((conv >= 0).sum()),
                          samples=len(ds.samples), vertices_per_sample=ds.samples[0]["V"], seconds=tgen,
                          median_fire_steps=float(np.median(conv[conv >= 0])), inScale=inScale.tolist(),
                          outScale=outScale)), flush=True)

    if args.mode == "compression":   # every 5th phi step held out
        val = ds.samples[4::5]
        trn = [d for i, d in enumerate(ds.samples) if i % 5 != 4]
    else:
        last = runs[-1][0]
        val = [d for d in ds.samples if d["sys"] is last] if args.batches > 1 else []
        trn = [d for d in ds.samples if d["sys"] is not last] if args.batches > 1 else ds.samples
    t1 = time.perf_counter()
    params, losses, vlosses = W.train(trn, W.init_params(0), args.epochs, lr=args.lr, log=log, val=val,
                                      lr_final=args.lr_final)
    ttrain = time.perf_counter() - t1
    W.save_weights(args.out, params, inScale, outScale, 4, mode=args.mode, N=args.N, n=args.n,
                   trajectories=int(len(conv)), held_out=len(val), captures=caps, dtMax=args.dt_max,
                   epochs=args.epochs, lr=args.lr, lr_final=args.lr_final, seed=args.seed,
                   replicas=args.replicas if args.mode == "compression" else 0,
                   phi_steps=args.phi_steps if args.mode == "compression" else 0)
    print(json.dumps(dict(phase="train", steps=args.epochs * len(trn), seconds=ttrain, train_loss=losses,
                          validation_loss=vlosses, zero_prediction_loss=2.0, weights=args.out)), flush=True)

    # free the training data before the comparison
    del ds, runs, trn, val
    if args.eval_packings > 0 and args.mode == "fresh":
        pk = W.make_c4_packings(args.eval_packings, args.eval_seed, N=args.N, n=args.n)
        for dt0 in (float(v) for v in args.eval_dt0.split(",")):
            res = W.ab_compare(W.load_weights(args.out), pk, maxSteps=args.max_steps,
                               applications=args.applications, log=log, dtMax=args.dt_max, dt0=dt0)
            res["phase"] = "ab"
            print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
