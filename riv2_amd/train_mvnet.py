"""Trains the learned form's MV-proposal network on pairs from the plain GPU search, then compares
the learned encoder with the plain one on held-out sequences.

    python -m riv2_amd.train_mvnet --seeds 0:64 --pframes 4 --epochs 2 \
        --eval-seeds 1000:1008 --eval-frames 10 --out riv2_amd/mvnet_weights.npz --report eval.json

Training pairs: every block of 1 I + `pframes` P-frames of each training seed's 1080p synthetic
sequence (global motion up to 8 px per frame, 12 rectangles moving up to 16 px), the target being
the plain search's integer vector. Held-out seeds are disjoint from the training seeds.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from riv2_amd import learned  # noqa: E402


def seed_range(s):
    a, b = s.split(":")
    return list(range(int(a), int(b)))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--seeds", default="0:64")
    ap.add_argument("--pframes", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--init-seed", type=int, default=1)
    ap.add_argument("--confidence", type=float, default=0.5)
    ap.add_argument("--eval-seeds", default="1000:1008")
    ap.add_argument("--eval-frames", type=int, default=10)
    ap.add_argument("--qualities", default="70,75,80,85")
    ap.add_argument("--out", default=learned.WEIGHTS_PATH)
    ap.add_argument("--report", default=None)
    ap.add_argument("--weights", default=None, help="evaluate these trained parameters instead of training")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)

    if a.weights:
        net = learned.MvNet.load(a.weights)
        qs = tuple(int(q) for q in a.qualities.split(","))
        ev = learned.compare(net, seed_range(a.eval_seeds), a.eval_frames, qs, a.confidence, log=log)
        report = {"weights": os.path.relpath(a.weights, ROOT),
                  "eval": {"seeds": a.eval_seeds, "frames": a.eval_frames, "confidence": a.confidence, **ev}}
        log(json.dumps(report["eval"]))
        if a.report:
            with open(a.report, "w") as f:
                json.dump(report, f, indent=1)
        return report
    t0 = time.time()
    ps = learned.collect_pairs(seed_range(a.seeds), a.pframes, log=log)
    t1 = time.time()
    log(f"{len(ps.samples)} pairs from {len(ps.sources)} P-frames in {t1 - t0:.1f} s")
    net = learned.MvNet(0, a.init_seed)
    losses = learned.train(net, ps, a.epochs, a.batch, a.lr, log=log)
    t2 = time.time()
    acc = learned.proposal_accuracy(net, ps, a.confidence)
    log(f"trained in {t2 - t1:.1f} s; training pairs: nearest proposal within 1 px {acc[0]:.3f}, confident {acc[1]:.3f}, "
        f"within 1 px when confident {acc[2]:.3f}")
    net.save(a.out, train_seeds=a.seeds, pframes=a.pframes, epochs=a.epochs, batch=a.batch, lr=a.lr)
    npairs = len(ps.samples)
    del ps
    report = {"train": {"seeds": a.seeds, "pframes": a.pframes, "pairs": int(npairs),
                        "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "seconds": round(t2 - t1, 1),
                        "loss": [(int(s), round(float(l), 4)) for s, l in losses],
                        "train_within_1px": acc[0], "train_confident": acc[1], "train_within_1px_confident": acc[2]}}
    if a.eval_seeds:
        qs = tuple(int(q) for q in a.qualities.split(","))
        ev = learned.compare(net, seed_range(a.eval_seeds), a.eval_frames, qs, a.confidence, log=log)
        report["eval"] = {"seeds": a.eval_seeds, "frames": a.eval_frames, "confidence": a.confidence, **ev}
        log(json.dumps(report["eval"]))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    return report


if __name__ == "__main__":
    main()
