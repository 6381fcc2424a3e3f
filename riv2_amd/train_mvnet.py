"""Trains the learned form's MV-proposal network on pairs from the plain GPU search, then compares
the learned encoder with the plain one on held-out sequences.

    python -m riv2_amd.train_mvnet --train cif:0:1000:5 --train hd:0:64:4 --epochs 6 \
        --eval cif:1000:1050:30 --eval cif1080:1000:1008:10 --eval hd:1000:1008:10 --passes 4 \
        --out riv2_amd/mvnet_weights.npz --report eval.json

Training pairs: every block of 1 I + P P-frames of each training seed's synthetic sequence, the
target being the plain search's integer vector; each --train is generator:first:end:P. Generators:
`cif` (352x288, global motion up to 3 px per frame at half-pel steps, 4 rectangles of 32-96 px
moving up to 6 px), `cif1080` (the same motion at 1920x1080) and `hd` (1920x1080, 8 px, 12
rectangles of 64-320 px moving up to 16 px). Each --eval is generator:first:end:frames over
held-out seeds, disjoint from the training seeds; the learned encoder runs `passes` relaxation
passes per motion search (0: to the plain form's fixed point).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from riv2_amd import learned, synth  # noqa: E402

GENERATORS = {
    "cif": lambda seed, n: synth.cif_params(seed=seed, frames=n),
    "cif1080": lambda seed, n: synth.custom_params(1920, 1080, n, seed=seed, global_motion=3.0, n_rects=4, rect_min=32,
                                                   rect_max=96, rect_velocity=6.0),
    "hd": lambda seed, n: synth.hd_params(seed=seed, frames=n),
}


def seed_range(s):
    a, b = s.split(":")
    return list(range(int(a), int(b)))


def evaluate(net, specs, qualities, confidence, passes, log):
    out = []
    for spec in specs:
        gen, a, b, frames = spec.split(":")
        ev = learned.compare(net, list(range(int(a), int(b))), int(frames), qualities, confidence, log=log, passes=passes,
                             params=GENERATORS[gen])
        out.append({"generator": gen, "seeds": f"{a}:{b}", "frames": int(frames), "confidence": confidence, "passes": passes, **ev})
        log(json.dumps(out[-1]))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--train", action="append", default=None, help="generator:first:end:pframes (repeatable)")
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--init-seed", type=int, default=1)
    ap.add_argument("--confidence", type=float, default=0.5)
    ap.add_argument("--eval", action="append", default=None, help="generator:first:end:frames (repeatable)")
    ap.add_argument("--passes", type=int, default=4)
    ap.add_argument("--qualities", default="70,75,80,85")
    ap.add_argument("--out", default=learned.WEIGHTS_PATH)
    ap.add_argument("--report", default=None)
    ap.add_argument("--weights", default=None, help="evaluate these trained parameters instead of training")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)

    specs = a.eval if a.eval is not None else ["cif:1000:1050:30", "hd:1000:1008:10"]
    qs = tuple(int(q) for q in a.qualities.split(","))
    if a.weights:
        net = learned.MvNet.load(a.weights)
        report = {"weights": os.path.relpath(a.weights, ROOT), "eval": evaluate(net, specs, qs, a.confidence, a.passes, log)}
        if a.report:
            with open(a.report, "w") as f:
                json.dump(report, f, indent=1)
        return report
    train_specs = a.train if a.train is not None else ["cif:0:1000:5", "hd:0:64:4"]
    t0 = time.time()
    parts = []
    for spec in train_specs:
        gen, lo, hi, pf = spec.split(":")
        parts.append(learned.collect_pairs(list(range(int(lo), int(hi))), int(pf), log=log, params=GENERATORS[gen]))
    ps = learned.merge_pairs(parts)
    t1 = time.time()
    log(f"{len(ps.samples)} pairs from {len(ps.sources)} P-frames in {t1 - t0:.1f} s")
    net = learned.MvNet(0, a.init_seed)
    losses = learned.train(net, ps, a.epochs, a.batch, a.lr, log=log)
    t2 = time.time()
    acc = learned.proposal_accuracy(net, ps, a.confidence)
    log(f"trained in {t2 - t1:.1f} s; training pairs: nearest proposal within 1 px {acc[0]:.3f}, confident {acc[1]:.3f}, "
        f"within 1 px when confident {acc[2]:.3f}")
    net.save(a.out, train=",".join(train_specs), epochs=a.epochs, batch=a.batch, lr=a.lr)
    npairs = len(ps.samples)
    del ps
    report = {"train": {"sets": train_specs, "pairs": int(npairs),
                        "collect_seconds": round(t1 - t0, 1),
                        "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "seconds": round(t2 - t1, 1),
                        "loss": [(int(s), round(float(l), 4)) for s, l in losses],
                        "train_within_1px": acc[0], "train_confident": acc[1], "train_within_1px_confident": acc[2]}}
    if specs:
        report["eval"] = evaluate(net, specs, qs, a.confidence, a.passes, log)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    return report


if __name__ == "__main__":
    main()
