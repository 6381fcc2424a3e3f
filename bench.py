#!/usr/bin/env python3
"""1080p P-frame encode throughput of the gfx950 RIV encoder (plain form, bit-exact bitstream; the
learned form on the same frames in the line's `learned` object).

Workload: 1920x1080 RGB24 synthetic sequence (seeded generator: 1/f^2 background, global motion up
to 8 px/frame at half-pel steps, 12 textured rectangles of 64-320 px at up to 16 px/frame, noise
sigma 2), intra/inter quality 85/80, search range 12, skip threshold 3, one I-frame then P-frames.
A step is one P-frame encode_frame call (GPU analysis + host rANS + packing) with its RGB input
already resident in HBM; the default run is the 50-frame sequence (1 I + 4 warm-up P + 45 timed
P). Prints one JSON line.

The roofline entry is for the dominant stage, the motion search (relaxation passes and mode
decision): achieved = 2*Ws*Hs algorithmic bytes (current + reference luma, each read once) / its
average duration from HIP events on the encoder's stream over the timed frames. cpu_baseline
times the oracle (a C++ restatement of the reference CPU encoder) on the host cores over a
bounded sample.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ME_KERNELS = ["k_me_search"]   # the motion-search stage (search passes and mode decision), one launch per frame
W, H = 1920, 1080
SW, SH = 1920, 1088


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor()


def make_config(riv2_amd, depth):
    return riv2_amd.EncoderConfig(W, H, 50).with_intra_quality(85).with_inter_quality(80).with_search_range(12) \
        .with_skip_threshold(3).with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(1000)).with_pipeline_depth(depth)


PROFILE_EVERY = 4   # timed frames per motion-search launch timed by events
LEARNED_PASSES = 4   # the learned form's relaxation passes per motion search (profiles/learned_mv_eval.json)


def run_gpu(frames, steps, warmup, depth, net=None, d_frames=None, confidence=0.5):
    """Times `steps` P-frames (after one I-frame and `warmup` P-frames) through the pipelined
    encoder, inputs resident in HBM; the motion-search launch of every PROFILE_EVERY-th timed frame
    records HIP events at its start and end on the encoder's stream. With `net`, the learned form (the network's proposals seed the search).
    Returns elapsed seconds, per-frame stats, stage times, the stream, the device frames and the
    learned-form counts."""
    import torch
    import riv2_amd

    dev = torch.device("cuda:0")
    if d_frames is None:
        d_frames = torch.from_numpy(frames).to(dev)
    torch.cuda.synchronize()
    fbytes = W * H * 3
    enc = riv2_amd.Encoder(make_config(riv2_amd, depth))
    if net is not None:
        enc.set_mvnet(net, confidence)
        enc.set_learned_passes(LEARNED_PASSES)
    ptr = lambda i: d_frames[i].data_ptr()
    enc.encode_frame_device(ptr(0), fbytes, W, H, 0)
    for i in range(1, 1 + warmup):
        enc.encode_frame_device(ptr(i), fbytes, W, H, i * 20)
    enc.flush()
    enc.pop_stats(4096)
    # the motion search's two events on every PROFILE_EVERY-th frame (the sample of its launch
    # duration): each event holds the stream ~4.6 us, a few percent of a frame
    enc.set_profiling(1)
    enc.set_profiling_sample(PROFILE_EVERY)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(1 + warmup, 1 + warmup + steps):
        enc.encode_frame_device(ptr(i), fbytes, W, H, i * 20)
    enc.flush()   # every timed frame entropy-coded and written
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stats = enc.pop_stats(4096)
    stages = enc.stage_times()
    counts = enc.learned_stats()
    enc.finish()
    data = enc.output()
    enc.close()
    return t1 - t0, stats, stages, data, d_frames, counts


def mean_pframe_psnr_y(frames, stream):
    """Mean PSNR-Y (BT.601 luma, as ri-quality computes it) of a stream's decoded P-frames."""
    import riv2_amd
    from riv2_amd import learned
    dec = riv2_amd.Decoder(stream)
    ps = []
    for k in range(len(frames)):
        f = dec.decode_frame()
        if k > 0:
            ps.append(learned.psnr_y(frames[k], f.data))
    return sum(ps) / len(ps)


def run_decode(stream, nframes, cpu_sample):
    """Decoder throughput on the plain bench stream: the GPU decoder (host MV and residual rANS
    decode feeding the GPU reconstruction, RGB24 back in host memory) over every frame, and the
    oracle's CPU decoder (one thread) over the first `cpu_sample` frames. Frames per second of
    decode_frame calls, as ri-quality times decoding (reitero_video_quality_test/src/main.rs:1417-1426)."""
    import riv2_amd
    import oracle
    dec = riv2_amd.Decoder(stream)
    dec.decode_frame()   # the I-frame, untimed
    t0 = time.perf_counter()
    for _ in range(1, nframes):
        dec.decode_frame()
    gpu = (nframes - 1) / (time.perf_counter() - t0)
    dec.close()
    rdec = oracle.OracleDecoder(stream)
    rdec.decode()
    t0 = time.perf_counter()
    for _ in range(cpu_sample):
        rdec.decode()
    cpu = cpu_sample / (time.perf_counter() - t0)
    return {"value": round(gpu, 3), "unit": "fps", "frames": nframes - 1,
            "cpu_baseline": {"value": round(cpu, 3), "unit": "fps", "cores": 1, "kind": "port",
                             "sample": f"the first {cpu_sample} P-frames of the same stream, oracle C++ restatement of the reference decoder"}}


def run_learned(frames, d_frames, steps, warmup, depth, plain_stream, plain_fps):
    """The learned form on the same frames, timed like the plain form, with the trained network
    committed beside the package (riv2_amd/mvnet_weights.npz, made by riv2_amd/train_mvnet.py)."""
    from riv2_amd import learned
    if not os.path.exists(learned.WEIGHTS_PATH):
        return {"value": None, "reason": "no trained weights at riv2_amd/mvnet_weights.npz"}
    net = learned.MvNet.load(learned.WEIGHTS_PATH)
    elapsed, stats, stages, stream, _, counts = run_gpu(frames, steps, warmup, depth, net, d_frames)
    fps = steps / elapsed
    me_ms, me_n = stages["motion_search"]
    p_plain = mean_pframe_psnr_y(frames, plain_stream)
    p_learned = mean_pframe_psnr_y(frames, stream)
    net.close()
    return {"value": round(fps, 3), "unit": "fps", "vs_plain": round(fps / plain_fps, 4),
            "ms_per_step": round(1000.0 * elapsed / steps, 4),
            "motion_search_avg_ms": round(me_ms / max(1, me_n), 4),
            "bytes_vs_plain": round(len(stream) / len(plain_stream), 5),
            "psnr_y": round(p_learned, 4), "plain_psnr_y": round(p_plain, 4),
            "fallback_rate": round(1.0 - counts[2] / max(1, counts[0]), 4),
            "confident_rate": round(counts[1] / max(1, counts[0]), 4),
            "note": "same frames and settings as the plain line; bytes and PSNR-Y over the whole 50-frame stream at q80 "
                    "(rate at equal PSNR over several qualities: profiles/learned_mv_eval.json)",
            "relaxation_passes": LEARNED_PASSES}


def stage_breakdown(d_frames, n, depth):
    """Per-stage device times from a separate, untimed pass with events around every stage. The
    pass is synchronous (each frame completes before the next is submitted), so a stage's time is
    its own and not a wait for another stream's work."""
    import riv2_amd
    enc = riv2_amd.Encoder(make_config(riv2_amd, 0))
    enc.encode_frame_device(d_frames[0].data_ptr(), W * H * 3, W, H, 0)
    enc.flush()
    enc.set_profiling(2)
    for i in range(1, n):
        enc.encode_frame_device(d_frames[i].data_ptr(), W * H * 3, W, H, i * 20)
    enc.flush()
    st = enc.stage_times()
    frames = [s.gpu_ms for s in enc.pop_stats(4096) if s.gpu_ms > 0]
    enc.close()
    return {k: round(v[0] / v[1], 4) for k, v in st.items() if v[1]}, (sum(frames) / len(frames) if frames else None)


GOLDEN = os.path.join(ROOT, "tests", "golden")


def sha256(b):
    import hashlib
    return hashlib.sha256(b).hexdigest()


def split_frames(stream):
    """Frame records of a .riv stream (36-byte header, then intra / inter records)."""
    out, p = [], 36
    while p < len(stream):
        q = p + 10
        if stream[p + 8] == 1:
            q += 4 + int.from_bytes(stream[q:q + 4], "little")
        else:
            q += 3
            q += 4 + int.from_bytes(stream[q:q + 4], "little")
            q += 4 + int.from_bytes(stream[q:q + 4], "little")
        out.append(stream[p:q])
        p = q
    return out


def check_against_fixture(seed, gpu_stream, nframes):
    """Every frame record of the stream against the oracle's digests of the same 1080p sequence
    (tests/golden/hd_config2.json: seed 1, 50 frames, timestamps 20 * index); any prefix of the
    sequence encodes to the same records. Returns (frames checked, all equal), or None when the
    run is not on the fixture's sequence."""
    path = os.path.join(GOLDEN, "hd_config2.json")
    if seed != 1 or nframes > 50 or not os.path.exists(path):
        return None
    with open(path) as f:
        want = json.load(f)["frames"]
    recs = split_frames(gpu_stream)
    ok = len(recs) == nframes and all(sha256(r) == want[i]["frame_sha256"] for i, r in enumerate(recs))
    return nframes, bool(ok)


def check_prefix_bit_exact(frames, gpu_stream, nframes):
    """Off the fixture's sequence: re-encodes the first frames with the oracle and compares the
    leading frame records."""
    import oracle
    ref = oracle.OracleEncoder(oracle.config(W, H))
    for i in range(nframes):
        ref.encode(frames[i], i * 20)
    ref_stream = ref.finish()
    body = len(ref_stream) - 36
    return gpu_stream[36:36 + body] == ref_stream[36:]


def run_drop_in(frames, steps, warmup, depth, pinned=False):
    """The drop-in boundary: riv_encode_frame with host RGB24 (pageable numpy frames, or frames in
    pinned host memory when `pinned`), timed like the resident-input line, so the H2D copy of every
    frame is inside the timed region."""
    import riv2_amd
    if pinned:
        import torch
        held = []
        for f in frames[:1 + warmup + steps]:
            t = torch.empty(f.size, dtype=torch.uint8, pin_memory=True)
            a = t.numpy().reshape(f.shape)
            a[...] = f
            held.append((t, a))
        frames = [a for _, a in held]
    enc = riv2_amd.Encoder(make_config(riv2_amd, depth))
    for i in range(1 + warmup):
        enc.encode_frame(riv2_amd.Frame(frames[i], W, H, i * 20))
    enc.flush()
    t0 = time.perf_counter()
    for i in range(1 + warmup, 1 + warmup + steps):
        enc.encode_frame(riv2_amd.Frame(frames[i], W, H, i * 20))
    enc.flush()
    elapsed = time.perf_counter() - t0
    enc.finish()
    data = enc.output()
    enc.close()
    return steps / elapsed, data


def cpu_baseline(frames, pframes, threads):
    """Oracle encoder on the host cores: serial-raster ME (bit-exact mode) is the reference library
    default; threads > 0 emulates its `threads` feature (rows_per_strip = ceil(MB_rows / (4 T)))."""
    import oracle
    enc = oracle.OracleEncoder(oracle.config(W, H, me_threads=threads))
    enc.encode(frames[0], 0)
    t0 = time.perf_counter()
    for i in range(1, 1 + pframes):
        enc.encode(frames[i], i * 20)
    return pframes / (time.perf_counter() - t0)


BATCH_SIZES = [(352, 288), (640, 360), (1280, 720), (1920, 1080)]


def _batch_config(riv2_amd, w, h, depth):
    return riv2_amd.EncoderConfig(w, h, 30).with_intra_quality(85).with_inter_quality(80).with_search_range(12) \
        .with_skip_threshold(3).with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(1000)).with_pipeline_depth(depth)


def _storage_luma_bytes(streams):
    return sum(((w + 15) // 16 * 16) * ((h + 15) // 16 * 16) for w, h in streams)


def _timed_batch(encs, frames, p_frames):
    """Batch steps over device frames: the I step and the autotuner's trial steps untimed, then the
    remaining P steps timed with the leader's whole search launch bracketed by HIP events (stage 7).
    Returns (timed steps, untimed steps, seconds, average search ms)."""
    import torch
    import riv2_amd
    step = lambda i: riv2_amd.encode_batch_step(encs, [f[i].data_ptr() for f in frames], [i * 33] * len(encs), on_device=True)
    step(0)
    first = 1
    while first < p_frames - 2:
        step(first)
        first += 1
        for e in encs:
            e.flush()
        if encs[0].batch_tuning()["settled"]:
            break
    for e in encs:
        e.flush()
    encs[0].set_profiling(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(first, 1 + p_frames):
        step(i)
    for e in encs:
        e.flush()
    elapsed = time.perf_counter() - t0
    ms, n = encs[0].stage_times()["batch_motion_search"]
    return 1 + p_frames - first, first, elapsed, ms / max(1, n)


def _batch_roofline(streams, search_ms):
    nbytes = 2 * _storage_luma_bytes(streams)   # current + reference luma of every stream, each read once
    achieved = nbytes / (search_ms * 1e-3) / 1e9 if search_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": None, "kernel": "motion search (k_me_search, one launch over every stream of a batch step)",
            "algorithmic_bytes_per_launch": nbytes, "avg_launch_ms": round(search_ms, 4)}


def _cpu_baseline_mix(make_frames, threads, pframes=2):
    """The oracle on one stream of each size (its strip-threaded search on `threads` threads), 1 I +
    pframes P-frames each, timed from the first P-frame: the P-frames/s it would sustain over the
    mix (one frame of every size per step)."""
    import oracle
    per = []
    for w, h in BATCH_SIZES:
        f = make_frames(w, h, 1 + pframes)
        enc = oracle.OracleEncoder(oracle.config(w, h, me_threads=threads))
        enc.encode(f[0], 0)
        t0 = time.perf_counter()
        for i in range(1, 1 + pframes):
            enc.encode(f[i], i * 33)
        per.append((time.perf_counter() - t0) / pframes)
    return len(BATCH_SIZES) / sum(per)


def run_batched(per_size, p_frames, depth, with_cpu):
    """The batched workload: per_size streams of each of BATCH_SIZES, each 1 I + p_frames P-frames
    of the CIF reference case's generator (global motion up to 3 px, 4 rectangles of 32-96 px at up
    to 6 px/frame) with its own seed, frames resident in HBM. One batch step encodes one frame of
    every stream. The I step and the P steps the launch-shape autotuner spends on its trials are
    untimed; the other P steps are timed. Every stream is then checked against the oracle's
    digests of its solo encode (tests/golden/batch_config3.json) when the workload is that one."""
    import torch
    import riv2_amd
    from riv2_amd import synth

    dev = torch.device("cuda:0")
    params = lambda w, h, n, seed: synth.custom_params(w, h, n, seed=seed, global_motion=3.0, n_rects=4, rect_min=32, rect_max=96,
                                                         rect_velocity=6.0)
    streams, frames = [], []
    for k, (w, h) in enumerate(BATCH_SIZES):
        for j in range(per_size):
            seq = synth.Sequence(params(w, h, 1 + p_frames, 1000 + 100 * k + j))
            frames.append(torch.from_numpy(seq.frames(0, 1 + p_frames)).to(dev))
            streams.append((w, h))
    torch.cuda.synchronize()
    encs = [riv2_amd.Encoder(_batch_config(riv2_amd, w, h, depth)) for w, h in streams]
    steps, first, elapsed, search_ms = _timed_batch(encs, frames, p_frames)
    tuning = encs[0].batch_tuning()
    outs = []
    for e in encs:
        e.finish()
        outs.append(e.output())
        e.close()
    del frames
    checked = None
    gpath = os.path.join(GOLDEN, "batch_config3.json")
    if per_size == 16 and p_frames == 20 and os.path.exists(gpath):
        with open(gpath) as f:
            want = json.load(f)["streams"]
        checked = all(len(o) == x["stream_bytes"] and sha256(o) == x["stream_sha256"] for o, x in zip(outs, want))
    cpu = None
    if with_cpu:
        cores = host_cores()
        print(f"[bench] batched: CPU baseline, oracle on {cores} threads", file=sys.stderr, flush=True)
        v = _cpu_baseline_mix(lambda w, h, n: synth.Sequence(params(w, h, n, 999)).frames(0, n), cores)
        cpu = {"value": round(v, 3), "unit": "fps", "cores": cores, "kind": "port",
               "sample": "one stream of each size, 1 I + 2 P-frames each, oracle C++ restatement of the reference encoder with its "
                         f"`threads` strip ME on {cores} threads ({cpu_model()}); P-frames/s over the size mix"}
    mbs = _storage_luma_bytes(streams) // 256
    return {"metric": "batched P-frame encode fps/GPU (64 mixed-resolution streams, one batch step per frame of every stream)",
            "value": round(steps * len(streams) / elapsed, 3), "unit": "fps", "n_gpus": 1, "steps": steps, "warmup": first,
            "ms_per_step": round(1000 * elapsed / steps, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u8", "data": "synthetic",
            "config": {"workload": f"{len(streams)} streams ({per_size} each of " + ", ".join(f"{w}x{h}" for w, h in BATCH_SIZES) +
                                   f"), 1 I + {p_frames} P each, q85/80, sr 12, inputs resident in HBM", "pipeline_depth": depth,
                       "bit_exact_vs_oracle": checked, "streams_checked_vs_oracle": len(outs) if checked is not None else 0},
            "roofline": _batch_roofline(streams, search_ms), "cpu_baseline": cpu,
            "fps_1080p_equivalent": round(steps * mbs / 8160 / elapsed, 3), "stream_bytes": sum(len(o) for o in outs),
            "search_shape": tuning}


def run_stress(per_size, p_frames, depth, with_cpu):
    """The learned + batched stress workload: per_size streams of each of BATCH_SIZES, each 1 I +
    p_frames P-frames of high-motion content (riv2_amd.synth.stress_params: global motion up to
    24 px, rectangles up to 32 px per frame that zoom and rotate, a scene cut every 25 frames,
    noise sigma 4), encoded by batch steps twice, plain and learned, frames resident in HBM. The
    P steps after the autotuner has settled are timed. Reports both forms' fps, the learned form's
    fallback rate, bytes and PSNR-Y against plain, and decodes every learned stream with the GPU
    decoder and the oracle decoder."""
    import torch
    import riv2_amd
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    from riv2_amd import synth, learned

    dev = torch.device("cuda:0")
    streams, host = [], []
    for k, (w, h) in enumerate(BATCH_SIZES):
        for j in range(per_size):
            host.append(synth.Sequence(synth.stress_params(w, h, 1 + p_frames, seed=5000 + 100 * k + j)).frames(0, 1 + p_frames))
            streams.append((w, h))
    frames = [torch.from_numpy(f).to(dev) for f in host]
    torch.cuda.synchronize()
    net = learned.MvNet.load(learned.WEIGHTS_PATH)

    def encode(with_net):
        encs = [riv2_amd.Encoder(_batch_config(riv2_amd, w, h, depth)) for w, h in streams]
        if with_net:
            for e in encs:
                e.set_mvnet(net, 0.5)
                e.set_learned_passes(LEARNED_PASSES)
        steps, _, elapsed, search_ms = _timed_batch(encs, frames, p_frames)
        counts = encs[0].learned_stats()
        outs = []
        for e in encs:
            e.finish()
            outs.append(e.output())
            e.close()
        return steps * len(encs) / elapsed, outs, counts, search_ms

    print("[bench] stress: plain", file=sys.stderr, flush=True)
    plain_fps, plain_out, _, plain_ms = encode(False)
    print("[bench] stress: learned", file=sys.stderr, flush=True)
    learned_fps, learned_out, counts, learned_ms = encode(True)
    net.close()
    del frames

    def psnr(outs):
        vals = []
        for f, data in zip(host, outs):
            dec = riv2_amd.Decoder(data)
            for k in range(len(f)):
                d = dec.decode_frame().data
                if k > 0:
                    vals.append(learned.psnr_y(f[k], d))
            dec.close()
        return sum(vals) / len(vals)

    print("[bench] stress: PSNR-Y of every decoded P-frame", file=sys.stderr, flush=True)
    p_plain, p_learned = psnr(plain_out), psnr(learned_out)
    print("[bench] stress: every learned stream through the oracle decoder", file=sys.stderr, flush=True)

    def oracle_digests(data):
        rdec = oracle.OracleDecoder(data)
        return [sha256(np.ascontiguousarray(rdec.decode()[0]).tobytes()) for _ in range(1 + p_frames)]

    with ThreadPoolExecutor(16) as ex:   # the oracle's ctypes calls release the GIL
        want = list(ex.map(oracle_digests, learned_out))
    exact = True
    for data, wd in zip(learned_out, want):
        dec = riv2_amd.Decoder(data)
        exact &= all(sha256(np.ascontiguousarray(dec.decode_frame().data).tobytes()) == d for d in wd)
        dec.close()
    cpu = None
    if with_cpu:
        cores = host_cores()
        print(f"[bench] stress: CPU baseline, oracle on {cores} threads", file=sys.stderr, flush=True)
        v = _cpu_baseline_mix(lambda w, h, n: synth.Sequence(synth.stress_params(w, h, n, seed=4999)).frames(0, n), cores)
        cpu = {"value": round(v, 3), "unit": "fps", "cores": cores, "kind": "port",
               "sample": "one high-motion stream of each size, 1 I + 2 P-frames each, oracle C++ restatement of the reference "
                         f"encoder (plain form) with its `threads` strip ME on {cores} threads ({cpu_model()}); P-frames/s over the size mix"}
    pb, lb = sum(len(o) for o in plain_out), sum(len(o) for o in learned_out)
    return {"metric": "learned + batched stress: P-frame encode fps/GPU (64 mixed-resolution streams, high motion)",
            "value": round(learned_fps, 3), "unit": "fps", "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u8", "data": "synthetic",
            "config": {"workload": f"{len(streams)} streams ({per_size} each of " + ", ".join(f"{w}x{h}" for w, h in BATCH_SIZES) +
                                   f"), 1 I + {p_frames} P each, high motion (global 24 px, rectangles 32 px/frame, zoom 0.5%, "
                                   "rotation 1 deg per frame, scene cut every 25 frames, sigma 4), q85/80, sr 12", "pipeline_depth": depth},
            "roofline": _batch_roofline(streams, learned_ms), "cpu_baseline": cpu,
            "plain_fps": round(plain_fps, 3), "learned_vs_plain_fps": round(learned_fps / plain_fps, 4),
            "plain_search_avg_ms": round(plain_ms, 4),
            "fallback_rate": round(1.0 - counts[2] / max(1, counts[0]), 4), "confident_rate": round(counts[1] / max(1, counts[0]), 4),
            "bytes_vs_plain": round(lb / pb, 5), "psnr_y": round(p_learned, 4), "plain_psnr_y": round(p_plain, 4),
            "decode_bit_exact_vs_oracle_decoder": bool(exact), "decode_checked_streams": len(learned_out)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=45)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-pframes", type=int, default=8)
    ap.add_argument("--pipeline-depth", type=int, default=32,
                    help="frames in flight (GPU analysis of frame k+1 overlaps host entropy coding of earlier frames)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned form's line")
    ap.add_argument("--cpu-decode-frames", type=int, default=6, help="P-frames the oracle decoder is timed on")
    ap.add_argument("--batched", action="store_true", help="measure the batched 64-stream workload instead")
    ap.add_argument("--batch-per-size", type=int, default=16)
    ap.add_argument("--batch-p-frames", type=int, default=20)
    ap.add_argument("--stress", action="store_true", help="with --batched: the learned + batched high-motion workload")
    ap.add_argument("--stress-p-frames", type=int, default=49)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain 256 MiB device copy kernel (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box here has a single MI355X; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is described in DESIGN.md, not measured).", file=sys.stderr)
        sys.exit(2)

    if args.batched and args.stress:
        print(f"[bench] stress: {4 * args.batch_per_size} streams, 1 I + {args.stress_p_frames} P each", file=sys.stderr, flush=True)
        print(json.dumps(run_stress(args.batch_per_size, args.stress_p_frames, max(2, min(args.pipeline_depth, 4)), not args.no_cpu_baseline)))
        return
    if args.batched:
        print(f"[bench] batched: {4 * args.batch_per_size} streams, 1 I + {args.batch_p_frames} P each", file=sys.stderr, flush=True)
        print(json.dumps(run_batched(args.batch_per_size, args.batch_p_frames, max(2, min(args.pipeline_depth, 4)), not args.no_cpu_baseline)))
        return

    from riv2_amd import synth
    nframes = 1 + args.warmup + args.steps
    print(f"[bench] generating {nframes} 1080p frames", file=sys.stderr, flush=True)
    frames = synth.Sequence(synth.hd_params(seed=args.seed, frames=max(nframes, 50))).frames(0, nframes)

    print(f"[bench] GPU encode: 1 I + {args.warmup} warmup + {args.steps} timed P-frames", file=sys.stderr, flush=True)
    elapsed, stats, stages, stream, d_frames, _ = run_gpu(frames, args.steps, args.warmup, args.pipeline_depth)
    stage_avg, gpu_frame_ms = stage_breakdown(d_frames, min(nframes, 12), args.pipeline_depth)
    fps = args.steps / elapsed
    learned_leg = None
    if not args.no_learned:
        print("[bench] learned form on the same frames", file=sys.stderr, flush=True)
        learned_leg = run_learned(frames, d_frames, args.steps, args.warmup, args.pipeline_depth, stream, fps)
    del d_frames
    ms_per_step = 1000.0 * elapsed / args.steps

    me_ms, me_n = stages["motion_search"]
    me_avg_ms = me_ms / max(1, me_n)
    me_bytes = 2 * SW * SH   # current + reference luma planes, each read once
    achieved = me_bytes / (me_avg_ms * 1e-3) / 1e9 if me_avg_ms > 0 else 0.0

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "motion_search_1080p_traffic.json")
    if os.path.exists(tpath):
        try:
            with open(tpath) as f:
                tj = json.load(f)
            # only a measurement of the current kernels counts
            traffic = tj.get("hbm_bytes_per_launch") if tj.get("kernels") == ME_KERNELS else None
        except Exception:
            traffic = None

    if args.copy_calibration:
        import torch
        src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda:0").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    print("[bench] every frame against the oracle's digests", file=sys.stderr, flush=True)
    fixture = check_against_fixture(args.seed, stream, nframes)
    if fixture is None:
        checked, bit_exact = 3, check_prefix_bit_exact(frames, stream, 3)
    else:
        checked, bit_exact = fixture
    print("[bench] drop-in boundary: host RGB in", file=sys.stderr, flush=True)
    drop_fps, drop_stream = run_drop_in(frames, args.steps, args.warmup, args.pipeline_depth)
    drop_same = drop_stream == stream
    pin_fps, pin_stream = run_drop_in(frames, args.steps, args.warmup, args.pipeline_depth, pinned=True)
    drop_same = drop_same and pin_stream == stream
    print("[bench] decode", file=sys.stderr, flush=True)
    decode = run_decode(stream, nframes, max(1, min(args.cpu_decode_frames, nframes - 1)))

    cpu = None
    if not args.no_cpu_baseline:
        cores = host_cores()
        print(f"[bench] CPU baseline: oracle, {args.cpu_pframes} P-frames, {cores} threads", file=sys.stderr, flush=True)
        cpu_fps = cpu_baseline(frames, args.cpu_pframes, cores)
        cpu = {"value": round(cpu_fps, 3), "unit": "fps", "cores": cores, "kind": "port",
               "sample": f"1920x1080 synthetic, 1 I + {args.cpu_pframes} P-frames timed, oracle C++ restatement of the "
                         f"reference encoder with its `threads` strip ME on {cores} threads ({cpu_model()})"}

    out = {
        "metric": "1080p P-frame encode fps/GPU: plain bit-exact vs learned-MV at ≤1% bitrate loss",
        "value": round(fps, 3),
        "unit": "fps",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "1080p P-frame encode, plain form: 1920x1080 RGB24, q85/80, search range 12, skip 3, "
                               "1 I + P-frames, inputs resident in HBM", "width": W, "height": H, "form": "plain",
                   "pipeline_depth": args.pipeline_depth, "bit_exact_vs_oracle": bool(bit_exact),
                   "frames_checked_vs_oracle": checked},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                     "kernel": "motion search (" + " + ".join(ME_KERNELS) + ", one frame)",
                     "algorithmic_bytes_per_launch": me_bytes, "avg_launch_ms": round(me_avg_ms, 4)},
        "cpu_baseline": cpu,
        "gpu_stage_avg_ms": stage_avg,
        "gpu_frame_avg_ms": round(gpu_frame_ms, 4) if gpu_frame_ms else None,
        "host_entropy_avg_ms": round(sum(s.host_entropy_ms for s in stats) / len(stats), 4),
        "host_tail_avg_ms": round(sum(s.host_tail_ms for s in stats) / len(stats), 4),
        "me_passes_avg": round(sum(s.me_passes for s in stats) / len(stats), 2),
        "bytes_per_pframe": round(sum(s.total_bytes for s in stats) / len(stats), 1),
        "learned": learned_leg,
        "drop_in": {"value": round(drop_fps, 3), "unit": "fps", "same_stream": bool(drop_same),
                    "pinned_value": round(pin_fps, 3),
                    "note": "riv_encode_frame with host RGB24 in pageable memory (the reference API's input), same frames "
                            "and timing; the H2D copy of each frame is inside the timed region. pinned_value: the same "
                            "with the frames in pinned host memory (the copy is then asynchronous for the caller)"},
        "decode": decode,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(fps / cpu["value"], 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
