"""A motion search that does not settle fails its frame: forced from the host by capping the
relaxation passes (the kernel still exits normally), the frame returns RIV_E_GPU and not one byte
of it reaches the writer, in synchronous and pipelined mode, for a single encoder and for a batch
step."""
import pytest

import riv2_amd
from riv2_amd import synth

pytestmark = pytest.mark.gpu

W, H = 352, 288


def _frames(n=3):
    return synth.Sequence(synth.cif_params(seed=11, frames=n)).frames()


def _cfg(depth):
    return riv2_amd.EncoderConfig(W, H, 30).with_intra_quality(85).with_inter_quality(80) \
        .with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(1000)).with_pipeline_depth(depth)


class CountingWriter(riv2_amd.VecWriter):
    def __init__(self):
        super().__init__()
        self.writes = []

    def write(self, data):
        self.writes.append(len(data))
        super().write(data)


@pytest.mark.parametrize("depth", [0, 3])
def test_unsettled_search_writes_no_bytes(depth):
    frames = _frames()
    w = CountingWriter()
    enc = riv2_amd.Encoder(_cfg(depth), writer=w)
    print(f"[failure depth {depth}] I-frame", flush=True)
    enc.encode_frame(riv2_amd.Frame(frames[0], W, H, 0))
    enc.flush()
    after_intra = len(w.buffer)
    assert after_intra > 36
    enc.debug_pass_limit(1)   # one pass cannot settle a moving frame's field
    print(f"[failure depth {depth}] capped P-frame", flush=True)
    with pytest.raises(riv2_amd.EncodeError) as ei:
        enc.encode_frame(riv2_amd.Frame(frames[1], W, H, 33))
        enc.flush()
    assert ei.value.kind == "Gpu" and "pass limit" in str(ei.value)
    assert len(w.buffer) == after_intra, "bytes of the failed frame reached the writer"
    with pytest.raises(riv2_amd.EncodeError):   # the encoder stays failed
        enc.encode_frame(riv2_amd.Frame(frames[2], W, H, 66))
    enc.close()


def test_unsettled_batch_search_fails_every_inter_stream():
    frames = _frames()
    writers = [CountingWriter(), CountingWriter()]
    encs = [riv2_amd.Encoder(_cfg(2), writer=wr) for wr in writers]
    riv2_amd.encode_batch_step(encs, [frames[0], frames[0]], [0, 0])
    for e in encs:
        e.flush()
    sizes = [len(wr.buffer) for wr in writers]
    encs[0].debug_pass_limit(1)   # the first encoder leads the batch's one search launch
    print("[failure batch] capped P step", flush=True)
    try:
        riv2_amd.encode_batch_step(encs, [frames[1], frames[1]], [33, 33])
    except riv2_amd.EncodeError:
        pass
    for e, n in zip(encs, sizes):
        with pytest.raises(riv2_amd.EncodeError):
            e.flush()
    assert [len(wr.buffer) for wr in writers] == sizes
    for e in encs:
        e.close()
