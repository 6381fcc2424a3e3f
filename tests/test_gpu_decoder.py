"""The decoder's look-ahead (in-memory streams: records parsed and vectors decoded in order on one
thread, residuals decoded ahead on worker threads): output equal to the oracle decoder's, a decoder
closed while the look-ahead waits for free slots, and a stream cut inside a later frame (the
frames before it decode, that frame reports EndOfStream)."""
import time

import numpy as np
import pytest

import oracle
import riv2_amd
from riv2_amd import synth

pytestmark = pytest.mark.gpu


def _stream(n=20, seed=61):
    frames = synth.Sequence(synth.cif_params(seed=seed, frames=n)).frames()
    cfg = riv2_amd.EncoderConfig(352, 288, 30).with_intra_quality(85).with_inter_quality(80) \
        .with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(7))
    enc = riv2_amd.Encoder(cfg)
    for i, f in enumerate(frames):
        enc.encode_frame(riv2_amd.Frame(f, 352, 288, 33 * i))
    enc.finish()
    return enc.output(), frames


def test_lookahead_output_equals_the_oracle_decoder():
    print("[decoder] look-ahead vs oracle", flush=True)
    data, frames = _stream()
    dec, rdec = riv2_amd.Decoder(data), oracle.OracleDecoder(data)
    for k in range(len(frames)):
        a = dec.decode_frame()
        b, _ = rdec.decode()
        assert a.frame_index == k and np.array_equal(a.data, b), f"frame {k}"
    assert not dec.has_more_frames()
    dec.close()


def test_close_while_the_lookahead_waits():
    print("[decoder] early close", flush=True)
    data, _ = _stream()
    dec = riv2_amd.Decoder(data)
    dec.decode_frame()
    dec.decode_frame()
    time.sleep(0.2)   # the look-ahead fills its slots and waits
    t0 = time.perf_counter()
    dec.close()
    assert time.perf_counter() - t0 < 5.0


def test_stream_cut_inside_a_later_frame():
    print("[decoder] truncated stream", flush=True)
    data, frames = _stream(12, seed=62)
    # the end of frame 5's record: walk the records (format.rs:171-205)
    pos = 36
    ends = []
    for _ in range(len(frames)):
        ft = data[pos + 8]
        if ft == 1:
            n = int.from_bytes(data[pos + 10:pos + 14], "little")
            pos += 14 + n
        else:
            m = int.from_bytes(data[pos + 13:pos + 17], "little")
            r = int.from_bytes(data[pos + 17 + m:pos + 21 + m], "little")
            pos += 21 + m + r
        ends.append(pos)
    assert ends[-1] == len(data)
    cut = data[:ends[5] - 3]
    dec, rdec = riv2_amd.Decoder(cut), oracle.OracleDecoder(data)
    for k in range(5):
        a = dec.decode_frame()
        b, _ = rdec.decode()
        assert np.array_equal(a.data, b), f"frame {k}"
    for _ in range(2):   # the failed frame keeps failing
        with pytest.raises(riv2_amd.DecodeError) as ei:
            dec.decode_frame()
        assert ei.value.kind == "EndOfStream"
    dec.close()
