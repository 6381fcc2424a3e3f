"""Error behaviour of the C ABI, mirroring EncodeError / DecodeError of the reference
(reitero_encode/src/error.rs:4-19, reitero_decode/src/error.rs:5-20).

Header validation runs before any device is touched, so those cases run on the CPU; the rest need
an encoder or decoder on the GPU and are marked accordingly."""
import struct

import numpy as np
import pytest

import oracle
import riv2_amd
from riv2_amd import synth


def _small_stream(frames=3, w=64, h=48):
    """A valid .riv stream from the oracle (CPU)."""
    seq = synth.Sequence(synth.custom_params(w, h, frames, seed=2, global_motion=2.0, n_rects=1, rect_min=8, rect_max=16,
                                             rect_velocity=3.0)).frames()
    enc = oracle.OracleEncoder(oracle.config(w, h))
    for i, f in enumerate(seq):
        enc.encode(f, i * 33)
    return enc.finish(), seq


def _decode_error(data):
    with pytest.raises(riv2_amd.DecodeError) as ei:
        riv2_amd.Decoder(data)
    return ei.value.kind


def test_header_errors_on_cpu():
    # Decoder::new reads and validates the 36-byte header (decoder.rs:112-171)
    data, _ = _small_stream()
    assert data[:4] == b"RIV\0" and struct.unpack_from("<I", data, 4)[0] == 5
    assert _decode_error(data[:20]) == "EndOfStream"                        # truncated header
    assert _decode_error(b"RIF\0" + data[4:]) == "InvalidHeader"            # magic
    assert _decode_error(data[:4] + struct.pack("<I", 4) + data[8:]) == "InvalidHeader"   # version
    bad = bytearray(data)
    struct.pack_into("<I", bad, 16, 63)                                      # storage width not a multiple of 16
    assert _decode_error(bytes(bad)) == "InvalidHeader"
    bad = bytearray(data)
    struct.pack_into("<I", bad, 8, 4096)                                     # display wider than storage
    assert _decode_error(bytes(bad)) == "InvalidHeader"


def test_config_errors_on_cpu():
    # EncoderConfig::validate (config.rs:129-164): zero sizes, zero fps
    for w, h, fps in ((0, 48, 30), (64, 0, 30), (64, 48, 0)):
        with pytest.raises(riv2_amd.EncodeError) as ei:
            riv2_amd.EncoderConfig(w, h, fps).validate()
        assert ei.value.kind == "InvalidConfig"


@pytest.mark.gpu
def test_encoder_frame_and_config_errors():
    cfg = riv2_amd.EncoderConfig(64, 48, 30)
    enc = riv2_amd.Encoder(cfg)
    # Frame of another size: EncodeError::FrameError (encoder.rs:237-246)
    with pytest.raises(riv2_amd.EncodeError) as ei:
        enc.encode_frame(riv2_amd.Frame(np.zeros((32, 64, 3), np.uint8), 64, 32, 0))
    assert ei.value.kind == "FrameError"
    # the encoder stays usable after a rejected frame
    enc.encode_frame(riv2_amd.Frame(np.zeros((48, 64, 3), np.uint8), 64, 48, 0))
    enc.finish()
    assert enc.frame_count() == 1
    enc.close()
    # a valid reference config outside this backend's limits is refused, not mis-encoded
    with pytest.raises(riv2_amd.EncodeError) as ei:
        riv2_amd.Encoder(riv2_amd.EncoderConfig(64, 48, 30).with_search_range(64))
    assert ei.value.kind == "Unsupported"


@pytest.mark.gpu
def test_decoder_stream_errors():
    data, seq = _small_stream()
    dec = riv2_amd.Decoder(data)
    for _ in range(len(seq)):
        dec.decode_frame()
    assert not dec.has_more_frames()
    # past the last frame: DecodeError::EndOfStream (decoder.rs:175-177)
    with pytest.raises(riv2_amd.DecodeError) as ei:
        dec.decode_frame()
    assert ei.value.kind == "EndOfStream"
    dec.close()
    # a frame record with an unknown frame type: DecodeError::InvalidFrame
    bad = bytearray(data)
    bad[36 + 8] = 7
    dec = riv2_amd.Decoder(bytes(bad))
    with pytest.raises(riv2_amd.DecodeError) as ei:
        dec.decode_frame()
    assert ei.value.kind == "InvalidFrame"
    dec.close()
    # a stream cut inside the first frame's payload: EndOfStream
    dec = riv2_amd.Decoder(data[:36 + 20])
    with pytest.raises(riv2_amd.DecodeError) as ei:
        dec.decode_frame()
    assert ei.value.kind == "EndOfStream"
    dec.close()
