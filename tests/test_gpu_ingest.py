"""
GPU frame ingest (daoIngestFrames / daoPadFrames through
storm_analysis_amd.sa_library.dao_ingest) against the oracle's restatement of
the reference's frame preparation (oracle/ingest_oracle.py): raw counts ->
photo-electrons -> mirror padding, bit-identical.
"""
import os
import sys

import numpy
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import ingest_oracle as IO  # noqa: E402

from storm_analysis_amd.sa_library import dao_ingest  # noqa: E402
from storm_analysis_amd.sa_library.dao_fit_c import MultiFitterException  # noqa: E402


def raw_frames(rng, n, h, w, lo=80, hi=5000):
    return rng.integers(lo, hi, size=(n, h, w), dtype=numpy.uint16)


@pytest.mark.gpu
@pytest.mark.parametrize("h,w,pad,big_endian", [(40, 56, 0, False), (40, 56, 7, False), (37, 23, 23, True),
                                                (64, 64, 10, True), (1, 9, 1, False)])
def test_std_camera_frames_match_oracle(h, w, pad, big_endian):
    rng = numpy.random.default_rng(1000 + h + w + pad)
    raw = raw_frames(rng, 4, h, w)
    reader = dao_ingest.FrameReaderStd(camera_gain=2.0, camera_offset=100.0)
    got = reader.loadFrames(raw, pad=pad, big_endian=big_endian).cpu().numpy()
    want = IO.ingest(raw, reader.offset, reader.gain, reader.rqe, pad, big_endian)
    numpy.testing.assert_array_equal(got, want)
    if pad:
        assert (got >= 1.0).all()


@pytest.mark.gpu
def test_scmos_maps_and_device_input_match_oracle():
    import torch
    rng = numpy.random.default_rng(2024)
    h, w, pad = 48, 33, 9
    raw = raw_frames(rng, 6, h, w, lo=0, hi=65535)
    offset = rng.normal(100.0, 5.0, size=(h, w))
    gain = rng.normal(2.0, 0.1, size=(h, w))
    rqe = rng.uniform(0.8, 1.2, size=(h, w))
    reader = dao_ingest.FrameReaderSCMOS(offset, gain, rqe)
    want = IO.ingest(raw, reader.offset, reader.gain, reader.rqe, pad, True)
    got = reader.loadFrames(raw, pad=pad, big_endian=True).cpu().numpy()
    numpy.testing.assert_array_equal(got, want)
    # counts already in device memory
    draw = torch.from_numpy(raw.astype(numpy.int32)).to(torch.uint16).cuda()
    got = reader.loadFrames(draw, pad=pad, big_endian=True).cpu().numpy()
    numpy.testing.assert_array_equal(got, want)


@pytest.mark.gpu
def test_pad_arrays_match_oracle():
    rng = numpy.random.default_rng(77)
    frames = rng.uniform(-3.0, 300.0, size=(3, 20, 31))
    for pad in (0, 1, 20):
        got = dao_ingest.padArrays(frames, pad).cpu().numpy()
        want = numpy.array([IO.pad_array(f, pad) for f in frames])
        numpy.testing.assert_array_equal(got, want)


@pytest.mark.gpu
def test_bad_sizes_raise():
    reader = dao_ingest.FrameReaderStd(camera_gain=1.0, camera_offset=0.0)
    with pytest.raises(MultiFitterException):
        reader.loadFrames(numpy.zeros((1, 4, 5), dtype=numpy.uint16), pad=5)


@pytest.mark.gpu
def test_ingested_frames_feed_the_fitter_without_host_copy():
    """Device-ingested frames fit exactly as the same frames prepared on the host."""
    from storm_analysis_amd.sa_library import dao_batch
    from storm_analysis_amd.diagnostics import synth
    rng = numpy.random.default_rng(synth.SEED_BASE + 60)
    fr = synth.frame_2d_dense(rng, size=64, n=40)
    # turn the synthetic photo-electron frame into camera counts (gain 2 ADU/e-, offset 100)
    raw = numpy.clip(numpy.round(fr.image * 2.0 + 100.0), 0, 65535).astype(numpy.uint16)[None]
    reader = dao_ingest.FrameReaderStd(camera_gain=2.0, camera_offset=100.0)
    dev = reader.loadFrames(raw, pad=0)
    host = IO.ingest(raw, reader.offset, reader.gain, reader.rqe, 0)
    kw = {"roi_size": 18, "sigma_range": [0.96, 8.0]}
    out = []
    for images in (dev, [host[0]]):
        bf = dao_batch.BatchFitter([(64, 64)], [40], model="2d", **kw)
        try:
            bf.newImages(images)
            bf.newBackgrounds([numpy.full((64, 64), 100.0)])
            bf.newPeaks([fr.peaks], "hdf5")
            bf.doFit()
            out.append(bf.getPeaks(0))
        finally:
            bf.cleanup()
    numpy.testing.assert_array_equal(out[0]["status"], out[1]["status"])
    numpy.testing.assert_array_equal(out[0]["x"], out[1]["x"])
    numpy.testing.assert_array_equal(out[0]["height"], out[1]["height"])
