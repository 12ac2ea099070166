"""
Pins of the frame-ingest oracle (oracle/ingest_oracle.py) against closed forms
and numpy's own routines (CPU).  The oracle restates the reference's frame
preparation: DaxReader.loadAFrame (datareader.py:234-243), FrameReader.loadAFrame
(analysis_io.py:243-260) and padArray (fitting.py:112-132).
"""
import os
import sys

import numpy
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import ingest_oracle as IO  # noqa: E402


@pytest.mark.parametrize("shape,pad", [((6, 9), 0), ((6, 9), 1), ((6, 9), 4), ((6, 9), 6), ((1, 1), 1),
                                       ((13, 5), 5)])
def test_pad_array_is_the_symmetric_mirror(shape, pad):
    """padArray repeats the edge pixel: numpy.pad's 'symmetric' mode, not 'reflect'."""
    a = numpy.random.default_rng(3).uniform(size=shape)
    got = IO.pad_array(a, pad)
    numpy.testing.assert_array_equal(got, numpy.pad(a, pad, mode="symmetric"))
    if pad > 0 and min(shape) > 1:
        assert not numpy.array_equal(got, numpy.pad(a, pad, mode="reflect", reflect_type="even"))


def test_pad_array_corner_by_hand():
    a = numpy.array([[1.0, 2.0], [3.0, 4.0]])
    want = numpy.array([[4, 3, 3, 4, 4, 3],
                        [2, 1, 1, 2, 2, 1],
                        [2, 1, 1, 2, 2, 1],
                        [4, 3, 3, 4, 4, 3],
                        [4, 3, 3, 4, 4, 3],
                        [2, 1, 1, 2, 2, 1]], dtype=float)
    numpy.testing.assert_array_equal(IO.pad_array(a, 2), want)


def test_load_a_frame_by_hand():
    """(frame - offset) * (gain * rqe); values below 1.0 become 1.0, exactly 1.0 stays."""
    frame = numpy.array([[100, 102], [150, 99]], dtype=numpy.uint16)
    got = IO.load_a_frame(frame, 100.0, 0.5, 1.0)
    numpy.testing.assert_array_equal(got, [[1.0, 1.0], [25.0, 1.0]])
    # sCMOS maps: per-pixel offset, gain and rqe
    off = numpy.array([[90.0, 100.0], [100.0, 110.0]])
    gain = numpy.array([[0.5, 0.25], [1.0, 2.0]])
    rqe = numpy.array([[1.0, 2.0], [0.5, 1.0]])
    got = IO.load_a_frame(frame, off, gain, rqe)
    numpy.testing.assert_array_equal(got, [[5.0, 1.0], [25.0, 1.0]])


def test_raw_counts_byte_order():
    """A big-endian movie's bytes read as native uint16 and swapped equal numpy's '>u2' view."""
    b = numpy.random.default_rng(5).integers(0, 256, size=64, dtype=numpy.uint8).tobytes()
    native = numpy.frombuffer(b, dtype=numpy.uint16)
    numpy.testing.assert_array_equal(IO.raw_counts(native, True), numpy.frombuffer(b, dtype=">u2"))
    numpy.testing.assert_array_equal(IO.raw_counts(native, False), numpy.frombuffer(b, dtype="<u2"))
    assert IO.raw_counts(numpy.array([0x0201], dtype=numpy.uint16), True)[0] == 0x0102


def test_ingest_composes_the_three_steps():
    rng = numpy.random.default_rng(7)
    raw = rng.integers(90, 400, size=(3, 5, 8), dtype=numpy.uint16)
    got = IO.ingest(raw, 100.0, 0.5, 1.0, 2)
    assert got.shape == (3, 9, 12)
    for k in range(3):
        conv = numpy.maximum((raw[k].astype(float) - 100.0) * 0.5, 1.0)
        numpy.testing.assert_array_equal(got[k], numpy.pad(conv, 2, mode="symmetric"))
