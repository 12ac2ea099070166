"""
Frame-ingest oracle -- test infrastructure only.

A numpy restatement of the reference's per-frame preparation, used by tests/
(and nothing else) to check the device ingest kernel (daoIngestFrames /
daoPadFrames) bit for bit:

- raw counts as DaxReader.loadAFrame returns them
  (storm_analysis/sa_library/datareader.py:234-243): uint16, byte-swapped
  for big-endian movies;
- FrameReader.loadAFrame (storm_analysis/sa_library/analysis_io.py:243-260):
  (frame - offset) * (gain * rqe), then every value below 1.0 set to 1.0;
  FrameReaderStd sets gain = 1 / camera_gain, rqe = 1 (analysis_io.py:263-278),
  FrameReaderSCMOS gain = 1 / gain map, rqe = 1 / rqe map (analysis_io.py:281-295);
- padArray (storm_analysis/sa_library/fitting.py:112-132): rows mirrored
  (numpy.flipud of the rows next to each edge), then columns (numpy.fliplr),
  into an array initialised with ones.
"""
import numpy


def raw_counts(buf, big_endian):
    """uint16 counts of a frame buffer read as the movie's byte order (datareader.py:239-242)."""
    a = numpy.asarray(buf, dtype=numpy.uint16)
    if big_endian:
        a = a.byteswap()
    return a


def load_a_frame(frame, offset, gain, rqe):
    """FrameReader.loadAFrame's conversion (analysis_io.py:243-260)."""
    frame = (frame - offset) * (gain * rqe)
    mask = (frame < 1.0)
    if numpy.sum(mask) > 0:
        frame[mask] = 1.0
    return frame


def pad_array(ori_array, pad_size):
    """padArray (fitting.py:112-132)."""
    if pad_size > 0:
        [x_size, y_size] = ori_array.shape
        lg = numpy.ones((x_size + 2 * pad_size, y_size + 2 * pad_size))
        lg[pad_size:(x_size + pad_size), pad_size:(y_size + pad_size)] = ori_array.astype(numpy.float64)
        lg[0:pad_size, :] = numpy.flipud(lg[pad_size:2 * pad_size, :])
        lg[(x_size + pad_size):(x_size + 2 * pad_size), :] = numpy.flipud(lg[x_size:(x_size + pad_size), :])
        lg[:, 0:pad_size] = numpy.fliplr(lg[:, pad_size:2 * pad_size])
        lg[:, (y_size + pad_size):(y_size + 2 * pad_size)] = numpy.fliplr(lg[:, y_size:(y_size + pad_size)])
        return lg
    return ori_array


def ingest(raw, offset, gain, rqe, pad, big_endian=False):
    """n raw frames (n x H x W uint16) -> n padded photo-electron frames, as the reference."""
    out = []
    for f in raw:
        out.append(pad_array(load_a_frame(raw_counts(f, big_endian), offset, gain, rqe), pad))
    return numpy.array(out)
