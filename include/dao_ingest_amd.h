/*
 * dao_ingest_amd.h -- frame ingest entry points of libdao_fit.so.
 *
 * The reference prepares every frame on the host before the fitter sees it
 * (storm_analysis/sa_library/fitting.py:934-937): the movie reader returns
 * uint16 counts (DaxReader.loadAFrame, storm_analysis/sa_library/datareader.py:234-243,
 * byte-swapped for big-endian movies), the frame reader converts them to
 * photo-electrons (FrameReader.loadAFrame, storm_analysis/sa_library/analysis_io.py:243-260)
 * and padArray (storm_analysis/sa_library/fitting.py:112-132) mirror-pads the
 * result by the finder's margin.  These functions do the same for a batch of
 * frames on the GPU, with the reference's arithmetic, and leave the result in
 * device memory, where daoBatchNewImages (on_device = 1) takes it.
 *
 * Frames are row-major, height x width, concatenated.  Output frames are
 * (height + 2 pad) x (width + 2 pad) float64 in device memory; pad must not
 * exceed width or height (the reference's mirror needs as many rows and
 * columns as it copies), and a call takes at most 65535 frames.  Calls are
 * synchronous.  Return 0 on success, -1 for
 * bad sizes or null pointers, -2 for a device error.
 */
#ifndef DAO_INGEST_AMD_H
#define DAO_INGEST_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* n raw frames -> (raw - offset) * (gain * rqe), values below 1.0 set to 1.0,
 * then mirror-padded.  offset, gain and rqe are the reader's attributes
 * (analysis_io.py:263-295: gain = 1 / camera gain in ADU per photo-electron,
 * rqe = 1 / relative quantum efficiency): single values, or height x width
 * maps when per_pixel (sCMOS).  raw (and the calibration) may be host or
 * device arrays (raw_on_device, calib_on_device); big_endian swaps the two
 * bytes of every count first. */
int daoIngestFrames(const uint16_t *raw, int raw_on_device, int n, int width, int height, int big_endian,
                    const double *offset, const double *gain, const double *rqe, int per_pixel,
                    int calib_on_device, int pad, double *out);

/* padArray (fitting.py:112-132) of n float64 frames (backgrounds, variance
 * maps) into device memory. */
int daoPadFrames(const double *in, int in_on_device, int n, int width, int height, int pad, double *out);

#ifdef __cplusplus
}
#endif

#endif /* DAO_INGEST_AMD_H */
