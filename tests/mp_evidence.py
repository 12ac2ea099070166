"""
Evidence for the multiplane pins (test_mp_pins.py), sharing no code with
the multiplane oracle (oracle/mp_oracle.c) or the HIP library: per channel,
the model, deviance, J and Gauss-Newton H of one peak come from
lm_evidence.Frame (closed forms and complex steps), the damped solve from
numpy.linalg.solve.  What it adds is the channel geometry: affine maps
between channel 0 and channel c (x_c = a + b x_0 + c y_0, in image
coordinates, as mpSetTransforms takes them) and a renderer of channel
images from emitters placed in channel 0.
"""
import numpy

from lm_evidence import Frame


class Affine(object):
    """x_c = xt . (1, x0, y0), y_c = yt . (1, x0, y0); inverse given explicitly."""
    def __init__(self, xt, yt, xt_inv, yt_inv):
        self.xt, self.yt = numpy.asarray(xt, float), numpy.asarray(yt, float)
        self.xt_inv, self.yt_inv = numpy.asarray(xt_inv, float), numpy.asarray(yt_inv, float)
        # the inverse really is one (a slip in a scenario's table would make every pin vacuous)
        for p in ((3.0, 5.0), (-7.0, 2.5), (11.0, -4.0)):
            q = self(*p)
            back = (self.xt_inv @ (1.0, q[0], q[1]), self.yt_inv @ (1.0, q[0], q[1]))
            assert numpy.allclose(back, p, atol=1e-12)

    def __call__(self, x0, y0):
        return self.xt @ (1.0, x0, y0), self.yt @ (1.0, x0, y0)


IDENTITY = Affine([0, 1, 0], [0, 0, 1], [0, 1, 0], [0, 0, 1])


def tables(maps):
    """The four [n_channels, 3] transform tables of mpSetTransforms."""
    return (numpy.array([m.xt for m in maps]), numpy.array([m.yt for m in maps]),
            numpy.array([m.xt_inv for m in maps]), numpy.array([m.yt_inv for m in maps]))


class WFrame(Frame):
    """Frame of the 2D model whose y width may differ from the fitted x width (theta[3])."""
    yw = None

    def widths(self, th):
        return th[3], (th[3] if self.yw is None else self.yw)


def render(shape, emitters, rqe, scmos, sigma):
    """Mean image h psf rqe + bg rqe + scmos; emitters (x, y, h) in this channel's coordinates."""
    yy, xx = numpy.mgrid[0:shape[0], 0:shape[1]]
    w = 1.0 / (2.0 * sigma ** 2)
    f = numpy.zeros(shape)
    for x, y, h in emitters:
        f = f + h * numpy.exp(-((xx - x) ** 2 + (yy - y) ** 2) * w)
    return f * rqe


def solve_step(fr, th, lam):
    """delta of (H + lam diag H) delta = J at th (the turn-start J, H)."""
    xi, yi = fr.origin(th)
    J, H = fr.jh(th, xi, yi)
    A = H.copy()
    A[numpy.diag_indices_from(A)] *= (1.0 + lam)
    return numpy.linalg.solve(A, J), J, H


def coupled_step(frames, maps, start, lam, wx, wy, wmin, wmax):
    """
    One damped step of a peak group from its channel-0 start
    (x0, y0, h, W, bg per channel): each channel's own step, then heights
    floored at 0.01, x widths clipped to [wmin, wmax], the channel-0 x and
    y moved by the weighted (wx / wy times new height) mean of the
    channel steps brought back to channel 0's axes, and every channel
    placed by its map.  Returns per-channel new thetas and the deltas.
    """
    x0, y0 = start["x"], start["y"]
    n = len(frames)
    ths, deltas = [], []
    for c in range(n):
        xc, yc = maps[c](x0, y0)
        th = numpy.array([start["h"][c], xc - frames[c].xoff, yc - frames[c].xoff, start["W"], start["bg"][c]])
        d, _, _ = solve_step(frames[c], th, lam)
        ths.append(th)
        deltas.append(d)
    h_new = numpy.array([max(ths[c][0] - deltas[c][0], 0.01) for c in range(n)])
    w_new = numpy.array([min(max(ths[c][3] - deltas[c][3], wmin), wmax) for c in range(n)])
    dx = sum(wx[c] * h_new[c] * (maps[c].xt_inv[1] * deltas[c][1] + maps[c].xt_inv[2] * deltas[c][2])
             for c in range(n)) / sum(wx[c] * h_new[c] for c in range(n))
    dy = sum(wy[c] * h_new[c] * (maps[c].yt_inv[1] * deltas[c][1] + maps[c].yt_inv[2] * deltas[c][2])
             for c in range(n)) / sum(wy[c] * h_new[c] for c in range(n))
    x0n, y0n = x0 - dx, y0 - dy
    out = []
    for c in range(n):
        xc, yc = maps[c](x0n, y0n)
        out.append(numpy.array([h_new[c], xc - frames[c].xoff, yc - frames[c].xoff, w_new[c],
                                ths[c][4] - deltas[c][4]]))
    return ths, out, deltas


def group_deviance(frames, thetas, origins, yw):
    """Summed deviance of the channels' peaks at the given ROI origins, y widths fixed at yw."""
    tot = 0.0
    for fr, th, (xi, yi) in zip(frames, thetas, origins):
        fr.yw = yw
        d = fr.deviance(th, xi, yi)
        fr.yw = None
        if d is None:
            return None
        tot += d
    return tot
