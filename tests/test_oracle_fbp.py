"""Pins the FBP spatial filter of the CPU oracle (ApplyFilterInplace modes 0
and 1, filter_prj.cpp:137-178, 266-446) before the GPU is checked against
it.  OpenCV is absent, so the reference's float DFT rounding is parity
unpinned; the convolution itself is pinned here by:
  - the kernels' frequency responses against their closed forms;
  - a numpy-FFT restatement of the reference's literal steps (cvRect
    padding and flips, mirror fill to the optimal DFT size, spectrum
    product, inverse DFT scaled by 1/N, the two crops);
  - the impulse response, which is the tap vector at the right shift;
  - cvGetOptimalDFTSize against a brute-force enumeration of 2^a 3^b 5^c.
"""
import numpy as np
import pytest


def smooth_numbers(limit):
    out = set()
    a = 1
    while a <= limit:
        b = a
        while b <= limit:
            c = b
            while c <= limit:
                out.add(c)
                c *= 5
            b *= 3
        a *= 2
    return sorted(out)


def test_optimal_dft_size(orc):
    nums = smooth_numbers(20000)
    for n in [1, 2, 7, 97, 251, 292, 552, 1001, 1276, 2300, 4348, 8443, 16384]:
        want = next(m for m in nums if m >= n)
        assert orc.optimal_dft_size(n) == want, n


@pytest.mark.parametrize("mode", [0, 1])
def test_taps_frequency_response(orc, mode):
    """Shepp-Logan: H(w) = |sin(w/2)|; Ram-Lak: H(w) = |w|/2 on [-pi, pi]
    (infinite kernels); 251 taps truncate them by about 2.5e-3."""
    h = orc.spatial_taps(mode).astype(np.float64)
    assert h.size == 251
    np.testing.assert_array_equal(h, h[::-1])
    n = np.arange(-125, 126)
    for w in np.linspace(0.05, np.pi, 40):
        H = (h * np.cos(w * n)).sum()
        want = abs(np.sin(w / 2)) if mode == 0 else abs(w) / 2
        assert abs(H - want) < 5e-3, (mode, w, H, want)


def reference_steps(row, taps):
    """The reference's FileterProjectionSymmetrize on one row, step by step
    with numpy's FFT (rows never interact, so one row stands for all)."""
    X = row.size
    L = taps.size
    row_pad = 2
    n_x_pad = X + row_pad
    N = next(m for m in smooth_numbers(4 * (n_x_pad + L)) if m >= n_x_pad + L - 1)
    pad = np.zeros(n_x_pad)
    rx = row_pad // 2
    pad[rx:rx + X] = row
    # symmetrize_IplImage_values_CvRect, x part: left block [0, rx) is the
    # flip of [rx, 2 rx), right block [rx + X, n_x_pad) the flip of the
    # same width just inside the ROI
    pad[0:rx] = pad[rx:2 * rx][::-1]
    w = n_x_pad - (rx + X)
    pad[rx + X:] = pad[rx + X - w:rx + X][::-1]
    conv = np.zeros(N)
    conv[:n_x_pad] = pad
    hi = N if n_x_pad >= N - n_x_pad else 2 * n_x_pad
    for i in range(n_x_pad, hi):
        conv[i] = conv[2 * n_x_pad - 1 - i]
    f = np.zeros(N)
    f[:L] = taps
    y = np.fft.ifft(np.fft.fft(conv) * np.fft.fft(f)).real
    pad2 = y[L // 2:L // 2 + n_x_pad]
    return pad2[rx:rx + X]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("X", [40, 300, 1024])
def test_filter_matches_reference_steps(orc, mode, X):
    rng = np.random.default_rng(X + mode)
    rows = rng.standard_normal((3, X)) + np.linspace(0, 5, X)
    got = orc.fbp_filter_f64(rows, mode)
    taps = orc.spatial_taps(mode).astype(np.float64)
    for r in range(3):
        want = reference_steps(rows[r], taps)
        np.testing.assert_allclose(got[r], want, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("mode", [0, 1])
def test_impulse_response(orc, mode):
    X, j0 = 600, 300
    row = np.zeros((1, X))
    row[0, j0] = 1.0
    got = orc.fbp_filter_f64(row, mode)[0]
    h = orc.spatial_taps(mode).astype(np.float64)
    want = np.zeros(X)
    for i in range(X):
        if abs(i - j0) <= 125:
            want[i] = h[i - j0 + 125]
    np.testing.assert_array_equal(got, want)


# ---------------------------------------------------------------- f2b

def test_f2b_hand_values(orc):
    """{0, 2}: mean 1, sd 1, z = -1 and 1: (9/20) 255 = 114.75 -> 114 and
    (11/20) 255 = 140.25 -> 140."""
    b, (m, s) = orc.float_to_byte(np.array([0.0, 2.0], np.float32))
    assert (m, s) == (1.0, 1.0)
    assert list(b) == [114, 140]


def test_f2b_clamp_and_centre(orc):
    """The mean maps to 127 (127.5 truncated); 10 sd and beyond map to the
    ends; a constant volume (sd 0) gives NaN, which the clamp turns into 255."""
    assert list(orc.bytes_from_stats(np.array([5.0, 5.0 + 10 * 2, 5.0 - 10 * 2, 100.0, -100.0], np.float32),
                                     5.0, 2.0)) == [127, 255, 0, 255, 0]
    b, (m, s) = orc.float_to_byte(np.full(7, 3.5, np.float32))
    assert m == 3.5 and s == 0.0 and list(b) == [255] * 7


def test_f2b_moments(orc):
    v = np.random.default_rng(3).standard_normal(10001).astype(np.float32) * 4 + 9
    b, (m, s) = orc.float_to_byte(v)
    assert m == np.float32(v.astype(np.float64).mean())
    assert abs(s - v.astype(np.float64).std()) <= 1e-6 * s
    assert np.all(np.diff(b[np.argsort(v, kind="stable")].astype(int)) >= 0)  # monotone
