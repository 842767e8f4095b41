"""Pins of the CPU oracle (CPU only).

Each test checks an oracle function against something that does not retype
its formula: a closed form, brute force written differently (padding instead
of clamping, numpy cumsums, explicit pair loops), an invariant, or a crafted
input on which a plausible mistake (wrong lane assignment, wrong tree or block
order, left-aligned instead of centred boxes, rounding instead of truncation,
dividing by the summed count instead of the gulp) gives a different answer.
"""
import math

import numpy as np
import pytest

TWO53 = float(2**53)


# ---------------------------------------------------------------- canonical sum
def test_csum_exact_on_integers(orc):
    rng = np.random.default_rng(1)
    for n in (0, 1, 255, 256, 4095, 4096, 4097, 3 * 4096 + 17):
        x = rng.integers(-1000, 1000, n).astype(np.float64)
        assert orc.csum(x) == float(x.sum())


def test_csum_close_to_exact_sum(orc):
    x = np.random.default_rng(2).standard_normal(50_000)
    assert abs(orc.csum(x) - math.fsum(x)) <= 1e-12 * np.abs(x).sum()


def test_csum_lanes_are_strided(orc):
    # lane 0 holds x[0] and x[256]; lane 1 holds x[1].  Strided lanes cancel
    # the 2^53 pair exactly (result 1); contiguous lanes or a sequential sum
    # absorb the 1 into 2^53 first (result 0).
    x = np.zeros(4096)
    x[0], x[1], x[256] = TWO53, 1.0, -TWO53
    assert orc.csum(x) == 1.0


def test_csum_tree_halves_first(orc):
    # lanes 0, 1, 128 = 2^53, 1, -2^53: the tree p[l] += p[l+128] first cancels
    # lanes 0 and 128 (result 1); adjacent pairing or a lane-sequential sum gives 0.
    x = np.zeros(4096)
    x[0], x[1], x[128] = TWO53, 1.0, -TWO53
    assert orc.csum(x) == 1.0


def test_csum_blocks_ascending(orc):
    # block sums 2^53, 1, -2^53 added in ascending block order give 0;
    # descending order would give 1.
    x = np.zeros(3 * 4096)
    x[0], x[4096], x[2 * 4096] = TWO53, 1.0, -TWO53
    assert orc.csum(x) == 0.0


# ---------------------------------------------------------------- statistics
def test_stats_mean_of_row_moments(orc):
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((7, 512)) * 3 + rng.standard_normal((7, 1)) * 5).astype(np.float32)
    mu, var = orc.stats(x)
    xd = x.astype(np.float64)
    mu_ref = np.mean(xd.mean(axis=1))                  # mean of row means
    var_ref = np.mean(((xd - mu_ref) ** 2).mean(axis=1))  # mean of row variances about mu
    assert abs(mu - mu_ref) <= 1e-12 * (1 + abs(mu_ref))
    assert abs(var - var_ref) <= 1e-10 * var_ref


def test_stats_column_window(orc):
    rng = np.random.default_rng(4)
    x = rng.standard_normal((3, 300)).astype(np.float32)
    y = x.copy()
    y[:, 256:] = 1e6          # columns past 256*floor(w/256) are not used
    assert orc.stats(x) == orc.stats(y)
    small = rng.standard_normal((2, 100)).astype(np.float32)  # w < 256: all columns
    mu, var = orc.stats(small)
    d = small.astype(np.float64)
    assert abs(mu - d.mean()) < 1e-12
    assert abs(var - np.mean(((d - d.mean()) ** 2).mean(axis=1))) < 1e-10


def _reference_float_row_reduce(v):
    """One 256-thread row block of computeRowMeans/computeRowVariances
    (normalization.cu:19-122) in float32: thread l sums v[l], v[l+256], ... in
    order; each 32-lane warp reduces by shuffle-down 16, 8, 4, 2, 1; thread 0
    adds the other warps' sums in order."""
    n = v.shape[-1]
    lanes = np.zeros(v.shape[:-1] + (256,), np.float32)
    for k in range(0, n, 256):
        lanes = (lanes + v[..., k:k + 256]).astype(np.float32)
    warps = lanes.reshape(v.shape[:-1] + (8, 32)).copy()
    for i in (16, 8, 4, 2, 1):
        warps[..., :i] = (warps[..., :i] + warps[..., i:2 * i]).astype(np.float32)
    acc = warps[..., 0, 0].copy()
    for w in range(1, 8):
        acc = (acc + warps[..., w, 0]).astype(np.float32)
    return acc


def _reference_float_stats(x):
    """calculate_stddev (normalization.cu:125-162) in float32, host loops in order."""
    h, w = x.shape
    wp = 256 * (w // 256)
    xs = x[:, :wp].astype(np.float32)
    row_means = (_reference_float_row_reduce(xs) / np.float32(wp)).astype(np.float32)
    mean = np.float32(0)
    for r in range(h):
        mean = np.float32(mean + row_means[r])
    mean = np.float32(mean / np.float32(h))
    d = (xs - mean).astype(np.float32)
    row_vars = (_reference_float_row_reduce((d * d).astype(np.float32)) / np.float32(wp)).astype(np.float32)
    var = np.float32(0)
    for r in range(h):
        var = np.float32(var + row_vars[r])
    var = np.float32(var / np.float32(h))
    return mean, np.float32(np.sqrt(var))


@pytest.mark.parametrize("shape,loc,scale", [((64, 4096), 0.0, 1.0), ((256, 8192), 20.0, 3.0),
                                             ((1889, 2048), 5000.0, 60.0)])
def test_stats_within_reference_float_arithmetic(orc, shape, loc, scale):
    """The oracle's double canonical statistics agree with the reference's own
    float32 reduction well inside the 1e-5 relative bar on normalised data:
    mean and sigma each within 1e-5 of the reference's sigma."""
    x = (loc + scale * np.random.default_rng(16).standard_normal(shape)).astype(np.float32)
    mf, sf = orc.stats_f(x)
    mr, sr = _reference_float_stats(x)
    assert abs(sf - sr) <= 1e-5 * sr
    assert abs(mf - mr) <= 1e-5 * sr


def test_stats_f_rounding(orc):
    x = np.random.default_rng(5).standard_normal((4, 1024)).astype(np.float32)
    mu, var = orc.stats(x)
    mf, sf = orc.stats_f(x)
    assert mf == np.float32(mu) and sf == np.float32(math.sqrt(var))


# ---------------------------------------------------------------- dedispersion
def _dedisp_bruteforce(data_tf, delta, nout, killmask=None):
    ndm, nch = delta.shape
    out = np.zeros((ndm, nout), np.uint64)
    live = np.ones(nch, bool) if killmask is None else killmask.astype(bool)
    t = np.arange(nout)
    for d in range(ndm):
        for c in np.nonzero(live)[0]:
            out[d] += data_tf[t + delta[d, c], c]
    return out.astype(np.uint32)


def test_dedisperse_bruteforce(orc):
    rng = np.random.default_rng(6)
    nch, ndm, nout = 24, 9, 300
    delta = np.sort(rng.integers(0, 80, (ndm, nch)), axis=0).astype(np.int32)
    data = rng.integers(0, 256, (nout + 80, nch)).astype(np.uint8)
    km = (rng.random(nch) > 0.3).astype(np.uint8)
    np.testing.assert_array_equal(orc.dedisperse_tf(data, delta, nout),
                                  _dedisp_bruteforce(data, delta, nout))
    np.testing.assert_array_equal(orc.dedisperse_tf(data, delta, nout, km),
                                  _dedisp_bruteforce(data, delta, nout, km))
    # the channel-major entry reads the same samples
    np.testing.assert_array_equal(orc.dedisperse_ft(data.T.copy(), delta, nout, km),
                                  _dedisp_bruteforce(data, delta, nout, km))


def test_dedisperse_closed_forms(orc):
    nch, nout = 16, 100
    delta = np.array([np.arange(nch) * k for k in range(4)], np.int32)
    const = np.full((nout + 60, nch), 7, np.uint8)
    assert np.all(orc.dedisperse_tf(const, delta, nout) == 7 * nch)
    # a dispersed impulse: sample p + delta[2][c] in channel c lines up at p in row 2
    imp = np.zeros((nout + 60, nch), np.uint8)
    p = 40
    imp[p + delta[2], np.arange(nch)] = 1
    out = orc.dedisperse_tf(imp, delta, nout)
    assert out[2, p] == nch and out[2].sum() == nch
    assert out[0].max() < nch and out[3].max() < nch


# ---------------------------------------------------------------- flagger stages
def test_bandpass_divides_by_gulp(orc):
    x = np.full((3, 100), 2.0, np.float32)  # 96 = 32*floor(100/32) samples summed
    np.testing.assert_array_equal(orc.bandpass(x), np.float32(2.0 * 96 / 100))
    x = np.random.default_rng(7).standard_normal((5, 256)).astype(np.float32)
    np.testing.assert_allclose(orc.bandpass(x), x.astype(np.float64).mean(axis=1), atol=1e-7)


def test_divide_bp_zero_guard(orc):
    x = np.full((3, 8), 6.0, np.float32)
    out = orc.divide_bp(x, np.array([2.0, 0.0, -1.0], np.float32))
    assert np.all(out[0] == 3.0) and np.all(out[1] == 0.0) and np.all(out[2] == 0.0)


def test_normalize_unit_moments(orc):
    rng = np.random.default_rng(8)
    x = (rng.standard_normal((8, 1024)) * 4 + 10).astype(np.float32)
    y = orc.normalize(x).astype(np.float64)
    mu = np.mean(y.mean(axis=1))
    var = np.mean(((y - mu) ** 2).mean(axis=1))
    assert abs(mu) < 1e-6 and abs(var - 1) < 1e-5


def test_spec_flag(orc):
    x = np.zeros((3, 64), np.float32)
    x[1] = 0.5   # channel mean 0.5 is out of [-0.05, 0.15]
    x[2] = 0.1
    y = orc.spec_flag(x, -0.05, 0.15)
    assert np.all(y[0] == 1.0) and np.all(y[1] == 1.0)
    np.testing.assert_array_equal(y[2], np.float32(0.1) + np.float32(1.0))


def _box_mean_padded(x, ts, fs):
    """box mean with the (ts/2, fs/2) anchor written as an edge-padded sliding sum"""
    ax, ay = ts // 2, fs // 2
    p = np.pad(x.astype(np.float64), ((ay, fs - 1 - ay), (ax, ts - 1 - ax)), mode="edge")
    nch, W = x.shape
    out = np.zeros((nch, W))
    for i in range(fs):
        for j in range(ts):
            out += p[i:i + nch, j:j + W]
    return out / (ts * fs)


@pytest.mark.parametrize("ts,fs", [(1, 1), (4, 2), (3, 3), (8, 1), (2, 5)])
def test_box_mean_anchor_and_border(orc, ts, fs):
    x = np.random.default_rng(9).standard_normal((11, 40)).astype(np.float32)
    np.testing.assert_allclose(orc.box_mean(x, ts, fs), _box_mean_padded(x, ts, fs),
                               rtol=1e-5, atol=1e-6)
    # impulse: output (c, t) sees source rows c-fs/2.. and columns t-ts/2..
    imp = np.zeros((11, 40), np.float32)
    imp[5, 20] = 1.0
    nz = np.argwhere(orc.box_mean(imp, ts, fs) != 0)
    assert nz[:, 0].min() == 5 + fs // 2 - fs + 1 and nz[:, 0].max() == 5 + fs // 2
    assert nz[:, 1].min() == 20 + ts // 2 - ts + 1 and nz[:, 1].max() == 20 + ts // 2


@pytest.mark.parametrize("ts,fs,f1,f2", [(2, 2, -1, -1), (4, 1, 3, 5), (1, 1, 0, 1), (3, 2, -1, -1)])
def test_scrunch_flag_against_padded_form(orc, ts, fs, f1, f2):
    rng = np.random.default_rng(10)
    # dyadic values keep every box sum exact in both formulations
    x = (rng.integers(-8, 9, (9, 48)) / 4.0).astype(np.float32)
    x[4, 10:14] = 12.0
    thresh = 3.0
    S = _box_mean_padded(x, ts, fs)
    thr = np.float32(thresh / math.sqrt(ts * fs))
    M = (S > thr) | (S < -thr)
    ax, ay = ts // 2, fs // 2
    Mp = np.pad(M, ((ay, fs - 1 - ay), (ax, ts - 1 - ax)), mode="edge")
    G = np.zeros_like(M)
    for i in range(fs):
        for j in range(ts):
            G |= Mp[i:i + 9, j:j + 48]
    fixed = np.zeros(9, bool)
    if f1 >= 0:
        fixed[f1:f2] = True
    ref = np.where(G | fixed[:, None], np.float32(0), x) + np.float32(1)
    y, cnt = orc.scrunch_flag(x, ts, fs, thresh, f1, f2)
    np.testing.assert_array_equal(y, ref.astype(np.float32))
    np.testing.assert_array_equal(cnt, G.sum(axis=1))


@pytest.mark.parametrize("ts,fs,W", [(1, 1, 300), (3, 2, 300), (4, 1, 48), (2, 5, 520)])
def test_flag_spectrum_is_the_grown_mask_mean(orc, ts, fs, W):
    """flagSpec (flagger.cu:333-334): each scrunch adds, per channel, the mean
    of G = box mean of the threshold mask over the statistics columns (the
    first 256*floor(W/256), all W below 256); written here as the masked
    pixels of each edge-padded box counted directly."""
    rng = np.random.default_rng(ts * 100 + fs * 10 + W)
    nch = 9
    x = (rng.integers(-8, 9, (nch, W)) / 4.0).astype(np.float32)
    x[4, 10:14] = 12.0
    x[7, W - 20:] = -9.0          # flags past the statistics columns do not count
    thresh = 3.0
    S = _box_mean_padded(x, ts, fs)
    thr = np.float32(thresh / math.sqrt(ts * fs))
    M = ((S > thr) | (S < -thr)).astype(np.int64)
    ax, ay = ts // 2, fs // 2
    Mp = np.pad(M, ((ay, fs - 1 - ay), (ax, ts - 1 - ax)), mode="edge")
    N = np.zeros((nch, W), np.int64)
    for i in range(fs):
        for j in range(ts):
            N += Mp[i:i + nch, j:j + W]
    wp = (W // 256) * 256 or W
    want = (N[:, :wp].sum(axis=1) / (ts * fs) / wp).astype(np.float32)
    assert want.max() > 0
    _, _, spec = orc.scrunch_flag(x, ts, fs, thresh, want_spec=True)
    np.testing.assert_array_equal(spec, want)
    # a second scrunch accumulates in float
    spec2 = spec.copy()
    L = orc.L
    y = np.array(x, np.float32, copy=True)
    cnt = np.zeros(nch, np.uint32)
    L.orc_scrunch_flag(y.ctypes.data, nch, W, ts, fs, thresh, -1, -1, cnt.ctypes.data,
                       spec2.ctypes.data)
    np.testing.assert_array_equal(spec2, (want + want).astype(np.float32))


@pytest.mark.parametrize("W", [100, 300, 777])
def test_robust_normalise_brute_force(orc, W):
    """orc_normalize_robust against numpy sorts: per channel the lower median
    of the statistics columns, the lower median of |x - m| (float), the scale
    1.4826 x MAD, and the normalised plane; a constant channel gives zeros"""
    rng = np.random.default_rng(W)
    nch = 6
    x = (rng.standard_normal((nch, W)) * np.arange(1, nch + 1)[:, None] + 3).astype(np.float32)
    x[2, :] = 5.0                                  # MAD 0: the channel becomes 0
    x[3, ::7] = 1e6                                # outliers move mean and sigma, not the median
    y, ms = orc.normalize_robust(x)
    wp = (W // 256) * 256 or W
    k = (wp - 1) // 2
    for c in range(nch):
        m = np.sort(x[c, :wp])[k]
        d = np.sort(np.abs(x[c, :wp] - m))[k]
        s = np.float32(1.4826 * np.float64(d))
        assert ms[c, 0] == m and ms[c, 1] == s
        want = np.zeros(W, np.float32) if s == 0 else \
            ((x[c].astype(np.float64) - np.float64(m)) * (1.0 / np.float64(s))).astype(np.float32)
        np.testing.assert_array_equal(y[c], want)
    assert not y[2].any()
    # closed form: the first 256 columns holding 0..255 in any order have
    # lower median 127 and |x - 127| lower median 64 (ranks 0..127 of the
    # deviations 0, 1, 1, 2, 2, ..., 127, 127, 128)
    z = np.r_[rng.permutation(256), [999.0]].astype(np.float32)[None, :]
    _, ms = orc.normalize_robust(np.repeat(z, 2, axis=0))
    assert ms[0, 0] == 127.0 and ms[0, 1] == np.float32(1.4826 * 64.0)


def test_robust_flag_ignores_a_loud_minority(orc):
    """with robust statistics a channel whose few samples are huge keeps its
    noise level after normalisation (median and MAD ignore them), where the
    plane's sigma would be inflated; the default mode is unchanged"""
    rng = np.random.default_rng(21)
    nch, W = 16, 4096
    x = (100 + 10 * rng.standard_normal((nch, W))).astype(np.float32)
    x[5, ::50] = 1e5
    y, ms = orc.normalize_robust(x)
    clean = np.delete(np.arange(W), np.arange(0, W, 50))
    assert 0.8 < np.std(y[5, clean]) < 1.25
    assert abs(ms[5, 0] - 100) < 2 and abs(ms[5, 1] - 10) < 1.5
    a, _ = orc.flag(x, robust=False)
    b, _ = orc.flag(x, robust=True)
    assert not np.array_equal(a, b)


def test_ts_correct(orc):
    x = np.vstack([np.full(16, 2.0), np.full(16, 4.0)]).astype(np.float32)
    y, ts = orc.ts_correct(x)
    assert np.all(ts == 3.0)
    np.testing.assert_array_equal(y, (x / np.float32(3.0)).astype(np.float32))


def test_chan_sum_partitions(orc):
    # 32 channels -> 16 partitions of 2.  2^53 (partition 0) and -2^53
    # (partition 1) meet at the last tree step, after the 1 of partition 2 was
    # absorbed into 2^53: result 0, where a sequential channel sum gives 1.
    x = np.zeros((32, 3), np.float32)
    x[0, 0], x[2, 0], x[4, 0] = TWO53, -TWO53, 1.0
    assert orc.chan_sum(x, 0) == 0.0
    # partitions 0 and 8 cancel in the first tree step (h = 8), so the 1 of
    # partition 1 survives: result 1, where summing the partials in order gives 0
    x[:, 1] = 0
    x[0, 1], x[16, 1], x[2, 1] = TWO53, -TWO53, 1.0
    assert orc.chan_sum(x, 1) == 1.0
    # exact on small integers, any channel count (empty partitions when nch < 16)
    for nch in (1, 5, 16, 17, 100, 1024):
        y = np.random.default_rng(nch).integers(-50, 50, (nch, 2)).astype(np.float32)
        assert orc.chan_sum(y, 1) == float(y[:, 1].astype(np.float64).sum())


def test_quantize_edges(orc):
    x = np.array([[-4.0, 10.0, 3.0, np.nan, np.inf, -np.inf, -100.0, 0.0]], np.float32)
    q = orc.quantize(x)[0]
    # (x + 4) * (255.f/14.f) in float, truncated: 3 -> 127.5 -> 127 (rounding
    # would give 128); 10 -> 14 * 18.214285f = 254.99998 -> 254 (a double
    # constant would give 255); +inf clamps to 255
    assert list(q[:7]) == [0, 254, 127, 0, 255, 0, 0]
    assert q[7] == int(np.float32(4.0) * np.float32(255.0 / 14.0))


def test_ingest_transposes(orc):
    rng = np.random.default_rng(11)
    a = rng.integers(0, 256, (50, 7)).astype(np.uint8)
    np.testing.assert_array_equal(orc.ingest_u8_tf(a), a.T.astype(np.float32))
    b = rng.standard_normal((50, 7)).astype(np.float32)
    np.testing.assert_array_equal(orc.ingest_f32_tf(b), b.T)


@pytest.mark.parametrize("scr,robust", [((), False), (((1, 1, 5.0),), False),
                                        (((1, 1, 4.0), (4, 2, 3.0)), False),
                                        (((2, 3, 3.5),), True)])
def test_flag_is_the_stages_in_reference_order(orc, scr, robust):
    """orc_flag equals its individually pinned stages composed in the order
    of fast_flagger (flagger.cu:345-411): bandpass_flag = bandpass, divide,
    normalise, spectral flag (flagger.cu:216-242); per scrunch bandpass,
    divide, normalise, flag (flagger.cu:365-384); then bandpass, divide,
    ts_correct, normalise (apply_scrunch(8,8,100,0,1), flagger.cu:385-401);
    quantise (transpose.cu:65-75).  Dropping or reordering a stage changes
    the bytes."""
    rng = np.random.default_rng(33)
    nch, W = 24, 1000
    x = (50 + 5 * rng.standard_normal((nch, W)) * np.linspace(0.5, 2, nch)[:, None])
    x[3] += 40 * rng.standard_normal(W)
    x[:, 300:305] += 30
    x = x.astype(np.float32)
    norm = (lambda y: orc.normalize_robust(y)[0]) if robust else orc.normalize

    def stages(y, drop=None):
        y = orc.divide_bp(y, orc.bandpass(y))
        y = norm(y)
        y = orc.spec_flag(y, -0.05, 0.15)
        for ts, fs, th in scr:
            y = orc.divide_bp(y, orc.bandpass(y))
            y = norm(y)
            y, _ = orc.scrunch_flag(y, ts, fs, th)
        y = orc.divide_bp(y, orc.bandpass(y))
        if drop != "ts":
            y, _ = orc.ts_correct(y)
        y = norm(y)
        return orc.quantize(y)

    got, _ = orc.flag(x, -0.05, 0.15, -1, -1, scr, robust=robust)
    np.testing.assert_array_equal(got, stages(x))
    assert not np.array_equal(got, stages(x, drop="ts"))


@pytest.mark.parametrize("W", [4096, 4000, 300])
def test_flag_byte_histogram_order(orc, W):
    """8-bit input with 1 x 1 scrunches: the byte-histogram order (each
    channel sum before ts_correct a 256-term sum of count x value) is the
    same arithmetic in another summation order, so the flagged bytes agree
    with the plane order to within one quantisation level on a tiny fraction
    of samples and the flag counts and spectrum agree exactly; with a box
    scrunch (no longer a function of the byte) the plane order is taken"""
    rng = np.random.default_rng(W)
    nch = 64
    g = np.clip(np.rint(rng.normal(128, 16, (W, nch))), 0, 255).astype(np.uint8)
    g[:, 5] = np.clip(np.rint(rng.normal(128, 160, W)), 0, 255).astype(np.uint8)
    g[W // 3:W // 3 + 6, :] = 250
    scr = ((1, 1, 5.0), (1, 1, 3.0))
    sa, sb = np.zeros(nch, np.float32), np.zeros(nch, np.float32)
    a, ca = orc.flag_u8(g, -0.05, 0.15, 3, 5, scr, spec=sa)
    b, cb = orc.flag(orc.ingest_u8_tf(g), -0.05, 0.15, 3, 5, scr, spec=sb)
    d = np.abs(a.astype(int) - b.astype(int))
    assert d.max() <= 1 and (d > 0).mean() < 1e-3
    np.testing.assert_array_equal(ca, cb)
    np.testing.assert_allclose(sa, sb, rtol=1e-6)
    assert ca[5] > 0
    box = ((1, 1, 5.0), (4, 2, 3.0))
    np.testing.assert_array_equal(orc.flag_u8(g, scrunches=box)[0],
                                  orc.flag(orc.ingest_u8_tf(g), scrunches=box)[0])


def test_flag_removes_bandpass_and_flags_rfi(orc):
    rng = np.random.default_rng(12)
    nch, W = 32, 4096
    gain = np.linspace(0.5, 2.0, nch)[:, None]
    x = (gain * (20 + 3 * rng.standard_normal((nch, W)))).astype(np.float32)
    x[7] += (40 * rng.standard_normal(W)).astype(np.float32)  # one loud channel
    out, cnt = orc.flag(x, scrunches=[(1, 1, 6.0), (8, 1, 5.0)])
    assert out.dtype == np.uint8
    clean = np.delete(np.arange(nch), 7)
    means = out[clean].astype(np.float64).mean(axis=1)
    # gain removed: every clean channel sits at the quantiser's zero point
    assert np.all(np.abs(means - 4 * 255 / 14) < 3.0)
    assert cnt[7] > 10 * np.median(cnt[clean]) + 50


# ---------------------------------------------------------------- search
def test_boxcar_planes_are_centred_box_sums(orc):
    rng = np.random.default_rng(13)
    ndm, nout, maxW = 6, 512, 8
    D = rng.integers(1000, 1100, (ndm, nout + maxW)).astype(np.uint32)
    r = orc.search(D, nout, maxW, snr=100.0, want_planes=True)
    P = np.concatenate([np.zeros((ndm, 1), np.int64), np.cumsum(D.astype(np.int64), axis=1)], axis=1)
    for k, w in enumerate([1, 2, 4, 8]):
        s = np.arange(nout)
        a = s + maxW // 2 - w // 2
        raw = (P[:, a + w] - P[:, a]).astype(np.float64)
        mu = np.mean(raw.mean(axis=1))
        var = np.mean(((raw - mu) ** 2).mean(axis=1))
        assert abs(r["mu"][k] - mu) <= 1e-6 * mu
        assert abs(r["sigma"][k] - math.sqrt(var)) <= 1e-5 * math.sqrt(var)
        np.testing.assert_allclose(r["planes"][k], (raw - mu) / r["sigma"][k], rtol=1e-5, atol=1e-5)


def test_boxcar_spike_position_and_order(orc):
    rng = np.random.default_rng(14)
    ndm, nout, maxW = 8, 1024, 16
    D = rng.integers(900, 1100, (ndm, nout + maxW)).astype(np.uint32)
    D[3, 600] += 5000
    r = orc.search(D, nout, maxW, snr=8.0)
    c = r["cands"]
    w1 = c[c["width"] == 0]
    assert list(zip(w1["dm_idx"], w1["samp"])) == [(3, 600 - maxW // 2)]
    keys = list(zip(c["width"], c["dm_idx"], c["samp"]))
    assert keys == sorted(keys) and r["count"] == len(c)
    # the width-w box centred at s covers 600 for s in [600-8-(w-w/2-1), 600-8+w/2]
    for k, w in enumerate([1, 2, 4, 8, 16]):
        ck = c[(c["width"] == k) & (c["dm_idx"] == 3)]["samp"]
        assert ck.min() >= 600 - maxW // 2 - (w - w // 2 - 1)
        assert ck.max() <= 600 - maxW // 2 + w // 2


def test_search_threshold_and_cap(orc):
    rng = np.random.default_rng(15)
    D = rng.integers(0, 2000, (10, 2048 + 4)).astype(np.uint32)
    r = orc.search(D, 2048, 4, snr=1.5, want_planes=True)
    thr = np.float32(1.5) * np.float32(r["sigma0"])
    assert r["count"] == int((r["planes"] > thr).sum())
    capped = orc.search(D, 2048, 4, snr=1.5, max_giants=17)
    assert capped["count"] == r["count"] and len(capped["cands"]) == 17
    np.testing.assert_array_equal(capped["cands"], r["cands"][:17])


def test_search_gates_flat_plane(orc):
    D = np.full((4, 300), 50, np.uint32)
    r = orc.search(D, 296, 4, snr=5.0)
    assert r["gated"] and r["count"] == 0


def test_hella_kernel_shape(orc):
    for sm in (1, 2, 4, 8):
        k = orc.hella_kernel(sm).astype(np.float64)
        assert abs(k.sum() - 1) < 1e-5
        np.testing.assert_allclose(k[0], k[2])
        np.testing.assert_allclose(k[0], 0.3 * k[1], rtol=1e-6)
        np.testing.assert_allclose(k[:, ::-1], k, rtol=1e-6)
        assert k[1].argmax() == sm


def test_hella_impulse_response(orc):
    ndm, nout, maxW = 9, 256, 8   # HELLA: scales 1, 2, 4
    D = np.full((ndm, nout + maxW), 100, np.uint32)
    D[4, 120] += 1000
    r = orc.search(D, nout, maxW, minW=1, mode=orc.MODE_HELLA, snr=1e9, want_planes=True)
    for k, sm in enumerate([1, 2, 4]):
        K = orc.hella_kernel(sm).astype(np.float64)
        raw = r["planes"][k].astype(np.float64) * r["sigma"][k] + r["mu"][k]
        base = np.median(raw)
        # plane row pr reads D rows pr..pr+2 and columns s+maxW/2 .. s+maxW/2+2sm
        for i in range(3):
            pr = 4 - i
            for j in range(2 * sm + 1):
                s = 120 - maxW // 2 - j
                assert abs(raw[pr, s] - base - 1000 * K[i, j]) < 1e-2 * (1 + 1000 * K[i, j])


# ---------------------------------------------------------------- clustering
def _cluster_bruteforce(c, ct, cd):
    n = len(c)
    cell = [(int(c[i]["dm_idx"]) // cd, int(c[i]["samp"]) // ct) for i in range(n)]
    par = list(range(n))

    def find(a):
        while par[a] != a:
            a = par[a]
        return a

    for i in range(n):
        for j in range(i + 1, n):
            if abs(cell[i][0] - cell[j][0]) <= 1 and abs(cell[i][1] - cell[j][1]) <= 1:
                a, b = find(i), find(j)
                if a != b:
                    par[max(a, b)] = min(a, b)
    groups = {}
    for i in range(n):
        groups.setdefault(find(i), []).append(i)
    out = []
    for g in groups.values():
        rep = g[0]
        for i in g:
            if c[i]["snr"] > c[rep]["snr"]:
                rep = i
        out.append(dict(rep=rep, count=len(g),
                        samp_min=int(min(c[i]["samp"] for i in g)),
                        samp_max=int(max(c[i]["samp"] for i in g)),
                        dm_min=int(min(c[i]["dm_idx"] for i in g)),
                        dm_max=int(max(c[i]["dm_idx"] for i in g))))
    return sorted(out, key=lambda d: d["rep"])


@pytest.mark.parametrize("seed", range(6))
def test_cluster_bruteforce(orc, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 60))
    c = np.zeros(n, dtype=[("snr", "f4"), ("samp", "i4"), ("width", "i4"), ("dm_idx", "i4")])
    c["snr"] = rng.integers(7, 12, n)   # ties on purpose
    c["samp"] = rng.integers(0, 400, n)
    c["dm_idx"] = rng.integers(0, 60, n)
    assert orc.cluster(c, 16, 4) == _cluster_bruteforce(c, 16, 4)
