"""CPU checks of the FP64 U-Net oracle (oracle/rc_oracle_unet.py): its analytic
gradient against central finite differences, the adjoint pairs it uses, the
bfloat16 rounding, and that its parameter layout matches the library's."""
import numpy as np

from oracle import rc_oracle_unet as U
from retrocam_amd import _native as N


def small_problem(seed=0, n=1, h=16, w=16):
    rng = np.random.default_rng(seed)
    raw = np.concatenate([rng.uniform(0, 3, (n, h, w, 3)), rng.uniform(0, 1, (n, h, w, 3)),
                          rng.standard_normal((n, h, w, 3)), rng.uniform(0, 10, (n, h, w, 1)),
                          rng.uniform(0, 9, (n, h, w, 1)), np.full((n, h, w, 1), 8.0)], axis=-1)
    f = U.features(raw)
    tgt = rng.uniform(0, 3, (n, h, w, 3))
    return f, tgt


def test_param_count_matches_library():
    assert U.param_count() == 482787 + 2 * 9 * 32
    assert N.hip_lib().rc_net_param_count() == U.param_count()


def test_gradient_matches_finite_differences():
    f, tgt = small_problem()
    th = U.he_init(1) * 0.5
    _, g = U.loss_and_grad(th, f, tgt)
    rng = np.random.default_rng(3)
    errs, off = [], 0
    for _, ws, bs in U.layer_shapes():
        nw = int(np.prod(ws))
        for i in list(off + rng.integers(0, nw, 4)) + [off + nw + int(rng.integers(0, bs[0]))]:
            e = 1e-6
            tp, tm = th.copy(), th.copy()
            tp[i] += e
            tm[i] -= e
            fd = (U.loss_and_grad(tp, f, tgt)[0] - U.loss_and_grad(tm, f, tgt)[0]) / (2 * e)
            errs.append(abs(fd - g[i]) / max(abs(fd), abs(g[i]), 1e-6))
        off += nw + bs[0]
    errs = np.array(errs)
    # a ReLU kink inside +-1e-6 may spoil a sample; the bulk must agree to 1e-5
    assert np.median(errs) < 1e-6
    assert np.mean(errs < 1e-5) >= 0.9


def test_head_bias_gradient_is_exact():
    # the head bias gradient is the sum of dL/dy; a halved value would be caught
    f, tgt = small_problem(1)
    th = U.he_init(2) * 0.5
    _, g = U.loss_and_grad(th, f, tgt)
    p = U.unpack(th)
    # the last-layer bias gradient equals the sum of dL/dy: check against FD
    off = sum(int(np.prod(ws)) + bs[0] for _, ws, bs in U.layer_shapes()[:-1]) + 96
    e = 1e-6
    tp, tm = th.copy(), th.copy()
    tp[off] += e
    tm[off] -= e
    fd = (U.loss_and_grad(tp, f, tgt)[0] - U.loss_and_grad(tm, f, tgt)[0]) / (2 * e)
    assert abs(fd - g[off]) < 1e-6 * max(1.0, abs(fd))
    assert abs(fd - 0.5 * g[off]) > 1e-3 * abs(fd)  # a halved gradient would be caught
    assert p["out"][0].shape == (3, 1, 1, 32)


def test_pool_up_adjoints():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((2, 8, 6, 4))
    y = rng.standard_normal((2, 4, 3, 4))
    assert np.isclose(np.sum(U.pool(x) * y), np.sum(x * U.pool_backward(y)))
    assert np.isclose(np.sum(U.up(y) * x), np.sum(y * U.up_backward(x)))


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 3.0e38, 1e-40])
    r = U.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even
    assert r[2] == 1.0 + 2 ** -6            # tie -> even (up)
    assert r[3] == 1.0 + 2 ** -7            # exact
    assert r[4] == -2.5
    assert np.all(np.abs(r[:5] - v[:5]) <= np.abs(v[:5]) * 2 ** -8)


def test_features_closed_forms():
    raw = np.zeros((1, 12))
    raw[0, 0:3] = [0.5, 1.0, 2.0]
    raw[0, 3:6] = [0.5, 0.25, 0.0]
    raw[0, 6:9] = [0.0, 1.0, 0.0]
    raw[0, 9] = np.e - 1.0
    luma = 0.5 * 0.2126 + 1.0 * 0.7152 + 2.0 * 0.0722
    raw[0, 10] = luma * luma + 4.0 * 16.0   # sample variance of luma 64 over 16 samples -> se 2
    raw[0, 11] = 16.0
    f = U.features(raw)[0]
    assert np.allclose(f[0:3], np.log1p([1.0, 4.0, 200.0]))   # demodulated by max(albedo, 0.01)
    assert np.allclose(f[9], 0.25)
    assert np.isclose(f[10], np.log1p(2.0))
    assert np.isclose(f[11], 0.4)


def test_independent_level_inputs_are_segment_means():
    """Inputs taken as differences of two reads of one FP64 accumulation equal
    the raw inputs of the later samples on their own (mean, mean luma^2, count)."""
    from retrocam_amd.denoiser import _raw_between, _raw_from
    rng = np.random.default_rng(5)
    w, h, n0, n1 = 4, 2, 3, 7
    s = rng.uniform(0, 4, (w * h, n1, 3))
    l2 = ((s[..., 0] * 0.2126 + s[..., 1] * 0.7152) + s[..., 2] * 0.0722) ** 2
    gbuf = rng.uniform(-1, 1, (w * h, 4)).astype(np.float16).view(np.uint16)
    galb = rng.uniform(0, 1, (w * h, 4)).astype(np.float16).view(np.uint16)

    def acc(a, b):
        out = np.empty((w * h, 4))
        out[:, 0:3] = s[:, a:b].sum(axis=1)
        out[:, 3] = b - a
        return out

    got = _raw_between(acc(0, n0), l2[:, :n0].sum(1).astype(np.float32), acc(0, n1),
                       l2[:, :n1].sum(1).astype(np.float32), gbuf, galb, w, h)
    want = _raw_from(acc(n0, n1), gbuf, galb, l2[:, n0:n1].sum(1), w, h)
    assert np.allclose(got[..., 11], n1 - n0)
    assert np.allclose(got[..., 0:3], s[:, n0:n1].mean(1).reshape(h, w, 3), rtol=1e-6)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6)
