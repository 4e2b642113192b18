"""FP64 numpy oracle of the denoiser network - TEST INFRASTRUCTURE ONLY.

Restates the network of custom-ray-tracing-shader_amd/csrc/denoiser.hip (layer list,
padded parameter layout [9][coutPad][cinPad] + bias[coutPad] per layer, 2x2 average
pooling, nearest upsampling, concatenation order [upsampled | skip], albedo-demodulated
log-space residual output and L1 loss) in double precision, with a hand-written backward
pass.

Pinned by: conv2d forward against scipy.signal.correlate (an independent library routine),
every backward (layer and whole network) against central finite differences in FP64
(tests/test_oracle_denoiser.py).  The reference's own denoiser is OptiX (closed), so there
is no reference output to match; the network is judged by image error instead.
"""
import numpy as np

SPEC = [(10, 32, True), (32, 32, True), (32, 64, True), (64, 64, True), (64, 64, True), (64, 64, True),
        (64, 96, True), (96, 96, True), (96, 128, True), (128, 128, True), (224, 96, True), (96, 96, True),
        (160, 64, True), (64, 64, True), (128, 64, True), (64, 64, True), (96, 32, True), (32, 32, True),
        (32, 3, False)]
NAMES = ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "ba", "bb", "d3a", "d3b", "d2a", "d2b", "d1a", "d1b",
         "d0a", "d0b", "out"]


def pad32(c):
    return (c + 31) // 32 * 32


def layout():
    """[(name, cin, cout, cinPad, coutPad, relu, wOff, bOff)] and the parameter count."""
    out, off = [], 0
    for name, (cin, cout, relu) in zip(NAMES, SPEC):
        ci, co = pad32(cin), pad32(cout)
        w_off = off
        off += 9 * co * ci
        b_off = off
        off += co
        out.append((name, cin, cout, ci, co, relu, w_off, b_off))
    return out, off


def unpack(params):
    lay, n = layout()
    assert params.size == n
    layers = {}
    for name, cin, cout, ci, co, relu, w_off, b_off in lay:
        layers[name] = (params[w_off:w_off + 9 * co * ci].reshape(9, co, ci).astype(np.float64),
                        params[b_off:b_off + co].astype(np.float64), relu, ci, co)
    return layers


# ------------------------------------------------------------------ layer primitives
def conv3x3(x, w, b):
    """x [n,h,w,ci], w [9,co,ci], b [co]; zero padding; tap = 3*(dy+1) + (dx+1)."""
    n, h, wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((n, h, wd, w.shape[1]))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        out += xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + wd, :] @ w[tap].T
    return out + b


def conv3x3_backward(x, w, g):
    """Gradients of conv3x3 w.r.t. x, w and b for upstream gradient g [n,h,w,co]."""
    n, h, wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    gxp = np.zeros_like(xp)
    gw = np.zeros_like(w)
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        sl = (slice(None), slice(1 + dy, 1 + dy + h), slice(1 + dx, 1 + dx + wd), slice(None))
        gxp[sl] += g @ w[tap]
        gw[tap] = np.einsum("nhwo,nhwc->oc", g, xp[sl])
    return gxp[:, 1:-1, 1:-1, :], gw, g.sum(axis=(0, 1, 2))


def pool(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool_backward(g):
    return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2) * 0.25


def upsample(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def upsample_backward(g):
    n, h, w, c = g.shape
    return g.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


DEMOD_EPS = 0.01


def features(raw):
    """raw [n,h,w,10] (radiance, albedo, normal, depth) -> features [n,h,w,32], demodulated
    log radiance, modulation factors (albedo + eps)."""
    f = np.zeros(raw.shape[:3] + (32,))
    m = np.maximum(raw[..., 3:6], 0) + DEMOD_EPS
    f[..., 0:3] = np.log1p(np.maximum(raw[..., 0:3], 0) / m)
    f[..., 3:9] = raw[..., 3:9]
    f[..., 9] = np.log1p(np.maximum(raw[..., 9], 0))
    return f, f[..., 0:3].copy(), m


# ------------------------------------------------------------------ network
def forward(params, raw):
    L = unpack(params)
    x0, noisy_log, demod = features(raw.astype(np.float64))
    cache = {"x0": x0, "demod": demod}

    def c(name, x):
        w, b, relu, _, _ = L[name]
        y = conv3x3(x, w, b)
        if relu:
            y = np.maximum(y, 0)
        cache[name + "_in"] = x
        cache[name] = y
        return y

    e0 = c("e0b", c("e0a", x0))
    e1 = c("e1b", c("e1a", pool(e0)))
    e2 = c("e2b", c("e2a", pool(e1)))
    e3 = c("e3b", c("e3a", pool(e2)))
    bb = c("bb", c("ba", pool(e3)))
    d3 = c("d3b", c("d3a", np.concatenate([upsample(bb), e3], axis=-1)))
    d2 = c("d2b", c("d2a", np.concatenate([upsample(d3), e2], axis=-1)))
    d1 = c("d1b", c("d1a", np.concatenate([upsample(d2), e1], axis=-1)))
    d0 = c("d0b", c("d0a", np.concatenate([upsample(d1), e0], axis=-1)))
    out = c("out", d0)
    cache["noisy_log"] = noisy_log
    return out[..., :3], cache


def infer(params, raw):
    out, cache = forward(params, raw)
    return np.maximum(np.expm1(cache["noisy_log"] + out), 0) * cache["demod"]


def display_value(L):
    """Display transform of Native/renderer.cu:86-90 at exposure 1 before 8-bit rounding,
    and its derivative w.r.t. L."""
    e = np.exp(-L)
    t = 1 - e
    lin = t <= 0.0031308
    tp = np.where(lin, 1.0, t)
    value = np.where(lin, 12.92 * t, 1.055 * tp ** (1 / 2.4) - 0.055)
    slope = np.where(lin, 12.92 * e, 1.055 / 2.4 * tp ** (1 / 2.4 - 1) * e)
    return value, slope


def loss_and_grads(params, raw, target, display_weight=0.0):
    """Mean over pixels and RGB of |y - log target| + display_weight * |D(L(y)) - D(target)|
    (y the demodulated log prediction, L(y) = max(expm1(y), 0) * demod, D the display
    transform) and its gradient w.r.t. params."""
    L = unpack(params)
    out, cache = forward(params, raw)
    tgt_log = np.log1p(np.maximum(target.astype(np.float64), 0) / cache["demod"])
    y = cache["noisy_log"] + out
    d = y - tgt_log
    count = d.size
    loss = np.abs(d).sum() / count
    g_y = np.sign(d) / count
    if display_weight > 0:
        m = cache["demod"]
        em = np.expm1(y)
        dp, sp = display_value(np.maximum(em, 0) * m)
        dt, _ = display_value(np.maximum(np.expm1(tgt_log), 0) * m)
        dd = dp - dt
        loss += display_weight * np.abs(dd).sum() / count
        g_y = g_y + np.where(em > 0, np.sign(dd) * display_weight * sp * (em + 1) * m, 0.0) / count
    grads = np.zeros(params.size)
    lay = {row[0]: row for row in layout()[0]}

    def back(name, g):
        w, b, relu, ci, co = L[name]
        if relu:
            g = g * (cache[name] > 0)
        gx, gw, gb = conv3x3_backward(cache[name + "_in"], w, g)
        _, _, _, _, _, _, w_off, b_off = lay[name]
        grads[w_off:w_off + gw.size] = gw.reshape(-1)
        grads[b_off:b_off + gb.size] = gb
        return gx

    g_out = np.zeros(out.shape[:3] + (32,))
    g_out[..., :3] = g_y
    g = back("d0a", back("d0b", back("out", g_out)))             # g w.r.t. concat [up(d1) | e0]
    g_e0 = g[..., 64:96]
    g = back("d1a", back("d1b", upsample_backward(g[..., :64])))
    g_e1 = g[..., 64:128]
    g = back("d2a", back("d2b", upsample_backward(g[..., :64])))
    g_e2 = g[..., 96:160]
    g = back("d3a", back("d3b", upsample_backward(g[..., :96])))
    g_e3 = g[..., 128:224]
    g = back("ba", back("bb", upsample_backward(g[..., :128])))
    g_e3 = g_e3 + pool_backward(g)
    g = back("e3a", back("e3b", g_e3))
    g_e2 = g_e2 + pool_backward(g)
    g = back("e2a", back("e2b", g_e2))
    g_e1 = g_e1 + pool_backward(g)
    g = back("e1a", back("e1b", g_e1))
    g_e0 = g_e0 + pool_backward(g)
    back("e0a", back("e0b", g_e0))
    return loss, grads


def init_params(seed=1):
    """He-normal weights over the real fan-in, zero biases and zero padding (like the GPU)."""
    rng = np.random.default_rng(seed)
    lay, n = layout()
    p = np.zeros(n, np.float32)
    for name, cin, cout, ci, co, relu, w_off, b_off in lay:
        std = np.sqrt(2.0 / (9 * cin)) * (0.1 if name == "out" else 1.0)
        w = np.zeros((9, co, ci), np.float32)
        w[:, :cout, :cin] = rng.normal(0, std, (9, cout, cin))
        p[w_off:w_off + w.size] = w.reshape(-1)
    return p


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)
