"""CPU restatement of the learned coarse-step corrector, FP64 numpy.

TEST INFRASTRUCTURE ONLY: the checker for the corrector entry points of
libfilament_hip.so (fil_corrector_*).  Only tests/ import it.

The corrector has no counterpart in the reference, so nothing of the
reference pins it.  Its checks are (tests/test_corrector_oracle.py):
  - conv3x3 against a brute-force loop over pixels, taps and channels;
  - loss_and_grad against central finite differences of the loss;
  - features against closed forms (a plane carrier exp(-i w0 t) times a
    constant gives A = constant; block averages of known ramps);
  - bf16_round against the bit-level definition (round to nearest even of the
    upper 16 bits).

Conventions (same as the device):
  X [B][4][H][W]: Re A / a_s, Im A / a_s, |E|^2 / a_s^2, rho / rho_s, each
  averaged over pr x pt blocks of the window rows [0, nr) and samples
  [t0, t0 + nt), with A = E exp(i w0 t_j), t_j = tmin + j dt.
  Network: 4 layers of 3x3 convolutions with zero padding,
  4 -> 32 -> 32 -> 32 -> 2 channels, ReLU after the first three; weights
  rounded to bf16, everything else in FP64 here (FP32 on the device).
  Loss: mean over B*2*H*W of (out - T / y_s)^2.
"""
import numpy as np

CH = 32
LAYERS = [(4, CH), (CH, CH), (CH, CH), (CH, 2)]   # (cin, cout)


def nparam():
    return sum(co * ci * 9 + co for ci, co in LAYERS)


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, round_weights=True):
    """Flat parameter vector -> [(W [cout][cin][3][3], b [cout])] per layer."""
    p = np.asarray(params, dtype=np.float64)
    out, o = [], 0
    for ci, co in LAYERS:
        n = co * ci * 9
        W = p[o:o + n]
        if round_weights:
            W = bf16_round(W)
        out.append((W.reshape(co, ci, 3, 3), p[o + n:o + n + co].copy()))
        o += n + co
    return out


def conv3x3(x, W, b):
    """x [B][cin][H][W] -> [B][cout][H][W], zero padding."""
    B, ci, H, Wd = x.shape
    xp = np.zeros((B, ci, H + 2, Wd + 2))
    xp[:, :, 1:-1, 1:-1] = x
    out = np.zeros((B, W.shape[0], H, Wd)) + b[None, :, None, None]
    for dy in range(3):
        for dx in range(3):
            out += np.einsum("oc,bchw->bohw", W[:, :, dy, dx], xp[:, :, dy:dy + H, dx:dx + Wd])
    return out


def conv3x3_transpose(g, W):
    """Adjoint of conv3x3 in its input: g [B][cout][H][W] -> [B][cin][H][W]."""
    B, co, H, Wd = g.shape
    out = np.zeros((B, W.shape[1], H + 2, Wd + 2))
    for dy in range(3):
        for dx in range(3):
            out[:, :, dy:dy + H, dx:dx + Wd] += np.einsum("oc,bohw->bchw", W[:, :, dy, dx], g)
    return out[:, :, 1:-1, 1:-1]


def conv3x3_weight_grad(g, x):
    """d/dW of sum(g * conv3x3(x, W, 0)): [cout][cin][3][3]."""
    B, ci, H, Wd = x.shape
    xp = np.zeros((B, ci, H + 2, Wd + 2))
    xp[:, :, 1:-1, 1:-1] = x
    dW = np.zeros((g.shape[1], ci, 3, 3))
    for dy in range(3):
        for dx in range(3):
            dW[:, :, dy, dx] = np.einsum("bohw,bchw->oc", g, xp[:, :, dy:dy + H, dx:dx + Wd])
    return dW


def forward(params, X, round_weights=True):
    """Network output [B][2][H][W] and the list of layer inputs."""
    layers = unpack(params, round_weights)
    a = np.asarray(X, dtype=np.float64)
    ins = []
    for l, (W, b) in enumerate(layers):
        ins.append(a)
        z = conv3x3(a, W, b)
        a = np.maximum(z, 0.0) if l < len(layers) - 1 else z
    return a, ins


def loss_and_grad(params, X, T, y_scale, round_weights=True):
    """MSE loss and its gradient in the flat parameter layout.  With rounded
    weights the gradient is taken at the rounded values (the device trains
    the FP32 master copy with these gradients)."""
    layers = unpack(params, round_weights)
    out, ins = forward(params, X, round_weights)
    d = out - np.asarray(T, dtype=np.float64) / y_scale
    n = d.size
    loss = float(np.sum(d * d) / n)
    g = 2.0 * d / n
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads[l] = (conv3x3_weight_grad(g, ins[l]), g.sum(axis=(0, 2, 3)))
        if l > 0:
            g = conv3x3_transpose(g, W) * (ins[l] > 0)
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return loss, flat


def features(E, rho, geom):
    """E [Nr][Nt] complex time-domain field, rho [Nr][Nt] (or None) -> X [4][H][W]."""
    nr, t0, nt, pr, pt = geom["nr"], geom["t0"], geom["nt"], geom["pool_r"], geom["pool_t"]
    H, W = nr // pr, nt // pt
    j = np.arange(t0, t0 + nt)
    t = geom["tmin"] + j * geom["dt"]
    Ew = np.asarray(E)[:nr, t0:t0 + nt]
    A = Ew * np.exp(1j * geom["w0"] * t)[None, :]
    I = np.abs(Ew) ** 2
    R = np.zeros_like(I) if rho is None else np.asarray(rho)[:nr, t0:t0 + nt]

    def pool(a):
        return a.reshape(H, pr, W, pt).mean(axis=(1, 3))

    a_s, r_s = geom["a_scale"], geom["rho_scale"]
    return np.stack([pool(A.real) / a_s, pool(A.imag) / a_s, pool(I) / a_s ** 2, pool(R) / r_s])


def upsample(D, geom):
    """Bilinear interpolation of pooled planes D [c][H][W] to the window
    samples [c][nr][nt], from the block centres, clamped at the edges."""
    nr, nt, pr, pt = geom["nr"], geom["nt"], geom["pool_r"], geom["pool_t"]
    H, W = nr // pr, nt // pt
    y = np.clip((np.arange(nr) + 0.5) / pr - 0.5, 0, H - 1)
    x = np.clip((np.arange(nt) + 0.5) / pt - 0.5, 0, W - 1)
    y0 = np.floor(y).astype(int)
    x0 = np.floor(x).astype(int)
    y1, x1 = np.minimum(y0 + 1, H - 1), np.minimum(x0 + 1, W - 1)
    fy, fx = (y - y0)[:, None], (x - x0)[None, :]
    D = np.asarray(D, dtype=np.float64)
    a = D[:, y0][:, :, x0] + fx * (D[:, y0][:, :, x1] - D[:, y0][:, :, x0])
    b = D[:, y1][:, :, x0] + fx * (D[:, y1][:, :, x1] - D[:, y1][:, :, x0])
    return a + fy * (b - a)


def apply(E, D, geom, scale):
    """Time-domain field with the correction scale * (D0 + i D1) exp(-i w0 t)
    added in the window (before the projection onto positive frequencies)."""
    E = np.array(E, dtype=np.complex128)
    nr, t0, nt = geom["nr"], geom["t0"], geom["nt"]
    up = upsample(D, geom)
    t = geom["tmin"] + np.arange(t0, t0 + nt) * geom["dt"]
    E[:nr, t0:t0 + nt] += scale * (up[0] + 1j * up[1]) * np.exp(-1j * geom["w0"] * t)[None, :]
    return E
