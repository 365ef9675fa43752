"""CPU restatement of the learned flux closure (test infrastructure only).

Only tests/ may import this module; it is the reference the HIP closure
kernels are compared with, never part of the product path.  The closure has no counterpart in
the reference: it corrects a coarse run of the per-block RK update
(src/scheme.rs:323-349) by a flux divergence predicted from the state at the
start of each step.  Everything here is FP64 numpy.

Net (grid arrays are [i][j][channel], i the x index, j the y index, periodic):
  X0 = features(u)                      4 channels, zero-padded to 32
  A1 = relu(conv(X0, W0) + b0)          32 channels
  A2 = relu(conv(A1, W1) + b1)
  A3 = relu(conv(A2, W2) + b2)
  Y  = conv(A3, W3) + b3                6 channels: dFx(S, Sx, Sy), dFy(S, Sx, Sy)
  conv(X, W)[i, j, co] = sum_{ky, kx, ci} X[i + ky - 1, j + kx - 1, ci] W[co, ci, ky, kx]
Correction of one step of length dt (v = 0, 1, 2: Sigma, Sx, Sy):
  r_v = -s_v ((Y_v[i, j] - Y_v[i-1, j]) / dx + (Y_{3+v}[i, j] - Y_{3+v}[i, j-1]) / dy)
  u_v += dt r_v
with each face's three corrections scaled by a positivity limiter when one
is set (`limiter`); training uses the unlimited r.
Training loss on a sample (features of the base state, target T, cell
weights c = 1 where the base density is at least the mask density, else 0):
  L = (1/P) sum_{cells, v} w_v c^2 (r_v - T_v)^2
"""
import numpy as np

HIDDEN = 32
LAYERS = [(4, 32), (32, 32), (32, 32), (32, 6)]   # (cin, cout) per conv layer


def param_shapes():
    out = []
    for cin, cout in LAYERS:
        out.append(("W", (cout, cin, 3, 3)))
        out.append(("b", (cout,)))
    return out


def num_params():
    return sum(int(np.prod(s)) for _, s in param_shapes())


def unpack(flat):
    """Canonical flat parameter vector -> [(W, b)] per layer."""
    flat = np.asarray(flat, dtype=np.float64)
    out, k = [], 0
    for cin, cout in LAYERS:
        n = cout * cin * 9
        W = flat[k:k + n].reshape(cout, cin, 3, 3)
        k += n
        b = flat[k:k + cout]
        k += cout
        out.append((W, b))
    assert k == flat.size
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def features(u_global, nv, gamma, scales, bodies=None, R=None, rs2=None):
    """u_global: [N][N][nv] conserved.  Channels: a0 ln Sigma, a1 vx, a1 vy,
    and a2 ln(p / Sigma) (Euler) or a2 ln(-phi) (isothermal, 0 without
    `bodies`): phi = -sum_k m_k / sqrt(|x - x_k|^2 + rs2) at the cell centres
    x = -R + (i + 1/2) dx, bodies = ((m1, x1, y1), (m2, x2, y2)).  Arguments
    of the logs are floored at 1e-30 and every feature is clamped to
    [-16, 16]."""
    a0, a1, a2 = scales[0], scales[1], scales[2]
    sig = u_global[..., 0]
    vx = u_global[..., 1] / sig
    vy = u_global[..., 2] / sig
    f = np.zeros(u_global.shape[:2] + (4,))
    f[..., 0] = a0 * np.log(np.maximum(sig, 1e-30))
    f[..., 1] = a1 * vx
    f[..., 2] = a1 * vy
    if nv == 4:
        p = (u_global[..., 3] - 0.5 * sig * (vx * vx + vy * vy)) * (gamma - 1.0)
        f[..., 3] = a2 * np.log(np.maximum(p / sig, 1e-30))
    elif bodies is not None:
        N = sig.shape[0]
        dx = 2.0 * R / N
        c = -R + (np.arange(N) + 0.5) * dx
        x, y = np.meshgrid(c, c, indexing="ij")
        mphi = sum(m / np.sqrt((x - bx) ** 2 + (y - by) ** 2 + rs2) for m, bx, by in bodies)
        f[..., 3] = a2 * np.log(np.maximum(mphi, 1e-30))
    return np.clip(f, -16.0, 16.0)


def conv(x, W, b=None):
    """3x3 periodic convolution: x [N][M][cin], W [cout][cin][3][3]."""
    out = np.zeros(x.shape[:2] + (W.shape[0],))
    for ky in range(3):
        for kx in range(3):
            xs = np.roll(np.roll(x, -(ky - 1), axis=0), -(kx - 1), axis=1)
            out += xs @ W[:, :, ky, kx].T
    if b is not None:
        out += b
    return out


def conv_input_grad(g, W):
    """d conv / d x contracted with g: the transposed periodic convolution."""
    out = np.zeros(g.shape[:2] + (W.shape[1],))
    for ky in range(3):
        for kx in range(3):
            t = g @ W[:, :, ky, kx]
            out += np.roll(np.roll(t, ky - 1, axis=0), kx - 1, axis=1)
    return out


def conv_weight_grad(g, x):
    """dW[co, ci, ky, kx] = sum_p g[p, co] x[p + (ky-1, kx-1), ci]; db = sum_p g."""
    cout, cin = g.shape[-1], x.shape[-1]
    dW = np.zeros((cout, cin, 3, 3))
    for ky in range(3):
        for kx in range(3):
            xs = np.roll(np.roll(x, -(ky - 1), axis=0), -(kx - 1), axis=1)
            dW[:, :, ky, kx] = np.einsum("ijc,ijd->cd", g, xs)
    return dW, g.sum(axis=(0, 1))


def forward(layers, x0, rnd=None):
    """rnd (optional) rounds each stored hidden activation, as a device that
    keeps them in bf16 does (the output Y is not rounded)."""
    acts = [x0]
    h = x0
    for k, (W, b) in enumerate(layers):
        z = conv(h, W, b)
        if k < len(layers) - 1:
            h = np.maximum(z, 0.0)
            if rnd is not None:
                h = rnd(h)
        else:
            h = z
        acts.append(h)
    return acts          # [X0, A1, A2, A3, Y]


def limiter(Y, s, dx, dy, dt, u, kappa, nv=4, gamma=5.0 / 3.0, mask=0.0):
    """Face factors (x+, y+ face of each cell): the largest lambda <= 1 with
    dt |s_0 Y_Sigma| / h <= kappa min(Sigma_L, Sigma_R) and, for the momentum
    channels, dt |s_v Y_v| / h <= kappa min(Sigma_L, Sigma_R) max(a_L, a_R),
    a = |v| + c_s the cells' signal speed (u: [N][N][nv] conserved); faces
    where min(Sigma_L, Sigma_R) < mask get 0."""
    lam = np.ones(Y.shape[:2] + (2,))
    sig = u[..., 0]
    for d, axis in enumerate((0, 1)):
        lam[..., d] = np.where(np.minimum(sig, np.roll(sig, -1, axis=axis)) < mask, 0.0, 1.0)
    if kappa <= 0:
        return lam
    vx, vy = u[..., 1] / sig, u[..., 2] / sig
    c2 = np.zeros_like(sig)
    if nv == 4:
        c2 = gamma * (u[..., 3] - 0.5 * sig * (vx * vx + vy * vy)) * (gamma - 1.0) / sig
    a = np.sqrt(vx * vx + vy * vy) + np.sqrt(np.maximum(c2, 0.0))
    for d, (axis, h) in enumerate(((0, dx), (1, dy))):
        cap = kappa * np.maximum(np.minimum(sig, np.roll(sig, -1, axis=axis)), 0.0)
        capm = cap * np.maximum(a, np.roll(a, -1, axis=axis))
        l = lam[..., d].copy()
        for v in range(3):
            m = dt * np.abs(s[v] * Y[..., 3 * d + v]) / h
            c = cap if v == 0 else capm
            l = np.where(m > c, np.minimum(l, c / np.where(m > c, m, 1.0)), l)
        lam[..., d] = l
    return lam


def correction(Y, s, dx, dy, lam=None):
    """r_v, v = 0..2, from the 6 output channels (face factors lam optional)."""
    r = np.zeros(Y.shape[:2] + (3,))
    lx = 1.0 if lam is None else lam[..., 0]
    ly = 1.0 if lam is None else lam[..., 1]
    for v in range(3):
        fx, fy = lx * Y[..., v], ly * Y[..., 3 + v]
        r[..., v] = -s[v] * ((fx - np.roll(fx, 1, axis=0)) / dx + (fy - np.roll(fy, 1, axis=1)) / dy)
    return r


def loss_and_grads(layers, x0, T, s, w, dx, dy, rnd=None, cw=None):
    """Loss and canonical flat gradient (analytic backward).  rnd (optional)
    rounds the stored activations, dL/dY and every propagated activation
    gradient, emulating a device that keeps them in bf16.  cw (optional,
    [N][N]) weights each cell's error (the loss over masked cells)."""
    acts = forward(layers, x0, rnd)
    Y = acts[-1]
    P = Y.shape[0] * Y.shape[1]
    e = correction(Y, s, dx, dy) - T[..., :3]
    if cw is not None:
        e = e * np.asarray(cw)[..., None]
    L = float(np.sum(np.asarray(w)[None, None, :] * e * e) / P)
    g = 2.0 * np.asarray(w) * e / P                 # dL/dr (times the cell weight)
    if cw is not None:
        g = g * np.asarray(cw)[..., None]
    gY = np.zeros_like(Y)
    for v in range(3):
        # r_v[i,j] = -s_v (fx[i,j] - fx[i-1,j]) / dx - ...
        gY[..., v] = -s[v] / dx * (g[..., v] - np.roll(g[..., v], -1, axis=0))
        gY[..., 3 + v] = -s[v] / dy * (g[..., v] - np.roll(g[..., v], -1, axis=1))
    grads = [None] * len(layers)
    gh = gY if rnd is None else rnd(gY)
    for k in range(len(layers) - 1, -1, -1):
        W, _ = layers[k]
        dW, db = conv_weight_grad(gh, acts[k])
        grads[k] = (dW, db)
        if k > 0:
            gh = conv_input_grad(gh, W) * (acts[k] > 0.0)
            if rnd is not None:
                gh = rnd(gh)
    return L, pack(grads), Y


def adam(params, grads, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1.0 - b1) * grads
    v = b2 * v + (1.0 - b2) * grads * grads
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32).copy()
    bits = a.view(np.uint32).astype(np.uint64)
    bits = (bits + 0x7FFF + ((bits >> 16) & 1)) >> 16 << 16
    return bits.astype(np.uint32).view(np.float32).astype(np.float64)
