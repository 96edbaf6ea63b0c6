"""CPU restatement of the learned preconditioner of the implicit field solve.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's CPU baseline as
the checker of the HIP kernels, never by the library.

The reference has no preconditioner: field.f90:319-461 (cgm) runs plain CG
on (f4 I - N4) phi = f5 gkl per component, N4 = sum of the four periodic
neighbours.  This module states the learned addition in numpy:

* ``forward`` -- a four-level linear convolutional net (V-cycle shaped):
      x0 = r / s                           s = rms(r) per image
      a0 = lift(x0)        3x3, 1 -> C0    (fine level, FP32, kept unrounded)
      d1 = pool(a0)                         2x2 mean, rounded to BF16
      a1 = conv(d1; D1)    C0 -> C1        level 1
      d2 = pool(a1); a2 = conv(d2; D2)      C1 -> C2, level 2
      d3 = pool(a2); a3 = conv(d3; D3)      C2 -> C2, level 3
      u2 = a2 + up(a3);  b2 = conv(u2; U2)  C2 -> C1
      u1 = a1 + up(b2);  b1 = conv(u1; U1)  C1 -> C0
      u0 = a0 + up(b1);  y = proj(u0)       3x3, C0 -> 1 (FP32)
      z  = s * y
  Every 3x3 convolution is periodic.  The MFMA convolutions take BF16
  operands and accumulate in FP32; their outputs are rounded to BF16, as are
  pool and up+add outputs.  ``bf16=False`` runs the same graph in float64
  without rounding (used to pin the gradients by finite differences).
* ``backward`` -- gradients of L = mean((y - e/s)^2) over all pixels of the
  batch with respect to every weight, with the rounding points of the device
  kernels (incoming gradients of an MFMA convolution are rounded to BF16).
* ``adam`` -- the optimiser step.
* ``fcg`` -- flexible CG with the net as preconditioner and plain CG as the
  fallback and as the judge of convergence.

Parity pins (tests/test_pc_oracle.py): finite differences of the float64
graph pin ``backward``; a net set to the identity stencil reduces ``fcg`` to
textbook CG; the trained net's iteration gain is measured against the plain
solver of oracle/wm_oracle.c (or_cgm).
"""
import numpy as np

C0, C1, C2 = 16, 16, 32
# (name, cin, cout); lift and proj are the FP32 fine-level convolutions,
# the rest run on MFMA with BF16 operands
LAYERS = [("lift", 1, C0), ("d1", C0, C1), ("d2", C1, C2), ("d3", C2, C2),
          ("u2", C2, C1), ("u1", C1, C0), ("proj", C0, 1)]
MFMA_LAYERS = ("d1", "d2", "d3", "u2", "u1")


def layout():
    """Offsets of each layer's weights in the flat FP32 parameter vector.
    Every layer is stored [tap][cout][cin], tap = (dy+1)*3 + (dx+1)."""
    off, o = {}, 0
    for name, ci, co in LAYERS:
        off[name] = (o, ci, co)
        o += 9 * ci * co
    return off, o


OFFSETS, NPARAM = layout()


def unpack(theta):
    return {n: theta[o:o + 9 * ci * co].reshape(9, co, ci) for n, (o, ci, co) in OFFSETS.items()}


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), returned as
    float32."""
    x = np.asarray(x, dtype=np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def _shift(x, dy, dx):
    """x[b, j+dy, i+dx] with periodic wrap (x is [B, H, W, ...])."""
    return np.roll(x, shift=(-dy, -dx), axis=(1, 2))


TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def conv(x, w, acc=np.float64):
    """Periodic 3x3 convolution: out[b,j,i,co] = sum_tap sum_ci
    x[b, j+dy, i+dx, ci] * w[tap, co, ci]."""
    out = 0
    for t, (dy, dx) in enumerate(TAPS):
        out = out + _shift(x, dy, dx).astype(acc) @ w[t].T.astype(acc)
    return out


def conv_t(g, w, acc=np.float64):
    """Adjoint of conv with respect to its input."""
    out = 0
    for t, (dy, dx) in enumerate(TAPS):
        out = out + _shift(g.astype(acc) @ w[t].astype(acc), -dy, -dx)
    return out


def conv_dw(x, g, acc=np.float64):
    """Adjoint of conv with respect to its weights: [tap, cout, cin]."""
    x = x.astype(acc)
    g = g.astype(acc)
    B, H, W, ci = x.shape
    co = g.shape[-1]
    gf = g.reshape(-1, co)
    return np.stack([gf.T @ _shift(x, dy, dx).reshape(-1, ci) for dy, dx in TAPS])


def pool(x):
    """2x2 mean: ((x00 + x01) + (x10 + x11)) * 0.25."""
    return ((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + (x[:, 1::2, 0::2] + x[:, 1::2, 1::2])) * 0.25


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_t(g):
    """Adjoint of up: sum over each 2x2 block, ((g00 + g01) + (g10 + g11))."""
    return (g[:, 0::2, 0::2] + g[:, 0::2, 1::2]) + (g[:, 1::2, 0::2] + g[:, 1::2, 1::2])


def scales(r):
    """Per-image rms of r ([B, H, W]); 1 for an all-zero image."""
    s = np.sqrt((r.astype(np.float64) ** 2).reshape(r.shape[0], -1).mean(axis=1))
    return np.where(s > 0, s, 1.0)


def forward(theta, r, bf16_path=True):
    """Net output z ([B, H, W] float64) for residuals r and the cached
    intermediates the backward pass needs."""
    W = unpack(np.asarray(theta, dtype=np.float32 if bf16_path else np.float64))
    R = (lambda v: bf16(v)) if bf16_path else (lambda v: v)
    ft = np.float32 if bf16_path else np.float64
    s = scales(r)
    x0 = (r / s[:, None, None]).astype(ft)[..., None]
    a0 = conv(x0, W["lift"]).astype(ft)
    Wm = {n: (bf16(W[n]) if bf16_path else W[n]) for n in MFMA_LAYERS}
    d1 = R(pool(a0).astype(ft))
    a1 = R(conv(d1, Wm["d1"]).astype(ft))
    d2 = R(pool(a1).astype(ft))
    a2 = R(conv(d2, Wm["d2"]).astype(ft))
    d3 = R(pool(a2).astype(ft))
    a3 = R(conv(d3, Wm["d3"]).astype(ft))
    u2 = R((a2 + up(a3)).astype(ft))
    b2 = R(conv(u2, Wm["u2"]).astype(ft))
    u1 = R((a1 + up(b2)).astype(ft))
    b1 = R(conv(u1, Wm["u1"]).astype(ft))
    u0 = (a0 + up(b1)).astype(ft)
    y = conv(u0, W["proj"]).astype(ft)[..., 0]
    z = y.astype(np.float64) * s[:, None, None]
    cache = dict(s=s, x0=x0, a0=a0, d1=d1, a1=a1, d2=d2, a2=a2, d3=d3, a3=a3, u2=u2, b2=b2,
                 u1=u1, b1=b1, u0=u0, y=y, W=W, Wm=Wm, R=R, ft=ft)
    return z, cache


def loss_grad(theta, r, e, bf16_path=True):
    """L = mean over all pixels of (y - e/s)^2 and dL/dtheta (flat)."""
    z, c = forward(theta, r, bf16_path)
    R, ft = c["R"], c["ft"]
    s = c["s"]
    tgt = (e / s[:, None, None])
    diff = c["y"].astype(np.float64) - tgt
    n = diff.size
    L = float((diff ** 2).sum() / n)
    dy = (2.0 * diff / n).astype(ft)[..., None]
    g = {}
    g["proj"] = conv_dw(c["u0"], dy)
    du0 = conv_t(dy, c["W"]["proj"]).astype(ft)
    db1 = R(up_t(du0).astype(ft))
    g["u1"] = conv_dw(c["u1"], db1)
    du1 = conv_t(db1, c["Wm"]["u1"]).astype(ft)
    db2 = R(up_t(du1).astype(ft))
    g["u2"] = conv_dw(c["u2"], db2)
    du2 = conv_t(db2, c["Wm"]["u2"]).astype(ft)
    da3 = R(up_t(du2).astype(ft))
    g["d3"] = conv_dw(c["d3"], da3)
    dd3 = conv_t(da3, c["Wm"]["d3"]).astype(ft)
    da2 = R((du2 + 0.25 * up(dd3)).astype(ft))
    g["d2"] = conv_dw(c["d2"], da2)
    dd2 = conv_t(da2, c["Wm"]["d2"]).astype(ft)
    da1 = R((du1 + 0.25 * up(dd2)).astype(ft))
    g["d1"] = conv_dw(c["d1"], da1)
    dd1 = conv_t(da1, c["Wm"]["d1"]).astype(ft)
    da0 = (du0 + 0.25 * up(dd1)).astype(ft)
    g["lift"] = conv_dw(c["x0"], da0)
    flat = np.zeros(NPARAM, dtype=np.float64)
    for name, (o, ci, co) in OFFSETS.items():
        flat[o:o + 9 * ci * co] = g[name].reshape(-1)
    return L, flat


def init_params(f4, seed=0):
    """Initial weights: lift/proj carry the 1/f4 identity stencil through
    channel 0, the rest small random values (fan-in scaled)."""
    rng = np.random.default_rng(seed)
    th = np.zeros(NPARAM, dtype=np.float32)
    for name, (o, ci, co) in OFFSETS.items():
        w = rng.standard_normal((9, co, ci)) * (0.1 / np.sqrt(9 * ci))
        if name == "lift":
            w[4, 0, 0] = 1.0
        if name == "proj":
            w[4, 0, 0] = 1.0 / f4
        th[o:o + 9 * ci * co] = w.reshape(-1)
    return th


class Adam:
    """Adam (Kingma & Ba 2015) with bias correction, FP32 state."""

    def __init__(self, n, lr=2e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n, np.float32)
        self.v = np.zeros(n, np.float32)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, theta, g):
        self.t += 1
        g = g.astype(np.float32)
        self.m = (self.b1 * self.m + (1 - self.b1) * g).astype(np.float32)
        self.v = (self.b2 * self.v + (1 - self.b2) * g * g).astype(np.float32)
        mh = self.m / np.float32(1 - self.b1 ** self.t)
        vh = self.v / np.float32(1 - self.b2 ** self.t)
        return (theta - np.float32(self.lr) * mh / (np.sqrt(vh) + np.float32(self.eps))).astype(np.float32)


# ------------------------------------------------------------------ solver
def apply_a(x, f4):
    """(f4 I - N4) x on the periodic grid (field.f90:400-402)."""
    return f4 * x - (np.roll(x, 1, 1) + np.roll(x, -1, 1) + np.roll(x, 1, 2) + np.roll(x, -1, 2))


def cg(b, x, f4, tol, maxit=1000):
    """Textbook CG per image; stops when |r| <= tol*|b|.  Returns x,
    iteration counts and the (r_k, x_k) snapshots."""
    B = b.shape[0]
    its = np.zeros(B, int)
    snaps = []
    for k in range(B):
        xk = x[k].copy()
        r = b[k] - apply_a(xk[None], f4)[0]
        eps = tol * np.linalg.norm(b[k])
        p = r.copy()
        rr = (r * r).sum()
        while np.sqrt(rr) > eps and its[k] < maxit:
            snaps.append((k, r.copy(), xk.copy()))
            q = apply_a(p[None], f4)[0]
            a = rr / (p * q).sum()
            xk += a * p
            r -= a * q
            rn = (r * r).sum()
            p = r + (rn / rr) * p
            rr = rn
            its[k] += 1
        x[k] = xk
    return x, its, snaps


def fcg(b, x, f4, tol, precond, maxit=200):
    """Flexible CG (Notay 2000, Polak-Ribiere beta) per image with a
    preconditioner callable precond(r[None]) -> z[None]; stops when
    |r| <= tol*|b|.  Returns x and iteration counts."""
    B = b.shape[0]
    its = np.zeros(B, int)
    for k in range(B):
        xk = x[k].copy()
        r = b[k] - apply_a(xk[None], f4)[0]
        eps = tol * np.linalg.norm(b[k])
        zp = rp = p = None
        while np.linalg.norm(r) > eps and its[k] < maxit:
            z = precond(r[None])[0]
            if p is None:
                p = z.copy()
            else:
                beta = (z * (r - rp)).sum() / (zp * rp).sum()
                p = z + beta * p
            q = apply_a(p[None], f4)[0]
            a = (r * z).sum() / (p * q).sum()
            xk += a * p
            rp, zp = r.copy(), z
            r = r - a * q
            its[k] += 1
        x[k] = xk
    return x, its
