"""Independent numpy statements of the mathematics the oracle must satisfy:
direct lapped-transform sums, the bicomplex algebra, Bayer masks.  Used by
the oracle pin tests (not a restatement of the oracle's fold/butterfly
algorithm)."""
import numpy as np


def bicomplex_mul(a, b):
    """Product in the algebra with basis (1, i, j, ij), i^2 = j^2 = -1, ij = ji;
    component p of an element is a[p] (bit 0: i, bit 1: j)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    out = np.zeros(np.broadcast_shapes(a.shape, b.shape))
    for p in range(4):
        for q in range(4):
            sgn = (-1) ** ((p & 1) & (q & 1)) * (-1) ** ((p >> 1) & (q >> 1))
            out[p ^ q] += sgn * a[p] * b[q]
    return out


def bicomplex_conj(a):
    """Conjugate both i and j (ij keeps its sign)."""
    a = np.array(a, np.float64)
    a[1] = -a[1]
    a[2] = -a[2]
    return a


def bayer_mask(color, rows, cols):
    """GR/BG mosaic, red at (row 0, col 1): R=0, B=1, G=2 (absolute coordinates)."""
    ay = np.asarray(rows)[:, None]
    ax = np.asarray(cols)[None, :]
    if color == 0:
        return (ay % 2 == 0) & (ax % 2 == 1)
    if color == 1:
        return (ay % 2 == 1) & (ax % 2 == 0)
    return (ay + ax) % 2 == 0


def clamp_bayer(v, n):
    v = np.asarray(v)
    return np.where(v >= n, n - 2 + (v & 1), np.where(v < 0, v & 1, v))


def lapped_basis(s):
    """[2][8][16] (cos, sin) rows: 1/2 w(n+s) T(pi/8 (n+4.5)(k+1/2)), w(x) = sin(pi (x+1/2)/16)."""
    n = np.arange(16)[None, :]
    k = np.arange(8)[:, None]
    w = np.sin(np.pi * (n + 0.5 + s) / 16)
    arg = np.pi / 8 * (n + 4.5) * (k + 0.5)
    return np.stack([0.5 * w * np.cos(arg), 0.5 * w * np.sin(arg)])


def mclt_direct(patch, sh, sv):
    """4 quadrants [q][l][k] of the shifted-window lapped transform of a
    16x16 patch: q bit 0 = horizontal sine, bit 1 = vertical sine."""
    bh = lapped_basis(sh)
    bv = lapped_basis(sv)
    q = np.zeros((4, 8, 8))
    for qq in range(4):
        q[qq] = bv[qq >> 1] @ patch @ bh[qq & 1].T
    return q


def phasor(sh, sv):
    """Bicomplex e^{i pi sh (2k+1)/16} e^{j pi sv (2l+1)/16} as [4][8][8]."""
    k = np.arange(8)
    th = np.pi * sh * (2 * k + 1) / 16
    tv = np.pi * sv * (2 * k + 1) / 16
    ph = np.zeros((4, 8, 8))
    ph[0] = np.cos(th)[None, :]
    ph[1] = np.sin(th)[None, :]
    pv = np.zeros((4, 8, 8))
    pv[0] = np.cos(tv)[:, None]
    pv[2] = np.sin(tv)[:, None]
    return bicomplex_mul(ph, pv)


def forward_direct(img, color, itlx, itly, sh, sv, kernel=None):
    """Expected convert_correct_tiles output for one color, given the window
    position and residual shifts (float64)."""
    H, W = img.shape
    r = np.arange(16)
    Y = clamp_bayer(itly + r, H)
    X = clamp_bayer(itlx + r, W)
    patch = img[np.ix_(Y, X)].astype(np.float64) * bayer_mask(color, itly + r, itlx + r)
    q = mclt_direct(patch, sh, sv)
    if kernel is not None:
        q = bicomplex_mul(q, np.asarray(kernel, np.float64).reshape(4, 8, 8))
    return bicomplex_mul(q, phasor(sh, sv))


def smooth_scene(seed, n=24, fmax=0.12):
    """Band-limited gray scene f(x, y) = sum a cos(2 pi (u x + v y) + phi)."""
    rng = np.random.default_rng(seed)
    u, v = rng.uniform(-fmax, fmax, (2, n))
    a = rng.uniform(0.5, 1.0, n)
    ph = rng.uniform(0, 2 * np.pi, n)

    def f(x, y):
        x = np.asarray(x, np.float64)[..., None]
        y = np.asarray(y, np.float64)[..., None]
        return (a * np.cos(2 * np.pi * (u * x + v * y) + ph)).sum(-1)
    return f


def render_cameras(geom, disparity, seed=0, fmax=0.3, ports=((-0.5, -0.5), (0.5, -0.5), (-0.5, 0.5), (0.5, 0.5))):
    """4 gray cameras of a band-limited scene; camera i sees the content
    displaced by disparity * port_i (disparity: scalar or [H][W] map)."""
    f = smooth_scene(seed, n=64, fmax=fmax)
    y, x = np.mgrid[0:geom.img_height, 0:geom.img_width].astype(np.float64)
    d = np.broadcast_to(np.asarray(disparity, np.float64), y.shape)
    return np.stack([f(x - d * px, y - d * py) for px, py in ports]).astype(np.float32)
