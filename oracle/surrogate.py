"""CPU oracle of the learned far-cascade radiance surrogate.

TEST INFRASTRUCTURE ONLY: tests, the smoke check and the benchmark's CPU
baseline leg may import this module, as the checker.  The product path is the
HIP library (bevy_penumbra_amd/csrc/lux_surrogate.hip) and never calls this.

The reference has no learned component (README.md:384-385 defers "online
neural radiance caching"), so nothing here follows a reference file.  The
network is pinned by mathematics instead (tests/test_oracle_surrogate.py):
the real spherical harmonics against scipy and by quadrature orthonormality,
the hash/dense grid index against its closed form, trilinear interpolation
against linear fields it must reproduce, the MLP against an explicit f64
GEMM, and every gradient against central finite differences.

Two evaluation modes share one code path:
  * ``exact=True`` reproduces the GPU's arithmetic: encoder in IEEE f32 with
    the GPU's operation order (bit-exact), BF16 rounding of X, W, H1, H2 and
    the back-propagated activations at the points where the GPU rounds them;
    products are then summed in f64 (the GPU sums in f32 on MFMA, so results
    agree to the f32 accumulation error).
  * ``exact=False`` is plain f64 throughout (the finite-difference target).

Network: input 128 = 16 hash levels x 2 features | 64 real SH coefficients
(bands 0-7) of the direction | normal (3) | constant 1 | zero padding;
MLP 128 -> 64 -> 64 -> 16, ReLU, no biases; outputs 0-2 radiance, output 3
the log of the predicted squared error.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

LEVELS = 16
FEATURES = 2
LOG2_TABLE = 19
TABLE = 1 << LOG2_TABLE
IN, HID, OUT = 128, 64, 16
SH_BANDS = 8
SH_COEFFS = SH_BANDS * SH_BANDS
PRIMES = (1, 2654435761, 805459861)
W_SIZES = (IN * HID, HID * HID, HID * OUT)
MLP_PARAMS = sum(W_SIZES)
F32 = np.float32


# --------------------------------------------------------------------- BF16
def bf16_round(x) -> np.ndarray:
    """Round f32 values to the nearest BF16 (ties to even), returned as f32."""
    u = np.ascontiguousarray(x, dtype=F32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(F32)


def bf16_bits(x) -> np.ndarray:
    """BF16 bit patterns (uint16) of f32 values already representable in BF16."""
    return (np.ascontiguousarray(x, dtype=F32).view(np.uint32) >> 16).astype(np.uint16)


# --------------------------------------------------------------------- grid
def level_resolutions(base: float = 16.0, growth: float = 1.5):
    """res_l = floor(base * growth^l) (f64), dense when (res+1)^3 <= 2^19."""
    b, g = float(F32(base)), float(F32(growth))
    res = [int(math.floor(b * math.pow(g, float(l)))) for l in range(LEVELS)]
    dense = [(r + 1) ** 3 <= TABLE for r in res]
    return res, dense


def grid_index(x, y, z, res: int, dense: bool) -> np.ndarray:
    """Dense: x + (res+1)(y + (res+1) z).  Hashed: (x*1 ^ y*p1 ^ z*p2) mod 2^19
    with u32 wrap-around (only the low 19 bits survive, so u64 products give
    the same bits)."""
    x = np.asarray(x, np.uint64)
    y = np.asarray(y, np.uint64)
    z = np.asarray(z, np.uint64)
    if dense:
        r1 = np.uint64(res + 1)
        return (x + r1 * (y + r1 * z)).astype(np.int64)
    h = (x * np.uint64(PRIMES[0])) ^ (y * np.uint64(PRIMES[1])) ^ (z * np.uint64(PRIMES[2]))
    return (h & np.uint64(TABLE - 1)).astype(np.int64)


def sh_constants() -> np.ndarray:
    """Real-SH normalisation sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!), x sqrt(2) for m != 0,
    rounded to f32 as the library uploads them."""
    c = np.zeros(SH_COEFFS, np.float64)
    for l in range(SH_BANDS):
        for m in range(l + 1):
            k = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))
            if m == 0:
                c[l * l + l] = k
            else:
                c[l * l + l + m] = math.sqrt(2.0) * k
                c[l * l + l - m] = math.sqrt(2.0) * k
    return c.astype(F32)


def sh_encode(d: np.ndarray, exact: bool = True) -> np.ndarray:
    """Real SH of unit directions d (n, 3), index l*l + l + m, polar axis z,
    no Condon-Shortley phase.  exact=True follows the GPU's f32 recurrence:
    (x + i y)^m by complex products, P_m^m = (2m-1)!!, P_{m+1}^m = z (2m+1) P_m^m,
    (l-m) P_l^m = (2l-1) z P_{l-1}^m - (l+m-1) P_{l-2}^m."""
    T = F32 if exact else np.float64
    d = np.asarray(d)
    x, y, z = (d[:, i].astype(T) for i in range(3))
    c = sh_constants().astype(T) if exact else sh_constants_f64()
    n = d.shape[0]
    out = np.zeros((n, SH_COEFFS), T)
    re = [np.ones(n, T)]
    im = [np.zeros(n, T)]
    for m in range(1, SH_BANDS):
        re.append(re[m - 1] * x - im[m - 1] * y)
        im.append(re[m - 1] * y + im[m - 1] * x)
    for m in range(SH_BANDS):
        pmm = T(1.0)
        for k in range(1, m + 1):
            pmm = T(pmm * T(2 * k - 1))
        pmm = np.full(n, pmm, T)
        p_prev = np.zeros(n, T)
        p_cur = pmm
        for l in range(m, SH_BANDS):
            if l == m:
                p = pmm
            elif l == m + 1:
                p = (z * T(2 * m + 1)) * pmm
            else:
                p = ((T(2 * l - 1) * z) * p_cur - T(l + m - 1) * p_prev) / T(l - m)
            if l > m:
                p_prev, p_cur = p_cur, p
            base = l * l + l
            if m == 0:
                out[:, base] = c[base] * p
            else:
                out[:, base + m] = (c[base + m] * re[m]) * p
                out[:, base - m] = (c[base - m] * im[m]) * p
    return out


def sh_constants_f64() -> np.ndarray:
    c = np.zeros(SH_COEFFS, np.float64)
    for l in range(SH_BANDS):
        for m in range(l + 1):
            k = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))
            if m == 0:
                c[l * l + l] = k
            else:
                c[l * l + l + m] = c[l * l + l - m] = math.sqrt(2.0) * k
    return c


@dataclass
class Encoder:
    box_min: np.ndarray
    box_max: np.ndarray
    base: float = 16.0
    growth: float = 1.5

    def __post_init__(self):
        self.box_min = np.asarray(self.box_min, F32)
        self.box_max = np.asarray(self.box_max, F32)
        ext = (self.box_max - self.box_min).astype(F32)
        with np.errstate(divide="ignore"):
            self.box_inv = np.where(ext > 0, F32(1.0) / np.where(ext > 0, ext, F32(1.0)), F32(0.0)).astype(F32)
        self.res, self.dense = level_resolutions(self.base, self.growth)

    def unit_position(self, pos: np.ndarray) -> np.ndarray:
        """Clamp((pos - min) * inv, 0, 1 - 2^-24) in f32; NaN maps to 0."""
        with np.errstate(invalid="ignore"):
            q = (np.asarray(pos, F32) - self.box_min) * self.box_inv
            return np.where(q > F32(0.99999994), F32(0.99999994), np.where(q >= 0, q, F32(0.0))).astype(F32)

    def corners(self, pos: np.ndarray, level: int):
        """Per corner c = bx + 2 by + 4 bz: table index (n,) and f32 weight (n,)."""
        p = self.unit_position(pos)
        res = F32(self.res[level])
        g = p * res
        fl = np.floor(g)
        cell = fl.astype(np.int64)
        t = (g - fl).astype(F32)
        out = []
        for c in range(8):
            b = (c & 1, (c >> 1) & 1, c >> 2)
            w = [t[:, a] if b[a] else F32(1.0) - t[:, a] for a in range(3)]
            wt = (w[0] * w[1]) * w[2]
            idx = grid_index(cell[:, 0] + b[0], cell[:, 1] + b[1], cell[:, 2] + b[2],
                             self.res[level], self.dense[level])
            out.append((idx, wt.astype(F32)))
        return out

    def hash_features(self, pos: np.ndarray, table: np.ndarray, exact: bool = True) -> np.ndarray:
        """(n, 32) trilinear features.  exact: table as FP16, f32 accumulation
        in corner order (the GPU's arithmetic); else f64 table and sums."""
        n = pos.shape[0]
        T = F32 if exact else np.float64
        tab = table.astype(np.float16).astype(F32) if exact else table.astype(np.float64)
        out = np.zeros((n, 2 * LEVELS), T)
        for l in range(LEVELS):
            f0 = np.zeros(n, T)
            f1 = np.zeros(n, T)
            for idx, w in self.corners(pos, l):
                w = w.astype(T)
                f0 = f0 + w * tab[l, idx, 0]
                f1 = f1 + w * tab[l, idx, 1]
            out[:, 2 * l] = f0
            out[:, 2 * l + 1] = f1
        return out

    def encode(self, queries: np.ndarray, table: np.ndarray, exact: bool = True) -> np.ndarray:
        """queries (n, 9) = position, direction, normal -> X (n, 128)."""
        q = np.asarray(queries)
        n = q.shape[0]
        X = np.zeros((n, IN), np.float64)
        X[:, 0:32] = self.hash_features(q[:, 0:3], table, exact)
        X[:, 32:96] = sh_encode(q[:, 3:6], exact)
        X[:, 96:99] = q[:, 6:9]
        X[:, 99] = 1.0
        if exact:
            X = bf16_round(X.astype(F32)).astype(np.float64)
        return X

    def scatter(self, pos: np.ndarray, dXh: np.ndarray) -> np.ndarray:
        """Table gradient: trilinear scatter of dX[:, 0:32] (f64)."""
        g = np.zeros((LEVELS, TABLE, FEATURES), np.float64)
        for l in range(LEVELS):
            for idx, w in self.corners(pos, l):
                w = w.astype(np.float64)
                np.add.at(g[l, :, 0], idx, w * dXh[:, 2 * l])
                np.add.at(g[l, :, 1], idx, w * dXh[:, 2 * l + 1])
        return g


# --------------------------------------------------------------------- MLP
def split_mlp(mlp: np.ndarray):
    """Concatenated W1 (128x64) | W2 (64x64) | W3 (64x16), each [in][out] row-major."""
    a, b, _ = W_SIZES
    m = np.asarray(mlp, np.float64)
    return m[:a].reshape(IN, HID), m[a:a + b].reshape(HID, HID), m[a + b:].reshape(HID, OUT)


def _rb(x, exact):
    return bf16_round(x.astype(F32)).astype(np.float64) if exact else x


def forward(X: np.ndarray, mlp: np.ndarray, exact: bool = True):
    W1, W2, W3 = (_rb(w, exact) for w in split_mlp(mlp))
    H1 = _rb(np.maximum(X @ W1, 0.0), exact)
    H2 = _rb(np.maximum(H1 @ W2, 0.0), exact)
    Y = H2 @ W3
    return H1, H2, Y


def loss_grad_y(Y: np.ndarray, target: np.ndarray, err_weight: float, fixed=None):
    """Relative L2 on RGB with the weight 1/(y^2 + 0.01) held constant, plus
    err_weight * (y3 - log(mse + 1e-4))^2 with its target held constant; both
    averaged over n samples (RGB also over 3 channels).  Samples with a
    non-finite target contribute nothing.  Returns (loss, dY, fixed) where
    fixed = (w, target_e) are the held constants."""
    n = Y.shape[0]
    ok = np.all(np.isfinite(target), axis=1)
    t = np.where(ok[:, None], target, 0.0)
    y = Y[:, 0:3]
    diff = y - t
    if fixed is None:
        w = 1.0 / (y * y + 0.01)
        se = np.sum(diff * diff, axis=1)
        te = np.log(se / 3.0 + 1e-4)
        fixed = (w, te)
    w, te = fixed
    inv = 1.0 / (3.0 * n)
    de = Y[:, 3] - te
    per = np.sum(diff * diff * w, axis=1) * inv + err_weight * de * de / n
    loss = float(np.sum(np.where(ok, per, 0.0)))
    dY = np.zeros_like(Y)
    dY[:, 0:3] = 2.0 * diff * w * inv
    dY[:, 3] = 2.0 * err_weight * de / n
    dY[~ok] = 0.0
    return loss, dY, fixed


def backward(X, H1, H2, dY, mlp, exact: bool = True):
    """Returns (dW1, dW2, dW3, dXh) with dXh = dL/dX[:, 0:32].  The error head
    (output 3) trains its column of W3 only: its gradient is stopped before
    the shared layers."""
    W1, W2, W3 = (_rb(w, exact) for w in split_mlp(mlp))
    dYb = _rb(dY, exact)
    trunk = dYb.copy()
    trunk[:, 3] = 0.0
    dH2 = _rb((trunk @ W3.T) * (H2 > 0), exact)
    dH1 = _rb((dH2 @ W2.T) * (H1 > 0), exact)
    dXh = (dH1 @ W1.T)[:, 0:32]
    return X.T @ dH1, H1.T @ dH2, H2.T @ dYb, dXh


def train_gradients(enc: Encoder, table, mlp, samples, err_weight: float, exact: bool = True,
                    fixed=None):
    """One training step's loss and gradients over samples (n, 12) =
    position, direction, normal, target RGB.  Returns (loss, mlp_grad, table_grad, fixed)."""
    s = np.asarray(samples, np.float64)
    X = enc.encode(s[:, 0:9].astype(F32), table, exact)
    H1, H2, Y = forward(X, mlp, exact)
    loss, dY, fixed = loss_grad_y(Y, s[:, 9:12], err_weight, fixed)
    dW1, dW2, dW3, dXh = backward(X, H1, H2, dY, mlp, exact)
    mg = np.concatenate([dW1.ravel(), dW2.ravel(), dW3.ravel()])
    tg = enc.scatter(s[:, 0:3].astype(F32), dXh)
    return loss, mg, tg, fixed


def loss_only(enc: Encoder, table, mlp, samples, err_weight, fixed):
    s = np.asarray(samples, np.float64)
    X = enc.encode(s[:, 0:9].astype(F32), table, exact=False)
    _, _, Y = forward(X, mlp, exact=False)
    return loss_grad_y(Y, s[:, 9:12], err_weight, fixed)[0]


def adam(param, grad, m, v, step: int, lr: float, b1=0.9, b2=0.99, eps=1e-15):
    """Adam with bias correction; returns (param, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return param - lr * mh / (np.sqrt(vh) + eps), m, v
