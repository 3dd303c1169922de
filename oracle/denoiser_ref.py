"""Oracle of the learned denoiser: TEST INFRASTRUCTURE ONLY.

A numpy restatement of the U-Net the HIP kernels run (oidn-protoray_amd/csrc/denoise.hip):
input log(1 + colour), albedo, normal (9 channels, padded to 16); encoder convolutions 9->32,
32->48, 48->64, 64->96, 96->128 with 2x2 max pooling between levels; decoder levels that
concatenate the nearest 2x upsampling of the level below with the encoder skip and convolve
(224->96, 160->64, 112->48, 80->32); a final 32->3 convolution; all 3x3 with zero padding and
ReLU except the last; L1 loss of log(1 + colour).  Parameters use the HIP layout (per layer
weights [cout][cin][3][3] then bias [cout]).

With bf16=True every stored activation and activation gradient is rounded to BF16 (round to
nearest even) where the HIP path stores BF16, convolution products are summed in FP32, and the
weights the convolutions see are BF16; with bf16=False everything is float64 (the reference
for the central-difference check of this backward pass).
"""
import numpy as np

LAYERS = [(9, 32), (32, 48), (48, 64), (64, 96), (96, 128), (224, 96), (160, 64), (112, 48), (80, 32), (32, 3)]
ENC_C = [32, 48, 64, 96, 128]
DEC_C = [32, 48, 64, 96]


def param_count():
    return sum(co * ci * 9 + co for ci, co in LAYERS)


def bf16(x):
    """Round float32 to the nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(params):
    out, off = [], 0
    for ci, co in LAYERS:
        w = params[off:off + co * ci * 9].reshape(co, ci, 3, 3)
        off += co * ci * 9
        b = params[off:off + co]
        off += co
        out.append((w, b))
    return out


def _patches(x):
    """x (N, H, W, C) -> (N, H, W, 9, C), zero padded, tap = 3 * (dy + 1) + (dx + 1)."""
    N, H, W, C = x.shape
    xp = np.zeros((N, H + 2, W + 2, C), x.dtype)
    xp[:, 1:-1, 1:-1] = x
    return np.stack([xp[:, dy:dy + H, dx:dx + W] for dy in range(3) for dx in range(3)], axis=3)


class Net:
    def __init__(self, params, bf16_mode=True):
        self.q = bf16_mode
        self.dt = np.float32 if bf16_mode else np.float64
        self.layers = []
        for w, b in unpack(np.asarray(params, np.float64)):
            wm = w.transpose(2, 3, 1, 0).reshape(9 * w.shape[1], w.shape[0])  # (9 cin, cout), tap-major
            if self.q:
                wm = bf16(wm.astype(np.float32))
            self.layers.append((wm.astype(self.dt), np.asarray(b, self.dt)))

    def r(self, x):
        return bf16(x) if self.q else x

    def conv(self, i, x, relu=True):
        wm, b = self.layers[i]
        N, H, W, C = x.shape
        P = _patches(x).reshape(N * H * W, 9 * C)
        y = (P @ wm).reshape(N, H, W, -1) + b
        return np.maximum(y, 0) if relu else y

    def conv_dx(self, i, dy, shape):
        wm, _ = self.layers[i]
        N, H, W, C = shape
        g = (dy.reshape(-1, dy.shape[-1]) @ wm.T).reshape(N, H, W, 3, 3, C)
        dx = np.zeros((N, H + 2, W + 2, C), self.dt)
        for dyy in range(3):
            for dxx in range(3):
                dx[:, dyy:dyy + H, dxx:dxx + W] += g[:, :, :, dyy, dxx]
        return dx[:, 1:-1, 1:-1]

    def conv_dw(self, x, dy):
        N, H, W, C = x.shape
        P = _patches(x).reshape(N * H * W, 9 * C)
        dw = P.T @ dy.reshape(-1, dy.shape[-1])                     # (9 cin, cout)
        db = dy.reshape(-1, dy.shape[-1]).sum(0)
        cout = dy.shape[-1]
        return dw.reshape(3, 3, C, cout).transpose(3, 2, 0, 1), db  # [cout][cin][3][3]

    @staticmethod
    def pool(x):
        N, H, W, C = x.shape
        return x.reshape(N, H // 2, 2, W // 2, 2, C).max(axis=(2, 4))

    @staticmethod
    def pool_bwd(x, dp):
        """Gradient to the first maximum of each 2x2 window in order (0,0), (0,1), (1,0), (1,1)."""
        N, H, W, C = x.shape
        win = x.reshape(N, H // 2, 2, W // 2, 2, C).transpose(0, 1, 3, 5, 2, 4).reshape(N, H // 2, W // 2, C, 4)
        first = win.argmax(-1)
        g = np.zeros_like(win)
        np.put_along_axis(g, first[..., None], dp[..., None], axis=-1)
        return g.reshape(N, H // 2, W // 2, C, 2, 2).transpose(0, 1, 4, 2, 5, 3).reshape(N, H, W, C)

    @staticmethod
    def up(x):
        return x.repeat(2, axis=1).repeat(2, axis=2)

    def prep(self, in9):
        x = np.zeros(in9.shape[:3] + (9,), self.dt)
        x[..., :3] = np.log1p(np.maximum(in9[..., :3], 0))
        x[..., 3:] = in9[..., 3:9]
        return self.r(x.astype(self.dt))

    def forward(self, in9):
        s = {"x": self.prep(in9)}
        h = s["x"]
        for l in range(5):
            e = self.r(self.conv(l, h))
            s["e%d" % l] = e
            if l < 4:
                h = self.pool(e)
                s["p%d" % l] = h
        low = s["e4"]
        for l in range(3, -1, -1):
            cat = np.concatenate([self.up(low), s["e%d" % l]], axis=-1)
            s["c%d" % l] = cat
            d = self.r(self.conv(8 - l, cat))
            s["d%d" % l] = d
            low = d
        s["pred"] = self.conv(9, s["d0"], relu=False)
        return s

    def loss_grad(self, in9, target):
        """(L1 loss of log(1 + colour), parameter gradient in the HIP layout, forward state)."""
        s = self.forward(in9)
        t = np.log1p(np.maximum(np.asarray(target, self.dt), 0))
        diff = s["pred"] - t
        n = diff.size
        loss = np.abs(diff).sum() / n
        grads = [None] * 10
        dout = self.r(np.sign(diff) / n)
        grads[9] = self.conv_dw(s["d0"], dout)
        dd = self.r(self.conv_dx(9, dout, s["d0"].shape) * (s["d0"] > 0))
        dcat = {}
        for l in range(4):
            layer = 8 - l
            grads[layer] = self.conv_dw(s["c%d" % l], dd)
            dc = self.r(self.conv_dx(layer, dd, s["c%d" % l].shape) * (s["c%d" % l] > 0))
            dcat[l] = dc
            lowC = 128 if l == 3 else DEC_C[l + 1]
            up = dc[..., :lowC]
            N, H, W, _ = up.shape
            dd = self.r(up.reshape(N, H // 2, 2, W // 2, 2, lowC).sum(axis=(2, 4)))
        de = dd  # gradient of e4
        for l in range(4, -1, -1):
            inp = s["x"] if l == 0 else s["p%d" % (l - 1)]
            grads[l] = self.conv_dw(inp, de)
            if l > 0:
                dp = self.r(self.conv_dx(l, de, inp.shape))
                e = s["e%d" % (l - 1)]
                lowC = 128 if l - 1 == 3 else DEC_C[l]
                g = self.pool_bwd(e, dp) + dcat[l - 1][..., lowC:]
                de = self.r(g * (e > 0))
        flat = np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in grads])
        return loss, flat, s


def he_init(seed=12345):
    """The HIP denoiser's initial parameters: uniform He bounds from the same 32-bit LCG."""
    out = []
    s = seed
    for ci, co in LAYERS:
        bound = np.float32(np.sqrt(np.float32(6.0) / np.float32(9 * ci)))
        n = co * ci * 9
        vals = np.empty(n, np.float32)
        for k in range(n):
            s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
            u = (np.float32(s >> 8) + np.float32(0.5)) * np.float32(1.0 / 16777216.0)
            vals[k] = (np.float32(2.0) * u - np.float32(1.0)) * bound
        out += [vals, np.zeros(co, np.float32)]
    return np.concatenate(out)
