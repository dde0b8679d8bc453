"""A plain PyTorch FP32 restatement of the learned radiance cache's forward
pass (helpers for the GPU tests, not tests): the hash-grid encoding
(Mueller, Evans, Schied and Keller 2022: trilinear blend of the 8 corner
entries per level, dense indexing while (res + 1)^3 fits the table, the
Teschner et al. primes above), the network input layout, and the 5 x 64
ReLU MLP with an RGB output.  Where the kernels keep BF16 copies (the table,
the weights, the network input and every hidden activation) the restatement
rounds to BF16 on the way forward and passes gradients straight through, so
autograd gives the gradients the kernels compute up to their own BF16
rounding of the deltas."""
import numpy as np
import torch

WIDTH, HIDDEN, OUT_PAD = 64, 5, 16


def bf16(x):
    return x.to(torch.bfloat16).to(torch.float32)


def ste_bf16(x):
    """BF16 forward, identity backward."""
    return x + (bf16(x) - x).detach()


def random_inputs(n, lo, hi, rng, margin=0.1):
    """rtx_nrc_input rows (n, 16): positions a little past the box (clamped by
    the encoding), unit directions and normals, roughness and albedo in [0, 1]."""
    lo, hi = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    ext = hi - lo
    x = np.zeros((n, 16), np.float32)
    x[:, 0:3] = rng.uniform(lo - margin * ext, hi + margin * ext, (n, 3))
    for c0 in (4, 8):
        v = rng.normal(size=(n, 3))
        x[:, c0:c0 + 3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    x[:, 3] = rng.uniform(0, 1, n)
    x[:, 12:15] = rng.uniform(0, 1, (n, 3))
    return x


class Reference:
    def __init__(self, info):
        self.L = int(info.levels)
        self.T = int(info.tableSize)
        self.res = [np.float32(info.resolution[l]) for l in range(self.L)]
        self.dense = [bool(info.dense[l]) for l in range(self.L)]
        self.lo = torch.tensor(list(info.lo), dtype=torch.float32)
        self.inv = torch.tensor(list(info.inv), dtype=torch.float32)

    def encode(self, inputs, table):
        """inputs (n, 16) float32 tensor; table (L, T, 2) float32 tensor."""
        tb = ste_bf16(table)
        n = inputs.shape[0]
        p = ((inputs[:, 0:3] - self.lo) * self.inv).clamp(0.0, 1.0)
        feats = []
        for l in range(self.L):
            r = torch.tensor(self.res[l])
            x = p * r
            f = torch.clamp(torch.floor(x), max=float(self.res[l]) - 1.0).clamp(min=0.0)
            t = x - f
            ci = f.to(torch.int64)
            f0 = torch.zeros(n)
            f1 = torch.zeros(n)
            for c in range(8):
                bits = [(c >> a) & 1 for a in range(3)]
                w = torch.ones(n)
                for a in range(3):
                    w = w * (t[:, a] if bits[a] else 1.0 - t[:, a])
                ix, iy, iz = (ci[:, a] + bits[a] for a in range(3))
                if self.dense[l]:
                    n1 = int(self.res[l]) + 1
                    idx = ix + n1 * (iy + n1 * iz)
                else:
                    idx = (ix ^ ((iy * 2654435761) & 0xFFFFFFFF) ^ ((iz * 805459861) & 0xFFFFFFFF)) & (self.T - 1)
                v = tb[l, idx]
                f0 = f0 + w * v[:, 0]
                f1 = f1 + w * v[:, 1]
            feats += [f0, f1]
        enc = torch.stack(feats, 1)
        pad = torch.zeros(n, 32 - 2 * self.L)
        rest = torch.cat([inputs[:, 4:7], inputs[:, 8:11], inputs[:, 3:4], inputs[:, 12:15], torch.ones(n, 1),
                          torch.zeros(n, 21)], 1)
        return ste_bf16(torch.cat([enc, pad, rest], 1))

    @staticmethod
    def weights(mlp):
        """The MLP vector (out-major per layer) as 5 (64, 64) matrices and the (16, 64) output matrix."""
        ws = [mlp[l * WIDTH * WIDTH:(l + 1) * WIDTH * WIDTH].reshape(WIDTH, WIDTH) for l in range(HIDDEN)]
        ws.append(mlp[HIDDEN * WIDTH * WIDTH:].reshape(OUT_PAD, WIDTH))
        return ws

    def forward(self, inputs, table, mlp):
        h = self.encode(inputs, table)
        ws = self.weights(mlp)
        for l in range(HIDDEN):
            h = ste_bf16(torch.relu(h @ ste_bf16(ws[l]).T))
        return h @ ste_bf16(ws[HIDDEN][:3]).T

    @staticmethod
    def loss(pred, target):
        """Relative L2 (Mueller et al. 2021), mean over rows and channels."""
        return ((pred - target) ** 2 / (pred.detach() ** 2 + 0.01)).mean()
