"""FP64 CPU restatement of the learned-start network, for gradient checks.

Same parameter layout and arithmetic as accelerint_amd/csrc/mlp.hip: every
matrix product rounds both operands to bfloat16 (round to nearest even) and
accumulates exactly (FP64 here, FP32 on the GPU); biases, GELU (tanh form),
the loss and its derivative stay unrounded.  ``bf16=False`` gives the plain
FP64 network, whose gradient the tests compare with central differences.
"""
import numpy as np

H = 128
K0, K1 = 0.7978845608028654, 0.044715


def pad16(n):
    return (n + 15) // 16 * 16


def bf16(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(K0 * (x + K1 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(K0 * (x + K1 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * K0 * (1.0 + 3.0 * K1 * x * x)


class Net:
    def __init__(self, n_states):
        self.n = n_states
        self.inn, self.out = n_states + 2, 3 * n_states
        self.inp, self.outp = pad16(self.inn), pad16(self.out)
        self.shapes = [(self.inp, H), (H,), (H, H), (H,), (H, self.outp), (self.outp,)]
        self.count = sum(int(np.prod(s)) for s in self.shapes)

    def unpack(self, p):
        out, o = [], 0
        for s in self.shapes:
            k = int(np.prod(s))
            out.append(np.asarray(p[o:o + k], dtype=np.float64).reshape(s))
            o += k
        return out

    def pack(self, parts):
        return np.concatenate([np.asarray(a).reshape(-1) for a in parts])

    def forward(self, p, x, bf=True):
        W1, b1, W2, b2, W3, b3 = self.unpack(p)
        r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
        B = x.shape[0]
        X = np.zeros((B, self.inp))
        X[:, :self.inn] = x
        Z1 = r(X) @ r(W1) + b1
        A1 = gelu(Z1)
        Z2 = r(A1) @ r(W2) + b2
        A2 = gelu(Z2)
        Y = r(A2) @ r(W3) + b3
        Y[:, self.out:] = 0.0
        return X, Z1, A1, Z2, A2, Y

    def predict(self, p, x, bf=True):
        return self.forward(p, x, bf)[-1][:, :self.out]

    def loss(self, p, x, t, bf=True):
        Y = self.forward(p, x, bf)[-1][:, :self.out]
        return float(np.mean((Y - t) ** 2))

    def loss_grad(self, p, x, t, bf=True):
        W1, b1, W2, b2, W3, b3 = self.unpack(p)
        r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
        X, Z1, A1, Z2, A2, Y = self.forward(p, x, bf)
        B = x.shape[0]
        dY = np.zeros_like(Y)
        e = Y[:, :self.out] - t
        dY[:, :self.out] = 2.0 * e / (B * self.out)
        gW3 = r(A2).T @ r(dY)
        gb3 = dY.sum(0)
        dZ2 = (r(dY) @ r(W3).T) * gelu_grad(Z2)
        gW2 = r(A1).T @ r(dZ2)
        gb2 = dZ2.sum(0)
        dZ1 = (r(dZ2) @ r(W2).T) * gelu_grad(Z1)
        gW1 = r(X).T @ r(dZ1)
        gb1 = dZ1.sum(0)
        return float(np.mean(e ** 2)), self.pack([gW1, gb1, gW2, gb2, gW3, gb3])
