"""FP64 CPU restatement of the cost model's MLP (the checker of the GPU kernels in
periodicdmk_amd/csrc/hip/costmodel.hip): 7 -> 64 -> 64 -> 64 -> 2, GELU(z) = z Phi(z),
parameters laid out layer by layer as W_l ([out][in], row-major) then b_l; loss = masked
mean squared error sum M (Y - T)^2 / sum M.  Gradients by backpropagation in FP64."""
import math

import numpy as np

DIMS = (7, 64, 64, 64, 2)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = [], 0
    for i in range(4):
        nout, nin = DIMS[i + 1], DIMS[i]
        W = p[o:o + nout * nin].reshape(nout, nin)
        o += nout * nin
        b = p[o:o + nout]
        o += nout
        out.append((W, b))
    assert o == p.size
    return out


_erf = np.vectorize(math.erf)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(p, X):
    h = np.asarray(X, dtype=np.float64)
    layers = unpack(p)
    zs, hs = [], [h]
    for i, (W, b) in enumerate(layers):
        z = h @ W.T + b
        if i < 3:
            zs.append(z)
            h = gelu(z)
            hs.append(h)
        else:
            h = z
    return h, zs, hs


def loss_grad(p, X, T, M):
    Y, zs, hs = forward(p, X)
    M = np.asarray(M, dtype=np.float64)
    cnt = M.sum()
    d = M * (Y - np.asarray(T, dtype=np.float64))
    loss = float(np.sum(d * (Y - T)) / cnt)
    dz = 2.0 * d / cnt
    layers = unpack(p)
    grads = [None] * 4
    for i in range(3, -1, -1):
        W, _ = layers[i]
        grads[i] = (dz.T @ hs[i], dz.sum(0))
        if i > 0:
            dz = (dz @ W) * gelu_grad(zs[i - 1])
    g = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, g
