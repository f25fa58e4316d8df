"""CPU restatements of the learned nest surrogate's network (the network of
csrc/surrogate.hip, described in include/cliffs_surrogate.h), the checkers
of its GPU forward and backward kernels.  Test infrastructure only: nothing
in the product path imports this module.

* `loss_grad_bf16`: numpy with the GPU's BF16 rounding points (every GEMM
  operand rounded to BF16, FP32 accumulation) and backpropagation written
  out layer by layer -- convolution windows, the time pool, col2im of the
  window gradients, the two heads meeting in the trunk -- a tight check of
  the kernels' indexing, layouts and ReLU masks;
* `loss_grad_autograd`: torch FP32 autograd of the same network without
  any rounding (convolutions as torch's conv1d) -- an independent check of
  the backpropagation itself (looser, as BF16 operands differ from FP32 by
  ~0.4 % each).

The network: x (B, 512 samples, 64 gauges) -> four convolutions along time
(E1 kernel 4 stride 4; E2 kernel 5 stride 1; E3, E4 kernel 5 stride 2; zero
padding 2 for kernel 5) with ReLU -> maximum and mean over the 32 remaining
time rows -> D1, D2 (512 -> 512, ReLU) -> run-up head HR (512 -> 1) and map
head M1 (512 -> 4096, ReLU) -> six 2 x 2 stride-2 transposed convolutions
(ReLU but the last) to a 512 x 512 map in quadtree order.
"""
import numpy as np

# (kernel, stride, padding, channels in, out)
CONVS = [(4, 4, 0, 64, 128), (5, 1, 2, 128, 128), (5, 2, 2, 128, 256), (5, 2, 2, 256, 256)]
# (rows per sample, in, out, ReLU) of D1, D2, HR, M1, T1..T6
DENSE = [(1, 512, 512, 1), (1, 512, 512, 1), (1, 512, 1, 0), (1, 512, 4096, 1), (64, 64, 128, 1),
         (256, 32, 64, 1), (1024, 16, 32, 1), (4096, 8, 32, 1), (16384, 8, 16, 1), (65536, 4, 4, 0)]
D1, D2, HR, M1 = 0, 1, 2, 3
# (in, out) of every layer's weight matrix, in parameter order
SHAPES = [(k * ci, co) for k, _, _, ci, co in CONVS] + [(i, o) for _, i, o, _ in DENSE]


def bf16(x):
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    r = (u + (((u >> 16) & 1) + 0x7FFF).astype(np.uint32)) & np.uint32(0xFFFF0000)
    return r.view(np.float32)


def unpack(p):
    out, o = [], 0
    for i, n in SHAPES:
        W = p[o:o + i * n].reshape(i, n)
        o += i * n
        b = p[o:o + n]
        o += n
        out.append((W, b))
    return out


def _mm(a, b):
    """FP32 result of a product accumulated in FP64."""
    return (a.astype(np.float64) @ b.astype(np.float64)).astype(np.float32)


def _windows(h, k, s, pad):
    """(B, T, C) -> zero-padded windows (B, T_out, k*C), rows time-major."""
    B, T, C = h.shape
    hp = np.zeros((B, T + 2 * pad, C), dtype=np.float32)
    hp[:, pad:pad + T] = h
    n = (T + 2 * pad - k) // s + 1
    idx = np.arange(n)[:, None] * s + np.arange(k)[None, :]
    return hp[:, idx].reshape(B, n, k * C), hp.shape[1]


def loss_grad_bf16(p, x, y, yr, w_map, w_runup, rounding=True, plant=None):
    """rounding=False: the same computation without BF16 rounding, which
    must equal FP32 autograd up to summation order.  plant: a deliberate
    mistake for the tests that show the autograd check catches it --
    "pool_row" (the maximum's gradient one time row late) or "col2im"
    (window gradients summed one input row off)."""
    rb = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float32))
    p = np.asarray(p, dtype=np.float32)
    Wb = unpack(p)
    B = x.shape[0]
    # encoder
    h = rb(x.reshape(B, 512, 64))
    enc_in, wins = [], []
    for l, (k, s, pad, ci, co) in enumerate(CONVS):
        win, _ = _windows(h, k, s, pad)
        W, b = Wb[l]
        z = _mm(win.reshape(-1, k * ci), rb(W)) + b
        enc_in.append(h)
        wins.append(win)
        h = rb(np.maximum(z, 0.0)).reshape(B, -1, co)
    h4 = h                                              # (B, 32, 256)
    arg = h4.argmax(axis=1)
    feat = rb(np.concatenate([h4.max(axis=1), h4.sum(axis=1, dtype=np.float32) / np.float32(32)], axis=1))
    # trunk and heads
    acts = {}
    d = feat
    for l in (D1, D2):
        W, b = Wb[4 + l]
        d = rb(np.maximum(_mm(d, rb(W)) + b, 0.0))
        acts[l] = d
    W, b = Wb[4 + HR]
    r = (_mm(acts[D2], rb(W)) + b)[:, 0]
    a = acts[D2]
    dec = []
    for l in range(M1, len(DENSE)):
        rows, i, n, relu = DENSE[l]
        W, b = Wb[4 + l]
        dec.append(a.reshape(B * rows, i))
        z = _mm(a.reshape(B * rows, i), rb(W)) + b
        a = rb(np.maximum(z, 0.0)) if relu else z
    pred = a.reshape(B, -1)
    dm = pred - y
    dr = r - yr
    loss = float(w_map * (dm.astype(np.float64) ** 2).mean() + w_runup * (dr.astype(np.float64) ** 2).mean())
    grads = [None] * len(SHAPES)

    def dense_back(j, X, dz, mask):
        grads[j] = (_mm(X.T, rb(dz)), dz.sum(axis=0, dtype=np.float64).astype(np.float32))
        W, _ = Wb[j]
        dx = _mm(rb(dz), rb(W).T)
        return np.where(mask > 0, dx, 0.0).astype(np.float32) if mask is not None else dx

    # decoder
    dz = (2.0 * w_map * dm / dm.size).astype(np.float32).reshape(-1, 4)
    for l in range(len(DENSE) - 1, M1 - 1, -1):
        X = dec[l - M1]
        dz = dense_back(4 + l, X, dz.reshape(X.shape[0], -1), X)
    dH = dz.reshape(B, 512)
    # run-up head, added to the trunk gradient
    drz = (2.0 * w_runup * dr / B).astype(np.float32)[:, None]
    dH = (dense_back(4 + HR, acts[D2], drz, acts[D2]) + dH).astype(np.float32)
    dz = dense_back(4 + D2, acts[D1], dH, acts[D1])
    dfeat = dense_back(4 + D1, feat, dz, None)
    # pool
    dz = np.where(h4 > 0, dfeat[:, None, 256:] / np.float32(32), 0.0).astype(np.float32)
    bi, ci = np.meshgrid(np.arange(B), np.arange(256), indexing="ij")
    ra = np.minimum(arg + 1, h4.shape[1] - 1) if plant == "pool_row" else arg
    dz[bi, ra, ci] += np.where(h4[bi, ra, ci] > 0, dfeat[:, :256], 0.0).astype(np.float32)
    # encoder
    for l in range(3, -1, -1):
        k, s, pad, ci_, co = CONVS[l]
        win = wins[l].reshape(-1, k * ci_)
        dzm = dz.reshape(-1, co)
        grads[l] = (_mm(win.T, rb(dzm)), dzm.sum(axis=0, dtype=np.float64).astype(np.float32))
        if l == 0:
            break
        W, _ = Wb[l]
        dcols = _mm(rb(dzm), rb(W).T).reshape(B, -1, k, ci_)
        X = enc_in[l]
        T = X.shape[1]
        acc = np.zeros((B, T + 2 * pad + 1, ci_), dtype=np.float32)
        n = dcols.shape[1]
        off = 1 if plant == "col2im" else 0
        for dt in range(k):
            acc[:, dt + off:dt + off + s * n:s] += dcols[:, :, dt]
        dz = np.where(X > 0, acc[:, pad:pad + T], 0.0).astype(np.float32)
    out = []
    for dW, db in grads:
        out += [dW.reshape(-1), db]
    return loss, np.concatenate(out)


def loss_grad_autograd(p, x, y, yr, w_map, w_runup):
    import torch
    import torch.nn.functional as F
    B = x.shape[0]
    t = [torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, requires_grad=True)
         for Wb in unpack(np.asarray(p, dtype=np.float32)) for a in Wb]
    h = torch.tensor(np.ascontiguousarray(x), dtype=torch.float32).reshape(B, 512, 64).transpose(1, 2)
    for l, (k, s, pad, ci, co) in enumerate(CONVS):
        W = t[2 * l].reshape(k, ci, co).permute(2, 1, 0)            # (co, ci, k)
        h = torch.relu(F.conv1d(h, W, t[2 * l + 1], stride=s, padding=pad))
    feat = torch.cat([h.max(dim=2)[0], h.mean(dim=2)], dim=1)
    j = 8
    d = feat
    for _ in (D1, D2):
        d = torch.relu(d @ t[j] + t[j + 1])
        j += 2
    r = (d @ t[j] + t[j + 1])[:, 0]
    j += 2
    a = d
    for l in range(M1, len(DENSE)):
        rows, i, n, relu = DENSE[l]
        z = a.reshape(B * rows, i) @ t[j] + t[j + 1]
        j += 2
        a = torch.relu(z) if relu else z
    pred = a.reshape(B, -1)
    loss = w_map * ((pred - torch.tensor(y)) ** 2).mean() + w_runup * ((r - torch.tensor(yr)) ** 2).mean()
    loss.backward()
    return float(loss.detach()), np.concatenate([a.grad.numpy().reshape(-1) for a in t])
