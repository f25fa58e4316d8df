"""CPU restatements of the learned nest surrogate's network (the network of
csrc/surrogate.hip, described in include/cliffs_surrogate.h), the checkers
of its GPU forward and backward kernels.  Test infrastructure only: nothing
in the product path imports this module.

* `loss_grad_bf16`: numpy FP32 with the GPU's BF16 rounding points (every
  GEMM operand rounded to BF16, FP32 accumulation) and backpropagation
  written out layer by layer -- a tight check of the kernels' indexing,
  layouts and ReLU masks;
* `loss_grad_autograd`: torch FP32 autograd of the same network without any
  rounding -- an independent check of the backpropagation itself (looser,
  as BF16 operands differ from FP32 by ~0.4 % each).
"""
import numpy as np

# (rows per sample, in, out) of layers L1..L7
LAYERS = [(32, 512, 128), (8, 512, 256), (1, 2048, 512), (1, 512, 4096),
          (64, 64, 128), (256, 32, 64), (1024, 16, 4)]


def bf16(x):
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    r = (u + (((u >> 16) & 1) + 0x7FFF).astype(np.uint32)) & np.uint32(0xFFFF0000)
    return r.view(np.float32)


def unpack(p):
    out, o = [], 0
    for _, i, n in LAYERS:
        W = p[o:o + i * n].reshape(i, n)
        o += i * n
        b = p[o:o + n]
        o += n
        out.append((W, b))
    return out


def loss_grad_bf16(p, x, y, rounding=True):
    """rounding=False: the same computation without BF16 rounding, which
    must equal FP32 autograd up to summation order."""
    rb = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float32))
    p = np.asarray(p, dtype=np.float32)
    B = x.shape[0]
    Wb = unpack(p)
    acts = [rb(x.reshape(-1))]
    z = None
    for l, ((rows, i, n), (W, b)) in enumerate(zip(LAYERS, Wb)):
        a = acts[-1].reshape(B * rows, i)
        z = (a.astype(np.float64) @ rb(W).astype(np.float64)).astype(np.float32) + b
        if l < 6:
            acts.append(rb(np.maximum(z, 0.0)).reshape(-1))
    pred = z.reshape(B, -1)
    d = pred - y
    loss = float((d.astype(np.float64) ** 2).mean())
    dz = (2.0 * d / d.size).astype(np.float32).reshape(-1)
    grads = []
    for l in range(6, -1, -1):
        rows, i, n = LAYERS[l]
        W, _ = Wb[l]
        dzm = dz.reshape(B * rows, n)
        db = dzm.sum(axis=0, dtype=np.float64).astype(np.float32)
        X = acts[l].reshape(B * rows, i)
        dW = (X.T.astype(np.float64) @ rb(dzm).astype(np.float64)).astype(np.float32)
        grads.append((dW, db))
        if l > 0:
            dx = (rb(dzm).astype(np.float64) @ rb(W).T.astype(np.float64)).astype(np.float32)
            dz = np.where(acts[l].reshape(B * rows, i) > 0, dx, 0.0).astype(np.float32).reshape(-1)
    out = []
    for dW, db in reversed(grads):
        out += [dW.reshape(-1), db]
    return loss, np.concatenate(out)


def loss_grad_autograd(p, x, y):
    import torch
    B = x.shape[0]
    params = [torch.tensor(np.ascontiguousarray(t), dtype=torch.float32, requires_grad=True)
              for Wb in unpack(np.asarray(p, dtype=np.float32)) for t in Wb]
    a = torch.tensor(np.ascontiguousarray(x), dtype=torch.float32)
    for l, (rows, i, n) in enumerate(LAYERS):
        a = a.reshape(B * rows, i) @ params[2 * l] + params[2 * l + 1]
        if l < 6:
            a = torch.relu(a)
    loss = ((a.reshape(B, -1) - torch.tensor(y)) ** 2).mean()
    loss.backward()
    return float(loss.detach()), np.concatenate([t.grad.numpy().reshape(-1) for t in params])
