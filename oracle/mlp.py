"""CPU oracle of the device multilayer perceptron (test infrastructure only).

Numpy restatement of the network in eprem_amd/csrc/ep_mlp.hip: layers
Z = X W + b, ReLU on hidden layers, mean squared error.  Each product
rounds both operands to bf16 (round to nearest even) as the matrix cores
see them and accumulates in float64, so the device (fp32 accumulation)
agrees to accumulation rounding.  `exact=True` skips the bf16 rounding
(used to pin the gradient formulas by finite differences)."""
import numpy as np


def to_bf16(a):
    """Round float32 values to bf16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def unpack(p, sizes):
    out, off = [], 0
    for a, b in zip(sizes[:-1], sizes[1:]):
        W = p[off:off + a * b].reshape(a, b)
        off += a * b
        bb = p[off:off + b]
        off += b
        out.append((W, bb))
    return out


def _mm(a, b, exact):
    if exact:
        return np.asarray(a, np.float64) @ np.asarray(b, np.float64)
    return to_bf16(a) @ to_bf16(b)


def forward(p, sizes, x, exact=False):
    acts = [np.asarray(x, np.float64)]
    layers = unpack(np.asarray(p, np.float64), sizes)
    for li, (W, b) in enumerate(layers):
        z = _mm(acts[-1], W, exact) + b
        if li < len(layers) - 1:
            z = np.maximum(z, 0.0)
        if not exact:
            z = z.astype(np.float32).astype(np.float64)  # activations stored fp32
        acts.append(z)
    return acts


def loss_grad(p, sizes, x, y, exact=False):
    acts = forward(p, sizes, x, exact)
    layers = unpack(np.asarray(p, np.float64), sizes)
    out = acts[-1]
    d = out - np.asarray(y, np.float64)
    loss = float(np.mean(d * d))
    if exact:
        dz = 2.0 * d / d.size
    else:  # the device forms the output gradient in fp32
        d32 = out.astype(np.float32) - np.asarray(y, np.float32)
        dz = (np.float32(2.0 / d.size) * d32).astype(np.float64)
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        if li < len(layers) - 1:
            dz = np.where(acts[li + 1] > 0, dz, 0.0)
        gW = _mm(acts[li].T, dz, exact)
        gb = dz.sum(axis=0)
        grads.append((gW, gb))
        if li > 0:
            dz = _mm(dz, W.T, exact)
            if not exact:
                dz = dz.astype(np.float32).astype(np.float64)
    flat = []
    for gW, gb in reversed(grads):
        flat.append(gW.ravel())
        flat.append(gb)
    return loss, np.concatenate(flat)
