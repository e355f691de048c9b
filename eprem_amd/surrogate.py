"""Learned large-step surrogate of streaming (the MLP form of the hot path).

One network application advances the mu row of one kept node of one
(stream, energy) slab by K streaming-and-relaxation subcycles of the plain
path (DiffuseStreamData, src/energeticParticles.c:2352-2430) and the
focusing of the same time (:1683-1896).  Inputs and targets come from
`EpSolver.surrogate_samples` (per mu cell: ln f at the previous node minus
ln f at the node, ln f at the node minus its mean over mu, ln f at the
next node minus ln f at the node; plus six scalars; target ln f after the
K subcycles and the focusing, relative to the same mean); the network is `runtime.Mlp` (matrix-core
products, bf16 operands, fp32 accumulation), trained on the device.

The large step is what a single node can do alone: the exact relaxation
of the anisotropy over K dtProp, f -> iso + (f - iso) exp(-K dtProp / tau),
with no transport; it is also the method without the network.  The
network is trained on, and in the EP update applies, the correction
ln f - ln f_large (`corrections`, `predict`), so that its bf16 arithmetic
only touches a small quantity.
Errors are stated on the omnidirectional flux 2 E <f>_mu at the node
(computeFlux, src/unifiedOutput.c:153-180), the quantity the accuracy
gate is defined on.
"""
import numpy as np

from . import runtime

LAYERS = (128, 128, 128)


def relaxation_step(x, M):
    """Target-space prediction of the relaxation-only large step from the
    features (ln f at the node relative to its mean, and K dtProp/tau)."""
    lf = x[:, M:2 * M].astype(np.float64)
    f = np.exp(lf)
    iso = f.mean(axis=1, keepdims=True)
    q = np.exp(-x[:, 3 * M + 1].astype(np.float64))[:, None]
    return np.log(iso + (f - iso) * q)


def corrections(x, y, M):
    """Training targets of the corrected large step: ln f after the plain
    subcycles minus ln f of the large step (both relative to the node's
    mean ln f, so the mean cancels)."""
    return (np.asarray(y, np.float64) - relaxation_step(x, M)).astype(np.float32)


def predict(net, x, M):
    """ln f (relative to the node's mean ln f) of the corrected large step:
    the large step plus the network's correction."""
    return relaxation_step(x, M) + net.forward(x).astype(np.float64)


def flux_error(pred, y):
    """Relative error of <f>_mu (the node's omnidirectional flux up to the
    factor 2E) between predicted and true ln f rows."""
    fp = np.exp(np.asarray(pred, np.float64)).mean(axis=1)
    ft = np.exp(np.asarray(y, np.float64)).mean(axis=1)
    return np.abs(fp - ft) / ft


def summary(err):
    return {"median": float(np.median(err)), "p95": float(np.percentile(err, 95)),
            "max": float(err.max()), "within_1pct": float(np.mean(err <= 0.01))}


def train(x, y, epochs=20, batch=4096, lr=2e-3, seed=1, layers=LAYERS):
    """A fresh network trained on (x, y); returns (net, loss history)."""
    sizes = [x.shape[1], *layers, y.shape[1]]
    net = runtime.Mlp(sizes, seed=seed)
    hist = net.train(x, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return net, hist
