"""ORACLE - test infrastructure only: the surrogate network's forward and backward pass in numpy FP64.

Restates the dense ReLU network of chaos_amd/surrogate.py (Z = W A + b per layer, feature-major
activations, BCE-with-logits on output 0 plus a weighted half squared error on output 1, batch mean) with
no BF16 rounding of activations, so the GPU kernels' gradients are checked against exact arithmetic on the
same BF16-valued weights and inputs. Its own check is a central-difference test (tests/test_mlp_oracle.py)."""
import numpy as np


def _same(a):
    return a


def forward(Ws, bs, X, rnd=_same):
    """Activations per layer; `rnd` rounds each hidden activation (identity: exact FP64)."""
    acts = [X]
    A = X
    for l, (W, b) in enumerate(zip(Ws, bs)):
        Z = W @ A + b[:, None]
        A = Z if l == len(Ws) - 1 else rnd(np.maximum(Z, 0.0))
        acts.append(A)
    return acts


def loss_and_dout(out, y, t, w, wreg):
    z = out[0]
    sp = np.where(z > 0, z + np.log1p(np.exp(-z)), np.log1p(np.exp(z)))
    B = out.shape[1]
    lb = np.sum(sp - y * z) / B
    e = out[1] - t
    lr = np.sum(0.5 * wreg * w * e * e) / B
    d = np.stack([(1 / (1 + np.exp(-z)) - y) / B, wreg * w * e / B])
    return lb, lr, d


def backward(Ws, acts, dout, rnd=_same):
    """Weight and bias gradients; `rnd` rounds dout and each back-propagated error (identity: exact FP64).
    With a BF16 rounding it models the error budget of a BF16-storage network, which bounds the GPU's error."""
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    dZ = rnd(dout)
    for l in reversed(range(len(Ws))):
        gW[l] = dZ @ acts[l].T
        gb[l] = dZ.sum(axis=1)
        if l > 0:
            dZ = rnd((Ws[l].T @ dZ) * (acts[l] > 0))
    return gW, gb
