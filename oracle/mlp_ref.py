"""CPU oracle of the seeding MLP -- TEST INFRASTRUCTURE ONLY.

numpy forward and backward of the 192 -> 256 -> 256 -> 256 -> 4097 network (ReLU hidden layers,
softmax cross-entropy averaged over the batch). With bf16=True it rounds what the GPU rounds
(weights, hidden activations, the logit gradient and the hidden-layer data gradients) to bf16
with round-to-nearest-even and accumulates in float64, so the GPU kernels (fp32 accumulation)
must agree within accumulation-order error. With bf16=False everything is float64: that mode is
pinned by central finite differences (tests/test_mlp_oracle.py). Weights are W^T [out][in]."""
import numpy as np


def bf16_round(a):
    """Round float32 values to the nearest bf16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits_to_f64(bits):
    return (np.ascontiguousarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def forward(x0, params, bf16=True):
    """x0 [n][192] float64 inputs; params = [(W^T, b)] * 4. Returns activations and logits."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    acts = [np.asarray(x0, np.float64)]
    h = acts[0]
    for layer, (wt, b) in enumerate(params):
        z = h @ rnd(wt).T + np.asarray(b, np.float64)
        if layer < 3:
            h = rnd(np.maximum(z, 0.0).astype(np.float32)) if bf16 else np.maximum(z, 0.0)
            acts.append(h)
        else:
            return acts, z


def loss_and_grads(x0, params, labels, bf16=True):
    """Mean softmax cross-entropy and [(gW^T, gb)] per layer."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    acts, z = forward(x0, params, bf16)
    n = len(labels)
    zm = z - z.max(axis=1, keepdims=True)
    p = np.exp(zm)
    p /= p.sum(axis=1, keepdims=True)
    loss = float(np.mean(-np.log(p[np.arange(n), labels])))
    dz = p.copy()
    dz[np.arange(n), labels] -= 1.0
    dz /= n
    dy = rnd(dz.astype(np.float32)) if bf16 else dz
    grads = [None] * 4
    for layer in range(3, -1, -1):
        x = acts[layer]
        grads[layer] = (dy.T @ x, dy.sum(axis=0))
        if layer == 0:
            break
        dx = dy @ rnd(params[layer][0])
        dx = dx * (acts[layer] > 0)
        dy = rnd(dx.astype(np.float32)) if bf16 else dx
    return loss, grads
