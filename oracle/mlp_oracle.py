"""CPU oracle of the learned-proposal network (test infrastructure only).

Restates, in float64 numpy, the 3-layer MLP of the learned branch-length /
composition move: forward pass, Gaussian negative log likelihood, backward
pass and one Adam step.  The reference (p4) has no learned proposal, so
there is no reference file to follow; the oracle is pinned instead by the
mathematics:
  * gradients against central finite differences of the float64 loss
    (tests/test_learned.py::test_oracle_gradients_finite_difference);
  * the loss against a direct evaluation of the Gaussian log density
    (scipy.stats.norm.logpdf) on random inputs;
  * Adam against its closed form on the first step (|step| = lr per
    coordinate where the gradient is nonzero).

With bf16=True every GEMM operand is rounded to bfloat16 (round to nearest
even) exactly where the GPU kernels round them, so that the GPU result
differs from the oracle only by float32 accumulation order.

Parameter vector layout: W1 [H][F], b1 [H], W2 [H][H], b2 [H], W3 [O][H],
b3 [O]; O = 2D (means, then log standard deviations).
"""
import numpy


def bf16_round(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned
    as float64."""
    a = numpy.ascontiguousarray(x, dtype=numpy.float32)
    u = a.view(numpy.uint32).astype(numpy.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(numpy.uint32).view(numpy.float32).astype(numpy.float64)


def unpack(theta, F, H, O):
    theta = numpy.asarray(theta, dtype=numpy.float64)
    i = 0
    out = []
    for shape in ((H, F), (H,), (H, H), (H,), (O, H), (O,)):
        n = int(numpy.prod(shape))
        out.append(theta[i:i + n].reshape(shape))
        i += n
    if i != theta.size:
        raise ValueError("parameter vector has %d values, expected %d" % (theta.size, i))
    return out


def n_params(F, H, O):
    return H * F + H + H * H + H + O * H + O


def _r(x, bf16):
    return bf16_round(x) if bf16 else numpy.asarray(x, dtype=numpy.float64)


def forward(theta, X, F, H, O, bf16=False):
    """Returns (OUT, cache)."""
    W1, b1, W2, b2, W3, b3 = unpack(theta, F, H, O)
    X = numpy.asarray(X, dtype=numpy.float64)
    Z1 = _r(X, bf16) @ _r(W1, bf16).T + b1
    H1 = numpy.maximum(Z1, 0.0)
    Z2 = _r(H1, bf16) @ _r(W2, bf16).T + b2
    H2 = numpy.maximum(Z2, 0.0)
    OUT = _r(H2, bf16) @ _r(W3, bf16).T + b3
    return OUT, (X, Z1, H1, Z2, H2)


def nll(OUT, Y, mask):
    """Sum over rows and masked coordinates of s + (y - mu)^2 e^{-2s} / 2
    (the Gaussian negative log density without its constant)."""
    D = Y.shape[1]
    mu, s = OUT[:, :D], OUT[:, D:2 * D]
    r = Y - mu
    return float((mask * (s + 0.5 * r * r * numpy.exp(-2.0 * s))).sum())


def loss_and_grad(theta, X, Y, mask, F, H, O, bf16=False):
    """(loss summed over rows, gradient of loss / n)."""
    W1, b1, W2, b2, W3, b3 = unpack(theta, F, H, O)
    n = X.shape[0]
    D = O // 2
    OUT, (X, Z1, H1, Z2, H2) = forward(theta, X, F, H, O, bf16)
    Y = numpy.asarray(Y, dtype=numpy.float64)
    mask = numpy.asarray(mask, dtype=numpy.float64)
    L = nll(OUT, Y, mask)
    mu, s = OUT[:, :D], OUT[:, D:]
    r = Y - mu
    w = numpy.exp(-2.0 * s)
    dOUT = numpy.concatenate([-mask * r * w, mask * (1.0 - r * r * w)], axis=1) / n
    if bf16:
        # the kernels keep dOUT and the backpropagated activations in float32
        dOUT = dOUT.astype(numpy.float32).astype(numpy.float64)
    gW3 = _r(dOUT, bf16).T @ _r(H2, bf16)
    gb3 = dOUT.sum(axis=0)
    dH2 = _r(dOUT, bf16) @ _r(W3, bf16)
    dZ2 = dH2 * (Z2 > 0)
    gW2 = _r(dZ2, bf16).T @ _r(H1, bf16)
    gb2 = dZ2.sum(axis=0)
    dH1 = _r(dZ2, bf16) @ _r(W2, bf16)
    dZ1 = dH1 * (Z1 > 0)
    gW1 = _r(dZ1, bf16).T @ _r(X, bf16)
    gb1 = dZ1.sum(axis=0)
    g = numpy.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return L, g


def adam_step(theta, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1); step counts from 1."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return theta - lr * mh / (numpy.sqrt(vh) + eps), m, v


def init_like_library(F, H, O, rng):
    """A random parameter vector of the same scale as p4g_mlpNew's."""
    W1 = rng.normal(0, numpy.sqrt(2.0 / F), (H, F))
    W2 = rng.normal(0, numpy.sqrt(2.0 / H), (H, H))
    W3 = rng.normal(0, 0.1 * numpy.sqrt(2.0 / H), (O, H))
    return numpy.concatenate([W1.ravel(), rng.normal(0, 0.1, H), W2.ravel(), rng.normal(0, 0.1, H),
                              W3.ravel(), rng.normal(0, 0.1, O)])
