"""CPU oracle of the learned directional proposal -- TEST INFRASTRUCTURE ONLY.

A von Mises-Fisher (vMF) mixture of 8 lobes predicted per shading point by a
5-layer MLP (96 -> 64 -> 64 -> 64 -> 64 -> 48, ReLU), trained online by
minimising the KL divergence from the incident-radiance distribution.  No
reference counterpart exists (the reference has no path guiding); this is
the specification the HIP kernels are checked against, in float64 with the
same bfloat16 roundings the device applies (weights, network inputs and the
hidden activations are bfloat16; accumulation is float32 or wider).

Layout (shared with the device):
  features (96): position normalised to the scene box, encoded as
    sin(2^k pi x), cos(2^k pi x) for k = 0..11 and each axis (index
    24 axis + 2k, +1), then wo (72..74), shading normal (75..77),
    roughness (78), zeros (79..95).
  parameters (float32 masters): W1[64][96] b1[64] W2[64][64] b2 W3 b3 W4 b4
    W5[48][64] b5[48] -- 21,808 values.
  outputs (48): lobe j uses 5j (weight logit), 5j+1 (log kappa), 5j+2..5j+4
    (mean direction, unnormalised); 40..47 are unused.
  mixture: w = softmax(logits), kappa = exp(-3 + 10 sigmoid(raw)) (a
    smooth map onto [e^-3, e^7] whose gradient never vanishes, so no lobe
    is frozen at a bound), mu = m / max(|m|, 1e-6).
  vMF density: kappa / (2 pi (1 - exp(-2 kappa))) * exp(kappa (mu.w - 1)).
  sampling (Jakob 2012, "Numerically stable sampling of the von
    Mises-Fisher distribution on S^2"): cos = 1 + log(u1 + (1 - u1)
    exp(-2 kappa)) / kappa, phi = 2 pi u2, rotated into a frame of mu.
  loss of one record (features x, direction d, weight c >= 0):
    -c log q(d); the batch loss is the sum divided by the sum of c.
  Adam: beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected.
"""
import numpy as np

N_IN, N_HID, N_OUT, N_LOBES, N_FREQ = 96, 64, 48, 8, 12
LAYERS = ((96, 64), (64, 64), (64, 64), (64, 64), (64, 48))
KAPPA_LO, KAPPA_HI = -3.0, 7.0


def n_params():
    return sum(o * i + o for i, o in LAYERS)


def split(theta):
    """Views (W, b) per layer of a flat parameter vector."""
    out, k = [], 0
    for i, o in LAYERS:
        W = theta[k:k + o * i].reshape(o, i)
        k += o * i
        b = theta[k:k + o]
        k += o
        out.append((W, b))
    return out


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), as floats."""
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    nan = np.isnan(x)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(nan, x, out)


def init_params(seed):
    """He-uniform weights (last layer x0.1), zero biases, from a seeded
    generator (any deterministic initialisation; the device copies it)."""
    rng = np.random.default_rng(seed)
    parts = []
    for li, (i, o) in enumerate(LAYERS):
        a = np.sqrt(6.0 / i) * (0.1 if li == len(LAYERS) - 1 else 1.0)
        parts.append(rng.uniform(-a, a, size=(o, i)).astype(np.float32).reshape(-1))
        parts.append(np.zeros(o, np.float32))
    return np.concatenate(parts)


def features(pos_n, wo, n, rough):
    """(N, 96) float32 network inputs from normalised positions (N, 3),
    outgoing directions, shading normals and roughness (N,)."""
    pos_n = np.asarray(pos_n, np.float64)
    N = pos_n.shape[0]
    f = np.zeros((N, N_IN), np.float64)
    for a in range(3):
        for k in range(N_FREQ):
            ang = (2.0 ** k) * np.pi * pos_n[:, a]
            f[:, 24 * a + 2 * k] = np.sin(ang)
            f[:, 24 * a + 2 * k + 1] = np.cos(ang)
    f[:, 72:75] = wo
    f[:, 75:78] = n
    f[:, 78] = rough
    return f.astype(np.float32)


def forward(theta, x, keep=False, exact=False):
    """Raw outputs (N, 48) in float64 from bfloat16 weights and inputs, with
    bfloat16 hidden activations (exact: no rounding at all, the smooth
    function whose gradients finite differences can check).  keep: also
    return the layer inputs and pre-activations."""
    rnd = (lambda a: np.asarray(a, np.float64)) if exact else (lambda a: bf16(a).astype(np.float64))
    theta = np.asarray(theta, np.float64 if exact else np.float32)
    layers = split(theta)
    h = rnd(x)
    acts = [h]
    pre = []
    for li, (W, b) in enumerate(layers):
        z = h @ rnd(W).T + np.asarray(b, np.float64)
        pre.append(z)
        if li < len(layers) - 1:
            r = np.maximum(z, 0.0)
            h = r if exact else bf16(r.astype(np.float32)).astype(np.float64)
            acts.append(h)
        else:
            h = z
    return (h, acts, pre) if keep else h


def decode(out):
    """(w, kappa, mu, norms, s) of raw outputs (N, 48); s = sigmoid of the
    raw concentration output."""
    out = np.asarray(out, np.float64)
    N = out.shape[0]
    o = out[:, :40].reshape(N, N_LOBES, 5)
    logits = o[:, :, 0]
    z = logits - logits.max(axis=1, keepdims=True)
    e = np.exp(z)
    w = e / e.sum(axis=1, keepdims=True)
    s = 1.0 / (1.0 + np.exp(-o[:, :, 1]))
    kappa = np.exp(KAPPA_LO + (KAPPA_HI - KAPPA_LO) * s)
    m = o[:, :, 2:5]
    nm = np.maximum(np.linalg.norm(m, axis=2), 1e-6)
    mu = m / nm[:, :, None]
    return w, kappa, mu, nm, s


def vmf_pdf(kappa, mu, d):
    """Density of vMF(mu, kappa) at unit directions d (broadcasting)."""
    kappa = np.asarray(kappa, np.float64)
    c = kappa / (2 * np.pi * (-np.expm1(-2 * kappa)))
    return c * np.exp(kappa * (np.sum(mu * d, axis=-1) - 1.0))


def mixture_pdf(w, kappa, mu, d):
    """q(d) for mixtures (N, 8) at directions d (N, 3)."""
    return np.sum(w * vmf_pdf(kappa, mu, d[:, None, :]), axis=1)


def vmf_sample(kappa, mu, u1, u2):
    """Jakob 2012 sampler: directions (N, 3) for lobes (kappa (N,), mu (N, 3))."""
    kappa = np.asarray(kappa, np.float64)
    c = 1.0 + np.log(u1 + (1.0 - u1) * np.exp(-2.0 * kappa)) / kappa
    c = np.clip(c, -1.0, 1.0)
    s = np.sqrt(np.maximum(0.0, 1.0 - c * c))
    phi = 2 * np.pi * u2
    local = np.stack([s * np.cos(phi), s * np.sin(phi), c], axis=-1)
    return frame_to_world(mu, local)


def frame_of(mu):
    """Orthonormal (t, b) completing mu (Duff et al. 2017 branchless basis)."""
    sign = np.where(mu[..., 2] >= 0, 1.0, -1.0)
    a = -1.0 / (sign + mu[..., 2])
    b = mu[..., 0] * mu[..., 1] * a
    t = np.stack([1.0 + sign * mu[..., 0] ** 2 * a, sign * b, -sign * mu[..., 0]], axis=-1)
    bb = np.stack([b, sign + mu[..., 1] ** 2 * a, -mu[..., 1]], axis=-1)
    return t, bb


def frame_to_world(mu, local):
    t, b = frame_of(mu)
    return t * local[..., 0:1] + b * local[..., 1:2] + mu * local[..., 2:3]


def loss_and_grad(theta, x, d, c, exact=False):
    """Batch loss sum(-c log q(d)) / sum(c) and its gradient wrt theta
    (float64, analytic; backprop through bfloat16-rounded activations
    treated as the identity, as the device does; exact: the unrounded
    network)."""
    theta = np.asarray(theta, np.float64 if exact else np.float32)
    layers = split(theta)
    out, acts, pre = forward(theta, x, keep=True, exact=exact)
    w, kappa, mu, nm, sg = decode(out)
    d = np.asarray(d, np.float64)
    c = np.asarray(c, np.float64)
    N = out.shape[0]
    f = vmf_pdf(kappa, mu, d[:, None, :])            # (N, 8)
    q = np.sum(w * f, axis=1)
    C = c.sum()
    loss = float(np.sum(-c * np.log(np.maximum(q, 1e-30))) / C)
    g_q = -c / np.maximum(q, 1e-30) / C              # dL/dq
    wf = w * f
    g_logit = g_q[:, None] * (wf - w * q[:, None])
    dot = np.sum(mu * d[:, None, :], axis=2)
    dlogf_dk = 1.0 / kappa - 2.0 * np.exp(-2 * kappa) / (-np.expm1(-2 * kappa)) + (dot - 1.0)
    g_kraw = g_q[:, None] * wf * dlogf_dk * kappa * (KAPPA_HI - KAPPA_LO) * sg * (1.0 - sg)
    proj = d[:, None, :] - dot[:, :, None] * mu
    g_m = (g_q[:, None] * wf * kappa)[:, :, None] * proj / nm[:, :, None]
    g_out = np.zeros((N, N_OUT))
    go = g_out[:, :40].reshape(N, N_LOBES, 5)
    go[:, :, 0] = g_logit
    go[:, :, 1] = g_kraw
    go[:, :, 2:5] = g_m
    g_out[:, :40] = go.reshape(N, 40)
    grads = []
    g = g_out
    for li in range(len(layers) - 1, -1, -1):
        W, b = layers[li]
        a_in = acts[li]
        gW = g.T @ a_in
        gb = g.sum(axis=0)
        grads.append((gW, gb))
        if li > 0:
            Wr = np.asarray(W, np.float64) if exact else bf16(W).astype(np.float64)
            g = (g @ Wr) * (pre[li - 1] > 0)
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])
    return loss, flat


class Adam:
    def __init__(self, n, lr=1e-2, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, theta, g):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return (np.asarray(theta, np.float64) - self.lr * mh / (np.sqrt(vh) + self.eps)).astype(np.float32)
