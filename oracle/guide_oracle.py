"""FP64 CPU oracle of the path-guiding network.

TEST INFRASTRUCTURE ONLY (the checker of the GPU guiding kernels).  The
reference renderer has no learned component; this restates the network the
GPU path defines, in float64 numpy, with an analytic backward pass that the
tests pin by finite differences.

Network (parameters in "output-major" layout, W[out, in]):
  features x (80) = 12-frequency encoding of the position normalised to the
                    scene box (3 x 12 x {sin, cos}, index 24 d + 2 k + s),
                    shading normal (72..74), outgoing direction (75..77), 0, 0
  h1 = relu(W1 x + b1)   W1 64 x 80
  h2 = relu(W2 h1 + b2)  W2 64 x 64
  h3 = relu(W3 h2 + b3)  W3 64 x 64
  y  = W4 h3 + b4        W4 48 x 64
Head, lobe j = 0..7 (y[5j .. 5j+4]):
  mu_j = m_j / |m_j|, kappa_j = exp(clip(k_j, KLO, KHI)), pi = softmax(a)
  alpha = 0.05 + 0.9 sigmoid(y[40])      (probability of sampling the guide)
  vMF density v(w) = kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1))
  q(w) = sum_j pi_j v_j(w);   p_mix(w) = alpha q(w) + (1 - alpha) p_bsdf(w)
Loss over a batch: L = -sum_i w_i log p_mix_i(omega_i) / sum_i w_i.
"""
import numpy as np

N_IN, N_H, N_OUT, N_LOBES, N_FREQ = 80, 64, 48, 8, 12
KLO, KHI = -3.0, 10.0
LAYERS = [(N_H, N_IN), (N_H, N_H), (N_H, N_H), (N_OUT, N_H)]
N_PARAMS = sum(o * i + o for o, i in LAYERS)


def unpack(theta):
    """Flat parameter vector -> [(W, b)] in the order W1, b1, W2, b2, ..."""
    out, off = [], 0
    for o, i in LAYERS:
        W = theta[off:off + o * i].reshape(o, i)
        off += o * i
        b = theta[off:off + o]
        off += o
        out.append((W, b))
    return out


def init_params(seed):
    """Xavier-uniform weights, zero biases (float32 master copy)."""
    rng = np.random.default_rng(seed)
    parts = []
    for o, i in LAYERS:
        lim = np.sqrt(6.0 / (i + o))
        parts.append(rng.uniform(-lim, lim, o * i))
        parts.append(np.zeros(o))
    return np.concatenate(parts).astype(np.float32)


def features(pos01, n, wo):
    """pos01 (B,3) in [0,1], n (B,3), wo (B,3) -> (B, 80)."""
    B = len(pos01)
    x = np.zeros((B, N_IN))
    k = np.arange(N_FREQ)
    for d in range(3):
        ang = (2.0 ** k)[None, :] * np.pi * pos01[:, d:d + 1]
        x[:, 24 * d + 2 * k] = np.sin(ang)
        x[:, 24 * d + 2 * k + 1] = np.cos(ang)
    x[:, 72:75] = n
    x[:, 75:78] = wo
    return x


def forward(theta, x, fp16_acts=False):
    """Returns y (B,48) and the cache of activations.  fp16_acts rounds the
    hidden activations and the input to float16 as the GPU kernels store them."""
    P = unpack(np.asarray(theta, np.float64))
    r = (lambda a: a.astype(np.float16).astype(np.float64)) if fp16_acts else (lambda a: a)
    acts = [r(x)]
    zs = []
    h = acts[0]
    for li, (W, b) in enumerate(P):
        z = h @ W.T + b
        zs.append(z)
        if li < 3:
            h = r(np.maximum(z, 0.0))
            acts.append(h)
        else:
            h = z
    return h, (acts, zs)


def head(y):
    m = y[:, :40].reshape(-1, N_LOBES, 5)
    mraw = m[:, :, 0:3]
    mlen = np.maximum(np.linalg.norm(mraw, axis=2, keepdims=True), 1e-6)
    mu = mraw / mlen
    kraw = m[:, :, 3]
    kc = np.clip(kraw, KLO, KHI)
    kappa = np.exp(kc)
    a = m[:, :, 4]
    a = a - a.max(axis=1, keepdims=True)
    pi = np.exp(a)
    pi /= pi.sum(axis=1, keepdims=True)
    s = 1.0 / (1.0 + np.exp(-y[:, 40]))
    alpha = 0.05 + 0.9 * s
    return dict(mu=mu, mlen=mlen, kraw=kraw, kappa=kappa, pi=pi, s=s, alpha=alpha)


def vmf(mu, kappa, w):
    """mu (B,J,3), kappa (B,J), w (B,3) -> densities (B,J)."""
    c = kappa / (2 * np.pi * (-np.expm1(-2 * kappa)))
    return c * np.exp(kappa * (np.einsum("bjk,bk->bj", mu, w) - 1.0))


def loss_and_grad_y(y, omega, p_bsdf, w):
    """Batch loss and dL/dy (B,48)."""
    H = head(y)
    v = vmf(H["mu"], H["kappa"], omega)
    q = (H["pi"] * v).sum(axis=1)
    alpha = H["alpha"]
    pm = alpha * q + (1 - alpha) * p_bsdf
    W = w.sum()
    L = -(w * np.log(pm)).sum() / W
    c = -w / (pm * W)
    dq = c * alpha
    dalpha = c * (q - p_bsdf)
    dy = np.zeros_like(y)
    pi, kappa, mu = H["pi"], H["kappa"], H["mu"]
    da = dq[:, None] * pi * (v - q[:, None])
    dv = dq[:, None] * pi
    cosw = np.einsum("bjk,bk->bj", mu, omega)
    e = np.exp(-2 * kappa)
    dlogc = 1.0 - 2 * kappa * e / (-np.expm1(-2 * kappa))  # kappa * d log C / d kappa
    inside = (H["kraw"] > KLO) & (H["kraw"] < KHI)
    dk = dv * v * (dlogc + kappa * (cosw - 1.0)) * inside
    dm = (dv * v * kappa)[:, :, None] * (omega[:, None, :] - cosw[:, :, None] * mu) / H["mlen"]
    g = np.zeros((len(y), N_LOBES, 5))
    g[:, :, 0:3] = dm
    g[:, :, 3] = dk
    g[:, :, 4] = da
    dy[:, :40] = g.reshape(len(y), 40)
    dy[:, 40] = dalpha * 0.9 * H["s"] * (1 - H["s"])
    return L, dy


def loss_and_grad(theta, x, omega, p_bsdf, w, fp16_acts=False):
    """Loss and gradient w.r.t. the flat parameter vector."""
    y, (acts, zs) = forward(theta, x, fp16_acts)
    L, dy = loss_and_grad_y(y, omega, p_bsdf, w)
    P = unpack(np.asarray(theta, np.float64))
    grads = [None] * 4
    d = dy
    for li in range(3, -1, -1):
        W, b = P[li]
        grads[li] = (d.T @ acts[li], d.sum(axis=0))
        if li > 0:
            d = (d @ W) * (zs[li - 1] > 0)
    flat = np.concatenate([np.concatenate([gw.reshape(-1), gb]) for gw, gb in grads])
    return L, flat, y


def adam_step(theta, m, v, g, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) with bias correction; t >= 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def mixture_pdf(y, omega, p_bsdf):
    H = head(y)
    q = (H["pi"] * vmf(H["mu"], H["kappa"], omega)).sum(axis=1)
    return H["alpha"] * q + (1 - H["alpha"]) * p_bsdf, q, H["alpha"]
