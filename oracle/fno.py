"""CPU oracle of the learned warm start: a Fourier neural operator (FNO) in
float64 numpy, forward and hand-written backward.  Test infrastructure only
(the GPU path trains and runs the network with its own BF16 MFMA kernels);
the spectral convolution here uses numpy's FFT, an independent route from
the GPU's truncated-DFT matrix products.

Network (per sample, fields on an nx x ny periodic grid, channels last):
  inputs  a = [h / h_rms, log10(p0) + 2.5]          (2 channels)
  lift    v0 = a P + bP                             (2 -> C, pointwise)
  layers  z_l = K_l(v_{l-1}) + v_{l-1} W_l + b_l,   v_l = gelu(z_l)
          (l = 1..L, no gelu after the last layer)
  K_l(v)  = irfft2( R_l(kx, ky) . rfft2(v) ) on the retained modes
          kx in [-M, M), ky in [0, M] (the 2M x 2M block of the full
          spectrum; R_l complex C x C per mode), zero elsewhere
  project u = v_L Q + bQ                            (C -> 3, pointwise)
gelu(x) = x Phi(x) (exact, erf form).  Loss: mean over samples, sites and
components of (u - u_target)^2.

Parity pins (tests/test_oracle.py): finite differences of every parameter
group and of the input; the spectral convolution against a direct
(brute-force) DFT on a tiny grid.
"""
import math

import numpy as np

try:
    from scipy.special import erf as _erf
except ImportError:  # pragma: no cover
    _erf = np.vectorize(math.erf)

SQRT2 = math.sqrt(2.0)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / SQRT2))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / SQRT2)) + x * np.exp(-0.5 * x * x) / math.sqrt(
        2.0 * math.pi)


def mode_indices(nx, M):
    """Grid indices of the retained kx modes [-M, M) in rfft2 order."""
    return np.concatenate([np.arange(0, M), np.arange(nx - M, nx)])


def init_params(rng, C=32, M=16, L=4, cin=2, cout=3):
    """Parameter dict with the shapes the GPU network uses."""
    p = {"P": rng.standard_normal((cin, C)) / math.sqrt(cin),
         "bP": np.zeros(C),
         "Q": rng.standard_normal((C, cout)) / math.sqrt(C),
         "bQ": np.zeros(cout)}
    for l in range(L):
        p["W%d" % l] = rng.standard_normal((C, C)) / math.sqrt(C)
        p["b%d" % l] = np.zeros(C)
        s = 1.0 / (C * C)
        p["R%d" % l] = s * (rng.random((2 * M, M + 1, C, C)) +
                            1j * rng.random((2 * M, M + 1, C, C)))
    return p


def spectral(v, R, M):
    """K(v) for v (B, nx, ny, C): rfft2 over the grid, mode-wise channel
    mixing with R (2M, M+1, C, C) on the retained block, irfft2."""
    B, nx, ny, C = v.shape
    V = np.fft.rfft2(v, axes=(1, 2))
    kx = mode_indices(nx, M)
    out = np.zeros_like(V)
    out[:, kx, :M + 1, :] = np.einsum("bxyi,xyio->bxyo", V[:, kx, :M + 1, :], R)
    return np.fft.irfft2(out, s=(nx, ny), axes=(1, 2))


def _spectral_back(v, g, R, M):
    """Gradients of <g, spectral(v, R)> w.r.t. v and R (v real)."""
    B, nx, ny, C = v.shape
    kx = mode_indices(nx, M)
    # irfft2 adjoint: the real inverse transform y = irfft2(Y) has
    # dL/dY = rfft2(g) / (nx ny) * w_ky with w = 1 at ky = 0 and at the
    # Nyquist column, 2 elsewhere (conjugate-symmetric pairs), as the
    # derivative with respect to Re Y and Im Y packed as a complex number
    G = np.fft.rfft2(g, axes=(1, 2)) / (nx * ny)
    w = np.full(ny // 2 + 1, 2.0)
    w[0] = 1.0
    if ny % 2 == 0:
        w[-1] = 1.0
    G = G * w[None, None, :, None]
    Gk = G[:, kx, :M + 1, :]                      # dL/d(out) (conj form)
    V = np.fft.rfft2(v, axes=(1, 2))
    Vk = V[:, kx, :M + 1, :]
    # out = V R  ->  dL/dR = sum_b conj(V)^T dL/dout (complex, conj form)
    dR = np.einsum("bxyi,bxyo->xyio", np.conj(Vk), Gk)
    # dL/dV = dL/dout R^H
    dVk = np.einsum("bxyo,xyio->bxyi", Gk, np.conj(R))
    dV = np.zeros_like(V)
    dV[:, kx, :M + 1, :] = dVk
    # rfft2 adjoint for real v: dL/dv = Re(sum over the full spectrum) ->
    # with V = rfft2(v): dL/dv = irfft2-like synthesis of dV with weights
    # 1 (the derivative of Re/Im pairs of a real transform)
    dv = np.real(_rfft2_adjoint(dV, nx, ny))
    return dv, dR


def _rfft2_adjoint(dV, nx, ny):
    """Adjoint of v -> rfft2(v) (real v) for dL/dV given in conj form
    (dL/dRe + i dL/dIm): sum_k Re(conj(dV_k) ... ) expressed as a synthesis
    e^{+i q r} with no Hermitian doubling."""
    full = np.zeros(dV.shape[:1] + (nx, ny) + dV.shape[3:], dtype=complex)
    full[:, :, :dV.shape[2], :] = dV
    return np.fft.ifft2(full, axes=(1, 2)) * (nx * ny)


def forward(p, a, M=16, L=4, keep=False):
    """a: (B, nx, ny, 2) inputs -> u (B, nx, ny, 3) (and the saved
    activations when keep)."""
    v = a @ p["P"] + p["bP"]
    acts = [v]
    zs = []
    for l in range(L):
        z = spectral(v, p["R%d" % l], M) + v @ p["W%d" % l] + p["b%d" % l]
        zs.append(z)
        v = gelu(z) if l < L - 1 else z
        acts.append(v)
    u = v @ p["Q"] + p["bQ"]
    if keep:
        return u, (a, acts, zs)
    return u


def loss_and_grads(p, a, target, M=16, L=4):
    """Mean-squared loss and its gradients (dict like p, plus 'input')."""
    u, (a, acts, zs) = forward(p, a, M, L, keep=True)
    n = u.size
    diff = u - target
    loss = float(np.sum(diff * diff) / n)
    du = 2.0 * diff / n
    g = {}
    vL = acts[-1]
    g["Q"] = np.einsum("bxyc,bxyo->co", vL, du)
    g["bQ"] = du.sum(axis=(0, 1, 2))
    dv = du @ p["Q"].T
    for l in reversed(range(L)):
        dz = dv * gelu_grad(zs[l]) if l < L - 1 else dv
        vin = acts[l]
        g["W%d" % l] = np.einsum("bxyc,bxyo->co", vin, dz)
        g["b%d" % l] = dz.sum(axis=(0, 1, 2))
        dv_sp, dR = _spectral_back(vin, dz, p["R%d" % l], M)
        g["R%d" % l] = dR
        dv = dz @ p["W%d" % l].T + dv_sp
    g["P"] = np.einsum("bxyc,bxyo->co", a, dv)
    g["bP"] = dv.sum(axis=(0, 1, 2))
    g["input"] = dv @ p["P"].T
    return loss, g


def features(h, p0):
    """Network inputs for height maps h (B, nx, ny) and loads p0 (B,)."""
    h = np.asarray(h, dtype=np.float64)
    rms = np.sqrt(np.mean(h * h, axis=(1, 2), keepdims=True))
    rms = np.where(rms > 0, rms, 1.0)
    a = np.empty(h.shape + (2,))
    a[..., 0] = h / rms
    a[..., 1] = (np.log10(np.asarray(p0, dtype=np.float64)) + 2.5)[:, None, None]
    return a


# parameter order of the flat vector shared with the GPU network
def param_layout(C=32, M=16, L=4, cin=2, cout=3):
    lay = [("P", (cin, C)), ("bP", (C,))]
    for l in range(L):
        lay += [("W%d" % l, (C, C)), ("b%d" % l, (C,)),
                ("R%d" % l, (2 * M, M + 1, C, C, 2))]
    lay += [("Q", (C, cout)), ("bQ", (cout,))]
    return lay


def flatten(p, C=32, M=16, L=4):
    out = []
    for name, shape in param_layout(C, M, L):
        x = p[name]
        if name.startswith("R"):
            x = np.stack([x.real, x.imag], axis=-1)
        out.append(np.asarray(x, dtype=np.float64).reshape(-1))
    return np.concatenate(out)


def unflatten(vec, C=32, M=16, L=4):
    p, off = {}, 0
    for name, shape in param_layout(C, M, L):
        k = int(np.prod(shape))
        x = vec[off:off + k].reshape(shape)
        off += k
        if name.startswith("R"):
            x = x[..., 0] + 1j * x[..., 1]
        p[name] = x
    assert off == vec.size
    return p
