"""Learned warm start on the device: the BF16 MFMA matrix-product kernel
against a float64 product of the same BF16-rounded operands, and the
network's forward and gradients against the float64 CPU oracle
(oracle/fno.py, itself pinned by finite differences).

Tolerances: GEMM 1e-5 relative (FP32 accumulation of exact BF16
products); network output 5e-2 of max |u| (max-norm) and 2e-2 (rms) --
BF16 operands in every product of four layers; gradients per parameter
group within 8e-2 (norm of the difference / norm) and cosine > 0.995.
"""
import ctypes

import numpy as np
import pytest

import gfmd_hip
import fno as F

pytestmark = pytest.mark.gpu


def bf16_bits(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bits_to_f64(b):
    return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("case", ["nn", "tn", "nt", "ragged", "batch_acc"])
def test_gemm_kernel_matches_host_product(case):
    import torch
    rng = np.random.default_rng(hash(case) % 1000)
    if case == "nn":
        M, N, K, B = 130, 96, 200, 1
    elif case == "tn":
        M, N, K, B = 32, 32, 1000, 3
    elif case == "nt":
        M, N, K, B = 257, 32, 32, 1
    elif case == "ragged":
        M, N, K, B = 34, 544, 37, 2
    else:
        M, N, K, B = 64, 70, 64, 4
    A = rng.standard_normal((B, M, K)).astype(np.float32)
    Bm = rng.standard_normal((B, K, N)).astype(np.float32)
    Ab, Bb = bf16_bits(A), bf16_bits(Bm)
    # storage layouts: A row-major or transposed, B row-major or transposed
    if case in ("tn",):
        As = np.ascontiguousarray(np.transpose(Ab, (0, 2, 1)))  # (B, K, M)
        sa = [M * K, 1, M]
    else:
        As = Ab
        sa = [M * K, K, 1]
    if case in ("nt",):
        Bs = np.ascontiguousarray(np.transpose(Bb, (0, 2, 1)))  # (B, N, K)
        sb = [K * N, 1, K]
    else:
        Bs = Bb
        sb = [K * N, N, 1]
    ref = np.einsum("bmk,bkn->bmn", bits_to_f64(Ab), bits_to_f64(Bb))
    C0 = rng.standard_normal((B, M, N)).astype(np.float32)
    acc = 1 if case == "batch_acc" else 0
    if acc:
        ref = ref + C0
    dA = torch.tensor(As.view(np.int16), device="cuda")
    dB = torch.tensor(Bs.view(np.int16), device="cuda")
    dC = torch.tensor(C0, device="cuda")
    torch.cuda.synchronize()
    L = gfmd_hip.load()
    arr = lambda v: (ctypes.c_longlong * 3)(*v)
    gfmd_hip.check(L.gfmd_fno_gemm_test(M, N, K, B, dA.data_ptr(), arr(sa),
                                        dB.data_ptr(), arr(sb), dC.data_ptr(),
                                        arr([M * N, N, 1]), acc))
    got = dC.cpu().numpy().astype(np.float64)
    assert np.abs(got - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max()) * \
        np.sqrt(K / 32)


def small_problem(seed=3, B=2, nx=40, ny=36):
    rng = np.random.default_rng(seed)
    p = F.init_params(rng)
    # larger mode weights so the spectral path matters
    for l in range(4):
        p["R%d" % l] = p["R%d" % l] * 8.0
        p["b%d" % l] = 0.1 * rng.standard_normal(32)
    a = np.stack([rng.standard_normal((B, nx, ny)),
                  np.full((B, nx, ny), 0.3)], axis=-1)
    t = 0.5 * rng.standard_normal((B, nx, ny, 3))
    return p, a, t


def test_fno_forward_matches_oracle():
    p, a, _ = small_problem()
    B, nx, ny = a.shape[:3]
    net = gfmd_hip.FNO(nx, ny, B)
    net.set_params(F.flatten(p).astype(np.float32))
    got = net.forward(a.astype(np.float32)).astype(np.float64)
    ref = F.forward(p, a)
    err = np.abs(got - ref)
    print("FNO forward: max err", err.max() / np.abs(ref).max(),
          "rms err", np.sqrt(np.mean(err ** 2) / np.mean(ref ** 2)), flush=True)
    assert err.max() <= 5e-2 * np.abs(ref).max()
    assert np.sqrt(np.mean(err ** 2) / np.mean(ref ** 2)) <= 2e-2
    net.close()


def test_fno_gradients_match_oracle():
    p, a, t = small_problem(seed=5)
    B, nx, ny = a.shape[:3]
    net = gfmd_hip.FNO(nx, ny, B)
    net.set_params(F.flatten(p).astype(np.float32))
    loss, g = net.gradients(a.astype(np.float32), t.astype(np.float32))
    lref, gref = F.loss_and_grads(p, a, t)
    assert abs(loss - lref) <= 2e-2 * lref
    gflat = F.flatten({k: gref[k] for k in p}).astype(np.float64)
    off = 0
    for name, shape in F.param_layout():
        k = int(np.prod(shape))
        x, y = g[off:off + k].astype(np.float64), gflat[off:off + k]
        off += k
        rel = np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-30)
        cos = x @ y / max(np.linalg.norm(x) * np.linalg.norm(y), 1e-30)
        print("grad", name, "rel", rel, "cos", cos, flush=True)
        assert rel <= 8e-2 and cos >= 0.995, name
    net.close()


def test_fno_training_reduces_loss():
    p, a, t = small_problem(seed=8)
    B, nx, ny = a.shape[:3]
    net = gfmd_hip.FNO(nx, ny, B)
    net.set_params(F.flatten(p).astype(np.float32))
    a32, t32 = a.astype(np.float32), t.astype(np.float32)
    l0 = net.train_step(a32, t32, 1e-3)
    for _ in range(300):
        l = net.train_step(a32, t32, 1e-3)
    print("FNO overfit: loss", l0, "->", l, flush=True)
    assert l < 0.2 * l0
    net.close()
