"""Pins the float64 restatement of the learned preconditioner's network
(oracle/fno_oracle.py) before it checks the HIP kernels: its hand-written
backward pass against central finite differences of its loss, along random
directions in the whole parameter space and in each parameter block (a
transposed operand, a dropped c_m = 2 mode factor or a sign slip in the complex
mixing fails one of them), and the spectral path's DFT / inverse DFT against
numpy's FFT on a uniform grid."""
import numpy as np
import pytest

import fno_oracle as F


def _samples(rng, sizes=(40, 70)):
    out = []
    for N in sizes:
        t = np.sort(rng.uniform(0, 2 * np.pi, N))
        out.append({"X": rng.standard_normal((N, F.CIN)) * 0.5, "mphys": rng.uniform(0.2, 1.0, N),
                    "t": t, "wt": np.full(N, 2 * np.pi / N),
                    "v": rng.standard_normal(N) + 1j * rng.standard_normal(N),
                    "target": rng.standard_normal(N) + 1j * rng.standard_normal(N)})
    return out


def _params(rng):
    p = np.zeros(F.NPARAMS)
    for name, (o, n) in F.LAY.items():
        scale = 0.02 if name.startswith("R") else 0.2
        p[o:o + n] = scale * rng.standard_normal(n)
    return p


@pytest.mark.parametrize("block", ["all", "wl", "bl", "R0", "R3", "W1", "b2", "wp1", "bp1", "wp2", "bp2"])
def test_backward_matches_finite_differences(block):
    rng = np.random.default_rng(3)
    samples = _samples(rng)
    p = _params(rng)
    loss, g = F.loss_grad(p, samples)
    for trial in range(2):
        d = np.zeros_like(p)
        if block == "all":
            d = rng.standard_normal(p.size)
        else:
            o, n = F.LAY[block]
            d[o:o + n] = rng.standard_normal(n)
        h = 1e-6 / np.linalg.norm(d) * max(1.0, np.linalg.norm(p))
        lp, _ = F.loss_grad(p + h * d, samples, grad=False)
        lm, _ = F.loss_grad(p - h * d, samples, grad=False)
        fd = (lp - lm) / (2 * h)
        an = float(np.dot(g, d))
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3 * abs(loss)), (block, trial, fd, an)


def test_unused_outputs_and_padding_carry_no_gradient():
    rng = np.random.default_rng(4)
    p = _params(rng)
    _, g = F.loss_grad(p, _samples(rng))
    P = F.unpack(g)
    assert np.all(P["wp2"][:, 2:] == 0) and np.all(P["bp2"][2:] == 0)
    assert np.all(P["wl"][F.CIN:] == 0)  # padded input channels are zero
    # mode 0 of a real signal is real (Im Hhat_0 = 0) and Im G_0 is dropped, so the
    # imaginary part of mode 0's mixing matrix never reaches the output
    assert np.all(P["R0"][0, :, :, 1] == 0) and np.any(P["R0"][0, :, :, 0] != 0)


def test_spectral_path_is_a_truncated_fourier_series():
    """On uniform nodes (wt = 2 pi / N) the DFT rows are numpy's FFT / N and
    T (c_m-weighted) rebuilds a band-limited signal exactly."""
    N = 128
    t = 2 * np.pi * np.arange(N) / N
    T, Tw = F.bases(t, np.full(N, 2 * np.pi / N), emulate=False)
    rng = np.random.default_rng(0)
    h = rng.standard_normal(N)
    Hh = Tw.T @ h
    f = np.fft.fft(h) / N
    assert np.allclose(Hh[0::2], f[:F.MODES].real, atol=1e-14)
    assert np.allclose(Hh[1::2], f[:F.MODES].imag, atol=1e-14)
    # band-limited signal (modes < 32): G = c_m Hhat, Im G_0 = 0 rebuilds it
    c = rng.standard_normal(F.MODES) + 1j * rng.standard_normal(F.MODES)
    c[0] = c[0].real
    sig = np.real(sum(c[m] * np.exp(1j * m * t) * (1 if m == 0 else 2) for m in range(F.MODES)))
    Hs = Tw.T @ sig
    G = np.empty(F.S)
    G[0::2] = F.CM * Hs[0::2]
    G[1::2] = F.CM * Hs[1::2]
    G[1] = 0.0
    assert np.allclose(T @ G, sig, atol=1e-12)


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 3.140625, 1e-3])
    y = F.bf16(x)
    assert y[0] == 1.0
    assert y[1] == 1.0                    # tie -> even (mantissa ...0)
    assert y[2] == 1.0 + 2 * 2.0 ** -7    # tie -> even (rounds up)
    assert y[3] == -2.5
    assert abs(y[5] - 1e-3) < 1e-3 * 2.0 ** -8
