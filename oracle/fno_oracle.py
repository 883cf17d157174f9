"""CPU restatement of the learned preconditioner's network (float64 numpy).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the HIP forward
and backward kernels (ifgf2d_amd/csrc/fno.hip), never by the product. The
network has no reference counterpart (the reference's GMRES takes an identity
preconditioner, ForwardMap/opGMRES.hpp:103,153,251,274).

forward/backward restate the device network exactly: lifting of the 11
features (padded to 32), four Fourier layers P_l = IDFT(mix_l(DFT(H_l))) +
H_l W_l + b_l with GELU between them, projection to the complex multiplier m =
m_phys + O_0 + i O_1, output z = m v, loss ||z - target||^2 / ||target||^2
averaged over the samples. With emulate=True every GEMM operand is rounded to
BF16 (round to nearest even), as the device stages its MFMA operands, and the
spectral weights R are rounded where the mixing reads them; the products and
sums stay in float64. The backward pass is written out by hand; it is pinned
by central finite differences of the float64 loss (tests/test_fno_oracle.py).
"""
import numpy as np
from scipy.special import erf

W, MODES, S, L, CIN, CINPAD, OUTPAD = 64, 32, 64, 4, 11, 32, 16


def layout():
    o = 0
    lay = {}

    def take(name, n):
        nonlocal o
        lay[name] = (o, n)
        o += n
    take("wl", CINPAD * W)
    take("bl", W)
    for l_ in range(L):
        take(f"R{l_}", MODES * W * W * 2)
        take(f"W{l_}", W * W)
        take(f"b{l_}", W)
    take("wp1", W * W)
    take("bp1", W)
    take("wp2", W * OUTPAD)
    take("bp2", OUTPAD)
    return lay, o


LAY, NPARAMS = layout()


def bf16(x):
    """float64 -> float32 -> BF16 (round to nearest even) -> float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)

    def g(name, shape):
        o, n = LAY[name]
        return p[o:o + n].reshape(shape)
    P = {"wl": g("wl", (CINPAD, W)), "bl": g("bl", (W,)), "wp1": g("wp1", (W, W)), "bp1": g("bp1", (W,)),
         "wp2": g("wp2", (W, OUTPAD)), "bp2": g("bp2", (OUTPAD,))}
    for l_ in range(L):
        P[f"R{l_}"] = g(f"R{l_}", (MODES, W, W, 2))
        P[f"W{l_}"] = g(f"W{l_}", (W, W))
        P[f"b{l_}"] = g(f"b{l_}", (W,))
    return P


def pack(G):
    out = np.zeros(NPARAMS)
    for name, (o, n) in LAY.items():
        out[o:o + n] = G[name].ravel()
    return out


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def bases(t, wt, emulate):
    m = np.arange(MODES)
    T = np.empty((t.size, S))
    T[:, 0::2] = np.cos(np.outer(t, m))
    T[:, 1::2] = -np.sin(np.outer(t, m))
    Tw = T * (wt / (2.0 * np.pi))[:, None]
    if emulate:
        # the device evaluates the bases in FP64 and rounds them to BF16
        T, Tw = bf16(T), bf16(Tw)
    return T, Tw


CM = np.where(np.arange(MODES) == 0, 1.0, 2.0)


def _mm(a, b, q):
    return q(a) @ q(b)


def forward_sample(P, X, T, Tw, emulate):
    """Activations of one sample (N nodes)."""
    q = bf16 if emulate else (lambda a: a)
    N = X.shape[0]
    X32 = np.zeros((N, CINPAD))
    X32[:, :CIN] = X
    act = {"X": X32}
    H = _mm(X32, P["wl"], q) + P["bl"]
    act["H"] = [H]
    act["P"], act["Hh"] = [], []
    for l_ in range(L):
        Hh = _mm(Tw.T, H, q)  # (S, W)
        act["Hh"].append(Hh)
        a, b = Hh[0::2], Hh[1::2]
        R = q(P[f"R{l_}"])
        Rr, Ri = R[..., 0], R[..., 1]
        gr = np.einsum("mi,mio->mo", a, Rr) - np.einsum("mi,mio->mo", b, Ri)
        gi = np.einsum("mi,mio->mo", a, Ri) + np.einsum("mi,mio->mo", b, Rr)
        G = np.empty((S, W))
        G[0::2] = CM[:, None] * gr
        G[1::2] = CM[:, None] * gi
        G[1] = 0.0
        Pl = _mm(T, G, q) + _mm(H, P[f"W{l_}"], q) + P[f"b{l_}"]
        act["P"].append(Pl)
        H = gelu(Pl) if l_ < L - 1 else Pl
        act["H"].append(H)
    Qp = _mm(H, P["wp1"], q) + P["bp1"]
    Q = gelu(Qp)
    O = _mm(Q, P["wp2"], q) + P["bp2"]
    act.update(Qp=Qp, Q=Q, O=O)
    return act


def loss_grad(params, samples, emulate=False, grad=True):
    """samples: list of dicts with X (N x 11), mphys, t, wt, v, target.
    Returns (mean loss, flat gradient or None)."""
    P = unpack(params)
    q = bf16 if emulate else (lambda a: a)
    nb = len(samples)
    G = {k: np.zeros_like(v) for k, v in P.items()}
    total = 0.0
    for smp in samples:
        T, Tw = bases(smp["t"], smp["wt"], emulate)
        act = forward_sample(P, np.asarray(smp["X"], dtype=np.float64), T, Tw, emulate)
        O = act["O"]
        v, tg = smp["v"], smp["target"]
        m = np.asarray(smp["mphys"], dtype=np.float64) + O[:, 0] + 1j * O[:, 1]
        r = m * v - tg
        inv = 1.0 / np.sum(np.abs(tg) ** 2)
        total += np.sum(np.abs(r) ** 2) * inv
        if not grad:
            continue
        cr = np.conj(r) * v
        dO = np.zeros_like(O)
        dO[:, 0] = 2.0 * cr.real * inv / nb
        dO[:, 1] = -2.0 * cr.imag * inv / nb
        G["wp2"] += _mm(act["Q"].T, dO, q)
        G["bp2"] += dO.sum(0)
        dQ = _mm(dO, P["wp2"].T, q) * gelu_d(act["Qp"])
        G["wp1"] += _mm(act["H"][L].T, dQ, q)
        G["bp1"] += dQ.sum(0)
        dH = _mm(dQ, P["wp1"].T, q)
        for l_ in range(L - 1, -1, -1):
            dP = dH * gelu_d(act["P"][l_]) if l_ < L - 1 else dH
            G[f"W{l_}"] += _mm(act["H"][l_].T, dP, q)
            G[f"b{l_}"] += dP.sum(0)
            dH = _mm(dP, P[f"W{l_}"].T, q)
            dG = _mm(T.T, dP, q)
            dgr = CM[:, None] * dG[0::2]
            dgi = CM[:, None] * dG[1::2]
            dgi[0] = 0.0
            Hh = act["Hh"][l_]
            a, b = Hh[0::2], Hh[1::2]
            G[f"R{l_}"][..., 0] += np.einsum("mi,mo->mio", a, dgr) + np.einsum("mi,mo->mio", b, dgi)
            G[f"R{l_}"][..., 1] += -np.einsum("mi,mo->mio", b, dgr) + np.einsum("mi,mo->mio", a, dgi)
            R = q(P[f"R{l_}"])
            Rr, Ri = R[..., 0], R[..., 1]
            dHh = np.empty((S, W))
            dHh[0::2] = np.einsum("mio,mo->mi", Rr, dgr) + np.einsum("mio,mo->mi", Ri, dgi)
            dHh[1::2] = -np.einsum("mio,mo->mi", Ri, dgr) + np.einsum("mio,mo->mi", Rr, dgi)
            dH = dH + _mm(Tw, dHh, q)
        G["wl"] += _mm(act["X"].T, dH, q)
        G["bl"] += dH.sum(0)
    return total / nb, (pack(G) if grad else None)


def apply(params, X, mphys, t, wt, v, emulate=False):
    """z = M(v) for one curve."""
    P = unpack(params)
    T, Tw = bases(t, wt, emulate)
    O = forward_sample(P, np.asarray(X, dtype=np.float64), T, Tw, emulate)["O"]
    return (np.asarray(mphys, dtype=np.float64) + O[:, 0] + 1j * O[:, 1]) * v
