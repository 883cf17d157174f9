"""The learned preconditioner's HIP kernels (ifgf2d_amd/csrc/fno.hip) through
the C ABI:

* the BF16 MFMA GEMM against numpy on BF16-rounded operands;
* the per-node features against their definition (plane-wave tangential
  wavenumber d.tau, curvature of a circle, normalised geometry);
* the loss and its gradient against the float64 restatement
  (oracle/fno_oracle.py, with the device's BF16 operand rounding emulated);
* training on densities the solver produced for random stars lowers the loss,
  and on held-out stars every learned mode reaches the FP64 true residual
  1e-8 (iterations and wall time against plain GMRES are printed);
* a deliberately poor network as the flexible preconditioner stagnates, the
  solve falls back to plain GMRES and still converges (all kernels exit
  normally).
"""
import time

import numpy as np
import pytest

import fno_oracle as FO
import ifgf2d_amd as I
from ifgf2d_amd import configs

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("M,N,K,ta,tb", [(64, 64, 32, 0, 0), (200, 64, 64, 0, 0), (1000, 16, 64, 0, 1),
                                         (64, 64, 5000, 1, 0), (32, 64, 777, 1, 0), (130, 70, 45, 0, 1),
                                         (3001, 64, 16, 0, 1)])
def test_mfma_gemm_matches_bf16_numpy(M, N, K, ta, tb):
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    C = I.fno_gemm_check(A, B, bool(ta), bool(tb))
    Ar, Br = FO.bf16(A), FO.bf16(B)
    ref = (Ar.T if ta else Ar) @ (Br.T if tb else Br)
    e = np.abs(C - ref).max() / np.abs(ref).max()
    print(f"gemm {M}x{N}x{K} ta={ta} tb={tb}: max err {e:.1e}", flush=True)
    assert e < 1e-5


def _star_map(k=60.0, npatch=300, seed=0, **kw):
    rng = np.random.default_rng(seed)
    a = [0.0, 0.0] + list(rng.uniform(-0.05, 0.05, 7))
    b = [0.0, 0.0] + list(rng.uniform(-0.05, 0.05, 7))
    R = rng.uniform(0.5, 1.5)
    curve = I.star(R=R, a=a, b=b)
    nlev = max(3, int(np.round(np.log2(npatch * 10 / 100))) + 1)
    return I.ForwardMap(curve, I.equal_parameter_lims(npatch), k, nlev, p=6.0, device_setup=True, **kw)


def test_features_match_their_definition():
    fno = I.LearnedPreconditioner(seed=1)
    # circle of radius 1.3: curvature 1 / 1.3 (feature 4 = kappa Rs / 4 = 1 / 4)
    fc = I.ForwardMap(I.circle(0.2, -0.1, 1.3), I.equal_parameter_lims(200), 30.0, 5, p=6.0)
    xc, yc, nxc, nyc, _ = fc.points()
    d = (np.cos(0.7), np.sin(0.7))
    v = -configs.plane_wave(xc, yc, 30.0, d)
    fno.add_sample(fc, v, v)
    s = fno.sample_data(0)
    X = s["X"].astype(np.float64)
    assert np.abs(X[:, 4] - 0.25).max() < 1e-4
    assert np.abs(X[:, 2] - nxc).max() < 1e-6 and np.abs(X[:, 3] - nyc).max() < 1e-6
    assert np.abs(X[:, 0] - (xc - 0.2) / 1.3).max() < 1e-5
    # plane wave: |v| / rms = 1; Re w = d.tau / 1.01 (tau = (-n_y, n_x)), Im w = 0
    dtau = -d[0] * nyc + d[1] * nxc
    print("w error", np.abs(X[:, 8] - dtau / 1.01).max(), np.abs(X[:, 9]).max(), flush=True)
    assert np.abs(X[:, 7] - 1).max() < 1e-5
    assert np.abs(X[:, 8] - dtau / 1.01).max() < 2e-3 and np.abs(X[:, 9]).max() < 2e-3
    # the fitted direction: e = d / 1.01, so e.n = d.n / 1.01 (lit side negative)
    dn = d[0] * nxc + d[1] * nyc
    assert np.abs(X[:, 10] - dn / 1.01).max() < 2e-3
    c = np.sqrt(np.maximum(0, 1 - X[:, 8] ** 2))
    assert np.abs(s["mphys"] - 2 * c / (1 + c)).max() < 1e-5
    # arclength fraction increases along the curve from 0 to 1
    assert np.all(np.diff(X[:, 5]) > 0) and X[0, 5] > 0 and X[-1, 5] < 1


def _samples(fno, maps_dirs):
    out = []
    for fm, th in maps_dirs:
        x, y, *_ = fm.points()
        b = -configs.plane_wave(x, y, fm.k, (np.cos(th), np.sin(th)))
        dens, it, rr, _ = fm.solve(b, restart=50, tol=1e-10, maxit=500)
        sid = fno.add_sample(fm, b, dens)
        out.append((sid, b, dens))
    return out


def test_loss_and_gradient_match_restatement():
    fno = I.LearnedPreconditioner(seed=3)
    # perturb the weights away from the small-output initialisation so every block matters
    p = fno.get_params()
    rng = np.random.default_rng(0)
    for name, (o, n) in FO.LAY.items():
        if name.startswith("wp2") or name.startswith("b"):
            p[o:o + n] += (0.05 * rng.standard_normal(n)).astype(np.float32)
    fno.set_params(p)
    fa, fb = _star_map(seed=1, npatch=120), _star_map(seed=2, npatch=170)
    smp = _samples(fno, [(fa, 0.3), (fb, 2.1)])
    loss, g = fno.loss_grad([s[0] for s in smp])
    ref_s = []
    for sid, b, dens in smp:
        d = fno.sample_data(sid)
        ref_s.append(dict(d, v=b, target=dens))
    lr_, gr = FO.loss_grad(fno.get_params().astype(np.float64), ref_s, emulate=True)
    print(f"loss GPU {loss:.6e} restatement {lr_:.6e}", flush=True)
    assert abs(loss - lr_) < 1e-4 * lr_
    worst = 0.0
    for name, (o, n) in FO.LAY.items():
        ref = gr[o:o + n]
        if np.linalg.norm(ref) == 0:
            assert np.abs(g[o:o + n]).max() == 0, name
            continue
        e = np.linalg.norm(g[o:o + n] - ref) / np.linalg.norm(ref)
        worst = max(worst, e)
        print(f"  grad {name}: rel diff {e:.2e}", flush=True)
        assert e < 1e-2, name
    cos = np.dot(g, gr) / (np.linalg.norm(g) * np.linalg.norm(gr))
    print(f"gradient worst block {worst:.2e}, cosine {cos:.8f}", flush=True)
    assert cos > 0.9999
    # the device's inference equals the restatement's forward
    z = fno.apply(fa, smp[0][1])
    d0 = fno.sample_data(smp[0][0])
    zr = FO.apply(fno.get_params().astype(np.float64), d0["X"], d0["mphys"], d0["t"], d0["wt"], smp[0][1],
                  emulate=True)
    assert rel(z, zr) < 1e-3


def _random_star_plan(rng, k, nodes_per_wavelength=10):
    R = rng.uniform(0.5, 1.5)
    a = [0.0, 0.0] + list(rng.uniform(-0.05, 0.05, 7))
    b = [0.0, 0.0] + list(rng.uniform(-0.05, 0.05, 7))
    npatch = int(np.ceil(1.1 * R * k * 10 / nodes_per_wavelength))
    nlev = max(3, int(np.round(np.log2(npatch * 10 / 100))) + 1)
    return I.ForwardMap(I.star(R=R, a=a, b=b), I.equal_parameter_lims(npatch), k, nlev, p=6.0, device_setup=True)


def test_training_on_solver_densities_and_held_out_solves():
    """Training on (right-hand side, density) pairs of random stars lowers the
    loss severalfold; on held-out stars the trained network's initial density
    has a residual well below the zero guess's, and every learned mode reaches
    the FP64 true residual 1e-8. The iteration counts and wall times of plain
    GMRES and the learned modes are printed: at this training scale the
    learned initial density costs a few more GMRES iterations than it saves
    (its residual is rougher than the right-hand side), and the learned
    multiplier as a flexible preconditioner stagnates and falls back."""
    rng = np.random.default_rng(11)
    fno = I.LearnedPreconditioner(seed=5)
    train = []
    t0 = time.time()
    for o in range(24):
        fm = _random_star_plan(rng, rng.uniform(50, 150))
        x, y, *_ = fm.points()
        th = rng.uniform(0, 2 * np.pi, 4)
        B = np.stack([-configs.plane_wave(x, y, fm.k, (np.cos(t), np.sin(t))) for t in th], axis=1)
        for j in range(4):
            d, it, rr, _ = fm.solve(B[:, j], restart=50, tol=1e-10, maxit=500)
            train.append(fno.add_sample(fm, B[:, j], d))
    print(f"{len(train)} training samples in {time.time() - t0:.1f} s", flush=True)
    l0, _ = fno.loss_grad(train, grad=False)
    t0 = time.time()
    hist = fno.train(train, epochs=40, batch=8, lr=1e-3, seed=1)
    print(f"training: loss {l0:.3f} -> {hist[-1]:.3f} in {time.time() - t0:.1f} s; "
          f"epochs {np.round(hist[::8], 3).tolist()}", flush=True)
    assert hist[-1] < l0 / 3
    tot = {"plain": [0, 0.0], "guess": [0, 0.0], "guess+adaptive": [0, 0.0]}
    for o in range(6):
        fm = _random_star_plan(rng, rng.uniform(50, 150))
        x, y, *_ = fm.points()
        th = rng.uniform(0, 2 * np.pi)
        b = -configs.plane_wave(x, y, fm.k, (np.cos(th), np.sin(th)))
        x0 = fno.apply(fm, b)
        r_guess = np.linalg.norm(b - fm.compute_Ax_acc(x0)) / np.linalg.norm(b)
        assert r_guess < 0.5
        runs = {"plain": lambda: fm.solve(b, restart=50, tol=1e-8, maxit=500),
                "guess": lambda: fm.solve_learned(b, fno, restart=50, tol=1e-8, maxit=500),
                "guess+adaptive": lambda: fm.solve_learned(b, fno, restart=50, tol=1e-8, maxit=500,
                                                           precondition=True)}
        line = []
        for name, fn in runs.items():
            fn()  # warm
            t1 = time.perf_counter()
            out = fn()
            dt = time.perf_counter() - t1
            it, rr = (out[1], out[2]) if name == "plain" else (out[1], out[3])
            assert rr < 1e-8, name
            tot[name][0] += it
            tot[name][1] += dt
            line.append(f"{name} {it} it {dt * 1e3:.1f} ms" + (f" (fallback {out[2]})" if name != "plain" else ""))
        print(f"held-out N={fm.N} k={fm.k:.0f}, initial density relres {r_guess:.2f}: " + ", ".join(line), flush=True)
    print("held-out totals: " + ", ".join(f"{k_} {v[0]} it {v[1] * 1e3:.1f} ms" for k_, v in tot.items()), flush=True)


def test_poor_network_falls_back_to_plain_gmres():
    fno = I.LearnedPreconditioner(seed=9)
    p = fno.get_params()
    o, n = FO.LAY["bp2"]
    p[o] = -40.0  # multiplier m_phys - 40: a preconditioner that amplifies and flips
    o, n = FO.LAY["wp2"]
    p[o:o + n] = np.random.default_rng(0).standard_normal(n).astype(np.float32) * 50.0
    fno.set_params(p)
    fm = _star_map(seed=4, npatch=200)
    x, y, *_ = fm.points()
    b = -configs.plane_wave(x, y, fm.k, (1.0, 0.0))
    d0, i0, r0, _ = fm.solve(b, restart=50, tol=1e-8, maxit=500)
    d, it, fb, rr, hist = fm.solve_learned(b, fno, restart=30, tol=1e-8, maxit=500, guess=False, precondition=True,
                                           window=4, ratio=0.5)
    print(f"poor network: {it} iterations, fallback {fb}, true relres {rr:.1e} (plain GMRES {i0})", flush=True)
    assert fb and rr < 1e-8 and rel(d, d0) < 1e-6
