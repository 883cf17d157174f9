"""ifgf_propagate_noalloc (the field at off-surface points through the
library's fixed device scratch, allocating nothing) against ifgf_propagate,
through the C ABI: two calls in a row on one device, the second with a smaller
plan and fewer targets, so the scratch still holds the first call's longer
density when the second runs."""
import ctypes

import numpy as np
import pytest

import ifgf2d_amd as I

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _call(fm, tx, ty, dens):
    import torch
    dev = torch.device("cuda:0")
    txd = torch.from_numpy(tx).to(dev)
    tyd = torch.from_numpy(ty).to(dev)
    dd = torch.from_numpy(dens).to(dev)
    out = torch.full((tx.size,), np.nan + 0j, dtype=torch.complex128, device=dev)
    torch.cuda.synchronize()
    fm.propagate_noalloc_device(txd.data_ptr(), tyd.data_ptr(), tx.size, dd.data_ptr(), out.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_field_matches_allocating_entry_point():
    rng = np.random.default_rng(4)
    cases = [
        # (map kwargs, targets): the second smaller in both points and targets
        (dict(curve=I.kite(), patch_lims=I.equal_parameter_lims(400), k=40.0, nlevels=6, p=6.0), 300),
        (dict(curve=I.circle(), patch_lims=I.equal_parameter_lims(40), k=10.0, nlevels=3, p=6.0), 64),
    ]
    for kw, nt in cases:
        fm = I.ForwardMap(**kw)
        th = 2 * np.pi * rng.random(nt)
        r = 2.5 + rng.random(nt)
        tx, ty = r * np.cos(th), r * np.sin(th)
        dens = rng.standard_normal(fm.N) + 1j * rng.standard_normal(fm.N)
        print(f"noalloc field: N={fm.N} targets={nt}", flush=True)
        ref = fm.propagate_unacc(tx, ty, dens)
        got = _call(fm, tx, ty, dens)
        e = rel(got, ref)
        print(f"  rel diff vs ifgf_propagate {e:.2e}", flush=True)
        assert np.all(np.isfinite(got)) and e < 1e-12
        # the same call again gives the same bits (the scratch is rewritten)
        assert np.array_equal(_call(fm, tx, ty, dens), got)


def test_noalloc_rejects_bad_arguments():
    fm = I.ForwardMap(curve=I.circle(), patch_lims=I.equal_parameter_lims(40), k=10.0, nlevels=3, p=6.0)
    rc = I.lib().ifgf_propagate_noalloc(fm._h, None, None, -1, None, None, None)
    assert rc == 1
    host = I.ForwardMap(curve=I.circle(), patch_lims=I.equal_parameter_lims(40), k=10.0, nlevels=3, p=6.0,
                        host_only=True)
    d = ctypes.c_void_p(1)
    assert I.lib().ifgf_propagate_noalloc(host._h, d, d, 4, d, d, None) == 5
