"""Allocation-free FFSL sweep (sisl_ffsl_1d_f64_noalloc) against the
caller-scratch entry point sisl_ffsl_1d_f64, through the C ABI.

The two run the same kernel on the same inputs, so the bar is bit equality:
two calls in a row (the second on a smaller input, so stale scratch from the
first would show), closed and periodic lines, every sweep direction's
memory layout.  Inputs beyond the fixed scratch are rejected on the host
before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from suetes_amd import _lib

pytestmark = pytest.mark.gpu
DEV = "cuda"
MAX_FACES = 1 << 23


def _inputs(outer, n, inner, periodic, seed):
    rng = np.random.default_rng(seed)
    mass = torch.from_numpy(rng.uniform(0.5, 2.0, (outer, n, inner))).to(DEV)
    cfl = torch.from_numpy(rng.uniform(-1.7, 1.7, (outer, n + 1, inner))).to(DEV)
    if periodic:
        cfl[:, n] = cfl[:, 0]
    else:
        cfl[:, 0] = 0.0
        cfl[:, n] = 0.0
    return mass.contiguous(), cfl.contiguous()


def _with_scratch(mass, cfl, outer, n, inner, periodic):
    L = _lib.lib()
    out = torch.empty_like(mass)
    scr = torch.empty(outer * (n + 1) * inner, dtype=torch.float64, device=DEV)
    _lib.check(L.sisl_ffsl_1d_f64(mass.data_ptr(), cfl.data_ptr(), out.data_ptr(), scr.data_ptr(), outer, n,
                                  inner, int(periodic), _lib.stream_ptr()), "sisl_ffsl_1d_f64")
    return out


def _no_alloc(mass, cfl, outer, n, inner, periodic):
    L = _lib.lib()
    out = torch.empty_like(mass)
    _lib.check(L.sisl_ffsl_1d_f64_noalloc(mass.data_ptr(), cfl.data_ptr(), out.data_ptr(), outer, n, inner,
                                          int(periodic), _lib.stream_ptr()), "sisl_ffsl_1d_f64_noalloc")
    return out


@pytest.mark.parametrize("periodic", [False, True])
def test_no_alloc_sweep_matches_scratch_entry_point(periodic):
    # first call: a z sweep of a 64 x 48 x 40 field (contiguous lines across threads)
    # second call: a smaller x sweep (inner = 1), so its lines overlap the first call's scratch
    shapes = [(1, 40, 64 * 48), (9 * 7, 23, 1)]
    for call, (outer, n, inner) in enumerate(shapes):
        print(f"[case] no-alloc ffsl call {call}: outer={outer} n={n} inner={inner} periodic={periodic}",
              flush=True)
        mass, cfl = _inputs(outer, n, inner, periodic, seed=10 * call + int(periodic))
        a = _no_alloc(mass, cfl, outer, n, inner, periodic)
        b = _with_scratch(mass, cfl, outer, n, inner, periodic)
        torch.cuda.synchronize()
        assert torch.equal(a, b), float((a - b).abs().max())
        tot = float(mass.sum())
        assert abs(float(a.sum()) - tot) <= 1e-12 * tot    # mass conserved on every line


def test_no_alloc_sweep_rejects_inputs_beyond_its_scratch():
    L = _lib.lib()
    mass = torch.ones(4, dtype=torch.float64, device=DEV)
    # (n + 1) * inner just above the fixed scratch; checked before any launch
    n, inner = 1023, MAX_FACES // 1024 + 1
    rc = L.sisl_ffsl_1d_f64_noalloc(mass.data_ptr(), mass.data_ptr(), mass.data_ptr(), 1, n, inner, 0,
                                    _lib.stream_ptr())
    assert rc == -4 and "fixed scratch" in _lib.ERRORS[rc]
    # an input that fills the scratch to within one line is accepted (8191 lines of 1023 cells)
    n_ok, inner_ok = 1023, MAX_FACES // 1024 - 1
    m = torch.ones(n_ok * inner_ok, dtype=torch.float64, device=DEV)
    c = torch.zeros((n_ok + 1) * inner_ok, dtype=torch.float64, device=DEV)
    out = torch.empty_like(m)
    rc = L.sisl_ffsl_1d_f64_noalloc(m.data_ptr(), c.data_ptr(), out.data_ptr(), 1, n_ok, inner_ok, 0,
                                    _lib.stream_ptr())
    assert rc == 0
    torch.cuda.synchronize()
    assert torch.equal(out, m)                 # zero Courant numbers: the field is unchanged
