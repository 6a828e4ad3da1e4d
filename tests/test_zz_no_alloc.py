"""fs_eval_flux_noalloc: the flux / dFdPh evaluation staged in the library's
fixed device array instead of per-call device allocations.  Checked through
the C ABI against the allocating entry points (fs_eval_flux, fs_eval_dfdph),
bitwise, on consecutive calls of decreasing size (the second call runs over
what the first left in the shared array)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_N = 32768  # 2 MiB of staging / 8 doubles per entry


def _args(seed, n):
    rng = np.random.default_rng(seed)
    E = rng.standard_normal(n) * 10.0 ** rng.uniform(-14, 5, n)
    h = rng.uniform(0.2, 3.0, n)
    q = rng.choice([-1.0, 1.0], n)
    mu = 10.0 ** rng.uniform(-3, 3, n)
    D = 10.0 ** rng.uniform(-2, 4, n)
    nl = 10.0 ** rng.uniform(-4, 8, n)
    nr = 10.0 ** rng.uniform(-4, 8, n)
    return np.ascontiguousarray(np.stack([E, h, q, mu, D, nl, nr], 1))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _call(L, fn, a, *extra):
    out = np.full(a.shape[0], -7.0)
    st = fn(a.shape[0], _dp(a), _dp(out), *extra)
    assert st == 0, L.fs_last_error()
    return out


def _same(a, b):
    # bitwise, NaN and inf included (the same inputs overflow on both paths)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("dfdph", [0, 1])
def test_noalloc_equals_allocating_entry_point(dfdph):
    from fluidsim_amd import lib

    L = lib.load()
    ref_fn = L.fs_eval_dfdph if dfdph else L.fs_eval_flux
    for seed, n in ((11 + dfdph, MAX_N), (21 + dfdph, 1537)):
        a = _args(seed, n)
        ref = _call(L, ref_fn, a)
        got = _call(L, L.fs_eval_flux_noalloc, a, dfdph)
        assert _same(got, ref), (seed, n, int(np.sum(got != ref)))


def test_noalloc_rejects_oversized_input_and_keeps_working():
    from fluidsim_amd import lib

    L = lib.load()
    a = _args(5, MAX_N + 1)
    out = np.zeros(MAX_N + 1)
    assert L.fs_eval_flux_noalloc(MAX_N + 1, _dp(a), _dp(out), 0) == -1  # FS_E_ARG
    assert b"0..32768" in L.fs_last_error()
    assert not out.any()
    small = _args(6, 100)
    assert _same(_call(L, L.fs_eval_flux_noalloc, small, 0), _call(L, L.fs_eval_flux, small))
