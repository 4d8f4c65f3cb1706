"""The no-allocation lower bound (rumgpu_lb64_noalloc) against rumgpu_lb64, through the C ABI.

rumgpu_lb64_noalloc takes its rows, y and p in device memory and keeps every buffer in one fixed
device array of the library. It runs the same kernels in the same order as rumgpu_lb64, so the two
lower bounds agree to the last bit up to the final square root, which runs on the device in one and
on the host in the other: the test allows 1 ulp (relative 2^-52). Two calls in a row, the second
on a smaller lattice with fewer rows, check that the array's leftover contents from the first call
do not leak into the second.
"""
import ctypes

import numpy as np
import pytest
import torch

from solver import _lib
from solver.bm_admm.engine import Handle

pytestmark = pytest.mark.gpu

ULP = 2.0 ** -52


def _rows(n, cells, rng):
    seen, om, ox = set(), [], []
    while len(om) < cells:
        m = int(rng.integers(0, 1 << n))
        if bin(m).count("1") < 2 or m in seen:
            continue
        seen.add(m)
        for a in range(n):
            if (m >> a) & 1:
                om.append(m)
                ox.append(a)
    menu, alt = np.array(om, dtype=np.uint64), np.array(ox, dtype=np.int32)
    p = rng.random(menu.size)
    _, inv = np.unique(menu, return_inverse=True)
    p = p / np.bincount(inv, weights=p)[inv]
    # y: an ADMM iterate near the projection of p, where the bound is positive (random p on many
    # menus lies far from the RUM polytope)
    h = Handle(n, menu, alt, p, dtype="float64", beta=1.0)
    h.run(200)
    y = np.ascontiguousarray(h.x_obs())
    h.close()
    return menu, alt, y, p


def _lb_host(ctx, n, menu, alt, y, p):
    out = ctypes.c_double()
    _lib.check(ctx.lib.rumgpu_lb64(ctx.h, n, _lib.ptr(menu), _lib.ptr(alt), menu.size, _lib.ptr(y), _lib.ptr(p),
                                   ctypes.byref(out)))
    return out.value


def _lb_noalloc(ctx, n, menu, alt, y, p):
    dev = torch.device("cuda", ctx.device)
    d_menu = torch.from_numpy(menu.view(np.int64).copy()).to(dev)
    d_alt = torch.from_numpy(alt.copy()).to(dev)
    d_y = torch.from_numpy(y.copy()).to(dev)
    d_p = torch.from_numpy(p.copy()).to(dev)
    d_lb = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
    torch.cuda.synchronize(dev)
    st = ctx.lib.rumgpu_lb64_noalloc(ctx.h, n, ctypes.c_void_p(d_menu.data_ptr()), ctypes.c_void_p(d_alt.data_ptr()),
                                     menu.size, ctypes.c_void_p(d_y.data_ptr()), ctypes.c_void_p(d_p.data_ptr()),
                                     ctypes.c_void_p(d_lb.data_ptr()))
    _lib.check(st, "lb64_noalloc")
    return float(d_lb.cpu().item())


def test_noalloc_lb_equals_allocating_lb_twice_smaller_second():
    ctx = _lib.context(0)
    rng = np.random.default_rng(7)
    for n, cells in ((16, 20000), (11, 900)):            # second call: smaller lattice, fewer rows
        menu, alt, y, p = _rows(n, cells, rng)
        print("[case] lb64_noalloc n=%d M=%d" % (n, menu.size), flush=True)
        ref = _lb_host(ctx, n, menu, alt, y, p)
        got = _lb_noalloc(ctx, n, menu, alt, y, p)
        assert ref > 0.01
        assert abs(got - ref) <= ULP * ref, (n, got, ref)


def test_noalloc_lb_at_the_maximum_n_and_repeat():
    ctx = _lib.context(0)
    rng = np.random.default_rng(8)
    n = 18
    menu, alt, y, p = _rows(n, 30000, rng)
    print("[case] lb64_noalloc n=%d M=%d" % (n, menu.size), flush=True)
    ref = _lb_host(ctx, n, menu, alt, y, p)
    first = _lb_noalloc(ctx, n, menu, alt, y, p)
    assert ref > 0.01
    assert abs(first - ref) <= ULP * ref, (first, ref)
    small = _rows(9, 200, rng)
    print("[case] lb64_noalloc n=9 M=%d" % small[0].size, flush=True)
    assert abs(_lb_noalloc(ctx, 9, *small) - _lb_host(ctx, 9, *small)) <= ULP * _lb_host(ctx, 9, *small)
    assert _lb_noalloc(ctx, n, menu, alt, y, p) == first          # same input again: same bits


def test_noalloc_lb_rejects_sizes_and_bad_rows():
    ctx = _lib.context(0)
    rng = np.random.default_rng(9)
    menu, alt, y, p = _rows(19, 100, rng)
    with pytest.raises(_lib.GPUError, match="outside 2..18"):
        _lb_noalloc(ctx, 19, menu, alt, y, p)
    menu, alt, y, p = _rows(8, 100, rng)
    bad = alt.copy()
    j = next(j for j in range(menu.size) if bin(int(menu[j])).count("1") < 8)
    bad[j] = next(a for a in range(8) if not (int(menu[j]) >> a) & 1)   # an alternative outside its menu
    assert np.isnan(_lb_noalloc(ctx, 8, menu, bad, y, p))
    # the flag does not stick: the next good call is exact again
    ref = _lb_host(ctx, 8, menu, alt, y, p)
    assert abs(_lb_noalloc(ctx, 8, menu, alt, y, p) - ref) <= ULP * ref
