"""The no-allocation entry point, mbgpu_ring_spring_dev_noalloc, against the entry it varies,
mbgpu_ring_spring_dev, through the C ABI.

* Two calls in a row, the second with a smaller input (fewer beads and atoms): forces,
  spring energy and centroid-virial term bit-identical to the allocating entry's, whose
  kernels it shares.
* It allocates nothing: the library's allocation counter does not move across its calls on
  a fresh context (the allocating entry's first call does allocate).
* A call that needs more scratch than the fixed array holds is refused before any launch.
"""
import ctypes
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def alloc_count():
    mb = importlib.import_module("dlpoly2-mbpol-heat_amd.mbgpu")
    fn = mb.load_library().mbgpu_debug_alloc_serial
    fn.restype = ctypes.c_long
    return fn()


def case(rng, nb, n):
    mass = rng.uniform(1.0, 16.0, n)
    x = rng.normal(size=(3, n))[None] + 0.05 * rng.normal(size=(nb, 3, n))
    f = rng.normal(size=(nb, 3, n))
    xc = x.mean(0)
    return mass, x, f, xc


def test_noalloc_equals_the_allocating_entry(pkg):
    import torch
    rng = np.random.default_rng(23)
    dev = dict(dtype=torch.float64, device="cuda")
    ctx_a = pkg.Context(8, 32)            # allocating entry
    ctx_b = pkg.Context(8, 32)            # no-allocation entry, fresh context
    for nb, n in ((16, 5000), (4, 700)):   # the second call smaller
        mass, x, f, xc = case(rng, nb, n)
        md, xd, xcd = (torch.as_tensor(np.ascontiguousarray(a), **dev) for a in (mass, x, xc))
        fa, fb = (torch.as_tensor(f.copy(), **dev) for _ in range(2))
        torch.cuda.synchronize()
        print(f"ring springs, {nb} beads x {n} atoms, both entries", flush=True)
        before = alloc_count()
        eb = ctx_b.ring_spring_dev_noalloc(nb, n, 37.5, 0.75, 2.0, md.data_ptr(), xd.data_ptr(), xcd.data_ptr(),
                                           fb.data_ptr())
        assert alloc_count() == before
        ea = ctx_a.ring_spring_dev(nb, n, 37.5, 0.75, 2.0, md.data_ptr(), xd.data_ptr(), xcd.data_ptr(),
                                   fa.data_ptr())
        assert eb == ea
        assert torch.equal(fa, fb)
        assert not torch.equal(fb, torch.as_tensor(f, **dev))          # the springs did act


def test_noalloc_refuses_more_than_its_scratch(pkg):
    import torch
    ctx = pkg.Context(8, 32)
    buf = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = buf.data_ptr()
    torch.cuda.synchronize()
    # 2 * 64 beads * 300 blocks of 256 atoms + 2 > 32768: refused on the host, nothing launched
    with pytest.raises(pkg.MbgpuError, match="scratch"):
        ctx.ring_spring_dev_noalloc(64, 256 * 300, 1.0, 1.0, 1.0, p, p, p, p)
