"""cim_local_box_stack_noalloc (working memory in the library's static
scratch array, no device allocation) against cim_local_box_stack_dev on the
same inputs: identical boxes and records on two calls in a row, the second
smaller, and the same refusals of inputs outside the image."""
import numpy as np
import pytest

from cryoim_amd import lib, synth

pytestmark = pytest.mark.gpu


def _case(box, n, seed, nops_sym):
    rng = np.random.default_rng(seed)
    ph = synth.make_phantom(box, radius=(box * 0.3, box * 0.4), sigma=(box / 64, box / 32), seed=seed)
    ps = synth.make_particles(ph, n, seed=seed + 1, shift_max=3.0, snr=1.0, ctf=True, apix=1.5)
    P = np.zeros((n, lib.PART_STRIDE), np.float32)
    P[:, 0:3] = ps.angles
    P[:, 3:5] = ps.shifts
    ops = lib.local_box_sym_ops(nops_sym)
    x0 = (rng.uniform(-5, 5), rng.uniform(-5, 5), box * 0.35)
    return ps.images, P, ops, x0


def _both(ctx, images, P, ops, x0, sub, box):
    from devmem import DevArray
    n, nops = len(images), len(ops)
    dimg = DevArray.from_numpy(images)
    d1, d2 = DevArray(n * nops * sub * sub), DevArray(n * nops * sub * sub)
    try:
        _, r1 = ctx.local_box_stack(dimg, P, ops, x0, sub, 1.5, boxes=d1, box=box)
        _, r2 = ctx.local_box_stack_noalloc(dimg, P, ops, x0, sub, 1.5, boxes=d2, box=box)
        return d1.numpy(), r1, d2.numpy(), r2
    finally:
        for d in (dimg, d1, d2):
            d.free()


def test_noalloc_stack_equals_device_stack(ctx):
    # box 384, 12 operators: 70 particles take three passes through the 64 MiB array
    for box, n, sub, sym, seed in ((384, 70, 96, "i5", 5), (128, 5, 48, "c5", 6)):
        images, P, ops, x0 = _case(box, n, seed, sym)
        print(f"noalloc local_box stack: box {box}, {n} particles, {len(ops)} operators, sub {sub}", flush=True)
        b1, r1, b2, r2 = _both(ctx, images, P, ops, x0, sub, box)
        assert np.abs(b1).max() > 0
        assert np.array_equal(b1, b2)
        assert np.array_equal(r1, r2)


def test_noalloc_stack_refuses_what_the_device_stack_refuses(ctx):
    from devmem import DevArray
    images, P, ops, _ = _case(64, 2, 7, "c5")
    dimg, dbox = DevArray.from_numpy(images), DevArray(2 * len(ops) * 24 * 24)
    try:
        with pytest.raises(lib.CimError):
            ctx.local_box_stack_noalloc(dimg, P, ops, (0.0, 0.0, 63.0), 24, 1.5, boxes=dbox, box=64)
        with pytest.raises(lib.CimError):
            ctx.local_box_stack_noalloc(dimg, P, ops, (0.0, 0.0, 5.0), 80, 1.5, boxes=DevArray(2 * len(ops) * 80 * 80),
                                        box=64)
    finally:
        dimg.free()
        dbox.free()
