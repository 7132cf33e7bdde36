"""spark_pack_noalloc: the pack entry point that allocates no device memory.

It stages the host blocks in a fixed per-device area instead of a buffer it
allocates, and must leave the device state exactly as spark_pack leaves it:
the unpacked state and one RK3 step from it equal spark_pack's, bit for bit,
on two calls in a row (the second with a smaller block set, so stale
contents of the shared area from the first call would show), and a block
larger than the area is refused before any device work.
"""
import numpy as np
import pytest

from sparkhip import capi, problems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    capi.init(0)
    yield


def _after_pack(prob, noalloc):
    """Unpacked state right after the pack, and after one RK3 step."""
    b = capi.Batch(prob.cfg, prob.blocks)
    try:
        if noalloc:
            b.pack_noalloc(prob.arrays)
        else:
            b.pack(prob.arrays)
        packed = [np.full_like(a, np.nan) for a in prob.arrays]
        b.unpack(packed)
        dts = b.advance(1)
        stepped = [np.full_like(a, np.nan) for a in prob.arrays]
        b.unpack(stepped)
    finally:
        b.close()
    return packed, dts, stepped


def _same(xs, ys):
    return len(xs) == len(ys) and all(np.array_equal(x, y, equal_nan=True) for x, y in zip(xs, ys))


def test_pack_noalloc_equals_pack_on_two_calls():
    # first call: 40 blocks of 8^3 / 16^3 / 32^3 with 10 guard cells, several
    # staging rounds (a 32^3 block alone is ~21 MB of unk); second call: one
    # 8^3-interior block, far smaller, reusing the same area
    big = problems.batched(nblocks=40, nguard=10)
    small = problems.batched(nblocks=1, nguard=10, seed_field=7, sizes=(8,), probs=(1.0,))
    for name, prob in (("40 mixed blocks", big), ("1 block", small)):
        print(f"[no_alloc] {name}: {len(prob.blocks)} blocks", flush=True)
        p_ref, d_ref, s_ref = _after_pack(prob, noalloc=False)
        p_na, d_na, s_na = _after_pack(prob, noalloc=True)
        assert _same(p_na, p_ref), name
        assert np.array_equal(d_na, d_ref), name
        assert _same(s_na, s_ref), name
        # the packed state is the host state on every mapped variable
        assert _same([a[..., :capi.SPARK_NPROP] for a in p_na], [a[..., :capi.SPARK_NPROP] for a in prob.arrays])


def test_pack_noalloc_refuses_a_block_larger_than_the_area():
    cfg = problems.base_cfg(ndim=3, nguard=10)
    blk = dict(nx=72, ny=72, nz=72, level=1, dx=1.0 / 72, dy=1.0 / 72, dz=1.0 / 72)
    NZ, NY, NX = problems.gc_shape(cfg, blk)
    arr = np.zeros((NZ, NY, NX, capi.SPARK_NPROP))  # 92^3 x 11 doubles = 68.5 MB > 64 MiB
    b = capi.Batch(cfg, [blk])
    try:
        with pytest.raises(capi.SparkError, match="staging area"):
            b.pack_noalloc([arr])
    finally:
        b.close()
