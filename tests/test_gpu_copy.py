"""xrh_copy_d2d: aligned, unaligned and ragged sizes are copied byte for byte."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbytes,src_off,dst_off", [(0, 0, 0), (1, 0, 0), (15, 0, 0), (16, 0, 0),
                                                    (4096 + 7, 0, 0), (1 << 20, 0, 0), (1000, 3, 5),
                                                    (65536 + 9, 16, 32)])
def test_copy_d2d(ctx, nbytes, src_off, dst_off):
    print(f"case: copy_d2d n={nbytes} src_off={src_off} dst_off={dst_off}", flush=True)
    rng = np.random.default_rng(nbytes)
    total = nbytes + 64
    data = rng.integers(0, 256, total, dtype=np.uint8)
    src = ctx.alloc(total)
    dst = ctx.alloc(total)
    src.upload(data)
    dst.upload(np.zeros(total, np.uint8))
    dst.copy_from(src, nbytes, dst_offset=dst_off, src_offset=src_off)
    got = dst.download(np.uint8)
    want = np.zeros(total, np.uint8)
    want[dst_off:dst_off + nbytes] = data[src_off:src_off + nbytes]
    assert np.array_equal(got, want)
    src.free()
    dst.free()
