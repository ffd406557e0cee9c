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


def test_buffer_ranges_are_checked_before_any_launch(ctx):
    """A DRR, moving-image or copy range that starts inside one of the context's allocations
    and runs past its end is refused with XRH_ERR_INVALID before anything reaches the GPU;
    in-range offsets work."""
    import types

    from xreg_amd import _xrh as X
    from xreg_amd import synth

    print("case: range checks", flush=True)
    w = synth.config("c1", n_poses=2, small=True)
    cam = w.cams[0]
    npx = cam.num_det_rows * cam.num_det_cols
    vol = X.Volume(ctx, w.vol, w.idx_to_phys)
    buf = ctx.alloc(4 * npx * 2)  # room for two DRRs
    poses = w.poses.astype(np.float32)
    with pytest.raises(X.XrhError) as e:
        X.line_int(ctx, vol, [cam], poses[:2], np.zeros(2, np.int64), w.step,
                   types.SimpleNamespace(ptr=buf.ptr + 4 * npx))  # second half holds only one
    assert e.value.status == X.XRH_ERR_INVALID
    X.line_int(ctx, vol, [cam], poses[:1], np.zeros(1, np.int64), w.step, types.SimpleNamespace(ptr=buf.ptr + 4 * npx))
    fx = X.Fixed(ctx, np.ones((cam.num_det_rows, cam.num_det_cols), np.float32), X.XRH_SIM_NCC)
    with pytest.raises(X.XrhError):
        fx.compute(buf, 2, offset_imgs=1)
    assert fx.compute(buf, 1, offset_imgs=1).shape == (1,)
    with pytest.raises(X.XrhError):
        buf.download(np.float32, count=npx + 1, offset_bytes=4 * npx)
    fx.close()
    buf.free()
    vol.close()


def test_context_hbm_budget(ctx):
    """Everything a context allocates is charged to its budget: an allocation or a volume that
    would pass it fails with XRH_ERR_BUDGET and leaves the charge unchanged; frees return it."""
    from xreg_amd import _xrh as X
    from xreg_amd import synth

    print("case: hbm budget", flush=True)
    c = X.Context(0, hbm_budget_bytes=64 << 20)
    try:
        base = c.bytes_in_use()
        a = c.alloc(32 << 20)
        assert c.bytes_in_use() == base + (32 << 20)
        with pytest.raises(X.XrhError) as e:
            c.alloc(48 << 20)
        assert e.value.status == X.XRH_ERR_BUDGET
        assert c.bytes_in_use() == base + (32 << 20)
        w = synth.config("c2", n_poses=1)  # 256x256x200: a 215 MB layout
        with pytest.raises(X.XrhError) as e:
            X.Volume(c, w.vol, w.idx_to_phys)
        assert e.value.status == X.XRH_ERR_BUDGET
        assert c.bytes_in_use() == base + (32 << 20)
        a.free()
        assert c.bytes_in_use() == base
        small = synth.config("c1", n_poses=1, small=True)
        v = X.Volume(c, small.vol, small.idx_to_phys)
        assert 0 < v.nbytes <= c.bytes_in_use() - base
        v.close()
        c.synchronize()
        assert c.bytes_in_use() == base
    finally:
        c.close()
