"""The allocation-free LDPCA entry point (dvc_ldpca_decode_batch_noalloc)
against dvc_ldpca_decode_batch on the same bitplanes: decoded bits and every
statistic equal (save the scheduling-dependent count of speculative
iterations), on two calls in a row (the second smaller, so the shared
scratch holds the first call's contents), through the C ABI with the inputs
and outputs in device memory; and a batch beyond the fixed scratch rejected
on the host."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(ctx, planes):
    from dvc_amd import _lib
    keep = []
    items = (_lib.LdpcaItem * len(planes))()
    for i, p in enumerate(planes):
        lr = _lib.DeviceBuffer.from_array(ctx, np.ascontiguousarray(p["LR"], np.float32))
        acc = _lib.DeviceBuffer.from_array(ctx, np.ascontiguousarray(p["acc"], np.uint8))
        dec = _lib.DeviceBuffer(ctx, p["N"])
        keep.append((lr, acc, dec))
        items[i] = _lib.LdpcaItem(p["N"], lr.ptr, acc.ptr, int(p["crc"]), -1, dec.ptr, _lib.BpStats())
    _lib._check(_lib.lib.dvc_ldpca_decode_batch_noalloc(ctx.h, len(planes), items), "dvc_ldpca_decode_batch_noalloc")
    out = [(keep[i][2].download(p["N"], np.uint8), items[i].stats.as_dict()) for i, p in enumerate(planes)]
    for b in keep:
        for x in b:
            x.free()
    return out


def test_noalloc_equals_batch_on_two_calls(ctx):
    from dvc_amd import workload as W
    planes = W.ldpca_bitplane_set(n_cif=24, n_qcif=16, seed=71)
    for k, subset in enumerate((planes, planes[1::3])):
        print(f"no-alloc call {k}: {len(subset)} bitplanes", flush=True)
        decs, sts = ctx.ldpca_decode_batch([p["LR"] for p in subset], [p["acc"] for p in subset],
                                           [p["crc"] for p in subset])
        got = _noalloc(ctx, subset)
        for i, (p, d, st) in enumerate(zip(subset, decs, sts)):
            gd, gst = got[i]
            assert np.array_equal(gd, d), (k, i)
            # iterations_run counts speculative codes above the accepted one,
            # cancelled when they notice it: scheduling decides how far they
            # got, in either entry point; every other statistic is exact
            ex = lambda s: {kk: v for kk, v in s.items() if kk != "iterations_run"}
            assert ex(gst) == ex(st), (k, i, gst, st)
            assert gst["iterations_run"] >= gst["iterations"], (k, i)
            assert np.array_equal(gd, p["bits"]), (k, i)


def test_noalloc_rejects_batch_beyond_scratch(ctx):
    from dvc_amd import _lib
    n = 200  # 200 CIF bitplanes need more than the fixed 64 MiB
    lr = _lib.DeviceBuffer(ctx, 4 * 6336)
    acc = _lib.DeviceBuffer(ctx, 6336 // 8)
    dec = _lib.DeviceBuffer(ctx, 6336)
    items = (_lib.LdpcaItem * n)()
    for i in range(n):
        items[i] = _lib.LdpcaItem(6336, lr.ptr, acc.ptr, 0, -1, dec.ptr, _lib.BpStats())
    rc = _lib.lib.dvc_ldpca_decode_batch_noalloc(ctx.h, n, items)
    assert rc == _lib.DVC_EINVAL
    assert b"scratch" in _lib.lib.dvc_last_error()
