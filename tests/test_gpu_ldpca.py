"""LDPCA on the GPU against the oracle: decoded bits, accepted code, entropy,
start code and iteration counts must be identical (ParallelMode 0 semantics),
for the plain sequential path and the speculative batched path."""
import numpy as np
import pytest

import oracle_lib as O
from dvc_amd import workload as W

pytestmark = pytest.mark.gpu


def _check(dec, st, pl, odec, ost, oent):
    assert np.array_equal(dec, odec), "decoded bitplane differs from the oracle"
    assert np.array_equal(dec, pl["bits"]), "decoded bitplane differs from the source"
    assert st["entropy"] == np.float32(oent)
    assert st["start_level"] == ost["start_level"]
    assert st["level"] == ost["level"]
    assert st["rate"] == ost["rate"]
    assert st["iterations"] == ost["total_iters"]
    assert st["used_inverse"] == ost["used_inverse"]


@pytest.mark.parametrize("n", [1584, 6336])
def test_plain_matches_oracle(ctx, n):
    planes = W.ldpca_bitplane_set(n_cif=6 if n == 6336 else 0, n_qcif=10 if n == 1584 else 0, seed=11)
    for pl in planes:
        dec, st = ctx.ldpca_decode(pl["LR"], pl["acc"], pl["crc"])
        odec, ost, oent = O.ldpca_decode(pl["LR"], pl["acc"], pl["crc"])
        print(n, st, ost, flush=True)
        _check(dec, st, pl, odec, ost, oent)


def test_batch_matches_oracle_mixed(ctx):
    planes = W.ldpca_bitplane_set(n_cif=12, n_qcif=12, seed=12)
    decs, sts = ctx.ldpca_decode_batch([p["LR"] for p in planes], [p["acc"] for p in planes],
                                       [p["crc"] for p in planes])
    for pl, dec, st in zip(planes, decs, sts):
        odec, ost, oent = O.ldpca_decode(pl["LR"], pl["acc"], pl["crc"])
        _check(dec, st, pl, odec, ost, oent)


def test_batch_equals_plain(ctx):
    planes = W.ldpca_bitplane_set(n_cif=8, n_qcif=8, seed=13)
    decs, sts = ctx.ldpca_decode_batch([p["LR"] for p in planes], [p["acc"] for p in planes],
                                       [p["crc"] for p in planes])
    for pl, dec, st in zip(planes, decs, sts):
        pdec, pst = ctx.ldpca_decode(pl["LR"], pl["acc"], pl["crc"])
        assert np.array_equal(dec, pdec)
        assert pst["iterations_run"] == pst["iterations"]  # the plain search runs only the codes it tries
        assert st["iterations_run"] >= st["iterations"]
        st.pop("iterations_run"), pst.pop("iterations_run")
        assert st == pst


@pytest.mark.parametrize("n", [1584, 6336])
def test_forced_inverse_on_bad_crc(ctx, n):
    """A wrong CRC makes every BP code fail the CRC test; the search ends at
    the full-rate inverse (LdpcaDecoder.cpp:695-716), which is not
    CRC-checked and still returns the source."""
    pl = W.ldpca_bitplane_set(n_cif=int(n == 6336), n_qcif=int(n == 1584), seed=14)[0]
    dec, st = ctx.ldpca_decode(pl["LR"], pl["acc"], (pl["crc"] + 1) & 0xFF)
    odec, ost, _ = O.ldpca_decode(pl["LR"], pl["acc"], (pl["crc"] + 1) & 0xFF)
    assert st["used_inverse"] == 1 and ost["used_inverse"] == 1
    assert st["level"] == 64 and st["rate"] == n + 8
    assert np.array_equal(dec, pl["bits"]) and np.array_equal(dec, odec)
    assert st["iterations"] == ost["total_iters"]


def test_batched_inverse_many_planes(ctx):
    """Ten bitplanes of each length need the full-rate inverse in one batch
    (more than the inverse kernel's eight plane columns, so columns take
    several), beside two of each that converge: every plane equals the
    oracle and the source."""
    planes = W.ldpca_bitplane_set(n_cif=12, n_qcif=12, seed=16)
    seen = {1584: 0, 6336: 0}
    crcs = []
    for p in planes:
        n = len(p["LR"])
        crcs.append((p["crc"] + (seen[n] < 10)) & 0xFF)
        seen[n] += 1
    decs, sts = ctx.ldpca_decode_batch([p["LR"] for p in planes], [p["acc"] for p in planes], crcs)
    inv = 0
    for pl, crc, dec, st in zip(planes, crcs, decs, sts):
        odec, ost, oent = O.ldpca_decode(pl["LR"], pl["acc"], crc)
        assert st["used_inverse"] == ost["used_inverse"]
        assert (st["level"], st["rate"], st["iterations"]) == (ost["level"], ost["rate"], ost["total_iters"])
        assert np.array_equal(dec, odec) and np.array_equal(dec, pl["bits"])
        inv += st["used_inverse"]
    assert inv == 20


def test_fixed_start_level(ctx):
    pl = W.ldpca_bitplane_set(n_cif=0, n_qcif=1, seed=15)[0]
    dec, st = ctx.ldpca_decode(pl["LR"], pl["acc"], pl["crc"], start_level=0)
    assert st["start_level"] == 0
    assert np.array_equal(dec, pl["bits"])


def test_autotune_and_every_tuning_decode_alike():
    """The autotuner picks valid launch variants; decoding under other
    variants (threads per workgroup, QCIF codewords per CU, poll interval)
    gives the same bits, codes and iteration counts as the default."""
    from dvc_amd import _lib
    planes = W.ldpca_bitplane_set(n_cif=24, n_qcif=24, seed=41)
    args = ([p["LR"] for p in planes], [p["acc"] for p in planes], [p["crc"] for p in planes])
    base = _lib.Context(max_iter=50)
    ref_dec, ref_st = base.ldpca_decode_batch(*args)
    t = base.autotune(64)
    print("tuning", t, flush=True)
    assert t["cif_threads"] in (512, 704, 1024) and t["qcif_threads"] in (256, 512)
    assert 1 <= t["qcif_wg_per_cu"] <= 4 and 1 <= t["poll"] <= 8 and t["cif_ms"] > 0
    for var in [dict(cif_threads=704, qcif_threads=512, qcif_wg_per_cu=2, poll=4),
                dict(cif_threads=512, qcif_threads=256, qcif_wg_per_cu=1, poll=2)]:
        base.set_tuning(**var)
        dec, st = base.ldpca_decode_batch(*args)
        for d0, d1, s0, s1, p in zip(ref_dec, dec, ref_st, st, planes):
            assert np.array_equal(d0, d1) and np.array_equal(d1, p["bits"])
            assert (s0["level"], s0["iterations"]) == (s1["level"], s1["iterations"])
    base.close()


def test_workspace_reused_across_call_sizes(ctx):
    """The context's LDPCA workspace grows to the largest call and is reused:
    a large batch, a small one, a plain call and the large batch again give
    the oracle's results every time, and the repeat equals the first."""
    big = W.ldpca_bitplane_set(n_cif=8, n_qcif=8, seed=17)
    small = W.ldpca_bitplane_set(n_cif=2, n_qcif=1, seed=18)
    args = lambda ps: ([p["LR"] for p in ps], [p["acc"] for p in ps], [p["crc"] for p in ps])
    d1, s1 = ctx.ldpca_decode_batch(*args(big))
    d2, s2 = ctx.ldpca_decode_batch(*args(small))
    for pl, dec, st in zip(small, d2, s2):
        odec, ost, oent = O.ldpca_decode(pl["LR"], pl["acc"], pl["crc"])
        _check(dec, st, pl, odec, ost, oent)
    pl = small[0]
    dec, st = ctx.ldpca_decode(pl["LR"], pl["acc"], pl["crc"])
    assert np.array_equal(dec, d2[0]) and st["level"] == s2[0]["level"]
    d3, s3 = ctx.ldpca_decode_batch(*args(big))
    for pl, a, b, sa, sb in zip(big, d1, d3, s1, s3):
        # iterations_run includes cancelled speculative work, which depends on scheduling
        sa.pop("iterations_run"), sb.pop("iterations_run")
        assert np.array_equal(a, b) and sa == sb
        assert np.array_equal(a, pl["bits"])
