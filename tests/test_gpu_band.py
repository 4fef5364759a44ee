"""The per-band entry points (dvc_band_frame_init, dvc_noise_distribution,
dvc_cond_bit_prob, dvc_update_si, dvc_update_smf, dvc_forward_me) through the
C ABI against the oracle's restatements of the CPU functions they replace
(noiseModel.cpp:77-132, 191-310; sideInfoCreation.cpp:61-145, 598-736;
motionLearning.cpp:70-127), on the reference's data layouts (SMF, SearchInfo,
Trans, MV of global.h).  Inputs are a seeded sequence's key frames, the
oracle's side information, noise parameters, SMF and over-complete
transform, and the decoded-so-far indices of the WZ frame's true bitplanes;
every output must be bit-exact."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(176, 144, 4, 1200), (352, 288, 8, 1201)]


def _setup(w, h, q, seed):
    import oracle_lib as O
    from dvc_amd import workload as W
    g = O.geom(w, h)
    vid = W.synthetic_video(w, h, 3, 3 if w == 176 else 6, seed)
    past, cur, fut = [np.ascontiguousarray(v) for v in vid]
    lib = O.orc()
    plp, flp = np.zeros_like(past), np.zeros_like(past)
    ph, fh = np.zeros((2 * h, 2 * w), np.uint8), np.zeros((2 * h, 2 * w), np.uint8)
    lib.orc_lowpass(O.P(past), w, h, O.P(plp))
    lib.orc_lowpass(O.P(fut), w, h, O.P(flp))
    lib.orc_fir(O.P(past), w, h, O.P(ph))
    lib.orc_fir(O.P(fut), w, h, O.P(fh))
    si = np.zeros((h, w), np.uint8)
    res = np.zeros((h, w), np.int32)
    mbmv = np.zeros(g.b1616n, O.MV_DTYPE)
    bmv = np.zeros(g.b88n, O.MV_DTYPE)
    lib.orc_create_side_info(ctypes.byref(g), O.P(plp), O.P(flp), O.P(ph), O.P(fh), 0, 1, 2, O.P(si), O.P(res),
                             O.P(mbmv), O.P(bmv), 0)
    st = np.zeros((h, w), np.int32)
    lib.orc_forward_transform(ctypes.byref(g), O.P(res), O.P(st))
    alpha = np.zeros((h, w), np.float32)
    lib.orc_noise_parameter(ctypes.byref(g), q, O.P(st), O.P(alpha), None)
    sir = np.zeros(g.b44n, O.SEARCH_INFO_DTYPE)
    nsir = lib.orc_init_sir(ctypes.byref(g), None, O.P(sir))
    smf = np.zeros(nsir, O.SMF_DTYPE)
    lib.orc_init_smf(ctypes.byref(g), O.P(smf), O.P(sir), nsir)
    # a partially reconstructed frame: the original with seeded noise
    rng = np.random.default_rng(seed)
    recons = np.clip(cur.astype(np.int32) + rng.integers(-6, 7, cur.shape), 0, 255).astype(np.uint8)
    lib.orc_update_smf(ctypes.byref(g), O.P(smf), O.P(recons), O.P(si), O.P(sir), nsir, 0)
    oct_ = np.zeros((h - 3) * (w - 3) * 16, np.int32)
    lib.orc_overcomplete(ctypes.byref(g), O.P(si), O.P(oct_), 0)
    wz = W.encode_wz_frame(cur, q)
    return dict(g=g, vid=(past, cur, fut), plp=plp, flp=flp, ph=ph, fh=fh, si=si, alpha=alpha, sir=sir, nsir=nsir,
                smf=smf, recons=recons, oct=oct_, wz=wz, lib=lib, O=O, W=W)


def _band_params(q, ac_range, index):
    """(step, range) of a band (DVC_Decoder.cpp:139-152)."""
    from dvc_amd import workload as W
    qt = [int(v) for v in W.QTABLE[q]]
    if index == 0:
        return 1024 // qt[0], 1024
    step = int(np.ceil(float((int(ac_range[index]) + 1) << 1) / float(qt[index])))
    return step, ((qt[index] - 2) // 2 + 1) * step - 1


def _recons_trans(s, q, band_pos, index, bit):
    """Decoded-so-far indices of coefficient `index` before bit `bit`: the
    band's first (qbits - bit) bitplanes of the WZ frame, at every block."""
    from dvc_amd import workload as W
    h, w = s["vid"][0].shape
    qbits = int(W.QBITS[q][index])
    first = int(sum(W.QBITS[q][W.ZIGZAG[b]] for b in range(band_pos)))
    planes = s["wz"].bitplanes[first:first + qbits - bit].astype(np.int32)
    val = np.zeros(planes.shape[1], np.int32)
    for p in planes:
        val = (val << 1) | p
    rt = np.zeros((h, w), np.int32)
    bw = w // 4
    for b in range(val.size):
        rt[(b // bw) * 4 + (index >> 2), (b % bw) * 4 + (index & 3)] = val[b]
    return rt


@pytest.mark.parametrize("w,h,q,seed", CASES)
def test_forward_me_matches_oracle(ctx, w, h, q, seed):
    from dvc_amd import _lib
    s = _setup(w, h, q, seed)
    O, g = s["O"], s["g"]
    si1616 = np.zeros(g.b1616n, O.SEARCH_INFO_DTYPE)
    s["lib"].orc_init_1616(ctypes.byref(g), O.P(si1616))
    mb = np.zeros(g.b1616n, O.MV_DTYPE)
    holes = np.zeros(2 * g.b1616n, np.float64)
    s["lib"].orc_forward_me(ctypes.byref(g), O.P(si1616), O.P(s["flp"]), O.P(s["plp"]), 0, 1, 2, O.P(mb), O.P(holes), 0)
    got = _lib.forward_me(ctx, s["flp"], s["plp"], (0, 1, 2))
    assert got.tobytes() == mb.tobytes()


@pytest.mark.parametrize("w,h,q,seed", CASES)
def test_noise_distribution_and_cond_bit_prob_match_oracle(ctx, w, h, q, seed):
    from dvc_amd import _lib
    s = _setup(w, h, q, seed)
    O, g, lib, W = s["O"], s["g"], s["lib"], s["W"]
    sess = _lib.BandSession(ctx, s["vid"][0], s["vid"][2], q, (0, 1, 2), sir=s["sir"][:s["nsir"]])
    for band_pos in (0, 1, 2):  # DC and two AC bands
        index = int(W.ZIGZAG[band_pos])
        step, rng = _band_params(q, s["wz"].ac_range, index)
        per = 2 * rng + 1 if index else rng
        odist = np.zeros((s["nsir"], per), np.float32)
        n = lib.orc_noise_distribution(ctypes.byref(g), rng, index, O.P(s["alpha"]), O.P(s["smf"]), O.P(s["oct"]),
                                       O.P(s["sir"]), s["nsir"], O.P(odist), 0)
        assert n == per
        gd = sess.noise_distribution(rng, index, s["alpha"], s["smf"], s["oct"])
        assert gd.view(np.uint32).tobytes() == odist.view(np.uint32).tobytes(), f"band {band_pos}"
        qbits = int(W.QBITS[q][index])
        for bit in sorted({qbits, max(1, qbits - 2), 1}, reverse=True):
            rt = _recons_trans(s, q, band_pos, index, bit)
            olr = np.zeros(s["nsir"], np.float32)
            lib.orc_cond_bit_prob(ctypes.byref(g), q, O.P(olr), index, bit, step, O.P(rt), rng, O.P(odist),
                                  O.P(s["sir"]), s["nsir"], 0)
            glr = sess.cond_bit_prob(index, bit, step, rt, rng)
            assert glr.view(np.uint32).tobytes() == olr.view(np.uint32).tobytes(), f"band {band_pos} bit {bit}"


@pytest.mark.parametrize("w,h,q,seed", CASES)
def test_update_si_and_smf_match_oracle(ctx, w, h, q, seed):
    from dvc_amd import _lib
    s = _setup(w, h, q, seed)
    O, g, lib = s["O"], s["g"], s["lib"]
    past, _, fut = s["vid"]
    sess = _lib.BandSession(ctx, past, fut, q, (0, 1, 2))
    osi = s["si"].copy()
    lib.orc_update_si(ctypes.byref(g), O.P(s["recons"]), O.P(past), O.P(fut), O.P(s["ph"]), O.P(s["fh"]), 0, 1, 2,
                      O.P(s["sir"]), s["nsir"], O.P(osi), None, 0)
    gsi = sess.update_si(s["si"], s["recons"])
    assert np.array_equal(gsi, osi)
    osmf = s["smf"].copy()
    lib.orc_update_smf(ctypes.byref(g), O.P(osmf), O.P(s["recons"]), O.P(osi), O.P(s["sir"]), s["nsir"], 0)
    gsmf = sess.update_smf(s["smf"], s["recons"], osi)
    for b in range(s["nsir"]):
        n = int(osmf["probLen"][b])
        assert gsmf["prob"][b, :n].view(np.uint32).tobytes() == osmf["prob"][b, :n].view(np.uint32).tobytes(), b


def test_band_calls_check_their_order(ctx):
    from dvc_amd import _lib
    s = _setup(176, 144, 4, 1200)
    sess = _lib.BandSession(ctx, s["vid"][0], s["vid"][2], 4, (0, 1, 2))
    rt = np.zeros((144, 176), np.int32)
    with pytest.raises(_lib.DvcError) as e:
        sess.cond_bit_prob(0, 5, 32, rt, 1024)  # no distribution of this band yet
    assert e.value.rc == _lib.DVC_ESTATE
    bad = s["sir"][:s["nsir"]].copy()
    bad["blkIdx"][3] = 7
    with pytest.raises(_lib.DvcError) as e:
        _lib.BandSession(ctx, s["vid"][0], s["vid"][2], 4, (0, 1, 2), sir=bad)
    assert e.value.rc == _lib.DVC_EINVAL


@pytest.mark.parametrize("w,h,q,seed", CASES)
def test_band_calls_on_a_sir_subset_match_oracle(ctx, w, h, q, seed):
    """Skip and intra blocks: initSIR lists only the WZ 8x8 blocks' 4x4 blocks
    (sideInfoCreation.cpp:554-596), so the SMF arrays and the distribution are
    per SIR entry and condBitProb writes only those blocks' likelihood ratios
    (the others keep decodeGOP's 0.9999/0.0001, DVC_Decoder.cpp:203-207)."""
    from dvc_amd import _lib
    s = _setup(w, h, q, seed)
    O, g, lib, W = s["O"], s["g"], s["lib"], s["W"]
    past, _, fut = s["vid"]
    flags = (np.random.default_rng(seed + 7).random(g.b88n) < 0.3).astype(np.uint8) * 2  # about 30% skip blocks
    sir = np.zeros(g.b44n, O.SEARCH_INFO_DTYPE)
    nsir = lib.orc_init_sir(ctypes.byref(g), O.P(flags), O.P(sir))
    assert 0 < nsir < g.b44n
    smf = np.zeros(nsir, O.SMF_DTYPE)
    lib.orc_init_smf(ctypes.byref(g), O.P(smf), O.P(sir), nsir)
    lib.orc_update_smf(ctypes.byref(g), O.P(smf), O.P(s["recons"]), O.P(s["si"]), O.P(sir), nsir, 0)
    sess = _lib.BandSession(ctx, past, fut, q, (0, 1, 2), sir=sir[:nsir])
    for band_pos in (0, 1):
        index = int(W.ZIGZAG[band_pos])
        step, rng = _band_params(q, s["wz"].ac_range, index)
        per = 2 * rng + 1 if index else rng
        odist = np.zeros((nsir, per), np.float32)
        lib.orc_noise_distribution(ctypes.byref(g), rng, index, O.P(s["alpha"]), O.P(smf), O.P(s["oct"]),
                                   O.P(sir), nsir, O.P(odist), 0)
        gd = sess.noise_distribution(rng, index, s["alpha"], smf, s["oct"])
        assert gd.view(np.uint32).tobytes() == odist.view(np.uint32).tobytes(), f"band {band_pos}"
        qbits = int(W.QBITS[q][index])
        for bit in sorted({qbits, 1}, reverse=True):
            rt = _recons_trans(s, q, band_pos, index, bit)
            olr = np.full(g.b44n, np.float32(0.9999) / np.float32(0.0001), np.float32)
            lib.orc_cond_bit_prob(ctypes.byref(g), q, O.P(olr), index, bit, step, O.P(rt), rng, O.P(odist),
                                  O.P(sir), nsir, 0)
            glr = sess.cond_bit_prob(index, bit, step, rt, rng)
            assert glr.view(np.uint32).tobytes() == olr.view(np.uint32).tobytes(), f"band {band_pos} bit {bit}"
    osi = s["si"].copy()
    lib.orc_update_si(ctypes.byref(g), O.P(s["recons"]), O.P(past), O.P(fut), O.P(s["ph"]), O.P(s["fh"]), 0, 1, 2,
                      O.P(sir), nsir, O.P(osi), None, 0)
    gsi = sess.update_si(s["si"], s["recons"])
    assert np.array_equal(gsi, osi)
    osmf = smf.copy()
    lib.orc_update_smf(ctypes.byref(g), O.P(osmf), O.P(s["recons"]), O.P(osi), O.P(sir), nsir, 0)
    gsmf = sess.update_smf(smf, s["recons"], osi)
    for i in range(nsir):
        n = int(osmf["probLen"][i])
        assert gsmf["prob"][i, :n].view(np.uint32).tobytes() == osmf["prob"][i, :n].view(np.uint32).tobytes(), i
