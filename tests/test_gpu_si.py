"""Batched side information and noise parameters (dvc_side_info_batch)
against the oracle's createSideInfo, forwardTransform and noiseParameter:
SI and residue bit-exact, alpha bit-exact (the stated bar is 1e-5 relative),
for QCIF and CIF, several quantisers and unequal interpolation weights."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
from dvc_amd import workload as W

pytestmark = pytest.mark.gpu


def oracle_si_alpha(past, fut, idx, qindex):
    h, w = past.shape
    g = O.geom(w, h)
    lib = O.orc()
    past = np.ascontiguousarray(past)
    fut = np.ascontiguousarray(fut)
    plp, flp = np.zeros_like(past), np.zeros_like(fut)
    ph, fh = np.zeros((2 * h, 2 * w), np.uint8), np.zeros((2 * h, 2 * w), np.uint8)
    lib.orc_lowpass(O.P(past), w, h, O.P(plp)); lib.orc_lowpass(O.P(fut), w, h, O.P(flp))
    lib.orc_fir(O.P(past), w, h, O.P(ph)); lib.orc_fir(O.P(fut), w, h, O.P(fh))
    si = np.zeros((h, w), np.uint8)
    res = np.zeros((h, w), np.int32)
    mbmv = np.zeros(g.b1616n, O.MV_DTYPE)
    bmv = np.zeros(g.b88n, O.MV_DTYPE)
    lib.orc_create_side_info(ctypes.byref(g), O.P(plp), O.P(flp), O.P(ph), O.P(fh), idx[0], idx[1], idx[2],
                             O.P(si), O.P(res), O.P(mbmv), O.P(bmv), 0)
    trans = np.zeros((h, w), np.int32)
    lib.orc_forward_transform(ctypes.byref(g), O.P(res), O.P(trans))
    alpha = np.zeros((h, w), np.float32)
    lib.orc_noise_parameter(ctypes.byref(g), qindex, O.P(trans), O.P(alpha), None)
    return si, res, alpha


def test_side_info_batch_matches_oracle(ctx):
    cases = [((176, 144, 3), 4, (0, 1, 2)), ((352, 288, 6), 8, (0, 1, 2)), ((176, 144, 3), 1, (0, 1, 4)),
             ((352, 288, 6), 6, (0, 3, 4))]
    reqs = []
    for k, ((w, h, sh), q, idx) in enumerate(cases):
        vid = W.synthetic_video(w, h, 5, sh, seed=1200 + k)
        reqs.append(dict(past=vid[idx[0]], fut=vid[idx[2]], idx=idx, qindex=q))
    outs = ctx.side_info_batch(reqs)
    for d, (si, res, al) in zip(reqs, outs):
        osi, ores, oal = oracle_si_alpha(d["past"], d["fut"], d["idx"], d["qindex"])
        assert np.array_equal(si, osi)
        assert np.array_equal(res, ores)
        assert np.array_equal(al.view(np.uint32), oal.view(np.uint32))
        assert np.count_nonzero(al) > 0
