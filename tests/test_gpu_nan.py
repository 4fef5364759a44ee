"""Bitplanes whose likelihood ratios hold NaN (condBitProb's p0/(p0+p1) when
both interval sums underflow to zero, as in the CIF Q8 GOP 4 golden frame)
against the reference's LdpcaDecoder on the same inputs
(tests/golden/ldpca_nan.json, ldpca_nan_mode1.json): accepted code, entropy
bits (x86's default NaN), decoded bitplane and, in ParallelMode 1, the BP
iterations the early jump leaves; plain (one code per launch) and batched
(speculative) decodes, in both ParallelModes."""
import hashlib
import json
import struct
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def _fbits(x):
    return struct.unpack("<I", struct.pack("<f", float(x)))[0]


def _inputs(fx):
    from dvc_amd import workload as W
    d = json.loads((GOLD / fx).read_text())
    sets = {}
    out = []
    for it in d["items"]:
        key = tuple(it["set"])
        if key not in sets:
            sets[key] = W.ldpca_bitplane_set(n_cif=key[0], n_qcif=key[1], seed=key[2])
        pl = sets[key][it["index"]]
        lr = np.array(pl["LR"], np.float32)
        lr.view(np.uint32)[it["nan_positions"]] = d["nan_bits"]
        out.append((it, pl, lr))
    return d, out


@pytest.mark.parametrize("fx", ["ldpca_nan.json", "ldpca_nan_mode1.json"])
def test_nan_bitplanes_match_reference(fx):
    from dvc_amd import _lib
    d, cases = _inputs(fx)
    with _lib.Context(max_iter=d["max_iter"], parallel_mode=d["mode"]) as c:
        plain = [c.ldpca_decode(lr, pl["acc"], pl["crc"], carry=np.full(pl["N"], 0xAA, np.uint8))
                 for it, pl, lr in cases]
        bdec, bst = c.ldpca_decode_batch([lr for _, _, lr in cases], [pl["acc"] for _, pl, _ in cases],
                                         [pl["crc"] for _, pl, _ in cases])
    for k, (it, pl, lr) in enumerate(cases):
        for how, dec, st in (("plain", plain[k][0], plain[k][1]), ("batched", bdec[k], bst[k])):
            where = f"{fx} item {k} {how}"
            assert st["level"] == it["level"], where
            assert _fbits(st["entropy"]) == it["entropy_bits"], where
            assert hashlib.sha256(np.ascontiguousarray(dec).tobytes()).hexdigest() == it["decoded_sha256"], where
            if it["bp_iterations"] is not None:
                # the reference's counter leaves out the last iteration of each
                # code that ends with a break (LdpcaDecoder.cpp:186): here every
                # code ends with the early jump's break
                assert st["iterations"] == it["bp_iterations"] + (it["level"] - st["start_level"]), where
            else:
                assert st["iterations"] == d["max_iter"] * (it["level"] - st["start_level"]), where
