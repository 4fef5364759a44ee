"""The batched LDPCA workload at its full shape: 512 seeded bitplanes
(384 x N=6336, 128 x N=1584, the bench's ldpca_batch set) decoded in one
speculative persistent launch at max_iter 100.  Every bitplane's accepted
code, rate, iteration count and bits must equal the sequential rate search
of the plain drop-in (dvc_ldpca_decode, LdpcaDecoder.cpp:646-771 one code at
a time), every decode must return its source bits, and a seeded sample of
eight must equal the oracle (the CPU restatement of LdpcaDecoder) at the same
iteration cap."""
import numpy as np
import pytest

import oracle_lib as O
from dvc_amd import workload as W

pytestmark = pytest.mark.gpu


def test_full_batch_equals_sequential_search_and_oracle_sample():
    from dvc_amd import _lib
    planes = W.ldpca_bitplane_set(n_cif=384, n_qcif=128, seed=2)
    assert len(planes) == 512
    c = _lib.Context(max_iter=100)
    try:
        decs, sts = c.ldpca_decode_batch([p["LR"] for p in planes], [p["acc"] for p in planes],
                                         [p["crc"] for p in planes])
        print("batch decoded", flush=True)
        for k, (pl, dec, st) in enumerate(zip(planes, decs, sts)):
            assert np.array_equal(dec, pl["bits"]), k
            pdec, pst = c.ldpca_decode(pl["LR"], pl["acc"], pl["crc"])
            assert np.array_equal(dec, pdec), k
            for key in ("start_level", "level", "rate", "iterations", "used_inverse"):
                assert st[key] == pst[key], (k, key, st[key], pst[key])
        print("sequential search equal on all 512", flush=True)
        rng = np.random.default_rng(77)
        cif = [k for k, p in enumerate(planes) if p["N"] == 6336]
        qcif = [k for k, p in enumerate(planes) if p["N"] == 1584]
        for k in list(rng.choice(cif, 4, replace=False)) + list(rng.choice(qcif, 4, replace=False)):
            pl = planes[k]
            odec, ost, _ = O.ldpca_decode(pl["LR"], pl["acc"], pl["crc"], max_iter=100)
            assert np.array_equal(decs[k], odec), k
            assert sts[k]["level"] == ost["level"] and sts[k]["rate"] == ost["rate"], k
            assert sts[k]["iterations"] == ost["total_iters"] and sts[k]["used_inverse"] == ost["used_inverse"], k
            print(f"oracle sample {k} (N={pl['N']}) equal", flush=True)
    finally:
        c.close()
