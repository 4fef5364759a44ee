"""Writes the compact LDPCA ladder packs used by the library and the tests.

A pack holds what is needed to rebuild every decoder code of the regular
ladder (ReadLdpcaLadderFile_LdpcaDecoder.cpp:131-250): the encoder parity
matrix (column pointers jc and encoder rows, release/DvcEnc/LDPCA_LadderFile/
RegEncoder<N>.lad) and the per-code transmission pattern (bitsPerPeriod, idx)
that RegDecoder1584.lad carries for its 65 codes.  The 6336 decoder ladder is
not shipped by the reference; it uses the same pattern, so both packs are
derived from the 1584 pattern and the matching encoder ladder.

Layout (little endian): b"DVCL", u32 version=1, u32 N, u32 E, u32 period,
u32 numCodes, u32 jc[N+1], u16 enc_ir[E], u16 bitsPerPeriod[numCodes],
u16 idx[numCodes*period] (unused tail entries 0).

Run in a checkout that has the reference tree:
    python dvc_amd/data/make_ladder_pack.py /root/reference
"""
import struct
import sys
from pathlib import Path


def read_ints(path):
    return [int(t) for t in Path(path).read_text().split()]


def main(ref):
    ref = Path(ref)
    dec = read_ints(ref / "release/DvcDec/LDPCA_LadderFile/RegDecoder1584.lad")
    num_codes, n, e, period = dec[:4]
    pos = 4 + n + 1
    bits, idx = [], []
    for _ in range(num_codes):
        b = dec[pos]; pos += 1
        bits.append(b)
        idx.append(dec[pos:pos + b] + [0] * (period - b)); pos += b
        pos += e  # decoder rows, rebuilt from the encoder rows at load time
    assert pos == len(dec), "RegDecoder1584.lad has trailing data"
    out_dir = Path(__file__).resolve().parent
    for size in (1584, 6336):
        enc = read_ints(ref / f"release/DvcEnc/LDPCA_LadderFile/RegEncoder{size}.lad")
        en, ee = enc[:2]
        jc = enc[2:2 + en + 1]
        ir = enc[2 + en + 1:2 + en + 1 + ee]
        assert len(ir) == ee and en == size and en % period == 0
        blob = b"DVCL" + struct.pack("<5I", 1, en, ee, period, num_codes)
        blob += struct.pack(f"<{en + 1}I", *jc)
        blob += struct.pack(f"<{ee}H", *ir)
        blob += struct.pack(f"<{num_codes}H", *bits)
        blob += struct.pack(f"<{num_codes * period}H", *[v for row in idx for v in row])
        (out_dir / f"reg{size}.ladder").write_bytes(blob)
        print(f"reg{size}.ladder: N={en} E={ee} codes={num_codes} bytes={len(blob)}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
