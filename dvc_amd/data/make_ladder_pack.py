"""Writes the compact LDPCA ladder packs used by the library and the tests.

A pack holds what is needed to rebuild every decoder code of a ladder
(ReadLdpcaLadderFile_LdpcaDecoder.cpp:131-250): the encoder parity matrix
(column pointers jc and encoder rows, release/DvcEnc/LDPCA_LadderFile/
RegEncoder<N>.lad) and the per-code transmission pattern (bitsPerPeriod, idx)
that RegDecoder1584.lad carries for its 65 codes.  The 6336 decoder ladder is
not shipped by the reference; it uses the same pattern, so both regular packs
are derived from the 1584 pattern and the matching encoder ladder.  The
irregular QCIF pack (irreg1584.ladder) takes irRegEncoder1584.lad and the 66
codes' pattern of irReg1584.lad.

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


def pattern(path):
    """(numCodes, period, bitsPerPeriod, idx rows) of a decoder ladder file."""
    dec = read_ints(path)
    num_codes, n, e, period = dec[:4]
    pos = 4 + n + 1
    bits, idx = [], []
    for _ in range(num_codes):
        b = dec[pos]; pos += 1
        bits.append(b)
        idx.append(dec[pos:pos + b] + [0] * (period - b)); pos += b
        pos += e  # decoder rows, rebuilt from the encoder rows at load time
    assert pos == len(dec), f"{path} has trailing data"
    return num_codes, period, bits, idx


def write_pack(out, enc_path, num_codes, period, bits, idx, size):
    enc = read_ints(enc_path)
    en, ee = enc[:2]
    jc = enc[2:2 + en + 1]
    ir = enc[2 + en + 1:2 + en + 1 + ee]
    assert len(ir) == ee and en == size and en % period == 0
    blob = b"DVCL" + struct.pack("<5I", 1, en, ee, period, num_codes)
    blob += struct.pack(f"<{en + 1}I", *jc)
    blob += struct.pack(f"<{ee}H", *ir)
    blob += struct.pack(f"<{num_codes}H", *bits)
    blob += struct.pack(f"<{num_codes * period}H", *[v for row in idx for v in row])
    out.write_bytes(blob)
    print(f"{out.name}: N={en} E={ee} codes={num_codes} bytes={len(blob)}")


def main(ref):
    ref = Path(ref)
    out_dir = Path(__file__).resolve().parent
    lad = ref / "release/DvcDec/LDPCA_LadderFile"
    enc = ref / "release/DvcEnc/LDPCA_LadderFile"
    reg = pattern(lad / "RegDecoder1584.lad")
    for size in (1584, 6336):
        write_pack(out_dir / f"reg{size}.ladder", enc / f"RegEncoder{size}.lad", *reg, size)
    # the irregular QCIF ladder (66 codes, column degrees 2..21), the
    # reference's alternative at DvcDec/src/main.cpp:84-85
    write_pack(out_dir / "irreg1584.ladder", enc / "irRegEncoder1584.lad", *pattern(lad / "irReg1584.lad"), 1584)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
