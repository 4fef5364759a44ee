"""Writes the golden fixtures under tests/golden/ from the reference's own CPU
sources (compiled into oracle/_ref by `make -C oracle ref`) and data files.
Run in a checkout that has /root/reference:

    python tests/golden/make_golden.py

Fixtures (all small JSON):
  ladder.json     hashes of the reference's RegDecoder1584.lad decoder rows and
                  Inverse_Matrix_H_Reg1584.bin, and the row-degree histograms
                  the reference lists for N=6336 codes (LDPCA_kernel.cu comments)
  encoder.json    reference encoder (DvcEnc forwardTransform/quantize/
                  extractBitplane/LdpcaEncoder) outputs for seeded frames
  ldpca.json      reference LdpcaDecoder (ParallelMode 0) results for seeded
                  bitplanes: start code, accepted code, rate, entropy bits,
                  decoded-bitplane hash
  frames.json     reference decodeGOP results for seeded WZ frames: accepted
                  code per bitplane, entropy bits, bitrate, PSNR, hashes of
                  the side information and of the reconstruction
  frames_mode1.json  the same with ParallelMode 1
  ldpca_dynstep.json, frames_dynstep.json  the ldpca and frames fixtures from
                  the decoder built with DYNAMICSTEP defined in
                  LdpcaDecoder.cpp (DVC_GOLDEN_ONLY=dynstep writes just these)
  ldpca_irreg.json, frames_irreg.json, frames_irreg_mode1.json  the irregular
                  QCIF ladder (irReg1584.lad): reference LdpcaDecoder on
                  seeded bitplanes and decodeGOP on QCIF frames, ParallelMode
                  0 and 1 (DVC_GOLDEN_ONLY=irreg writes just these)
  ldpca_nan.json, ldpca_nan_mode1.json  reference LdpcaDecoder, ParallelMode
                  0 and 1, on seeded bitplanes with NaN likelihood ratios
                  (DVC_GOLDEN_ONLY=nan writes just these)
  frames_gop.json, frames_gop_mode1.json  reference decodeGOP(0, gop) for
                  GOP 4 and 8 (decoded WZ frames are references of later
                  ones), ParallelMode 0 and 1 (DVC_GOLDEN_ONLY=gop writes
                  just these)
Inputs are regenerated from their seeds by dvc_amd.workload, so a fixture
holds only seeds, sizes and expected outputs.
"""
import ctypes
import hashlib
import json
import os
import struct
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle_lib as O  # noqa: E402
from dvc_amd import workload as W  # noqa: E402

REF = Path("/root/reference")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def fbits(x):
    return struct.unpack("<I", struct.pack("<f", float(x)))[0]


def ladder_fixture():
    L = O.Ladder(lad_path=REF / "release/DvcDec/LDPCA_LadderFile/RegDecoder1584.lad",
                 inv_path=REF / "release/DvcDec/LDPCA_LadderFile/Inverse_Matrix_H_Reg1584.bin")
    ir = np.stack([L.ir(c) for c in range(L.num_codes)]).astype(np.int32)
    inv = (REF / "release/DvcDec/LDPCA_LadderFile/Inverse_Matrix_H_Reg1584.bin").read_bytes()
    return {
        "N": 1584, "num_codes": int(L.num_codes), "E": int(L.E),
        "synd_length": L.synd_length().tolist(),
        "decoder_ir_sha256": sha(ir),
        "inverse_sha256": hashlib.sha256(inv).hexdigest(),
        # Row-degree histograms of the N=6336 decoder codes, as listed in the
        # reference's kernel comments (DvcDec/src/LDPCA_kernel.cu:321,364,465,584):
        # {code: {degree: rows}}
        # the irregular QCIF ladder irReg1584.lad (main.cpp:84-85): decoder rows
        # of its 66 codes; the reference ships no inverse for it
        "irreg": irreg_ladder_facts(),
        "cif_row_degrees": {
            "1": {"48": 96, "50": 5, "51": 84, "52": 7, "98": 10, "99": 78, "100": 8},
            "3": {"23": 5, "24": 84, "25": 7, "27": 96, "48": 192, "50": 10, "51": 78, "52": 8},
            "30": {"5": 15, "6": 2850, "7": 15, "9": 192},
            "63": {"2": 15, "3": 6114, "4": 15, "6": 96},
        },
    }


def irreg_ladder_facts():
    L = O.Ladder(lad_path=REF / "release/DvcDec/LDPCA_LadderFile/irReg1584.lad")
    ir = np.stack([L.ir(c) for c in range(L.num_codes)]).astype(np.int32)
    return {"N": 1584, "num_codes": int(L.num_codes), "E": int(L.E), "synd_length": L.synd_length().tolist(),
            "decoder_ir_sha256": sha(ir)}


def encoder_fixture():
    enc = O.ref_enc()
    cases = []
    for (w, h, q, seed, frame) in [(176, 144, 4, 1000, 1), (176, 144, 8, 1001, 3), (352, 288, 8, 1002, 1),
                                   (352, 288, 1, 1003, 1)]:
        vid = W.synthetic_video(w, h, frame + 2, 3 if w == 176 else 6, seed)
        n = (w // 4) * (h // 4)
        lad = REF / f"release/DvcEnc/LDPCA_LadderFile/RegEncoder{n}.lad"
        nbp = enc.refenc_setup(w, h, q, str(lad).encode())
        acr = np.zeros(16, np.int32)
        crc = np.zeros(nbp, np.uint8)
        acc = np.zeros((nbp, n // 8), np.uint8)
        enc.refenc_encode(O.P(np.ascontiguousarray(vid[frame])), O.P(acr), O.P(crc), O.P(acc))
        cases.append({"W": w, "H": h, "qindex": q, "seed": seed, "frame": frame,
                      "shift": 3 if w == 176 else 6, "nbitplanes": int(nbp),
                      "ac_range": acr.tolist(), "crc": crc.tolist(), "acc_sha256": sha(acc)})
    return {"cases": cases}


def ldpca_fixture():
    R = O.ref_dec()
    out = []
    for (n_cif, n_qcif, seed) in [(0, 24, 31), (8, 0, 32)]:
        w, h = (352, 288) if n_cif else (176, 144)
        O.RefDecoder(w, h, 4)
        planes = W.ldpca_bitplane_set(n_cif=n_cif, n_qcif=n_qcif, seed=seed)
        for i, pl in enumerate(planes):
            lr = pl["LR"].copy()
            ent = R.ref_compute_ce(O.P(lr))
            acc_bits = np.unpackbits(pl["acc"]).astype(np.uint8)
            dec = np.full(pl["N"], 0xAA, np.uint8)
            rate = ctypes.c_int()
            R.ref_ldpca(O.P(lr), O.P(acc_bits), ctypes.c_uint8(pl["crc"]), ctypes.c_float(ent), O.P(dec),
                        ctypes.byref(rate))
            level = [(rate.value - 8) // (pl["N"] // 66) - 2]
            out.append({"set": [n_cif, n_qcif, seed], "index": i, "N": pl["N"], "rate": rate.value,
                        "level": level[0] if level else -1, "entropy_bits": fbits(ent),
                        "decoded_sha256": sha(dec), "source_ok": bool(np.array_equal(dec, pl["bits"]))})
    return {"max_iter": 50, "mode": 0, "items": out}


def ldpca_dynstep_fixture():
    """The reference LdpcaDecoder built with DYNAMICSTEP (oracle/_ref/
    libdvcref_dyn.so) on the bitplanes of ldpca_fixture."""
    R = O.ref_dec(dynstep=True)
    out = []
    for (n_cif, n_qcif, seed) in [(0, 24, 31), (8, 0, 32)]:
        w, h = (352, 288) if n_cif else (176, 144)
        O.RefDecoder(w, h, 4, dynstep=True)
        planes = W.ldpca_bitplane_set(n_cif=n_cif, n_qcif=n_qcif, seed=seed)
        for i, pl in enumerate(planes):
            lr = pl["LR"].copy()
            ent = R.ref_compute_ce(O.P(lr))
            acc_bits = np.unpackbits(pl["acc"]).astype(np.uint8)
            dec = np.full(pl["N"], 0xAA, np.uint8)
            rate = ctypes.c_int()
            R.ref_ldpca(O.P(lr), O.P(acc_bits), ctypes.c_uint8(pl["crc"]), ctypes.c_float(ent), O.P(dec),
                        ctypes.byref(rate))
            out.append({"set": [n_cif, n_qcif, seed], "index": i, "N": pl["N"], "rate": rate.value,
                        "level": (rate.value - 8) // (pl["N"] // 66) - 2, "entropy_bits": fbits(ent),
                        "decoded_sha256": sha(dec), "source_ok": bool(np.array_equal(dec, pl["bits"]))})
    return {"max_iter": 50, "mode": 0, "rate_step": 1, "items": out}


FRAME_CASES = [
    # (W, H, qindex, seed, shift, frames, wz frame index)  -- config-0 style QCIF and two CIF frames
    (176, 144, 4, 1000, 3, 3, 1),
    (176, 144, 8, 1004, 3, 3, 1),
    (352, 288, 2, 1005, 6, 3, 1),
    (352, 288, 8, 1006, 6, 3, 1),
]


def frames_fixture(mode=0, dynstep=False):
    out = []
    for (w, h, q, seed, shift, nf, fi) in FRAME_CASES:
        vid = W.synthetic_video(w, h, nf, shift, seed)
        wz = W.encode_wz_frame(vid[fi], q)
        R = O.RefDecoder(w, h, q, mode=mode, nthreads=1, dynstep=dynstep)
        # side information of the reference (createSideInfo)
        si = np.zeros((h, w), np.uint8)
        res = np.zeros((h, w), np.int32)
        R.lib.ref_set_frame(0, O.P(np.ascontiguousarray(vid[0])), O.P(np.ascontiguousarray(vid[fi])))
        R.lib.ref_set_frame(2, O.P(np.ascontiguousarray(vid[2])), O.P(np.ascontiguousarray(vid[fi])))
        R.lib.ref_create_side_info(0, 2, 1, O.P(si), O.P(res), None, None)
        R = O.RefDecoder(w, h, q, mode=mode, nthreads=1, dynstep=dynstep)
        rec, info = R.decode_wz_frame(vid[0], vid[2], vid[fi], wz)
        out.append({"W": w, "H": h, "qindex": q, "seed": seed, "shift": shift, "frames": nf, "wz": fi,
                    "mode": mode, "levels": info["levels"], "rates": info["rates"],
                    "entropy_bits": [fbits(e) for e in info["entropy"]], "bitrate": info["bitrate"],
                    "psnr": info["psnr"], "si_sha256": sha(si), "residue_sha256": sha(res),
                    "recons_sha256": sha(rec)})
        print(f"frame case {w}x{h} q{q} mode{mode}: psnr {info['psnr']:.4f} bitrate {info['bitrate']}", flush=True)
    return {"max_iter": 50, "cases": out} if not dynstep else {"max_iter": 50, "rate_step": 1, "cases": out}


def ldpca_irreg_fixture():
    """Reference LdpcaDecoder (ParallelMode 0) with the irregular QCIF ladder
    irReg1584.lad (main.cpp:84-85) on seeded bitplanes whose syndromes come
    from irRegEncoder1584.lad."""
    R = O.ref_dec()
    O.RefDecoder(176, 144, 4, kind="irreg")
    out = []
    for i, pl in enumerate(W.ldpca_bitplane_set(n_cif=0, n_qcif=24, seed=35, kind="irreg")):
        lr = pl["LR"].copy()
        ent = R.ref_compute_ce(O.P(lr))
        acc_bits = np.unpackbits(pl["acc"]).astype(np.uint8)
        dec = np.full(pl["N"], 0xAA, np.uint8)
        rate = ctypes.c_int()
        R.ref_ldpca(O.P(lr), O.P(acc_bits), ctypes.c_uint8(pl["crc"]), ctypes.c_float(ent), O.P(dec),
                    ctypes.byref(rate))
        out.append({"set": [0, 24, 35], "index": i, "N": pl["N"], "rate": rate.value,
                    "level": (rate.value - 8) // (pl["N"] // 66) - 1, "entropy_bits": fbits(ent),
                    "decoded_sha256": sha(dec), "source_ok": bool(np.array_equal(dec, pl["bits"]))})
    return {"max_iter": 50, "mode": 0, "ladder": "irreg", "items": out}


IRREG_FRAME_CASES = [(176, 144, 4, 1020, 3, 3, 1), (176, 144, 8, 1021, 3, 3, 1)]


def frames_irreg_fixture(mode=0):
    """Reference decodeGOP on QCIF WZ frames with the irregular ladder."""
    out = []
    for (w, h, q, seed, shift, nf, fi) in IRREG_FRAME_CASES:
        vid = W.synthetic_video(w, h, nf, shift, seed)
        wz = W.encode_wz_frame(vid[fi], q, kind="irreg")
        R = O.RefDecoder(w, h, q, mode=mode, nthreads=1, kind="irreg")
        rec, info = R.decode_wz_frame(vid[0], vid[2], vid[fi], wz)
        out.append({"W": w, "H": h, "qindex": q, "seed": seed, "shift": shift, "frames": nf, "wz": fi,
                    "mode": mode, "levels": info["levels"], "rates": info["rates"],
                    "entropy_bits": [fbits(e) for e in info["entropy"]], "bitrate": info["bitrate"],
                    "psnr": info["psnr"], "recons_sha256": sha(rec)})
        print(f"irregular frame {w}x{h} q{q} mode{mode}: psnr {info['psnr']:.4f} bitrate {info['bitrate']}",
              flush=True)
    return {"max_iter": 50, "ladder": "irreg", "cases": out}


NAN_BITS = 0xFFC00000  # x86's default NaN: what 0/0 yields on the reference's CPU path


def nan_planes():
    """Seeded bitplanes (workload.ldpca_bitplane_set) with 1 or 3 likelihood
    ratios replaced by NaN, as condBitProb yields them when both of a block's
    interval sums underflow to zero (p0/(p0+p1) = 0/0)."""
    rng = np.random.default_rng(77)
    out = []
    for (n_cif, n_qcif, seed) in [(0, 12, 61), (3, 0, 62)]:
        for i, pl in enumerate(W.ldpca_bitplane_set(n_cif=n_cif, n_qcif=n_qcif, seed=seed)):
            nn = 1 if i % 2 == 0 else 3
            pos = sorted(int(v) for v in rng.choice(pl["N"], nn, replace=False))
            out.append(((n_cif, n_qcif, seed), i, pl, pos))
    return out


def with_nans(lr, pos):
    lr = np.array(lr, np.float32)
    lr.view(np.uint32)[pos] = NAN_BITS
    return lr


def ldpca_nan_fixture(mode):
    """Reference LdpcaDecoder (ParallelMode `mode`, one thread) on bitplanes
    with NaN likelihood ratios: accepted code, rate, entropy bits, decoded
    hash, and (ParallelMode 1) the BP iteration count."""
    R = O.ref_dec()
    items = []
    for key, i, pl, pos in nan_planes():
        w, h = (352, 288) if pl["N"] == 6336 else (176, 144)
        O.RefDecoder(w, h, 4, mode=mode, nthreads=1)
        lr = with_nans(pl["LR"], pos)
        ent = R.ref_compute_ce(O.P(lr.copy())) if mode == 0 else R.ref_compute_ce_omp(O.P(lr.copy()))
        acc_bits = np.unpackbits(pl["acc"]).astype(np.uint8)
        dec = np.full(pl["N"], 0xAA, np.uint8)
        rate = ctypes.c_int()
        it0 = R.ref_bp_iterations()
        R.ref_ldpca(O.P(lr), O.P(acc_bits), ctypes.c_uint8(pl["crc"]), ctypes.c_float(ent), O.P(dec), ctypes.byref(rate))
        items.append({"set": list(key), "index": i, "N": pl["N"], "nan_positions": pos, "rate": rate.value,
                      "level": (rate.value - 8) // (pl["N"] // 66) - 2, "entropy_bits": fbits(ent),
                      "decoded_sha256": sha(dec), "source_ok": bool(np.array_equal(dec, pl["bits"])),
                      "bp_iterations": (R.ref_bp_iterations() - it0) if mode == 1 else None})
    return {"max_iter": 50, "mode": mode, "nan_bits": NAN_BITS, "items": items}


GOP_CASES = [
    # (W, H, qindex, seed, shift, gop): frames 0..gop of a seeded video, key
    # frames at 0 and gop, the WZ frames in between decoded hierarchically
    (176, 144, 4, 1010, 3, 4),
    (176, 144, 4, 1011, 3, 8),
    (352, 288, 8, 1012, 6, 4),
    (352, 288, 8, 1013, 6, 8),
]
GOP_CASES_MODE1 = [(176, 144, 4, 1011, 3, 8)]


def frames_gop_fixture(mode=0):
    """Reference decodeGOP(0, gop) for GOP 4 and 8: per WZ frame in the
    reference's recursion order (DVC_Decoder.cpp:471-472), its references,
    accepted code per bitplane, entropy bits, bitrate, PSNR and the hash of
    its reconstruction (which later frames of the GOP take as a reference)."""
    out = []
    for (w, h, q, seed, shift, gop) in (GOP_CASES if mode == 0 else GOP_CASES_MODE1):
        vid = W.synthetic_video(w, h, gop + 1, shift, seed)
        wz = {k: W.encode_wz_frame(vid[k], q) for k in range(1, gop)}
        R = O.RefDecoder(w, h, q, mode=mode, nthreads=1, nframes=gop + 1)
        res, order = R.decode_gop(vid, wz, 0, gop)
        frames = []
        for cur, past, fut in order:
            rec, info = res[cur]
            frames.append({"cur": cur, "past": past, "fut": fut, "levels": info["levels"],
                           "entropy_bits": [fbits(e) for e in info["entropy"]], "bitrate": info["bitrate"],
                           "psnr": info["psnr"], "recons_sha256": sha(rec)})
            print(f"gop case {w}x{h} q{q} gop{gop} mode{mode} frame {cur} ({past},{fut}): "
                  f"psnr {info['psnr']:.4f} bitrate {info['bitrate']}", flush=True)
        out.append({"W": w, "H": h, "qindex": q, "seed": seed, "shift": shift, "gop": gop, "mode": mode,
                    "frames": frames})
    return {"max_iter": 50, "cases": out}


def main():
    assert O.have_reference_build(), "run `make -C oracle ref` first"
    if os.environ.get("DVC_GOLDEN_ONLY") == "ladder":
        (HERE / "ladder.json").write_text(json.dumps(ladder_fixture(), indent=1))
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "irreg":
        (HERE / "ldpca_irreg.json").write_text(json.dumps(ldpca_irreg_fixture(), indent=1))
        (HERE / "frames_irreg.json").write_text(json.dumps(frames_irreg_fixture(0), indent=1))
        env = dict(os.environ, DVC_GOLDEN_ONLY="irreg_mode1", MALLOC_PERTURB_="85", OMP_NUM_THREADS="1")
        subprocess.run([sys.executable, __file__], env=env, check=True)
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "irreg_mode1":
        (HERE / "frames_irreg_mode1.json").write_text(json.dumps(frames_irreg_fixture(1), indent=1))
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "nan":
        (HERE / "ldpca_nan.json").write_text(json.dumps(ldpca_nan_fixture(0), indent=1))
        env = dict(os.environ, DVC_GOLDEN_ONLY="nan_mode1", MALLOC_PERTURB_="85", OMP_NUM_THREADS="1")
        subprocess.run([sys.executable, __file__], env=env, check=True)
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "nan_mode1":
        (HERE / "ldpca_nan_mode1.json").write_text(json.dumps(ldpca_nan_fixture(1), indent=1))
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "gop":
        (HERE / "frames_gop.json").write_text(json.dumps(frames_gop_fixture(mode=0), indent=1))
        env = dict(os.environ, DVC_GOLDEN_ONLY="gop_mode1", MALLOC_PERTURB_="85", OMP_NUM_THREADS="1")
        subprocess.run([sys.executable, __file__], env=env, check=True)
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "gop_mode1":
        (HERE / "frames_gop_mode1.json").write_text(json.dumps(frames_gop_fixture(mode=1), indent=1))
        return
    if os.environ.get("DVC_GOLDEN_ONLY") == "dynstep":
        # the DYNAMICSTEP build's fixtures alone (the others are unchanged by it)
        (HERE / "ldpca_dynstep.json").write_text(json.dumps(ldpca_dynstep_fixture(), indent=1))
        (HERE / "frames_dynstep.json").write_text(json.dumps(frames_fixture(mode=0, dynstep=True), indent=1))
        return
    (HERE / "ladder.json").write_text(json.dumps(ladder_fixture(), indent=1))
    (HERE / "encoder.json").write_text(json.dumps(encoder_fixture(), indent=1))
    (HERE / "ldpca.json").write_text(json.dumps(ldpca_fixture(), indent=1))
    if os.environ.get("DVC_GOLDEN_CHILD") == "mode1":
        (HERE / "frames_mode1.json").write_text(json.dumps(frames_fixture(mode=1), indent=1))
        return
    (HERE / "frames.json").write_text(json.dumps(frames_fixture(mode=0), indent=1))
    # ParallelMode 1 keeps the BP decision buffer across calls; its first use
    # reads fresh malloc memory, made deterministic (0xAA) with MALLOC_PERTURB_.
    env = dict(os.environ, DVC_GOLDEN_CHILD="mode1", MALLOC_PERTURB_="85", OMP_NUM_THREADS="1")
    subprocess.run([sys.executable, __file__], env=env, check=True)


if __name__ == "__main__":
    main()
