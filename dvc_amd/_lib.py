"""ctypes binding of libdvc_accel.so (include/dvc_accel.h).

Importing this module loads the library; a missing library raises instead of
falling back to any CPU path.
"""
import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("DVC_ACCEL_LIB", PKG / "libdvc_accel.so"))
DATA_DIR = PKG / "data"

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python -m dvc_amd.build` "
        "(hipcc --offload-arch=gfx950); there is no CPU fallback")

lib = ctypes.CDLL(str(LIB_PATH))

DVC_OK, DVC_EINVAL, DVC_EDEVICE, DVC_EIO, DVC_ENOMEM, DVC_ESTATE = 0, -1, -2, -3, -4, -5


class DvcError(RuntimeError):
    def __init__(self, rc, where):
        msg = lib.dvc_last_error().decode(errors="replace")
        super().__init__(f"{where} failed ({rc}): {msg}")
        self.rc = rc


class Params(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("max_iter", ctypes.c_int),
                ("data_dir", ctypes.c_char_p), ("bp_threads", ctypes.c_int), ("parallel_mode", ctypes.c_int),
                ("rate_step", ctypes.c_int), ("qcif_ladder", ctypes.c_int)]


class BpStats(ctypes.Structure):
    _fields_ = [("start_level", ctypes.c_int), ("level", ctypes.c_int), ("rate", ctypes.c_int),
                ("requests", ctypes.c_int), ("iterations", ctypes.c_int),
                ("used_inverse", ctypes.c_int), ("entropy", ctypes.c_float), ("iterations_run", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LdpcaItem(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("LR", ctypes.c_void_p), ("acc", ctypes.c_void_p),
                ("crc", ctypes.c_uint8), ("start_level", ctypes.c_int),
                ("decoded", ctypes.c_void_p), ("stats", BpStats)]


class FrameStats(ctypes.Structure):
    _fields_ = [("nbitplanes", ctypes.c_int), ("bitrate", ctypes.c_int), ("levels", ctypes.c_int * 128),
                ("start_levels", ctypes.c_int * 128), ("iterations", ctypes.c_int * 128),
                ("used_inverse", ctypes.c_int * 128), ("entropy", ctypes.c_float * 128),
                ("psnr", ctypes.c_double)]

    def as_dict(self):
        n = self.nbitplanes
        return dict(nbitplanes=n, bitrate=self.bitrate, levels=list(self.levels[:n]),
                    start_levels=list(self.start_levels[:n]), iterations=list(self.iterations[:n]),
                    used_inverse=list(self.used_inverse[:n]), entropy=list(self.entropy[:n]), psnr=self.psnr)


class WZFrameSpec(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("qindex", ctypes.c_int),
                ("past_idx", ctypes.c_int), ("cur_idx", ctypes.c_int), ("fut_idx", ctypes.c_int),
                ("past", ctypes.c_void_p), ("fut", ctypes.c_void_p), ("orig", ctypes.c_void_p),
                ("ac_range", ctypes.c_void_p), ("crc", ctypes.c_void_p), ("acc", ctypes.c_void_p),
                ("recons", ctypes.c_void_p), ("stats", ctypes.POINTER(FrameStats)),
                ("block_mode", ctypes.c_void_p), ("mode_bits", ctypes.c_int), ("intra_coeffs", ctypes.c_void_p),
                ("deblocking", ctypes.c_int), ("intra_qp", ctypes.c_int), ("past_is_wz", ctypes.c_int),
                ("past_block_mode", ctypes.c_void_p)]


class Tuning(ctypes.Structure):
    _fields_ = [("cif_threads", ctypes.c_int), ("qcif_threads", ctypes.c_int), ("qcif_wg_per_cu", ctypes.c_int),
                ("poll", ctypes.c_int), ("cif_ms", ctypes.c_double), ("qcif_ms", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SIRequest(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("qindex", ctypes.c_int),
                ("past_idx", ctypes.c_int), ("cur_idx", ctypes.c_int), ("fut_idx", ctypes.c_int),
                ("past", ctypes.c_void_p), ("fut", ctypes.c_void_p), ("si", ctypes.c_void_p),
                ("residue", ctypes.c_void_p), ("alpha", ctypes.c_void_p)]


class EngineParams(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_inflight", ctypes.c_int), ("sequential_bp", ctypes.c_int),
                ("lanes", ctypes.c_int), ("eager_steps", ctypes.c_int), ("device_buffers", ctypes.c_int),
                ("lane_steps", ctypes.c_int)]


# Every symbol include/dvc_accel.h declares, with its ctypes signature.
_c = ctypes
class WzStreamInfo(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("qindex", ctypes.c_int), ("intra_qp", ctypes.c_int),
                ("gop", ctypes.c_int), ("frames", ctypes.c_int), ("deblocking", ctypes.c_int),
                ("wz_frames", ctypes.c_int), ("nbitplanes", ctypes.c_int), ("bytes_read", ctypes.c_size_t)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class WzStreamFrame(ctypes.Structure):
    _fields_ = [("past_idx", ctypes.c_int), ("cur_idx", ctypes.c_int), ("fut_idx", ctypes.c_int),
                ("ac_range", ctypes.c_int * 16), ("skip_mode", ctypes.c_int), ("intra_mode", ctypes.c_int),
                ("mode_bits", ctypes.c_int), ("block_mode", ctypes.c_void_p), ("crc", ctypes.c_void_p),
                ("acc", ctypes.c_void_p), ("intra_coeffs", ctypes.c_void_p), ("past_is_wz", ctypes.c_int),
                ("past_block_mode", ctypes.c_void_p)]


SIGNATURES = {
    "dvc_ctx_create": (_c.c_int, [_c.POINTER(Params), _c.POINTER(_c.c_void_p)]),
    "dvc_ctx_destroy": (_c.c_int, [_c.c_void_p]),
    "dvc_last_error": (_c.c_char_p, []),
    "dvc_prepare_data": (_c.c_int, [_c.c_char_p]),
    "dvc_profile": (_c.c_int, [_c.c_void_p, _c.c_int]),
    "dvc_profile_get": (_c.c_int, [_c.c_void_p, _c.c_char_p, _c.POINTER(_c.c_double), _c.POINTER(_c.c_long)]),
    "dvc_ladder_info": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_void_p, _c.c_void_p, _c.c_void_p]),
    "dvc_ladder_export": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_void_p, _c.c_void_p]),
    "dvc_ldpca_decode": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_void_p, _c.c_void_p, _c.c_uint8,
                                    _c.c_int, _c.c_void_p, _c.POINTER(BpStats)]),
    "dvc_ldpca_decode_batch": (_c.c_int, [_c.c_void_p, _c.c_int, _c.POINTER(LdpcaItem)]),
    "dvc_ldpca_decode_batch_noalloc": (_c.c_int, [_c.c_void_p, _c.c_int, _c.POINTER(LdpcaItem)]),
    "dvc_selftest_unary": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_void_p, _c.c_void_p, _c.c_int]),
    "dvc_autotune": (_c.c_int, [_c.c_void_p, _c.c_int, _c.POINTER(Tuning)]),
    "dvc_get_tuning": (_c.c_int, [_c.c_void_p, _c.POINTER(Tuning)]),
    "dvc_set_tuning": (_c.c_int, [_c.c_void_p, _c.POINTER(Tuning)]),
    "dvc_selftest_phi_sweep": (_c.c_int, [_c.c_void_p, _c.c_uint32, _c.c_uint32, _c.POINTER(_c.c_ulonglong)]),
    "dvc_selftest_valu_rate": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_int, _c.POINTER(_c.c_double),
                                          _c.POINTER(_c.c_double)]),
    "dvc_decode_wz_frames": (_c.c_int, [_c.c_void_p, _c.c_int, _c.POINTER(WZFrameSpec), _c.c_int]),
    "dvc_side_info_batch": (_c.c_int, [_c.c_void_p, _c.c_int, _c.POINTER(SIRequest)]),
    "dvc_wz_stream_open": (_c.c_int, [_c.c_void_p, _c.c_size_t, _c.POINTER(_c.c_void_p)]),
    "dvc_wz_stream_info_get": (_c.c_int, [_c.c_void_p, _c.POINTER(WzStreamInfo)]),
    "dvc_wz_stream_frame_get": (_c.c_int, [_c.c_void_p, _c.c_int, _c.POINTER(WzStreamFrame)]),
    "dvc_wz_stream_close": (_c.c_int, [_c.c_void_p]),
    "dvc_wz_stream_attach_intra": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_size_t]),
    "dvc_intra_parse": (_c.c_int, [_c.c_void_p, _c.c_size_t, _c.POINTER(_c.c_size_t), _c.c_void_p, _c.c_int, _c.c_int,
                                   _c.c_int, _c.c_void_p]),
    "dvc_jbig_decode_map": (_c.c_int, [_c.c_void_p, _c.c_size_t, _c.c_int, _c.c_int, _c.c_void_p,
                                       _c.POINTER(_c.c_size_t)]),
    "dvc_jbig_decode": (_c.c_int, [_c.c_void_p, _c.c_size_t, _c.c_void_p, _c.c_size_t, _c.POINTER(_c.c_uint32),
                                   _c.POINTER(_c.c_uint32), _c.POINTER(_c.c_size_t)]),
    "dvc_buffer_alloc": (_c.c_int, [_c.c_void_p, _c.c_size_t, _c.POINTER(_c.c_void_p)]),
    "dvc_buffer_free": (_c.c_int, [_c.c_void_p, _c.c_void_p]),
    "dvc_buffer_copy": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_size_t, _c.c_int]),
    "dvc_band_frame_init": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_int, _c.c_void_p, _c.c_void_p,
                                       _c.c_int, _c.c_int, _c.c_int, _c.c_void_p, _c.c_int]),
    "dvc_noise_distribution": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_void_p, _c.c_void_p, _c.c_void_p,
                                          _c.c_void_p]),
    "dvc_cond_bit_prob": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_int, _c.c_int, _c.c_int, _c.c_void_p, _c.c_int]),
    "dvc_update_si": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p]),
    "dvc_update_smf": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_void_p]),
    "dvc_forward_me": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_void_p, _c.c_void_p, _c.c_int, _c.c_int,
                                  _c.c_int, _c.c_void_p]),
    "dvc_engine_create": (_c.c_int, [_c.c_void_p, _c.POINTER(EngineParams), _c.POINTER(_c.c_void_p)]),
    "dvc_engine_destroy": (_c.c_int, [_c.c_void_p]),
    "dvc_engine_submit": (_c.c_int, [_c.c_void_p, _c.POINTER(WZFrameSpec), _c.c_uint64, _c.c_uint64,
                                     _c.POINTER(_c.c_uint64)]),
    "dvc_engine_step": (_c.c_int, [_c.c_void_p, _c.POINTER(_c.c_int)]),
    "dvc_engine_status": (_c.c_int, [_c.c_void_p, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int), _c.POINTER(_c.c_uint64)]),
    "dvc_engine_done": (_c.c_int, [_c.c_void_p, _c.c_uint64, _c.POINTER(_c.c_int)]),
    "dvc_engine_release_key": (_c.c_int, [_c.c_void_p, _c.c_uint64]),
}
for _name, (_res, _args) in SIGNATURES.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def _check(rc, where):
    if rc != DVC_OK:
        raise DvcError(rc, where)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def prepare_data(data_dir=None):
    _check(lib.dvc_prepare_data(str(data_dir).encode() if data_dir else None), "dvc_prepare_data")


class DeviceBuffer:
    """A device allocation of the context's device (dvc_buffer_*)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib.dvc_buffer_alloc(ctx.h, self.nbytes, ctypes.byref(p)), "dvc_buffer_alloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, ctx, a):
        a = np.ascontiguousarray(a)
        b = cls(ctx, a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib.dvc_buffer_copy(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes, 0), "dvc_buffer_copy")

    def download(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _check(lib.dvc_buffer_copy(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes, 1), "dvc_buffer_copy")
        return out

    def copy_from(self, other, nbytes=None):
        n = self.nbytes if nbytes is None else int(nbytes)
        _check(lib.dvc_buffer_copy(self.ctx.h, self.ptr, other.ptr, n, 2), "dvc_buffer_copy")

    def free(self):
        if self.ptr and self.ctx.h:
            lib.dvc_buffer_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """Owns a dvc_ctx (device, both ladders and their inverses resident)."""

    def __init__(self, device=0, max_iter=50, bp_threads=0, data_dir=None, parallel_mode=0, rate_step=0,
                 qcif_ladder="reg"):
        """qcif_ladder: "reg" (RegDecoder1584.lad) or "irreg" (irReg1584.lad)."""
        assert qcif_ladder in ("reg", "irreg")
        p = Params(device, max_iter, str(data_dir).encode() if data_dir else None, bp_threads, parallel_mode,
                   rate_step, 1 if qcif_ladder == "irreg" else 0)
        h = ctypes.c_void_p()
        _check(lib.dvc_ctx_create(ctypes.byref(p), ctypes.byref(h)), "dvc_ctx_create")
        self.h = h
        self.max_iter = max_iter

    def close(self):
        if self.h:
            lib.dvc_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def profile(self, enable=True):
        _check(lib.dvc_profile(self.h, 1 if enable else 0), "dvc_profile")

    def profile_get(self, name):
        ms, n = ctypes.c_double(), ctypes.c_long()
        _check(lib.dvc_profile_get(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n)), "dvc_profile_get")
        return ms.value, n.value

    def ladder_info(self, n):
        nc, ne = ctypes.c_int(), ctypes.c_int()
        synd = np.zeros(128, np.int32)
        _check(lib.dvc_ladder_info(self.h, n, ctypes.byref(nc), ctypes.byref(ne), _ptr(synd)), "dvc_ladder_info")
        return nc.value, ne.value, synd[:nc.value].copy()

    def ladder_export(self, n, code):
        nc, ne, _ = self.ladder_info(n)
        ir = np.zeros(ne, np.int32)
        inv = np.zeros(n * 4 * ((n + 31) // 32), np.uint8)
        _check(lib.dvc_ladder_export(self.h, n, code, _ptr(ir), _ptr(inv)), "dvc_ladder_export")
        return ir, inv

    def ldpca_decode(self, lr, acc, crc, start_level=-1, carry=None):
        """Sequential rate search for one bitplane (host arrays).  Returns
        (decoded uint8[N], stats dict).  carry: with parallel_mode 1, the
        decisions carried into the first code (default: none match)."""
        lr = np.ascontiguousarray(lr, np.float32)
        acc = np.ascontiguousarray(acc, np.uint8)
        n = lr.size
        out = np.full(n, 0xAA, np.uint8) if carry is None else np.array(carry, np.uint8)
        st = BpStats()
        _check(lib.dvc_ldpca_decode(self.h, n, _ptr(lr), _ptr(acc), int(crc), int(start_level), _ptr(out),
                                    ctypes.byref(st)), "dvc_ldpca_decode")
        return out, st.as_dict()

    def selftest(self, func, x):
        """Device evaluation of a scalar routine (see dvc_selftest_unary)."""
        x = np.ascontiguousarray(x)
        out = np.zeros(x.size, np.float64 if func in (4, 5) else np.float32)
        _check(lib.dvc_selftest_unary(self.h, func, _ptr(x), _ptr(out), x.size), "dvc_selftest_unary")
        return out

    def autotune(self, planes=0):
        """Times the BP launch variants and keeps the fastest (dvc_autotune)."""
        t = Tuning()
        _check(lib.dvc_autotune(self.h, int(planes), ctypes.byref(t)), "dvc_autotune")
        return t.as_dict()

    def tuning(self):
        t = Tuning()
        _check(lib.dvc_get_tuning(self.h, ctypes.byref(t)), "dvc_get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw):
        t = Tuning()
        _check(lib.dvc_get_tuning(self.h, ctypes.byref(t)), "dvc_get_tuning")
        for k, v in kw.items():
            setattr(t, k, v)
        _check(lib.dvc_set_tuning(self.h, ctypes.byref(t)), "dvc_set_tuning")

    def phi_sweep(self, lo, hi):
        """(compared, mismatches, lowest mismatching pattern) of the BP
        kernel's phi against the direct restatement over [lo, hi]."""
        c = (ctypes.c_ulonglong * 3)()
        _check(lib.dvc_selftest_phi_sweep(self.h, lo, hi, c), "dvc_selftest_phi_sweep")
        return int(c[0]), int(c[1]), int(c[2])

    def decode_wz_frames(self, frames, batched=True):
        """frames: list of dicts with past, fut (uint8 HxW), orig (optional),
        wz (workload.WZFrame), and optional idx=(past, cur, fut).  Returns a
        list of (recons uint8 HxW, stats dict)."""
        k = len(frames)
        specs = (WZFrameSpec * max(k, 1))()
        stats = [FrameStats() for _ in range(k)]
        keep, outs = [], []
        for i, fr in enumerate(frames):
            past = np.ascontiguousarray(fr["past"], np.uint8)
            fut = np.ascontiguousarray(fr["fut"], np.uint8)
            orig = None if fr.get("orig") is None else np.ascontiguousarray(fr["orig"], np.uint8)
            wz = fr["wz"]
            acr = np.ascontiguousarray(wz.ac_range, np.int32)
            crc = np.ascontiguousarray(wz.crc, np.uint8)
            acc = np.ascontiguousarray(wz.acc, np.uint8)
            rec = np.zeros_like(past)
            mf, mkeep = _mode_fields(wz)
            keep += [past, fut, orig, acr, crc, acc] + mkeep
            outs.append(rec)
            pi, ci, fi = fr.get("idx", (0, 1, 2))
            h, w = past.shape
            specs[i] = WZFrameSpec(w, h, wz.qindex, pi, ci, fi, past.ctypes.data, fut.ctypes.data,
                                   None if orig is None else orig.ctypes.data, acr.ctypes.data, crc.ctypes.data,
                                   acc.ctypes.data, rec.ctypes.data, ctypes.pointer(stats[i]), *mf)
        _check(lib.dvc_decode_wz_frames(self.h, k, specs, 1 if batched else 0), "dvc_decode_wz_frames")
        return [(outs[i], stats[i].as_dict()) for i in range(k)]

    def side_info_batch(self, pairs, qindex=8):
        """pairs: list of (past, fut) uint8 HxW key frames, or of dicts with
        past, fut and optional idx=(past, cur, fut) and qindex.  Returns a list
        of (si uint8 HxW, residue int32 HxW, alpha float32 HxW)."""
        k = len(pairs)
        reqs = (SIRequest * max(k, 1))()
        keep, outs = [], []
        for i, pr in enumerate(pairs):
            d = pr if isinstance(pr, dict) else dict(past=pr[0], fut=pr[1])
            past = np.ascontiguousarray(d["past"], np.uint8)
            fut = np.ascontiguousarray(d["fut"], np.uint8)
            h, w = past.shape
            si = np.zeros((h, w), np.uint8)
            res = np.zeros((h, w), np.int32)
            al = np.zeros((h, w), np.float32)
            keep += [past, fut]
            outs.append((si, res, al))
            pi, ci, fi = d.get("idx", (0, 1, 2))
            reqs[i] = SIRequest(w, h, int(d.get("qindex", qindex)), pi, ci, fi, past.ctypes.data, fut.ctypes.data,
                                si.ctypes.data, res.ctypes.data, al.ctypes.data)
        _check(lib.dvc_side_info_batch(self.h, k, reqs), "dvc_side_info_batch")
        return outs

    def ldpca_decode_batch(self, lrs, accs, crcs, start_levels=None, carries=None):
        """Speculative batched decode of many bitplanes (mixed N allowed).
        carries: with parallel_mode 1, per bitplane the decisions carried into
        its first code (default: none match)."""
        k = len(lrs)
        items = (LdpcaItem * max(k, 1))()
        keep, outs = [], []
        for i in range(k):
            lr = np.ascontiguousarray(lrs[i], np.float32)
            acc = np.ascontiguousarray(accs[i], np.uint8)
            out = np.full(lr.size, 0xAA, np.uint8) if carries is None else np.array(carries[i], np.uint8)
            keep += [lr, acc]
            outs.append(out)
            sl = -1 if start_levels is None else int(start_levels[i])
            items[i] = LdpcaItem(lr.size, lr.ctypes.data, acc.ctypes.data, int(crcs[i]), sl, out.ctypes.data, BpStats())
        _check(lib.dvc_ldpca_decode_batch(self.h, k, items), "dvc_ldpca_decode_batch")
        return outs, [items[i].stats.as_dict() for i in range(k)]


# Reference data layouts of the per-band entry points (DvcDec/inc/global.h):
# SMF, SearchInfo and MV records.
SMF_DTYPE = np.dtype([("prob", "<f4", 81), ("searchTop", "<i4"), ("searchLeft", "<i4"), ("searchRight", "<i4"),
                      ("searchBottom", "<i4"), ("searchWidth", "<i4"), ("blkIdx", "<i4"), ("probLen", "<i4")])
SEARCH_INFO_DTYPE = np.dtype([(n, "<i4") for n in ("searchTop", "searchBottom", "searchLeft", "searchRight", "top",
                                                     "left", "searchWidth", "searchRange", "blkIdx")])
MV_DTYPE = np.dtype([("past", "<f8", 2), ("future", "<f8", 2), ("vector", "<f8", 2)])


class BandSession:
    """The per-band entry points (dvc_band_frame_init, dvc_noise_distribution,
    dvc_cond_bit_prob, dvc_update_si, dvc_update_smf) on one WZ frame, with
    numpy arrays in the reference's layouts."""

    def __init__(self, ctx, past, fut, qindex, idx=(0, 1, 2), sir=None):
        self.ctx = ctx
        past = np.ascontiguousarray(past, np.uint8)
        fut = np.ascontiguousarray(fut, np.uint8)
        self.h, self.w = past.shape
        sp = None if sir is None else np.ascontiguousarray(sir, SEARCH_INFO_DTYPE)
        _check(lib.dvc_band_frame_init(ctx.h, self.w, self.h, qindex, past.ctypes.data, fut.ctypes.data, idx[0],
                                       idx[1], idx[2], None if sp is None else sp.ctypes.data,
                                       0 if sp is None else sp.size), "dvc_band_frame_init")

    def noise_distribution(self, rng, index, alpha, smf, oct_, want=True):
        alpha = np.ascontiguousarray(alpha, np.float32)
        smf = np.ascontiguousarray(smf, SMF_DTYPE)
        oct_ = np.ascontiguousarray(oct_, np.int32)
        n = smf.size
        per = 2 * rng + 1 if index else rng
        out = np.zeros((n, per), np.float32) if want else None
        _check(lib.dvc_noise_distribution(self.ctx.h, rng, index, alpha.ctypes.data, smf.ctypes.data, oct_.ctypes.data,
                                          None if out is None else out.ctypes.data), "dvc_noise_distribution")
        return out

    def cond_bit_prob(self, index, bit, step, recons_trans, rng):
        rt = np.ascontiguousarray(recons_trans, np.int32)
        lr = np.zeros(self.w * self.h // 16, np.float32)
        _check(lib.dvc_cond_bit_prob(self.ctx.h, lr.ctypes.data, index, bit, step, rt.ctypes.data, rng),
               "dvc_cond_bit_prob")
        return lr

    def update_si(self, si, recons):
        si = np.array(si, np.uint8, copy=True, order="C")
        recons = np.ascontiguousarray(recons, np.uint8)
        _check(lib.dvc_update_si(self.ctx.h, si.ctypes.data, recons.ctypes.data), "dvc_update_si")
        return si

    def update_smf(self, smf, recons, si):
        smf = np.array(smf, SMF_DTYPE, copy=True)
        recons = np.ascontiguousarray(recons, np.uint8)
        si = np.ascontiguousarray(si, np.uint8)
        _check(lib.dvc_update_smf(self.ctx.h, smf.ctypes.data, recons.ctypes.data, si.ctypes.data), "dvc_update_smf")
        return smf


def forward_me(ctx, fut_lp, past_lp, idx=(0, 1, 2)):
    """dvc_forward_me: macroblock MVs (MV_DTYPE records) of the WZ frame."""
    fut_lp = np.ascontiguousarray(fut_lp, np.uint8)
    past_lp = np.ascontiguousarray(past_lp, np.uint8)
    h, w = fut_lp.shape
    out = np.zeros((h // 16) * (w // 16), MV_DTYPE)
    _check(lib.dvc_forward_me(ctx.h, w, h, fut_lp.ctypes.data, past_lp.ctypes.data, idx[0], idx[1], idx[2],
                              out.ctypes.data), "dvc_forward_me")
    return out


def _frame_spec(fr, stats):
    """Builds a WZFrameSpec; returns (spec, arrays to keep alive, recons)."""
    past = np.ascontiguousarray(fr["past"], np.uint8)
    fut = np.ascontiguousarray(fr["fut"], np.uint8)
    orig = None if fr.get("orig") is None else np.ascontiguousarray(fr["orig"], np.uint8)
    wz = fr["wz"]
    acr = np.ascontiguousarray(wz.ac_range, np.int32)
    crc = np.ascontiguousarray(wz.crc, np.uint8)
    acc = np.ascontiguousarray(wz.acc, np.uint8)
    rec = np.zeros_like(past)
    pi, ci, fi = fr.get("idx", (0, 1, 2))
    h, w = past.shape
    mf, mkeep = _mode_fields(wz)
    spec = WZFrameSpec(w, h, wz.qindex, pi, ci, fi, past.ctypes.data, fut.ctypes.data,
                       None if orig is None else orig.ctypes.data, acr.ctypes.data, crc.ctypes.data,
                       acc.ctypes.data, rec.ctypes.data, ctypes.pointer(stats), *mf)
    return spec, [past, fut, orig, acr, crc, acc, rec] + mkeep, rec


def _block_mode(wz):
    m = getattr(wz, "block_mode", None)
    return None if m is None else np.ascontiguousarray(m, np.uint8)


def _mode_fields(wz):
    """(ctypes field values from block_mode on, arrays to keep alive) of a WZFrame."""
    mode = _block_mode(wz)
    ic = getattr(wz, "intra_coeffs", None)
    ic = None if ic is None else np.ascontiguousarray(ic, np.int32)
    pm = getattr(wz, "past_block_mode", None)
    pm = None if pm is None else np.ascontiguousarray(pm, np.uint8)
    vals = (None if mode is None else mode.ctypes.data, int(getattr(wz, "mode_bits", 0)),
            None if ic is None else ic.ctypes.data, int(getattr(wz, "deblocking", 0)), int(getattr(wz, "intra_qp", 0)),
            int(getattr(wz, "past_is_wz", 0)), None if pm is None else pm.ctypes.data)
    return vals, [mode, ic, pm]


class Engine:
    """Continuous-batching decoder (dvc_engine_*): submit frames at any time,
    step, and collect results as their tickets complete."""

    def __init__(self, ctx, max_inflight=64, hbm_budget_bytes=0, sequential_bp=False, lanes=0, eager_steps=False,
                 device_buffers=False, lane_steps=0):
        p = EngineParams(int(hbm_budget_bytes), int(max_inflight), 1 if sequential_bp else 0, int(lanes),
                         1 if eager_steps else 0, 1 if device_buffers else 0, int(lane_steps))
        h = ctypes.c_void_p()
        _check(lib.dvc_engine_create(ctx.h, ctypes.byref(p), ctypes.byref(h)), "dvc_engine_create")
        self.h = h
        self.ctx = ctx
        self.live = {}

    def submit(self, fr, past_key=0, fut_key=0):
        stats = FrameStats()
        spec, keep, rec = _frame_spec(fr, stats)
        t = ctypes.c_uint64()
        _check(lib.dvc_engine_submit(self.h, ctypes.byref(spec), past_key, fut_key, ctypes.byref(t)), "dvc_engine_submit")
        self.live[t.value] = (spec, keep, rec, stats)
        return t.value

    def submit_device(self, width, height, wz, past_ptr, fut_ptr, acc_ptr, recons_ptr, idx=(0, 1, 2),
                      past_key=0, fut_key=0):
        """Submits a frame whose key frames, accumulated syndromes and output
        live in device memory (an engine made with device_buffers=True); the
        pointers are integers (e.g. torch tensors' data_ptr()).  ac_range and
        crc come from the host-side workload.WZFrame `wz`."""
        stats = FrameStats()
        acr = np.ascontiguousarray(wz.ac_range, np.int32)
        crc = np.ascontiguousarray(wz.crc, np.uint8)
        mf, mkeep = _mode_fields(wz)
        spec = WZFrameSpec(width, height, wz.qindex, idx[0], idx[1], idx[2], past_ptr, fut_ptr, None,
                           acr.ctypes.data, crc.ctypes.data, acc_ptr, recons_ptr, ctypes.pointer(stats), *mf)
        t = ctypes.c_uint64()
        _check(lib.dvc_engine_submit(self.h, ctypes.byref(spec), past_key, fut_key, ctypes.byref(t)), "dvc_engine_submit")
        self.live[t.value] = (spec, [acr, crc] + mkeep, None, stats)
        return t.value

    def step(self):
        n = ctypes.c_int()
        _check(lib.dvc_engine_step(self.h, ctypes.byref(n)), "dvc_engine_step")
        return n.value

    def done(self, ticket):
        d = ctypes.c_int()
        _check(lib.dvc_engine_done(self.h, ticket, ctypes.byref(d)), "dvc_engine_done")
        return bool(d.value)

    def result(self, ticket):
        _, _, rec, stats = self.live.pop(ticket)
        return rec, stats.as_dict()

    def status(self):
        q, i, u = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint64()
        _check(lib.dvc_engine_status(self.h, ctypes.byref(q), ctypes.byref(i), ctypes.byref(u)), "dvc_engine_status")
        return q.value, i.value, u.value

    def decode_all(self, frames):
        """Submits every frame, steps until all are done, and returns their
        (recons, stats) in submission order."""
        tickets = [self.submit(fr) for fr in frames]
        pending = set(tickets)
        while pending:
            self.step()
            for t in list(pending):
                if self.done(t):
                    pending.discard(t)
        return [self.result(t) for t in tickets]

    def release_key(self, key):
        _check(lib.dvc_engine_release_key(self.h, key), "dvc_engine_release_key")

    def close(self):
        if self.h:
            lib.dvc_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def jbig_decode_map(data, width, height):
    """One block-mode map's JBIG data -> (uint8[height, width] of 0/1, bytes consumed)."""
    buf = np.frombuffer(bytes(data), np.uint8)
    out = np.zeros((height, width), np.uint8)
    used = ctypes.c_size_t()
    _check(lib.dvc_jbig_decode_map(buf.ctypes.data, len(buf), width, height, out.ctypes.data, ctypes.byref(used)),
           "dvc_jbig_decode_map")
    return out, used.value


def jbig_decode(bie):
    """A JBIG BIE (header + data) -> (uint8[yd, xd] image, bytes consumed)."""
    buf = np.frombuffer(bytes(bie), np.uint8)
    w, h, used = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_size_t()
    _check(lib.dvc_jbig_decode(buf.ctypes.data, len(buf), None, 0, ctypes.byref(w), ctypes.byref(h), ctypes.byref(used)),
           "dvc_jbig_decode")
    out = np.zeros((h.value, w.value), np.uint8)
    _check(lib.dvc_jbig_decode(buf.ctypes.data, len(buf), out.ctypes.data, out.size, ctypes.byref(w), ctypes.byref(h),
                               ctypes.byref(used)), "dvc_jbig_decode")
    return out, used.value


def read_wz_stream(data, intra=None):
    """Parses a WZ bitstream (dvc_wz_stream_*) and, when given, its intra bitstream: (info dict, list of
    per-WZ-frame dicts in decoding order with past/cur/fut indices and a workload.WZFrame carrying every field
    dvc_wz_frame takes from the stream)."""
    from .workload import WZFrame
    buf = np.frombuffer(bytes(data), np.uint8)
    h = ctypes.c_void_p()
    _check(lib.dvc_wz_stream_open(buf.ctypes.data, len(buf), ctypes.byref(h)), "dvc_wz_stream_open")
    try:
        if intra is not None:
            ib = np.frombuffer(bytes(intra) or b"\0", np.uint8)
            _check(lib.dvc_wz_stream_attach_intra(h, ib.ctypes.data, len(bytes(intra))), "dvc_wz_stream_attach_intra")
        info = WzStreamInfo()
        _check(lib.dvc_wz_stream_info_get(h, ctypes.byref(info)), "dvc_wz_stream_info_get")
        n8 = (info.width // 8) * (info.height // 8)
        synd = (info.width // 4) * (info.height // 4) // 8
        frames = []
        for k in range(info.wz_frames):
            f = WzStreamFrame()
            _check(lib.dvc_wz_stream_frame_get(h, k, ctypes.byref(f)), "dvc_wz_stream_frame_get")
            mode = np.ctypeslib.as_array((ctypes.c_uint8 * n8).from_address(f.block_mode)).copy()
            crc = np.ctypeslib.as_array((ctypes.c_uint8 * info.nbitplanes).from_address(f.crc)).copy()
            acc = np.ctypeslib.as_array((ctypes.c_uint8 * (info.nbitplanes * synd)).from_address(f.acc)).copy()
            nblk = (info.width // 4) * (info.height // 4)
            ic = (np.ctypeslib.as_array((ctypes.c_int32 * (16 * nblk)).from_address(f.intra_coeffs)).copy()
                  .reshape(nblk, 16) if f.intra_coeffs else None)
            pm = (np.ctypeslib.as_array((ctypes.c_uint8 * n8).from_address(f.past_block_mode)).copy()
                  if f.past_block_mode else None)
            wz = WZFrame(info.qindex, np.array(list(f.ac_range), np.int32), crc, acc.reshape(info.nbitplanes, synd),
                         None, mode if mode.any() else None, f.mode_bits, ic, info.deblocking, info.intra_qp,
                         f.past_is_wz, pm)
            frames.append(dict(idx=(f.past_idx, f.cur_idx, f.fut_idx), skip_mode=f.skip_mode, intra_mode=f.intra_mode,
                               wz=wz))
        return info.as_dict(), frames
    finally:
        lib.dvc_wz_stream_close(h)
