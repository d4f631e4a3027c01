"""ctypes binding of libsteppe_hip.so — the Python face of the C ABI.

Mirrors the reference's nanobind module ``steppe._core`` for the f-statistics
path (bindings/bind_f2handle.cpp:14-47, bindings/bind_fstats.cpp:131-215):
``F2Handle``, ``read_f2``, ``run_extract_f2`` and the fit entry points, with the
same argument names and error behaviour (ValueError for invalid input or device
faults, KeyError for unknown population names).

There is no CPU fallback: importing works anywhere, but every call needs the
built HIP library and a visible GPU, and raises when either is missing.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsteppe_hip.so")

STP_PREC_FP64 = 0
STP_PREC_SPLIT_BF16 = 1

_PRECISIONS = {
    None: STP_PREC_SPLIT_BF16,
    "split_bf16": STP_PREC_SPLIT_BF16, "split": STP_PREC_SPLIT_BF16,
    "emu": STP_PREC_SPLIT_BF16, "emu40": STP_PREC_SPLIT_BF16, "emulated_fp64": STP_PREC_SPLIT_BF16,
    "fp64": STP_PREC_FP64, "native": STP_PREC_FP64,
}
PRECISION_NAMES = {STP_PREC_FP64: "fp64", STP_PREC_SPLIT_BF16: "split_bf16"}


class Tile(C.Structure):
    """stp_tile (include/steppe_hip.h) — the reference's DecodeTileView."""
    _fields_ = [
        ("packed", C.c_void_p),
        ("bytes_per_record", C.c_size_t),
        ("n_snp", C.c_size_t),
        ("n_individuals", C.c_size_t),
        ("pop_offsets", C.c_void_p),
        ("n_pop", C.c_int32),
        ("sample_ploidy", C.c_void_p),
        ("ploidy", C.c_int32),
        ("detect_ploidy_on_device", C.c_int32),
        ("packed_on_device", C.c_int32),
    ]


QPADM_MAX_NL = 8
QPADM_MAX_RANK = QPADM_MAX_NL
QPADM_MAX_POPDROP = QPADM_MAX_NL + 1
STATUS_NAMES = {0: "ok", 1: "device_oom", 2: "rank_deficient", 3: "non_spd_covariance",
                4: "chisq_undefined", 5: "invalid_config"}


class QpAdmResult(C.Structure):
    """stp_qpadm_result (include/stp_qpadm_result.h)."""
    _fields_ = [
        ("status", C.c_int32), ("model_index", C.c_int32), ("nl", C.c_int32), ("dof", C.c_int32),
        ("est_rank", C.c_int32), ("f4rank", C.c_int32), ("n_rank", C.c_int32), ("n_popdrop", C.c_int32),
        ("p", C.c_double), ("chisq", C.c_double),
        ("weight", C.c_double * QPADM_MAX_NL), ("se", C.c_double * QPADM_MAX_NL),
        ("z", C.c_double * QPADM_MAX_NL),
        ("rank_chisq", C.c_double * QPADM_MAX_RANK), ("rank_dof", C.c_int32 * QPADM_MAX_RANK),
        ("rankdrop_f4rank", C.c_int32 * QPADM_MAX_RANK), ("rankdrop_dof", C.c_int32 * QPADM_MAX_RANK),
        ("rankdrop_dofdiff", C.c_int32 * QPADM_MAX_RANK),
        ("rankdrop_chisq", C.c_double * QPADM_MAX_RANK), ("rankdrop_p", C.c_double * QPADM_MAX_RANK),
        ("rankdrop_chisqdiff", C.c_double * QPADM_MAX_RANK),
        ("rankdrop_p_nested", C.c_double * QPADM_MAX_RANK),
        ("popdrop_pat", (C.c_char * (QPADM_MAX_NL + 1)) * QPADM_MAX_POPDROP),
        ("popdrop_wt", C.c_int32 * QPADM_MAX_POPDROP), ("popdrop_dof", C.c_int32 * QPADM_MAX_POPDROP),
        ("popdrop_f4rank", C.c_int32 * QPADM_MAX_POPDROP),
        ("popdrop_feasible", C.c_int32 * QPADM_MAX_POPDROP),
        ("popdrop_chisq", C.c_double * QPADM_MAX_POPDROP), ("popdrop_p", C.c_double * QPADM_MAX_POPDROP),
    ]

    def to_dict(self, precision: str = "fp64") -> dict:
        """The reference's result dict (bindings/internal/bind_common.hpp:179-212)."""
        nl, nk, npd = self.nl, self.n_rank, self.n_popdrop
        d = {
            "weight": list(self.weight[:nl]), "se": list(self.se[:nl]), "z": list(self.z[:nl]),
            "p": self.p, "chisq": self.chisq, "dof": self.dof, "est_rank": self.est_rank,
            "f4rank": self.f4rank,
            "rankdrop_f4rank": list(self.rankdrop_f4rank[:nk]), "rankdrop_dof": list(self.rankdrop_dof[:nk]),
            "rankdrop_dofdiff": list(self.rankdrop_dofdiff[:nk]),
            "rankdrop_chisq": list(self.rankdrop_chisq[:nk]), "rankdrop_p": list(self.rankdrop_p[:nk]),
            "rankdrop_chisqdiff": list(self.rankdrop_chisqdiff[:nk]),
            "rankdrop_p_nested": list(self.rankdrop_p_nested[:nk]),
            "popdrop_pat": [self.popdrop_pat[k].value.decode() for k in range(npd)],
            "popdrop_wt": list(self.popdrop_wt[:npd]), "popdrop_dof": list(self.popdrop_dof[:npd]),
            "popdrop_f4rank": list(self.popdrop_f4rank[:npd]),
            "popdrop_chisq": list(self.popdrop_chisq[:npd]), "popdrop_p": list(self.popdrop_p[:npd]),
            "popdrop_feasible": [bool(v) for v in self.popdrop_feasible[:npd]],
            "status": STATUS_NAMES.get(self.status, str(self.status)), "precision": precision,
            "model_index": self.model_index,
        }
        return d


class QpAdmOptions(C.Structure):
    """stp_qpadm_options (include/steppe_hip.h); reference: steppe::QpAdmOptions."""
    _fields_ = [("fudge", C.c_double), ("als_iterations", C.c_int32), ("rank", C.c_int32),
                ("rank_alpha", C.c_double), ("jackknife", C.c_int32), ("se_require_p", C.c_int32),
                ("p_se_threshold", C.c_double)]


class ExtractRequest(C.Structure):
    """stp_extract_request (include/steppe_hip.h)."""
    _fields_ = [("geno", C.c_char_p), ("snp", C.c_char_p), ("ind", C.c_char_p), ("pop_mode", C.c_int32),
                ("pops", C.POINTER(C.c_char_p)), ("n_pops", C.c_int32), ("top_k", C.c_int64), ("min_n", C.c_int64),
                ("maf_min", C.c_double), ("maxmiss", C.c_double), ("autosomes_only", C.c_int32),
                ("drop_monomorphic", C.c_int32), ("transversions_only", C.c_int32), ("strand_mode", C.c_int32),
                ("blgsize", C.c_double), ("ploidy", C.c_int32), ("precision", C.c_int32), ("tier", C.c_int32),
                ("disk_path", C.c_char_p), ("chunk_blocks", C.c_int64)]


class ExtractInfo(C.Structure):
    _fields_ = [("n_snp_total", C.c_int64), ("n_snp_kept", C.c_int64), ("n_pseudo_haploid", C.c_int64),
                ("n_diploid", C.c_int64), ("P", C.c_int32), ("n_block", C.c_int32)]


class SweepRequest(C.Structure):
    """stp_sweep_request (include/steppe_hip.h); reference: steppe::SweepRequest."""
    _fields_ = [("k", C.c_int32), ("filter", C.c_int32), ("min_z", C.c_double), ("top_k", C.c_int64),
                ("sure", C.c_int32), ("method", C.c_int32), ("subset", C.c_void_p), ("n_subset", C.c_int32)]


# The reference's jackknife policy names (bindings/internal/bind_common.hpp:168-178).
JACKKNIFE = {"none": 0, "feasible": 1, "feasible_only": 1, "all": 2}


_lib = None


def lib():
    """Load the engine; raise loudly when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"steppe_amd: HIP engine not built ({LIB_PATH} missing); "
                           "run `python -m steppe_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, sz, dp = C.c_void_p, C.c_int32, C.c_int64, C.c_size_t, C.c_double
    sig = {
        "stp_open": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "stp_close": (C.c_int, [vp]),
        "stp_last_error": (C.c_char_p, []),
        "stp_device_info": (C.c_int, [vp, C.c_char_p, C.c_int, C.POINTER(i32)]),
        "stp_detect_ploidy": (C.c_int, [vp, C.POINTER(Tile), vp]),
        "stp_decode_counts": (C.c_int, [vp, C.POINTER(Tile), vp, vp, i64]),
        "stp_decode_counts_noalloc": (C.c_int, [vp, C.POINTER(Tile), vp, vp, i64]),
        "stp_decode_af": (C.c_int, [vp, C.POINTER(Tile), vp, vp, vp]),
        "stp_f2_blocks_from_counts": (C.c_int, [vp, vp, vp, i64, i32, i64, vp, i32, i32, vp, vp]),
        "stp_compute_f2_blocks": (C.c_int, [vp, vp, vp, vp, i32, i64, vp, i32, i32, vp, vp, vp]),
        "stp_f2_autotune": (C.c_int, [vp, vp, vp, i64, i32, i64, vp, i32, i32, vp, vp, C.c_char_p, i32,
                                      C.POINTER(i32)]),
        "stp_extract_f2": (C.c_int, [vp, C.POINTER(Tile), vp, i32, i32, C.POINTER(vp)]),
        "stp_extract_f2_tiered": (C.c_int, [vp, C.POINTER(Tile), vp, i32, i32, i32, C.c_char_p, i64, C.POINTER(vp)]),
        "stp_f2_graph_create": (C.c_int, [vp, C.POINTER(Tile), vp, i32, i32, vp, vp, i64, vp, vp, C.POINTER(vp)]),
        "stp_f2_graph_launch": (C.c_int, [vp]),
        "stp_f2_graph_free": (C.c_int, [vp]),
        "stp_f2_handle_tier": (C.c_int, [vp, C.POINTER(i32)]),
        "stp_f2_handle_upload": (C.c_int, [vp, vp, vp, vp, i32, i32, C.POINTER(vp)]),
        "stp_f2_handle_shape": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i32)]),
        "stp_f2_handle_block_sizes": (C.c_int, [vp, vp]),
        "stp_f2_handle_f2_device": (vp, [vp]),
        "stp_f2_handle_vpair_device": (vp, [vp]),
        "stp_f2_handle_download": (C.c_int, [vp, vp, vp]),
        "stp_f2_handle_free": (C.c_int, [vp]),
        "stp_f2bin_write": (C.c_int, [C.c_char_p, vp, vp, vp, i32, i32]),
        "stp_f2bin_read_header": (C.c_int, [C.c_char_p, C.POINTER(i32), C.POINTER(i32)]),
        "stp_f2bin_read": (C.c_int, [C.c_char_p, vp, vp, vp]),
        "stp_qpadm_batch": (C.c_int, [vp, vp, i32, vp, vp, vp, vp, vp, C.POINTER(QpAdmOptions),
                                      C.POINTER(QpAdmResult)]),
        "stp_extract_f2_files": (C.c_int, [vp, C.POINTER(ExtractRequest), C.POINTER(vp), C.POINTER(ExtractInfo)]),
        "stp_f2_handle_label": (C.c_int, [vp, i32, C.c_char_p, i32, C.POINTER(i32)]),
        "stp_f4": (C.c_int, [vp, vp, vp, i64, vp, vp, vp, vp]),
        "stp_f3": (C.c_int, [vp, vp, vp, i64, vp, vp, vp, vp]),
        "stp_f4ratio": (C.c_int, [vp, vp, vp, i64, vp, vp, vp]),
        "stp_fstat_sweep": (C.c_int, [vp, vp, C.POINTER(SweepRequest), i64, vp, vp, vp, vp, vp, C.POINTER(i64),
                                      C.POINTER(C.c_uint64), C.POINTER(i32), C.POINTER(i32)]),
        "stpx_f2_tier_choice": (C.c_int, [i32, i64, vp, i32, i32, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64),
                                          C.POINTER(i32)]),
        "stpx_timing": (C.c_int, [vp, i32]),
        "stpx_timing_read": (C.c_int, [vp, vp, vp, i32]),
        "stpx_synth_panel": (C.c_int, [vp, C.c_uint64, i32, i64, i64, i64, vp, vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    del sz, dp
    _lib = L
    return L


def check(rc: int, what: str = ""):
    if rc == 0:
        return
    msg = lib().stp_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == -3:
        raise MemoryError(text)
    if rc in (-4, -5):
        raise ValueError(text)
    raise ValueError(text)


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


class Context:
    """One engine context per HIP device (the reference's per-GPU backend)."""
    _by_device: dict = {}

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().stp_open(int(device), C.byref(h)), f"open device {device}")
        self.h = h
        self.device = int(device)

    @classmethod
    def get(cls, device: int = 0) -> "Context":
        ctx = cls._by_device.get(int(device))
        if ctx is None:
            ctx = cls(device)
            cls._by_device[int(device)] = ctx
        return ctx

    STAGES = ("decode", "prep", "f2", "count", "qpadm", "qpgraph", "fstat", "frontend", "fixup")

    def timing(self, enable: bool = True):
        check(lib().stpx_timing(self.h, 1 if enable else 0))

    def read_timing(self, reset: bool = False) -> dict:
        ms = np.zeros(len(self.STAGES), np.float64)
        cnt = np.zeros(len(self.STAGES), np.int64)
        check(lib().stpx_timing_read(self.h, _ptr(ms), _ptr(cnt), 1 if reset else 0))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(self.STAGES)}

    def info(self):
        buf = C.create_string_buffer(64)
        n = C.c_int32()
        check(lib().stp_device_info(self.h, buf, 64, C.byref(n)))
        return buf.value.decode(), int(n.value)


def make_tile(packed, n_snp: int, pop_offsets, sample_ploidy=None, ploidy: int = 2,
              detect: bool = False, bytes_per_record: int | None = None,
              n_individuals: int | None = None, device_ptr: int | None = None):
    """Build a Tile.  ``packed`` is a uint8 [n_ind, bytes_per_record] host array, or
    None with ``device_ptr`` (plus bytes_per_record, n_individuals) for an
    HBM-resident panel.  Returns (tile, keepalive)."""
    keep = []
    offs = np.ascontiguousarray(pop_offsets, dtype=np.uint64)
    keep.append(offs)
    t = Tile()
    if device_ptr is not None:
        t.packed = C.c_void_p(int(device_ptr))
        t.bytes_per_record = int(bytes_per_record)
        t.n_individuals = int(n_individuals)
        t.packed_on_device = 1
    else:
        arr = np.ascontiguousarray(packed, dtype=np.uint8)
        keep.append(arr)
        t.packed = _ptr(arr)
        t.bytes_per_record = int(arr.shape[1]) if bytes_per_record is None else int(bytes_per_record)
        t.n_individuals = int(arr.shape[0])
        t.packed_on_device = 0
    t.n_snp = int(n_snp)
    t.pop_offsets = _ptr(offs)
    t.n_pop = int(offs.shape[0] - 1)
    if sample_ploidy is not None:
        sp = np.ascontiguousarray(sample_ploidy, dtype=np.int32)
        keep.append(sp)
        t.sample_ploidy = _ptr(sp)
    else:
        t.sample_ploidy = None
    t.ploidy = int(ploidy)
    t.detect_ploidy_on_device = 1 if detect else 0
    return t, keep


def tile_from_panel(panel, ploidy="sample"):
    if ploidy == "sample":
        return make_tile(panel.packed, panel.n_snp, panel.pop_offsets, panel.sample_ploidy)
    if ploidy == "auto":
        return make_tile(panel.packed, panel.n_snp, panel.pop_offsets, None, detect=True)
    return make_tile(panel.packed, panel.n_snp, panel.pop_offsets, None, ploidy=int(ploidy))


def f2_autotune(ctx: "Context", a2: int, n: int, ld: int, P: int, M: int, block_id, n_block: int, precision,
                f2: int, vpair: int) -> dict:
    """Time every launch configuration of the f2 contraction on these device counts
    (stp_f2_autotune) and keep the fastest for the workload; returns the sweep table."""
    bid = np.ascontiguousarray(block_id, dtype=np.int32)
    need = C.c_int32(0)
    buf = C.create_string_buffer(1 << 16)
    check(lib().stp_f2_autotune(ctx.h, C.c_void_p(int(a2)), C.c_void_p(int(n)), int(ld), int(P), int(M), _ptr(bid),
                                int(n_block), parse_precision(precision), C.c_void_p(int(f2)), C.c_void_p(int(vpair)),
                                buf, len(buf), C.byref(need)), "f2 autotune")
    return json.loads(buf.value.decode())


class F2Graph:
    """A resident extraction of one fixed shape captured as a HIP graph
    (stp_f2_graph_create): ``launch()`` replays ploidy detection, decode, prep, pair
    counts, contraction and fix-up into the device buffers given at creation.

    ``tile`` must be device-resident; ``block_id`` is the SNP -> block map; ``f2`` and
    ``vpair`` are device pointers to [P x P x n_block] FP64 buffers; ``a2`` / ``n`` are
    device pointers to uint16 [P x ld] counts, or None for counts held by the graph."""

    def __init__(self, ctx: "Context", tile, block_id, n_block: int, precision, f2: int, vpair: int,
                 a2: int | None = None, n: int | None = None, ld: int = 0):
        self.ctx = ctx
        self._bid = np.ascontiguousarray(block_id, dtype=np.int32)
        self._tile = tile
        h = C.c_void_p()
        check(lib().stp_f2_graph_create(ctx.h, C.byref(tile), _ptr(self._bid), int(n_block), parse_precision(precision),
                                        C.c_void_p(a2) if a2 else None, C.c_void_p(n) if n else None, int(ld),
                                        C.c_void_p(int(f2)), C.c_void_p(int(vpair)), C.byref(h)), "f2 graph create")
        self.h = h

    def launch(self):
        check(lib().stp_f2_graph_launch(self.h), "f2 graph launch")

    def close(self):
        if getattr(self, "h", None):
            lib().stp_f2_graph_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class F2Handle:
    """Resident f2 tensor plus population labels (bindings/bind_f2handle.cpp:24-47).

    Unlike the reference's handle, which keeps a host tensor and uploads it on every
    fit, this one keeps the tensor resident in HBM; fits read it in place."""

    def __init__(self, raw: C.c_void_p, pops, device: int):
        self._h = raw
        self._pops = list(pops)
        self._device = int(device)
        P, nb = C.c_int32(), C.c_int32()
        check(lib().stp_f2_handle_shape(raw, C.byref(P), C.byref(nb)))
        self._P, self._nb = int(P.value), int(nb.value)
        if len(self._pops) != self._P:
            raise ValueError(f"pops has {len(self._pops)} labels but P = {self._P}")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.stp_f2_handle_free(h)
            self._h = None

    @property
    def pops(self):
        return list(self._pops)

    @property
    def P(self) -> int:
        return self._P

    @property
    def n_block(self) -> int:
        return self._nb

    @property
    def device(self) -> int:
        return self._device

    @property
    def block_sizes(self):
        out = np.zeros(self._nb, np.int32)
        check(lib().stp_f2_handle_block_sizes(self._h, _ptr(out)))
        return out.tolist()

    @property
    def raw(self):
        return self._h

    @property
    def tier(self) -> str:
        """Where the tensor lives: "resident" (HBM), "host" or "disk" (an f2.bin file)."""
        t = C.c_int32()
        check(lib().stp_f2_handle_tier(self._h, C.byref(t)))
        return TIER_NAMES[int(t.value)]

    def device_pointers(self):
        L = lib()
        return int(L.stp_f2_handle_f2_device(self._h) or 0), int(L.stp_f2_handle_vpair_device(self._h) or 0)

    def _download(self, which: str) -> np.ndarray:
        n = self._P * self._P * self._nb
        out = np.empty(n, np.float64)
        if which == "f2":
            check(lib().stp_f2_handle_download(self._h, _ptr(out), None))
        else:
            check(lib().stp_f2_handle_download(self._h, None, _ptr(out)))
        return out.reshape((self._P, self._P, self._nb), order="F")

    def _f2_numpy(self) -> np.ndarray:
        """f2 as float64, F-contiguous (P, P, n_block). COPY."""
        return self._download("f2")

    def _vpair_numpy(self) -> np.ndarray:
        """vpair as float64, F-contiguous (P, P, n_block). COPY."""
        return self._download("vpair")


def upload_f2(f2, vpair, block_sizes, pops, device: int = 0) -> F2Handle:
    """Host tensor -> resident handle (the reference's upload_f2_blocks_to_device)."""
    f2 = np.asarray(f2, dtype=np.float64)
    if f2.ndim != 3 or f2.shape[0] != f2.shape[1]:
        raise ValueError("f2 must have shape (P, P, n_block)")
    P, _, nb = f2.shape
    fl = np.asfortranarray(f2).ravel(order="F")
    vp = None if vpair is None else np.asfortranarray(np.asarray(vpair, np.float64)).ravel(order="F")
    bs = np.ascontiguousarray(block_sizes, dtype=np.int32)
    if bs.shape != (nb,):
        raise ValueError("block_sizes must have n_block entries")
    ctx = Context.get(device)
    raw = C.c_void_p()
    check(lib().stp_f2_handle_upload(ctx.h, _ptr(fl), None if vp is None else _ptr(vp), _ptr(bs),
                                     int(P), int(nb), C.byref(raw)), "upload")
    return F2Handle(raw, pops, device)


def _sha256(path: str) -> str:
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def write_f2_dir(dir_path: str, f2, vpair, block_sizes, pops, meta: dict | None = None) -> str:
    """Write an STPF2BK1 f2 directory: f2.bin, pops.txt, meta.json
    (src/app/f2_dir_writer.cpp:306-440)."""
    f2 = np.asarray(f2, dtype=np.float64)
    P, _, nb = f2.shape
    if len(pops) != P:
        raise ValueError(f"write_f2_dir: pop_labels has {len(pops)} entries but f2 tensor P={P}")
    os.makedirs(dir_path, exist_ok=True)
    fl = np.asfortranarray(f2).ravel(order="F")
    vp = np.asfortranarray(np.asarray(vpair, np.float64)).ravel(order="F")
    bs = np.ascontiguousarray(block_sizes, dtype=np.int32)
    bin_path = os.path.join(dir_path, "f2.bin")
    check(lib().stp_f2bin_write(bin_path.encode(), _ptr(fl), _ptr(vp), _ptr(bs), int(P), int(nb)),
          "write_f2_dir")
    pops_path = os.path.join(dir_path, "pops.txt")
    with open(pops_path, "w") as f:
        for s in pops:
            f.write(f"{s}\n")
    m = {"format": "STPF2BK1", "meta_schema_version": 1, "steppe_version": "steppe_amd",
         "P": int(P), "n_block": int(nb)}
    m.update(meta or {})
    m["pops_sha256"] = "sha256:" + _sha256(pops_path)
    m["f2_cache_id"] = "sha256:" + _sha256(bin_path)
    with open(os.path.join(dir_path, "meta.json"), "w") as f:
        json.dump(m, f, indent=2)
        f.write("\n")
    return dir_path


def read_f2_host(dir_path: str):
    """Read f2.bin + pops.txt into host arrays (src/app/f2_dir_io.cpp:52-158)."""
    bin_path = os.path.join(dir_path, "f2.bin")
    if not os.path.exists(bin_path):
        raise ValueError(f"read_f2: no f2.bin in {dir_path}")
    P, nb = C.c_int32(), C.c_int32()
    check(lib().stp_f2bin_read_header(bin_path.encode(), C.byref(P), C.byref(nb)), "read_f2")
    P, nb = int(P.value), int(nb.value)
    f2 = np.empty(P * P * nb, np.float64)
    vp = np.empty_like(f2)
    bs = np.empty(nb, np.int32)
    check(lib().stp_f2bin_read(bin_path.encode(), _ptr(f2), _ptr(vp), _ptr(bs)), "read_f2")
    pops_path = os.path.join(dir_path, "pops.txt")
    if os.path.exists(pops_path):
        with open(pops_path) as f:
            pops = [ln.rstrip("\n") for ln in f if ln.strip()]
    else:
        pops = [f"pop{i}" for i in range(P)]
    if len(pops) != P:
        raise ValueError(f"read_f2: pops.txt has {len(pops)} labels but f2.bin has P={P}")
    return (f2.reshape((P, P, nb), order="F"), vp.reshape((P, P, nb), order="F"), bs, pops)


def read_f2(dir: str, device: int = 0) -> F2Handle:
    """Load an f2 directory into a resident F2Handle (bindings/bind_f2handle.cpp:14-22)."""
    f2, vp, bs, pops = read_f2_host(dir)
    return upload_f2(f2, vp, bs, pops, device)


def parse_precision(precision) -> int:
    if isinstance(precision, (int, np.integer)) and not isinstance(precision, bool) and \
            int(precision) in (STP_PREC_FP64, STP_PREC_SPLIT_BF16):
        return int(precision)
    if precision not in _PRECISIONS:
        raise ValueError("precision must be one of 'split_bf16'/'emu', 'fp64'/'native' "
                         f"(got '{precision}')")
    return _PRECISIONS[precision]


# Output tiers (include/steppe/config.hpp:121 ForceTier, include/steppe/extract.hpp:33).
TIERS = {"auto": -1, "resident": 0, "host": 1, "hostram": 1, "disk": 2}
TIER_NAMES = {0: "resident", 1: "host", 2: "disk"}


def _tier(tier, disk_path):
    t = TIERS.get(str(tier).lower())
    if t is None:
        raise ValueError(f"tier must be one of 'auto', 'resident', 'host', 'disk' (got {tier!r})")
    if t == 2 and not disk_path:
        raise ValueError("the disk tier needs disk_path (the f2.bin file to write)")
    return t


_PLOIDY = {"auto": 0, "1": 1, "pseudo": 1, "pseudo_haploid": 1, "2": 2, "diploid": 2}
_STRAND = {"drop": 0, "keep": 1, "flip": 2}


def extract_f2_files(geno: str, snp: str, ind: str, pops=None, *, top_k: int = 0, min_n: int = 0,
                     out: str = "", device: int = 0, blgsize: float = 0.05, maf: float = 0.0,
                     maxmiss: float = 0.0, autosomes_only: bool = True, drop_monomorphic: bool = False,
                     transversions_only: bool = False, ploidy="auto", strand_mode: str = "drop",
                     precision=None, tier="auto", disk_path: str = "", chunk_blocks: int = 0):
    """Packed genotype triple (TGENO / GENO .geno + .snp + .ind, or PLINK .bed + .bim +
    .fam, told apart by the first bytes of ``geno``) -> f2 (src/app/extract_f2_core.cpp:96-407): pops
    (explicit labels), or the top_k largest populations, or every population with at least
    min_n individuals.  tier: "auto" (HBM if the tensor fits 0.70 of free device memory,
    else host memory, else the f2.bin file disk_path), "resident", "host" or "disk";
    chunk_blocks: blocks per streamed chunk of the host / disk tiers (0 = auto).  Returns
    (F2Handle, info dict), or (out path, info) when ``out`` is set (an STPF2BK1 directory
    with the reference's meta fields)."""
    if pops:
        mode, labels = 0, [str(x) for x in pops]
    elif top_k > 0:
        mode, labels = 1, []
    elif min_n > 0:
        mode, labels = 2, []
    else:
        raise ValueError("extract_f2: pops needs at least one population name")
    pl = _PLOIDY.get(str(ploidy))
    if pl is None:
        raise ValueError(f"extract_f2: ploidy must be one of 'auto', '1'/'pseudo', '2'/'diploid' (got {ploidy!r})")
    if strand_mode not in _STRAND:
        raise ValueError(f"extract_f2: strand_mode must be one of 'drop', 'keep', 'flip' (got {strand_mode!r})")
    for path in (geno, snp, ind):
        if not os.path.exists(path):
            raise ValueError(f"extract_f2: missing genotype file {path}")
    prec = parse_precision(precision)
    tr = _tier(tier, disk_path)
    arr = (C.c_char_p * max(len(labels), 1))(*[x.encode() for x in labels])
    req = ExtractRequest(str(geno).encode(), str(snp).encode(), str(ind).encode(), mode, arr, len(labels),
                         int(top_k), int(min_n), float(maf), float(maxmiss), int(bool(autosomes_only)),
                         int(bool(drop_monomorphic)), int(bool(transversions_only)), _STRAND[strand_mode],
                         float(blgsize), pl, prec, tr, str(disk_path).encode(), int(chunk_blocks))
    ctx = Context.get(device)
    raw = C.c_void_p()
    info = ExtractInfo()
    check(lib().stp_extract_f2_files(ctx.h, C.byref(req), C.byref(raw), C.byref(info)), "extract_f2")
    names = []
    buf = C.create_string_buffer(4096)
    n = C.c_int32(0)
    for i in range(info.P):
        check(lib().stp_f2_handle_label(raw, i, buf, 4096, C.byref(n)), "extract_f2")
        names.append(buf.value.decode())
    h = F2Handle(raw, names, device)
    inf = {"n_snp_total": int(info.n_snp_total), "n_snp_kept": int(info.n_snp_kept),
           "n_pseudo_haploid": int(info.n_pseudo_haploid), "n_diploid": int(info.n_diploid),
           "P": int(info.P), "n_block": int(info.n_block)}
    if out:
        meta = {"precision_tag": PRECISION_NAMES[prec], "precision_mantissa_bits": 53 if prec == STP_PREC_FP64 else 48,
                "blgsize_cm": blgsize * 100.0, "n_snp_total": inf["n_snp_total"], "n_snp_kept": inf["n_snp_kept"],
                "filters": {"maf_min": maf, "maxmiss": maxmiss, "mind_max_missing": 1.0,
                            "autosomes_only": bool(autosomes_only), "drop_monomorphic": bool(drop_monomorphic),
                            "transversions_only": bool(transversions_only)},
                "pop_selection": "explicit" if mode == 0 else ("top_k" if mode == 1 else "min_n"),
                "source": {"geno": str(geno), "snp": str(snp), "ind": str(ind), "source_hash_computed": False,
                           "geno_sha256": "", "snp_sha256": "", "ind_sha256": ""}}
        write_f2_dir(out, h._f2_numpy(), h._vpair_numpy(), h.block_sizes, h.pops, meta)
        return out, inf
    return h, inf


def genotype_triple(prefix: str):
    """The files of a genotype prefix (src/io/genotype_source.cpp:15-30): <prefix>.geno /
    .snp / .ind when the .geno exists, else the PLINK <prefix>.bed / .bim / .fam."""
    if not os.path.exists(prefix + ".geno") and os.path.exists(prefix + ".bed"):
        return prefix + ".bed", prefix + ".bim", prefix + ".fam"
    return prefix + ".geno", prefix + ".snp", prefix + ".ind"


def run_extract_f2(packed, *args, **kw):
    """Genotypes -> f2.  With a path prefix as the first argument this is the reference's
    binding (bindings/bind_fstats.cpp:131-215): run_extract_f2(prefix, pops, out="",
    device=0, blgsize=0.05, maf=0.0, maxmiss=0.0, autosomes_only=True,
    drop_monomorphic=False, transversions_only=False, ploidy="auto", strand_mode="drop",
    precision=None) reads <prefix>.{geno,snp,ind} (or the PLINK <prefix>.{bed,bim,fam}) and
    returns an F2Handle (or the out path).
    Otherwise it takes an in-memory tile (run_extract_f2_tile)."""
    if isinstance(packed, (str, os.PathLike)):
        names = ("pops", "out", "device", "blgsize", "maf", "maxmiss", "autosomes_only", "drop_monomorphic",
                 "transversions_only", "ploidy", "strand_mode", "precision")
        kw.update(dict(zip(names, args)))
        prefix = str(packed)
        pops = kw.pop("pops", None)
        if not pops:
            raise ValueError("extract_f2: pops needs at least one population name")
        h, _ = extract_f2_files(*genotype_triple(prefix), pops,
                                **{k: v for k, v in kw.items() if v is not None or k == "precision"})
        return h
    return run_extract_f2_tile(packed, *args, **kw)


def run_extract_f2_tile(packed, n_snp: int, pop_offsets, block_id, n_block: int, pops=None,
                        out: str = "", device: int = 0, ploidy="auto", precision=None,
                        sample_ploidy=None, tier="resident", disk_path: str = "", chunk_blocks: int = 0):
    """Genotype tile -> f2 (the compute core of bindings/bind_fstats.cpp:131-215): a packed
    individual-major tile, population offsets and the jackknife block partition.  ploidy:
    "auto" (detect), 1 or 2, or "sample" with sample_ploidy.  tier / disk_path /
    chunk_blocks choose where the tensor lives, as in extract_f2_files.  Returns an
    F2Handle, or the directory path when ``out`` is set."""
    n_pop = len(pop_offsets) - 1
    pops = list(pops) if pops is not None else [f"pop{i}" for i in range(n_pop)]
    if sample_ploidy is not None:
        tile, keep = make_tile(packed, n_snp, pop_offsets, sample_ploidy)
    elif ploidy == "auto":
        tile, keep = make_tile(packed, n_snp, pop_offsets, None, detect=True)
    elif ploidy in (1, 2, "1", "2"):
        tile, keep = make_tile(packed, n_snp, pop_offsets, None, ploidy=int(ploidy))
    else:
        raise ValueError(f"ploidy must be 'auto', 1 or 2 (got {ploidy!r})")
    prec = parse_precision(precision)
    bid = np.ascontiguousarray(block_id, dtype=np.int32)
    if bid.shape != (int(n_snp),):
        raise ValueError("block_id must have one entry per SNP")
    ctx = Context.get(device)
    raw = C.c_void_p()
    tr = _tier(tier, disk_path)
    check(lib().stp_extract_f2_tiered(ctx.h, C.byref(tile), _ptr(bid), int(n_block), prec, tr,
                                      str(disk_path).encode(), int(chunk_blocks), C.byref(raw)),
          "run_extract_f2")
    del keep
    h = F2Handle(raw, pops, device)
    if out:
        write_f2_dir(out, h._f2_numpy(), h._vpair_numpy(), h.block_sizes, h.pops,
                     {"precision_tag": PRECISION_NAMES[prec], "n_snp_total": int(n_snp),
                      "n_snp_kept": int(n_snp)})
        return out
    return h


# ---------------------------------------------------------------------------------------
# qpAdm on a resident f2 handle (bindings/bind_qpadm.cpp:14-134)
# ---------------------------------------------------------------------------------------
def _resolve(f2: F2Handle, name, what: str) -> int:
    """A population name (or index) -> index; unknown names raise KeyError, as the
    reference's PopResolver does (bindings/bind_qpadm.cpp:24-31)."""
    if isinstance(name, (int, np.integer)):
        i = int(name)
        if not 0 <= i < f2.P:
            raise KeyError(f"{what}: index {i} out of range")
        return i
    try:
        return f2.pops.index(str(name))
    except ValueError:
        raise KeyError(f"{what}: unknown population {name!r}") from None


def _options(fudge, als_iterations, rank, rank_alpha, jackknife, se_require_p=False,
             p_se_threshold=0.05) -> QpAdmOptions:
    if isinstance(jackknife, str):
        if jackknife not in JACKKNIFE:
            raise ValueError(f"jackknife must be one of {sorted(JACKKNIFE)}")
        jackknife = JACKKNIFE[jackknife]
    return QpAdmOptions(float(fudge), int(als_iterations), int(rank), float(rank_alpha), int(jackknife),
                        1 if se_require_p else 0, float(p_se_threshold))


def qpadm_models(f2: F2Handle, models, rank=-1, fudge=1e-4, als_iterations=20, rank_alpha=0.05,
                 jackknife="all", se_require_p=False, p_se_threshold=0.05, raw=False):
    """Fit a batch of (target, left, right) models in one device call; results in
    input order (the reference's fit_models_batched, src/device/backend.hpp:1636-1647).
    ``raw=True`` returns the QpAdmResult records instead of dicts."""
    tg, loff, lf, roff, rt = [], [0], [], [0], []
    for i, (target, left, right) in enumerate(models):
        if not left:
            raise ValueError(f"qpadm: model {i} has an empty left source list")
        if not right:
            raise ValueError(f"qpadm: model {i} needs at least one right population")
        tg.append(_resolve(f2, target, "target"))
        lf.extend(_resolve(f2, x, "left") for x in left)
        rt.extend(_resolve(f2, x, "right") for x in right)
        loff.append(len(lf))
        roff.append(len(rt))
    n = len(tg)
    out = (QpAdmResult * max(n, 1))()
    if n:
        arr = [np.ascontiguousarray(a, np.int32) for a in (tg, loff, lf, roff, rt)]
        opts = _options(fudge, als_iterations, rank, rank_alpha, jackknife, se_require_p, p_se_threshold)
        ctx = Context.get(f2.device)
        check(lib().stp_qpadm_batch(ctx.h, f2.raw, n, *[_ptr(a) for a in arr], C.byref(opts), out), "qpadm")
    recs = list(out)[:n]
    return recs if raw else [r.to_dict() for r in recs]


def run_qpadm(f2: F2Handle, target, left, right, rank=-1, fudge=1e-4, als_iterations=20, rank_alpha=0.05,
              allow_negative_weights=True) -> dict:
    """Single-model qpAdm (bindings/bind_qpadm.cpp:14-42); SEs for every model."""
    if target is None or target == "":
        raise ValueError("qpadm: target is required")
    if not left:
        raise ValueError("qpadm: left needs at least one source population")
    if not right:
        raise ValueError("qpadm: right needs at least one outgroup population")
    return qpadm_models(f2, [(target, list(left), list(right))], rank=rank, fudge=fudge,
                        als_iterations=als_iterations, rank_alpha=rank_alpha, jackknife="all")[0]


def run_qpadm_search(f2: F2Handle, target, lefts, right, rank=-1, fudge=1e-4, als_iterations=20,
                     rank_alpha=0.05, allow_negative_weights=True, jackknife="all") -> list:
    """Batched qpAdm over left-source sets with a shared target and right set
    (bindings/bind_qpadm.cpp:72-113); per-model dicts in input order."""
    if target is None or target == "":
        raise ValueError("qpadm_search: target is required")
    if not lefts:
        raise ValueError("qpadm_search: models needs at least one (left) entry")
    if not right:
        raise ValueError("qpadm_search: right needs at least one outgroup population")
    for i, l in enumerate(lefts):
        if not l:
            raise ValueError(f"qpadm_search: model {i} has an empty left source list")
    return qpadm_models(f2, [(target, list(l), list(right)) for l in lefts], rank=rank, fudge=fudge,
                        als_iterations=als_iterations, rank_alpha=rank_alpha, jackknife=jackknife)


# ---------------------------------------------------------------------------------------
# standalone f-statistics (bindings/bind_fstats.cpp:29-120, bindings/steppe/__init__.py:865-1102)
# ---------------------------------------------------------------------------------------
def _tuples(f2: F2Handle, rows, width: int, what: str) -> np.ndarray:
    """Rows of population names / indices / {"pop1": ..} dicts -> (n, width) int32 indices;
    unknown names and out-of-range indices raise KeyError (bindings/bind_fstats.cpp)."""
    if isinstance(rows, np.ndarray) and rows.dtype.kind in "iu":
        idx = np.ascontiguousarray(rows.reshape(-1, width) if rows.size else rows.reshape(0, width), np.int32)
        if idx.size and (idx.min() < 0 or idx.max() >= f2.P):
            bad = int(np.argmax((idx < 0).any(axis=1) | (idx >= f2.P).any(axis=1)))
            raise KeyError(f"{what}: index out of range in row {bad}")
        return idx
    lookup = {name: i for i, name in enumerate(f2.pops)}
    idx = np.zeros((len(rows), width), np.int32)
    for i, r in enumerate(rows):
        if isinstance(r, dict):
            r = [r[f"pop{j + 1}"] for j in range(width)]
        r = list(r)
        if len(r) != width:
            raise ValueError(f"{what}: row {i} needs {width} populations, got {len(r)}")
        for j, name in enumerate(r):
            if isinstance(name, (int, np.integer)):
                if not 0 <= int(name) < f2.P:
                    raise KeyError(f"{what}: index {int(name)} out of range")
                idx[i, j] = int(name)
            else:
                v = lookup.get(str(name))
                if v is None:
                    raise KeyError(f"{what}: unknown population {name!r}")
                idx[i, j] = v
    return idx


def _names(f2: F2Handle, idx: np.ndarray, width: int) -> dict:
    pops = np.asarray(f2.pops, dtype=object) if f2.pops else None
    return {f"pop{j + 1}": (pops[idx[:, j]].tolist() if pops is not None else idx[:, j].tolist())
            for j in range(width)}


def _fstat(f2: F2Handle, rows, width: int, fn: str, what: str) -> dict:
    idx = _tuples(f2, rows, width, what)
    n = idx.shape[0]
    est, se, z, p = (np.zeros(max(n, 1)) for _ in range(4))
    if n:
        ctx = Context.get(f2.device)
        check(getattr(lib(), fn)(ctx.h, f2.raw, _ptr(idx), n, _ptr(est), _ptr(se), _ptr(z), _ptr(p)), what)
    d = _names(f2, idx, width)
    d.update({"est": est[:n].tolist(), "se": se[:n].tolist(), "z": z[:n].tolist(), "p": p[:n].tolist(),
              "status": "ok", "precision": "fp64"})
    return d


def run_f4(f2: F2Handle, quartets) -> dict:
    """f4(p1,p2;p3,p4) per quartet: est / se / z / p (reference: run_f4, src/core/qpadm/f4.cpp)."""
    return _fstat(f2, quartets, 4, "stp_f4", "f4")


def run_qpdstat(f2: F2Handle, quartets) -> dict:
    """qpDstat on the f2 path: the same statistic as f4 (bindings/bind_fstats.cpp:52-75)."""
    return _fstat(f2, quartets, 4, "stp_f4", "qpdstat")


def run_f3(f2: F2Handle, triples) -> dict:
    """f3(C;A,B) per triple: est / se / z / p (reference: run_f3, src/core/qpadm/f3.cpp)."""
    return _fstat(f2, triples, 3, "stp_f3", "f3")


def run_f4ratio(f2: F2Handle, tuples) -> dict:
    """alpha = f4(p1,p2;p3,p4) / f4(p1,p2;p5,p4) per 5-tuple with the ratio block jackknife
    (reference: run_f4ratio, src/core/qpadm/f4ratio.cpp)."""
    idx = _tuples(f2, tuples, 5, "f4ratio")
    n = idx.shape[0]
    alpha, se, z = (np.zeros(max(n, 1)) for _ in range(3))
    if n:
        ctx = Context.get(f2.device)
        check(lib().stp_f4ratio(ctx.h, f2.raw, _ptr(idx), n, _ptr(alpha), _ptr(se), _ptr(z)), "f4ratio")
    d = _names(f2, idx, 5)
    d.update({"alpha": alpha[:n].tolist(), "se": se[:n].tolist(), "z": z[:n].tolist(), "status": "ok",
              "precision": "fp64"})
    return d


SWEEP_METHODS = {"auto": 0, "direct": 1, "gram": 2}


def run_fstat_sweep(f2: F2Handle, k: int, pops=None, min_z: float = 3.0, top_k: int = 0, sure: bool = False,
                    all_combinations: bool = False, method: str = "auto") -> dict:
    """All-combination f4 (k=4) / f3 (k=3) sweep (reference: run_f4_sweep / run_f3_sweep,
    src/core/qpadm/fstat_sweep.cpp, driven as cmd_fstat_sweep.cpp:96-110 drives it):
    top_k > 0 keeps the top_k largest |z|; all_combinations keeps up to 1e6 by |z|;
    otherwise every combination with |z| >= min_z (at most 1e6).  Rows are ordered by
    |z| descending; keys name the populations (k columns)."""
    if k not in (3, 4):
        raise ValueError("fstat_sweep: k must be 3 or 4")
    if method not in SWEEP_METHODS:
        raise ValueError(f"fstat_sweep: method must be one of {sorted(SWEEP_METHODS)}")
    sub = None if not pops else np.ascontiguousarray([_resolve(f2, x, "pops") for x in pops], np.int32)
    if top_k > 0:
        filt, K = 1, int(top_k)
    elif all_combinations:
        filt, K = 1, 1000000
    else:
        filt, K = 0, 1000000
    from math import comb
    rng = f2.P if sub is None else len(sub)
    cap = max(1, min(K, comb(rng, k)))
    req = SweepRequest(k, filt, float(min_z), K, 1 if sure else 0, SWEEP_METHODS[method],
                       None if sub is None else sub.ctypes.data_as(C.c_void_p), 0 if sub is None else len(sub))
    keys = np.zeros((cap, 4), np.int32)
    est, se, z, p = (np.zeros(cap) for _ in range(4))
    n_out, enum = C.c_int64(0), C.c_uint64(0)
    capped, status = C.c_int32(0), C.c_int32(0)
    ctx = Context.get(f2.device)
    check(lib().stp_fstat_sweep(ctx.h, f2.raw, C.byref(req), cap, _ptr(keys), _ptr(est), _ptr(se), _ptr(z), _ptr(p),
                                C.byref(n_out), C.byref(enum), C.byref(capped), C.byref(status)), "fstat_sweep")
    n = min(int(n_out.value), cap)
    d = _names(f2, keys[:n], k)
    d.update({"keys": keys[:n, :k].copy(), "est": est[:n], "se": se[:n], "z": z[:n], "p": p[:n],
              "enumerated": int(enum.value), "survivors": int(n_out.value), "capped": bool(capped.value),
              "status": STATUS_NAMES.get(int(status.value), str(status.value)), "precision": "fp64"})
    return d


def run_qpwave(f2: F2Handle, left, right, fudge: float = 1e-4, rank_alpha: float = 0.05) -> dict:
    """qpWave rank sweep with left[0] as the reference row (reference: run_qpwave,
    src/core/qpadm/qpadm_fit.cpp:186-214).  Its f4 basis, covariance and rank sweep are
    those of a qpAdm model with target left[0] and sources left[1:], so it runs through the
    batched device fit without the leave-one-block-out refits."""
    left, right = list(left), list(right)
    if len(left) < 2:
        raise ValueError("qpwave: left needs at least two populations")
    if len(right) < 2:
        raise ValueError("qpwave: right needs at least two populations")
    r = qpadm_models(f2, [(left[0], left[1:], right)], fudge=fudge, rank_alpha=rank_alpha, jackknife="none")[0]
    by_rank = {fr: (c, pv) for fr, c, pv in zip(r["rankdrop_f4rank"], r["rankdrop_chisq"], r["rankdrop_p"])}
    n = len(r["rankdrop_f4rank"])
    status = r["status"] if r["status"] in ("rank_deficient", "non_spd_covariance") else "ok"
    return {
        "rank_chisq": [by_rank[i][0] for i in range(n)],
        "rank_dof": [d for _, d in sorted(zip(r["rankdrop_f4rank"], r["rankdrop_dof"]))],
        "rank_p": [by_rank[i][1] for i in range(n)],
        "rankdrop_f4rank": r["rankdrop_f4rank"], "rankdrop_dof": r["rankdrop_dof"],
        "rankdrop_dofdiff": r["rankdrop_dofdiff"], "rankdrop_chisq": r["rankdrop_chisq"],
        "rankdrop_p": r["rankdrop_p"], "rankdrop_chisqdiff": r["rankdrop_chisqdiff"],
        "rankdrop_p_nested": r["rankdrop_p_nested"],
        "f4rank": r["f4rank"], "est_rank": r["f4rank"], "status": status, "precision": "fp64",
    }
