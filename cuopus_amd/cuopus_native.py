"""cuopus_native -- Python binding of libcuopus.so (MI355X / ROCm).

Public API of the reference's python/cuopus_native.py (:30-180), same names,
arguments and ``CuopusError(rc)`` semantics: ``probe``,
``decode_ogg_to_cuda_tensor``, ``decode_ogg_multi``, ``warmup``,
``workspace_bytes``, ``last_arena_overflow``, ``release_workspaces``.
Tensors are torch ROCm tensors (device "cuda" is the HIP device); every decode
runs on torch's current stream, which the library synchronises before the
call returns.

How it is built here: one table of C signatures, one call helper that owns
the error buffer and turns a non-zero return code into ``CuopusError``, and
parsed-file handles (``cuopus_file_open``) so a stream is demuxed once per
decode: the output tensor is sized from the handle and the same handle is
decoded.

Additions: ``decode_ogg_to_s16_tensor`` (playable int16 48 kHz),
``decode_ogg_lossy`` (per-packet loss mask with in-band FEC),
``decode_ogg_noalloc`` (the allocation-free entry point), ``log_mel``
(log-mel features of a resident f32 16 kHz tensor), ``packet_count``,
``chain_table``, ``packets``, ``demux_stats``, the CUOP/CUC2 dump writers,
``gate_chains`` (per-packet sample counts and final ranges), ``Batch``
(resident inputs, repeated decodes) and ``Server`` (continuous batching).

The library is the only decode path: if it is missing or no GPU is present,
calls raise; nothing falls back to another implementation.
"""
import ctypes
import os

import numpy as np

_ERRCAP = 512
_RC_NAMES = {1: "soft", 2: "gpu", 3: "caller"}


class CuopusError(RuntimeError):
    """A libcuopus call failed; ``rc`` is its return code (1 soft failure,
    the GPU is healthy; 2 GPU/HIP error; 3 caller error)."""

    def __init__(self, rc, msg):
        super().__init__("cuopus[%s]: %s" % (_RC_NAMES.get(rc, rc), msg))
        self.rc = rc


# ------------------------------------------------------------------ C signatures
_c = ctypes
VP, SZ, I32, I64, U32, U64 = _c.c_void_p, _c.c_size_t, _c.c_int, _c.c_int64, _c.c_uint32, _c.c_uint64
BUF = _c.c_char_p
P_I64 = _c.POINTER(I64)


class ServerOpts(_c.Structure):
    """cuopus_server_opts (include/cuopus.h)."""
    _fields_ = [("n_streams", I32), ("n_waves", I32), ("mem_budget", _c.c_longlong),
                ("max_chains_per_launch", I32)]


class BatchOpts(_c.Structure):
    """cuopus_batch_opts (include/cuopus.h)."""
    _fields_ = [("mode", I32), ("n_slots", I32), ("lpt_order", I32), ("use_graph", I32), ("profile", I32),
                ("rs_variant", I32)]


# name -> (restype, argtypes without the trailing (errbuf, errcap) pair, takes_err)
_SIGS = {
    "cuopus_probe": (I32, [BUF, SZ, P_I64], True),
    "cuopus_probe_packets": (I32, [BUF, SZ, P_I64, P_I64], True),
    "cuopus_decode_f16k": (I32, [BUF, SZ, VP, SZ, P_I64], True),
    "cuopus_decode_f16k_on": (I32, [BUF, SZ, VP, SZ, P_I64, VP], True),
    "cuopus_decode_f16k_noalloc_on": (I32, [BUF, SZ, VP, SZ, P_I64, VP], True),
    "cuopus_decode_s16_on": (I32, [BUF, SZ, VP, SZ, P_I64, VP], True),
    "cuopus_decode_lossy_f16k_on": (I32, [BUF, SZ, VP, SZ, VP, SZ, P_I64, VP], True),
    "cuopus_decode_f16k_multi": (I32, [_c.POINTER(BUF), _c.POINTER(SZ), I32, _c.POINTER(VP), _c.POINTER(SZ),
                                       P_I64, VP], True),
    "cuopus_file_open": (I32, [BUF, SZ, _c.POINTER(VP), P_I64], True),
    "cuopus_file_decode_f16k_on": (I32, [VP, VP, SZ, P_I64, VP], True),
    "cuopus_files_decode_f16k_multi": (I32, [_c.POINTER(VP), I32, _c.POINTER(VP), _c.POINTER(SZ), P_I64, VP],
                                       True),
    "cuopus_file_close": (None, [VP], False),
    "cuopus_gate_chains": (I32, [VP, SZ, VP, VP, U32, VP, VP, VP, VP, VP, U32, VP, VP, VP, VP, VP], True),
    "cuopus_gate_chains_ch": (I32, [VP, SZ, VP, VP, U32, VP, VP, VP, VP, VP, U32, I32, I32, VP, VP, VP, VP, VP],
                              True),
    "cuopus_resample_s16_to_f16k": (I32, [VP, I64, VP, SZ, P_I64, VP], True),
    "cuopus_log_mel": (I32, [VP, I64, VP, SZ, P_I64, I32, VP], True),
    "cuopus_mel_filters": (I32, [I32, VP, SZ], False),
    "cuopus_tuned_defaults": (I32, [_c.POINTER(I32), _c.POINTER(I32), _c.POINTER(I32)], False),
    "cuopus_dump_chains": (I32, [BUF, SZ, VP, SZ, P_I64], True),
    "cuopus_dump_packets": (I32, [BUF, SZ, VP, SZ, VP, SZ, P_I64, P_I64], True),
    "cuopus_demux_stats": (I32, [BUF, SZ, VP], True),
    "cuopus_warmup": (I32, [], True),
    "cuopus_last_arena_overflow": (I32, [], False),
    "cuopus_workspace_bytes": (_c.c_longlong, [], False),
    "cuopus_release_workspaces": (None, [], False),
    "cuopus_prof_read": (I32, [_c.POINTER(_c.c_ulonglong), I32], False),
    "cuopus_server_create": (I32, [_c.POINTER(ServerOpts), _c.POINTER(VP)], True),
    "cuopus_server_submit": (I32, [VP, BUF, SZ, VP, SZ, VP, SZ, VP, _c.POINTER(U64)], True),
    "cuopus_server_poll": (I32, [VP, _c.POINTER(U64), _c.POINTER(I32), I32, _c.POINTER(I32)], True),
    "cuopus_server_stats": (I32, [VP, P_I64], False),
    "cuopus_server_destroy": (None, [VP], False),
    "cuopus_batch_create": (I32, [_c.POINTER(BUF), _c.POINTER(SZ), I32, _c.POINTER(VP), _c.POINTER(SZ),
                                  _c.POINTER(VP), _c.POINTER(SZ), _c.POINTER(BatchOpts), _c.POINTER(VP)], True),
    "cuopus_batch_run": (I32, [VP, VP], True),
    "cuopus_batch_check": (I32, [VP, VP, P_I64], True),
    "cuopus_batch_info": (I32, [VP, P_I64], False),
    "cuopus_batch_kernel_ms": (I32, [VP, _c.POINTER(_c.c_float), _c.POINTER(_c.c_float)], False),
    "cuopus_batch_chain_times": (I32, [VP, VP, VP, VP, I64], False),
    "cuopus_batch_destroy": (None, [VP], False),
}


def library_path():
    """libcuopus.so: $CUOPUS_LIB, else the one built into this package."""
    path = os.environ.get("CUOPUS_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                        "libcuopus.so")
    if not os.path.exists(path):
        raise OSError("libcuopus.so not found at %s (build it with __graft_entry__.build() or set CUOPUS_LIB)"
                      % path)
    return path


class _Native:
    """The loaded library: typed entry points plus the error-checking call."""

    def __init__(self, path):
        self.dll = _c.CDLL(path)
        self.takes_err = {}
        for name, (res, args, err) in _SIGS.items():
            fn = getattr(self.dll, name)
            fn.restype = res
            fn.argtypes = args + ([BUF, SZ] if err else [])
            self.takes_err[name] = err

    def raw(self, name):
        return getattr(self.dll, name)

    def __call__(self, name, *args, ok=(0,)):
        """Call entry point `name`; a return code outside `ok` raises
        CuopusError with the library's message.  Returns the code."""
        errbuf = _c.create_string_buffer(_ERRCAP)
        rc = getattr(self.dll, name)(*args, errbuf, _ERRCAP)
        if rc not in ok:
            raise CuopusError(rc, errbuf.value.decode(errors="replace") or name)
        return rc


_NATIVE = None


def _native():
    global _NATIVE
    if _NATIVE is None:
        _NATIVE = _Native(library_path())
    return _NATIVE


def _load():
    """The ctypes library object (tests bind the C ABI through it)."""
    return _native().dll


def _bytes_of(src):
    """An Ogg/Opus stream given as bytes-like or as a path."""
    if isinstance(src, (bytes, bytearray, memoryview)):
        return bytes(src)
    with open(src, "rb") as fh:
        return fh.read()


def _cur_stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def _ptr_array(ctype, values):
    return (ctype * len(values))(*values)


class _Parsed:
    """A demuxed stream (cuopus_file_open); closed on exit."""

    FIELDS = ("n16", "play48", "channels", "packets", "chains", "gain_q8")

    def __init__(self, src):
        self.data = _bytes_of(src)
        self.h = VP(0)
        info = (I64 * 6)()
        _native()("cuopus_file_open", self.data, len(self.data), _c.byref(self.h), info)
        for k, v in zip(self.FIELDS, info):
            setattr(self, k, int(v))

    def close(self):
        if self.h:
            _native().raw("cuopus_file_close")(self.h)
            self.h = VP(0)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _f32_cuda(n):
    import torch
    return torch.empty(n, dtype=torch.float32, device="cuda")


# ------------------------------------------------------------------ the reference's API
def probe(path_or_bytes) -> int:
    """Exact float32 16 kHz output length, max(0, (playable48 - 48) // 3);
    host only."""
    data = _bytes_of(path_or_bytes)
    n16 = I64(0)
    _native()("cuopus_probe", data, len(data), _c.byref(n16))
    return n16.value


def decode_ogg_to_cuda_tensor(path_or_bytes):
    """Decode one Ogg/Opus stream to a new float32 16 kHz tensor on the
    current device and stream (finished when this returns)."""
    with _Parsed(path_or_bytes) as f:
        out = _f32_cuda(f.n16)
        if f.n16:
            wrote = I64(0)
            _native()("cuopus_file_decode_f16k_on", f.h, out.data_ptr(), f.n16, _c.byref(wrote), _cur_stream())
            if wrote.value != f.n16:
                raise CuopusError(1, "output length %d differs from the probed %d" % (wrote.value, f.n16))
    return out


def _parse_all(items):
    """_Parsed handles of items, demuxed side by side on a few threads (the C
    call releases the interpreter lock); on an error the handles opened so
    far are closed and the first failing item's error is raised."""
    items = list(items)
    if len(items) <= 1:
        return [_Parsed(src) for src in items]
    from concurrent.futures import ThreadPoolExecutor

    def one(src):
        try:
            return _Parsed(src), None
        except Exception as exc:  # noqa: BLE001 -- re-raised below in item order
            return None, exc

    with ThreadPoolExecutor(max_workers=min(len(items), os.cpu_count() or 1, 16)) as ex:
        res = list(ex.map(one, items))
    err = next((exc for _, exc in res if exc is not None), None)
    if err is not None:
        for f, _ in res:
            if f is not None:
                f.close()
        raise err
    return [f for f, _ in res]


def decode_ogg_multi(items):
    """Decode several streams in one decode launch and one resample launch;
    returns one float32 16 kHz tensor per item, each identical to what
    decode_ogg_to_cuda_tensor gives for it."""
    parsed = []
    try:
        parsed = _parse_all(items)
        if not parsed:
            return []
        outs = [_f32_cuda(f.n16) for f in parsed]
        k = len(parsed)
        lens = (I64 * k)()
        _native()("cuopus_files_decode_f16k_multi", _ptr_array(VP, [f.h for f in parsed]), k,
                  _ptr_array(VP, [o.data_ptr() for o in outs]), _ptr_array(SZ, [f.n16 for f in parsed]), lens,
                  _cur_stream())
        bad = [i for i, f in enumerate(parsed) if lens[i] != f.n16]
        if bad:
            i = bad[0]
            raise CuopusError(1, "item %d: output length %d differs from the probed %d" % (i, lens[i], parsed[i].n16))
        return outs
    finally:
        for f in parsed:
            f.close()


def warmup():
    """Load the kernels and resolve device state once (optional)."""
    _native()("cuopus_warmup")


def workspace_bytes() -> int:
    return int(_native().raw("cuopus_workspace_bytes")())


def last_arena_overflow() -> bool:
    return bool(_native().raw("cuopus_last_arena_overflow")())


def release_workspaces():
    _native().raw("cuopus_release_workspaces")()


# ------------------------------------------------------------------ additions
def decode_ogg_noalloc(path_or_bytes, out=None):
    """cuopus_decode_f16k_noalloc_on: the same decode with every device
    buffer in the library's static scratch (no device allocation by the
    library).  `out` may be a preallocated float32 tensor of at least
    probe() samples."""
    data = _bytes_of(path_or_bytes)
    n16 = probe(data)
    if out is None:
        out = _f32_cuda(n16)
    wrote = I64(0)
    _native()("cuopus_decode_f16k_noalloc_on", data, len(data), out.data_ptr() if n16 else None, out.numel(),
              _c.byref(wrote), _cur_stream())
    return out[:wrote.value]


def packet_count(path_or_bytes):
    """(audio packets, chains) of a stream (host only)."""
    with _Parsed(path_or_bytes) as f:
        return f.packets, f.chains


def file_info(path_or_bytes):
    """Dict of the parsed stream: n16, play48, channels, packets, chains,
    gain_q8 (host only)."""
    with _Parsed(path_or_bytes) as f:
        return {k: getattr(f, k) for k in _Parsed.FIELDS}


def chain_table(path_or_bytes):
    """(n, 4) uint32 array of (first, npkts, trim, keep) per chain (host only)."""
    data = _bytes_of(path_or_bytes)
    n = I64(0)
    _native()("cuopus_dump_chains", data, len(data), None, 0, _c.byref(n))
    table = np.zeros((n.value, 4), np.uint32)
    _native()("cuopus_dump_chains", data, len(data), table.ctypes.data, n.value, _c.byref(n))
    return table


def packets(path_or_bytes):
    """(packet bytes uint8, lengths uint32) of the audio packets (host only):
    the reference CLI's --dump-blob payload."""
    data = _bytes_of(path_or_bytes)
    npk, nb = I64(0), I64(0)
    _native()("cuopus_dump_packets", data, len(data), None, 0, None, 0, _c.byref(npk), _c.byref(nb))
    blob = np.zeros(max(1, nb.value), np.uint8)
    lens = np.zeros(max(1, npk.value), np.uint32)
    _native()("cuopus_dump_packets", data, len(data), blob.ctypes.data, blob.size, lens.ctypes.data, lens.size,
              _c.byref(npk), _c.byref(nb))
    return blob[:nb.value], lens[:npk.value]


DEMUX_STATS = ("pages", "links", "crc_bad", "foreign_pages", "dropped_packets", "granule_warnings",
               "granule_checked", "skipped_bytes", "packets", "play48")


def demux_stats(path_or_bytes):
    """Demuxer report as a dict (keys DEMUX_STATS).  Honours CUOPUS_CHECK_CRC
    and CUOPUS_GRANULE like every entry point."""
    data = _bytes_of(path_or_bytes)
    s = (I64 * len(DEMUX_STATS))()
    _native()("cuopus_demux_stats", data, len(data), s)
    return dict(zip(DEMUX_STATS, (int(v) for v in s)))


def dump_blob(path_or_bytes, out_path):
    """Write the CUOP packet dump (cuopus_demux.h:387-397 layout: 'CUOP', u32
    packet count, u32 lengths, packet bytes)."""
    blob, lens = packets(path_or_bytes)
    with open(out_path, "wb") as fh:
        fh.write(b"CUOP" + np.uint32(lens.size).tobytes() + lens.astype("<u4").tobytes() + blob.tobytes())


def dump_chains(path_or_bytes, out_path):
    """Write the CUC2 chain dump (cuopus_demux.h:399-411 layout: 'CUC2', u32
    chain count, 4 x u32 first/npkts/trim/keep per chain)."""
    table = chain_table(path_or_bytes)
    with open(out_path, "wb") as fh:
        fh.write(b"CUC2" + np.uint32(len(table)).tobytes() + table.astype("<u4").tobytes())


def decode_ogg_lossy(path_or_bytes, lost_mask):
    """Decode with a per-packet loss mask (lost packets come from the next
    packet's in-band FEC when it has one, else concealment)."""
    data = _bytes_of(path_or_bytes)
    mask = np.ascontiguousarray(np.asarray(lost_mask, dtype=np.uint8))
    n16 = probe(data)
    out = _f32_cuda(n16)
    if n16:
        wrote = I64(0)
        _native()("cuopus_decode_lossy_f16k_on", data, len(data), mask.ctypes.data, mask.size, out.data_ptr(), n16,
                  _c.byref(wrote), _cur_stream())
    return out


def decode_ogg_to_s16_tensor(path_or_bytes):
    """Playable int16 48 kHz PCM (preroll and pre-skip removed) as a device
    tensor: shape (n,) for a mono stream, (n, 2) interleaved L/R for stereo."""
    import torch
    with _Parsed(path_or_bytes) as f:
        out = torch.empty(max(1, f.play48 * f.channels), dtype=torch.int16, device="cuda")
        wrote = I64(0)
        _native()("cuopus_decode_s16_on", f.data, len(f.data), out.data_ptr(), out.numel(), _c.byref(wrote),
                  _cur_stream())
        n, ch = wrote.value, f.channels
    return out[:n] if ch == 1 else out[:n * ch].view(n, ch)


def resample_s16(pcm48):
    """int16 48 kHz device tensor -> float32 16 kHz tensor with the library's
    resampler."""
    n16 = max(0, (pcm48.numel() - 48) // 3)
    out = _f32_cuda(n16)
    wrote = I64(0)
    _native()("cuopus_resample_s16_to_f16k", pcm48.data_ptr(), pcm48.numel(), out.data_ptr(), n16,
              _c.byref(wrote), _cur_stream())
    return out[:wrote.value]


MEL_N_FFT, MEL_WIN, MEL_HOP = 512, 400, 160


def log_mel_frames(n16):
    """Frames of the log-mel front-end for n16 input samples."""
    return 0 if n16 < MEL_WIN else 1 + (n16 - MEL_WIN) // MEL_HOP


def log_mel(f16k, n_mels=80):
    """Log-mel features of a float32 16 kHz device tensor (cuopus_log_mel):
    (frames, n_mels) float32, frames of 400 samples every 160 (25 ms / 10 ms),
    periodic Hann window, 512-point power spectrum, Slaney mel filters
    0-8 kHz, natural log with a 1e-10 floor."""
    import torch
    n = f16k.numel()
    frames = log_mel_frames(n)
    out = torch.empty((frames, n_mels), dtype=torch.float32, device="cuda")
    got = I64(0)
    _native()("cuopus_log_mel", f16k.data_ptr() if n else None, n, out.data_ptr() if frames else None, out.numel(),
              _c.byref(got), n_mels, _cur_stream())
    return out


def tuned_defaults():
    """(slots, resampler variant, work order) the library uses on the current
    device when a caller does not choose."""
    v = [I32(0), I32(0), I32(0)]
    _native().raw("cuopus_tuned_defaults")(*[_c.byref(x) for x in v])
    return tuple(x.value for x in v)


def mel_filters(n_mels=80):
    """(n_mels, 257) float32 filter weights that log_mel applies (host only)."""
    out = np.zeros((n_mels, MEL_N_FFT // 2 + 1), np.float32)
    rc = _native().raw("cuopus_mel_filters")(n_mels, out.ctypes.data, out.size)
    if rc:
        raise CuopusError(rc, "n_mels %d outside 1..128" % n_mels)
    return out


def gate_chains(data, off, length, chains, lost=None, channels=1, stereo_build=False, lib=None):
    """Per-packet gate on the GPU.

    data: uint8 packet bytes; off/length: uint32 per packet; chains: int array
    (n, 4) of (first, npkts, trim, keep).  channels: decoder output channels
    (2: interleaved stereo); stereo_build: decode with the stereo-capable
    kernel build (implied by channels=2).  Returns (pcm int16 sum(keep) x
    channels, nsamp int32 per packet, rng uint32 per packet, stop int32 per
    chain, rc) where rc 1 means some chain stopped early (see stop).  lib: a
    ``_Native`` of another build of the library (tests), default the loaded one.
    """
    as_u32 = lambda a: np.ascontiguousarray(a, dtype=np.uint32)  # noqa: E731
    data = np.ascontiguousarray(data, dtype=np.uint8)
    off, length = as_u32(off), as_u32(length)
    table = as_u32(np.asarray(chains, dtype=np.int64).reshape(-1, 4))
    cols = [np.ascontiguousarray(table[:, i]) for i in range(4)]
    pcm = np.zeros(int(cols[3].sum(dtype=np.uint64)) * channels, np.int16)
    nsamp = np.zeros(off.size, np.int32)
    rng = np.zeros(off.size, np.uint32)
    stop = np.zeros(len(table), np.int32)
    mask = None if lost is None else np.ascontiguousarray(lost, dtype=np.uint8)
    rc = (lib or _native())("cuopus_gate_chains_ch", data.ctypes.data, data.size, off.ctypes.data, length.ctypes.data,
                   off.size, None if mask is None else mask.ctypes.data, *[c.ctypes.data for c in cols], len(table),
                   channels, int(bool(stereo_build)), pcm.ctypes.data, nsamp.ctypes.data, rng.ctypes.data,
                   stop.ctypes.data, _cur_stream(), ok=(0, 1))
    return pcm, nsamp, rng, stop, rc


PROF_SCOPES = ["chain", "frame", "celt", "celt_head", "celt_bands", "celt_synth", "celt_comb", "celt_deemph",
               "silk", "emit", "celt_lost", "celt_energy", "celt_alloc", "celt_tail", "pvq", "cwrsi", "exp_rot",
               "theta", "leaf_fold", "band_pre", "band_post", "ec_uint", "alloc_bisect", "alloc_tail",
               "qp_leaf_bits", "bands_prologue", "lowband_out", "qp_stage", "silk_indices", "silk_pulses",
               "silk_params", "silk_core", "silk_resamp", "silk_plc", "silk_nlsf", "silk_nlsf2a", "silk_cng",
               "silk_pulse_counts", "silk_shell", "silk_signs", "qp_leaf"]


def prof_read():
    """Per-scope shader-clock cycle sums of a -DCUO_PROF library (cleared on
    read); empty dict for a normal build."""
    sums = (_c.c_ulonglong * 64)()
    n = min(_native().raw("cuopus_prof_read")(sums, 64), len(PROF_SCOPES))
    return {PROF_SCOPES[i]: int(sums[i]) for i in range(max(n, 0)) if sums[i]}


MODE_PLAIN, MODE_ONE_LAUNCH, MODE_PERSISTENT = 0, 1, 2


class Batch:
    """A resident batch: packets demuxed and uploaded once, then decoded on the
    device as many times as asked (run), without host work per run.

    mode: MODE_PLAIN (one decode and one resample launch per file),
    MODE_ONE_LAUNCH (every chain in one decode launch) or MODE_PERSISTENT
    (n_slots wavefronts pulling chains from a device work queue).
    lpt_order: 0 file order, 1 (True) costliest chains first by the per-packet
    cost model, 2 longest first by samples and packets, 3 the library default
    (tuned).  rs_variant: -1 the
    library's default resampler shape, else a variant index.
    Outputs are float32 16 kHz device tensors, one per file (``outputs``).
    """

    def __init__(self, items, lost_masks=None, mode=MODE_PERSISTENT, n_slots=0, lpt_order=True,
                 use_graph=False, profile=False, outputs=True, rs_variant=-1):
        blobs = [_bytes_of(x) for x in items]
        k = len(blobs)
        if k == 0:
            raise CuopusError(3, "no files")
        self._h = VP(0)
        self.profiled = bool(profile) and not use_graph
        self.n16 = [probe(b) for b in blobs]
        self.outputs = [_f32_cuda(n) for n in self.n16] if outputs else None
        masks = None
        if lost_masks is not None:
            self._masks = [None if m is None else np.ascontiguousarray(np.asarray(m, dtype=np.uint8))
                           for m in lost_masks]
            masks = (_ptr_array(VP, [None if m is None else m.ctypes.data for m in self._masks]),
                     _ptr_array(SZ, [0 if m is None else m.size for m in self._masks]))
        opts = BatchOpts(mode, n_slots, int(lpt_order), int(bool(use_graph)), int(bool(profile)),
                         0 if rs_variant < 0 else rs_variant + 1)
        _native()("cuopus_batch_create", _ptr_array(BUF, blobs), _ptr_array(SZ, [len(b) for b in blobs]), k,
                  masks[0] if masks else None, masks[1] if masks else None,
                  _ptr_array(VP, [o.data_ptr() for o in self.outputs]) if outputs else None,
                  _ptr_array(SZ, self.n16) if outputs else None, _c.byref(opts), _c.byref(self._h))
        counts = (I64 * 8)()
        _native().raw("cuopus_batch_info")(self._h, counts)
        (self.chains, self.packets, self.playable, self.decoded, self.out16, self.packet_bytes,
         self.chain_packets, self.chain_packet_bytes) = (int(v) for v in counts)

    def run(self, stream=None):
        _native()("cuopus_batch_run", self._h, _cur_stream() if stream is None else stream)

    def check(self, stream=None):
        """Wait for the last run and verify every chain decoded fully; returns
        the output lengths."""
        lens = (I64 * len(self.n16))()
        _native()("cuopus_batch_check", self._h, _cur_stream() if stream is None else stream, lens)
        return list(lens)

    def kernel_ms(self):
        """(decode, resample) milliseconds of the last profiled run (events on
        the launch stream)."""
        d, r = _c.c_float(0), _c.c_float(0)
        rc = _native().raw("cuopus_batch_kernel_ms")(self._h, _c.byref(d), _c.byref(r))
        if rc:
            raise CuopusError(rc, "batch was not created with profile=True (or uses a graph)")
        return d.value, r.value

    def chain_times(self, stream=None):
        """(start, end) int64 arrays per chain (file order) of the last run, in
        10 ns ticks of the device's constant clock, start of the earliest
        chain = 0."""
        t0 = np.zeros(self.chains, np.uint32)
        t1 = np.zeros(self.chains, np.uint32)
        rc = _native().raw("cuopus_batch_chain_times")(self._h, _cur_stream() if stream is None else stream,
                                                       t0.ctypes.data, t1.ctypes.data, self.chains)
        if rc:
            raise CuopusError(rc, "chain times need a batch that has run")
        base = t0.min() if t0.size else 0
        s = (t0.astype(np.int64) - base) & 0xFFFFFFFF
        e = (t1.astype(np.int64) - base) & 0xFFFFFFFF
        return s, e

    def close(self):
        if self._h:
            _native().raw("cuopus_batch_destroy")(self._h)
            self._h = VP(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching decode server (cuopus_server_* in include/cuopus.h).

    submit(ogg) demuxes on the host, allocates the float32 16 kHz output
    tensor on the current stream and queues the request (the launch waits for
    that stream's work up to the submit); poll() admits queued requests as
    launch slots and device memory free up, and returns [(request id,
    tensor, rc)] for the requests that finished.  Never blocks.
    """

    def __init__(self, n_streams=4, n_waves=0, mem_budget=8 << 30, max_chains_per_launch=0):
        self._h = VP(0)
        _native()("cuopus_server_create", _c.byref(ServerOpts(n_streams, n_waves, mem_budget, max_chains_per_launch)),
                  _c.byref(self._h))
        self._out = {}
        self._ids = (U64 * 256)()
        self._rcs = (I32 * 256)()

    def submit(self, path_or_bytes, lost_mask=None):
        data = _bytes_of(path_or_bytes)
        n16 = probe(data)
        out = _f32_cuda(n16)
        mask = None if lost_mask is None else np.ascontiguousarray(np.asarray(lost_mask, dtype=np.uint8))
        rid = U64(0)
        _native()("cuopus_server_submit", self._h, data, len(data), None if mask is None else mask.ctypes.data,
                  0 if mask is None else mask.size, out.data_ptr() if n16 else None, n16, _cur_stream(),
                  _c.byref(rid))
        self._out[rid.value] = out
        return rid.value

    def poll(self):
        done = []
        while True:
            n = I32(0)
            _native()("cuopus_server_poll", self._h, self._ids, self._rcs, 256, _c.byref(n))
            done += [(int(self._ids[i]), self._out.pop(int(self._ids[i])), int(self._rcs[i])) for i in range(n.value)]
            if n.value < 256:
                return done

    def stats(self):
        s = (I64 * 4)()
        _native().raw("cuopus_server_stats")(self._h, s)
        return {"queued": s[0], "in_flight": s[1], "held_bytes": s[2], "finished": s[3]}

    def close(self):
        if self._h:
            _native().raw("cuopus_server_destroy")(self._h)
            self._h = VP(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
