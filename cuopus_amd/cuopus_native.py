"""cuopus_native -- ctypes binding of libcuopus.so (MI355X / ROCm).

Same API as the reference's python/cuopus_native.py (:30-180):
``probe``, ``decode_ogg_to_cuda_tensor``, ``decode_ogg_multi``, ``warmup``,
``workspace_bytes``, ``last_arena_overflow``, ``release_workspaces`` and
``CuopusError(rc)``.  Tensors are torch ROCm tensors (device "cuda" is the HIP
device); every call runs on torch's current stream and the library
synchronises that stream before returning.

Additions: ``decode_ogg_to_s16_tensor`` (playable int16 48 kHz),
``decode_ogg_lossy`` (per-packet loss mask with in-band FEC),
``packet_count`` and ``gate_chains`` (per-packet sample counts and final
ranges, the reference harness2.cu protocol).

The library is the only decode path: if it is missing or no GPU is present,
calls raise instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

_ERRCAP = 512
_RC_NAMES = {1: "soft", 2: "gpu", 3: "caller"}


class CuopusError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__("cuopus[%s]: %s" % (_RC_NAMES.get(rc, rc), msg))
        self.rc = rc


def _find_lib():
    env = os.environ.get("CUOPUS_LIB")
    if env:
        return env
    here = os.path.dirname(os.path.abspath(__file__))
    cand = os.path.join(here, "libcuopus.so")
    if os.path.exists(cand):
        return cand
    raise OSError("libcuopus.so not found (build it with __graft_entry__.build() or set CUOPUS_LIB)")


_lib = None
P = ctypes.c_void_p


class ServerOpts(ctypes.Structure):
    """Mirror of cuopus_server_opts (include/cuopus.h)."""
    _fields_ = [("n_streams", ctypes.c_int), ("n_waves", ctypes.c_int), ("mem_budget", ctypes.c_longlong),
                ("max_chains_per_launch", ctypes.c_int)]


class BatchOpts(ctypes.Structure):
    """Mirror of cuopus_batch_opts (include/cuopus.h)."""
    _fields_ = [("mode", ctypes.c_int), ("n_slots", ctypes.c_int), ("lpt_order", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("profile", ctypes.c_int)]
SZ = ctypes.c_size_t
I64P = ctypes.POINTER(ctypes.c_int64)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    lib = ctypes.CDLL(_find_lib())
    sig = {
        "cuopus_probe": (ctypes.c_int, [ctypes.c_char_p, SZ, I64P, ctypes.c_char_p, SZ]),
        "cuopus_probe_packets": (ctypes.c_int, [ctypes.c_char_p, SZ, I64P, I64P, ctypes.c_char_p, SZ]),
        "cuopus_decode_f16k": (ctypes.c_int, [ctypes.c_char_p, SZ, P, SZ, I64P, ctypes.c_char_p, SZ]),
        "cuopus_decode_f16k_on": (ctypes.c_int, [ctypes.c_char_p, SZ, P, SZ, I64P, P, ctypes.c_char_p, SZ]),
        "cuopus_decode_s16_on": (ctypes.c_int, [ctypes.c_char_p, SZ, P, SZ, I64P, P, ctypes.c_char_p, SZ]),
        "cuopus_decode_lossy_f16k_on": (ctypes.c_int, [ctypes.c_char_p, SZ, P, SZ, P, SZ, I64P, P,
                                                       ctypes.c_char_p, SZ]),
        "cuopus_decode_f16k_multi": (ctypes.c_int, [ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(SZ),
                                                    ctypes.c_int, ctypes.POINTER(P), ctypes.POINTER(SZ),
                                                    I64P, P, ctypes.c_char_p, SZ]),
        "cuopus_gate_chains": (ctypes.c_int, [P, SZ, P, P, ctypes.c_uint32, P, P, P, P, P, ctypes.c_uint32,
                                              P, P, P, P, P, ctypes.c_char_p, SZ]),
        "cuopus_resample_s16_to_f16k": (ctypes.c_int, [P, ctypes.c_int64, P, SZ, I64P, P, ctypes.c_char_p, SZ]),
        "cuopus_dump_chains": (ctypes.c_int, [ctypes.c_char_p, SZ, P, SZ, I64P, ctypes.c_char_p, SZ]),
        "cuopus_dump_packets": (ctypes.c_int, [ctypes.c_char_p, SZ, P, SZ, P, SZ, I64P, I64P, ctypes.c_char_p, SZ]),
        "cuopus_demux_stats": (ctypes.c_int, [ctypes.c_char_p, SZ, P, ctypes.c_char_p, SZ]),
        "cuopus_warmup": (ctypes.c_int, [ctypes.c_char_p, SZ]),
        "cuopus_last_arena_overflow": (ctypes.c_int, []),
        "cuopus_workspace_bytes": (ctypes.c_longlong, []),
        "cuopus_release_workspaces": (None, []),
        "cuopus_prof_read": (ctypes.c_int, [ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_int]),
        "cuopus_server_create": (ctypes.c_int, [ctypes.POINTER(ServerOpts), ctypes.POINTER(P), ctypes.c_char_p, SZ]),
        "cuopus_server_submit": (ctypes.c_int, [P, ctypes.c_char_p, SZ, P, SZ, P, SZ,
                                                ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p, SZ]),
        "cuopus_server_poll": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int),
                                              ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, SZ]),
        "cuopus_server_stats": (ctypes.c_int, [P, I64P]),
        "cuopus_server_destroy": (None, [P]),
        "cuopus_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(SZ), ctypes.c_int,
                                               ctypes.POINTER(P), ctypes.POINTER(SZ), ctypes.POINTER(P),
                                               ctypes.POINTER(SZ), ctypes.POINTER(BatchOpts),
                                               ctypes.POINTER(P), ctypes.c_char_p, SZ]),
        "cuopus_batch_run": (ctypes.c_int, [P, P, ctypes.c_char_p, SZ]),
        "cuopus_batch_check": (ctypes.c_int, [P, P, I64P, ctypes.c_char_p, SZ]),
        "cuopus_batch_info": (ctypes.c_int, [P, I64P]),
        "cuopus_batch_kernel_ms": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_float),
                                                  ctypes.POINTER(ctypes.c_float)]),
        "cuopus_batch_destroy": (None, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _as_bytes(path_or_bytes):
    if isinstance(path_or_bytes, (bytes, bytearray, memoryview)):
        return bytes(path_or_bytes)
    with open(path_or_bytes, "rb") as f:
        return f.read()


def _check(rc, err):
    if rc:
        raise CuopusError(rc, err.value.decode(errors="replace"))


def probe(path_or_bytes) -> int:
    """Exact output length in float32 16 kHz samples (host only)."""
    lib = _load()
    data = _as_bytes(path_or_bytes)
    n16 = ctypes.c_int64(0)
    err = ctypes.create_string_buffer(_ERRCAP)
    _check(lib.cuopus_probe(data, len(data), ctypes.byref(n16), err, _ERRCAP), err)
    return n16.value


def packet_count(path_or_bytes):
    """(audio packets, chains) of a stream (host only)."""
    lib = _load()
    data = _as_bytes(path_or_bytes)
    npk, nch = ctypes.c_int64(0), ctypes.c_int64(0)
    err = ctypes.create_string_buffer(_ERRCAP)
    _check(lib.cuopus_probe_packets(data, len(data), ctypes.byref(npk), ctypes.byref(nch), err, _ERRCAP), err)
    return npk.value, nch.value


def chain_table(path_or_bytes):
    """(n, 4) uint32 array of (first, npkts, trim, keep) per chain (host only)."""
    lib = _load()
    data = _as_bytes(path_or_bytes)
    n = ctypes.c_int64(0)
    err = ctypes.create_string_buffer(_ERRCAP)
    _check(lib.cuopus_dump_chains(data, len(data), None, 0, ctypes.byref(n), err, _ERRCAP), err)
    out = np.zeros((n.value, 4), np.uint32)
    _check(lib.cuopus_dump_chains(data, len(data), out.ctypes.data, n.value, ctypes.byref(n), err, _ERRCAP), err)
    return out


def packets(path_or_bytes):
    """(data uint8, lens uint32) of the audio packets (host only), the
    reference CLI's --dump-blob payload."""
    lib = _load()
    data = _as_bytes(path_or_bytes)
    npk, nb = ctypes.c_int64(0), ctypes.c_int64(0)
    err = ctypes.create_string_buffer(_ERRCAP)
    _check(lib.cuopus_dump_packets(data, len(data), None, 0, None, 0, ctypes.byref(npk), ctypes.byref(nb), err,
                                   _ERRCAP), err)
    buf = np.zeros(max(1, nb.value), np.uint8)
    lens = np.zeros(max(1, npk.value), np.uint32)
    _check(lib.cuopus_dump_packets(data, len(data), buf.ctypes.data, buf.size, lens.ctypes.data, lens.size,
                                   ctypes.byref(npk), ctypes.byref(nb), err, _ERRCAP), err)
    return buf[:nb.value], lens[:npk.value]


DEMUX_STATS = ("pages", "links", "crc_bad", "foreign_pages", "dropped_packets", "granule_warnings",
               "granule_checked", "skipped_bytes", "packets", "play48")


def demux_stats(path_or_bytes):
    """Demuxer report as a dict (keys DEMUX_STATS).  Honours CUOPUS_CHECK_CRC
    and CUOPUS_GRANULE like every entry point."""
    lib = _load()
    data = _as_bytes(path_or_bytes)
    s = (ctypes.c_int64 * 10)()
    err = ctypes.create_string_buffer(_ERRCAP)
    _check(lib.cuopus_demux_stats(data, len(data), s, err, _ERRCAP), err)
    return dict(zip(DEMUX_STATS, list(s)))


def dump_blob(path_or_bytes, out_path):
    """Write the CUOP packet dump (cuopus_demux.h:387-397 layout: 'CUOP', u32
    packet count, u32 lengths, packet bytes)."""
    data, lens = packets(path_or_bytes)
    with open(out_path, "wb") as f:
        f.write(b"CUOP")
        f.write(np.uint32(len(lens)).tobytes())
        f.write(lens.astype("<u4").tobytes())
        f.write(data.tobytes())


def dump_chains(path_or_bytes, out_path):
    """Write the CUC2 chain dump (cuopus_demux.h:399-411 layout: 'CUC2', u32
    chain count, 4 x u32 first/npkts/trim/keep per chain)."""
    ch = chain_table(path_or_bytes)
    with open(out_path, "wb") as f:
        f.write(b"CUC2")
        f.write(np.uint32(len(ch)).tobytes())
        f.write(ch.astype("<u4").tobytes())


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def decode_ogg_to_cuda_tensor(path_or_bytes):
    """Decode an Ogg/Opus file or bytes into a new float32 16 kHz device tensor."""
    import torch
    lib = _load()
    data = _as_bytes(path_or_bytes)
    n16 = probe(data)
    out = torch.empty(n16, dtype=torch.float32, device="cuda")
    if n16 == 0:
        return out
    out_len = ctypes.c_int64(0)
    err = ctypes.create_string_buffer(_ERRCAP)
    _check(lib.cuopus_decode_f16k_on(data, len(data), out.data_ptr(), n16, ctypes.byref(out_len),
                                     _stream(), err, _ERRCAP), err)
    if out_len.value != n16:
        raise CuopusError(1, "decode wrote %d samples, probe said %d" % (out_len.value, n16))
    return out


def decode_ogg_lossy(path_or_bytes, lost_mask):
    """Decode with a per-packet loss mask (in-band FEC when available)."""
    import torch
    lib = _load()
    data = _as_bytes(path_or_bytes)
    mask = np.ascontiguousarray(np.asarray(lost_mask, dtype=np.uint8))
    n16 = probe(data)
    out = torch.empty(n16, dtype=torch.float32, device="cuda")
    if n16 == 0:
        return out
    out_len = ctypes.c_int64(0)
    err = ctypes.create_string_buffer(_ERRCAP)
    _check(lib.cuopus_decode_lossy_f16k_on(data, len(data), mask.ctypes.data, mask.size, out.data_ptr(), n16,
                                           ctypes.byref(out_len), _stream(), err, _ERRCAP), err)
    return out


def decode_ogg_to_s16_tensor(path_or_bytes):
    """Playable int16 48 kHz PCM (preroll and pre-skip removed) as a device tensor."""
    import torch
    lib = _load()
    data = _as_bytes(path_or_bytes)
    npk, nch = packet_count(data)
    del npk, nch
    n16 = probe(data)
    # playable = 3*n16 + 48 when n16 > 0; size generously and trim
    cap = 3 * n16 + 48 + 3 if n16 > 0 else 48 + 3
    out = torch.empty(cap, dtype=torch.int16, device="cuda")
    out_len = ctypes.c_int64(0)
    err = ctypes.create_string_buffer(_ERRCAP)
    _check(lib.cuopus_decode_s16_on(data, len(data), out.data_ptr(), cap, ctypes.byref(out_len), _stream(),
                                    err, _ERRCAP), err)
    return out[:out_len.value]


def decode_ogg_multi(items):
    """Decode N files in one decode launch; returns a list of float32 16 kHz device tensors."""
    import torch
    lib = _load()
    blobs = [_as_bytes(x) for x in items]
    n = len(blobs)
    if n == 0:
        return []
    n16 = [probe(b) for b in blobs]
    outs = [torch.empty(k, dtype=torch.float32, device="cuda") for k in n16]
    ba = (ctypes.c_char_p * n)(*blobs)
    la = (SZ * n)(*[len(b) for b in blobs])
    oa = (P * n)(*[o.data_ptr() for o in outs])
    ca = (SZ * n)(*n16)
    ol = (ctypes.c_int64 * n)()
    err = ctypes.create_string_buffer(_ERRCAP)
    _check(lib.cuopus_decode_f16k_multi(ba, la, n, oa, ca, ol, _stream(), err, _ERRCAP), err)
    for i in range(n):
        if ol[i] != n16[i]:
            raise CuopusError(1, "file %d: decode wrote %d samples, probe said %d" % (i, ol[i], n16[i]))
    return outs


def gate_chains(data, off, length, chains, lost=None):
    """Per-packet gate on the GPU.

    data: uint8 packet bytes; off/length: uint32 per packet; chains: int array
    (n, 4) of (first, npkts, trim, keep).  Returns (pcm int16 sum(keep),
    nsamp int32 per packet, rng uint32 per packet, stop int32 per chain, rc).
    """
    lib = _load()
    data = np.ascontiguousarray(data, dtype=np.uint8)
    off = np.ascontiguousarray(off, dtype=np.uint32)
    length = np.ascontiguousarray(length, dtype=np.uint32)
    ch = np.ascontiguousarray(np.asarray(chains, dtype=np.uint32).reshape(-1, 4))
    first, npk, trim, keep = [np.ascontiguousarray(ch[:, i]) for i in range(4)]
    n_pk, n_ch = off.size, first.size
    pcm = np.zeros(int(keep.sum(dtype=np.uint64)), np.int16)
    nsamp = np.zeros(n_pk, np.int32)
    rng = np.zeros(n_pk, np.uint32)
    stop = np.zeros(n_ch, np.int32)
    lp = None
    if lost is not None:
        lost = np.ascontiguousarray(lost, dtype=np.uint8)
        lp = lost.ctypes.data
    err = ctypes.create_string_buffer(_ERRCAP)
    rc = lib.cuopus_gate_chains(data.ctypes.data, data.size, off.ctypes.data, length.ctypes.data, n_pk, lp,
                                first.ctypes.data, npk.ctypes.data, trim.ctypes.data, keep.ctypes.data, n_ch,
                                pcm.ctypes.data, nsamp.ctypes.data, rng.ctypes.data, stop.ctypes.data,
                                _stream(), err, _ERRCAP)
    if rc not in (0, 1):
        _check(rc, err)
    return pcm, nsamp, rng, stop, rc


def warmup():
    lib = _load()
    err = ctypes.create_string_buffer(_ERRCAP)
    _check(lib.cuopus_warmup(err, _ERRCAP), err)


def workspace_bytes() -> int:
    return int(_load().cuopus_workspace_bytes())


def last_arena_overflow() -> bool:
    return bool(_load().cuopus_last_arena_overflow())


def release_workspaces():
    _load().cuopus_release_workspaces()


PROF_SCOPES = ["chain", "frame", "celt", "celt_head", "celt_bands", "celt_synth", "celt_comb", "celt_deemph",
               "silk", "emit", "celt_lost", "celt_energy", "celt_alloc", "celt_tail", "pvq", "cwrsi", "exp_rot",
               "theta", "leaf_fold", "band_pre", "band_post", "ec_uint", "alloc_bisect", "alloc_tail",
               "qp_leaf_bits", "bands_prologue", "lowband_out", "qp_stage", "silk_indices", "silk_pulses",
               "silk_params", "silk_core", "silk_resamp", "silk_plc", "silk_nlsf", "silk_nlsf2a", "silk_cng",
               "silk_pulse_counts", "silk_shell", "silk_signs", "qp_leaf"]


def prof_read():
    """Per-scope shader-clock cycle sums of a -DCUO_PROF library (cleared on
    read); empty dict for a normal build."""
    buf = (ctypes.c_ulonglong * 64)()
    n = min(_load().cuopus_prof_read(buf, 64), len(PROF_SCOPES))
    return {PROF_SCOPES[i]: int(buf[i]) for i in range(max(n, 0)) if buf[i]}


MODE_PLAIN, MODE_ONE_LAUNCH, MODE_PERSISTENT = 0, 1, 2


class Batch:
    """A resident batch: packets demuxed and uploaded once, then decoded on the
    device as many times as asked (run), without host work per run.

    mode: MODE_PLAIN (one decode and one resample launch per file),
    MODE_ONE_LAUNCH (every chain in one decode launch) or MODE_PERSISTENT
    (n_slots wavefronts pulling chains from a device work queue).
    Outputs are float32 16 kHz device tensors, one per file (``outputs``).
    """

    def __init__(self, items, lost_masks=None, mode=MODE_PERSISTENT, n_slots=0, lpt_order=True,
                 use_graph=False, profile=False, outputs=True):
        import torch
        lib = _load()
        self._lib = lib
        self._h = P(0)
        self.profiled = bool(profile) and not use_graph
        blobs = [_as_bytes(x) for x in items]
        n = len(blobs)
        if n == 0:
            raise CuopusError(3, "no files")
        self.n16 = [probe(b) for b in blobs]
        self.outputs = [torch.empty(k, dtype=torch.float32, device="cuda") for k in self.n16] if outputs else None
        ba = (ctypes.c_char_p * n)(*blobs)
        la = (SZ * n)(*[len(b) for b in blobs])
        oa = (P * n)(*[o.data_ptr() for o in self.outputs]) if outputs else None
        ca = (SZ * n)(*self.n16) if outputs else None
        ma = mn = None
        self._masks = None
        if lost_masks is not None:
            self._masks = [None if m is None else np.ascontiguousarray(np.asarray(m, dtype=np.uint8))
                           for m in lost_masks]
            ma = (P * n)(*[None if m is None else m.ctypes.data for m in self._masks])
            mn = (SZ * n)(*[0 if m is None else m.size for m in self._masks])
        opts = BatchOpts(mode, n_slots, 1 if lpt_order else 0, 1 if use_graph else 0, 1 if profile else 0)
        err = ctypes.create_string_buffer(_ERRCAP)
        _check(lib.cuopus_batch_create(ba, la, n, ma, mn, oa, ca, ctypes.byref(opts), ctypes.byref(self._h),
                                       err, _ERRCAP), err)
        c = (ctypes.c_int64 * 8)()
        lib.cuopus_batch_info(self._h, c)
        (self.chains, self.packets, self.playable, self.decoded, self.out16, self.packet_bytes,
         self.chain_packets, self.chain_packet_bytes) = list(c)

    def run(self, stream=None):
        err = ctypes.create_string_buffer(_ERRCAP)
        _check(self._lib.cuopus_batch_run(self._h, stream if stream is not None else _stream(), err, _ERRCAP), err)

    def check(self, stream=None):
        """Wait for the last run and verify every chain decoded fully; returns output lengths."""
        n = len(self.n16)
        ol = (ctypes.c_int64 * n)()
        err = ctypes.create_string_buffer(_ERRCAP)
        _check(self._lib.cuopus_batch_check(self._h, stream if stream is not None else _stream(), ol, err,
                                            _ERRCAP), err)
        return list(ol)

    def kernel_ms(self):
        """(decode, resample) milliseconds of the last profiled run (events on the launch stream)."""
        d, r = ctypes.c_float(0), ctypes.c_float(0)
        rc = self._lib.cuopus_batch_kernel_ms(self._h, ctypes.byref(d), ctypes.byref(r))
        if rc:
            raise CuopusError(rc, "batch was not created with profile=True (or uses a graph)")
        return d.value, r.value

    def close(self):
        if self._h:
            self._lib.cuopus_batch_destroy(self._h)
            self._h = P(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching decode server (see cuopus_server_* in include/cuopus.h).

    submit(ogg) demuxes on the host, allocates the float32 16 kHz output
    tensor and queues the request; poll() admits queued requests as streams
    and device memory free up, and returns [(request id, tensor, rc)] for the
    requests that finished.  Never blocks.
    """

    def __init__(self, n_streams=4, n_waves=0, mem_budget=8 << 30, max_chains_per_launch=0):
        self._lib = _load()
        self._h = P(0)
        opts = ServerOpts(n_streams, n_waves, mem_budget, max_chains_per_launch)
        err = ctypes.create_string_buffer(_ERRCAP)
        _check(self._lib.cuopus_server_create(ctypes.byref(opts), ctypes.byref(self._h), err, _ERRCAP), err)
        self._out = {}
        self._ids = (ctypes.c_uint64 * 256)()
        self._rcs = (ctypes.c_int * 256)()

    def submit(self, path_or_bytes, lost_mask=None):
        import torch
        data = _as_bytes(path_or_bytes)
        n16 = probe(data)
        out = torch.empty(n16, dtype=torch.float32, device="cuda")
        mp, mn = None, 0
        if lost_mask is not None:
            m = np.ascontiguousarray(np.asarray(lost_mask, dtype=np.uint8))
            mp, mn = m.ctypes.data, m.size
        rid = ctypes.c_uint64(0)
        err = ctypes.create_string_buffer(_ERRCAP)
        _check(self._lib.cuopus_server_submit(self._h, data, len(data), mp, mn, out.data_ptr() if n16 else None,
                                              n16, ctypes.byref(rid), err, _ERRCAP), err)
        self._out[rid.value] = out
        return rid.value

    def poll(self):
        done = []
        while True:
            n = ctypes.c_int(0)
            err = ctypes.create_string_buffer(_ERRCAP)
            _check(self._lib.cuopus_server_poll(self._h, self._ids, self._rcs, 256, ctypes.byref(n), err, _ERRCAP),
                   err)
            for i in range(n.value):
                rid = int(self._ids[i])
                done.append((rid, self._out.pop(rid), int(self._rcs[i])))
            if n.value < 256:
                return done

    def stats(self):
        s = (ctypes.c_int64 * 4)()
        self._lib.cuopus_server_stats(self._h, s)
        return {"queued": s[0], "in_flight": s[1], "admitted_bytes": s[2], "finished": s[3]}

    def close(self):
        if self._h:
            self._lib.cuopus_server_destroy(self._h)
            self._h = P(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
