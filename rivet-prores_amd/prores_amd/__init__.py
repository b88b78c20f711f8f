"""prores_amd — the MI355X ProRes slice engine, Python face.

Mirrors the rivet-prores crate's API (src/lib.rs:52-56): ``Decoder``,
``Encoder``, ``Config``, ``Profile``, ``Frame``, ``FrameHeader``, ``Error``
(``Invalid`` / ``Unsupported`` / ``Config``), over the C ABI of
``include/prores_hip.h`` (``libprores_hip.so``). Every decode and encode runs
the gfx950 kernels; there is no CPU fallback: without the built library or a
GPU the calls raise.
"""

import ctypes as C
import dataclasses
import enum
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libprores_hip.so")

# ---------------------------------------------------------------------------
# Errors (src/error.rs:5-21). Equal when the kind and the message are equal,
# as the reference's `Error: PartialEq` is.


class Error(Exception):
    kind = "Error"
    prefix = ""

    def __init__(self, msg):
        super().__init__(msg)
        self.msg = msg

    def __eq__(self, other):
        return type(self) is type(other) and self.msg == other.msg

    def __hash__(self):
        return hash((self.kind, self.msg))

    def __str__(self):
        return self.prefix + self.msg


class Invalid(Error):
    kind = "Invalid"
    prefix = "invalid ProRes data: "


class Unsupported(Error):
    kind = "Unsupported"
    prefix = "unsupported ProRes feature: "


class ConfigError(Error):
    kind = "Config"
    prefix = "invalid ProRes configuration: "


class DeviceError(Error):
    """A HIP runtime failure (no device, out of memory): not a codec error."""
    kind = "Device"
    prefix = "device: "


_KINDS = {1: Invalid, 2: Unsupported, 3: ConfigError, 4: DeviceError}

# ---------------------------------------------------------------------------
# C structures (include/prores_hip.h).


class _Err(C.Structure):
    _fields_ = [("kind", C.c_int32), ("msg", C.c_char * 256)]


class _Meta(C.Structure):
    _fields_ = [("aspect_ratio", C.c_uint8), ("frame_rate_code", C.c_uint8),
                ("color_primaries", C.c_uint8), ("transfer_characteristic", C.c_uint8),
                ("matrix_coefficients", C.c_uint8)]


class _Plane(C.Structure):
    _fields_ = [("offset", C.c_uint64), ("width", C.c_uint32), ("height", C.c_uint32)]


class _Frame(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("bit_depth", C.c_uint32),
                ("chroma", C.c_uint32), ("interlace", C.c_uint32), ("metadata", _Meta),
                ("data", C.c_void_p), ("data_len", C.c_size_t), ("planes", _Plane * 3),
                ("alpha", C.c_void_p), ("alpha_len", C.c_size_t)]


class _Header(C.Structure):
    _fields_ = [("frame_size", C.c_uint32), ("header_size", C.c_uint16),
                ("bitstream_version", C.c_uint8), ("encoder_identifier", C.c_uint8 * 4),
                ("width", C.c_uint16), ("height", C.c_uint16), ("chroma", C.c_uint32),
                ("interlace", C.c_uint32), ("metadata", _Meta), ("alpha", C.c_uint32),
                ("has_luma_matrix", C.c_uint8), ("has_chroma_matrix", C.c_uint8),
                ("luma_matrix", C.c_uint8 * 64), ("chroma_matrix", C.c_uint8 * 64)]


class _Config(C.Structure):
    _fields_ = [("profile", C.c_uint32), ("log2_slice_mbs", C.c_uint32), ("alpha", C.c_uint32),
                ("has_luma_matrix", C.c_uint8), ("has_chroma_matrix", C.c_uint8),
                ("luma_matrix", C.c_uint8 * 64), ("chroma_matrix", C.c_uint8 * 64),
                ("target_frame_bytes", C.c_uint64), ("encoder_identifier", C.c_uint8 * 4),
                ("threads", C.c_uint32)]


class _Tuning(C.Structure):
    _fields_ = [("persistent", C.c_uint32), ("chunk", C.c_uint32), ("idct_blocks", C.c_uint32),
                ("entropy_wg", C.c_uint32), ("graph", C.c_uint32)]


class _ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_batch_frames", C.c_uint32), ("decode_bit_depth", C.c_uint32)]


class _ServerResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("kind", C.c_uint32), ("error", _Err), ("dev_out", C.c_void_p),
                ("out_bytes", C.c_size_t), ("dev_alpha", C.c_void_p), ("alpha_bytes", C.c_size_t),
                ("header", _Header), ("submit_us", C.c_double), ("start_us", C.c_double),
                ("done_us", C.c_double), ("batch_frames", C.c_uint32)]


class _ServerStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("submitted", "completed", "failed", "decode_batches", "encode_batches",
                                           "max_batch_frames", "held_bytes", "peak_held_bytes", "hbm_budget",
                                           "pending", "staged_ahead")]


class _DevDecodeJob(C.Structure):
    _fields_ = [("host_packet", C.c_void_p), ("len", C.c_size_t), ("dev_packet", C.c_void_p),
                ("dev_data", C.c_void_p), ("dev_alpha", C.c_void_p)]


class _DevEncodeJob(C.Structure):
    _fields_ = [("cfg", C.POINTER(_Config)), ("frame", _Frame), ("dev_out", C.c_void_p),
                ("out_len", C.c_size_t)]


# The exported symbols of include/prores_hip.h.
EXPORTS = [
    "prores_decoder_create", "prores_decoder_destroy", "prores_decode", "prores_decode_into",
    "prores_frame_free", "prores_frame_header_parse", "prores_frame_layout",
    "prores_config_default", "prores_encode", "prores_buffer_free",
    "prores_profile_target_frame_bytes", "prores_profile_fourcc", "prores_profile_from_fourcc",
    "prores_decode_device_batch", "prores_encode_device_batch", "prores_last_batch_timing",
    "prores_last_encode_timing", "prores_version", "prores_device_alloc", "prores_device_free",
    "prores_memcpy", "prores_decode_tuning_default", "prores_decode_device_batch_tuned",
    "prores_set_decode_tuning", "prores_get_decode_tuning", "prores_autotune_decode",
    "prores_server_config_default", "prores_server_create", "prores_server_destroy",
    "prores_server_submit_decode", "prores_server_submit_encode", "prores_server_poll",
    "prores_server_release", "prores_server_get_stats", "prores_last_encode_counters",
    "prores_entropy_tables", "prores_memcpy_error", "prores_decode_host_batch", "prores_find_frames",
    "prores_fnv1a", "prores_graph_counters", "prores_last_encode_stages", "prores_noalloc_init",
    "prores_noalloc_scratch_bytes", "prores_decode_device_batch_noalloc",
]

_lib = None


def lib():
    """The loaded C library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run rivet-prores_amd/build.py")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.prores_decoder_create.argtypes = [C.c_uint32, C.c_uint32, P(C.c_void_p), P(_Err)]
        L.prores_decoder_destroy.argtypes = [C.c_void_p]
        L.prores_decode.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t, P(_Frame), P(_Err)]
        L.prores_decode_into.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t, P(_Frame), P(_Err)]
        L.prores_frame_free.argtypes = [P(_Frame)]
        L.prores_decode_host_batch.argtypes = [C.c_void_p, C.c_size_t, P(C.c_char_p), P(C.c_size_t), P(_Frame),
                                               P(_Err)]
        L.prores_frame_header_parse.argtypes = [C.c_char_p, C.c_size_t, P(_Header), P(_Err)]
        L.prores_frame_layout.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, P(_Plane)]
        L.prores_frame_layout.restype = C.c_size_t
        L.prores_config_default.argtypes = [C.c_uint32, P(_Config)]
        L.prores_encode.argtypes = [P(_Config), P(_Frame), P(C.c_void_p), P(C.c_size_t), P(_Err)]
        L.prores_buffer_free.argtypes = [C.c_void_p]
        L.prores_profile_target_frame_bytes.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.prores_profile_target_frame_bytes.restype = C.c_uint64
        L.prores_profile_fourcc.argtypes = [C.c_uint32, C.c_char_p]
        L.prores_profile_from_fourcc.argtypes = [C.c_char_p]
        L.prores_profile_from_fourcc.restype = C.c_int32
        L.prores_decode_device_batch.argtypes = [C.c_void_p, C.c_size_t, P(_DevDecodeJob), P(_Err)]
        L.prores_decode_device_batch_tuned.argtypes = [C.c_void_p, C.c_size_t, P(_DevDecodeJob), P(_Tuning),
                                                       P(_Err)]
        L.prores_decode_tuning_default.argtypes = [P(_Tuning)]
        L.prores_set_decode_tuning.argtypes = [P(_Tuning), P(_Err)]
        L.prores_get_decode_tuning.argtypes = [P(_Tuning)]
        L.prores_server_config_default.argtypes = [P(_ServerConfig)]
        L.prores_server_create.argtypes = [P(_ServerConfig), P(C.c_void_p), P(_Err)]
        L.prores_server_destroy.argtypes = [C.c_void_p]
        L.prores_server_submit_decode.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
        L.prores_server_submit_decode.restype = C.c_uint64
        L.prores_server_submit_encode.argtypes = [C.c_void_p, P(_Config), P(_Frame)]
        L.prores_server_submit_encode.restype = C.c_uint64
        L.prores_server_poll.argtypes = [C.c_void_p, P(_ServerResult), C.c_size_t, C.c_int32]
        L.prores_server_poll.restype = C.c_size_t
        L.prores_server_release.argtypes = [C.c_void_p, C.c_uint64]
        L.prores_server_get_stats.argtypes = [C.c_void_p, P(_ServerStats)]
        L.prores_autotune_decode.argtypes = [C.c_void_p, C.c_size_t, P(_DevDecodeJob), C.c_uint32, P(_Tuning),
                                             P(C.c_double), P(_Err)]
        L.prores_encode_device_batch.argtypes = [C.c_size_t, P(_DevEncodeJob), P(_Err)]
        L.prores_last_batch_timing.argtypes = [P(C.c_double)]
        L.prores_last_encode_timing.argtypes = [P(C.c_double)]
        L.prores_last_encode_counters.argtypes = [P(C.c_uint64)]
        L.prores_graph_counters.argtypes = [P(C.c_uint64)]
        L.prores_last_encode_stages.argtypes = [P(C.c_double)]
        L.prores_noalloc_init.argtypes = [P(_Err)]
        L.prores_noalloc_scratch_bytes.restype = C.c_size_t
        L.prores_decode_device_batch_noalloc.argtypes = [C.c_void_p, C.c_size_t, P(_DevDecodeJob), P(_Err)]
        L.prores_version.restype = C.c_char_p
        L.prores_find_frames.argtypes = [C.c_char_p, C.c_size_t, P(C.c_uint64), P(C.c_uint64), C.c_size_t]
        L.prores_find_frames.restype = C.c_size_t
        L.prores_fnv1a.argtypes = [C.c_uint64, C.c_void_p, C.c_size_t]
        L.prores_fnv1a.restype = C.c_uint64
        L.prores_memcpy_error.restype = C.c_char_p
        L.prores_device_alloc.argtypes = [C.c_size_t]
        L.prores_device_alloc.restype = C.c_void_p
        L.prores_device_free.argtypes = [C.c_void_p]
        L.prores_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int32]
        _lib = L
    return _lib


def _raise(err):
    if err.kind:
        raise _KINDS.get(err.kind, Error)(err.msg.decode("utf-8", "replace"))


# ---------------------------------------------------------------------------
# Frame types (src/frame.rs).


class ChromaFormat(enum.IntEnum):
    Yuv422 = 2
    Yuv444 = 3

    def subsampling(self):
        return (2, 1) if self is ChromaFormat.Yuv422 else (1, 1)


class Interlace(enum.IntEnum):
    Progressive = 0
    TopFieldFirst = 1
    BottomFieldFirst = 2


class AlphaType(enum.IntEnum):
    None_ = 0
    Bits8 = 1
    Bits16 = 2

    def bits(self):
        return (0, 8, 16)[self]


class Metadata:
    """src/frame.rs:83-110; the default is every field unspecified."""

    def __init__(self, aspect_ratio=0, frame_rate_code=0, color_primaries=2,
                 transfer_characteristic=2, matrix_coefficients=2):
        self.aspect_ratio = aspect_ratio
        self.frame_rate_code = frame_rate_code
        self.color_primaries = color_primaries
        self.transfer_characteristic = transfer_characteristic
        self.matrix_coefficients = matrix_coefficients

    def astuple(self):
        return (self.aspect_ratio, self.frame_rate_code, self.color_primaries,
                self.transfer_characteristic, self.matrix_coefficients)

    def __eq__(self, o):
        return isinstance(o, Metadata) and self.astuple() == o.astuple()

    _RATES = {1: (24000, 1001), 2: (24, 1), 3: (25, 1), 4: (30000, 1001), 5: (30, 1), 6: (50, 1),
              7: (60000, 1001), 8: (60, 1), 9: (100, 1), 10: (120000, 1001), 11: (120, 1)}

    def frame_rate(self):
        return self._RATES.get(self.frame_rate_code)

    @staticmethod
    def frame_rate_code_for(num, den):
        for c, (n, d) in Metadata._RATES.items():
            if n * den == num * d:
                return c
        return 0


def layout(width, height, chroma):
    """Frame::layout (src/frame.rs:247-266): [(offset, width, height)] x 3."""
    cw = (width + 1) // 2 if int(chroma) == 2 else width
    y = (0, width, height)
    cb = (width * height, cw, height)
    cr = (cb[0] + cw * height, cw, height)
    return [y, cb, cr]


class Frame:
    """A picture (src/frame.rs:180-203): ``data`` holds Y' then Cb then Cr as
    u16, each tightly packed; ``alpha`` is a separate plane or None."""

    def __init__(self, width, height, chroma, bit_depth, data=None, alpha=None,
                 interlace=Interlace.Progressive, metadata=None):
        self.width, self.height = int(width), int(height)
        self.chroma = ChromaFormat(int(chroma))
        self.bit_depth = int(bit_depth)
        self.interlace = Interlace(int(interlace))
        self.metadata = metadata if metadata is not None else Metadata()
        self.planes = layout(self.width, self.height, self.chroma)
        total = self.planes[2][0] + self.planes[2][1] * self.planes[2][2]
        if data is None:
            data = np.full(total, 1 << (self.bit_depth - 1), dtype=np.uint16)
        self.data = np.ascontiguousarray(data, dtype=np.uint16)
        self.alpha = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.uint16)

    @staticmethod
    def new(width, height, chroma, bit_depth):
        """Frame::new (src/frame.rs:209-213): mid-grey, checked."""
        if width == 0 or height == 0 or width > 65535 or height > 65535:
            raise ConfigError(f"a frame of {width}×{height} cannot be coded (1–65535 each way)")
        if not 8 <= bit_depth <= 16:
            raise ConfigError(f"bit depth {bit_depth} is outside 8–16")
        return Frame(width, height, chroma, bit_depth)

    def plane(self, i):
        off, w, h = self.planes[i]
        return self.data[off:off + w * h].reshape(h, w)

    def packed(self):
        return self.data

    def to_le_bytes(self):
        return self.data.astype("<u2").tobytes()

    def _c(self):
        f = _Frame()
        f.width, f.height, f.bit_depth = self.width, self.height, self.bit_depth
        f.chroma, f.interlace = int(self.chroma), int(self.interlace)
        (f.metadata.aspect_ratio, f.metadata.frame_rate_code, f.metadata.color_primaries,
         f.metadata.transfer_characteristic, f.metadata.matrix_coefficients) = self.metadata.astuple()
        f.data = self.data.ctypes.data
        f.data_len = self.data.size
        for i, (o, w, h) in enumerate(self.planes):
            f.planes[i].offset, f.planes[i].width, f.planes[i].height = o, w, h
        if self.alpha is not None:
            f.alpha = self.alpha.ctypes.data
            f.alpha_len = self.alpha.size
        return f

    def __eq__(self, o):
        return (isinstance(o, Frame) and (self.width, self.height, self.bit_depth, self.chroma, self.interlace)
                == (o.width, o.height, o.bit_depth, o.chroma, o.interlace) and self.metadata == o.metadata
                and np.array_equal(self.data, o.data)
                and ((self.alpha is None and o.alpha is None)
                     or (self.alpha is not None and o.alpha is not None and np.array_equal(self.alpha, o.alpha))))


class FrameHeader:
    """FrameHeader::parse (src/header.rs:82-170)."""

    @staticmethod
    def parse(data):
        h = _Header()
        err = _Err()
        data = bytes(data)
        lib().prores_frame_header_parse(data, len(data), C.byref(h), C.byref(err))
        _raise(err)
        out = FrameHeader()
        out.frame_size, out.header_size, out.bitstream_version = h.frame_size, h.header_size, h.bitstream_version
        out.encoder_identifier = bytes(h.encoder_identifier)
        out.width, out.height = h.width, h.height
        out.chroma, out.interlace = ChromaFormat(h.chroma), Interlace(h.interlace)
        out.metadata = Metadata(h.metadata.aspect_ratio, h.metadata.frame_rate_code, h.metadata.color_primaries,
                                h.metadata.transfer_characteristic, h.metadata.matrix_coefficients)
        out.alpha = AlphaType(h.alpha)
        out.luma_matrix = bytes(h.luma_matrix) if h.has_luma_matrix else None
        out.chroma_matrix = bytes(h.chroma_matrix) if h.has_chroma_matrix else None
        return out

    def picture_count(self):
        return 1 if self.interlace == Interlace.Progressive else 2

    def picture_height(self, index):
        if self.interlace == Interlace.Progressive:
            return self.height
        top = (self.interlace == Interlace.TopFieldFirst) == (index == 0)
        return (self.height + 1) // 2 if top else self.height // 2


# ---------------------------------------------------------------------------
# Decoder (src/decode.rs:19-104).


class _FrameMemory:
    """Owns the planes of a frame the library allocated: freed with
    prores_frame_free once no array over them remains."""

    def __init__(self, f):
        self.f = f
        weakref.finalize(self, _free_frame, f)


def _free_frame(f):
    if _lib is not None:
        _lib.prores_frame_free(C.byref(f))


class _PlaneView:
    """An array interface over library memory that keeps its owner alive."""

    def __init__(self, owner, ptr, n):
        self.owner = owner
        self.__array_interface__ = {"data": (ptr, False), "shape": (n,), "typestr": "<u2", "version": 3}


def _owned_frame(f):
    """A Frame over a library-allocated prores_frame, without copying."""
    owner = _FrameMemory(f)
    data = np.asarray(_PlaneView(owner, f.data, f.data_len))
    alpha = np.asarray(_PlaneView(owner, f.alpha, f.alpha_len)) if f.alpha else None
    m = f.metadata
    return Frame(f.width, f.height, f.chroma, f.bit_depth, data, alpha, f.interlace,
                 Metadata(m.aspect_ratio, m.frame_rate_code, m.color_primaries, m.transfer_characteristic,
                          m.matrix_coefficients))


class Decoder:
    def __init__(self, bit_depth=0, threads=0):
        err = _Err()
        h = C.c_void_p()
        lib().prores_decoder_create(bit_depth, threads, C.byref(h), C.byref(err))
        _raise(err)
        self._h = h
        self.bit_depth = bit_depth
        self.threads = threads

    @staticmethod
    def new():
        return Decoder()

    @staticmethod
    def with_bit_depth(bit_depth):
        return Decoder(bit_depth)

    def with_threads(self, threads):
        return Decoder(self.bit_depth, threads)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.prores_decoder_destroy(self._h)
            self._h = None

    def decode(self, data):
        """Decodes one frame on the GPU; raises Invalid / Unsupported."""
        data = bytes(data)
        f = _Frame()
        err = _Err()
        lib().prores_decode(self._h, data, len(data), C.byref(f), C.byref(err))
        _raise(err)
        return _owned_frame(f)

    def decode_batch(self, packets):
        """Decodes many frames from host memory (prores_decode_host_batch):
        one Frame or one Error per packet, in order."""
        packets = [bytes(p) for p in packets]
        n = len(packets)
        ptrs = (C.c_char_p * n)(*packets)
        lens = (C.c_size_t * n)(*[len(p) for p in packets])
        frames = (_Frame * n)()
        errs = (_Err * n)()
        rc = lib().prores_decode_host_batch(self._h, n, ptrs, lens, frames, errs)
        out = []
        for i in range(n):
            if errs[i].kind == 0:
                f = _Frame()
                C.pointer(f)[0] = frames[i]
                out.append(_owned_frame(f))
            else:
                out.append(_KINDS.get(errs[i].kind, Error)(errs[i].msg.decode()))
        if rc == 4 and all(isinstance(o, Error) for o in out):
            raise out[0]
        return out

    def decode_device_batch(self, jobs, tuning=None):
        """Decodes device-resident packets: ``jobs`` is a DecodeJobs or a list
        of (host_packet bytes, dev_packet ptr, dev_data ptr, dev_alpha ptr or
        0). ``tuning`` (a DecodeTuning) overrides the device's batch tuning.
        Returns one Error or None per job."""
        if not isinstance(jobs, DecodeJobs):
            jobs = DecodeJobs(jobs)
        if tuning is None:
            rc = lib().prores_decode_device_batch(self._h, jobs.n, jobs.arr, jobs.errs)
        else:
            t = tuning._c()
            rc = lib().prores_decode_device_batch_tuned(self._h, jobs.n, jobs.arr, C.byref(t), jobs.errs)
        if rc == 4 or (rc and tuning is not None):
            _raise(jobs.errs[0])
        if not jobs._kinds.any():
            return [None] * jobs.n
        return [None if e.kind == 0 else _KINDS.get(e.kind, Error)(e.msg.decode()) for e in jobs.errs]

    def decode_device_batch_noalloc(self, jobs):
        """decode_device_batch through the allocation-free entry point
        (prores_decode_device_batch_noalloc): the batch is planned on the
        device and its workspace is the library's scratch array; call
        noalloc_init() on the device first. Returns one Error or None per
        job."""
        if not isinstance(jobs, DecodeJobs):
            jobs = DecodeJobs(jobs)
        rc = lib().prores_decode_device_batch_noalloc(self._h, jobs.n, jobs.arr, jobs.errs)
        if rc == 4:
            _raise(jobs.errs[0])
        return [None if e.kind == 0 else _KINDS.get(e.kind, Error)(e.msg.decode()) for e in jobs.errs]

    def autotune(self, jobs, reps=3):
        """Times every candidate DecodeTuning on this batch and makes the
        fastest the device's batch tuning; returns (tuning, median us)."""
        if not isinstance(jobs, DecodeJobs):
            jobs = DecodeJobs(jobs)
        t, us, err = _Tuning(), C.c_double(), _Err()
        if lib().prores_autotune_decode(self._h, jobs.n, jobs.arr, reps, C.byref(t), C.byref(us), C.byref(err)):
            _raise(err)
        return DecodeTuning._from(t), us.value


class DecodeJobs:
    """A prepared list of device decode jobs, for repeated batches. Each job
    is (packet, dev_packet, dev_data, dev_alpha): `packet` is the host copy of
    the packet (read for its headers only; kept referenced, not copied), or
    an int, the packet's length, when the packet is only in device memory
    and its headers are to be read there."""

    def __init__(self, jobs):
        self.n = len(jobs)
        self.arr = (_DevDecodeJob * self.n)()
        self.errs = (_Err * self.n)()
        # The result kinds, read in one step after each batch.
        self._kinds = np.frombuffer(self.errs, dtype=np.int32).reshape(self.n, C.sizeof(_Err) // 4)[:, 0]
        self._keep = []
        for i, (hp, dp, dd, da) in enumerate(jobs):
            if isinstance(hp, int):
                self.arr[i].host_packet = None
                self.arr[i].len = hp
            else:
                hp = bytes(hp)
                self._keep.append(hp)
                self.arr[i].host_packet = C.cast(C.c_char_p(hp), C.c_void_p)
                self.arr[i].len = len(hp)
            self.arr[i].dev_packet, self.arr[i].dev_data, self.arr[i].dev_alpha = dp, dd, da or None


@dataclasses.dataclass
class DecodeTuning:
    """How a decode batch is launched (prores_decode_tuning): every setting
    decodes the same samples."""
    persistent: int = 0
    chunk: int = 1
    idct_blocks: int = 96
    entropy_wg: int = 0
    graph: int = 0

    def _c(self):
        return _Tuning(self.persistent, self.chunk, self.idct_blocks, self.entropy_wg, self.graph)

    @classmethod
    def _from(cls, t):
        return cls(t.persistent, t.chunk, t.idct_blocks, t.entropy_wg, t.graph)


def noalloc_init():
    """Set-up of the allocation-free entry point on the current device."""
    err = _Err()
    if lib().prores_noalloc_init(C.byref(err)):
        _raise(err)


def noalloc_scratch_bytes():
    return int(lib().prores_noalloc_scratch_bytes())


def set_decode_tuning(tuning):
    err = _Err()
    t = tuning._c()
    if lib().prores_set_decode_tuning(C.byref(t), C.byref(err)):
        _raise(err)


def get_decode_tuning():
    t = _Tuning()
    lib().prores_get_decode_tuning(C.byref(t))
    return DecodeTuning._from(t)


def last_batch_timing():
    """[slice table, colour entropy, alpha, idct, all kernels] in microseconds."""
    us = (C.c_double * 5)()
    lib().prores_last_batch_timing(us)
    return list(us)


def last_encode_counters():
    """[share-out sizes past the window, most for one picture, pictures,
    mean uniform qi] of the last encode batch."""
    out = (C.c_uint64 * 4)()
    lib().prores_last_encode_counters(out)
    return list(out)


def graph_counters():
    """[captures made, captures that failed, graph launches] of the decode
    batches on the current device."""
    out = (C.c_uint64 * 3)()
    lib().prores_graph_counters(out)
    return list(out)


ENCODE_STAGES = ["fdct", "alpha", "probes", "window", "shareout", "offsets", "pack", "all"]


def last_encode_stages():
    """Microseconds of the last encode batch by kernel (ENCODE_STAGES)."""
    us = (C.c_double * 8)()
    lib().prores_last_encode_stages(us)
    return dict(zip(ENCODE_STAGES, us))


def last_encode_timing():
    us = (C.c_double * 4)()
    lib().prores_last_encode_timing(us)
    return list(us)


# ---------------------------------------------------------------------------
# Encoder (src/encode.rs:26-287).


class Profile(enum.IntEnum):
    Proxy = 0
    Lt = 1
    Standard = 2
    Hq = 3
    P4444 = 4
    P4444Xq = 5

    def fourcc(self):
        buf = C.create_string_buffer(5)
        lib().prores_profile_fourcc(int(self), buf)
        return buf.raw[:4]

    @staticmethod
    def from_fourcc(code):
        p = lib().prores_profile_from_fourcc(bytes(code))
        return None if p < 0 else Profile(p)

    def chroma(self):
        return ChromaFormat.Yuv444 if self >= Profile.P4444 else ChromaFormat.Yuv422

    def reference_bitrate(self):
        return (45_000_000, 102_000_000, 147_000_000, 220_000_000, 330_000_000, 500_000_000)[self]

    def target_frame_bytes(self, width, height):
        return int(lib().prores_profile_target_frame_bytes(int(self), width, height))


Profile.ALL = list(Profile)


class Config:
    """src/encode.rs:135-193; ``Config(profile)`` is ``Config::new``."""

    def __init__(self, profile, log2_slice_mbs=3, alpha=AlphaType.Bits16, luma_matrix=None,
                 chroma_matrix=None, target_frame_bytes=None, encoder_identifier=b"rivt", threads=0):
        self.profile = Profile(int(profile))
        self.log2_slice_mbs = log2_slice_mbs
        self.alpha = AlphaType(int(alpha))
        self.luma_matrix = luma_matrix
        self.chroma_matrix = chroma_matrix
        self.target_frame_bytes = target_frame_bytes
        self.encoder_identifier = bytes(encoder_identifier)
        self.threads = threads

    def _c(self):
        c = _Config()
        lib().prores_config_default(int(self.profile), C.byref(c))
        c.log2_slice_mbs = self.log2_slice_mbs
        c.alpha = int(self.alpha)
        for name, m in (("luma", self.luma_matrix), ("chroma", self.chroma_matrix)):
            if m is not None:
                setattr(c, f"has_{name}_matrix", 1)
                arr = getattr(c, f"{name}_matrix")
                for k in range(64):
                    arr[k] = int(m[k])
        c.target_frame_bytes = self.target_frame_bytes or 0
        for i in range(4):
            c.encoder_identifier[i] = self.encoder_identifier[i]
        c.threads = self.threads
        return c


class Encoder:
    def __init__(self, config):
        self._config = config

    def config(self):
        return self._config

    def encode(self, frame):
        """Encodes one frame on the GPU -> bytes; raises ConfigError."""
        c = self._config._c()
        f = frame._c()
        out = C.c_void_p()
        n = C.c_size_t()
        err = _Err()
        lib().prores_encode(C.byref(c), C.byref(f), C.byref(out), C.byref(n), C.byref(err))
        _raise(err)
        try:
            return C.string_at(out, n.value)
        finally:
            lib().prores_buffer_free(out)


def encode_device_batch(items):
    """Encodes device-resident frames. ``items``: list of (Config, Frame
    metadata, dev_data ptr, dev_alpha ptr or 0). Returns a list of
    (dev_out ptr, length) or an Error per item."""
    n = len(items)
    jobs = (_DevEncodeJob * n)()
    cfgs = []
    for i, (cfg, frame, dd, da) in enumerate(items):
        c = cfg._c()
        cfgs.append(c)
        jobs[i].cfg = C.pointer(c)
        f = frame._c()
        f.data = dd
        if frame.alpha is not None or da:
            f.alpha = da
            f.alpha_len = frame.width * frame.height
        else:
            f.alpha = None
            f.alpha_len = 0
        jobs[i].frame = f
    errs = (_Err * n)()
    rc = lib().prores_encode_device_batch(n, jobs, errs)
    if rc == 4:
        _raise(errs[0])
    out = []
    for i in range(n):
        if errs[i].kind:
            out.append(_KINDS.get(errs[i].kind, Error)(errs[i].msg.decode()))
        else:
            out.append((jobs[i].dev_out, jobs[i].out_len))
    return out


# ---------------------------------------------------------------------------
# Serving loop (prores_server_*).


class ServerResult:
    """A completed request. Decode results hold device planes in Frame layout
    (and alpha); encode results the device packet. They stay on the device
    until Server.release(id)."""

    def __init__(self, r):
        self.id = r.id
        self.kind = "decode" if r.kind == 1 else "encode"
        self.error = None if r.error.kind == 0 else _KINDS.get(r.error.kind, Error)(r.error.msg.decode())
        self.dev_out, self.out_bytes = r.dev_out, r.out_bytes
        self.dev_alpha, self.alpha_bytes = r.dev_alpha, r.alpha_bytes
        self.width, self.height, self.chroma = r.header.width, r.header.height, r.header.chroma
        self.submit_us, self.start_us, self.done_us = r.submit_us, r.start_us, r.done_us
        self.batch_frames = r.batch_frames

    @property
    def latency_us(self):
        return self.done_us - self.submit_us

    def samples(self):
        return download_ptr(self.dev_out, self.out_bytes).view(np.uint16)

    def alpha(self):
        return None if not self.dev_alpha else download_ptr(self.dev_alpha, self.alpha_bytes).view(np.uint16)

    def packet(self):
        return download_ptr(self.dev_out, self.out_bytes).tobytes()


class Server:
    """The serving loop: submit decode / encode requests from any thread,
    poll results, release them (prores_server_* in include/prores_hip.h)."""

    def __init__(self, hbm_budget=64 << 30, max_batch_frames=64, decode_bit_depth=0):
        c = _ServerConfig(hbm_budget, max_batch_frames, decode_bit_depth)
        h, err = C.c_void_p(), _Err()
        lib().prores_server_create(C.byref(c), C.byref(h), C.byref(err))
        _raise(err)
        self._h = h

    def submit_decode(self, packet):
        packet = bytes(packet)
        return lib().prores_server_submit_decode(self._h, packet, len(packet))

    def submit_encode(self, config, frame):
        c, f = config._c(), frame._c()
        return lib().prores_server_submit_encode(self._h, C.byref(c), C.byref(f))

    def poll(self, max_results=256, wait_ms=0):
        arr = (_ServerResult * max_results)()
        n = lib().prores_server_poll(self._h, arr, max_results, wait_ms)
        return [ServerResult(arr[i]) for i in range(n)]

    def release(self, request_id):
        if lib().prores_server_release(self._h, request_id):
            raise ConfigError(f"request {request_id} is unknown or not polled")

    def stats(self):
        st = _ServerStats()
        lib().prores_server_get_stats(self._h, C.byref(st))
        return {n: getattr(st, n) for n, _ in _ServerStats._fields_}

    def close(self):
        if self._h:
            lib().prores_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def version():
    return lib().prores_version().decode()


def find_frames(data):
    """(offset, size) of every whole frame in a MOV file's bytes, found by
    'icpf' and frame_size (prores_find_frames; tests/sample.rs:26-41)."""
    data = bytes(data)
    n = lib().prores_find_frames(data, len(data), None, None, 0)
    offs, sizes = (C.c_uint64 * max(n, 1))(), (C.c_uint64 * max(n, 1))()
    lib().prores_find_frames(data, len(data), offs, sizes, n)
    return [(int(offs[i]), int(sizes[i])) for i in range(n)]


def fnv1a64(h, b):
    """FNV-1a 64 over bytes (or an array's bytes), from state h
    (prores_fnv1a; examples/prores_bench.rs:96-101)."""
    a = np.ascontiguousarray(np.frombuffer(b, dtype=np.uint8) if isinstance(b, (bytes, bytearray)) else b)
    return int(lib().prores_fnv1a(h, a.ctypes.data, a.nbytes))


def decode_file(path_or_bytes, decoder=None):
    """Decodes every frame of a MOV file on the GPU, many a call
    (prores_decode_host_batch): returns (frame count, FNV-1a of every
    picture's samples as little-endian u16), as the reference's bench reports
    for a file (examples/prores_bench.rs:104-122)."""
    data = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else open(path_or_bytes, "rb").read()
    dec = decoder or Decoder()
    frames = find_frames(data)
    h = 0xCBF29CE484222325
    for k in range(0, len(frames), 16):
        part = [data[o:o + n] for o, n in frames[k:k + 16]]
        for f in dec.decode_batch(part):
            if isinstance(f, Exception):
                raise f
            h = fnv1a64(h, f.data.astype("<u2", copy=False))
    return len(frames), h


# ---------------------------------------------------------------------------
# Device buffers (plumbing for device-resident batches).


class DeviceBuffer:
    """A device allocation on the current device."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = lib().prores_device_alloc(self.nbytes)
        if not self.ptr:
            raise DeviceError(f"could not allocate {nbytes} bytes of device memory")

    def upload(self, arr, offset=0):
        a = np.ascontiguousarray(arr)
        assert offset + a.nbytes <= self.nbytes
        if lib().prores_memcpy(self.ptr + offset, a.ctypes.data, a.nbytes, 1):
            raise DeviceError("host-to-device copy failed: " + lib().prores_memcpy_error().decode())

    def download(self, nbytes=None, offset=0, dtype=np.uint8):
        n = self.nbytes - offset if nbytes is None else nbytes
        out = np.empty(n // np.dtype(dtype).itemsize, dtype=dtype)
        if n and lib().prores_memcpy(out.ctypes.data, self.ptr + offset, n, 2):
            raise DeviceError("device-to-host copy failed: " + lib().prores_memcpy_error().decode())
        return out

    def free(self):
        if self.ptr and _lib is not None:
            _lib.prores_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        self.free()


def copy_device(dst_ptr, src_ptr, nbytes):
    if nbytes and lib().prores_memcpy(dst_ptr, src_ptr, nbytes, 3):
        raise DeviceError("device-to-device copy failed: " + lib().prores_memcpy_error().decode())


def download_ptr(src_ptr, nbytes):
    out = np.empty(nbytes, dtype=np.uint8)
    if nbytes and lib().prores_memcpy(out.ctypes.data, src_ptr, nbytes, 2):
        raise DeviceError("device-to-host copy failed: " + lib().prores_memcpy_error().decode())
    return out
