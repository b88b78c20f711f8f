"""ctypes bindings for the CPU oracle (oracle/liboracle.so).

Test infrastructure: the oracle is the checker for the HIP path, never the
thing measured or shipped.
"""

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

KIND_NAMES = {1: "Invalid", 2: "Unsupported", 3: "Config"}
PROFILES = ["Proxy", "Lt", "Standard", "Hq", "P4444", "P4444Xq"]


class OFrame(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("bit_depth", C.c_uint32),
        ("chroma", C.c_uint32), ("interlace", C.c_uint32),
        ("aspect_ratio", C.c_uint8), ("frame_rate_code", C.c_uint8),
        ("color_primaries", C.c_uint8), ("transfer_characteristic", C.c_uint8),
        ("matrix_coefficients", C.c_uint8), ("has_alpha", C.c_uint8),
        ("data", C.POINTER(C.c_uint16)), ("data_len", C.c_size_t),
        ("alpha", C.POINTER(C.c_uint16)), ("alpha_len", C.c_size_t),
        ("plane_offset", C.c_uint64 * 3), ("plane_width", C.c_uint32 * 3),
        ("plane_height", C.c_uint32 * 3),
    ]


class OConfig(C.Structure):
    _fields_ = [
        ("profile", C.c_uint32), ("log2_slice_mbs", C.c_uint32), ("alpha", C.c_uint32),
        ("has_luma_matrix", C.c_uint8), ("has_chroma_matrix", C.c_uint8),
        ("luma_matrix", C.c_uint8 * 64), ("chroma_matrix", C.c_uint8 * 64),
        ("target_frame_bytes", C.c_uint64), ("encoder_identifier", C.c_uint8 * 4),
        ("threads", C.c_uint32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.oracle_decode.argtypes = [C.c_char_p, C.c_size_t, C.c_uint32, C.c_uint32, P(OFrame), C.c_char_p, C.c_size_t]
        L.oracle_frame_free.argtypes = [P(OFrame)]
        L.oracle_encode.argtypes = [P(OConfig), P(OFrame), P(P(C.c_uint8)), P(C.c_size_t), C.c_char_p, C.c_size_t]
        L.oracle_free.argtypes = [C.c_void_p]
        L.oracle_config_default.argtypes = [C.c_uint32, P(OConfig)]
        L.oracle_target_frame_bytes.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.oracle_target_frame_bytes.restype = C.c_uint64
        L.oracle_idct_put.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
        L.oracle_idct.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_fdct.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_fdct_load.argtypes = [C.c_void_p, C.c_float, C.c_void_p]
        L.oracle_quantise.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_float, C.c_void_p, C.c_void_p]
        L.oracle_basis.argtypes = [C.c_void_p]
        L.oracle_slice_sizes.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_size_t]
        L.oracle_slice_sizes.restype = C.c_size_t
        L.oracle_qscale.argtypes = [C.c_uint32]
        L.oracle_qscale.restype = C.c_uint32
        L.oracle_scan.argtypes = [C.c_uint32, C.c_void_p]
        L.oracle_codeword.argtypes = [C.c_uint32] * 5 + [C.c_char_p, C.c_size_t]
        L.oracle_codeword.restype = C.c_uint32
        L.oracle_code_len.argtypes = [C.c_uint32] * 5
        L.oracle_code_len.restype = C.c_uint32
        L.oracle_code_read.argtypes = [C.c_uint32] * 4 + [C.c_char_p, C.c_size_t, P(C.c_uint32), C.c_char_p, C.c_size_t]
        L.oracle_signed_to_symbol.argtypes = [C.c_int32]
        L.oracle_signed_to_symbol.restype = C.c_uint32
        L.oracle_symbol_to_signed.argtypes = [C.c_uint32]
        L.oracle_symbol_to_signed.restype = C.c_int32
        L.oracle_alpha_run_bits.argtypes = [C.c_uint32, C.c_char_p, C.c_size_t]
        L.oracle_alpha_run_bits.restype = C.c_uint32
        L.oracle_alpha_diff_bits.argtypes = [C.c_int32, C.c_uint32, C.c_char_p, C.c_size_t]
        L.oracle_alpha_diff_bits.restype = C.c_uint32
        L.oracle_rescale.argtypes = [C.c_uint32, C.c_uint32, C.c_uint16]
        L.oracle_rescale.restype = C.c_uint16
        L.oracle_rescale_many.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_size_t, C.c_void_p]
        L.oracle_encode_coefficients.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
        L.oracle_encode_coefficients.restype = C.c_size_t
        L.oracle_coefficient_bits.argtypes = [C.c_void_p, C.c_size_t]
        L.oracle_coefficient_bits.restype = C.c_size_t
        L.oracle_decode_coefficients.argtypes = [C.c_char_p, C.c_size_t, C.c_size_t, C.c_void_p, C.c_char_p, C.c_size_t]
        L.oracle_encode_alpha.argtypes = [C.c_void_p, C.c_size_t, C.c_uint32, C.c_void_p, C.c_size_t]
        L.oracle_encode_alpha.restype = C.c_size_t
        L.oracle_decode_alpha.argtypes = [C.c_char_p, C.c_size_t, C.c_uint32, C.c_size_t, C.c_size_t, C.c_void_p, P(C.c_size_t), C.c_char_p, C.c_size_t]
        L.oracle_test_frame.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32, C.c_void_p]
        L.oracle_resolve_threads.argtypes = [C.c_uint32]
        L.oracle_resolve_threads.restype = C.c_uint32
        _lib = L
    return _lib


class OracleError(Exception):
    def __init__(self, kind, msg):
        super().__init__(f"{KIND_NAMES.get(kind, kind)}: {msg}")
        self.kind = KIND_NAMES.get(kind, str(kind))
        self.msg = msg

    def as_tuple(self):
        return (self.kind, self.msg)


def layout(width, height, chroma):
    """Frame::layout, src/frame.rs:247-266: (offset, width, height) per plane."""
    cw = (width + 1) // 2 if chroma == 2 else width
    y = (0, width, height)
    cb = (width * height, cw, height)
    cr = (cb[0] + cw * height, cw, height)
    return [y, cb, cr]


class Frame:
    """A frame in the reference's layout: one u16 buffer Y'||Cb||Cr plus alpha."""

    def __init__(self, width, height, chroma, bit_depth, data=None, alpha=None, interlace=0,
                 metadata=(0, 0, 2, 2, 2)):
        self.width, self.height, self.chroma, self.bit_depth = width, height, chroma, bit_depth
        self.interlace = interlace
        self.metadata = tuple(metadata)
        self.planes = layout(width, height, chroma)
        total = self.planes[2][0] + self.planes[2][1] * self.planes[2][2]
        if data is None:
            data = np.full(total, 1 << (bit_depth - 1), dtype=np.uint16)
        self.data = np.ascontiguousarray(data, dtype=np.uint16)
        self.alpha = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.uint16)

    def plane(self, i):
        off, w, h = self.planes[i]
        return self.data[off:off + w * h].reshape(h, w)

    def to_c(self):
        f = OFrame()
        f.width, f.height, f.bit_depth = self.width, self.height, self.bit_depth
        f.chroma, f.interlace = self.chroma, self.interlace
        (f.aspect_ratio, f.frame_rate_code, f.color_primaries, f.transfer_characteristic,
         f.matrix_coefficients) = self.metadata
        f.data = self.data.ctypes.data_as(C.POINTER(C.c_uint16))
        f.data_len = self.data.size
        f.has_alpha = 1 if self.alpha is not None else 0
        if self.alpha is not None:
            f.alpha = self.alpha.ctypes.data_as(C.POINTER(C.c_uint16))
            f.alpha_len = self.alpha.size
        for i, (off, w, h) in enumerate(self.planes):
            f.plane_offset[i], f.plane_width[i], f.plane_height[i] = off, w, h
        return f


def test_frame(width, height, chroma, bit_depth, grain, seed):
    """tests/common/mod.rs:26-80."""
    f = Frame(width, height, chroma, bit_depth)
    lib().oracle_test_frame(width, height, chroma, bit_depth, grain, seed, f.data.ctypes.data)
    return f


def config(profile, log2_slice_mbs=3, alpha=2, luma=None, chroma=None, target=0,
           encoder_identifier=b"rivt", threads=0):
    c = OConfig()
    lib().oracle_config_default(profile, C.byref(c))
    c.log2_slice_mbs, c.alpha, c.target_frame_bytes, c.threads = log2_slice_mbs, alpha, target, threads
    if luma is not None:
        c.has_luma_matrix = 1
        for k in range(64):
            c.luma_matrix[k] = int(luma[k])
    if chroma is not None:
        c.has_chroma_matrix = 1
        for k in range(64):
            c.chroma_matrix[k] = int(chroma[k])
    for i in range(4):
        c.encoder_identifier[i] = encoder_identifier[i]
    return c


def encode(frame, cfg):
    """Encoder::encode, src/encode.rs:223-287 -> bytes, or raises OracleError."""
    L = lib()
    out = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    msg = C.create_string_buffer(512)
    cf = frame.to_c()
    rc = L.oracle_encode(C.byref(cfg), C.byref(cf), C.byref(out), C.byref(n), msg, 512)
    if rc:
        raise OracleError(rc, msg.value.decode())
    b = C.string_at(out, n.value)
    L.oracle_free(out)
    return b


def decode(data, bit_depth=0, threads=0):
    """Decoder::decode, src/decode.rs:61-104 -> Frame, or raises OracleError."""
    L = lib()
    f = OFrame()
    msg = C.create_string_buffer(512)
    rc = L.oracle_decode(bytes(data), len(data), bit_depth, threads, C.byref(f), msg, 512)
    if rc:
        raise OracleError(rc, msg.value.decode())
    data_arr = np.ctypeslib.as_array(f.data, shape=(f.data_len,)).copy()
    alpha = None
    if f.has_alpha:
        alpha = np.ctypeslib.as_array(f.alpha, shape=(f.alpha_len,)).copy()
    out = Frame(f.width, f.height, f.chroma, f.bit_depth, data_arr, alpha, f.interlace,
                (f.aspect_ratio, f.frame_rate_code, f.color_primaries, f.transfer_characteristic,
                 f.matrix_coefficients))
    L.oracle_frame_free(C.byref(f))
    return out


def decode_result(data, bit_depth=0, threads=0):
    """('Ok', frame) or (kind, message) — the Result the reference compares."""
    try:
        return ("Ok", decode(data, bit_depth, threads))
    except OracleError as e:
        return e.as_tuple()


def psnr_samples(a, b, bit_depth):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    mse = np.mean((a - b) ** 2)
    peak = float((1 << bit_depth) - 1)
    return float("inf") if mse == 0 else 10.0 * np.log10(peak * peak / mse)


def codeword(kind, a, b, c, n):
    buf = C.create_string_buffer(256)
    L = lib().oracle_codeword(kind, a, b, c, n, buf, 256)
    return buf.value.decode()[:L]


def target_frame_bytes(profile, w, h):
    return lib().oracle_target_frame_bytes(profile, w, h)
