"""Python view of libcomfec_gpu.so, shaped like the reference's NrLdpc / LdpcDecoder API.

It mirrors crates/com-fec/src/nr/ldpc.rs (NrLdpc::with_options, encode_block,
decode_block, app) and crates/com-fec/src/ldpc.rs (LdpcDecoder, CheckRule::from_name),
and adds the batched device entry point. Errors raise ComfecError with the kind and the
message of the library (the Display of com_core::Error). The library is required: there
is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libcomfec_gpu.so")

CF_OK, CF_CONFIG, CF_SHAPE, CF_NUMERICAL, CF_CONDITION, CF_IO, CF_DEVICE = range(7)
KINDS = {1: "Config", 2: "Shape", 3: "Numerical", 4: "Condition", 5: "Io", 6: "Device"}
SUM_PRODUCT, NORMALIZED_MIN_SUM, OFFSET_MIN_SUM = 0, 1, 2
MSG_F64, MSG_F32, MSG_F16 = 0, 1, 2


class ComfecError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(message)
        self.kind = KINDS.get(status, str(status))
        self.status = status


class Dims(C.Structure):
    _fields_ = [(name, C.c_uint32) for name in
                ("k_info", "e", "n", "k", "m", "edges", "base_graph", "z", "ils", "qm", "rv")]


_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")

# (name, restype, argtypes) of every entry point of include/comfec_gpu.h
SIGNATURES = [
    ("cf_last_error", C.c_char_p, []),
    ("cf_nr_create", C.c_int, [C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("cf_lifted_create", C.c_int, [C.c_int, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("cf_code_destroy", None, [C.c_void_p]),
    ("cf_code_dims_get", C.c_int, [C.c_void_p, C.POINTER(Dims)]),
    ("cf_bit_map", C.c_int, [C.c_void_p, _u32p, C.c_size_t]),
    ("cf_set_decoder", C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_uint32]),
    ("cf_encode_block", C.c_int, [C.c_void_p, _u8p, C.c_size_t, _u8p, C.c_size_t]),
    ("cf_decode_llr", C.c_int, [C.c_void_p, _f64p, C.c_size_t, _f64p, C.POINTER(C.c_uint32)]),
    ("cf_decode_block", C.c_int, [C.c_void_p, _f64p, C.c_size_t, _u8p, C.c_size_t]),
    ("cf_app_into", C.c_int, [C.c_void_p, _f64p, C.c_size_t, C.c_void_p, _f64p, C.c_size_t,
                              C.POINTER(C.c_uint32)]),
    ("cf_decode_batch", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p]),
]

_lib = None


def lib():
    """The loaded library; raises if it was not built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python com-rs_amd/build.py`")
        l = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _check(status):
    if status != CF_OK:
        raise ComfecError(status, lib().cf_last_error().decode())


def rule_from_name(rule, alpha):
    """CheckRule::from_name (crates/com-fec/src/ldpc.rs:113-123): (rule id, alpha)."""
    r = rule.lower()
    if r == "sum_product":
        return SUM_PRODUCT, alpha
    if r == "min_sum":
        return NORMALIZED_MIN_SUM, 1.0
    if r == "normalized_min_sum":
        return NORMALIZED_MIN_SUM, alpha
    if r == "offset_min_sum":
        return OFFSET_MIN_SUM, alpha
    raise ComfecError(CF_CONFIG, "invalid configuration: rule must be sum_product, min_sum, "
                                 f"normalized_min_sum or offset_min_sum, got {rule!r}")


def _stream_ptr(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return C.c_void_p(stream)


class NrLdpc:
    """NrLdpc of crates/com-fec/src/nr/ldpc.rs on the GPU."""

    def __init__(self, k_info, e, base_graph=None, rv=0, qm=1, *, _handle=None):
        h = C.c_void_p()
        if _handle is None:
            _check(lib().cf_nr_create(k_info, e, base_graph or 0, rv, qm, C.byref(h)))
        else:
            h = _handle
        self._h = h
        d = Dims()
        _check(lib().cf_code_dims_get(self._h, C.byref(d)))
        self.dims = d
        self._settings = (SUM_PRODUCT, 0.8, True, 50)

    @classmethod
    def lifted(cls, base_graph, z):
        """from_core_rows on the lifted base graph (ldpc.rs:270-287 on nr/ldpc.rs:124-133)."""
        h = C.c_void_p()
        _check(lib().cf_lifted_create(base_graph, z, C.byref(h)))
        return cls(0, 0, _handle=h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.cf_code_destroy(h)
            self._h = None

    # BlockCode / NrLdpc accessors
    @property
    def k(self):
        return self.dims.k_info

    @property
    def n(self):
        return self.dims.e

    def base_graph(self):
        return self.dims.base_graph

    def lifting_size(self):
        return self.dims.z

    def bit_map(self):
        out = np.zeros(self.dims.e, np.uint32)
        _check(lib().cf_bit_map(self._h, out, out.size))
        return out

    def set_decoder(self, rule="normalized_min_sum", alpha=0.75, layered=True, max_iterations=50):
        rid, a = rule_from_name(rule, alpha) if isinstance(rule, str) else (rule, alpha)
        _check(lib().cf_set_decoder(self._h, rid, a, int(bool(layered)), max_iterations))
        self._settings = (rid, a, bool(layered), max_iterations)
        return self

    def encode_block(self, info):
        info = np.ascontiguousarray(info, np.uint8)
        out = np.zeros(self.dims.e, np.uint8)
        _check(lib().cf_encode_block(self._h, info, info.size, out, out.size))
        return out

    def decode_llr(self, llr):
        """LdpcCode::decode_llr on the N LLRs of the lifted code: (post, iterations)."""
        llr = np.ascontiguousarray(llr, np.float64)
        post = np.zeros(self.dims.n)
        it = C.c_uint32()
        _check(lib().cf_decode_llr(self._h, llr, llr.size, post, C.byref(it)))
        return post, it.value

    def decode_block(self, llr):
        llr = np.ascontiguousarray(llr, np.float64)
        bits = np.zeros(self.dims.k_info, np.uint8)
        _check(lib().cf_decode_block(self._h, llr, llr.size, bits, bits.size))
        return bits

    def app(self, llr, apriori=None):
        llr = np.ascontiguousarray(llr, np.float64)
        out = np.zeros(self.dims.k_info)
        it = C.c_uint32()
        ap = None
        if apriori is not None:
            ap = np.ascontiguousarray(apriori, np.float64)
            if ap.size != self.dims.k_info:
                raise ComfecError(CF_SHAPE, f"inconsistent shapes: {ap.size} a priori LLRs for a block of "
                                            f"{self.dims.k_info}")
        _check(lib().cf_app_into(self._h, llr, llr.size, ap.ctypes.data if ap is not None else None, out,
                                 out.size, C.byref(it)))
        return out, it.value

    def decode_batch(self, llr, *, want_post=False, want_bits=True, stream=None):
        """Decode a (count, E) torch tensor on the GPU (float32 or float64 messages, the type of
        llr): returns (bits (count, K') uint8, iterations (count,) int32, post (count, N) or None)."""
        import torch
        if llr.dim() != 2 or llr.shape[1] != self.dims.e:
            raise ComfecError(CF_SHAPE, f"inconsistent shapes: {llr.shape[-1]} LLRs for a block of {self.dims.e}")
        if not llr.is_cuda or not llr.is_contiguous():
            raise ComfecError(CF_CONFIG, "invalid configuration: the LLRs must be a contiguous device tensor")
        msg = {torch.float64: MSG_F64, torch.float32: MSG_F32}[llr.dtype]
        count = llr.shape[0]
        dev = llr.device
        bits = torch.empty((count, self.dims.k_info), dtype=torch.uint8, device=dev) if want_bits else None
        its = torch.empty(count, dtype=torch.int32, device=dev)
        post = torch.empty((count, self.dims.n), dtype=llr.dtype, device=dev) if want_post else None
        with torch.cuda.device(dev):
            _check(lib().cf_decode_batch(self._h, msg, C.c_void_p(llr.data_ptr()), count,
                                         C.c_void_p(bits.data_ptr()) if bits is not None else None,
                                         C.c_void_p(post.data_ptr()) if post is not None else None,
                                         C.c_void_p(its.data_ptr()), _stream_ptr(stream)))
        return bits, its, post
