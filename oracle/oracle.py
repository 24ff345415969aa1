"""ctypes view of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product path never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


class OrRng(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4)]


class OrDec(C.Structure):
    _fields_ = [("rule", C.c_int), ("alpha", C.c_double), ("layered", C.c_int),
                ("max_iterations", C.c_int)]


SPA, NMS, OMS = 0, 1, 2


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    sig = {
        "or_splitmix64": (C.c_uint64, [C.POINTER(C.c_uint64)]),
        "or_fnv1a64": (C.c_uint64, [C.c_char_p, C.c_size_t]),
        "or_rng_new": (None, [C.POINTER(OrRng), C.c_uint64]),
        "or_rng_next": (C.c_uint64, [C.POINTER(OrRng)]),
        "or_rng_uniform": (C.c_double, [C.POINTER(OrRng)]),
        "or_rng_bit": (C.c_uint8, [C.POINTER(OrRng)]),
        "or_rng_gaussian": (C.c_double, [C.POINTER(OrRng)]),
        "or_rng_complex_gaussian": (None, [C.POINTER(OrRng), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "or_stream": (None, [C.POINTER(OrRng), C.c_uint64, C.c_char_p, C.c_uint64, C.c_uint64]),
        "or_code_from_rows": (C.c_void_p, [C.c_int32, C.c_int32, _i32p, _i32p]),
        "or_code_lifted": (C.c_void_p, [C.c_int, C.c_int32]),
        "or_nr_create": (C.c_void_p, [C.c_int32, C.c_int32, C.c_int, C.c_int, C.c_int32]),
        "or_code_free": (None, [C.c_void_p]),
        "or_last_error": (C.c_char_p, []),
        "or_code_info": (C.c_int64, [C.c_void_p, C.c_int]),
        "or_code_row_start": (C.POINTER(C.c_int32), [C.c_void_p]),
        "or_code_cols": (C.POINTER(C.c_int32), [C.c_void_p]),
        "or_code_bit_map": (C.POINTER(C.c_int32), [C.c_void_p]),
        "or_encode": (C.c_int, [C.c_void_p, _u8p, _u8p]),
        "or_nr_encode": (C.c_int, [C.c_void_p, _u8p, _u8p]),
        "or_nr_full_llr": (C.c_int, [C.c_void_p, _f64p, _f64p]),
        "or_is_codeword": (C.c_int, [C.c_void_p, _u8p]),
        "or_decode_llr": (C.c_int, [C.c_void_p, C.POINTER(OrDec), _f64p, _f64p]),
        "or_nr_decode_block": (C.c_int, [C.c_void_p, C.POINTER(OrDec), _f64p, _u8p]),
        "or_decode_many": (None, [C.c_void_p, C.POINTER(OrDec), _f64p, C.c_int64, _f64p, _i32p, C.c_int]),
        "or_trace_f32_f64": (C.c_int, [C.c_void_p, C.POINTER(OrDec), _f64p, _f64p]),
        "or_decode_llr_f32": (C.c_int, [C.c_void_p, C.POINTER(OrDec), _f32p, _f32p]),
        "or_decode_many_f32": (None, [C.c_void_p, C.POINTER(OrDec), _f32p, C.c_int64, _f32p, _i32p, C.c_int]),
        "or_f32_to_f16": (C.c_uint16, [C.c_float]),
        "or_f16_to_f32": (C.c_float, [C.c_uint16]),
        "or_decode_llr_f16": (C.c_int, [C.c_void_p, C.POINTER(OrDec), _f32p, _f32p]),
        "or_decode_many_f16": (None, [C.c_void_p, C.POINTER(OrDec), _f32p, C.c_int64, C.c_void_p, _f32p, _i32p,
                                      C.c_int]),
        "or_db2lin": (C.c_double, [C.c_double]),
        "or_bpsk_noise_variance": (C.c_double, [C.c_double, C.c_double]),
        "or_ebn0_noise_variance": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double]),
        "or_bpsk_awgn_llr": (None, [_u8p, C.c_int64, C.c_double, C.POINTER(OrRng), _f64p]),
        "or_qam_points": (None, [C.c_int, _f64p, _f64p]),
        "or_qam_llr": (None, [C.c_int, _f64p, _f64p, C.c_int64, C.c_double, _f64p]),
        "or_qam_awgn_llr": (None, [C.c_int, _u8p, C.c_int64, C.c_double, C.POINTER(OrRng), _f64p]),
        "or_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "or_philox_info_bits": (None, [C.c_uint64, C.c_uint32, C.c_uint64, C.c_int32, _u8p]),
        "or_philox_normals": (None, [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
        "or_philox_trial": (None, [C.c_void_p, C.c_int, C.c_double, C.c_uint64, C.c_uint32, C.c_uint64, _u8p,
                                   _f64p, _f64p, _f64p]),
        "or_label_halves_llr": (C.c_double, [_f64p, C.c_int, C.c_int]),
        "or_llr_from_metrics": (None, [_f64p, C.c_int, _f64p]),
        "or_trial_nr_qam_many": (None, [C.c_void_p, C.c_int, C.c_double, C.c_uint64, C.c_char_p, C.c_uint64,
                                        C.c_uint64, C.c_int64, _u8p, _f64p, C.c_int]),
        "or_erfc": (C.c_double, [C.c_double]),
        "or_erfcinv": (C.c_double, [C.c_double]),
        "or_qfuncinv": (C.c_double, [C.c_double]),
        "or_z_score": (C.c_double, [C.c_double]),
        "or_wilson": (None, [C.c_uint64, C.c_uint64, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "or_trial_bpsk_lifted": (None, [C.c_void_p, C.c_double, C.c_uint64, C.c_char_p, C.c_uint64,
                                        C.c_uint64, _u8p, _f64p]),
        "or_trial_nr_qam": (None, [C.c_void_p, C.c_int, C.c_double, C.c_uint64, C.c_char_p, C.c_uint64,
                                   C.c_uint64, _u8p, _f64p]),
        "or_learned_decode": (C.c_int, [C.c_void_p, C.c_int, _f64p, _f64p, _f64p, _f64p]),
        "or_learned_loss_grad": (C.c_double, [C.c_void_p, C.c_int, _f64p, _f64p, _f64p, _u8p, _f64p, _f64p]),
    }
    for name, (res, args) in sig.items():
        if not hasattr(lib, name):
            continue
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


class Rng:
    """xoshiro256** of crates/com-core/src/rng.rs."""

    def __init__(self, seed=0, state=None):
        self.r = OrRng()
        if state is None:
            lib.or_rng_new(C.byref(self.r), seed)

    @staticmethod
    def stream(base, name, cond, trial):
        r = Rng(state=True)
        lib.or_stream(C.byref(r.r), base, name.encode(), cond, trial)
        return r

    def next_u64(self):
        return lib.or_rng_next(C.byref(self.r))

    def uniform(self):
        return lib.or_rng_uniform(C.byref(self.r))

    def bit(self):
        return lib.or_rng_bit(C.byref(self.r))

    def gaussian(self):
        return lib.or_rng_gaussian(C.byref(self.r))


def fnv1a64(b):
    return lib.or_fnv1a64(b, len(b))


class Code:
    """An LDPC code of the oracle: NR (NrLdpc::with_options), lifted base graph, or rows."""

    def __init__(self, handle):
        if not handle:
            raise ValueError(lib.or_last_error().decode())
        self.h = handle
        self.n, self.m, self.k, self.k_info, self.e, self.bg, self.z, self.ils, self.edges = (
            int(lib.or_code_info(handle, i)) for i in range(9))

    def __del__(self):
        if getattr(self, "h", None):
            lib.or_code_free(self.h)
            self.h = None

    @classmethod
    def nr(cls, k_info, e, bg=0, rv=0, qm=1):
        return cls(lib.or_nr_create(k_info, e, bg, rv, qm))

    @classmethod
    def lifted(cls, bg, z):
        return cls(lib.or_code_lifted(bg, z))

    @classmethod
    def from_rows(cls, n, rows):
        start = np.zeros(len(rows) + 1, np.int32)
        start[1:] = np.cumsum([len(r) for r in rows])
        cols = np.array([v for r in rows for v in r], np.int32)
        return cls(lib.or_code_from_rows(n, len(rows), start, cols))

    def rows_csr(self):
        rs = np.ctypeslib.as_array(lib.or_code_row_start(self.h), (self.m + 1,)).copy()
        cols = np.ctypeslib.as_array(lib.or_code_cols(self.h), (self.edges,)).copy()
        return rs, cols

    def bit_map(self):
        return np.ctypeslib.as_array(lib.or_code_bit_map(self.h), (self.e,)).copy()

    def encode(self, u):
        cw = np.zeros(self.n, np.uint8)
        if lib.or_encode(self.h, np.ascontiguousarray(u, np.uint8), cw):
            raise ValueError(lib.or_last_error().decode())
        return cw

    def nr_encode(self, info):
        out = np.zeros(self.e, np.uint8)
        lib.or_nr_encode(self.h, np.ascontiguousarray(info, np.uint8), out)
        return out

    def full_llr(self, llr):
        full = np.zeros(self.n)
        lib.or_nr_full_llr(self.h, np.ascontiguousarray(llr, np.float64), full)
        return full

    def is_codeword(self, cw):
        return bool(lib.or_is_codeword(self.h, np.ascontiguousarray(cw, np.uint8)))

    def decode_llr(self, llr, rule=NMS, alpha=0.75, layered=True, max_iterations=20):
        d = OrDec(rule, alpha, int(layered), max_iterations)
        post = np.zeros(self.n)
        it = lib.or_decode_llr(self.h, C.byref(d), np.ascontiguousarray(llr, np.float64), post)
        return post, it

    def decode_block(self, llr, rule=NMS, alpha=0.75, layered=True, max_iterations=20):
        d = OrDec(rule, alpha, int(layered), max_iterations)
        bits = np.zeros(self.k_info, np.uint8)
        it = lib.or_nr_decode_block(self.h, C.byref(d), np.ascontiguousarray(llr, np.float64), bits)
        return bits, it

    def decode_many(self, llr, rule=NMS, alpha=0.75, layered=True, max_iterations=20, threads=1):
        llr = np.ascontiguousarray(llr, np.float64).reshape(-1, self.n)
        d = OrDec(rule, alpha, int(layered), max_iterations)
        post = np.zeros_like(llr)
        its = np.zeros(len(llr), np.int32)
        lib.or_decode_many(self.h, C.byref(d), llr, len(llr), post, its, threads)
        return post, its

    def decode_many_f32(self, llr, rule=NMS, alpha=0.75, layered=True, max_iterations=20, threads=1):
        """The decoder with FP32 messages (the same operations rounded to FP32)."""
        llr = np.ascontiguousarray(llr, np.float32).reshape(-1, self.n)
        d = OrDec(rule, alpha, int(layered), max_iterations)
        post = np.zeros_like(llr)
        its = np.zeros(len(llr), np.int32)
        lib.or_decode_many_f32(self.h, C.byref(d), llr, len(llr), post, its, threads)
        return post, its

    def trace_f32_f64(self, llr, rule=NMS, alpha=0.75, max_iterations=10):
        """The layered min-sum in f64 and FP32 side by side on the same LLRs (values exact in
        FP32): a dict of where the two runs part (or_trace_f32_f64)."""
        d = OrDec(rule, alpha, 1, max_iterations)
        out = np.zeros(10 + max_iterations)
        if lib.or_trace_f32_f64(self.h, C.byref(d), np.ascontiguousarray(llr, np.float64), out):
            raise ValueError("the side-by-side run takes the min-sum rules")
        keys = ("it64", "it32", "min_abs", "first_sign_abs", "first_sign_it", "sign_events", "argmin_events",
                "min_argmin_gap", "first_argmin_it", "decisions_differ")
        r = {k: float(out[i]) for i, k in enumerate(keys)}
        r["rel"] = out[10:].copy()
        return r

    def decode_many_f16(self, llr, rule=NMS, alpha=0.75, max_iterations=20, threads=1):
        """The layered min-sum with the messages stored in binary16 and FP32 arithmetic.
        max_iterations: one limit, or one per codeword. Returns the posteriors (the stored
        halves widened to float32) and the iterations."""
        llr = np.ascontiguousarray(llr, np.float32).reshape(-1, self.n)
        per = None
        lim = max_iterations
        if np.ndim(max_iterations):
            per = np.ascontiguousarray(max_iterations, np.int32)
            assert per.shape == (len(llr),)
            lim = 0
        d = OrDec(rule, alpha, 1, int(lim))
        post = np.zeros_like(llr)
        its = np.zeros(len(llr), np.int32)
        lib.or_decode_many_f16(self.h, C.byref(d), llr, len(llr),
                               None if per is None else per.ctypes.data_as(C.c_void_p), post, its, threads)
        return post, its

    def trial_bpsk_lifted(self, ebn0_db, seed, name, cond, trial):
        cw = np.zeros(self.n, np.uint8)
        llr = np.zeros(self.n)
        lib.or_trial_bpsk_lifted(self.h, ebn0_db, seed, name.encode(), cond, trial, cw, llr)
        return cw, llr

    def trial_nr_qam(self, m, noise_var, seed, name, cond, trial):
        info = np.zeros(self.k_info, np.uint8)
        llr = np.zeros(self.e)
        lib.or_trial_nr_qam(self.h, m, noise_var, seed, name.encode(), cond, trial, info, llr)
        return info, llr

    def trial_nr_qam_many(self, m, noise_var, seed, name, cond, trial0, count, threads=1):
        """trial_nr_qam for trials trial0 .. trial0 + count - 1: (info (count, K'), llr (count, E))."""
        info = np.zeros((count, self.k_info), np.uint8)
        llr = np.zeros((count, self.e))
        lib.or_trial_nr_qam_many(self.h, m, noise_var, seed, name.encode(), cond, trial0, count, info, llr, threads)
        return info, llr

    def philox_trial(self, bps, noise_var, seed, point, trial):
        """One trial of the device trial generator restated in f64: (info, y, llr); y is
        complex (BPSK: the real received values)."""
        info = np.zeros(self.k_info, np.uint8)
        nsym = self.e // bps
        yre, yim, llr = np.zeros(nsym), np.zeros(nsym), np.zeros(self.e)
        lib.or_philox_trial(self.h, bps, noise_var, seed, point, trial, info, yre, yim, llr)
        return info, yre + 1j * yim, llr

    def learned_decode(self, llr, alpha, beta):
        alpha = np.ascontiguousarray(alpha, np.float64)
        beta = np.ascontiguousarray(beta, np.float64)
        post = np.zeros(self.n)
        it = lib.or_learned_decode(self.h, alpha.shape[0], alpha.ravel(), beta.ravel(),
                                   np.ascontiguousarray(llr, np.float64), post)
        return post, it

    def learned_loss_grad(self, llr, cw, alpha, beta):
        alpha = np.ascontiguousarray(alpha, np.float64)
        beta = np.ascontiguousarray(beta, np.float64)
        ga = np.zeros_like(alpha)
        gb = np.zeros_like(beta)
        loss = lib.or_learned_loss_grad(self.h, alpha.shape[0], alpha.ravel(), beta.ravel(),
                                        np.ascontiguousarray(llr, np.float64),
                                        np.ascontiguousarray(cw, np.uint8), ga.ravel(), gb.ravel())
        return loss, ga, gb


def f32_to_f16_bits(x):
    """binary32 -> binary16 bits, round to nearest even (the C restatement)."""
    return np.array([lib.or_f32_to_f16(float(v)) for v in np.asarray(x, np.float32).ravel()], np.uint16)


def f16_bits_to_f32(h):
    return np.array([lib.or_f16_to_f32(int(v)) for v in np.asarray(h, np.uint16).ravel()], np.float32)


def wilson(errors, total, confidence=0.95):
    lo, hi = C.c_double(), C.c_double()
    lib.or_wilson(errors, total, confidence, C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.or_philox4x32_10(c, k, o)
    return list(o)


def philox_info_bits(seed, point, trial, k_info):
    bits = np.zeros(k_info, np.uint8)
    lib.or_philox_info_bits(seed, point, trial, k_info, bits)
    return bits


def philox_normals(seed, point, trial, index):
    g0, g1 = C.c_double(), C.c_double()
    lib.or_philox_normals(seed, point, trial, index, C.byref(g0), C.byref(g1))
    return g0.value, g1.value


def label_halves_llr(metric, shift):
    metric = np.ascontiguousarray(metric, np.float64)
    return lib.or_label_halves_llr(metric, metric.size, shift)


def llr_from_metrics(metric):
    metric = np.ascontiguousarray(metric, np.float64)
    out = np.zeros(int(np.log2(metric.size)))
    lib.or_llr_from_metrics(metric, metric.size, out)
    return out


def qam_llr(m, y, noise_var):
    y = np.asarray(y, np.complex128).ravel()
    b = int(np.log2(m))
    out = np.zeros(y.size * b)
    lib.or_qam_llr(m, np.ascontiguousarray(y.real), np.ascontiguousarray(y.imag), y.size, noise_var, out)
    return out


def complex_gaussians(seed, count):
    """count draws of Rng::complex_gaussian (rng.rs:161-167) from Rng::new(seed)."""
    r = Rng(seed)
    re, im = C.c_double(), C.c_double()
    out = np.zeros(count, np.complex128)
    for i in range(count):
        lib.or_rng_complex_gaussian(C.byref(r.r), C.byref(re), C.byref(im))
        out[i] = re.value + 1j * im.value
    return out


def qam_points(m):
    re, im = np.zeros(m), np.zeros(m)
    lib.or_qam_points(m, re, im)
    return re + 1j * im
