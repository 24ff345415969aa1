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
    ("cf_code_from_rows", C.c_int, [C.c_uint32, C.c_uint32, _u32p, _u32p, C.POINTER(C.c_void_p)]),
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
    ("cf_decode_batch_static", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]),
]

class StopNode(C.Structure):
    _fields_ = [("kind", C.c_int), ("n_children", C.c_uint32), ("count", C.c_uint64), ("value", C.c_double),
                ("confidence", C.c_double), ("metric", C.c_char_p)]


class GridAxis(C.Structure):
    _fields_ = [("name", C.c_char_p), ("values", C.POINTER(C.c_double)), ("n_values", C.c_size_t)]


class RunnerConfig(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("batch", C.c_uint64), ("confidence", C.c_double), ("fast_dev_run", C.c_int),
                ("threads", C.c_uint64), ("config_json", C.c_char_p), ("hooks", C.c_void_p)]


class RunInfoC(C.Structure):
    _fields_ = [("scenario", C.c_char_p), ("seed", C.c_uint64), ("batch", C.c_uint64), ("threads", C.c_uint64),
                ("stop", C.c_char_p), ("confidence", C.c_double), ("n_conditions", C.c_uint64),
                ("started_unix_s", C.c_uint64), ("version", C.c_char_p), ("commit", C.c_char_p),
                ("config", C.c_char_p)]


class ProgressC(C.Structure):
    _fields_ = [("condition", C.c_uint64), ("trials", C.c_uint64), ("elapsed_s", C.c_double),
                ("record", C.c_void_p)]


ON_RUN_START = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(RunInfoC))
ON_CONDITION_START = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_size_t, C.POINTER(C.c_double), C.c_size_t)
ON_BATCH_END = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(ProgressC))
ON_CONDITION_END = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)
ON_RUN_END = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)
LOG_START = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(RunInfoC))
LOG_LOG = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_double)
LOG_FINISH = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(RunInfoC), C.c_void_p, C.c_double)


class CallbackC(C.Structure):
    _fields_ = [("user", C.c_void_p), ("on_run_start", ON_RUN_START), ("on_condition_start", ON_CONDITION_START),
                ("on_batch_end", ON_BATCH_END), ("on_condition_end", ON_CONDITION_END), ("on_run_end", ON_RUN_END)]


class LoggerC(C.Structure):
    _fields_ = [("user", C.c_void_p), ("start", LOG_START), ("log", LOG_LOG), ("finish", LOG_FINISH)]


class NrLink(C.Structure):
    _fields_ = [("modulation", C.c_int), ("snr_kind", C.c_int), ("d_learned_params", C.c_void_p),
                ("learned_iterations", C.c_uint32), ("fallback", C.c_int)]


class Metric(C.Structure):
    _fields_ = [("value", C.c_double), ("lo", C.c_double), ("hi", C.c_double), ("count", C.c_uint64),
                ("total", C.c_uint64), ("is_rate", C.c_int), ("dropped", C.c_uint64)]


BATCH_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_size_t, C.POINTER(C.c_double), C.c_size_t, C.c_uint64,
                       C.c_uint64, C.c_void_p)

SIGNATURES += [
    ("cf_generate_batch", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_uint32, C.c_uint64,
                                    C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("cf_record_errors", C.c_int, [C.c_void_p, C.c_char_p, C.c_uint64, C.c_uint64]),
    ("cf_record_value", C.c_int, [C.c_void_p, C.c_char_p, C.c_double]),
    ("cf_record_fail", C.c_int, [C.c_void_p, C.c_int, C.c_char_p]),
    ("cf_run", C.c_int, [C.POINTER(RunnerConfig), C.POINTER(StopNode), C.c_size_t, C.POINTER(GridAxis), C.c_size_t,
                         C.c_char_p, BATCH_FN, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("cf_nr_campaign", C.c_int, [C.c_void_p, C.POINTER(NrLink), C.POINTER(RunnerConfig), C.POINTER(StopNode),
                                 C.c_size_t, C.POINTER(GridAxis), C.c_size_t, C.POINTER(C.c_void_p)]),
    ("cf_results_count", C.c_size_t, [C.c_void_p]),
    ("cf_results_condition", C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_double)]),
    ("cf_results_metric", C.c_int, [C.c_void_p, C.c_size_t, C.c_char_p, C.c_double, C.POINTER(Metric)]),
    ("cf_results_free", None, [C.c_void_p]),
    ("cf_results_values", C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(C.c_double),
                                    C.c_size_t, C.POINTER(C.c_size_t)]),
    ("cf_results_metric_count", C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("cf_results_metric_name", C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, C.POINTER(C.c_char_p)]),
    ("cf_record_metric", C.c_int, [C.c_void_p, C.c_char_p, C.c_double, C.POINTER(Metric)]),
    ("cf_hooks_create", C.c_int, [C.POINTER(C.c_void_p)]),
    ("cf_hooks_add_callback", C.c_int, [C.c_void_p, C.POINTER(CallbackC)]),
    ("cf_hooks_add_logger", C.c_int, [C.c_void_p, C.POINTER(LoggerC)]),
    ("cf_hooks_progress_bar", C.c_int, [C.c_void_p, C.c_double]),
    ("cf_hooks_skip_after", C.c_int, [C.c_void_p, C.c_char_p, C.c_double, C.c_char_p]),
    ("cf_hooks_csv_logger", C.c_int, [C.c_void_p, C.c_char_p]),
    ("cf_hooks_json_logger", C.c_int, [C.c_void_p, C.c_char_p]),
    ("cf_hooks_destroy", None, [C.c_void_p]),
    ("cf_wilson", C.c_int, [C.c_uint64, C.c_uint64, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("cf_constellation_points", C.c_int, [C.c_int, _f64p, _f64p, C.c_size_t]),
]

BPSK, QPSK, QAM16, QAM64 = 1, 2, 4, 6
EBN0_DB, ESN0_DB = 0, 1


def constellation_points(modulation):
    """The unit-energy Gray QAM of the device generator (Constellation::new(Qam, M)
    .normalized().points(), constellation.rs:168-205, 396-398) as complex128, point i the
    symbol of label i."""
    m = 1 << modulation
    re, im = np.zeros(m), np.zeros(m)
    _check(lib().cf_constellation_points(modulation, re, im, m))
    return re + 1j * im

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

    @classmethod
    def from_rows(cls, n, rows):
        """LdpcCode::from_rows (crates/com-fec/src/ldpc.rs:216-231): rows of H as lists of
        variables; K = N - rank H at the non-pivot positions, sent whole (E = N)."""
        start = np.zeros(len(rows) + 1, np.uint32)
        start[1:] = np.cumsum([len(r) for r in rows])
        cols = np.ascontiguousarray([v for r in rows for v in r], np.uint32)
        h = C.c_void_p()
        _check(lib().cf_code_from_rows(n, len(rows), start, cols, C.byref(h)))
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

    def decode_batch(self, llr, *, want_post=False, want_bits=True, half=False, stream=None):
        """Decode a (count, E) torch tensor on the GPU (float32 or float64 messages, the type of
        llr; half=True: float32 LLRs with the messages stored in binary16): returns
        (bits (count, K') uint8, iterations (count,) int32, post (count, N) or None)."""
        import torch
        if llr.dim() != 2 or llr.shape[1] != self.dims.e:
            raise ComfecError(CF_SHAPE, f"inconsistent shapes: {llr.shape[-1]} LLRs for a block of {self.dims.e}")
        if not llr.is_cuda or not llr.is_contiguous():
            raise ComfecError(CF_CONFIG, "invalid configuration: the LLRs must be a contiguous device tensor")
        msg = {torch.float64: MSG_F64, torch.float32: MSG_F32}[llr.dtype]
        if half:
            if llr.dtype != torch.float32:
                raise ComfecError(CF_CONFIG, "invalid configuration: FP16 storage takes float32 LLRs")
            msg = MSG_F16
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


# ---- campaigns: the Runner / Stop / Grid API of crates/com-sim -------------------------------

STOP_MAX_TRIALS, STOP_MIN_TRIALS, STOP_MIN_ERRORS, STOP_PRECISION, STOP_MAX_TIME, STOP_ANY, STOP_ALL = range(7)


class Stop:
    """stopping.rs:64-141: the rules and their `or` / `and` combinations."""

    def __init__(self, kind, count=0, value=0.0, confidence=0.95, metric=None, children=()):
        self.kind, self.count, self.value, self.confidence, self.metric = kind, count, value, confidence, metric
        self.children = list(children)

    @staticmethod
    def max_trials(n):
        return Stop(STOP_MAX_TRIALS, count=n)

    @staticmethod
    def min_trials(n):
        return Stop(STOP_MIN_TRIALS, count=n)

    @staticmethod
    def min_errors(metric, errors):
        return Stop(STOP_MIN_ERRORS, count=errors, metric=metric)

    @staticmethod
    def precision(metric, relative, confidence=0.95):
        return Stop(STOP_PRECISION, value=relative, confidence=confidence, metric=metric)

    @staticmethod
    def max_time(seconds):
        return Stop(STOP_MAX_TIME, value=seconds)

    def or_(self, other):
        return Stop(STOP_ANY, children=[self, other])

    def and_(self, other):
        return Stop(STOP_ALL, children=[self, other])

    __or__ = or_
    __and__ = and_

    def nodes(self):
        out = [StopNode(self.kind, len(self.children), int(self.count), float(self.value), float(self.confidence),
                        self.metric.encode() if self.metric else None)]
        for c in self.children:
            out.extend(c.nodes())
        return out


class Grid:
    """grid.rs:231-290: axes in order (the last fastest) and fixed values."""

    def __init__(self):
        self.parts = []

    def axis(self, name, values):
        self.parts.append((name, [float(v) for v in values]))
        return self

    def fixed(self, name, value):
        return self.axis(name, [value])

    def _axes(self):
        keep = []
        arr = (GridAxis * max(1, len(self.parts)))()
        for i, (name, values) in enumerate(self.parts):
            buf = (C.c_double * max(1, len(values)))(*values)
            keep.append(buf)
            arr[i] = GridAxis(name.encode(), C.cast(buf, C.POINTER(C.c_double)), len(values))
        return arr, keep


class Record:
    """The record a scenario fills for a batch of trials (scenario.rs:66-198)."""

    def __init__(self, handle):
        self._h = handle

    def errors(self, name, errors, total):
        _check(lib().cf_record_errors(self._h, name.encode(), int(errors), int(total)))

    def value(self, name, v):
        _check(lib().cf_record_value(self._h, name.encode(), float(v)))


STATUS = {0: "completed", 1: "stopped", 2: "skipped", 3: "dev_run"}


class Action:
    """What a callback asks the runner to do (callbacks.rs:50-61)."""
    CONTINUE, SKIP, STOP_CONDITION, STOP_RUN = 0, 1, 2, 3


class RunInfo:
    """What the callbacks and loggers know about the run (runner.rs:150-175)."""

    def __init__(self, c):
        d = lambda b: b.decode() if b is not None else None
        self.scenario, self.seed, self.batch, self.threads = d(c.scenario), c.seed, c.batch, c.threads
        self.stop, self.confidence, self.n_conditions = d(c.stop), c.confidence, c.n_conditions
        self.started_unix_s, self.version, self.commit, self.config = (c.started_unix_s, d(c.version), d(c.commit),
                                                                       d(c.config))


class Progress:
    """The state of a condition after a batch (stopping.rs:48-59); `metric` reads its merged
    record, valid during the callback only."""

    def __init__(self, c):
        self.condition, self.trials, self.elapsed_s, self._rec = c.condition, c.trials, c.elapsed_s, c.record

    def metric(self, name, confidence=0.95):
        m = Metric()
        if lib().cf_record_metric(self._rec, name.encode(), confidence, C.byref(m)) != CF_OK:
            return None
        return m


class ConditionResult:
    """runner.rs:80-104: a condition's status, trials, wall time and metrics."""

    def __init__(self, index, condition, status, trials, wall_s, metrics, confidence):
        self.index, self.condition, self.status, self.trials, self.wall_s = index, condition, status, trials, wall_s
        self.metrics, self.confidence = metrics, confidence

    def metric(self, name):
        return self.metrics.get(name)


def _result(res, i, keys, confidence):
    st, tr, wall = C.c_int(), C.c_uint64(), C.c_double()
    _check(lib().cf_results_condition(res, i, C.byref(st), C.byref(tr), C.byref(wall)))
    index, n = C.c_size_t(), C.c_size_t()
    vals = (C.c_double * max(1, len(keys)))()
    _check(lib().cf_results_values(res, i, C.byref(index), vals, len(keys), C.byref(n)))
    count = C.c_size_t()
    _check(lib().cf_results_metric_count(res, i, C.byref(count)))
    metrics = {}
    for j in range(count.value):
        name = C.c_char_p()
        _check(lib().cf_results_metric_name(res, i, j, C.byref(name)))
        m = Metric()
        _check(lib().cf_results_metric(res, i, name.value, confidence, C.byref(m)))
        metrics[name.value.decode()] = m
    cond = {k: vals[q] for q, k in enumerate(keys[:n.value])}
    return ConditionResult(index.value, cond, STATUS[st.value], tr.value, wall.value, metrics, confidence)


def _collect(res, keys, confidence):
    try:
        return [_result(res, i, keys, confidence) for i in range(lib().cf_results_count(res))]
    finally:
        lib().cf_results_free(res)


def _conditions(grid):
    conds = [dict()]
    for name, values in grid.parts:
        conds = [dict(c, **{name: v}) for c in conds for v in values]
    return conds


class ProgressBar:
    """callbacks.rs:85-166: a line per condition on the standard error, and a progress line
    within a condition every `every_s` seconds."""

    def __init__(self, every_s=5.0):
        self.every_s = float(every_s)

    def _install(self, hooks):
        _check(lib().cf_hooks_progress_bar(hooks, self.every_s))


class SkipAfter:
    """callbacks.rs:168-203: once `metric` falls below `threshold`, skip the later conditions
    equal to that one on every key but `along`."""

    def __init__(self, metric, threshold, along):
        self.metric, self.threshold, self.along = metric, float(threshold), along

    @staticmethod
    def below(metric, threshold, along):
        return SkipAfter(metric, threshold, along)

    def _install(self, hooks):
        _check(lib().cf_hooks_skip_after(hooks, self.metric.encode(), self.threshold, self.along.encode()))


class CsvLogger:
    """loggers.rs:107-215: one CSV row per condition, written and flushed as it ends."""

    def __init__(self, path):
        self.path = os.fspath(path)

    def _install(self, hooks):
        _check(lib().cf_hooks_csv_logger(hooks, self.path.encode()))


class JsonLogger:
    """loggers.rs:221-318: the full report as JSON at the end of the run."""

    def __init__(self, path):
        self.path = os.fspath(path)

    def _install(self, hooks):
        _check(lib().cf_hooks_json_logger(hooks, self.path.encode()))


class _Hooks:
    """The native hooks of one run: built-in observers, and adapters for Python objects with
    any of on_run_start(info), on_condition_start(index, condition) -> Action,
    on_batch_end(progress) -> Action, on_condition_end(result) -> Action, on_run_end(results)
    (a callback) or start(info), log(result, confidence), finish(info, results, wall_s) (a
    logger). An exception in Python ends the run (the condition is skipped or stopped, the
    run stopped, a logger fails) and is raised once the run returns."""

    def __init__(self, callbacks, loggers, keys, confidence):
        self.errors, self.keep = [], []
        self.h = C.c_void_p()
        _check(lib().cf_hooks_create(C.byref(self.h)))
        try:
            for c in callbacks:
                c._install(self.h) if hasattr(c, "_install") else self._callback(c, keys, confidence)
            for l in loggers:
                l._install(self.h) if hasattr(l, "_install") else self._logger(l, keys, confidence)
        except BaseException:
            self.close()
            raise

    def _guard(self, f, on_error):
        def g(*a):
            try:
                return f(*a)
            except BaseException as e:  # noqa: BLE001 - carried across the C boundary
                self.errors.append(e)
                return on_error
        return g

    def _callback(self, obj, keys, conf):
        fn = lambda name: getattr(obj, name, None)
        act = lambda r: Action.CONTINUE if r is None else int(r)
        rs, cs, be, ce, re_ = (fn("on_run_start"), fn("on_condition_start"), fn("on_batch_end"),
                               fn("on_condition_end"), fn("on_run_end"))
        t = CallbackC()
        if rs:
            t.on_run_start = ON_RUN_START(self._guard(lambda u, i: rs(RunInfo(i.contents)), None))
        if cs:
            t.on_condition_start = ON_CONDITION_START(self._guard(
                lambda u, i, v, n: act(cs(i, {keys[q]: v[q] for q in range(n)})), Action.SKIP))
        if be:
            t.on_batch_end = ON_BATCH_END(self._guard(lambda u, p: act(be(Progress(p.contents))),
                                                      Action.STOP_CONDITION))
        if ce:
            t.on_condition_end = ON_CONDITION_END(self._guard(
                lambda u, r: act(ce(_result(r, 0, keys, conf))), Action.STOP_RUN))
        if re_:
            t.on_run_end = ON_RUN_END(self._guard(
                lambda u, r: re_([_result(r, i, keys, conf) for i in range(lib().cf_results_count(r))]), None))
        self.keep.append(t)
        _check(lib().cf_hooks_add_callback(self.h, C.byref(t)))

    def _logger(self, obj, keys, conf):
        fn = lambda name: getattr(obj, name, None)
        st, lg, fi = fn("start"), fn("log"), fn("finish")
        ok = lambda f: (lambda *a: (f(*a), 0)[1])
        t = LoggerC()
        if st:
            t.start = LOG_START(self._guard(ok(lambda u, i: st(RunInfo(i.contents))), 1))
        if lg:
            t.log = LOG_LOG(self._guard(ok(lambda u, r, c: lg(_result(r, 0, keys, c), c)), 1))
        if fi:
            t.finish = LOG_FINISH(self._guard(ok(lambda u, i, r, w: fi(
                RunInfo(i.contents), [_result(r, q, keys, conf) for q in range(lib().cf_results_count(r))], w)), 1))
        self.keep.append(t)
        _check(lib().cf_hooks_add_logger(self.h, C.byref(t)))

    def close(self):
        if self.h:
            lib().cf_hooks_destroy(self.h)
            self.h = C.c_void_p()


class Runner:
    """runner.rs:188-473: seed, batch, threads, stop, confidence, fast_dev_run, config,
    callbacks and loggers; run(scenario, grid)."""

    def __init__(self):
        self._seed, self._batch, self._stop, self._conf, self._dev = 0, 32, Stop.max_trials(100), 0.95, False
        self._threads, self._config_json, self._callbacks, self._loggers = 0, None, [], []

    def seed(self, s):
        self._seed = s
        return self

    def batch(self, n):
        self._batch = max(1, int(n))
        return self

    def threads(self, n):
        self._threads = int(n)
        return self

    def stop(self, rule):
        self._stop = rule
        return self

    def confidence(self, c):
        self._conf = c
        return self

    def fast_dev_run(self, on=True):
        self._dev = bool(on)
        return self

    def config(self, json_doc):
        self._config_json = str(json_doc)
        return self

    def callback(self, c):
        self._callbacks.append(c)
        return self

    def logger(self, l):
        self._loggers.append(l)
        return self

    def _config(self, hooks=None):
        return RunnerConfig(self._seed, self._batch, self._conf, int(self._dev), self._threads,
                            self._config_json.encode() if self._config_json is not None else None,
                            hooks.h if hooks is not None else None)

    def _stop_arr(self):
        nodes = self._stop.nodes()
        arr = (StopNode * len(nodes))(*nodes)
        return arr, len(nodes)

    def _run(self, grid, call, errors=()):
        keys = [name for name, _ in grid.parts]
        hooks = _Hooks(self._callbacks, self._loggers, keys, self._conf)
        try:
            res = C.c_void_p()
            status = call(self._config(hooks), C.byref(res))
            failures = list(errors) + hooks.errors
            if failures:
                if status == CF_OK:
                    lib().cf_results_free(res)
                raise failures[0]
            _check(status)
            return _collect(res, keys, self._conf)
        finally:
            hooks.close()

    def run(self, name, batch_fn, grid, metrics=None):
        """Run a scenario given as batch_fn(index, condition: dict, t0, t1, record); every
        metric the scenario records is in the results (`metrics` is accepted for older
        callers and ignored)."""
        errors = []
        conds = _conditions(grid)

        def cb(user, index, values, n_values, t0, t1, rec):
            try:
                batch_fn(index, conds[index], t0, t1, Record(rec))
                return 0
            except ComfecError as e:
                lib().cf_record_fail(rec, e.status, str(e).split(": ", 1)[-1].encode())
                return 1
            except Exception as e:  # a Python failure stops the run with its message
                errors.append(e)
                lib().cf_record_fail(rec, CF_CONFIG, repr(e).encode())
                return 1

        fn = BATCH_FN(cb)
        axes, keep = grid._axes()
        stop, n_stop = self._stop_arr()
        return self._run(grid, lambda cfg, out: lib().cf_run(C.byref(cfg), stop, n_stop, axes, len(grid.parts),
                                                             name.encode(), fn, None, out), errors)

    def run_nr(self, code, grid, modulation=BPSK, snr_kind=EBN0_DB, learned_params=None, fallback=False,
               metrics=None):
        """Runner::run with the NR link on the GPU as the scenario (cf_nr_campaign). With
        learned_params (iterations, rows, 2), the learned decoder runs its iterations and, with
        fallback, the code's own min-sum on the codewords whose syndrome still fails."""
        link = NrLink(modulation, snr_kind, C.c_void_p(learned_params.data_ptr()) if learned_params is not None
                      else None, learned_params.shape[0] if learned_params is not None else 0, int(fallback))
        axes, keep = grid._axes()
        stop, n_stop = self._stop_arr()
        return self._run(grid, lambda cfg, out: lib().cf_nr_campaign(code._h, C.byref(link), C.byref(cfg), stop,
                                                                     n_stop, axes, len(grid.parts), out))


def wilson(errors, total, confidence=0.95):
    """ErrorRate::wilson (crates/com-sim/src/stats.rs:74-87)."""
    lo, hi = C.c_double(), C.c_double()
    _check(lib().cf_wilson(int(errors), int(total), confidence, C.byref(lo), C.byref(hi)))
    return lo.value, hi.value


def generate_batch(code, count, modulation, snr_db, snr_kind=EBN0_DB, seed=1, point=0, trial0=0, stream=None):
    """The device trial generator: (LLRs (count, E) float32, packed info bits (count, ceil(K'/32)) int32)."""
    import torch
    llr = torch.empty((count, code.dims.e), dtype=torch.float32, device="cuda")
    info = torch.empty((count, (code.dims.k_info + 31) // 32), dtype=torch.int32, device="cuda")
    _check(lib().cf_generate_batch(code._h, modulation, snr_kind, snr_db, seed, point, trial0, count,
                                   C.c_void_p(llr.data_ptr()), C.c_void_p(info.data_ptr()), _stream_ptr(stream)))
    return llr, info


def unpack_info(info, k_info):
    """Packed information bits (LSB first) to a (count, K') uint8 array."""
    w = np.ascontiguousarray(info.cpu().numpy() if hasattr(info, "cpu") else info).view(np.uint32)
    bits = (w[:, :, None] >> np.arange(32, dtype=np.uint32)) & 1
    return bits.reshape(w.shape[0], -1)[:, :k_info].astype(np.uint8)


# ---- the learned offset/scale min-sum ---------------------------------------------------------

SIGNATURES += [
    ("cf_generate_training_batch", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint64,
                                             C.c_uint32, C.c_uint64, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("cf_learned_loss_grad", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                       C.c_void_p, C.c_void_p, C.c_void_p]),
    ("cf_adam_step", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_uint32, C.c_double,
                               C.c_double, C.c_double, C.c_double, C.c_void_p]),
    ("cf_learned_train", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_double, C.c_int,
                                   C.c_int, C.c_double, C.c_double, C.c_uint64, C.c_void_p, C.c_void_p]),
    ("cf_learned_decode_batch", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_int, C.c_void_p, C.c_size_t,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
]


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def initial_params(code, iterations, alpha=0.75, beta=0.0):
    """(a, b) = (alpha, beta) for every iteration and base row: the plain min-sum."""
    import torch
    rows = 46 if code.dims.base_graph == 1 else 42
    p = torch.empty((iterations, rows, 2), dtype=torch.float32, device="cuda")
    p[..., 0] = alpha
    p[..., 1] = beta
    return p


def generate_training_batch(code, count, modulation, snr_lo, snr_hi, snr_kind=EBN0_DB, seed=1, point=0, trial0=0,
                            stream=None):
    """Training LLRs with per-trial SNR uniform in [lo, hi], and the packed codeword bits."""
    import torch
    llr = torch.empty((count, code.dims.e), dtype=torch.float32, device="cuda")
    cw = torch.empty((count, (code.dims.n + 31) // 32), dtype=torch.int32, device="cuda")
    _check(lib().cf_generate_training_batch(code._h, modulation, snr_kind, snr_lo, snr_hi, seed, point, trial0, count,
                                            _ptr(llr), _ptr(cw), _stream_ptr(stream)))
    return llr, cw


def learned_loss_grad(code, params, llr, cw, stream=None):
    import torch
    grad = torch.empty_like(params)
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    _check(lib().cf_learned_loss_grad(code._h, params.shape[0], _ptr(params), _ptr(llr), _ptr(cw), llr.shape[0],
                                      _ptr(grad), _ptr(loss), _stream_ptr(stream)))
    return loss, grad


def adam_step(params, grad, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
    _check(lib().cf_adam_step(_ptr(params), _ptr(grad), _ptr(m), _ptr(v), params.numel(), step, lr, beta1, beta2,
                              eps, _stream_ptr(stream)))


def learned_train(code, params, steps, batch, lr=1e-3, modulation=BPSK, snr_kind=EBN0_DB, snr_lo=0.0, snr_hi=3.0,
                  seed=1, stream=None):
    """Train in place; returns the loss of every step (numpy)."""
    hist = np.zeros(steps, np.float32)
    _check(lib().cf_learned_train(code._h, params.shape[0], _ptr(params), steps, batch, lr, modulation, snr_kind,
                                  snr_lo, snr_hi, seed, hist.ctypes.data_as(C.c_void_p), _stream_ptr(stream)))
    return hist


def learned_decode_batch(code, params, llr, fallback=True, stream=None):
    """(bits (count, K'), iterations (count,), ok (count,)) with the learned parameters."""
    import torch
    count = llr.shape[0]
    bits = torch.empty((count, code.dims.k_info), dtype=torch.uint8, device="cuda")
    its = torch.empty(count, dtype=torch.int32, device="cuda")
    ok = torch.empty(count, dtype=torch.uint8, device="cuda")
    _check(lib().cf_learned_decode_batch(code._h, params.shape[0], _ptr(params), int(fallback), _ptr(llr), count,
                                         _ptr(bits), _ptr(its), _ptr(ok), _stream_ptr(stream)))
    return bits, its, ok


# ---- batches of many codes ----------------------------------------------------------------------


class PlanDims(C.Structure):
    _fields_ = [("llr_total", C.c_uint64), ("bits_total", C.c_uint64), ("codewords", C.c_uint32),
                ("classes", C.c_uint32), ("tasks", C.c_uint32), ("threads", C.c_uint32), ("grid", C.c_uint32),
                ("lds_bytes", C.c_uint32), ("msg_type", C.c_int), ("segments", C.c_uint32)]


class PlanConfig(C.Structure):
    """cf_batch_plan_config: one launch at `threads` lanes, or (per_code_width) each code at
    its own width; grid_per_cu 0 = as many resident workgroups as fit."""
    _fields_ = [("threads", C.c_uint32), ("per_code_width", C.c_uint32), ("grid_per_cu", C.c_uint32),
                ("msg_type", C.c_int), ("overlap", C.c_uint32)]

    def __repr__(self):
        layout = "per-code widths" if self.per_code_width else f"{self.threads} lanes"
        if self.per_code_width:
            layout += ", overlapped" if self.overlap else ", in turn"
        grid = f"{self.grid_per_cu}/CU" if self.grid_per_cu else "all resident"
        return f"PlanConfig({layout}, grid {grid}, {'FP16' if self.msg_type == MSG_F16 else 'FP32'})"


class PlanSegment(C.Structure):
    _fields_ = [("threads", C.c_uint32), ("grid", C.c_uint32), ("lds_bytes", C.c_uint32), ("classes", C.c_uint32),
                ("tasks", C.c_uint32), ("codewords", C.c_uint32)]


class PlanTrial(C.Structure):
    _fields_ = [("config", PlanConfig), ("segments", C.c_uint32), ("ms", C.c_double)]


TUNE_STORAGE = 1

SIGNATURES += [
    ("cf_batch_plan_create", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
                                       C.POINTER(C.c_void_p)]),
    ("cf_batch_plan_destroy", None, [C.c_void_p]),
    ("cf_batch_plan_dims_get", C.c_int, [C.c_void_p, C.POINTER(PlanDims)]),
    ("cf_batch_plan_set_threads", C.c_int, [C.c_void_p, C.c_uint32]),
    ("cf_batch_plan_config_get", C.c_int, [C.c_void_p, C.POINTER(PlanConfig)]),
    ("cf_batch_plan_configure", C.c_int, [C.c_void_p, C.POINTER(PlanConfig)]),
    ("cf_batch_plan_segment_get", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(PlanSegment)]),
    ("cf_batch_plan_tune", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                     C.c_uint32, C.POINTER(PlanTrial), C.c_size_t, C.POINTER(C.c_size_t),
                                     C.POINTER(C.c_double), C.c_void_p]),
    ("cf_batch_plan_run", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                    C.c_void_p]),
    ("cf_decode_batch_uneven", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
]


def _plan_inputs(codes, llr_offsets, max_iterations):
    handles = (C.c_void_p * len(codes))(*[c._h for c in codes])
    offs = np.ascontiguousarray(llr_offsets, np.uint64)
    if offs.shape != (len(codes) + 1,):
        raise ComfecError(CF_SHAPE, f"inconsistent shapes: {offs.size} offsets for {len(codes)} codewords")
    lim = None if max_iterations is None else np.ascontiguousarray(max_iterations, np.uint32)
    return handles, offs, lim


class BatchPlan:
    """A decode plan for codewords of many codes (cf_batch_plan_*): codes[i] decodes
    llr[llr_offsets[i]:llr_offsets[i+1]]; the decisions are concatenated in codeword order."""

    def __init__(self, codes, llr_offsets, max_iterations=None, half=False):
        self.codes = list(codes)  # the handles outlive the plan
        handles, offs, lim = _plan_inputs(self.codes, llr_offsets, max_iterations)
        self._h = C.c_void_p()
        _check(lib().cf_batch_plan_create(handles, offs.ctypes.data_as(C.c_void_p),
                                          lim.ctypes.data_as(C.c_void_p) if lim is not None else None,
                                          len(self.codes), MSG_F16 if half else MSG_F32, C.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.cf_batch_plan_destroy(h)
            self._h = None

    @property
    def dims(self):
        d = PlanDims()
        _check(lib().cf_batch_plan_dims_get(self._h, C.byref(d)))
        return d

    def outputs(self):
        import torch
        d = self.dims
        return (torch.empty(d.bits_total, dtype=torch.uint8, device="cuda"),
                torch.empty(d.codewords, dtype=torch.int32, device="cuda"),
                torch.empty(d.codewords, dtype=torch.uint8, device="cuda"))

    def set_threads(self, threads):
        _check(lib().cf_batch_plan_set_threads(self._h, threads))
        return self

    @property
    def config(self):
        c = PlanConfig()
        _check(lib().cf_batch_plan_config_get(self._h, C.byref(c)))
        return c

    def configure(self, threads=None, per_code_width=None, grid_per_cu=None, half=None, overlap=None):
        """Change the configuration; the settings not given are kept."""
        c = self.config
        if overlap is not None:
            c.overlap = int(overlap)
        if threads is not None:
            c.threads = threads
        if per_code_width is not None:
            c.per_code_width = int(per_code_width)
        if grid_per_cu is not None:
            c.grid_per_cu = grid_per_cu
        if half is not None:
            c.msg_type = MSG_F16 if half else MSG_F32
        _check(lib().cf_batch_plan_configure(self._h, C.byref(c)))
        return self

    def segments(self):
        """The launches of a run (PlanSegment each)."""
        out = []
        for i in range(self.dims.segments):
            sg = PlanSegment()
            _check(lib().cf_batch_plan_segment_get(self._h, i, C.byref(sg)))
            out.append(sg)
        return out

    def tune(self, llr, out=None, repeats=3, storage=False, stream=None):
        """Autotune on this input: returns (best ms, outputs); the trials (PlanTrial each) are
        left in self.trials. storage: also try the other storage type."""
        bits, its, ok = out if out is not None else self.outputs()
        ms = C.c_double()
        cap = 64
        trials = (PlanTrial * cap)()
        n = C.c_size_t()
        _check(lib().cf_batch_plan_tune(self._h, _ptr(llr), _ptr(bits), _ptr(its), _ptr(ok), repeats,
                                        TUNE_STORAGE if storage else 0, trials, cap, C.byref(n), C.byref(ms),
                                        _stream_ptr(stream)))
        self.trials = [trials[i] for i in range(min(n.value, cap))]
        return ms.value, (bits, its, ok)

    def run(self, llr, out=None, graph=False, stream=None):
        """Decode; returns (bits (sum K',), iterations (count,), ok (count,))."""
        bits, its, ok = out if out is not None else self.outputs()
        _check(lib().cf_batch_plan_run(self._h, _ptr(llr), _ptr(bits), _ptr(its), _ptr(ok), int(graph),
                                       _stream_ptr(stream)))
        return bits, its, ok


def decode_batch_uneven(codes, llr, llr_offsets, max_iterations=None, half=False, stream=None):
    """One-shot cf_decode_batch_uneven: (bits (sum K',), iterations, ok)."""
    import torch
    handles, offs, lim = _plan_inputs(codes, llr_offsets, max_iterations)
    total = sum(c.dims.k_info for c in codes)
    bits = torch.empty(total, dtype=torch.uint8, device="cuda")
    its = torch.empty(len(codes), dtype=torch.int32, device="cuda")
    ok = torch.empty(len(codes), dtype=torch.uint8, device="cuda")
    _check(lib().cf_decode_batch_uneven(handles, offs.ctypes.data_as(C.c_void_p),
                                        lim.ctypes.data_as(C.c_void_p) if lim is not None else None, len(codes),
                                        MSG_F16 if half else MSG_F32, _ptr(llr), _ptr(bits), _ptr(its), _ptr(ok),
                                        _stream_ptr(stream)))
    return bits, its, ok
