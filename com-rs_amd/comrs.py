"""The reference's Python front end of the NR LDPC code, on the GPU decoder.

`comrs.NrLdpc(k, e, *, base_graph=None, rv=0, qm=1, rule="sum_product", alpha=0.8,
max_iterations=50, layered=True)` with the same constructor, properties (`k`, `n`,
`base_graph`, `rate`, `lifting_size`) and whole-block `encode` / `decode` as the binding of
bindings/com-py/src/fec/nr.rs:36-140, the defaults of crates/com-fec/src/ldpc.rs:75-84, and
the same errors: a com_core::Error of kind Config, Shape or Condition is a ValueError, Numerical
an ArithmeticError, Io an OSError (bindings/com-py/src/lib.rs:96-105), with the reference's
message ("inconsistent shapes: 301 LLRs is not a multiple of the block length 300",
crates/com-fec/src/block.rs:209-214).

The reference decodes a stream block by block (BlockCode::decode_stream, block.rs:81-89, on
one thread: bindings/com-py/src/fec/mod.rs:40-45). Here `decode` sends every block of the
call to the device at once and decodes them in one launch of cf_decode_batch, in f64 (the
reference's arithmetic: bit-identical decisions for the min-sum rules) unless the keyword
`precision="f32"` asks for the FP32 kernels.
"""
import numpy as np

import comfec as _F

_RULES = ("sum_product", "min_sum", "normalized_min_sum", "offset_min_sum")


def _py_error(e):
    """bindings/com-py/src/lib.rs:96-105: the Python exception of a com_core::Error."""
    if e.kind == "Numerical":
        return ArithmeticError(str(e))
    if e.kind == "Io":
        return OSError(str(e))
    return ValueError(str(e))


def _bits(x):
    """bindings/com-py/src/lib.rs:145-153: a 1-D array-like of bits."""
    a = np.ascontiguousarray(x, np.uint8)
    if a.ndim != 1:
        raise ValueError("expected a 1-D array of bits")
    if np.any(a > 1):
        raise ValueError("bits must be 0 or 1")
    return a


def _reals(x):
    """bindings/com-py/src/lib.rs:138-142: a 1-D array-like of real values."""
    a = np.ascontiguousarray(x, np.float64)
    if a.ndim != 1:
        raise ValueError("expected a 1-D array of real values")
    return a


def _whole_blocks(length, block, what):
    """crates/com-fec/src/block.rs:209-214."""
    if block == 0 or length % block != 0:
        raise ValueError(f"inconsistent shapes: {length} {what} is not a multiple of the block length {block}")


class NrLdpc:
    """The LDPC code of the data channels of 5G NR (3GPP TS 38.212, Sec. 5.3.2, 5.4.2): one
    code block of `k` information bits rate-matched to `e` bits, decoded by belief
    propagation on the GPU (bindings/com-py/src/fec/nr.rs:6-35)."""

    def __init__(self, k, e, *, base_graph=None, rv=0, qm=1, rule="sum_product", alpha=0.8, max_iterations=50,
                 layered=True, precision="f64"):
        if precision not in ("f64", "f32"):
            raise ValueError(f"invalid configuration: precision must be f64 or f32, got {precision!r}")
        try:
            code = _F.NrLdpc(int(k), int(e), base_graph, int(rv), int(qm))
            rid, a = _F.rule_from_name(rule, float(alpha))
            code.set_decoder(rid, a, bool(layered), int(max_iterations))
        except _F.ComfecError as err:
            raise _py_error(err) from None
        self._code = code
        self._precision = precision

    def __repr__(self):
        return (f"NrLdpc(k={self.k}, n={self.n}, base_graph={self.base_graph}, "
                f"lifting_size={self.lifting_size})")

    @property
    def k(self):
        """int: Information bits per block."""
        return self._code.dims.k_info

    @property
    def n(self):
        """int: Transmitted bits per block."""
        return self._code.dims.e

    @property
    def base_graph(self):
        """int: The base graph, 1 or 2."""
        return self._code.dims.base_graph

    @property
    def rate(self):
        """float: Code rate k / n."""
        return self.k / self.n

    @property
    def lifting_size(self):
        """int: The lifting size Z."""
        return self._code.dims.z

    def encode(self, bits):
        """Encode whole blocks: a multiple of `k` bits in, `n` bits per block out (uint8)."""
        b = _bits(bits)
        _whole_blocks(b.size, self.k, "bits")
        out = np.empty(b.size // self.k * self.n, np.uint8)
        try:
            for i in range(b.size // self.k):
                out[i * self.n:(i + 1) * self.n] = self._code.encode_block(b[i * self.k:(i + 1) * self.k])
        except _F.ComfecError as err:
            raise _py_error(err) from None
        return out

    def decode(self, llr):
        """Decode whole blocks of LLRs (positive for a 0): a multiple of `n` LLRs in, `k` bits
        per block out (uint8), every block in one device launch."""
        import torch
        l = _reals(llr)
        _whole_blocks(l.size, self.n, "LLRs")
        count = l.size // self.n
        if count == 0:
            return np.empty(0, np.uint8)
        dtype = torch.float64 if self._precision == "f64" else torch.float32
        dev = torch.from_numpy(l.reshape(count, self.n)).to(device="cuda", dtype=dtype).contiguous()
        try:
            bits, _, _ = self._code.decode_batch(dev)
        except _F.ComfecError as err:
            raise _py_error(err) from None
        return bits.cpu().numpy().reshape(-1)
