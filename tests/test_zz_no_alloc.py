"""cf_decode_batch_static (no counterpart in the reference: for callers that cannot allocate)
against cf_decode_batch, through the C ABI: the same decisions, iteration counts and
posteriors bit for bit, on two calls in a row (the second with fewer codewords, so that the
shared scratch holds the first call's contents), for the register-resident layered kernel
(FP32, NR BG1 Z=384) and the general kernel (f64, BG2 Z=52); a code whose tables are not on
the device yet is refused; and repeated calls take no device memory."""
import ctypes as C

import numpy as np
import pytest

import comfec as F

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def decode(entry, code, llr, msg):
    count = llr.shape[0]
    bits = torch.full((count, code.dims.k_info), 7, dtype=torch.uint8, device="cuda")
    its = torch.full((count,), -1, dtype=torch.int32, device="cuda")
    post = torch.full((count, code.dims.n), float("nan"), dtype=llr.dtype, device="cuda")
    status = entry(code._h, msg, C.c_void_p(llr.data_ptr()), count, C.c_void_p(bits.data_ptr()),
                   C.c_void_p(post.data_ptr()), C.c_void_p(its.data_ptr()),
                   C.c_void_p(torch.cuda.current_stream().cuda_stream))
    F._check(status)
    torch.cuda.synchronize()
    return bits.cpu().numpy(), its.cpu().numpy(), post.cpu().numpy()


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("case", ["layered f32 BG1 Z=384", "general f64 BG2 Z=52"])
def test_static_scratch_equals_the_allocating_entry_point(case):
    print(f"\n[no-alloc] {case}: two calls, the second smaller", flush=True)
    lib = F.lib()
    if case.startswith("layered"):
        code = F.NrLdpc(8448, 25344, 1, 0, 4).set_decoder("normalized_min_sum", 0.75, True, 10)
        llr, _ = F.generate_batch(code, 96, F.QAM16, 4.2, snr_kind=F.ESN0_DB, seed=31)
        msg = F.MSG_F32
    else:
        code = F.NrLdpc.lifted(2, 52).set_decoder("normalized_min_sum", 0.75, True, 20)
        llr, _ = F.generate_batch(code, 96, F.BPSK, 1.25, seed=32)
        llr = llr.double().contiguous()
        msg = F.MSG_F64
    want = decode(lib.cf_decode_batch, code, llr, msg)
    got = decode(lib.cf_decode_batch_static, code, llr, msg)
    assert same(got, want)
    small = llr[:17].contiguous()
    want2 = decode(lib.cf_decode_batch, code, small, msg)
    got2 = decode(lib.cf_decode_batch_static, code, small, msg)
    assert same(got2, want2)
    assert same(got2, tuple(x[:17] for x in want))
    # repeated calls take no device memory
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(3):
        again = decode(lib.cf_decode_batch_static, code, llr, msg)
    assert torch.cuda.mem_get_info()[0] == free0
    assert same(again, want)


def test_static_scratch_refuses_a_code_not_yet_on_the_device():
    print("\n[no-alloc] tables not resident", flush=True)
    lib = F.lib()
    code = F.NrLdpc(100, 300).set_decoder("normalized_min_sum", 0.75, True, 20)
    llr = torch.ones((4, 300), dtype=torch.float32, device="cuda")
    with pytest.raises(F.ComfecError) as e:
        decode(lib.cf_decode_batch_static, code, llr, F.MSG_F32)
    assert e.value.kind == "Config" and "tables are not on device" in str(e.value)
    want = decode(lib.cf_decode_batch, code, llr, F.MSG_F32)
    assert same(decode(lib.cf_decode_batch_static, code, llr, F.MSG_F32), want)
    assert (want[0] == 0).all() and (want[1] == 0).all()
