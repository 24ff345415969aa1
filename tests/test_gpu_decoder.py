"""Parity of the HIP decoder with the oracle, through the C ABI.

- f64 messages: bit-identical posteriors and iteration counts for the min-sum rules
  (the same IEEE operations in the same order, no FMA contraction).
- f32 messages: identical iteration counts and frame-error verdicts on every frame,
  identical hard decisions and bit-error counts on every frame without a recorded tie (a
  value whose FP32 sign differs from its f64 sign, recorded by the oracle's side-by-side
  run; TIE is the magnitude below which such a value is a tie on zero), posteriors within
  the relative tolerances written in each test.
"""
import numpy as np
import pytest

import comfec as F
import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SEED = 1
TIE = 1e-6


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def config1_llrs(count, ebn0_db=1.0, name="cfg1"):
    """BG2 Z=52 lifted, K=520, BPSK over AWGN, LLRs 2y/s2 of the 2600 sent bits."""
    ref = O.Code.lifted(2, 52)
    cws, llrs = [], []
    for t in range(count):
        cw, llr = ref.trial_bpsk_lifted(ebn0_db, SEED, name, 0, t)
        cws.append(cw)
        llrs.append(llr)
    return ref, np.array(cws), np.array(llrs)


RULES = [("nms", O.NMS, "normalized_min_sum", 0.75), ("oms", O.OMS, "offset_min_sum", 0.15)]


@pytest.mark.parametrize("layered", [True, False], ids=["layered", "flooding"])
@pytest.mark.parametrize("rule", RULES, ids=[r[0] for r in RULES])
def test_f64_batch_is_bit_exact(rule, layered):
    print(f"\n[f64 batch] rule={rule[0]} layered={layered}", flush=True)
    _, orule, name, alpha = rule
    ref, cws, full = config1_llrs(48)
    post_ref, it_ref = ref.decode_many(full, rule=orule, alpha=alpha, layered=layered, max_iterations=20,
                                       threads=8)
    code = F.NrLdpc.lifted(2, 52).set_decoder(name, alpha, layered, 20)
    bits, its, post = code.decode_batch(cuda(full[:, 104:]), want_post=True)
    torch.cuda.synchronize()
    post = post.cpu().numpy()
    assert np.array_equal(its.cpu().numpy(), it_ref)
    assert np.array_equal(post, post_ref)
    assert np.array_equal(bits.cpu().numpy(), (post_ref[:, :520] < 0).astype(np.uint8))
    # the workload is not trivial: the frames stop at many different iterations
    assert len(set(it_ref.tolist())) >= 4


def test_config1_full_batch_error_counts():
    """The CPU reference workload at its full size: BG2 Z=52 lifted (K=520, 2600 bits sent),
    BPSK at Eb/N0 1.0 dB, layered NMS 0.75, 20 iterations, 1024 codewords of the oracle's
    xoshiro streams. The GPU's f64 decisions, bit and frame error counts and iteration
    counts equal the oracle's on every codeword."""
    print("\n[config 1] 1024 codewords, f64", flush=True)
    ref, cws, full = config1_llrs(1024)
    post_ref, it_ref = ref.decode_many(full, rule=O.NMS, alpha=0.75, layered=True, max_iterations=20, threads=16)
    code = F.NrLdpc.lifted(2, 52).set_decoder("normalized_min_sum", 0.75, True, 20)
    bits, its, _ = code.decode_batch(cuda(full[:, 104:]))
    bits, its = bits.cpu().numpy(), its.cpu().numpy()
    bits_ref = (post_ref[:, :520] < 0).astype(np.uint8)
    assert np.array_equal(bits, bits_ref) and np.array_equal(its, it_ref)
    be, fe = int((bits != cws[:, :520]).sum()), int((bits != cws[:, :520]).any(1).sum())
    print(f"bit errors {be} of {1024 * 520}, frame errors {fe} of 1024, mean iterations {its.mean():.3f}",
          flush=True)
    # the oracle's counts for this seeded workload (FER 2.8 %, 8.98 iterations on average)
    assert (be, fe) == (2752, 29)


@pytest.mark.parametrize("k,e,rv,qm", [(100, 300, 0, 1), (500, 1200, 2, 4), (60, 1000, 1, 2), (1000, 1500, 3, 6)])
def test_host_entry_points_match_oracle(k, e, rv, qm):
    print(f"\n[host f64] k={k} e={e} rv={rv} qm={qm}", flush=True)
    code = F.NrLdpc(k, e, None, rv, qm).set_decoder("normalized_min_sum", 0.75, True, 25)
    ref = O.Code.nr(k, e, 0, rv, qm)
    rng = O.Rng.stream(SEED, "host", k, e)
    info = np.array([rng.bit() for _ in range(k)], np.uint8)
    tx = ref.nr_encode(info)
    s2 = O.lib.or_bpsk_noise_variance(1.5, k / e)
    llr = np.zeros(e)
    O.lib.or_bpsk_awgn_llr(tx, e, s2, rng.r, llr)
    bits_ref, it_ref = ref.decode_block(llr, rule=O.NMS, alpha=0.75, max_iterations=25)
    assert np.array_equal(code.decode_block(llr), bits_ref)
    full = ref.full_llr(llr)
    post_ref, it2 = ref.decode_llr(full, rule=O.NMS, alpha=0.75, max_iterations=25)
    post, it = code.decode_llr(full)
    assert it == it2 == it_ref and np.array_equal(post, post_ref)
    apri = np.linspace(-0.5, 0.5, k)
    app, it3 = code.app(llr, apri)
    full2 = full.copy()
    full2[:k] += apri
    post3, it3_ref = ref.decode_llr(full2, rule=O.NMS, alpha=0.75, max_iterations=25)
    assert it3 == it3_ref and np.array_equal(app, post3[:k])


def test_codeword_input_and_zero_iterations():
    print("\n[edge] codeword input, max_iterations 0, empty batch", flush=True)
    code = F.NrLdpc.lifted(2, 52).set_decoder("normalized_min_sum", 0.75, True, 20)
    ref = O.Code.lifted(2, 52)
    cw = ref.encode(np.random.default_rng(3).integers(0, 2, ref.k).astype(np.uint8))
    llr = np.where(cw == 0, 2.0, -2.0)
    post, it = code.decode_llr(llr)
    assert it == 0 and np.array_equal(post, llr)
    noisy = llr + np.random.default_rng(4).normal(0, 2.5, llr.size)
    code.set_decoder("normalized_min_sum", 0.75, True, 0)
    post, it = code.decode_llr(noisy)
    assert it == 0 and np.array_equal(post, noisy)
    bits, its, _ = code.decode_batch(cuda(np.zeros((0, code.n))))
    assert bits.shape == (0, 520)
    # -0.0 is a 0 decision, as in the reference (x < 0.0)
    z = np.zeros(ref.n)
    z[:] = -0.0
    code.set_decoder("normalized_min_sum", 0.75, True, 5)
    post, it = code.decode_llr(z)
    assert it == 0


def test_f64_sum_product_within_libm_rounding():
    print("\n[f64 spa]", flush=True)
    ref, cws, full = config1_llrs(16, ebn0_db=1.5, name="spa")
    post_ref, it_ref = ref.decode_many(full, rule=O.SPA, layered=True, max_iterations=20, threads=8)
    code = F.NrLdpc.lifted(2, 52).set_decoder("sum_product", 0.0, True, 20)
    bits, its, post = code.decode_batch(cuda(full[:, 104:]), want_post=True)
    post = post.cpu().numpy()
    its = its.cpu().numpy()
    same = its == it_ref
    assert same.mean() >= 0.9
    # device tanh/atanh differ from glibc's in the last ulps
    np.testing.assert_allclose(post[same], post_ref[same], rtol=1e-9, atol=1e-9)


def config2_llrs(esn0_db, count, k=8448, e=25344, qm=4, m=16):
    ref = O.Code.nr(k, e, 1, 0, qm)
    n0 = 1.0 / O.lib.or_db2lin(esn0_db)
    infos, llrs = [], []
    for t in range(count):
        info, llr = ref.trial_nr_qam(m, n0, SEED, "cfg2", int(esn0_db), t)
        infos.append(info)
        llrs.append(llr)
    return ref, np.array(infos), np.array(llrs)


@pytest.mark.parametrize("layered", [True, False], ids=["layered", "flooding"])
def test_f32_sum_product_bg1_z384_16qam(layered):
    """FP32 sum-product on BG1 Z=384, 16-QAM at Es/N0 4.2 dB, whose LLRs and posteriors run
    far past the point where tanh(l/2) rounds to 1 in FP32: against the FP32 oracle (the
    same log-domain check, device exp/expm1/log1p against glibc's) and the f64 reference
    decoder, the same iterations, decisions and frame errors, posteriors within 1e-4
    relative (1e-3 absolute near zero)."""
    print(f"\n[f32 spa BG1 Z=384 16-QAM] layered={layered}", flush=True)
    count, limit = 8, (20 if layered else 40)
    ref, infos, llrs = config2_llrs(4.2, count)
    llr32 = llrs.astype(np.float32)
    full = np.array([ref.full_llr(l.astype(np.float64)) for l in llr32])
    post64, it64 = ref.decode_many(full, rule=O.SPA, layered=layered, max_iterations=limit, threads=8)
    post32, it32 = ref.decode_many_f32(full.astype(np.float32), rule=O.SPA, layered=layered,
                                       max_iterations=limit, threads=8)
    code = F.NrLdpc(8448, 25344, 1, 0, 4).set_decoder("sum_product", 0.0, layered, limit)
    bits, its, post = code.decode_batch(cuda(llr32), want_post=True)
    bits, its, post = bits.cpu().numpy(), its.cpu().numpy(), post.cpu().numpy().astype(np.float64)
    assert not np.isnan(post).any() and np.abs(post64).max() > 60
    assert np.array_equal(its, it64) and np.array_equal(its, it32) and (its < limit).all()
    assert np.array_equal(bits, (post64[:, :8448] < 0).astype(np.uint8))
    assert not (bits != infos).any()
    np.testing.assert_allclose(post, post32, rtol=1e-4, atol=1e-3)
    np.testing.assert_allclose(post, post64, rtol=1e-4, atol=1e-3)


@pytest.mark.parametrize("rule", RULES, ids=[r[0] for r in RULES])
@pytest.mark.parametrize("layered", [True, False], ids=["layered", "flooding"])
def test_f32_bg2_z52_bit_exact_with_fp32_oracle(rule, layered):
    """BG2 Z=52 (a lifting size that is not a multiple of 32: the row-by-row syndrome)."""
    print(f"\n[f32 BG2 Z=52] rule={rule[0]} layered={layered}", flush=True)
    _, orule, name, alpha = rule
    ref, cws, full = config1_llrs(64, ebn0_db=1.25, name="f32")
    full32 = full.astype(np.float32)
    post_ref, it_ref = ref.decode_many_f32(full32, rule=orule, alpha=alpha, layered=layered, max_iterations=20,
                                           threads=8)
    code = F.NrLdpc.lifted(2, 52).set_decoder(name, alpha, layered, 20)
    bits, its, post = code.decode_batch(cuda(full32[:, 104:]), want_post=True)
    assert np.array_equal(its.cpu().numpy(), it_ref)
    assert np.array_equal(post.cpu().numpy(), post_ref)
    assert np.array_equal(bits.cpu().numpy(), (post_ref[:, :520] < 0).astype(np.uint8))


@pytest.mark.parametrize("esn0_db", [4.0, 4.5, 5.0, 7.0])
def test_f32_bg1_z384_bit_exact_with_fp32_oracle(esn0_db):
    """BG1 Z=384, 16-QAM with the NR bit interleaver: the packed-syndrome path."""
    print(f"\n[f32 BG1 Z=384 16-QAM] Es/N0={esn0_db}", flush=True)
    count = 16
    ref, infos, llrs = config2_llrs(esn0_db, count)
    llr32 = llrs.astype(np.float32)
    full32 = np.array([ref.full_llr(l.astype(np.float64)) for l in llr32]).astype(np.float32)
    post_ref, it_ref = ref.decode_many_f32(full32, rule=O.NMS, alpha=0.75, layered=True, max_iterations=10,
                                           threads=8)
    code = F.NrLdpc(8448, 25344, 1, 0, 4).set_decoder("normalized_min_sum", 0.75, True, 10)
    bits, its, post = code.decode_batch(cuda(llr32), want_post=True)
    assert np.array_equal(its.cpu().numpy(), it_ref)
    assert np.array_equal(post.cpu().numpy(), post_ref)
    assert np.array_equal(bits.cpu().numpy(), (post_ref[:, :8448] < 0).astype(np.uint8))


@pytest.mark.parametrize("esn0_db", [4.0, 4.5])
def test_f32_bg1_z384_offset_min_sum_bit_exact_with_fp32_oracle(esn0_db):
    """BG1 Z=384, 16-QAM, offset min-sum (beta 0.5): the two-wave-set kernel's general
    instantiation, which reads the code's (a, b) = (1, beta) in every layer, from the launch's
    arguments. Posteriors, decisions and iteration counts equal the FP32 oracle's."""
    print(f"\n[f32 BG1 Z=384 16-QAM OMS 0.5] Es/N0={esn0_db}", flush=True)
    count = 16
    ref, infos, llrs = config2_llrs(esn0_db, count)
    llr32 = llrs.astype(np.float32)
    full32 = np.array([ref.full_llr(l.astype(np.float64)) for l in llr32]).astype(np.float32)
    post_ref, it_ref = ref.decode_many_f32(full32, rule=O.OMS, alpha=0.5, layered=True, max_iterations=10,
                                           threads=8)
    assert it_ref.min() < 10  # the syndrome stop ends some frames early (at 4.0 dB some reach the limit)
    code = F.NrLdpc(8448, 25344, 1, 0, 4).set_decoder("offset_min_sum", 0.5, True, 10)
    bits, its, post = code.decode_batch(cuda(llr32), want_post=True)
    assert np.array_equal(its.cpu().numpy(), it_ref)
    assert np.array_equal(post.cpu().numpy(), post_ref)
    assert np.array_equal(bits.cpu().numpy(), (post_ref[:, :8448] < 0).astype(np.uint8))


@pytest.mark.parametrize("esn0_db", [4.0, 4.5, 5.0, 6.0, 7.0])
def test_f32_against_the_f64_reference_with_recorded_ties(esn0_db):
    """FP32 messages against the reference's f64 decoder on the same LLRs (BG1 Z=384,
    16-QAM, layered NMS 0.75, 10 iterations; config 2's Es/N0 grid {4, 5, 6, 7} dB and the
    4.5 dB point near FER 1e-3), 64 frames a point.

    Ties are recorded by the oracle's side-by-side run of the two arithmetics
    (or_trace_f32_f64): a frame has a recorded tie when a check-node input or a posterior
    takes another sign in FP32 than in f64, the only way their decisions can part; its f64
    magnitude is recorded too (below TIE on frames where one input sits on zero, up to
    ~1e-3 where the FP32 rounding has grown over several iterations in the waterfall).
    Every frame: the same iteration count and the same frame-error verdict. Frames without
    a recorded tie: identical decisions and bit-error counts. Frames with one: decisions
    may differ only on frames that fail in both. Posteriors: within 1e-4 relative (of
    max(|f64|, 1)) on every frame at Es/N0 >= 5 dB (measured <= 4.9e-5), within 2e-3 on
    converged frames below (measured <= 1.2e-3: the rounding difference grows about
    five-fold an iteration over the first four, then saturates)."""
    print(f"\n[f32 vs f64] BG1 Z=384 16-QAM {esn0_db} dB, recorded ties", flush=True)
    count = 64
    ref, infos, llrs = config2_llrs(esn0_db, count)
    llr32 = llrs.astype(np.float32)
    fulls = np.array([ref.full_llr(l.astype(np.float64)) for l in llr32])
    post64, it64 = ref.decode_many(fulls, rule=O.NMS, alpha=0.75, layered=True, max_iterations=10, threads=8)
    trace = [ref.trace_f32_f64(f, rule=O.NMS, alpha=0.75, max_iterations=10) for f in fulls]
    tied = np.array([t["sign_events"] > 0 for t in trace])
    below = sum(t["sign_events"] > 0 and t["first_sign_abs"] < TIE for t in trace)
    code = F.NrLdpc(8448, 25344, 1, 0, 4).set_decoder("normalized_min_sum", 0.75, True, 10)
    bits, its, post = code.decode_batch(cuda(llr32), want_post=True)
    bits, its, post = bits.cpu().numpy(), its.cpu().numpy(), post.cpu().numpy().astype(np.float64)
    bits64 = (post64[:, :8448] < 0).astype(np.uint8)
    fail32, fail64 = (bits != infos).any(1), (bits64 != infos).any(1)
    print(f"  frames {count}, converged {int((it64 < 10).sum())}, recorded ties {int(tied.sum())} "
          f"({below} first below {TIE:g})", flush=True)
    assert np.array_equal(its, it64)
    assert np.array_equal(its, np.array([int(t["it32"]) for t in trace]))
    assert np.array_equal(fail32, fail64)
    same = (bits == bits64).all(1)
    assert same[~tied].all()
    assert np.array_equal((bits != infos).sum(1)[~tied], (bits64 != infos).sum(1)[~tied])
    assert (same | (fail32 & fail64)).all()
    rel = np.abs(post - post64) / np.maximum(np.abs(post64), 1.0)
    conv = it64 < 10
    if esn0_db >= 5.0:
        assert rel.max() < 1e-4, rel.max()
    assert rel[conv].max() < 2e-3, rel[conv].max()


@pytest.mark.parametrize("limit", [1, 2, 3])
def test_f32_posteriors_within_1e4_of_f64_over_the_first_iterations(limit):
    """The FP32 kernel against the reference's f64 decoder, both stopped after 1, 2 or 3
    layered iterations (BG1 Z=384, 16-QAM, Es/N0 4.5 dB, 8 frames): every posterior within
    1e-4 relative (against max(|f64|, 1)) and every decision identical. Measured on these
    frames: 1.0e-6, 7.3e-6 and 3.1e-5; the rounding differences then grow with the
    iterations (1.4e-4 after 5 on the same frames, up to ~1.3e-3 on rare frames after 10),
    which the test above bounds at convergence."""
    print(f"\n[f32 vs f64] {limit} iterations", flush=True)
    ref, infos, llrs = config2_llrs(4.5, 8)
    llr32 = llrs.astype(np.float32)
    full = np.array([ref.full_llr(l.astype(np.float64)) for l in llr32])
    post64, it64 = ref.decode_many(full, rule=O.NMS, alpha=0.75, layered=True, max_iterations=limit, threads=8)
    code = F.NrLdpc(8448, 25344, 1, 0, 4).set_decoder("normalized_min_sum", 0.75, True, limit)
    bits, its, post = code.decode_batch(cuda(llr32), want_post=True)
    post = post.cpu().numpy().astype(np.float64)
    assert np.array_equal(its.cpu().numpy(), it64)
    assert np.array_equal(post < 0, post64 < 0)
    assert np.max(np.abs(post - post64) / np.maximum(np.abs(post64), 1.0)) < 1e-4


def test_host_entry_points_from_eight_threads_on_one_handle():
    """The runner calls decode_llr from rayon workers through &self (runner.rs:435-446): eight
    host threads call cf_decode_llr, cf_decode_block and cf_app_into on one handle at once
    (ctypes releases the interpreter lock during each call), every result bit-identical to
    the oracle's f64 decoder."""
    import threading
    print("\n[threads] 8 host threads, one handle, f64 host entry points", flush=True)
    ref, cws, full = config1_llrs(48, ebn0_db=1.25, name="threads")
    post_ref, it_ref = ref.decode_many(full, rule=O.NMS, alpha=0.75, layered=True, max_iterations=20, threads=8)
    code = F.NrLdpc.lifted(2, 52).set_decoder("normalized_min_sum", 0.75, True, 20)
    results, errors = {}, []

    def work(t):
        try:
            for i in range(t, 48, 8):
                if i % 3 == 0:
                    post, it = code.decode_llr(full[i])
                    results[i] = ("llr", post, it)
                elif i % 3 == 1:
                    results[i] = ("block", code.decode_block(full[i, 104:]), None)
                else:
                    app, it = code.app(full[i, 104:])
                    results[i] = ("app", app, it)
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    threads = [threading.Thread(target=work, args=(t,)) for t in range(8)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    assert len(results) == 48
    for i, (kind, value, it) in results.items():
        if kind == "llr":
            assert np.array_equal(value, post_ref[i]) and it == it_ref[i]
        elif kind == "block":
            assert np.array_equal(value, (post_ref[i, :520] < 0).astype(np.uint8))
        else:
            assert np.array_equal(value, post_ref[i, :520]) and it == it_ref[i]
