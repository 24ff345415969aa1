"""The FP16 storage mode, codewords packed per workgroup, and batches of many codes in one
launch, through the C ABI.

- FP16 storage: bit-identical to the oracle's binary16 restatement (posteriors, iterations,
  decisions) on BG2 Z=52 (row-by-row syndrome) and BG1 Z=384 16-QAM (packed syndrome),
  for the normalized and the offset min-sum; its decisions track the FP32 decoder.
- Small Z: the single-code path decodes floor(lanes / Z) codewords per workgroup:
  bit-identical to the FP32 and FP16 oracles at Z = 2, 16, 24, 32, 36, 352.
- Batches of many codes (mixed base graphs, lifting sizes, BPSK/QPSK, per-trial Eb/N0 and
  per-codeword iteration limits): every codeword's decisions, iteration count and syndrome
  verdict equal the oracle's (FP32 and FP16), on the 20 000-codeword workload too; the
  results do not depend on the workgroup width the autotuner picks, and a replayed HIP
  graph gives the eager results on every replay.
"""
import ctypes as C

import numpy as np
import pytest

import comfec as F
import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

LIFTS = [16, 18, 20, 22, 24, 26, 28, 30, 32, 36, 40, 44, 48, 52, 56, 60, 64, 72, 80, 88, 96, 104, 112, 120, 128,
         144, 160, 176, 192, 208, 224, 240, 256, 288, 320, 352, 384]
THREADS = 16


def lifted_llrs(code, count, lo=1.0, hi=2.0, modulation=F.BPSK, seed=3, point=0):
    """Device-generated LLRs of a lifted code (per-trial Eb/N0 uniform in [lo, hi]), and the
    full N-value LLRs the oracle decodes (the first 2Z columns are not sent: 0.0)."""
    llr, _ = F.generate_training_batch(code, count, modulation, lo, hi, snr_kind=F.EBN0_DB, seed=seed, point=point)
    host = llr.cpu().numpy()
    z = code.dims.z
    full = np.concatenate([np.zeros((count, 2 * z), np.float32), host], axis=1)
    return llr, full


RULES = [("nms", O.NMS, "normalized_min_sum", 0.75), ("oms", O.OMS, "offset_min_sum", 0.15)]


@pytest.mark.parametrize("rule", RULES, ids=[r[0] for r in RULES])
def test_f16_bg2_z52_bit_exact_with_the_half_oracle(rule):
    print(f"\n[f16 BG2 Z=52] rule={rule[0]}", flush=True)
    _, orule, name, alpha = rule
    code = F.NrLdpc.lifted(2, 52).set_decoder(name, alpha, True, 20)
    ref = O.Code.lifted(2, 52)
    llr, full = lifted_llrs(code, 96, 0.8, 1.6)
    bits, its, post = code.decode_batch(llr, want_post=True, half=True)
    post_ref, it_ref = ref.decode_many_f16(full, rule=orule, alpha=alpha, max_iterations=20, threads=THREADS)
    assert np.array_equal(its.cpu().numpy(), it_ref)
    assert np.array_equal(post.cpu().numpy().view(np.uint32), post_ref.view(np.uint32))
    assert np.array_equal(bits.cpu().numpy(), (post_ref[:, :520] < 0).astype(np.uint8))
    assert len(set(it_ref.tolist())) >= 4


@pytest.mark.parametrize("esn0_db", [4.0, 4.4, 6.0])
def test_f16_bg1_z384_bit_exact_with_the_half_oracle(esn0_db):
    print(f"\n[f16 BG1 Z=384 16-QAM] Es/N0={esn0_db}", flush=True)
    code = F.NrLdpc(8448, 25344, 1, 0, 4).set_decoder("normalized_min_sum", 0.75, True, 10)
    ref = O.Code.nr(8448, 25344, 1, 0, 4)
    llr, _ = F.generate_batch(code, 24, F.QAM16, esn0_db, snr_kind=F.ESN0_DB, seed=8)
    host = llr.cpu().numpy()
    full = np.array([ref.full_llr(l.astype(np.float64)) for l in host]).astype(np.float32)
    bits, its, post = code.decode_batch(llr, want_post=True, half=True)
    post_ref, it_ref = ref.decode_many_f16(full, rule=O.NMS, alpha=0.75, max_iterations=10, threads=THREADS)
    assert np.array_equal(its.cpu().numpy(), it_ref)
    assert np.array_equal(post.cpu().numpy().view(np.uint32), post_ref.view(np.uint32))
    assert np.array_equal(bits.cpu().numpy(), (post_ref[:, :8448] < 0).astype(np.uint8))


def test_f16_decisions_track_fp32():
    """FP16 storage against FP32 messages on 4096 codewords of BG1 Z=384 16-QAM at 4.4 dB:
    the frame error counts agree within their statistical spread and the mean iteration
    counts within 2 % (FP16 rounds each stored posterior to 11 significant bits)."""
    print("\n[f16 vs f32] BG1 Z=384 16-QAM 4.4 dB", flush=True)
    code = F.NrLdpc(8448, 25344, 1, 0, 4).set_decoder("normalized_min_sum", 0.75, True, 10)
    llr, info = F.generate_batch(code, 4096, F.QAM16, 4.4, snr_kind=F.ESN0_DB, seed=9)
    truth = torch.from_numpy(F.unpack_info(info, 8448)).cuda()
    b32, i32, _ = code.decode_batch(llr)
    b16, i16, _ = code.decode_batch(llr, half=True)
    fe32 = int((b32 != truth).any(1).sum())
    fe16 = int((b16 != truth).any(1).sum())
    lo, hi = F.wilson(fe32, 4096, 0.999)
    assert lo * 4096 - 3 <= fe16 <= hi * 4096 + 3
    assert abs(float(i16.float().mean()) / float(i32.float().mean()) - 1.0) < 0.02
    assert float((b32 == b16).all(1).float().mean()) > 0.98


@pytest.mark.parametrize("bg,z", [(1, 2), (2, 16), (1, 24), (2, 32), (1, 36), (2, 352), (2, 384), (1, 384)])
def test_codewords_packed_per_workgroup_bit_exact(bg, z):
    """floor(64k / Z) codewords per workgroup (Z = 2: 32 of them) decode as the oracle; with
    one codeword per workgroup the lanes beyond Z idle (Z = 36: 28 of 64, Z = 352: 32 of 384);
    Z = 384 runs the kernels with Z and the shifts compiled in, for both base graphs."""
    print(f"\n[packed workgroups] BG{bg} Z={z}", flush=True)
    code = F.NrLdpc.lifted(bg, z).set_decoder("normalized_min_sum", 0.75, True, 15)
    ref = O.Code.lifted(bg, z)
    count = 203  # not a multiple of the group: the last workgroup is partly empty
    llr, full = lifted_llrs(code, count, 0.5, 2.5)
    k = code.dims.k_info
    bits, its, post = code.decode_batch(llr, want_post=True)
    post_ref, it_ref = ref.decode_many_f32(full, rule=O.NMS, alpha=0.75, layered=True, max_iterations=15,
                                           threads=THREADS)
    assert np.array_equal(its.cpu().numpy(), it_ref)
    assert np.array_equal(post.cpu().numpy(), post_ref)
    assert np.array_equal(bits.cpu().numpy(), (post_ref[:, :k] < 0).astype(np.uint8))
    bits, its, post = code.decode_batch(llr, want_post=True, half=True)
    post_ref, it_ref = ref.decode_many_f16(full, rule=O.NMS, alpha=0.75, max_iterations=15, threads=THREADS)
    assert np.array_equal(its.cpu().numpy(), it_ref)
    assert np.array_equal(post.cpu().numpy().view(np.uint32), post_ref.view(np.uint32))


# ---- batches of many codes ---------------------------------------------------------------------

_codes = {}


def gpu_code(bg, z):
    if (bg, z) not in _codes:
        _codes[(bg, z)] = (F.NrLdpc.lifted(bg, z).set_decoder("normalized_min_sum", 0.75, True, 25),
                           O.Code.lifted(bg, z))
    return _codes[(bg, z)]


def uneven_workload(count, seed, lifts=LIFTS):
    """Descriptors drawn from a fixed seed: base graph, Z, BPSK or QPSK, iteration limit in
    5..25; LLRs from the device generator at per-trial Eb/N0 uniform in [0, 3] dB."""
    rng = np.random.default_rng(seed)
    bgs = rng.integers(1, 3, count)
    zs = np.array(lifts)[rng.integers(0, len(lifts), count)]
    mods = np.where(rng.integers(0, 2, count) == 0, F.BPSK, F.QPSK)
    limits = rng.integers(5, 26, count).astype(np.uint32)
    groups = {}
    for i in range(count):
        groups.setdefault((int(bgs[i]), int(zs[i]), int(mods[i])), []).append(i)
    per = [None] * count
    fulls = [None] * count
    for gi, ((bg, z, mod), idx) in enumerate(sorted(groups.items())):
        code, _ = gpu_code(bg, z)
        llr, full = lifted_llrs(code, len(idx), 0.0, 3.0, modulation=mod, seed=seed, point=gi)
        for j, i in enumerate(idx):
            per[i] = llr[j]
            fulls[i] = full[j]
    codes = [gpu_code(int(b), int(z))[0] for b, z in zip(bgs, zs)]
    offsets = np.zeros(count + 1, np.uint64)
    offsets[1:] = np.cumsum([c.dims.e for c in codes])
    flat = torch.cat(per).contiguous()
    return codes, (bgs, zs, limits), flat, offsets, fulls


def oracle_uneven(desc, fulls, half, idx=None):
    """Per codeword: decisions, iterations and syndrome verdict of the oracle (FP32 or FP16)."""
    bgs, zs, limits = desc
    idx = range(len(bgs)) if idx is None else idx
    groups = {}
    for i in idx:
        groups.setdefault((int(bgs[i]), int(zs[i])), []).append(i)
    out = {}
    for (bg, z), members in groups.items():
        _, ref = gpu_code(bg, z)
        full = np.array([fulls[i] for i in members])
        lim = limits[members].astype(np.int32)
        if half:
            post, its = ref.decode_many_f16(full, rule=O.NMS, alpha=0.75, max_iterations=lim, threads=THREADS)
        else:
            post = np.zeros_like(full)
            its = np.zeros(len(members), np.int32)
            for m in np.unique(lim):
                sel = np.nonzero(lim == m)[0]
                p, t = ref.decode_many_f32(full[sel], rule=O.NMS, alpha=0.75, layered=True, max_iterations=int(m),
                                           threads=THREADS)
                post[sel], its[sel] = p, t
        k = ref.k_info
        for j, i in enumerate(members):
            hard = (post[j] < 0).astype(np.uint8)
            out[i] = (hard[:k], int(its[j]), int(ref.is_codeword(hard)))
    return out


def split_bits(codes, bits):
    b = bits.cpu().numpy()
    out, at = [], 0
    for c in codes:
        out.append(b[at:at + c.dims.k_info])
        at += c.dims.k_info
    return out


def check_against_oracle(codes, desc, fulls, result, half, idx=None):
    bits, its, ok = result
    per = split_bits(codes, bits)
    its, ok = its.cpu().numpy(), ok.cpu().numpy()
    ref = oracle_uneven(desc, fulls, half, idx)
    for i, (rb, rit, rok) in ref.items():
        assert np.array_equal(per[i], rb), f"codeword {i}: decisions"
        assert its[i] == rit, f"codeword {i}: {its[i]} iterations, oracle {rit}"
        assert ok[i] == rok, f"codeword {i}: syndrome verdict"
    return ref


@pytest.mark.parametrize("half", [False, True], ids=["f32", "f16"])
def test_uneven_batch_bit_exact_per_codeword(half):
    print(f"\n[uneven batch] 600 codewords half={half}", flush=True)
    codes, desc, flat, offsets, fulls = uneven_workload(600, seed=11)
    plan = F.BatchPlan(codes, offsets, desc[2], half=half)
    d = plan.dims
    assert d.codewords == 600 and d.classes > 40 and d.llr_total == offsets[-1]
    result = plan.run(flat)
    ref = check_against_oracle(codes, desc, fulls, result, half)
    its = np.array([ref[i][1] for i in range(600)])
    assert (its == desc[2]).any() and (its < desc[2]).any() and (its == 0).sum() < 600
    # the one-shot entry point gives the same
    b2, i2, o2 = F.decode_batch_uneven(codes, flat, offsets, desc[2], half=half)
    assert torch.equal(b2, result[0]) and torch.equal(i2, result[1]) and torch.equal(o2, result[2])


@pytest.mark.parametrize("half", [True, False], ids=["f16", "f32"])
def test_uneven_20000_codewords_per_codeword(half):
    """The uneven workload at full size: 20 000 codewords, one persistent launch, FP16 or
    FP32 storage, every codeword compared with the oracle's binary16 or FP32 restatement."""
    print(f"\n[uneven batch] 20000 codewords half={half}", flush=True)
    codes, desc, flat, offsets, fulls = uneven_workload(20000, seed=7)
    plan = F.BatchPlan(codes, offsets, desc[2], half=half)
    result = plan.run(flat)
    torch.cuda.synchronize()
    print(f"  decoded; grid {plan.dims.grid} x {plan.dims.threads} lanes, {plan.dims.tasks} tasks", flush=True)
    check_against_oracle(codes, desc, fulls, result, half)


def test_width_autotuner_and_graph_replays():
    """Z <= 120: widths 128..384 are admissible; every width gives the same results, the
    tuner keeps one of them, and a HIP graph replays the eager decode exactly."""
    print("\n[uneven batch] widths, tuner, graph", flush=True)
    small = [z for z in LIFTS if z <= 120]
    codes, desc, flat, offsets, fulls = uneven_workload(900, seed=5, lifts=small)
    plan = F.BatchPlan(codes, offsets, desc[2])
    eager = [t.clone() for t in plan.run(flat)]
    for w in (128, 192, 256, 320, 384):
        plan.set_threads(w)
        got = plan.run(flat)
        assert all(torch.equal(a, b) for a, b in zip(got, eager)), f"width {w}"
    plan.configure(per_code_width=True)
    got = plan.run(flat)
    assert all(torch.equal(a, b) for a, b in zip(got, eager)), "per-code widths"
    ms, got = plan.tune(flat, repeats=2)
    print(f"  {len(plan.trials)} trials, chosen {plan.config!r}: {ms:.3f} ms", flush=True)
    assert ms > 0 and len(plan.trials) >= 6
    assert any(t.ms == ms for t in plan.trials)
    assert plan.config.msg_type == F.MSG_F32
    assert all(torch.equal(a, b) for a, b in zip(got, eager))
    out = plan.outputs()
    for rep in range(3):
        for t in out:
            t.fill_(0x5A if t.dtype == torch.uint8 else -1)
        torch.cuda.synchronize()
        plan.run(flat, out=out, graph=True)
        torch.cuda.synchronize()
        assert all(torch.equal(a, b) for a, b in zip(out, eager)), f"replay {rep}"
        _ = (flat * 2.0).sum().item()  # eager work between the replays
    check_against_oracle(codes, desc, fulls, eager, False, idx=range(0, 900, 7))


def test_plan_configurations_bit_identical():
    """Z <= 64: one launch at every width from 64 to 384, per-code widths (several launches in
    one run) and one resident workgroup per CU all give the same decisions, iterations and
    verdicts, equal to the oracle's."""
    print("\n[uneven batch] configurations, Z <= 64", flush=True)
    small = [z for z in LIFTS if z <= 64]
    codes, desc, flat, offsets, fulls = uneven_workload(700, seed=9, lifts=small)
    plan = F.BatchPlan(codes, offsets, desc[2])
    assert plan.dims.threads == 64 and plan.dims.segments == 1
    eager = [t.clone() for t in plan.run(flat)]
    for w in (64, 128, 192, 256, 320, 384):
        for per_cu in (0, 1):
            plan.configure(threads=w, per_code_width=False, grid_per_cu=per_cu)
            got = plan.run(flat)
            assert all(torch.equal(a, b) for a, b in zip(got, eager)), f"width {w}, {per_cu} per CU"
    plan.configure(per_code_width=True, grid_per_cu=0)
    segs = plan.segments()
    print("  per-code widths: " + ", ".join(f"{s.threads} lanes x {s.classes} codes" for s in segs), flush=True)
    assert sum(s.codewords for s in segs) == 700 and sum(s.tasks for s in segs) == plan.dims.tasks
    assert [s.threads for s in segs] == sorted({s.threads for s in segs}, reverse=True)
    got = plan.run(flat)
    assert all(torch.equal(a, b) for a, b in zip(got, eager)), "per-code widths"
    # the launches side by side, eagerly and through a replayed graph
    plan.configure(overlap=True)
    assert plan.config.overlap == 1 and plan.dims.segments == len(segs)
    got = plan.run(flat)
    assert all(torch.equal(a, b) for a, b in zip(got, eager)), "overlapped launches"
    out = plan.outputs()
    for rep in range(3):
        for t in out:
            t.fill_(0x3C if t.dtype == torch.uint8 else -1)
        torch.cuda.synchronize()
        plan.run(flat, out=out, graph=True)
        torch.cuda.synchronize()
        assert all(torch.equal(a, b) for a, b in zip(out, eager)), f"overlapped replay {rep}"
        _ = (flat * 3.0).sum().item()
    check_against_oracle(codes, desc, fulls, eager, False, idx=range(0, 700, 5))


def test_tuner_on_20000_codewords():
    """The 20 000-codeword plan with FP16 storage: the tuner times at least two candidates
    and keeps results bit-identical; allowed to change the storage, its results equal a plan
    made with the storage it chose."""
    print("\n[uneven batch] tuner, 20000 codewords", flush=True)
    codes, desc, flat, offsets, fulls = uneven_workload(20000, seed=7)
    plan = F.BatchPlan(codes, offsets, desc[2], half=True)
    eager = [t.clone() for t in plan.run(flat)]
    ms, got = plan.tune(flat, repeats=2)
    for t in plan.trials:
        print(f"  trial {t.config!r}: {t.segments} launch(es), {t.ms:.3f} ms", flush=True)
    print(f"  chosen {plan.config!r}: {ms:.3f} ms", flush=True)
    assert len(plan.trials) >= 2 and ms == min(t.ms for t in plan.trials)
    assert {t.config.overlap for t in plan.trials if t.segments > 1} == {0, 1}
    assert plan.config.msg_type == F.MSG_F16
    assert all(torch.equal(a, b) for a, b in zip(got, eager))
    assert all(torch.equal(a, b) for a, b in zip(plan.run(flat), eager))
    ms2, got = plan.tune(flat, repeats=2, storage=True)
    chosen = plan.config
    print(f"  with storage: {len(plan.trials)} trials, chosen {chosen!r}: {ms2:.3f} ms", flush=True)
    assert {t.config.msg_type for t in plan.trials} == {F.MSG_F32, F.MSG_F16}
    half = chosen.msg_type == F.MSG_F16
    direct = F.BatchPlan(codes, offsets, desc[2], half=half).run(flat)
    assert all(torch.equal(a, b) for a, b in zip(got, direct))
    check_against_oracle(codes, desc, fulls, got, half, idx=range(0, 20000, 41))


def test_plan_errors():
    print("\n[uneven batch] errors", flush=True)
    a, _ = gpu_code(1, 16)
    b, _ = gpu_code(2, 24)
    offs = np.array([0, a.dims.e, a.dims.e + b.dims.e + 1], np.uint64)
    with pytest.raises(F.ComfecError) as e:
        F.BatchPlan([a, b], offs)
    assert e.value.kind == "Shape"
    spa = F.NrLdpc.lifted(1, 16).set_decoder("sum_product", 0.0, True, 10)
    with pytest.raises(F.ComfecError) as e:
        F.BatchPlan([spa], np.array([0, spa.dims.e], np.uint64))
    assert e.value.kind == "Config"
    plan = F.BatchPlan([a, b], np.array([0, a.dims.e, a.dims.e + b.dims.e], np.uint64))
    with pytest.raises(F.ComfecError):
        plan.set_threads(96)
    with pytest.raises(F.ComfecError) as e:
        plan.configure(threads=0, per_code_width=False)
    assert e.value.kind == "Config"
    plan.configure(per_code_width=True)
    with pytest.raises(F.ComfecError) as e:
        plan.configure(per_code_width=2)
    assert e.value.kind == "Config"
    c = plan.config
    c.msg_type = F.MSG_F64
    assert F.lib().cf_batch_plan_configure(plan._h, C.byref(c)) == F.CF_CONFIG


@pytest.mark.parametrize("k,e,rv,qm", [(500, 1200, 2, 4), (60, 1000, 1, 2), (1000, 1500, 3, 6),
                                       (8448, 12672, 2, 6)])
def test_rate_matching_variants_on_the_fast_kernel(k, e, rv, qm):
    """The layered kernel's LLR placement on NR codes with fillers, repetitions (E > Ncb),
    redundancy versions 1-3 and Qm interleaving (nr/ldpc.rs:136-197), LLRs of the device
    generator: FP32 and FP16 storage bit-identical to their oracles, whose full LLRs are
    summed in FP32 in output order as the kernel sums them."""
    print(f"\n[rate matching] k={k} e={e} rv={rv} qm={qm}", flush=True)
    code = F.NrLdpc(k, e, None, rv, qm).set_decoder("normalized_min_sum", 0.75, True, 12)
    ref = O.Code.nr(k, e, 0, rv, qm)
    mod = {1: F.BPSK, 2: F.QPSK, 4: F.QAM16, 6: F.QAM64}[qm]
    llr, _ = F.generate_batch(code, 48, mod, 2.0 + qm, snr_kind=F.EBN0_DB, seed=12)
    host = llr.cpu().numpy()
    pos = ref.bit_map()
    full = np.zeros((48, ref.n), np.float32)
    for t in range(e):
        full[:, pos[t]] += host[:, t]  # float32 sums in output order
    full[:, ref.k_info:ref.k] = 30.0
    for half in (False, True):
        bits, its, post = code.decode_batch(llr, want_post=True, half=half)
        if half:
            pr, ir = ref.decode_many_f16(full, rule=O.NMS, alpha=0.75, max_iterations=12, threads=THREADS)
        else:
            pr, ir = ref.decode_many_f32(full, rule=O.NMS, alpha=0.75, layered=True, max_iterations=12,
                                         threads=THREADS)
        assert np.array_equal(its.cpu().numpy(), ir), half
        assert np.array_equal(post.cpu().numpy().view(np.uint32), pr.view(np.uint32)), half
        assert np.array_equal(bits.cpu().numpy(), (pr[:, :k] < 0).astype(np.uint8)), half
