// The C ABI of the learned offset/scale min-sum (include/comfec_gpu.h): training batches
// from the device trial generator, the loss and its gradient through the unrolled
// iterations, Adam, a training loop, and decoding with the learned parameters, codewords
// whose syndrome still fails being decoded again by the plain min-sum in the same launch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "kernels.hpp"
#include "nr_code.hpp"

using namespace comfec;
using namespace comfec::detail;

namespace {

enum { kTrainSlot = 1 };

struct TrainBuffers {
    float* ckpt;
    float* gpost;
    float* cbar;
    float* loss_cw;
    float* partial;
};

TrainBuffers train_buffers(WorkspaceSlot& slot, const NrCode& c, uint32_t T, size_t count) {
    const size_t edges = size_t(c.edges_b) * c.z;
    const size_t np = size_t(T) * c.rows_b * 2;
    const size_t n_ckpt = count * T * edges, n_gpost = count * T * c.n, n_cbar = count * edges, n_loss = count,
                 n_part = count * np;
    float* base = static_cast<float*>(slot.ws.get((n_ckpt + n_gpost + n_cbar + n_loss + n_part) * sizeof(float)));
    if (!base) fail(Kind::Device, "out of device memory for the training buffers");
    TrainBuffers b;
    b.ckpt = base;
    b.gpost = b.ckpt + n_ckpt;
    b.cbar = b.gpost + n_gpost;
    b.loss_cw = b.cbar + n_cbar;
    b.partial = b.loss_cw + n_loss;
    return b;
}

void loss_grad(const cf_code* h, int device, uint32_t T, const float* params, const float* llr, const uint32_t* cw,
               size_t count, float* grad, float* loss, hipStream_t s) {
    const NrCode& c = h->code;
    if (T < 1) fail(Kind::Config, "the learned decoder needs at least one iteration");
    WorkspaceSlot& slot = workspace(device, s, kTrainSlot);
    std::lock_guard<std::mutex> l(slot.mu);
    TrainBuffers b = train_buffers(slot, c, T, count);
    LearnArgs a{};
    a.code = dev_code(c, device);
    a.T = int(T);
    a.params = params;
    a.llr = llr;
    a.cw = cw;
    a.count = int64_t(count);
    a.scale = float(1.0 / (double(T) * double(c.n) * double(count)));
    a.ckpt = b.ckpt;
    a.gpost = b.gpost;
    a.cbar = b.cbar;
    a.loss_out = b.loss_cw;
    a.grad_partial = b.partial;
    a.grad = grad;
    a.loss = loss;
    check_hip(launch_learn_forward(a, s), "learned forward launch");
    check_hip(launch_learn_backward(a, s), "learned backward launch");
}

}  // namespace

extern "C" {

cf_status cf_generate_training_batch(const cf_code* h, int modulation, int snr_kind, double snr_lo_db,
                                     double snr_hi_db, uint64_t seed, uint32_t point, uint64_t trial0, size_t count,
                                     float* d_llr, uint32_t* d_codeword, void* stream) {
    return guarded([&] {
        if (!h || !d_llr) fail(Kind::Config, "null handle or LLR buffer");
        if (!(snr_hi_db >= snr_lo_db)) fail(Kind::Config, "the SNR range must have lo <= hi");
        TrialGenArgs a = make_generator_args(h->code, current_device(), modulation, snr_kind, snr_lo_db, seed, point);
        a.trial0 = trial0;
        a.count = int64_t(count);
        a.llr = d_llr;
        a.info = nullptr;
        a.codeword = d_codeword;
        a.snr_lo = float(snr_lo_db);
        a.snr_hi = float(snr_hi_db);
        a.snr_kind = snr_kind;
        a.rate = float(double(h->code.k_info) / double(h->code.e));
        check_hip(launch_generate(a, static_cast<hipStream_t>(stream)), "generator launch");
    });
}

cf_status cf_learned_loss_grad(const cf_code* h, uint32_t iterations, const float* d_params, const float* d_llr,
                               const uint32_t* d_codeword, size_t count, float* d_grad, float* d_loss, void* stream) {
    return guarded([&] {
        if (!h || !d_params || !d_llr || !d_codeword || !d_grad) fail(Kind::Config, "null handle or buffer");
        if (count == 0) fail(Kind::Config, "an empty training batch");
        loss_grad(h, current_device(), iterations, d_params, d_llr, d_codeword, count, d_grad, d_loss,
                  static_cast<hipStream_t>(stream));
    });
}

cf_status cf_adam_step(float* d_params, const float* d_grad, float* d_m, float* d_v, size_t n, uint32_t step,
                       double lr, double beta1, double beta2, double eps, void* stream) {
    return guarded([&] {
        if (!d_params || !d_grad || !d_m || !d_v) fail(Kind::Config, "null buffer");
        if (step < 1) fail(Kind::Config, "Adam steps count from 1");
        check_hip(launch_adam(d_params, d_grad, d_m, d_v, int(n), float(lr), float(beta1), float(beta2), float(eps),
                              int(step), static_cast<hipStream_t>(stream)),
                  "adam launch");
    });
}

cf_status cf_learned_decode_batch(const cf_code* h, uint32_t iterations, const float* d_params, int fallback,
                                  const float* d_llr, size_t count, uint8_t* d_bits, uint32_t* d_iterations,
                                  uint8_t* d_ok, void* stream) {
    return guarded([&] {
        if (!h || !d_params || !d_llr) fail(Kind::Config, "null handle or buffer");
        if (iterations < 1) fail(Kind::Config, "the learned decoder needs at least one iteration");
        if (count == 0) return;
        if (!layered_supported(h->code.bg, int(h->code.z)))
            fail(Kind::Config, "the learned decoder needs a lifting size of at most 384");
        if (fallback) {
            const DecoderSettings s = h->settings();
            if (!s.layered || (s.rule != 1 && s.rule != 2))
                fail(Kind::Config, "the fallback runs the code's layered normalized or offset min-sum");
        }
        // one launch: the learned iterations, then (fallback) the plain min-sum of the
        // code's settings on the codewords whose syndrome still fails, in the same workgroup
        LayeredRun r;
        r.llr = d_llr;
        r.count = int64_t(count);
        r.bits = d_bits;
        r.iterations = d_iterations;
        r.ok = d_ok;
        r.params = d_params;
        r.learned_iterations = iterations;
        r.fallback = fallback != 0;
        run_layered(h, current_device(), r, static_cast<hipStream_t>(stream));
    });
}

cf_status cf_learned_train(const cf_code* h, uint32_t iterations, float* d_params, uint32_t steps, uint32_t batch,
                           double lr, int modulation, int snr_kind, double snr_lo_db, double snr_hi_db, uint64_t seed,
                           float* loss_history_host, void* stream) {
    return guarded([&] {
        if (!h || !d_params) fail(Kind::Config, "null handle or parameters");
        if (batch == 0 || steps == 0) fail(Kind::Config, "training needs steps and a batch");
        const int device = current_device();
        hipStream_t s = static_cast<hipStream_t>(stream);
        const NrCode& c = h->code;
        const size_t np = size_t(iterations) * c.rows_b * 2;
        const size_t words = (c.n + 31) / 32;
        float* buf = nullptr;
        // regions at multiples of 64 floats (256 bytes), so the u64 fills stay aligned
        auto up = [](size_t x) { return (x + 63) & ~size_t(63); };
        const size_t at_cw = up(size_t(batch) * c.e), at_grad = at_cw + up(size_t(batch) * words);
        const size_t at_m = at_grad + up(np), at_losses = at_m + up(2 * np);
        check_hip(hipMalloc(&buf, (at_losses + up(steps)) * sizeof(float)), "hipMalloc");
        std::unique_ptr<float, void (*)(float*)> guard(buf, [](float* p) { (void)hipFree(p); });
        float* llr = buf;
        uint32_t* cw = reinterpret_cast<uint32_t*>(buf + at_cw);
        float* grad = buf + at_grad;
        float* m = buf + at_m;
        float* v = m + np;
        float* losses = buf + at_losses;
        // the Adam moments m and v (2 np floats, contiguous) start at zero
        check_hip(launch_zero(reinterpret_cast<uint64_t*>(m), int(np), s), "zero launch");
        for (uint32_t step = 1; step <= steps; ++step) {
            if (cf_generate_training_batch(h, modulation, snr_kind, snr_lo_db, snr_hi_db, seed, 0,
                                           uint64_t(step - 1) * batch, batch, llr, cw, s) != CF_OK)
                fail(Kind::Config, g_error);
            loss_grad(h, device, iterations, d_params, llr, cw, batch, grad, losses + (step - 1), s);
            check_hip(launch_adam(d_params, grad, m, v, int(np), float(lr), 0.9f, 0.999f, 1e-8f, int(step), s),
                      "adam launch");
        }
        if (loss_history_host)
            check_hip(hipMemcpyAsync(loss_history_host, losses, steps * sizeof(float), hipMemcpyDeviceToHost, s),
                      "copy");
        check_hip(hipStreamSynchronize(s), "training");
    });
}

}  // extern "C"
