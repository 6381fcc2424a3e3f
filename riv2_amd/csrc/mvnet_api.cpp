// Host driver of the learned MV-proposal network (mvnet.hip): parameters, training (forward,
// backward and Adam through the layer-wise kernels) and proposals for the encoder's motion search.
#include <cmath>
#include <algorithm>
#include <mutex>
#include <vector>
#include "host_common.hpp"
#include "mvnet.hpp"
#include "mvnet_host.hpp"

using namespace rivhost;
using namespace rivnet;

namespace {

struct Rng {   // splitmix64
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

uint16_t host_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// 1 where a parameter trains, 0 for the fixed zero padding (conv1 rows k >= 36, head columns >= 9)
std::vector<uint8_t> train_mask() {
    std::vector<uint8_t> m(NPARAM, 1);
    for (int k = 9 * rivnet::C0; k < rivnet::K1; k++)
        for (int n = 0; n < rivnet::F; n++) m[rivnet::OFF_W1 + k * rivnet::F + n] = 0;
    for (int k = 0; k < rivnet::FC_H; k++)
        for (int n = rivnet::OUT_USED; n < rivnet::OUT; n++) m[rivnet::OFF_W5 + k * rivnet::OUT + n] = 0;
    for (int n = rivnet::OUT_USED; n < rivnet::OUT; n++) m[rivnet::OFF_B5 + n] = 0;
    return m;
}

}  // namespace

struct riv_mvnet {
    int dev = 0;
    hipStream_t st = nullptr;
    DevBuf p32, p16, adam_m, adam_v, grad, mask;
    int step = 0;
    DevBuf srcs;
    int nsrc = 0;
    int bcap = 0;
    DevBuf samp, tgt, x0, c1, a1, c2, a2, c3, a3, a4, z1, z2, z3, z4, y, dy, dyb, da4, dz4, da3, dz3, dc3, dz2, dc2, dz1, part, lossb, prop;
    size_t part_n = 0;
    // pinned staging of a batch's samples and targets: the caller's arrays may go away as soon as
    // a step returns, and the previous step's upload must have left the buffer before it is refilled
    HostBuf stage;
    hipEvent_t staged = nullptr;

    // the fragment-order copy of p16 for the fused inference kernel, remade (on st, waited for)
    // the first time the weights are asked for after p16 changed
    mutable DevBuf pk;
    mutable std::mutex pk_mu;
    mutable bool pk_stale = true;

    // inferences queued on other streams (the encoders'): the last event recorded after one on each
    // such stream. Writes of p16, p32 or pk are queued on st behind all of them.
    mutable std::mutex use_mu;
    mutable std::vector<std::pair<hipStream_t, hipEvent_t>> uses;

    hipError_t note_use(hipStream_t s) const {
        std::lock_guard<std::mutex> lk(use_mu);
        hipEvent_t ev = nullptr;
        for (auto& u : uses) if (u.first == s) ev = u.second;
        if (!ev) {
            hipError_t r = hipEventCreateWithFlags(&ev, hipEventDisableTiming | hipEventDisableSystemFence);
            if (r != hipSuccess) return r;
            uses.push_back({s, ev});
        }
        return hipEventRecord(ev, s);
    }
    // st waits for every noted inference (an event whose stream has gone away has completed)
    hipError_t after_uses() const {
        std::lock_guard<std::mutex> lk(use_mu);
        for (auto& u : uses) {
            hipError_t r = hipStreamWaitEvent(st, u.second, 0);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    }

    // pk is null if the copy could not be made; infer then fails
    rivnet::Weights weights() const {
        const uint16_t* w = p16.as<uint16_t>();
        const float* b = p32.as<float>();
        rivnet::Weights wt{w + rivnet::OFF_W1, w + rivnet::OFF_W2, w + rivnet::OFF_W3, w + rivnet::OFF_W4, w + rivnet::OFF_W5,
                           b + rivnet::OFF_B1, b + rivnet::OFF_B2, b + rivnet::OFF_B3, b + rivnet::OFF_B4, b + rivnet::OFF_B5, nullptr};
        std::lock_guard<std::mutex> lk(pk_mu);
        if (pk_stale && after_uses() == hipSuccess && rivnet::pack_fragments(st, wt, pk.as<uint2>()) == hipSuccess &&
            hipStreamSynchronize(st) == hipSuccess)
            pk_stale = false;
        if (!pk_stale) wt.pk = pk.as<uint2>();
        return wt;
    }
    void weights_changed() {
        std::lock_guard<std::mutex> lk(pk_mu);
        pk_stale = true;
    }
};

namespace rivnet {
Weights net_weights(const riv_mvnet* n) { return n->weights(); }
hipError_t note_use(const riv_mvnet* n, hipStream_t st) { return n->note_use(st); }
int net_device(const riv_mvnet* n) { return n->dev; }
}  // namespace rivnet

namespace {

constexpr int kSplitRows = 4096;   // rows of the reduction per split of a weight-gradient GEMM

int upload_params(riv_mvnet* n, const float* host) {
    std::vector<uint16_t> h16(NPARAM);
    for (size_t i = 0; i < NPARAM; i++) h16[i] = host_bf16(host[i]);
    // inferences still queued on encoder streams read the old weights to the end
    RIV_HIP(n->after_uses());
    RIV_HIP(hipStreamSynchronize(n->st));
    RIV_HIP(hipMemcpy(n->p32.p, host, NPARAM * 4, hipMemcpyHostToDevice));
    RIV_HIP(hipMemcpy(n->p16.p, h16.data(), NPARAM * 2, hipMemcpyHostToDevice));
    RIV_HIP(hipMemset(n->adam_m.p, 0, NPARAM * 4));
    RIV_HIP(hipMemset(n->adam_v.p, 0, NPARAM * 4));
    n->step = 0;
    n->weights_changed();
    return RIV_OK;
}

int ensure_batch(riv_mvnet* n, int B) {
    if (B <= n->bcap) return RIV_OK;
    const size_t b = (size_t)B;
    RIV_HIP(n->samp.alloc(8 * b)); RIV_HIP(n->tgt.alloc(2 * b));
    RIV_HIP(n->x0.alloc(2 * b * 4096));
    RIV_HIP(n->c1.alloc(2 * b * 256 * K1)); RIV_HIP(n->z1.alloc(4 * b * 256 * F)); RIV_HIP(n->a1.alloc(2 * b * 256 * F));
    RIV_HIP(n->c2.alloc(2 * b * 64 * K2)); RIV_HIP(n->z2.alloc(4 * b * 64 * F)); RIV_HIP(n->a2.alloc(2 * b * 64 * F));
    RIV_HIP(n->c3.alloc(2 * b * 16 * K2)); RIV_HIP(n->z3.alloc(4 * b * FC_IN)); RIV_HIP(n->a3.alloc(2 * b * FC_IN));
    RIV_HIP(n->z4.alloc(4 * b * FC_H)); RIV_HIP(n->a4.alloc(2 * b * FC_H));
    RIV_HIP(n->y.alloc(4 * b * OUT)); RIV_HIP(n->dy.alloc(4 * b * OUT)); RIV_HIP(n->dyb.alloc(2 * b * OUT));
    RIV_HIP(n->da4.alloc(4 * b * FC_H)); RIV_HIP(n->dz4.alloc(2 * b * FC_H));
    RIV_HIP(n->da3.alloc(4 * b * FC_IN)); RIV_HIP(n->dz3.alloc(2 * b * 16 * F));
    RIV_HIP(n->dc3.alloc(4 * b * 16 * K2)); RIV_HIP(n->dz2.alloc(2 * b * 64 * F));
    RIV_HIP(n->dc2.alloc(4 * b * 64 * K2)); RIV_HIP(n->dz1.alloc(2 * b * 256 * F));
    RIV_HIP(n->lossb.alloc(4 * b)); RIV_HIP(n->prop.alloc(16 * b));
    if (n->staged) RIV_HIP(hipEventSynchronize(n->staged));
    RIV_HIP(n->stage.alloc(10 * b));
    if (!n->staged) RIV_HIP(hipEventCreateWithFlags(&n->staged, hipEventDisableTiming));
    // the largest split partials: conv1's weight gradient (K1 x F over B*256 rows) or fc1's
    const size_t splits1 = (b * 256 + kSplitRows - 1) / kSplitRows;
    n->part_n = std::max(splits1 * K1 * F, ((b + kSplitRows - 1) / kSplitRows) * (size_t)FC_IN * FC_H);
    n->part_n = std::max(n->part_n, ((b * 64 + kSplitRows - 1) / kSplitRows) * (size_t)K2 * F);
    RIV_HIP(n->part.alloc(4 * n->part_n));
    n->bcap = B;
    return RIV_OK;
}

Gemm gemm_fwd(const uint16_t* A, int lda, const uint16_t* W, int M, int N, int K, float* z, const float* bias, uint16_t* out, int relu,
              long long mg, long long ms, long long ldc) {
    Gemm g{};
    g.A = A; g.sam = lda; g.sak = 1;
    g.B = W; g.sbk = N; g.sbn = 1;
    g.M = M; g.N = N; g.K = K; g.kc = K; g.splits = 1;
    g.C = z; g.mg = mg; g.ms = ms; g.ldc = ldc; g.bias = bias; g.Cb = out; g.relu = relu;
    return g;
}

// The layer-wise forward pass over the n staged samples.
void forward(riv_mvnet* n, int B) {
    hipStream_t st = n->st;
    const Weights w = n->weights();
    gather(st, n->srcs.as<Source>(), n->samp.as<int>(), B, n->x0.as<uint16_t>(), n->a3.as<uint16_t>());
    im2col(st, n->x0.as<uint16_t>(), B, 32, C0, K1, n->c1.as<uint16_t>());
    gemm(st, gemm_fwd(n->c1.as<uint16_t>(), K1, w.w1, B * 256, F, K1, n->z1.as<float>(), w.b1, n->a1.as<uint16_t>(), 1, B * 256, 0, F));
    im2col(st, n->a1.as<uint16_t>(), B, 16, F, K2, n->c2.as<uint16_t>());
    gemm(st, gemm_fwd(n->c2.as<uint16_t>(), K2, w.w2, B * 64, F, K2, n->z2.as<float>(), w.b2, n->a2.as<uint16_t>(), 1, B * 64, 0, F));
    im2col(st, n->a2.as<uint16_t>(), B, 8, F, K2, n->c3.as<uint16_t>());
    // conv3's 16 pixels x 16 channels of sample b land in columns 0..255 of fc1 input row b
    gemm(st, gemm_fwd(n->c3.as<uint16_t>(), K2, w.w3, B * 16, F, K2, n->z3.as<float>(), w.b3, n->a3.as<uint16_t>(), 1, 16, FC_IN, F));
    gemm(st, gemm_fwd(n->a3.as<uint16_t>(), FC_IN, w.w4, B, FC_H, FC_IN, n->z4.as<float>(), w.b4, n->a4.as<uint16_t>(), 1, B, 0, FC_H));
    gemm(st, gemm_fwd(n->a4.as<uint16_t>(), FC_H, w.w5, B, OUT, FC_H, n->y.as<float>(), w.b5, nullptr, 0, B, 0, OUT));
}

// out[K][N] = A^T dZ over `rows` rows: A(r, k) = A[r * lda + k], dZ[r * N + n], in K splits of
// kSplitRows rows summed in order.
void weight_grad(riv_mvnet* n, const uint16_t* A, int lda, const uint16_t* dz, int K, int N, int rows, float* out) {
    Gemm g{};
    g.A = A; g.sam = 1; g.sak = lda;
    g.B = dz; g.sbk = N; g.sbn = 1;
    g.M = K; g.N = N; g.K = rows; g.kc = kSplitRows;
    g.splits = (rows + kSplitRows - 1) / kSplitRows;
    g.mg = K; g.ms = 0; g.ldc = N;
    if (g.splits > 1) {
        g.C = n->part.as<float>();
        gemm(n->st, g);
        sum_splits(n->st, n->part.as<float>(), g.splits, K * N, out);
    } else {
        g.C = out;
        gemm(n->st, g);
    }
}

// dA = dZ W^T: dA(m, k) = sum_n dZ[m * N + n] W[k * N + n], f32 at m * K + k.
void input_grad(riv_mvnet* n, const uint16_t* dz, const uint16_t* W, int M, int N, int K, float* out) {
    Gemm g{};
    g.A = dz; g.sam = N; g.sak = 1;
    g.B = W; g.sbk = 1; g.sbn = N;
    g.M = M; g.N = K; g.K = N; g.kc = N; g.splits = 1;
    g.C = out; g.mg = M; g.ms = 0; g.ldc = K;
    gemm(n->st, g);
}

void backward(riv_mvnet* n, int B) {
    hipStream_t st = n->st;
    const Weights w = n->weights();
    float* gr = n->grad.as<float>();
    loss(st, n->y.as<float>(), n->tgt.as<int8_t>(), B, n->dy.as<float>(), n->lossb.as<float>());
    to_bf16(st, n->dy.as<float>(), (size_t)B * OUT, n->dyb.as<uint16_t>());
    // head
    weight_grad(n, n->a4.as<uint16_t>(), FC_H, n->dyb.as<uint16_t>(), FC_H, OUT, B, gr + OFF_W5);
    colsum(st, n->dyb.as<uint16_t>(), B, OUT, gr + OFF_B5);
    input_grad(n, n->dyb.as<uint16_t>(), w.w5, B, OUT, FC_H, n->da4.as<float>());
    relu_bwd(st, n->da4.as<float>(), n->z4.as<float>(), B, FC_H, B, 0, FC_H, n->dz4.as<uint16_t>());
    // fc1
    weight_grad(n, n->a3.as<uint16_t>(), FC_IN, n->dz4.as<uint16_t>(), FC_IN, FC_H, B, gr + OFF_W4);
    colsum(st, n->dz4.as<uint16_t>(), B, FC_H, gr + OFF_B4);
    input_grad(n, n->dz4.as<uint16_t>(), w.w4, B, FC_H, FC_IN, n->da3.as<float>());
    relu_bwd(st, n->da3.as<float>(), n->z3.as<float>(), B * 16, F, 16, FC_IN, F, n->dz3.as<uint16_t>());
    // conv3
    weight_grad(n, n->c3.as<uint16_t>(), K2, n->dz3.as<uint16_t>(), K2, F, B * 16, gr + OFF_W3);
    colsum(st, n->dz3.as<uint16_t>(), B * 16, F, gr + OFF_B3);
    input_grad(n, n->dz3.as<uint16_t>(), w.w3, B * 16, F, K2, n->dc3.as<float>());
    col2im_relu(st, n->dc3.as<float>(), B, 8, F, K2, n->z2.as<float>(), n->dz2.as<uint16_t>());
    // conv2
    weight_grad(n, n->c2.as<uint16_t>(), K2, n->dz2.as<uint16_t>(), K2, F, B * 64, gr + OFF_W2);
    colsum(st, n->dz2.as<uint16_t>(), B * 64, F, gr + OFF_B2);
    input_grad(n, n->dz2.as<uint16_t>(), w.w2, B * 64, F, K2, n->dc2.as<float>());
    col2im_relu(st, n->dc2.as<float>(), B, 16, F, K2, n->z1.as<float>(), n->dz1.as<uint16_t>());
    // conv1
    weight_grad(n, n->c1.as<uint16_t>(), K1, n->dz1.as<uint16_t>(), K1, F, B * 256, gr + OFF_W1);
    colsum(st, n->dz1.as<uint16_t>(), B * 256, F, gr + OFF_B1);
}

int stage_batch(riv_mvnet* n, const int32_t* samples, const int8_t* targets, size_t count) {
    if (!n->nsrc) { set_error("mvnet: no sources set"); return RIV_E_ARG; }
    if (count == 0 || count > (1u << 20)) { set_error("mvnet: batch size must be in 1..=1048576"); return RIV_E_ARG; }
    for (size_t i = 0; i < count; i++) {
        if (samples[2 * i] < 0 || samples[2 * i] >= n->nsrc) { set_error("mvnet: sample source out of range"); return RIV_E_ARG; }
    }
    int r = ensure_batch(n, (int)count);
    if (r) return r;
    RIV_HIP(hipEventSynchronize(n->staged));   // the previous upload has read the staging buffer
    uint8_t* h = n->stage.as<uint8_t>();
    std::memcpy(h, samples, 8 * count);
    if (targets) std::memcpy(h + 8 * count, targets, 2 * count);
    RIV_HIP(hipMemcpyAsync(n->samp.p, h, 8 * count, hipMemcpyHostToDevice, n->st));
    if (targets) RIV_HIP(hipMemcpyAsync(n->tgt.p, h + 8 * count, 2 * count, hipMemcpyHostToDevice, n->st));
    RIV_HIP(hipEventRecord(n->staged, n->st));
    return RIV_OK;
}

int read_loss(riv_mvnet* n, size_t count, float* loss_out) {
    if (!loss_out) return RIV_OK;
    std::vector<float> l(count);
    RIV_HIP(hipMemcpyAsync(l.data(), n->lossb.p, 4 * count, hipMemcpyDeviceToHost, n->st));
    RIV_HIP(hipStreamSynchronize(n->st));
    double s = 0;
    for (float v : l) s += v;
    *loss_out = (float)(s / (double)count);
    return RIV_OK;
}

}  // namespace

float rivnet::conf_logit(float confidence) {
    if (!(confidence > 0.0f)) return -INFINITY;
    if (confidence >= 1.0f) return INFINITY;
    return std::log(confidence / (1.0f - confidence));
}

extern "C" {

size_t riv_mvnet_param_count(void) { return NPARAM; }

int riv_mvnet_create(int32_t device_id, uint64_t seed, riv_mvnet** out) {
    if (!out) return RIV_E_ARG;
    *out = nullptr;
    int r = ensure_device(device_id);
    if (r) return r;
    riv_mvnet* n = new riv_mvnet();
    n->dev = device_id;
    auto fail = [&](int code) { riv_mvnet_destroy(n); return code; };
    if (hipStreamCreateWithFlags(&n->st, hipStreamNonBlocking) != hipSuccess) { set_error("mvnet: stream"); return fail(RIV_E_GPU); }
    if (n->p32.alloc(NPARAM * 4) || n->p16.alloc(NPARAM * 2) || n->adam_m.alloc(NPARAM * 4) || n->adam_v.alloc(NPARAM * 4) ||
        n->grad.alloc(NPARAM * 4) || n->mask.alloc(NPARAM) || n->pk.alloc((size_t)rivnet::NFRAG * 64 * sizeof(uint2))) {
        set_error("mvnet: out of device memory");
        return fail(RIV_E_GPU);
    }
    // uniform initialisation, bound sqrt(6 / fan_in) for the ReLU layers and sqrt(3 / fan_in) for
    // the head; biases zero
    std::vector<float> h(NPARAM, 0.0f);
    Rng rng(seed);
    auto fill = [&](size_t off, int K, int N, int fan_in, double gain) {
        const double bound = std::sqrt(gain / fan_in);
        for (int k = 0; k < K; k++)
            for (int j = 0; j < N; j++) h[off + (size_t)k * N + j] = (float)((2.0 * rng.uniform() - 1.0) * bound);
    };
    fill(rivnet::OFF_W1, 9 * rivnet::C0, rivnet::F, 9 * 3, 6.0);   // three live input channels
    fill(rivnet::OFF_W2, rivnet::K2, rivnet::F, rivnet::K2, 6.0);
    fill(rivnet::OFF_W3, rivnet::K2, rivnet::F, rivnet::K2, 6.0);
    fill(rivnet::OFF_W4, rivnet::FC_IN, rivnet::FC_H, rivnet::FC_IN, 6.0);
    for (int k = 0; k < rivnet::FC_H; k++)
        for (int j = 0; j < rivnet::OUT_USED; j++)
            h[rivnet::OFF_W5 + (size_t)k * rivnet::OUT + j] = (float)((2.0 * rng.uniform() - 1.0) * std::sqrt(3.0 / rivnet::FC_H));
    const std::vector<uint8_t> m = train_mask();
    for (size_t i = 0; i < NPARAM; i++) if (!m[i]) h[i] = 0.0f;
    if (hipMemcpy(n->mask.p, m.data(), NPARAM, hipMemcpyHostToDevice) != hipSuccess) { set_error("mvnet: upload"); return fail(RIV_E_GPU); }
    r = upload_params(n, h.data());
    if (r) return fail(r);
    *out = n;
    return RIV_OK;
}

void riv_mvnet_destroy(riv_mvnet* n) {
    if (!n) return;
    (void)hipSetDevice(n->dev);
    if (n->st) { (void)hipStreamSynchronize(n->st); (void)hipStreamDestroy(n->st); }
    if (n->staged) (void)hipEventDestroy(n->staged);
    for (auto& u : n->uses) {
        (void)hipEventSynchronize(u.second);
        (void)hipEventDestroy(u.second);
    }
    delete n;
}

int riv_mvnet_get_params(riv_mvnet* n, float* out, size_t count) {
    if (!n || !out || count != NPARAM) { set_error("mvnet: parameter count mismatch"); return RIV_E_ARG; }
    RIV_HIP(hipSetDevice(n->dev));
    RIV_HIP(hipStreamSynchronize(n->st));
    RIV_HIP(hipMemcpy(out, n->p32.p, NPARAM * 4, hipMemcpyDeviceToHost));
    return RIV_OK;
}

int riv_mvnet_set_params(riv_mvnet* n, const float* in, size_t count) {
    if (!n || !in || count != NPARAM) { set_error("mvnet: parameter count mismatch"); return RIV_E_ARG; }
    RIV_HIP(hipSetDevice(n->dev));
    RIV_HIP(hipStreamSynchronize(n->st));
    std::vector<float> h(in, in + NPARAM);
    const std::vector<uint8_t> m = train_mask();
    for (size_t i = 0; i < NPARAM; i++) if (!m[i]) h[i] = 0.0f;
    return upload_params(n, h.data());
}

int riv_mvnet_set_sources(riv_mvnet* n, const riv_mv_source* src, size_t count) {
    if (!n || (!src && count)) return RIV_E_ARG;
    RIV_HIP(hipSetDevice(n->dev));
    std::vector<rivnet::Source> h(count);
    for (size_t i = 0; i < count; i++) {
        const riv_mv_source& s = src[i];
        if (!s.cur_y || !s.ref_y || s.width == 0 || s.height == 0 || s.width % 16 || s.height % 16) {
            set_error("mvnet: source planes must be non-null with dimensions that are multiples of 16");
            return RIV_E_ARG;
        }
        h[i] = {s.cur_y, s.ref_y, s.prev_mv, s.prev_stride, (int)s.width, (int)s.height, (int)s.width / 16, (int)s.height / 16};
    }
    RIV_HIP(hipStreamSynchronize(n->st));
    RIV_HIP(n->srcs.alloc(sizeof(rivnet::Source) * std::max<size_t>(1, count)));
    if (count) RIV_HIP(hipMemcpy(n->srcs.p, h.data(), sizeof(rivnet::Source) * count, hipMemcpyHostToDevice));
    n->nsrc = (int)count;
    return RIV_OK;
}

int riv_mvnet_train_step(riv_mvnet* n, const int32_t* samples, const int8_t* targets, size_t count, float lr, float* loss_out) {
    if (!n || !samples || !targets) return RIV_E_ARG;
    RIV_HIP(hipSetDevice(n->dev));
    int r = stage_batch(n, samples, targets, count);
    if (r) return r;
    forward(n, (int)count);
    backward(n, (int)count);
    n->step++;
    RIV_HIP(n->after_uses());   // the update waits for inferences queued on encoder streams
    const float b1 = 0.9f, b2 = 0.999f;
    rivnet::adam(n->st, n->p32.as<float>(), n->adam_m.as<float>(), n->adam_v.as<float>(), n->grad.as<float>(), n->mask.as<uint8_t>(), NPARAM, lr,
                 b1, b2, 1e-8f, 1.0f - std::pow(b1, (float)n->step), 1.0f - std::pow(b2, (float)n->step), n->p16.as<uint16_t>());
    n->weights_changed();
    RIV_HIP(hipGetLastError());
    return read_loss(n, count, loss_out);
}

int riv_mvnet_gradients(riv_mvnet* n, const int32_t* samples, const int8_t* targets, size_t count, float* grad_out, float* loss_out) {
    if (!n || !samples || !targets || !grad_out) return RIV_E_ARG;
    RIV_HIP(hipSetDevice(n->dev));
    int r = stage_batch(n, samples, targets, count);
    if (r) return r;
    forward(n, (int)count);
    backward(n, (int)count);
    RIV_HIP(hipGetLastError());
    RIV_HIP(hipMemcpyAsync(grad_out, n->grad.p, NPARAM * 4, hipMemcpyDeviceToHost, n->st));
    RIV_HIP(hipStreamSynchronize(n->st));
    return read_loss(n, count, loss_out);
}

int riv_mvnet_forward(riv_mvnet* n, const int32_t* samples, size_t count, int fused, float* out) {
    if (!n || !samples || !out) return RIV_E_ARG;
    RIV_HIP(hipSetDevice(n->dev));
    int r = stage_batch(n, samples, nullptr, count);
    if (r) return r;
    if (fused) {
        RIV_HIP(rivnet::infer(n->st, n->srcs.as<rivnet::Source>(), rivnet::Source{}, n->samp.as<int>(), (int)count, n->weights(), INFINITY, 0,
                              nullptr, n->y.as<float>()));
    } else {
        forward(n, (int)count);
    }
    RIV_HIP(hipGetLastError());
    RIV_HIP(hipMemcpyAsync(out, n->y.p, 4 * count * rivnet::OUT, hipMemcpyDeviceToHost, n->st));
    RIV_HIP(hipStreamSynchronize(n->st));
    return RIV_OK;
}

int riv_mvnet_propose(riv_mvnet* n, const int32_t* samples, size_t count, float confidence, int32_t sr, uint32_t* prop) {
    if (!n || !samples || !prop || sr < 0 || sr > 127) return RIV_E_ARG;
    RIV_HIP(hipSetDevice(n->dev));
    int r = stage_batch(n, samples, nullptr, count);
    if (r) return r;
    RIV_HIP(rivnet::infer(n->st, n->srcs.as<rivnet::Source>(), rivnet::Source{}, n->samp.as<int>(), (int)count, n->weights(),
                          rivnet::conf_logit(confidence), sr, n->prop.as<uint4>(), nullptr));
    RIV_HIP(hipMemcpyAsync(prop, n->prop.p, 16 * count, hipMemcpyDeviceToHost, n->st));
    RIV_HIP(hipStreamSynchronize(n->st));
    return RIV_OK;
}

}  // extern "C"
