// C-ABI of the learned surrogate (declarations in include/icephase.h): device
// state, parameter (un)packing, the training step and record serving.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "host_records.hpp"
#include "runtime_internal.hpp"
#include "surrogate.hpp"

namespace icephase {
hipError_t launch_cdf(const CdfJob* d_jobs, uint32_t njobs, uint32_t max_onp, bool any_coarse, uint32_t reduce_blocks,
                      hipStream_t s);
uint32_t cdf_partial_blocks(uint32_t reduce_blocks, uint32_t max_onp);
}  // namespace icephase

using namespace icephase;
using namespace icephase::sur;

namespace {

uint16_t host_bf16(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float host_f32(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

// condition inputs: alpha/90, (lambda-550)/150, (radius-1.5)/1.5, log2(aspect)
bool normalize_cond(const float* c, float* out)
{
    for (int k = 0; k < 4; ++k)
        if (!std::isfinite(c[k])) return false;
    if (!(c[1] > 0 && c[2] > 0 && c[3] > 0)) return false;
    out[0] = c[0] / 90.0f;
    out[1] = (c[1] - 550.0f) / 150.0f;
    out[2] = (c[2] - 1.5f) / 1.5f;
    out[3] = std::log2(c[3]);
    return true;
}

uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

struct iph_sur {
    iph_ctx* ctx = nullptr;
    iph_sur_config cfg{};
    Static st{};
    uint64_t step = 0;
    DevBuf<float> params, adam_m, adam_v, grad;
    DevBuf<__bf16> w, wt, xhi, xlo, Z, dZ;
    DevBuf<uint32_t> canon, mirr;
    DevBuf<float> pphi, ptheta, feat, tfeat, logw, logits, dF, lse, loss_part, loss, conds, part, part2, part_w5, part_b5, part_db;
    DevBuf<float2> lse_part;
    DevBuf<float> tv_part, tv_G;  // total-variation term: block sums of g p, their record sum
    float tv_weight = 0.0f;       // loss = cross entropy + tv_weight * total variation
    DevBuf<double> mass;
    // training set: probabilities [dataset][M], normalized conditions [dataset][4]
    DevBuf<float> data, data_conds;
    std::vector<uint8_t> filled;
    // predictions (FP32) and comparison scratch
    DevBuf<float> pred, traced;
    DevBuf<double> row_cos, cmp_part;
    // CDF serving state (built for the policy of the first iph_sur_records call)
    DevBuf<double> cdf_buf;
    DevBuf<CdfJob> d_jobs;
    std::vector<CdfJob> jobs;
    std::vector<double> edges;
    uint32_t cdf_n = 0, cdf_ont = 0, cdf_onp = 0;
    double cdf_l1 = 0, cdf_lost = 0, cdf_tv = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double ms[3] = {0, 0, 0};

    // weight-gradient split partials of layer l: [splits][256][in]
    float* part_layer(int l) const
    {
        const size_t per = (size_t)st.splits * 256;
        return part.p + (l == 0 ? 0 : per * (64 + (size_t)(l - 1) * 256));
    }
    Net net() const
    {
        Net n;
        n.w1 = params.p + kOffW1;
        n.pphi = pphi.p;
        n.ptheta = ptheta.p;
        for (int l = 0; l < 3; ++l) {
            n.w[l] = w.p + (size_t)l * 65536;
            n.wt[l] = wt.p + (size_t)l * 65536;
        }
        for (int l = 0; l < 4; ++l) n.b[l] = params.p + off_b(l);
        n.w5 = params.p + kOffW5;
        n.b5 = params.p + kOffB5;
        return n;
    }
    ~iph_sur()
    {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

static int sur_fail(iph_sur* s, const std::string& m) { return fail(s ? s->ctx : nullptr, IPH_E_INVALID, m); }

// packed <-> device parameter layouts
static void unpack(const float* packed, std::vector<float>& dev)
{
    dev.assign(kParamsPad, 0.0f);
    size_t p = 0;
    for (int o = 0; o < 256; ++o)
        for (int i = 0; i < kInputs; ++i) dev[kOffW1 + o * 64 + i] = packed[p++];
    std::memcpy(&dev[kOffB1], packed + p, (kParams - kOffB1) * sizeof(float));
}
static void pack_host(const std::vector<float>& dev, float* packed)
{
    size_t p = 0;
    for (int o = 0; o < 256; ++o)
        for (int i = 0; i < kInputs; ++i) packed[p++] = dev[kOffW1 + o * 64 + i];
    std::memcpy(packed + p, &dev[kOffB1], (kParams - kOffB1) * sizeof(float));
}

extern "C" {

uint32_t iph_sur_param_count(void) { return kPacked; }

int iph_sur_create(iph_ctx* ctx, const iph_sur_config* cfg, iph_sur** out)
{
    if (!ctx) return fail(nullptr, IPH_E_INVALID, "ctx must not be NULL.");
    if (!cfg || !out || !cfg->theta_count || !cfg->phi_count || !cfg->max_records ||
        (uint64_t)cfg->theta_count * cfg->phi_count > (1u << 26))
        return fail(ctx, IPH_E_INVALID, "Invalid surrogate configuration.");
    *out = nullptr;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    auto s = std::make_unique<iph_sur>();
    s->ctx = ctx;
    s->cfg = *cfg;
    Static& st = s->st;
    st.nt = cfg->theta_count;
    st.np = cfg->phi_count;
    if (st.np % 2) return fail(ctx, IPH_E_INVALID, "The surrogate needs an even phi_count (mirror-symmetric records).");
    st.M = st.nt * st.np;
    // folded grid: canonical column j pairs with mirror(j) = (np/2 - 1 - j) mod np
    std::vector<uint32_t> canon, mirr;
    for (uint32_t j = 0; j < st.np; ++j) {
        const uint32_t m = (uint32_t)(((int64_t)st.np / 2 - 1 - (int64_t)j + st.np) % st.np);
        if (j <= m) {
            canon.push_back(j);
            mirr.push_back(m);
        }
    }
    st.nq = (uint32_t)canon.size();
    st.Mh = st.nt * st.nq;
    st.Mpad = (st.Mh + 127) / 128 * 128;  // whole 128-row tiles (kTileRows, kFwdRows)
    // weight-gradient splits: rows per split a multiple of 32, ~256 splits
    st.rows_per_split = std::max<uint32_t>(32, (st.Mpad / 256 + 31) / 32 * 32);
    st.splits = (st.Mpad + st.rows_per_split - 1) / st.rows_per_split;
    // static inputs and row weights
    std::vector<double> ue;
    std::string err;
    if (!make_u_edges(st.nt, ue, err)) return fail(ctx, IPH_E_INVALID, err);
    std::vector<uint16_t> xh((size_t)st.Mpad * 64, 0), xl((size_t)st.Mpad * 64, 0);
    std::vector<float> lw(st.Mpad, -INFINITY), feat((size_t)st.nq * kFourier), tfeat((size_t)st.nt * kThetaIn);
    const double pi = 3.141592653589793238462643383280, two_pi = 2 * pi;
    // record-independent inputs in FP64: xt[i] = [t, sin(pi f t), cos(pi f t)],
    // xv[q] = [cos(2 pi k (v - 1/4))] of canonical column q
    std::vector<double> xt((size_t)st.nt * kThetaIn), xv((size_t)st.nq * kFourier);
    for (uint32_t i = 0; i < st.nt; ++i) {
        const double t = (i + 0.5) / st.nt;
        double* x = &xt[(size_t)i * kThetaIn];
        x[0] = t;
        for (int m = 0; m < kThetaFreqs; ++m) {
            x[1 + m] = std::sin(pi * kThetaFreq[m] * t);
            x[1 + kThetaFreqs + m] = std::cos(pi * kThetaFreq[m] * t);
        }
    }
    for (uint32_t q = 0; q < st.nq; ++q) {
        const double w = (canon[q] + 0.5) / st.np - 0.25;
        for (int k = 1; k <= kFourier; ++k) xv[(size_t)q * kFourier + k - 1] = std::cos(two_pi * k * w);
    }
    for (size_t e = 0; e < tfeat.size(); ++e) tfeat[e] = (float)xt[e];
    for (size_t e = 0; e < feat.size(); ++e) feat[e] = (float)xv[e];
    for (uint32_t i = 0; i < st.nt; ++i)
        for (uint32_t q = 0; q < st.nq; ++q) {
            const size_t r = (size_t)i * st.nq + q;
            double x[kStatic];
            x[0] = xt[(size_t)i * kThetaIn];
            for (int k = 0; k < kFourier; ++k) x[kColPhi + k] = xv[(size_t)q * kFourier + k];
            for (int k = 0; k < 2 * kThetaFreqs; ++k) x[kColTheta + k] = xt[(size_t)i * kThetaIn + 1 + k];
            for (int c = 0; c < kStatic; ++c) {
                const uint16_t h = host_bf16((float)x[c]);
                xh[r * 64 + c] = h;
                xl[r * 64 + c] = host_bf16((float)(x[c] - (double)host_f32(h)));
            }
            lw[r] = (float)(std::log(ue[i + 1] - ue[i]) + (canon[q] == mirr[q] ? 0.0 : std::log(2.0)));
        }
    IPH_HIP(ctx, s->canon.alloc(st.nq));
    IPH_HIP(ctx, s->mirr.alloc(st.nq));
    IPH_HIP(ctx, hipMemcpy(s->canon.p, canon.data(), st.nq * 4, hipMemcpyHostToDevice));
    IPH_HIP(ctx, hipMemcpy(s->mirr.p, mirr.data(), st.nq * 4, hipMemcpyHostToDevice));
    st.canon = s->canon.p;
    st.mirr = s->mirr.p;
    const size_t zsz = (size_t)4 * st.Mpad * 256;
    IPH_HIP(ctx, s->xhi.alloc(xh.size()));
    IPH_HIP(ctx, s->xlo.alloc(xl.size()));
    IPH_HIP(ctx, s->logw.alloc(st.Mpad));
    IPH_HIP(ctx, s->feat.alloc(feat.size()));
    IPH_HIP(ctx, hipMemcpy(s->feat.p, feat.data(), feat.size() * 4, hipMemcpyHostToDevice));
    IPH_HIP(ctx, s->tfeat.alloc(tfeat.size()));
    IPH_HIP(ctx, hipMemcpy(s->tfeat.p, tfeat.data(), tfeat.size() * 4, hipMemcpyHostToDevice));
    IPH_HIP(ctx, hipMemcpy(s->xhi.p, xh.data(), xh.size() * 2, hipMemcpyHostToDevice));
    IPH_HIP(ctx, hipMemcpy(s->xlo.p, xl.data(), xl.size() * 2, hipMemcpyHostToDevice));
    IPH_HIP(ctx, hipMemcpy(s->logw.p, lw.data(), lw.size() * 4, hipMemcpyHostToDevice));
    st.xhi = s->xhi.p;
    st.xlo = s->xlo.p;
    st.logw = s->logw.p;
    st.feat = s->feat.p;
    st.tfeat = s->tfeat.p;
    IPH_HIP(ctx, s->params.alloc(kParamsPad));
    IPH_HIP(ctx, s->adam_m.alloc(kParamsPad));
    IPH_HIP(ctx, s->adam_v.alloc(kParamsPad));
    IPH_HIP(ctx, s->grad.alloc(kParamsPad));
    IPH_HIP(ctx, s->pphi.alloc((size_t)st.nq * 256));
    IPH_HIP(ctx, s->ptheta.alloc((size_t)st.nt * 256));
    IPH_HIP(ctx, s->w.alloc(3 * 65536));
    IPH_HIP(ctx, s->wt.alloc(3 * 65536));
    IPH_HIP(ctx, s->Z.alloc(zsz));
    IPH_HIP(ctx, s->dZ.alloc(zsz));
    IPH_HIP(ctx, s->logits.alloc((size_t)cfg->max_records * st.Mpad));
    IPH_HIP(ctx, s->dF.alloc(st.Mpad));
    IPH_HIP(ctx, s->lse.alloc(cfg->max_records));
    IPH_HIP(ctx, s->lse_part.alloc((size_t)cfg->max_records * kLseParts));
    IPH_HIP(ctx, s->loss_part.alloc((st.Mpad + 255) / 256));
    IPH_HIP(ctx, s->tv_part.alloc((st.Mpad + 255) / 256));
    IPH_HIP(ctx, s->tv_G.alloc(1));
    IPH_HIP(ctx, s->loss.alloc(std::max<uint32_t>(cfg->max_records, 1024)));
    IPH_HIP(ctx, s->conds.alloc((size_t)4 * std::max<uint32_t>(cfg->max_records, 1024)));
    IPH_HIP(ctx, s->part.alloc((size_t)st.splits * 256 * (64 + 3 * 256)));
    IPH_HIP(ctx, s->part2.alloc((size_t)kReduceGroups * 256 * 256));
    IPH_HIP(ctx, s->part_w5.alloc((size_t)(st.Mpad / kTileRows) * 256));
    IPH_HIP(ctx, s->part_b5.alloc(st.Mpad / kTileRows));
    IPH_HIP(ctx, s->part_db.alloc((size_t)4 * (st.Mpad / kTileRows) * 256));
    IPH_HIP(ctx, s->mass.alloc((size_t)cfg->max_records * st.M));
    IPH_HIP(ctx, s->data.alloc((size_t)cfg->dataset * st.Mh));
    IPH_HIP(ctx, s->data_conds.alloc((size_t)4 * std::max<uint32_t>(cfg->dataset, 1)));
    s->filled.assign(cfg->dataset, 0);
    IPH_HIP(ctx, s->pred.alloc((size_t)cfg->max_records * st.M));
    IPH_HIP(ctx, s->traced.alloc(st.M));
    IPH_HIP(ctx, s->cmp_part.alloc((size_t)kCompareBlocks * 2));
    {
        std::vector<double> mc(st.nt);
        for (uint32_t i = 0; i < st.nt; ++i) mc[i] = 1.0 - (ue[i] + ue[i + 1]);
        IPH_HIP(ctx, s->row_cos.alloc(st.nt));
        IPH_HIP(ctx, hipMemcpy(s->row_cos.p, mc.data(), st.nt * 8, hipMemcpyHostToDevice));
    }
    IPH_HIP(ctx, hipMemset(s->adam_m.p, 0, kParamsPad * 4));
    IPH_HIP(ctx, hipMemset(s->adam_v.p, 0, kParamsPad * 4));
    IPH_HIP(ctx, hipMemset(s->grad.p, 0, kParamsPad * 4));
    for (auto& e : s->ev) IPH_HIP(ctx, hipEventCreate(&e));
    // initialization: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases,
    // output layer scaled by 0.1
    std::vector<float> packed(kPacked);
    uint64_t seed = cfg->seed;
    size_t p = 0;
    auto uni = [&](double a) { return (float)(a * (2.0 * ((splitmix(seed) >> 11) * 0x1p-53) - 1.0)); };
    const double a1 = 1 / std::sqrt((double)kInputs), a2 = 1 / std::sqrt(256.0);
    for (int k = 0; k < 256 * kInputs + 256; ++k) packed[p++] = uni(a1);
    for (int l = 0; l < 3; ++l)
        for (int k = 0; k < 65536 + 256; ++k) packed[p++] = uni(a2);
    for (int k = 0; k < 257; ++k) packed[p++] = uni(0.1 * a2);
    iph_sur* raw = s.release();
    ctx->surrogates.push_back(raw);
    const int rc = iph_sur_set_params(raw, packed.data());
    if (rc != IPH_OK) {
        iph_sur_destroy(raw);
        return rc;
    }
    *out = raw;
    return IPH_OK;
}

void iph_sur_destroy(iph_sur* s)
{
    if (!s) return;
    iph_ctx* ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    auto it = std::find(ctx->surrogates.begin(), ctx->surrogates.end(), s);
    if (it != ctx->surrogates.end()) ctx->surrogates.erase(it);
    delete s;
}

int iph_sur_get_params(iph_sur* s, float* packed)
{
    if (!s || !packed) return sur_fail(s, "Invalid arguments.");
    iph_ctx* ctx = s->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<float> dev(kParamsPad);
    IPH_HIP(ctx, hipMemcpy(dev.data(), s->params.p, kParamsPad * 4, hipMemcpyDeviceToHost));
    pack_host(dev, packed);
    return IPH_OK;
}

int iph_sur_set_params(iph_sur* s, const float* packed)
{
    if (!s || !packed) return sur_fail(s, "Invalid arguments.");
    iph_ctx* ctx = s->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<float> dev;
    unpack(packed, dev);
    for (float v : dev)
        if (!std::isfinite(v)) return sur_fail(s, "Parameters must be finite.");
    IPH_HIP(ctx, hipMemcpy(s->params.p, dev.data(), kParamsPad * 4, hipMemcpyHostToDevice));
    IPH_HIP(ctx, pack(s->params.p, s->w.p, s->wt.p, s->st, s->pphi.p, s->ptheta.p, ctx->stream));
    // new parameters start a new optimization: Adam moments and step count from zero
    IPH_HIP(ctx, hipMemsetAsync(s->adam_m.p, 0, kParamsPad * 4, ctx->stream));
    IPH_HIP(ctx, hipMemsetAsync(s->adam_v.p, 0, kParamsPad * 4, ctx->stream));
    s->step = 0;
    IPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return IPH_OK;
}

}  // extern "C"

// Forward + backward of training slot `slot`: the weight- and bias-gradient
// partials of its rows go into (accumulate: are added to) the per-layer split
// buffers; the loss goes to s->loss[k].
static int record_gradient(iph_sur* s, uint32_t slot, uint32_t k, bool accumulate)
{
    iph_ctx* ctx = s->ctx;
    hipStream_t q = ctx->stream;
    const Static& st = s->st;
    const Net n = s->net();
    const float* cond = s->data_conds.p + 4 * (size_t)slot;
    const float* d_target = s->data.p + (size_t)slot * st.Mh;
    IPH_HIP(ctx, forward(n, st, cond, 1, true, s->logits.p, s->Z.p, q));
    IPH_HIP(ctx, lse(st, s->logits.p, 1, s->lse_part.p, s->lse.p, q));
    IPH_HIP(ctx, dlogits(st, s->logits.p, s->lse.p, d_target, s->tv_weight, s->tv_part.p, s->tv_G.p, s->dF.p,
                         s->loss_part.p, s->loss.p + k, q));
    IPH_HIP(ctx, backward(n, st, s->dF.p, s->Z.p, s->dZ.p, s->part_w5.p, s->part_b5.p, s->part_db.p, accumulate, q));
    const size_t zl = (size_t)st.Mpad * 256;
    for (int l = 3; l >= 0; --l)
        IPH_HIP(ctx, wgrad(st, l, s->dZ.p + l * zl, l ? s->Z.p + (l - 1) * zl : nullptr, cond, s->part_layer(l),
                           accumulate, q));
    return IPH_OK;
}

// Fixed-order reductions of the split partials into s->grad (added).
static int reduce_gradient(iph_sur* s)
{
    iph_ctx* ctx = s->ctx;
    hipStream_t q = ctx->stream;
    const Static& st = s->st;
    const uint32_t tiles = st.Mpad / kTileRows;
    for (int l = 3; l >= 0; --l) {
        const uint32_t in = l == 0 ? 64 : 256;
        IPH_HIP(ctx, reduce(s->part_layer(l), st.splits, 256 * in, s->grad.p + off_w(l), true, s->part2.p, q));
        IPH_HIP(ctx, reduce(s->part_db.p + (size_t)l * tiles * 256, tiles, 256, s->grad.p + off_b(l), true, s->part2.p,
                            q));
    }
    IPH_HIP(ctx, reduce(s->part_w5.p, tiles, 256, s->grad.p + kOffW5, true, s->part2.p, q));
    IPH_HIP(ctx, reduce(s->part_b5.p, tiles, 1, s->grad.p + kOffB5, true, s->part2.p, q));
    return IPH_OK;
}

extern "C" {

static bool slot_ok(const iph_sur* s, uint32_t slot) { return slot < s->cfg.dataset && s->filled[slot]; }

int iph_sur_set_record(iph_sur* s, uint32_t slot, iph_batch* b, uint32_t rec, const float cond[4])
{
    if (!s || !b || !cond || slot >= s->cfg.dataset) return sur_fail(s, "Invalid arguments (slot beyond the training set?).");
    iph_ctx* ctx = s->ctx;
    float c[4];
    if (!normalize_cond(cond, c)) return sur_fail(s, "Invalid condition (alpha, wavelength, radius, aspect).");
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    uint32_t bnt = 0, bnp = 0;
    if (!batch_record_grid(b, rec, bnt, bnp) || bnt != s->st.nt || bnp != s->st.np)
        return sur_fail(s, "The traced record's grid differs from the surrogate's.");
    const int rc = iph_batch_probabilities(b, rec, s->traced.p);
    if (rc != IPH_OK) return fail(ctx, rc, std::string("iph_batch_probabilities: ") + iph_last_error(nullptr));
    IPH_HIP(ctx, fold(s->st, s->traced.p, s->data.p + (size_t)slot * s->st.Mh, ctx->stream));
    IPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    IPH_HIP(ctx, hipMemcpy(s->data_conds.p + 4 * (size_t)slot, c, sizeof c, hipMemcpyHostToDevice));
    s->filled[slot] = 1;
    return IPH_OK;
}

int iph_sur_set_target(iph_sur* s, uint32_t slot, const float* host_probs, const float cond[4])
{
    if (!s || !host_probs || !cond || slot >= s->cfg.dataset) return sur_fail(s, "Invalid arguments.");
    iph_ctx* ctx = s->ctx;
    float c[4];
    if (!normalize_cond(cond, c)) return sur_fail(s, "Invalid condition (alpha, wavelength, radius, aspect).");
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    IPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    IPH_HIP(ctx, hipMemcpy(s->traced.p, host_probs, (size_t)s->st.M * 4, hipMemcpyHostToDevice));
    IPH_HIP(ctx, fold(s->st, s->traced.p, s->data.p + (size_t)slot * s->st.Mh, ctx->stream));
    IPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    IPH_HIP(ctx, hipMemcpy(s->data_conds.p + 4 * (size_t)slot, c, sizeof c, hipMemcpyHostToDevice));
    s->filled[slot] = 1;
    return IPH_OK;
}

int iph_sur_set_loss(iph_sur* s, float tv_weight)
{
    if (!s) return sur_fail(s, "Invalid arguments.");
    if (!(std::isfinite(tv_weight) && tv_weight >= 0))
        return sur_fail(s, "The total-variation weight must be finite and non-negative.");
    s->tv_weight = tv_weight;
    return IPH_OK;
}

int iph_sur_gradient(iph_sur* s, uint32_t slot, float* loss, float* packed_grad)
{
    if (!s || !slot_ok(s, slot)) return sur_fail(s, "Invalid or empty training slot.");
    iph_ctx* ctx = s->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    IPH_HIP(ctx, hipMemsetAsync(s->grad.p, 0, kParamsPad * 4, ctx->stream));
    int rc = record_gradient(s, slot, 0, false);
    if (rc == IPH_OK) rc = reduce_gradient(s);
    if (rc != IPH_OK) return rc;
    std::vector<float> g(kParamsPad);
    float l = 0;
    IPH_HIP(ctx, hipMemcpyAsync(&l, s->loss.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    IPH_HIP(ctx, hipMemcpyAsync(g.data(), s->grad.p, kParamsPad * 4, hipMemcpyDeviceToHost, ctx->stream));
    IPH_HIP(ctx, hipMemsetAsync(s->grad.p, 0, kParamsPad * 4, ctx->stream));
    IPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (loss) *loss = l;
    if (packed_grad) pack_host(g, packed_grad);
    return IPH_OK;
}

int iph_sur_train_step(iph_sur* s, const uint32_t* slots, uint32_t n, float lr, float* loss)
{
    if (!s || !slots || !n || n > 1024) return sur_fail(s, "Invalid arguments.");
    if (!(std::isfinite(lr) && lr > 0)) return sur_fail(s, "Learning rate must be finite and positive.");
    for (uint32_t k = 0; k < n; ++k)
        if (!slot_ok(s, slots[k])) return sur_fail(s, "Invalid or empty training slot.");
    iph_ctx* ctx = s->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t q = ctx->stream;
    IPH_HIP(ctx, hipEventRecord(s->ev[0], q));
    for (uint32_t k = 0; k < n; ++k) {
        const int rc = record_gradient(s, slots[k], k, k > 0);
        if (rc != IPH_OK) return rc;
    }
    const int rc = reduce_gradient(s);
    if (rc != IPH_OK) return rc;
    ++s->step;
    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)s->step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)s->step)));
    IPH_HIP(ctx, adam(s->params.p, s->adam_m.p, s->adam_v.p, s->grad.p, kParams, lr, 1.0f / n, c1, c2, q));
    IPH_HIP(ctx, pack(s->params.p, s->w.p, s->wt.p, s->st, s->pphi.p, s->ptheta.p, q));
    IPH_HIP(ctx, hipEventRecord(s->ev[1], q));
    if (loss) IPH_HIP(ctx, hipMemcpyAsync(loss, s->loss.p, n * 4, hipMemcpyDeviceToHost, q));
    IPH_HIP(ctx, hipStreamSynchronize(q));
    float t = 0;
    IPH_HIP(ctx, hipEventElapsedTime(&t, s->ev[0], s->ev[1]));
    s->ms[2] = t;
    return IPH_OK;
}

// network part of predict/compare/records: probabilities of n records (FP64
// into s->mass, FP32 into s->pred)
static int run_network(iph_sur* s, const float* conds, uint32_t n)
{
    iph_ctx* ctx = s->ctx;
    hipStream_t q = ctx->stream;
    std::vector<float> c(4 * (size_t)n);
    for (uint32_t r = 0; r < n; ++r)
        if (!normalize_cond(conds + 4 * r, &c[4 * r])) return sur_fail(s, "Invalid condition.");
    IPH_HIP(ctx, hipStreamSynchronize(q));
    IPH_HIP(ctx, hipMemcpy(s->conds.p, c.data(), c.size() * 4, hipMemcpyHostToDevice));
    IPH_HIP(ctx, hipEventRecord(s->ev[0], q));
    IPH_HIP(ctx, forward(s->net(), s->st, s->conds.p, n, false, s->logits.p, nullptr, q));
    IPH_HIP(ctx, lse(s->st, s->logits.p, n, s->lse_part.p, s->lse.p, q));
    IPH_HIP(ctx, probs(s->st, s->logits.p, s->lse.p, n, s->mass.p, s->pred.p, q));
    IPH_HIP(ctx, hipEventRecord(s->ev[1], q));
    return IPH_OK;
}

int iph_sur_predict(iph_sur* s, const float* conds, uint32_t n, float* probs_out)
{
    if (!s || !conds || !n || n > s->cfg.max_records || !probs_out)
        return sur_fail(s, "Invalid arguments (n > max_records?).");
    iph_ctx* ctx = s->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    const int rc = run_network(s, conds, n);
    if (rc != IPH_OK) return rc;
    IPH_HIP(ctx, hipMemcpyAsync(probs_out, s->pred.p, (size_t)n * s->st.M * 4, hipMemcpyDeviceToHost, ctx->stream));
    IPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float t = 0;
    IPH_HIP(ctx, hipEventElapsedTime(&t, s->ev[0], s->ev[1]));
    s->ms[0] = t;
    return IPH_OK;
}

int iph_sur_compare(iph_sur* s, const float* conds, uint32_t n, iph_batch* b, const uint32_t* recs, double* tv,
                    double* dg)
{
    if (!s || !conds || !n || n > s->cfg.max_records || !b || !recs || !tv || !dg) return sur_fail(s, "Invalid arguments.");
    iph_ctx* ctx = s->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    int rc = run_network(s, conds, n);
    if (rc != IPH_OK) return rc;
    std::vector<double> part((size_t)kCompareBlocks * 2);
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t bnt = 0, bnp = 0;
        if (!batch_record_grid(b, recs[k], bnt, bnp) || bnt != s->st.nt || bnp != s->st.np)
            return sur_fail(s, "The traced record's grid differs from the surrogate's.");
        rc = iph_batch_probabilities(b, recs[k], s->traced.p);
        if (rc != IPH_OK) return fail(ctx, rc, std::string("iph_batch_probabilities: ") + iph_last_error(nullptr));
        IPH_HIP(ctx, compare(s->pred.p + (size_t)k * s->st.M, s->traced.p, s->st.M, s->st.np, s->row_cos.p,
                             s->cmp_part.p, ctx->stream));
        IPH_HIP(ctx, hipMemcpyAsync(part.data(), s->cmp_part.p, part.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
        IPH_HIP(ctx, hipStreamSynchronize(ctx->stream));
        double a = 0, g = 0;
        for (uint32_t j = 0; j < kCompareBlocks; ++j) {
            a += part[2 * j];
            g += part[2 * j + 1];
        }
        tv[k] = 0.5 * a;
        dg[k] = std::fabs(g);
    }
    return IPH_OK;
}

int iph_sur_records(iph_sur* s, const float* conds, uint32_t n, const iph_cdf_policy* policy, double* phi,
                    double* theta, iph_cdf_report* reports)
{
    if (!s || !conds || !n || n > s->cfg.max_records || !phi || !theta) return sur_fail(s, "Invalid arguments.");
    iph_ctx* ctx = s->ctx;
    IPH_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t q = ctx->stream;
    const Static& st = s->st;
    const uint32_t ont = policy && policy->out_theta ? policy->out_theta : st.nt;
    const uint32_t onp = policy && policy->out_phi ? policy->out_phi : st.np;
    if (ont > st.nt || onp > st.np || st.nt % ont || st.np % onp)
        return sur_fail(s, "Storage dimensions must divide histogram dimensions.");
    const double l1 = policy ? policy->max_l1 : 1e-10, lost = policy ? policy->max_lost_mass : 1e-12,
                 tv = policy ? policy->max_coarsening_tv : 1.0;
    const bool coarse = ont != st.nt || onp != st.np;
    const uint32_t red = 512;
    if (s->cdf_n != s->cfg.max_records || s->cdf_ont != ont || s->cdf_onp != onp) {
        // CDF jobs over the surrogate's FP64 mass buffer, one per record slot
        std::string err;
        if (!make_u_edges(st.nt, s->edges, err)) return sur_fail(s, err);
        const uint32_t pblocks = cdf_partial_blocks(red, onp);
        const size_t per = (onp + 1) + (size_t)onp * (ont + 1) + onp + kCdfScalars + (size_t)pblocks * 8 +
                           (coarse ? (size_t)ont * onp : 0);
        const size_t ne = s->edges.size() + ont + 1;
        IPH_HIP(ctx, s->cdf_buf.alloc(ne + per * s->cfg.max_records));
        std::vector<double> e2(s->edges);
        const uint32_t dt = st.nt / ont;
        for (uint32_t i = 0; i <= ont; ++i) e2.push_back(s->edges[(size_t)i * dt]);
        IPH_HIP(ctx, hipMemcpy(s->cdf_buf.p, e2.data(), e2.size() * 8, hipMemcpyHostToDevice));
        s->jobs.assign(s->cfg.max_records, CdfJob{});
        for (uint32_t r = 0; r < s->cfg.max_records; ++r) {
            CdfJob& J = s->jobs[r];
            double* p = s->cdf_buf.p + ne + per * r;
            J.nt = st.nt; J.np = st.np; J.ont = ont; J.onp = onp;
            J.bins = nullptr; J.stats = nullptr;
            J.mass = s->mass.p + (size_t)r * st.M;
            J.edges = s->cdf_buf.p; J.cedges = s->cdf_buf.p + s->edges.size();
            J.phi = p; p += onp + 1;
            J.theta = p; p += (size_t)onp * (ont + 1);
            J.cols = p; p += onp;
            J.scalars = p; p += kCdfScalars;
            J.partials = p; p += (size_t)pblocks * 8;
            J.coarse = coarse ? p : nullptr;
        }
        IPH_HIP(ctx, s->d_jobs.alloc(s->jobs.size()));
        IPH_HIP(ctx, hipMemcpy(s->d_jobs.p, s->jobs.data(), s->jobs.size() * sizeof(CdfJob), hipMemcpyHostToDevice));
        s->cdf_n = s->cfg.max_records;
        s->cdf_ont = ont;
        s->cdf_onp = onp;
    }
    const int rc = run_network(s, conds, n);
    if (rc != IPH_OK) return rc;
    IPH_HIP(ctx, launch_cdf(s->d_jobs.p, n, onp, coarse, red, q));
    IPH_HIP(ctx, hipEventRecord(s->ev[2], q));
    for (uint32_t r = 0; r < n; ++r) {
        const CdfJob& J = s->jobs[r];
        IPH_HIP(ctx, hipMemcpyAsync(phi + (size_t)r * (onp + 1), J.phi, (onp + 1) * 8, hipMemcpyDeviceToHost, q));
        IPH_HIP(ctx, hipMemcpyAsync(theta + (size_t)r * onp * (ont + 1), J.theta, (size_t)onp * (ont + 1) * 8,
                                    hipMemcpyDeviceToHost, q));
    }
    std::vector<double> sc((size_t)n * kCdfScalars);
    for (uint32_t r = 0; r < n; ++r)
        IPH_HIP(ctx, hipMemcpyAsync(&sc[(size_t)r * kCdfScalars], s->jobs[r].scalars, kCdfScalars * 8,
                                    hipMemcpyDeviceToHost, q));
    IPH_HIP(ctx, hipStreamSynchronize(q));
    float t0 = 0, t1 = 0;
    IPH_HIP(ctx, hipEventElapsedTime(&t0, s->ev[0], s->ev[1]));
    IPH_HIP(ctx, hipEventElapsedTime(&t1, s->ev[0], s->ev[2]));
    s->ms[0] = t0;
    s->ms[1] = t1;
    if (reports) {
        for (uint32_t r = 0; r < n; ++r) {
            const double* x = &sc[(size_t)r * kCdfScalars];
            iph_cdf_report& R = reports[r];
            std::memset(&R, 0, sizeof R);
            R.fine_mass = x[kCdfMassFine];
            R.stored_mass = x[kCdfMassStored];
            R.g_source = x[kCdfAxialFine] / x[kCdfMassFine];
            R.coarsening_tv = x[kCdfTv];
            R.l1_error = x[kCdfL1];
            R.lost_mass = x[kCdfLost];
            R.max_cell_error = x[kCdfMaxCell];
            R.cdf_mass = x[kCdfRecovered];
            R.g_stored = x[kCdfAxialCdf] / x[kCdfRecovered];
            R.lost_cells = (uint64_t)x[kCdfLostCells];
            R.malformed = (uint64_t)x[kCdfMalformed];
            R.status = (R.l1_error > l1 || R.lost_mass > lost || R.coarsening_tv > tv || R.malformed) ? 6 : 0;
        }
    }
    return IPH_OK;
}

int iph_sur_timing(const iph_sur* s, double* ms)
{
    if (!s || !ms) return fail(nullptr, IPH_E_INVALID, "Invalid arguments.");
    std::memcpy(ms, s->ms, sizeof s->ms);
    return IPH_OK;
}

}  // extern "C"
