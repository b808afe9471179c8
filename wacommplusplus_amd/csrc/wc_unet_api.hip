// C ABI of the concentration-reconstruction network (include/wacomm_hip.h,
// section 4): parameters, forward, gradients and Adam steps of a 4-level
// U-Net with 32 base channels, run by the kernels of wc_unet.hip.
//
// Network (channels; L0 = full resolution, Lk = 1/2^k):
//   in   = x as channel 0 of 32 (the rest zero)                  L0
//   enc1 = relu(conv(relu(conv(in))))            32              L0  (skip s1)
//   enc2 = 2 convs on avgpool(s1)                64              L1  (skip s2)
//   enc3 = 2 convs on avgpool(s2)                128             L2  (skip s3)
//   bott = 2 convs on avgpool(s3)                256             L3
//   dec3 = 2 convs on concat(s3, up(bott))       384 -> 128      L2
//   dec2 = 2 convs on concat(s2, up(dec3))       192 -> 64       L1
//   dec1 = 2 convs on concat(s1, up(dec2))       96 -> 32        L0
//   out  = x + head . dec1 + b_head (1x1 conv, residual)
// Parameters, FP32, in this order: per conv layer w[cout][3][3][cin] then
// b[cout]; then the head's 32 weights and its bias.  The loss is the mean
// squared error of out against the target.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "wc_launch.h"

using namespace wc;

namespace {

struct Layer {
    int cout, c1, c2;        // input channels c1 (+ c2 concatenated)
    size_t w_off, b_off;     // offsets into the FP32 parameters
    size_t wb_off;           // offset into the BF16 weight copy
};

constexpr int kBase = 32;
constexpr size_t kBiasPartFloats = (size_t)1024 * 256;       // mask workgroups x channels
constexpr size_t kHeadPartFloats = (size_t)256 * kHeadPartials;
// weight-gradient partial slices of one batch of layers (the batches run in
// order on one stream, so each reuses the space)
constexpr size_t kUnetPartFloats = 3 * kWgradPartFloats;
// The tiled weight gradients go out in batches, each launched once the
// masked gradient of its last layer is ready: the two layers of a level
// (13-12, 11-10, ..., 3-2), then 1 and 0 alone, as their masks come last.
// Small batches keep the second stream close behind the data-gradient chain
// (measured: 1.21 ms per step against 1.24 ms with three batches of 4-5
// layers).  A batch also closes when it is full or out of space.
bool batch_ends_at(int li) { return li % 2 == 0 || li == 1; }

}  // namespace

struct wc_unet {
    int device = 0;
    hipStream_t stream = nullptr;
    // the backward pass runs each layer's weight gradient (and its
    // reductions) on `side`, beside the data gradient on `stream`:
    // evMask[l] orders it after layer l's mask, evWgrad[l] marks it done
    hipStream_t side = nullptr;
    hipEvent_t evMask[kMaxConvLayers] = {}, evWgrad[kMaxConvLayers] = {};
    std::string err;
    std::vector<Layer> layers;
    size_t nparam = 0, nwb = 0, head_off = 0;
    float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
    uint16_t *d_wb = nullptr, *d_wf = nullptr;     // BF16 weights, and flipped for the data gradient
    WeightMap wmap{};
    float *d_loss = nullptr;
    float *d_part = nullptr;                     // weight-gradient partial sums (kUnetPartFloats)
    // tiled weight gradients waiting for their batch launch on `side`: the
    // jobs, the partial floats they take of d_part, the last one's layer
    WgradJobs batch{};
    size_t batchFloats = 0;
    int batchLast = -1;
    float *d_bpart = nullptr;                    // bias-gradient partials, kBiasPartFloats per layer
    float *d_hpart = nullptr;                    // the head's partials (kHeadPartFloats)
    SideSums pending{};                          // sums the next mask launch finishes
    int adam_t = 0;
    double fwd_ms = 0;
    // forward and backward of the gradient and training calls, each captured
    // once per batch shape (and loss kind) and replayed; between the two the
    // targets are copied in on `side` (evUp), beside the forward pass
    bool useGraph = true;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    hipEvent_t evUp = nullptr;
    int gN = 0, gH = 0, gW = 0;
    bool gWgt = false;
    // activations / gradients for the current shape
    int n = 0, h = 0, w = 0;
    std::vector<void *> bufs;
    uint16_t *in0 = nullptr, *a1 = nullptr, *s1 = nullptr, *p1 = nullptr, *a2 = nullptr, *s2 = nullptr,
             *p2 = nullptr, *a3 = nullptr, *s3 = nullptr, *p3 = nullptr, *a4 = nullptr, *b4 = nullptr,
             *u3 = nullptr, *d3a = nullptr, *d3 = nullptr, *u2 = nullptr, *d2a = nullptr, *d2 = nullptr,
             *u1 = nullptr, *d1a = nullptr, *d1 = nullptr;
    // each layer's ReLU-masked output gradient in BF16 (backward scratch),
    // gmaskCap[l] elements: one buffer per layer, so the data-gradient chain
    // never waits for a weight gradient that still reads an earlier mask
    uint16_t *gmask[kMaxConvLayers] = {};
    size_t gmaskCap[kMaxConvLayers] = {};
    float *x = nullptr, *out = nullptr, *tgt = nullptr;
    // pinned host staging of a batch, [x (P cin) | t (P) | wgt (P)]: the
    // caller's arrays are copied in on the host, then to the device
    // asynchronously (a copy from pageable memory would hold the host until
    // the device is idle)
    float *h_stage = nullptr;
    float *wgt = nullptr;     // per-pixel loss weights of the current batch
    bool useWgt = false;
    int cin = 1;                // input channels (x holds channel 0, xc all of them when cin > 1)
    float *xc = nullptr;
    float *g_d1 = nullptr, *g_d1a = nullptr, *g_s1 = nullptr, *g_u1 = nullptr, *g_d2 = nullptr, *g_d2a = nullptr,
          *g_s2 = nullptr, *g_u2 = nullptr, *g_d3 = nullptr, *g_d3a = nullptr, *g_s3 = nullptr, *g_u3 = nullptr,
          *g_b4 = nullptr, *g_a4 = nullptr, *g_p3 = nullptr, *g_a3 = nullptr, *g_p2 = nullptr, *g_a2 = nullptr,
          *g_p1 = nullptr, *g_a1 = nullptr;
};

namespace {

int ufail(wc_unet *u, int code, const char *fmt, ...) {
    if (u) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        u->err = buf;
    }
    return code;
}

#define UCHK(u, call)                                                                                \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return ufail(u, e_ == hipErrorOutOfMemory ? WC_ERR_NOMEM : WC_ERR_HIP, "%s: %s (%s:%d)", #call, \
                         hipGetErrorString(e_), __FILE__, __LINE__);                                 \
    } while (0)

void drop_graph(wc_unet *u) {
    for (hipGraphExec_t &g : u->gexec) {
        if (g) hipGraphExecDestroy(g);
        g = nullptr;
    }
}

void free_shape(wc_unet *u) {
    drop_graph(u);           // its kernels point into the buffers freed below
    for (void *b : u->bufs) hipFree(b);
    u->bufs.clear();
    if (u->h_stage) hipHostFree(u->h_stage);
    u->h_stage = nullptr;
    u->n = u->h = u->w = 0;
}

template <class T>
int alloc(wc_unet *u, T *&p, size_t count) {
    void *q = nullptr;
    UCHK(u, hipMalloc(&q, count * sizeof(T) + 256));
    u->bufs.push_back(q);
    p = (T *)q;
    return WC_OK;
}

int ensure_shape(wc_unet *u, int n, int h, int w) {
    if (n == u->n && h == u->h && w == u->w) return WC_OK;
    if (n < 1 || h < 8 || w < 8 || h % 8 || w % 8)
        return ufail(u, WC_ERR_INVALID, "images must be at least 8x8 with sides divisible by 8 (got %dx%d)", h, w);
    if ((long long)n * h * w * 384 > (1LL << 40)) return ufail(u, WC_ERR_INVALID, "batch too large");
    UCHK(u, hipStreamSynchronize(u->stream));
    UCHK(u, hipStreamSynchronize(u->side));
    free_shape(u);
    const size_t P0 = (size_t)n * h * w, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    const int C = kBase;
    int rc = 0;
    rc |= alloc(u, u->in0, P0 * C) | alloc(u, u->a1, P0 * C) | alloc(u, u->s1, P0 * C) | alloc(u, u->p1, P1 * C);
    rc |= alloc(u, u->a2, P1 * 2 * C) | alloc(u, u->s2, P1 * 2 * C) | alloc(u, u->p2, P2 * 2 * C);
    rc |= alloc(u, u->a3, P2 * 4 * C) | alloc(u, u->s3, P2 * 4 * C) | alloc(u, u->p3, P3 * 4 * C);
    rc |= alloc(u, u->a4, P3 * 8 * C) | alloc(u, u->b4, P3 * 8 * C) | alloc(u, u->u3, P2 * 8 * C);
    rc |= alloc(u, u->d3a, P2 * 4 * C) | alloc(u, u->d3, P2 * 4 * C) | alloc(u, u->u2, P1 * 4 * C);
    rc |= alloc(u, u->d2a, P1 * 2 * C) | alloc(u, u->d2, P1 * 2 * C) | alloc(u, u->u1, P0 * 2 * C);
    rc |= alloc(u, u->d1a, P0 * C) | alloc(u, u->d1, P0 * C);
    for (size_t l = 0; l < u->layers.size(); ++l) {     // a layer of cout = C 2^k runs at level k: P0 / 4^k pixels
        int k = 0;
        while ((C << k) < u->layers[l].cout) ++k;
        u->gmaskCap[l] = (P0 >> (2 * k)) * (size_t)u->layers[l].cout;
        rc |= alloc(u, u->gmask[l], u->gmaskCap[l]);
    }
    rc |= alloc(u, u->x, P0) | alloc(u, u->out, P0) | alloc(u, u->tgt, P0) | alloc(u, u->wgt, P0);
    if (u->cin > 1) rc |= alloc(u, u->xc, P0 * u->cin);
    if (!rc && hipHostMalloc((void **)&u->h_stage, P0 * (u->cin + 2) * sizeof(float), hipHostMallocDefault) != hipSuccess) {
        u->h_stage = nullptr;
        rc = ufail(u, WC_ERR_NOMEM, "pinned staging of %zu bytes", P0 * (u->cin + 2) * sizeof(float));
    }
    rc |= alloc(u, u->g_d1, P0 * C) | alloc(u, u->g_d1a, P0 * C) | alloc(u, u->g_s1, P0 * C) |
          alloc(u, u->g_u1, P0 * 2 * C);
    rc |= alloc(u, u->g_d2, P1 * 2 * C) | alloc(u, u->g_d2a, P1 * 2 * C) | alloc(u, u->g_s2, P1 * 2 * C) |
          alloc(u, u->g_u2, P1 * 4 * C);
    rc |= alloc(u, u->g_d3, P2 * 4 * C) | alloc(u, u->g_d3a, P2 * 4 * C) | alloc(u, u->g_s3, P2 * 4 * C) |
          alloc(u, u->g_u3, P2 * 8 * C);
    rc |= alloc(u, u->g_b4, P3 * 8 * C) | alloc(u, u->g_a4, P3 * 8 * C) | alloc(u, u->g_p3, P3 * 4 * C);
    rc |= alloc(u, u->g_a3, P2 * 4 * C) | alloc(u, u->g_p2, P2 * 2 * C) | alloc(u, u->g_a2, P1 * 2 * C);
    rc |= alloc(u, u->g_p1, P1 * C) | alloc(u, u->g_a1, P0 * C);
    if (rc) {
        free_shape(u);
        return rc;
    }
    u->n = n; u->h = h; u->w = w;
    return WC_OK;
}

// forward conv of layer L: y = relu(conv(concat(x1, x2)) + b)
int conv_fwd(wc_unet *u, const Layer &L, const uint16_t *x1, const uint16_t *x2, uint16_t *y, int h, int w) {
    ConvArgs a{};
    a.x1 = x1; a.x2 = x2; a.m1 = nullptr; a.c1 = L.c1; a.c2 = L.c2;
    a.w_ = u->d_wb + L.wb_off; a.bias = u->d_p + L.b_off;
    a.y1 = y; a.y2 = nullptr; a.cy1 = L.cout; a.cy2 = 0; a.acc1 = a.acc2 = false; a.mode = 0;
    a.n = u->n; a.h = h; a.w = w;
    UCHK(u, launch_conv3x3(a, false, u->stream));
    return WC_OK;
}

// The masked output gradient of layer li (P pixels, output y): the buffer
// that the kernel producing the gradient writes it to.
int mask_of(wc_unet *u, int li, const uint16_t *y, long long P, MaskOut &m) {
    if ((size_t)P * u->layers[li].cout > u->gmaskCap[li])
        return ufail(u, WC_ERR_INVALID, "layer %d: mask buffer too small", li);
    m = MaskOut{u->gmask[li], y};
    return WC_OK;
}

// The batched weight gradients on the side stream, after the mask of the
// batch's last layer (and every earlier one: the main stream is in order);
// sums left pending (the head's) are finished first.
int flush_wgrad(wc_unet *u) {
    if (u->batch.nj == 0) return WC_OK;
    UCHK(u, hipStreamWaitEvent(u->side, u->evMask[u->batchLast], 0));
    UCHK(u, launch_finish_sums(u->pending, u->side));
    u->pending = SideSums{};
    UCHK(u, launch_wgrad_batch(u->batch, u->side));
    UCHK(u, hipEventRecord(u->evWgrad[u->batchLast], u->side));
    u->batch = WgradJobs{};
    u->batchFloats = 0;
    u->batchLast = -1;
    return WC_OK;
}

// backward of layer L, whose ReLU-masked output gradient u->gmask[li] its
// producer has written: the weight and bias gradients (side stream, batched
// with neighbouring layers where the tiled path applies); when gx1 != null
// the input gradients (channels [0, c1) to gx1, [c1, c1+c2) to gx2, stored
// or accumulated), or, with next.gm set, the previous layer's masked
// gradient in place of gx1 (a single input, stored).  pooledFrom: the input
// was the 2x2 average pool of the previous layer's output, whose gradient
// so far is pooledFrom (FP32); next.gm then receives that gradient plus the
// pooling's backward of this one, masked (gx1: FP32 scratch).  up.gm: the
// second input was the 2x upsampling of an earlier layer's output up.y, whose
// masked gradient up.gm receives the upsampling's backward of the gx2 part
// (gx2 and upScratch, FP32 at half size, are scratch).
int conv_bwd(wc_unet *u, const Layer &L, const uint16_t *x1, const uint16_t *x2, float *gx1, bool acc1, float *gx2,
             bool acc2, int h, int w, const MaskOut &next = MaskOut{}, const float *pooledFrom = nullptr,
             const MaskOut &up = MaskOut{}, float *upScratch = nullptr) {
    const long long P = (long long)u->n * h * w;
    const int li = (int)(&L - u->layers.data());
    if ((size_t)P * L.cout > u->gmaskCap[li]) return ufail(u, WC_ERR_INVALID, "layer %d: mask buffer too small", li);
    if (next.gm && (!gx1 || li == 0 || L.c2 || acc1 || u->layers[li - 1].cout != L.c1))
        return ufail(u, WC_ERR_INVALID, "layer %d: a masked input gradient needs one stored input", li);
    if (up.gm && (!gx1 || !gx2 || !L.c2 || acc2 || !upScratch))
        return ufail(u, WC_ERR_INVALID, "layer %d: the upsampling's backward needs a second, stored input", li);
    uint16_t *gm = u->gmask[li];
    WgradArgs g{};
    g.x1 = x1; g.x2 = x2; g.c1 = L.c1; g.c2 = L.c2; g.cout = L.cout;
    g.dw = u->d_g + L.w_off; g.n = u->n; g.h = h; g.w = w;
    long long sp = P / 2048;
    g.splits = (int)(sp < 1 ? 1 : (sp > 1024 ? 1024 : sp));
    // the masked gradient in BF16 feeds the weight gradient and the data
    // gradient; the bias gradient is summed by the tiled weight gradient, or
    // from per-workgroup partials of the masked gradient (split-K path)
    g.dyb = gm;
    g.db = u->d_g + L.b_off;
    WgradJob job{};
    size_t need = 0;
    const bool tiled = wgrad_plan(g, job, &need) && need <= kUnetPartFloats;
    UCHK(u, hipEventRecord(u->evMask[li], u->stream));
    // the data gradient first: it is recorded as the mask's first dependent,
    // so a replayed graph keeps the data-gradient chain on one queue
    if (gx1) {
        ConvArgs a{};
        a.x1 = gm; a.x2 = nullptr; a.m1 = nullptr; a.c1 = L.cout; a.c2 = 0;
        a.w_ = u->d_wf + L.wb_off; a.bias = nullptr;
        a.y1 = gx1; a.y2 = gx2; a.cy1 = L.c1; a.cy2 = L.c2; a.acc1 = acc1; a.acc2 = acc2; a.mode = 1;
        a.n = u->n; a.h = h; a.w = w;
        a.mo = next.gm; a.my = next.y; a.pg = pooledFrom;
        a.uo = up.gm; a.uy = up.y; a.ud = upScratch;
        UCHK(u, launch_conv3x3(a, false, u->stream));
    }
    int rc;
    if (tiled) {
        if (u->batch.nj == kWgradBatchMax || u->batchFloats + need > kUnetPartFloats) {
            if ((rc = flush_wgrad(u))) return rc;
        }
        job.a.part = u->d_part + u->batchFloats;
        job.a.partCap = kUnetPartFloats - u->batchFloats;
        u->batch.job[u->batch.nj++] = job;
        u->batchFloats += need;
        u->batchLast = li;
        return batch_ends_at(li) ? flush_wgrad(u) : WC_OK;
    }
    // split-K path, in layer order after the batch so far
    if ((rc = flush_wgrad(u))) return rc;
    float *bpart = u->d_bpart + li * kBiasPartFloats;
    UCHK(u, hipStreamWaitEvent(u->side, u->evMask[li], 0));
    UCHK(u, launch_finish_sums(u->pending, u->side));
    u->pending = SideSums{};
    int nb = 0;
    UCHK(u, launch_mask_bf16(nullptr, nullptr, P * L.cout, L.cout, gm, bpart, kBiasPartFloats, &nb, SideSums{},
                             u->side));
    g.part = nullptr;               // the atomic path
    g.bpart = bpart;
    g.bnb = nb;
    UCHK(u, launch_wgrad(g, u->side));
    UCHK(u, hipEventRecord(u->evWgrad[li], u->side));
    return WC_OK;
}

int forward(wc_unet *u) {
    const int H = u->h, W = u->w, C = kBase, n = u->n;
    const auto &Ls = u->layers;
    int rc;
    UCHK(u, launch_unet_input(u->cin > 1 ? u->xc : u->x, (long long)n * H * W, H * W, u->cin, C, u->in0,
                              u->cin > 1 ? u->x : nullptr, u->stream));
    if ((rc = conv_fwd(u, Ls[0], u->in0, nullptr, u->a1, H, W)) || (rc = conv_fwd(u, Ls[1], u->a1, nullptr, u->s1, H, W)))
        return rc;
    UCHK(u, launch_pool_fwd(u->s1, n, H, W, C, u->p1, u->stream));
    if ((rc = conv_fwd(u, Ls[2], u->p1, nullptr, u->a2, H / 2, W / 2)) ||
        (rc = conv_fwd(u, Ls[3], u->a2, nullptr, u->s2, H / 2, W / 2)))
        return rc;
    UCHK(u, launch_pool_fwd(u->s2, n, H / 2, W / 2, 2 * C, u->p2, u->stream));
    if ((rc = conv_fwd(u, Ls[4], u->p2, nullptr, u->a3, H / 4, W / 4)) ||
        (rc = conv_fwd(u, Ls[5], u->a3, nullptr, u->s3, H / 4, W / 4)))
        return rc;
    UCHK(u, launch_pool_fwd(u->s3, n, H / 4, W / 4, 4 * C, u->p3, u->stream));
    if ((rc = conv_fwd(u, Ls[6], u->p3, nullptr, u->a4, H / 8, W / 8)) ||
        (rc = conv_fwd(u, Ls[7], u->a4, nullptr, u->b4, H / 8, W / 8)))
        return rc;
    UCHK(u, launch_up_fwd(u->b4, n, H / 8, W / 8, 8 * C, u->u3, u->stream));
    if ((rc = conv_fwd(u, Ls[8], u->s3, u->u3, u->d3a, H / 4, W / 4)) ||
        (rc = conv_fwd(u, Ls[9], u->d3a, nullptr, u->d3, H / 4, W / 4)))
        return rc;
    UCHK(u, launch_up_fwd(u->d3, n, H / 4, W / 4, 4 * C, u->u2, u->stream));
    if ((rc = conv_fwd(u, Ls[10], u->s2, u->u2, u->d2a, H / 2, W / 2)) ||
        (rc = conv_fwd(u, Ls[11], u->d2a, nullptr, u->d2, H / 2, W / 2)))
        return rc;
    UCHK(u, launch_up_fwd(u->d2, n, H / 2, W / 2, 2 * C, u->u1, u->stream));
    if ((rc = conv_fwd(u, Ls[12], u->s1, u->u1, u->d1a, H, W)) || (rc = conv_fwd(u, Ls[13], u->d1a, nullptr, u->d1, H, W)))
        return rc;
    UCHK(u, launch_head_fwd(u->d1, u->x, u->d_p + u->head_off, (long long)n * H * W, u->out, u->stream));
    return WC_OK;
}

// gradients of the mean squared error into d_g; loss in d_loss
int backward(wc_unet *u) {
    const int H = u->h, W = u->w, C = kBase, n = u->n;
    const auto &Ls = u->layers;
    int rc;
    u->pending = SideSums{};
    u->batch = WgradJobs{};
    u->batchFloats = 0;
    u->batchLast = -1;
    // every gradient and the loss are written (not accumulated) by the
    // launches below: no zeroing first
    const long long P0 = (long long)n * H * W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
    // each layer's masked output gradient is written by the kernel that
    // produces the gradient: the head, a data-gradient convolution, an
    // upsampling's or a pooling's backward
    MaskOut m{};
    int hb = 0;
    if ((rc = mask_of(u, 13, u->d1, P0, m))) return rc;
    UCHK(u, launch_head_bwd(u->d1, u->out, u->tgt, u->useWgt ? u->wgt : nullptr, u->d_p + u->head_off, P0, u->g_d1,
                            m.gm, u->d_hpart, kHeadPartFloats, &hb, u->stream));
    // the head's weight and bias gradients and the loss: finished on the
    // side stream before the first weight gradient
    u->pending.job[0] = PartialSums{u->d_hpart, hb, kHeadPartials - 1, kHeadPartials, u->d_g + u->head_off, 1.0f};
    u->pending.job[1] = PartialSums{u->d_hpart + kHeadPartials - 1, hb, 1, kHeadPartials, u->d_loss, 1.0f};
    // decoder, level 0
    if ((rc = mask_of(u, 12, u->d1a, P0, m)) ||
        (rc = conv_bwd(u, Ls[13], u->d1a, nullptr, u->g_d1a, false, nullptr, false, H, W, m)) ||
        (rc = mask_of(u, 11, u->d2, P1, m)) ||
        (rc = conv_bwd(u, Ls[12], u->s1, u->u1, u->g_s1, false, u->g_u1, false, H, W, MaskOut{}, nullptr, m, u->g_d2)))
        return rc;
    // level 1
    if ((rc = mask_of(u, 10, u->d2a, P1, m)) ||
        (rc = conv_bwd(u, Ls[11], u->d2a, nullptr, u->g_d2a, false, nullptr, false, H / 2, W / 2, m)) ||
        (rc = mask_of(u, 9, u->d3, P2, m)) ||
        (rc = conv_bwd(u, Ls[10], u->s2, u->u2, u->g_s2, false, u->g_u2, false, H / 2, W / 2, MaskOut{}, nullptr, m,
                       u->g_d3)))
        return rc;
    // level 2
    if ((rc = mask_of(u, 8, u->d3a, P2, m)) ||
        (rc = conv_bwd(u, Ls[9], u->d3a, nullptr, u->g_d3a, false, nullptr, false, H / 4, W / 4, m)) ||
        (rc = mask_of(u, 7, u->b4, P3, m)) ||
        (rc = conv_bwd(u, Ls[8], u->s3, u->u3, u->g_s3, false, u->g_u3, false, H / 4, W / 4, MaskOut{}, nullptr, m,
                       u->g_b4)))
        return rc;
    // bottleneck
    if ((rc = mask_of(u, 6, u->a4, P3, m)) ||
        (rc = conv_bwd(u, Ls[7], u->a4, nullptr, u->g_a4, false, nullptr, false, H / 8, W / 8, m)) ||
        (rc = mask_of(u, 5, u->s3, P2, m)) ||
        (rc = conv_bwd(u, Ls[6], u->p3, nullptr, u->g_p3, false, nullptr, false, H / 8, W / 8, m, u->g_s3)))
        return rc;
    // encoder
    if ((rc = mask_of(u, 4, u->a3, P2, m)) ||
        (rc = conv_bwd(u, Ls[5], u->a3, nullptr, u->g_a3, false, nullptr, false, H / 4, W / 4, m)) ||
        (rc = mask_of(u, 3, u->s2, P1, m)) ||
        (rc = conv_bwd(u, Ls[4], u->p2, nullptr, u->g_p2, false, nullptr, false, H / 4, W / 4, m, u->g_s2)))
        return rc;
    if ((rc = mask_of(u, 2, u->a2, P1, m)) ||
        (rc = conv_bwd(u, Ls[3], u->a2, nullptr, u->g_a2, false, nullptr, false, H / 2, W / 2, m)) ||
        (rc = mask_of(u, 1, u->s1, P0, m)) ||
        (rc = conv_bwd(u, Ls[2], u->p1, nullptr, u->g_p1, false, nullptr, false, H / 2, W / 2, m, u->g_s1)))
        return rc;
    if ((rc = mask_of(u, 0, u->a1, P0, m)) ||
        (rc = conv_bwd(u, Ls[1], u->a1, nullptr, u->g_a1, false, nullptr, false, H, W, m)) ||
        (rc = conv_bwd(u, Ls[0], u->in0, nullptr, nullptr, false, nullptr, false, H, W)))
        return rc;
    // the weight gradients ran in layer order on the side stream: the last one closes it
    if ((rc = flush_wgrad(u))) return rc;
    UCHK(u, hipStreamWaitEvent(u->stream, u->evWgrad[0], 0));
    return WC_OK;
}

// Captures body(u) on the main stream into a graph (`what` names it in errors).
int capture(wc_unet *u, int (*body)(wc_unet *), hipGraphExec_t &out, const char *what) {
    UCHK(u, hipStreamBeginCapture(u->stream, hipStreamCaptureModeThreadLocal));
    int rc = body(u);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(u->stream, &g);        // ends the capture on every path
    if (!rc && e != hipSuccess) rc = ufail(u, WC_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    if (!rc && hipGraphInstantiate(&out, g, nullptr, nullptr, 0) != hipSuccess) {
        out = nullptr;
        rc = ufail(u, WC_ERR_HIP, "hipGraphInstantiate of the %s failed", what);
    }
    if (g) hipGraphDestroy(g);
    return rc;
}

// The batch's inputs (and the loss kind): into the per-shape buffers on the
// main stream, before the forward pass.
int upload_inputs(wc_unet *u, const float *x, const float *wgt, int n, int h, int w) {
    int rc = ensure_shape(u, n, h, w);
    if (rc) return rc;
    const size_t P = (size_t)n * h * w;
    // the previous call's copies out of the staging are complete: every
    // call ends with the main stream synchronized, after the side stream's copies
    std::memcpy(u->h_stage, x, P * u->cin * sizeof(float));
    UCHK(u, hipMemcpyAsync(u->cin > 1 ? u->xc : u->x, u->h_stage, P * u->cin * 4, hipMemcpyHostToDevice, u->stream));
    u->useWgt = wgt != nullptr;
    return WC_OK;
}

// The targets (and loss weights), needed from the head's backward on: copied
// on the side stream while the forward pass runs; the main stream waits.
int upload_targets(wc_unet *u, const float *t, const float *wgt) {
    const size_t P = (size_t)u->n * u->h * u->w;
    float *st = u->h_stage + P * u->cin, *sw = st + P;
    std::memcpy(st, t, P * sizeof(float));
    UCHK(u, hipMemcpyAsync(u->tgt, st, P * 4, hipMemcpyHostToDevice, u->side));
    if (u->useWgt) {
        std::memcpy(sw, wgt, P * sizeof(float));
        UCHK(u, hipMemcpyAsync(u->wgt, sw, P * 4, hipMemcpyHostToDevice, u->side));
    }
    UCHK(u, hipEventRecord(u->evUp, u->side));
    UCHK(u, hipStreamWaitEvent(u->stream, u->evUp, 0));
    return WC_OK;
}

// forward(), the target upload, then backward() on the current batch.  With
// useGraph the two passes are captured into graphs the first time a batch
// shape (and loss kind) is seen and replayed afterwards: the same kernels
// with the same arguments, so the results are those of direct launches.  The
// batch buffers are fixed per shape and the uploads write into them.
int forward_backward(wc_unet *u, const float *t, const float *wgt) {
    int rc;
    if (!u->useGraph) {
        if ((rc = forward(u)) || (rc = upload_targets(u, t, wgt))) return rc;
        return backward(u);
    }
    if (!u->gexec[0] || u->gN != u->n || u->gH != u->h || u->gW != u->w || u->gWgt != u->useWgt) {
        drop_graph(u);
        if ((rc = capture(u, forward, u->gexec[0], "forward pass")) ||
            (rc = capture(u, backward, u->gexec[1], "backward pass"))) {
            drop_graph(u);
            return rc;
        }
        u->gN = u->n; u->gH = u->h; u->gW = u->w; u->gWgt = u->useWgt;
    }
    UCHK(u, hipGraphLaunch(u->gexec[0], u->stream));
    if ((rc = upload_targets(u, t, wgt))) return rc;
    UCHK(u, hipGraphLaunch(u->gexec[1], u->stream));
    return WC_OK;
}

int refresh_bf16(wc_unet *u) {
    UCHK(u, launch_weights_bf16(u->d_p, u->wmap, u->d_wb, u->d_wf, u->stream));
    return WC_OK;
}

}  // namespace

extern "C" {

int wc_unet_create(int device, uint64_t seed, wc_unet **out) { return wc_unet_create_ex(device, seed, 1, out); }

int wc_unet_create_ex(int device, uint64_t seed, int channels, wc_unet **out) {
    if (!out) return WC_ERR_INVALID;
    *out = nullptr;
    if (channels < 1 || channels > kBase) return WC_ERR_INVALID;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return WC_ERR_NO_DEVICE;
    auto *u = new wc_unet();
    u->device = device;
    u->cin = channels;
    const int C = kBase;
    const int spec[14][3] = {{C, C, 0},          {C, C, 0},         {2 * C, C, 0},     {2 * C, 2 * C, 0},
                             {4 * C, 2 * C, 0},  {4 * C, 4 * C, 0}, {8 * C, 4 * C, 0}, {8 * C, 8 * C, 0},
                             {4 * C, 4 * C, 8 * C}, {4 * C, 4 * C, 0}, {2 * C, 2 * C, 4 * C}, {2 * C, 2 * C, 0},
                             {C, C, 2 * C},      {C, C, 0}};
    size_t off = 0, woff = 0;
    for (auto &s : spec) {
        Layer L{s[0], s[1], s[2], 0, 0, 0};
        const size_t nw = (size_t)L.cout * 9 * (L.c1 + L.c2);
        L.w_off = off;
        off += nw;
        L.b_off = off;
        off += L.cout;
        L.wb_off = woff;
        woff += nw;
        u->layers.push_back(L);
    }
    u->wmap.layers = (int)u->layers.size();
    for (int l = 0; l < u->wmap.layers; ++l) {
        const Layer &L = u->layers[l];
        u->wmap.cout[l] = L.cout;
        u->wmap.cin[l] = L.c1 + L.c2;
        u->wmap.w_off[l] = (long long)L.w_off;
        u->wmap.wb_off[l] = (long long)L.wb_off;
    }
    u->head_off = off;
    off += C + 1;
    u->nparam = off;
    u->nwb = woff;
    // He-normal initialisation (fan-in of the real inputs; the first layer
    // sees `channels` real channels), zero biases, small head weights
    std::vector<float> p(u->nparam, 0.0f);
    std::mt19937_64 rng(seed);
    for (size_t l = 0; l < u->layers.size(); ++l) {
        const Layer &L = u->layers[l];
        const int cin = L.c1 + L.c2, fan = 9 * (l == 0 ? channels : cin);
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / fan));
        for (int co = 0; co < L.cout; ++co)
            for (int t = 0; t < 9; ++t)
                for (int ci = 0; ci < cin; ++ci)
                    p[L.w_off + ((size_t)co * 9 + t) * cin + ci] = (l == 0 && ci >= channels) ? 0.0f : nd(rng);
    }
    std::normal_distribution<float> hd(0.0f, 0.01f);
    for (int c = 0; c < C; ++c) p[u->head_off + c] = hd(rng);
    auto bad = [&](int code) {
        wc_unet_destroy(u);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&u->stream) != hipSuccess ||
        hipStreamCreate(&u->side) != hipSuccess)
        return bad(WC_ERR_HIP);
    for (int l = 0; l < kMaxConvLayers; ++l)
        if (hipEventCreateWithFlags(&u->evMask[l], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&u->evWgrad[l], hipEventDisableTiming) != hipSuccess)
            return bad(WC_ERR_HIP);
    if (hipEventCreateWithFlags(&u->evUp, hipEventDisableTiming) != hipSuccess) return bad(WC_ERR_HIP);
    if (hipMalloc(&u->d_p, u->nparam * 4) || hipMalloc(&u->d_g, u->nparam * 4) || hipMalloc(&u->d_m, u->nparam * 4) ||
        hipMalloc(&u->d_v, u->nparam * 4) || hipMalloc(&u->d_wb, u->nwb * 2 + 256) ||
        hipMalloc(&u->d_wf, u->nwb * 2 + 256) || hipMalloc(&u->d_loss, 256) ||
        hipMalloc(&u->d_part, kUnetPartFloats * sizeof(float)) ||
        hipMalloc(&u->d_bpart, kMaxConvLayers * kBiasPartFloats * sizeof(float)) || hipMalloc(&u->d_hpart, kHeadPartFloats * sizeof(float)))
        return bad(WC_ERR_NOMEM);
    if (hipMemcpy(u->d_p, p.data(), u->nparam * 4, hipMemcpyHostToDevice) != hipSuccess) return bad(WC_ERR_HIP);
    if (launch_fill_f32(u->d_m, (long long)u->nparam, 0.f, u->stream) != hipSuccess ||
        launch_fill_f32(u->d_v, (long long)u->nparam, 0.f, u->stream) != hipSuccess || refresh_bf16(u) ||
        hipStreamSynchronize(u->stream) != hipSuccess)
        return bad(WC_ERR_HIP);
    *out = u;
    return WC_OK;
}

int wc_unet_destroy(wc_unet *u) {
    if (!u) return WC_OK;
    hipSetDevice(u->device);
    if (u->stream) hipStreamSynchronize(u->stream);
    if (u->side) hipStreamSynchronize(u->side);
    free_shape(u);
    hipFree(u->d_p); hipFree(u->d_g); hipFree(u->d_m); hipFree(u->d_v); hipFree(u->d_wb); hipFree(u->d_wf);
    hipFree(u->d_loss); hipFree(u->d_part); hipFree(u->d_bpart); hipFree(u->d_hpart);
    for (int l = 0; l < kMaxConvLayers; ++l) {
        if (u->evMask[l]) hipEventDestroy(u->evMask[l]);
        if (u->evWgrad[l]) hipEventDestroy(u->evWgrad[l]);
    }
    if (u->evUp) hipEventDestroy(u->evUp);
    if (u->side) hipStreamDestroy(u->side);
    if (u->stream) hipStreamDestroy(u->stream);
    delete u;
    return WC_OK;
}

const char *wc_unet_last_error(const wc_unet *u) { return u ? u->err.c_str() : "null network"; }

int wc_unet_set_graph(wc_unet *u, int on) {
    if (!u) return WC_ERR_INVALID;
    UCHK(u, hipSetDevice(u->device));
    UCHK(u, hipStreamSynchronize(u->stream));
    UCHK(u, hipStreamSynchronize(u->side));
    drop_graph(u);
    u->useGraph = on != 0;
    return WC_OK;
}

size_t wc_unet_param_count(const wc_unet *u) { return u ? u->nparam : 0; }

int wc_unet_get_params(wc_unet *u, float *p) {
    if (!u || !p) return WC_ERR_INVALID;
    UCHK(u, hipSetDevice(u->device));
    UCHK(u, hipMemcpyAsync(p, u->d_p, u->nparam * 4, hipMemcpyDeviceToHost, u->stream));
    UCHK(u, hipStreamSynchronize(u->stream));
    return WC_OK;
}

int wc_unet_set_params(wc_unet *u, const float *p) {
    if (!u || !p) return WC_ERR_INVALID;
    UCHK(u, hipSetDevice(u->device));
    UCHK(u, hipMemcpyAsync(u->d_p, p, u->nparam * 4, hipMemcpyHostToDevice, u->stream));
    int rc = refresh_bf16(u);
    if (rc) return rc;
    UCHK(u, launch_fill_f32(u->d_m, (long long)u->nparam, 0.f, u->stream));
    UCHK(u, launch_fill_f32(u->d_v, (long long)u->nparam, 0.f, u->stream));
    u->adam_t = 0;
    UCHK(u, hipStreamSynchronize(u->stream));
    return WC_OK;
}

int wc_unet_forward(wc_unet *u, const float *x, int n, int h, int w, float *y, double *device_ms) {
    if (!u || !x || !y) return WC_ERR_INVALID;
    UCHK(u, hipSetDevice(u->device));
    int rc = upload_inputs(u, x, nullptr, n, h, w);
    if (rc) return rc;
    hipEvent_t e0, e1;
    UCHK(u, hipEventCreate(&e0));
    UCHK(u, hipEventCreate(&e1));
    UCHK(u, hipEventRecord(e0, u->stream));
    rc = forward(u);
    if (!rc) {
        UCHK(u, hipEventRecord(e1, u->stream));
        UCHK(u, hipMemcpyAsync(y, u->out, (size_t)n * h * w * 4, hipMemcpyDeviceToHost, u->stream));
        UCHK(u, hipStreamSynchronize(u->stream));
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        u->fwd_ms = ms;
        if (device_ms) *device_ms = ms;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

int wc_unet_gradients(wc_unet *u, const float *x, const float *t, int n, int h, int w, float *grad, float *loss) {
    return wc_unet_gradients_weighted(u, x, t, nullptr, n, h, w, grad, loss);
}

int wc_unet_gradients_weighted(wc_unet *u, const float *x, const float *t, const float *weight, int n, int h, int w,
                               float *grad, float *loss) {
    if (!u || !x || !t) return WC_ERR_INVALID;
    UCHK(u, hipSetDevice(u->device));
    int rc = upload_inputs(u, x, weight, n, h, w);
    if (!rc) rc = forward_backward(u, t, weight);
    if (rc) return rc;
    if (grad) UCHK(u, hipMemcpyAsync(grad, u->d_g, u->nparam * 4, hipMemcpyDeviceToHost, u->stream));
    float l = 0;
    UCHK(u, hipMemcpyAsync(&l, u->d_loss, 4, hipMemcpyDeviceToHost, u->stream));
    UCHK(u, hipStreamSynchronize(u->stream));
    if (loss) *loss = l;
    return WC_OK;
}

int wc_unet_train_step(wc_unet *u, const float *x, const float *t, int n, int h, int w, float lr, float *loss) {
    return wc_unet_train_step_weighted(u, x, t, nullptr, n, h, w, lr, loss);
}

int wc_unet_train_step_weighted(wc_unet *u, const float *x, const float *t, const float *weight, int n, int h, int w,
                                float lr, float *loss) {
    if (!u || !x || !t || !(lr > 0)) return WC_ERR_INVALID;
    UCHK(u, hipSetDevice(u->device));
    int rc = upload_inputs(u, x, weight, n, h, w);
    if (!rc) rc = forward_backward(u, t, weight);
    if (rc) return rc;
    u->adam_t += 1;
    UCHK(u, launch_adam_bf16(u->d_p, u->d_g, u->d_m, u->d_v, (long long)u->nparam, lr, 0.9f, 0.999f, 1e-8f,
                             u->adam_t, u->wmap, u->d_wb, u->d_wf, u->stream));
    float l = 0;
    UCHK(u, hipMemcpyAsync(&l, u->d_loss, 4, hipMemcpyDeviceToHost, u->stream));
    UCHK(u, hipStreamSynchronize(u->stream));
    if (loss) *loss = l;
    return WC_OK;
}

}  // extern "C"
