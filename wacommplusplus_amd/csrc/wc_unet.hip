// Kernels of the concentration-reconstruction network (a 4-level U-Net with
// 32 base channels) for gfx950: training and inference on BF16 matrix cores
// with FP32 accumulation.  No counterpart in the reference.
//
// Layout: activations NHWC in BF16 (round to nearest even), activation
// gradients NHWC in FP32, weights [cout][3][3][cin] in BF16 (FP32 master
// copies live in the optimiser).  Every channel count is a multiple of 32.
//
// Convolutions are 3x3, stride 1, zero padding 1, as implicit GEMMs
//   out[p][co] = sum_{tap, ci} in[p + tap][ci] * w[co][tap][ci]
// on v_mfma_f32_16x16x32_bf16.  The same products serve
//   - forward: two inputs concatenated along channels (decoder skip + up),
//     bias + ReLU, BF16 out;
//   - input gradient: the ReLU-masked output gradient (BF16, written once
//     per layer by the kernel that produces it: the head's backward, the
//     next layer's input-gradient epilogue, or an upsampling's backward)
//     convolved with the flipped, transposed weights, FP32 out split over
//     the concat's two inputs, stored or accumulated, or the previous
//     layer's masked gradient directly (through a pooling's backward where
//     the layer's input was pooled);
//   - weight gradient: dW[co][tap][ci] = sum_p dYm[p][co] X[p + tap][ci], a
//     GEMM whose reduction runs over pixels.
// conv3x3_tile_kernel (forward, input gradient) and wgrad_tile_kernel
// (weight gradient) stage their operands in LDS; conv3x3_kernel (operands
// straight from global memory) and wgrad_kernel (split-K with FP32 atomics)
// take the shapes whose stage does not fit.
// MFMA operand layout (16x16x32 bf16): lane l holds A[l%16][8*(l/16)..+7]
// and B[8*(l/16)..+7][l%16]; the accumulator holds D[4*(l/16)+r][l%16],
// r = 0..3.
#include "wc_launch.h"

#pragma clang fp contract(on)

namespace wc {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float x) {
    const uint32_t b = __float_as_uint(x);
    if ((b & 0x7F800000u) == 0x7F800000u) return (uint16_t)((b >> 16) | ((b & 0xFFFF) ? 0x40 : 0));   // inf/nan
    return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

union Pack8 {
    uint4 u;
    uint16_t h[8];
    bf16x8 v;
};

// bf16(y > 0 ? g : 0) of four channels (y: four BF16 values in a uint2)
__device__ __forceinline__ uint2 mask4(const float4 &g, const uint2 &y) {
    const uint16_t h0 = to_bf16(from_bf16((uint16_t)y.x) > 0.f ? g.x : 0.f);
    const uint16_t h1 = to_bf16(from_bf16((uint16_t)(y.x >> 16)) > 0.f ? g.y : 0.f);
    const uint16_t h2 = to_bf16(from_bf16((uint16_t)y.y) > 0.f ? g.z : 0.f);
    const uint16_t h3 = to_bf16(from_bf16((uint16_t)(y.y >> 16)) > 0.f ? g.w : 0.f);
    return make_uint2((uint32_t)h0 | (uint32_t)h1 << 16, (uint32_t)h2 | (uint32_t)h3 << 16);
}

// A PartialSums job by one workgroup of 256, in a fixed order: thread
// c + m s sums the partials of workgroups s, s + S, s + 2S, ... (S = 256 / m)
// of value c into four running sums (every fourth term each, so four loads
// are in flight), then thread c adds the S segment sums in order.  The
// partials come from an earlier launch on the stream.
__device__ void fixed_order_sum(const PartialSums &j) {
    __shared__ float seg[256];
    const int m = j.m, S = 256 / m, t = threadIdx.x;
    float s = 0.f;
    if (t < m * S) {
        const float *p = j.part + t % m;
        const size_t st = (size_t)S * j.stride;
        int b = t / m;
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
        for (; b + 3 * S < j.nb; b += 4 * S) {
            const float *q = p + (size_t)b * j.stride;
            s0 += q[0]; s1 += q[st]; s2 += q[2 * st]; s3 += q[3 * st];
        }
        for (; b < j.nb; b += S) s0 += p[(size_t)b * j.stride];
        s = (s0 + s1) + (s2 + s3);
    }
    seg[t] = s;
    __syncthreads();
    if (t < m) {
        float tot = 0.f;
        for (int q = 0; q < S; ++q) tot += seg[q * m + t];
        j.out[t] = j.scale * tot;
    }
}


// 8 consecutive channels of one pixel as BF16 (zero when `valid` is false),
// optionally masked by the BF16 ReLU output m (value kept where m > 0).
template <class In>
__device__ __forceinline__ bf16x8 load8(const In *p, const uint16_t *m, bool valid);

template <>
__device__ __forceinline__ bf16x8 load8<uint16_t>(const uint16_t *p, const uint16_t *m, bool valid) {
    Pack8 r;
    r.u = valid ? *(const uint4 *)p : make_uint4(0, 0, 0, 0);
    if (m && valid) {
        Pack8 mm;
        mm.u = *(const uint4 *)m;
        for (int q = 0; q < 8; ++q)
            if (!(from_bf16(mm.h[q]) > 0.f)) r.h[q] = 0;
    }
    return r.v;
}

template <>
__device__ __forceinline__ bf16x8 load8<float>(const float *p, const uint16_t *m, bool valid) {
    Pack8 r;
    if (!valid) {
        r.u = make_uint4(0, 0, 0, 0);
        return r.v;
    }
    const float4 a = *(const float4 *)p, b = *(const float4 *)(p + 4);
    const float f[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    Pack8 mm;
    if (m) mm.u = *(const uint4 *)m;
    for (int q = 0; q < 8; ++q) r.h[q] = (m && !(from_bf16(mm.h[q]) > 0.f)) ? (uint16_t)0 : to_bf16(f[q]);
    return r.v;
}

template <class In>
__global__ __launch_bounds__(256) void conv3x3_kernel(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int hw = a.h * a.w;
    const long long P = (long long)a.n * hw;
    const int nout = a.cy1 + a.cy2;
    // workgroup tile: 128 pixels (4 waves x 32) x 32 output channels
    const long long p0 = (long long)blockIdx.x * 128 + wave * 32;
    const int co0 = blockIdx.y * 32;
    const int cin = a.c1 + a.c2;
    const int row = lane & 15, kq = lane >> 4;      // MFMA row/column within a 16-tile, k-group
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // the two pixels this lane feeds as A rows
    int py[2], px[2], pn[2];
    bool pv[2];
    for (int i = 0; i < 2; ++i) {
        const long long p = p0 + i * 16 + row;
        pv[i] = p < P;
        const long long pc = pv[i] ? p : 0;
        pn[i] = (int)(pc / hw);
        const int r = (int)(pc - (long long)pn[i] * hw);
        py[i] = r / a.w;
        px[i] = r - py[i] * a.w;
    }
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        long long q[2];            // source pixel of each A row (0 when padding)
        bool inb[2];
        for (int i = 0; i < 2; ++i) {
            const int y = py[i] + dy, x = px[i] + dx;
            inb[i] = pv[i] && y >= 0 && y < a.h && x >= 0 && x < a.w;
            q[i] = inb[i] ? ((long long)pn[i] * hw + (long long)y * a.w + x) : 0;
        }
        for (int c0 = 0; c0 < cin; c0 += 32) {
            const bool first = c0 < a.c1;
            const In *base = (const In *)(first ? a.x1 : a.x2);
            const int cc = first ? a.c1 : a.c2;
            const int cl = (first ? c0 : c0 - a.c1) + kq * 8;      // channel within its tensor
            const uint16_t *mb = first ? a.m1 : nullptr;
            bf16x8 A[2];
            for (int i = 0; i < 2; ++i)
                A[i] = load8<In>(base + q[i] * cc + cl, mb ? mb + q[i] * cc + cl : nullptr, inb[i]);
            bf16x8 B[2];
            for (int j = 0; j < 2; ++j) {
                const int co = co0 + j * 16 + row;
                Pack8 w;
                w.u = co < nout ? *(const uint4 *)(a.w_ + ((size_t)co * 9 + tap) * cin + c0 + kq * 8)
                                : make_uint4(0, 0, 0, 0);
                B[j] = w.v;
            }
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[i], B[j], acc[i][j], 0, 0, 0);
        }
    }
    // epilogue: D[4*kq + r][row] of tile (i, j) = pixel p0 + 16 i + 4 kq + r, channel co0 + 16 j + row
    for (int j = 0; j < 2; ++j) {
        const int co = co0 + j * 16 + row;
        if (co >= nout) continue;
        const bool first = co < a.cy1;
        const int cl = first ? co : co - a.cy1;
        const int cc = first ? a.cy1 : a.cy2;
        for (int i = 0; i < 2; ++i)
            for (int r = 0; r < 4; ++r) {
                const long long p = p0 + i * 16 + kq * 4 + r;
                if (p >= P) continue;
                float v = acc[i][j][r];
                if (a.mode == 0) {                                  // forward: bias + ReLU, BF16
                    v += a.bias[co];
                    ((uint16_t *)a.y1)[p * a.cy1 + co] = to_bf16(v > 0.f ? v : 0.f);
                } else {
                    float *y = (float *)(first ? a.y1 : a.y2) + p * cc + cl;
                    const bool accum = first ? a.acc1 : a.acc2;
                    *y = accum ? *y + v : v;
                }
            }
    }
}

// LDS-tiled 3x3 convolution (BF16 input, no input mask).  A workgroup owns R
// whole rows of one image (R * w <= 256 pixels) and 32 output channels.  Per
// 32-channel input chunk the workgroup stages in LDS
//   sx[(rr, xx)][40]   input rows y0-1 .. y0+R, columns -1 .. w (zero outside
//                      the image), 32 channels padded to 40 (80 B per pixel:
//                      16-byte aligned, banks spread);
//   sw[tap][co][40]    the chunk's weights of the 32 output channels (WL;
//                      otherwise each wave keeps them in registers, for
//                      the wide rows whose stage leaves no room);
// every work-item issues up to kStageBatch 16-byte loads before it stores
// any, so a stage costs about one memory latency, not one per load.  Then
// each wave computes 16 NI pixels (NI MFMA row tiles) x 32 channels: per tap
// NI A and 2 B reads from LDS feed 2 NI MFMAs.  NI (4, 2 or 1) is the
// smallest that gives the four waves all R * w pixels, so the small images
// of the deep levels keep every wave busy.  The epilogue is the one of
// conv3x3_kernel.
constexpr int kCvPS = 40;
constexpr int kCvOutF = 36;               // FP32 epilogue row: 32 channels padded to 144 B
constexpr int kStageBatch = 8;
constexpr int kCvWRows = 9 * 32;          // staged weight rows per chunk

// Epilogue through LDS: D[4 kq + r][row] of tile (i, j) is stage pixel
// q = 16 NI wave + 16 i + 4 kq + r, channel co0 + 16 j + row.  The tile
// [q][32 channels] is written to LDS (the caller has synchronised after the
// last read of the stage), then stored with 16-byte accesses, 8 BF16 or 4
// FP32 channels of one pixel per access.  The stage holds whole rows, so
// stage pixel q is image pixel y0 w + q: stored while q < rv w (rows past the
// image's last are not).  The tile stays within the input stage's part of
// LDS.  Returns after the last LDS read.
template <int NI>
__device__ __forceinline__ void conv_epilogue(const ConvArgs &a, const f32x4 (&acc)[NI][2], uint16_t *smem, int rv,
                                              int img, int y0, int co0, bool active) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, row = lane & 15, kq = lane >> 4;
    const int W = a.w, nout = a.cy1 + a.cy2;
    const int qmax = rv * W;                              // stage pixels inside the image
    const long long p0 = (long long)img * a.h * W + (long long)y0 * W;
    if (a.mode == 0) {                                    // forward: bias + ReLU, BF16
        uint16_t *so = smem;                              // [q][kCvPS]
        if (active)
            for (int j = 0; j < 2; ++j) {
                const int col = 16 * j + row;
                const float b = co0 + col < nout ? a.bias[co0 + col] : 0.f;
                for (int i = 0; i < NI; ++i)
                    for (int r = 0; r < 4; ++r) {
                        const int q = 16 * NI * wave + 16 * i + 4 * kq + r;
                        const float v = acc[i][j][r] + b;
                        if (q < qmax) so[q * kCvPS + col] = to_bf16(v > 0.f ? v : 0.f);
                    }
            }
        __syncthreads();
        uint16_t *y = (uint16_t *)a.y1 + p0 * a.cy1 + co0;
        for (int c = tid; c < qmax * 4; c += 256) {
            const int q = c >> 2, part = c & 3;
            if (co0 + 8 * part < nout) *(uint4 *)(y + (size_t)q * a.cy1 + 8 * part) = *(const uint4 *)(so + q * kCvPS + 8 * part);
        }
    } else {                                              // input gradient: FP32, stored or accumulated
        float *so = (float *)smem;                        // [q][kCvOutF]
        if (active)
            for (int j = 0; j < 2; ++j)
                for (int i = 0; i < NI; ++i)
                    for (int r = 0; r < 4; ++r) {
                        const int q = 16 * NI * wave + 16 * i + 4 * kq + r;
                        if (q < qmax) so[q * kCvOutF + 16 * j + row] = acc[i][j][r];
                    }
        __syncthreads();
        const bool first = co0 < a.cy1;
        const int cc = first ? a.cy1 : a.cy2, cl0 = first ? co0 : co0 - a.cy1;
        if (a.uo && !first) {
            // the second input was the 2x nearest upsampling of the previous
            // layer's output (uy, half the height and width): each of its
            // pixels gets the sum of the four it fed, masked into uo (the
            // stage's rows pair up: R, h and w are even)
            const int Wl = W >> 1, ql = (rv >> 1) * Wl;
            const size_t lp0 = ((size_t)img * (a.h >> 1) + (y0 >> 1)) * Wl;
            for (int c = tid; c < ql * 8; c += 256) {
                const int q = c >> 3, part = c & 7;
                if (co0 + 4 * part >= nout) continue;
                const int lr = q / Wl, lx = q - lr * Wl, q00 = 2 * lr * W + 2 * lx;
                const float4 v0 = *(const float4 *)(so + q00 * kCvOutF + 4 * part);
                const float4 v1 = *(const float4 *)(so + (q00 + 1) * kCvOutF + 4 * part);
                const float4 v2 = *(const float4 *)(so + (q00 + W) * kCvOutF + 4 * part);
                const float4 v3 = *(const float4 *)(so + (q00 + W + 1) * kCvOutF + 4 * part);
                const float4 g = make_float4(v0.x + v1.x + v2.x + v3.x, v0.y + v1.y + v2.y + v3.y,
                                             v0.z + v1.z + v2.z + v3.z, v0.w + v1.w + v2.w + v3.w);
                const size_t o = (lp0 + q) * cc + cl0 + 4 * part;
                *(uint2 *)(a.uo + o) = mask4(g, *(const uint2 *)(a.uy + o));
            }
            return;
        }
        if (a.mo && a.pg) {
            // the input was the 2x2 average pool of the previous layer's
            // output: each of its four pixels gets pg + g / 4, masked into mo
            // (pg, mo, my at twice the height and width)
            const int W2 = 2 * W;
            for (int c = tid; c < qmax * 8; c += 256) {
                const int q = c >> 3, part = c & 7;
                if (co0 + 4 * part >= nout) continue;
                const float4 v = *(const float4 *)(so + q * kCvOutF + 4 * part);
                const float4 g = make_float4(0.25f * v.x, 0.25f * v.y, 0.25f * v.z, 0.25f * v.w);
                const int r = q / W, x = q - r * W;
                const size_t pix00 = ((size_t)img * 2 * a.h + 2 * (y0 + r)) * W2 + 2 * x;
                for (int k = 0; k < 4; ++k) {
                    const size_t o = (pix00 + (size_t)(k >> 1) * W2 + (k & 1)) * cc + cl0 + 4 * part;
                    const float4 s = *(const float4 *)(a.pg + o);
                    *(uint2 *)(a.mo + o) =
                        mask4(make_float4(s.x + g.x, s.y + g.y, s.z + g.z, s.w + g.w), *(const uint2 *)(a.my + o));
                }
            }
            return;
        }
        if (a.mo) {                                       // the next layer's masked gradient (one output, stored)
            const uint16_t *my = a.my + p0 * cc + cl0;
            uint16_t *mo = a.mo + p0 * cc + cl0;
            for (int c = tid; c < qmax * 8; c += 256) {
                const int q = c >> 3, part = c & 7;
                if (co0 + 4 * part >= nout) continue;
                const float4 v = *(const float4 *)(so + q * kCvOutF + 4 * part);
                *(uint2 *)(mo + (size_t)q * cc + 4 * part) = mask4(v, *(const uint2 *)(my + (size_t)q * cc + 4 * part));
            }
            return;
        }
        const bool accum = first ? a.acc1 : a.acc2;
        float *y = (float *)(first ? a.y1 : a.y2) + p0 * cc + cl0;
        for (int c = tid; c < qmax * 8; c += 256) {
            const int q = c >> 3, part = c & 7;
            if (co0 + 4 * part >= nout) continue;
            float4 v = *(const float4 *)(so + q * kCvOutF + 4 * part);
            float4 *dst = (float4 *)(y + (size_t)q * cc + 4 * part);
            if (accum) {
                const float4 o = *dst;
                v.x += o.x; v.y += o.y; v.z += o.z; v.w += o.w;
            }
            *dst = v;
        }
    }
}

template <int NI, bool WL>
__global__ __launch_bounds__(256) void conv3x3_tile_kernel(ConvArgs a, int R, int rowsPerImg) {
    extern __shared__ __align__(16) uint16_t smem[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int W = a.w, H = a.h, hw = H * W, Wp = W + 2;
    const int img = blockIdx.x / rowsPerImg, y0 = (blockIdx.x - img * rowsPerImg) * R;
    const int rv = min(R, H - y0), RW = R * W;
    const int co0 = blockIdx.y * 32;
    const int nout = a.cy1 + a.cy2, cin = a.c1 + a.c2;
    const int npos = (R + 2) * Wp;
    const uint16_t *sx = smem;                            // [(R+2) * (W+2)][kCvPS]
    const uint16_t *sw = smem + npos * kCvPS;             // [9][32][kCvPS]
    const int row = lane & 15, kq = lane >> 4;
    // this lane's A rows: stage pixel 16 NI wave + 16 i + row (its centre-tap
    // LDS offset; pixels past the stage read pixel 0 and are not stored)
    int base[NI];
    for (int i = 0; i < NI; ++i) {
        const int q = 16 * NI * wave + 16 * i + row;
        const int r = q < RW ? q / W : 0, x = q < RW ? q - r * W : 0;
        base[i] = ((r + 1) * Wp + (x + 1)) * kCvPS + 8 * kq;
    }
    const bool active = 16 * NI * wave < RW;
    f32x4 acc[NI][2];
    for (int i = 0; i < NI; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nx = npos * 4, nt = nx + (WL ? kCvWRows * 4 : 0);     // 16-byte groups: input, then weights
    // the staging position of group tid (row rr0, column xx0 of the stage);
    // later groups advance by 64 positions, followed by counting
    const int rr0 = (tid >> 2) / Wp, xx0 = (tid >> 2) - rr0 * Wp, g0 = tid & 3;
    for (int c0 = 0; c0 < cin; c0 += 32) {
        const bool first = c0 < a.c1;
        const int cc = first ? a.c1 : a.c2, cl0 = first ? c0 : c0 - a.c1;
        // this image's rows (offsets below stay within one image: h w c < 2^31)
        const uint16_t *xs = (const uint16_t *)(first ? a.x1 : a.x2) + (size_t)img * hw * cc + cl0 + 8 * g0;
        const uint16_t *ws = a.w_ + (size_t)co0 * 9 * cin + c0 + 8 * g0;
        __syncthreads();                                  // the previous chunk's reads are done
        constexpr int SB = WL ? kStageBatch : kStageBatch / 2;   // the register-weight kernel keeps fewer
        int rr = rr0, xx = xx0;
        for (int e0 = tid; e0 < nt; e0 += 256 * SB) {
            uint4 v[SB];
            int dst[SB];
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                const int e = e0 + 256 * u;
                v[u] = make_uint4(0, 0, 0, 0);
                dst[u] = -1;
                if (e < nx) {
                    const int y = y0 - 1 + rr, x = xx - 1;
                    dst[u] = (e >> 2) * kCvPS + 8 * g0;
                    if (y >= 0 && y < H && x >= 0 && x < W) v[u] = *(const uint4 *)(xs + (y * W + x) * cc);
                } else if (e < nt) {
                    const int tc = (e - nx) >> 2;                           // tap * 32 + local channel
                    const int col = tc & 31, tap = tc >> 5;
                    dst[u] = (npos + tc) * kCvPS + 8 * g0;
                    if (co0 + col < nout) v[u] = *(const uint4 *)(ws + (col * 9 + tap) * cin);
                }
                for (xx += 64; xx >= Wp; xx -= Wp) ++rr;
            }
#pragma unroll
            for (int u = 0; u < SB; ++u)
                if (dst[u] >= 0) *(uint4 *)(smem + dst[u]) = v[u];
        }
        // !WL: B of the 9 taps in registers (loaded after the stage, so the
        // two do not hold registers at once), channels c0 + 8 kq .. +7 of
        // output channels co0 + 16 j + row
        Pack8 Br[WL ? 1 : 9][2];
        if (!WL && active)
            for (int tap = 0; tap < (WL ? 1 : 9); ++tap)
                for (int j = 0; j < 2; ++j) {
                    const int co = co0 + 16 * j + row;
                    Br[tap][j].u = co < nout ? *(const uint4 *)(a.w_ + ((size_t)co * 9 + tap) * cin + c0 + 8 * kq)
                                             : make_uint4(0, 0, 0, 0);
                }
        __syncthreads();
        if (active)
            for (int tap = 0; tap < 9; ++tap) {
                const int off = ((tap / 3 - 1) * Wp + (tap % 3 - 1)) * kCvPS;
                Pack8 B[2];
                for (int j = 0; j < 2; ++j)
                    B[j].u = WL ? *(const uint4 *)(sw + (tap * 32 + 16 * j + row) * kCvPS + 8 * kq)
                                : Br[WL ? 0 : tap][j].u;
                for (int i = 0; i < NI; ++i) {
                    Pack8 A;
                    A.u = *(const uint4 *)(sx + base[i] + off);
                    for (int j = 0; j < 2; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.v, B[j].v, acc[i][j], 0, 0, 0);
                }
            }
    }
    __syncthreads();                                      // every wave is done with the stage
    conv_epilogue<NI>(a, acc, smem, rv, img, y0, co0, active);
}

// dW[co][tap][ci] += sum_p dY[p][co] * (Y[p][co] > 0) * X[p + tap][ci];
// X = concat(x1, x2) along channels.  Grid: (cout/32, cin/32, 9 * splits).
__global__ __launch_bounds__(256) void wgrad_kernel(WgradArgs a) {
    __shared__ uint16_t sdy[32][64 + 8];      // [co][pixel] (padded rows)
    __shared__ uint16_t sx[32][64 + 8];       // [ci][pixel]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int co0 = blockIdx.x * 32, ci0 = blockIdx.y * 32;
    const int tap = blockIdx.z % 9, split = blockIdx.z / 9;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int hw = a.h * a.w;
    const long long P = (long long)a.n * hw;
    const long long chunk = (P + a.splits - 1) / a.splits;
    const long long pbeg = split * chunk, pend = pbeg + chunk < P ? pbeg + chunk : P;
    const int cin = a.c1 + a.c2;
    const int row = lane & 15, kq = lane >> 4;
    const int ti = wave >> 1, tj = wave & 1;   // this wave's 16x16 tile (co block, ci block)
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long long pb = pbeg; pb < pend; pb += 64) {
        // stage 64 pixels x 32 channels of dYm and of the shifted X
        for (int e = tid; e < 64 * 32; e += 256) {
            const int pl = e >> 5, c = e & 31;
            const long long p = pb + pl;
            uint16_t gv = 0;
            if (p < pend && co0 + c < a.cout) {
                const size_t o = (size_t)p * a.cout + co0 + c;
                if (a.dyb) {
                    gv = a.dyb[o];
                } else {
                    float g = a.dy[o];
                    if (!(from_bf16(a.y[o]) > 0.f)) g = 0.f;
                    gv = to_bf16(g);
                }
            }
            sdy[c][pl] = gv;
            uint16_t xv = 0;
            if (p < pend && ci0 + c < cin) {
                const int nimg = (int)(p / hw), rr = (int)(p - (long long)nimg * hw);
                const int y = rr / a.w + dy, x = rr % a.w + dx;
                if (y >= 0 && y < a.h && x >= 0 && x < a.w) {
                    const long long q = (long long)nimg * hw + (long long)y * a.w + x;
                    const int ci = ci0 + c;
                    xv = ci < a.c1 ? a.x1[q * a.c1 + ci] : a.x2[q * a.c2 + (ci - a.c1)];
                }
            }
            sx[c][pl] = xv;
        }
        __syncthreads();
        for (int kk = 0; kk < 64; kk += 32) {
            Pack8 A, B;
            for (int q = 0; q < 8; ++q) {
                A.h[q] = sdy[ti * 16 + row][kk + kq * 8 + q];
                B.h[q] = sx[tj * 16 + row][kk + kq * 8 + q];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.v, B.v, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // D[4 kq + r][row]: co = co0 + 16 ti + 4 kq + r, ci = ci0 + 16 tj + row
    for (int r = 0; r < 4; ++r) {
        const int co = co0 + ti * 16 + kq * 4 + r, ci = ci0 + tj * 16 + row;
        if (co < a.cout && ci < cin && acc[r] != 0.f) atomicAdd(&a.dw[((size_t)co * 9 + tap) * cin + ci], acc[r]);
    }
}

// Tiled weight gradient for widths that are a multiple of 8.  A workgroup
// owns 32 output channels x 32 input channels and all 9 taps; it walks
// stages of R = 256 / w image rows (whole rows of one image).  Per stage it
// stages in LDS, channel-major with pixels contiguous:
//   sdy[co][r][x]      the ReLU-masked output gradient (BF16), R rows;
//   sx[ci][rr][8 + x]  the input rows y0-1 .. y0+R with an 8-element zero
//                      pad each side (rows outside the image are zero).
// Each wave takes one 16 x 16 (co, ci) tile and keeps 9 tap accumulators.
// For a chunk of 32 pixels a lane's MFMA k-group is 8 consecutive pixels of
// one row: A = 8 pixels of dY (one aligned 16-byte read); for each row
// offset dy, three aligned 16-byte reads of the X row give the windows for
// dx = -1, 0, +1 by byte alignment.  9 MFMAs per chunk.  The workgroups of
// the first input-channel block also sum the staged dY per output channel
// (the bias gradient).  Partial sums go to part[z] = {dW [co][tap][ci], db
// [co]} (no atomics); the reduction adds them into dW and db, which follow
// each other in the parameters.
constexpr int kWgPix = 256;
// workgroups of one layer's launch: fewer slices of partials against more
// parallelism; 256 measured best in the training step (1.18 ms, against
// 1.21 ms at 512 and 1.31 ms at 128)
constexpr int kWgTargetBlocks = 256;

__device__ __forceinline__ bf16x8 window(const uint4 &lo, const uint4 &mid, const uint4 &hi, int dx) {
    Pack8 r;
    if (dx == 0) {
        r.u = mid;
    } else if (dx < 0) {      // elements -1 .. 6 of mid
        r.u.x = __builtin_amdgcn_alignbyte(mid.x, lo.w, 2);
        r.u.y = __builtin_amdgcn_alignbyte(mid.y, mid.x, 2);
        r.u.z = __builtin_amdgcn_alignbyte(mid.z, mid.y, 2);
        r.u.w = __builtin_amdgcn_alignbyte(mid.w, mid.z, 2);
    } else {                  // elements 1 .. 8 of mid
        r.u.x = __builtin_amdgcn_alignbyte(mid.y, mid.x, 2);
        r.u.y = __builtin_amdgcn_alignbyte(mid.z, mid.y, 2);
        r.u.z = __builtin_amdgcn_alignbyte(mid.w, mid.z, 2);
        r.u.w = __builtin_amdgcn_alignbyte(hi.x, mid.w, 2);
    }
    return r.v;
}

// Transpose of an 8 x 8 block of 16-bit values: in[j] holds 8 channels of
// pixel j; dst + q plane receives channel q's 8 pixels as one 16-byte write.
__device__ __forceinline__ void transpose8_store(const uint4 (&in)[8], uint16_t *dst, int plane) {
    uint32_t w[8][4];
    for (int j = 0; j < 8; ++j) {
        w[j][0] = in[j].x;
        w[j][1] = in[j].y;
        w[j][2] = in[j].z;
        w[j][3] = in[j].w;
    }
    for (int q = 0; q < 8; ++q) {
        uint32_t o[4];
        for (int k = 0; k < 4; ++k) {          // pixels 2k, 2k + 1 of channel q
            const uint32_t lo = w[2 * k][q >> 1], hi = w[2 * k + 1][q >> 1];
            o[k] = (q & 1) ? ((lo >> 16) | (hi & 0xffff0000u)) : ((lo & 0xffffu) | (hi << 16));
        }
        *(uint4 *)(dst + (size_t)q * plane) = make_uint4(o[0], o[1], o[2], o[3]);
    }
}

// Workgroup (bx, by, bz) of zn slices; smem: the dynamic LDS of the launch.
__device__ __forceinline__ void wgrad_tile_body(const WgradArgs &a, int R, int rowsPerImg, int stages, int bx, int by,
                                                int bz, int zn, uint16_t *smem) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int co0 = bx * 32, ci0 = by * 32;
    const int W = a.w, H = a.h, hw = H * W;
    const int cin = a.c1 + a.c2;
    const bool first = ci0 < a.c1;
    const uint16_t *xs = first ? a.x1 : a.x2;
    const int cc = first ? a.c1 : a.c2, cl0 = first ? ci0 : ci0 - a.c1;
    const int RW = R * W, rowX = W + 16;
    const int planeD = RW + 8, planeX = (R + 2) * rowX + 8;      // padded channel planes
    uint16_t *sdy = smem;
    uint16_t *sx = smem + 32 * planeD;
    const int ti = wave >> 1, tj = wave & 1, row = lane & 15, kq = lane >> 4;
    f32x4 acc[9];
    for (int t = 0; t < 9; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // bias gradient (workgroups of the first input-channel block): channel
    // tid / 8 of the block, 16-byte groups tid % 8, tid % 8 + 8, ... of its row
    const bool doBias = a.db && by == 0;
    float bsum = 0.f;
    for (int st = bz; st < stages; st += zn) {
        const int img = st / rowsPerImg, y0 = (st - img * rowsPerImg) * R;
        const int rv = min(R, H - y0);                 // rows of this stage inside the image
        __syncthreads();                               // the previous stage's reads are done
        // 8 pixels x 8 channels per work-item: eight 16-byte loads (one per
        // pixel), a register transpose, eight 16-byte LDS writes (one per
        // channel).  W and the padded row rowX are multiples of 8, so an
        // octet never crosses a row and lies wholly inside or outside the
        // image.  dY: RW / 8 octets x 4 channel groups.
        for (int e = tid; e < RW / 2; e += 256) {
            const int c8 = (e & 3) * 8, p0 = (e >> 2) * 8;
            const int r = p0 / W, x = p0 - r * W;
            uint4 in[8];
            if (r < rv) {
                const uint16_t *src = a.dyb + ((size_t)img * hw + (size_t)(y0 + r) * W + x) * a.cout + co0 + c8;
                for (int j = 0; j < 8; ++j) in[j] = *(const uint4 *)(src + (size_t)j * a.cout);
            } else {
                for (int j = 0; j < 8; ++j) in[j] = make_uint4(0, 0, 0, 0);   // rows beyond the image
            }
            transpose8_store(in, sdy + c8 * planeD + p0, planeD);
        }
        // X rows y0-1 .. y0+R with 8 zero columns either side: (R + 2) rowX / 8 octets
        const int octX = (R + 2) * rowX / 8;
        for (int e = tid; e < octX * 4; e += 256) {
            const int c8 = (e & 3) * 8, q0 = (e >> 2) * 8;
            const int rr = q0 / rowX, xx = q0 - rr * rowX - 8, y = y0 - 1 + rr;
            uint4 in[8];
            if (y >= 0 && y < H && xx >= 0 && xx < W) {
                const uint16_t *src = xs + ((size_t)img * hw + (size_t)y * W + xx) * cc + cl0 + c8;
                for (int j = 0; j < 8; ++j) in[j] = *(const uint4 *)(src + (size_t)j * cc);
            } else {
                for (int j = 0; j < 8; ++j) in[j] = make_uint4(0, 0, 0, 0);
            }
            transpose8_store(in, sx + c8 * planeX + q0, planeX);
        }
        __syncthreads();
        if (doBias) {
            const uint16_t *bRow = sdy + (tid >> 3) * planeD;
            for (int g8 = 8 * (tid & 7); g8 < RW; g8 += 64) {
                Pack8 v;
                v.u = *(const uint4 *)(bRow + g8);
                for (int q = 0; q < 8; ++q) bsum += from_bf16(v.h[q]);
            }
        }
        const uint16_t *dRow = sdy + (ti * 16 + row) * planeD;
        const uint16_t *xPlane = sx + (tj * 16 + row) * planeX;
        int gr = (8 * kq) / W, gx = 8 * kq - gr * W;       // row and column of pixel pb + 8 kq
        for (int pb = 0; pb < RW; pb += 32, gx += 32) {
            for (; gx >= W; gx -= W) ++gr;
            // k-groups past the stage's pixels (R * W not a multiple of 32)
            // contribute A = 0 and read row 0
            const int g = pb + 8 * kq;
            const bool gv = g < RW;
            const int r = gv ? gr : 0, x = gv ? gx : 0;                // x is a multiple of 8
            Pack8 A;
            A.u = gv ? *(const uint4 *)(dRow + g) : make_uint4(0, 0, 0, 0);   // zero rows beyond the image
            for (int dy = 0; dy < 3; ++dy) {
                const uint16_t *xr = xPlane + (r + dy) * rowX + 8 + x;
                const uint4 lo = *(const uint4 *)(xr - 8), mid = *(const uint4 *)xr, hi = *(const uint4 *)(xr + 8);
                for (int dx = 0; dx < 3; ++dx)
                    acc[dy * 3 + dx] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.v, window(lo, mid, hi, dx - 1),
                                                                                acc[dy * 3 + dx], 0, 0, 0);
            }
        }
    }
    // D[4 kq + r][row]: co = co0 + 16 ti + 4 kq + r, ci = ci0 + 16 tj + row.
    // Slice z holds dW [cout][9][cin] then the bias gradient [cout].
    const size_t E = (size_t)a.cout * 9 * cin;
    float *dst = a.part + (size_t)bz * (E + (a.db ? a.cout : 0));
    for (int t = 0; t < 9; ++t)
        for (int r = 0; r < 4; ++r) {
            const int co = co0 + ti * 16 + kq * 4 + r, ci = ci0 + tj * 16 + row;
            dst[((size_t)co * 9 + t) * cin + ci] = acc[t][r];
        }
    if (doBias) {                    // the 8 work-items of a channel are lanes 8c .. 8c + 7 of a wave
        for (int off = 1; off < 8; off <<= 1) bsum += __shfl_xor(bsum, off, 64);
        if ((tid & 7) == 0) dst[E + co0 + (tid >> 3)] = bsum;
    }
}

__global__ __launch_bounds__(256) void wgrad_tile_kernel(WgradArgs a, int R, int rowsPerImg, int stages) {
    extern __shared__ __align__(16) uint16_t smem[];
    wgrad_tile_body(a, R, rowsPerImg, stages, blockIdx.x, blockIdx.y, blockIdx.z, gridDim.z, smem);
}

// Several layers' weight gradients in one launch: job j takes workgroups
// [job[j].first, job[j + 1].first), laid out as its own (gx, gy, zn) grid.
__global__ __launch_bounds__(256) void wgrad_batch_kernel(WgradJobs b) {
    extern __shared__ __align__(16) uint16_t smem[];
    int j = 0;
    while (j + 1 < b.nj && (int)blockIdx.x >= b.job[j + 1].first) ++j;
    const WgradJob &J = b.job[j];
    const int w = (int)blockIdx.x - J.first, bx = w % J.gx, byz = w / J.gx;
    wgrad_tile_body(J.a, J.R, J.rowsPerImg, J.stages, bx, byz % J.gy, byz / J.gy, J.zn, smem);
}

// Deterministic reduction of part[zn][E] into dW: dW[e] = sum_z part[z][e]
// (stored: dW needs no zeroing).  A workgroup takes `cols` float4 columns (4
// cols consecutive elements) and zg = 256 / cols groups of slices; work-item
// (col, g) sums slices g, g + zg, g + 2 zg, ... in that order (eight loads
// in flight), then the zg group sums of a column are added in group order.
// The order depends only on zn and zg, so the result is reproducible.
__device__ __forceinline__ void slice_sum_body(const float4 *__restrict__ part, int zn, int zg, size_t E4,
                                               float4 *__restrict__ dw, unsigned blk) {
    __shared__ float4 sums[256];
    const int tid = threadIdx.x, cols = 256 / zg, col = tid % cols, g = tid / cols;
    const size_t e4 = (size_t)blk * cols + col;
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    if (e4 < E4) {
        const float4 *p = part + e4;
        int z = g;
        for (; z + 7 * zg < zn; z += 8 * zg) {
            float4 v[8];
            for (int u = 0; u < 8; ++u) v[u] = p[(size_t)(z + u * zg) * E4];
            for (int u = 0; u < 8; ++u) {
                s.x += v[u].x;
                s.y += v[u].y;
                s.z += v[u].z;
                s.w += v[u].w;
            }
        }
        for (; z < zn; z += zg) {
            const float4 v = p[(size_t)z * E4];
            s.x += v.x;
            s.y += v.y;
            s.z += v.z;
            s.w += v.w;
        }
    }
    sums[tid] = s;
    __syncthreads();
    if (g == 0 && e4 < E4) {
        float4 t = sums[col];
        for (int k = 1; k < zg; ++k) {
            const float4 v = sums[k * cols + col];
            t.x += v.x;
            t.y += v.y;
            t.z += v.z;
            t.w += v.w;
        }
        dw[e4] = t;
    }
}

__global__ __launch_bounds__(256) void wgrad_sum_slices_kernel(const float4 *__restrict__ part, int zn, int zg,
                                                               size_t E4, float4 *__restrict__ dw) {
    slice_sum_body(part, zn, zg, E4, dw, blockIdx.x);
}

// the slice sums of a batch: entry j takes workgroups [s[j].first, s[j + 1].first)
__global__ __launch_bounds__(256) void wgrad_sum_batch_kernel(SliceSumJobs b) {
    int j = 0;
    while (j + 1 < b.n && (int)blockIdx.x >= b.s[j + 1].first) ++j;
    const SliceSumJob &S = b.s[j];
    slice_sum_body(S.part, S.zn, S.zg, S.E4, S.dw, blockIdx.x - S.first);
}

// db[c] = sum_b part[b][c]: one workgroup per channel, a fixed-order tree
// (the bias gradient beside the split-K weight gradient)
__global__ __launch_bounds__(256) void bias_reduce_kernel(const float *__restrict__ part, int nb, int cout,
                                                          float *__restrict__ db) {
    const int c = blockIdx.x, tid = threadIdx.x;
    float s = 0.f;
    for (int b = tid; b < nb; b += 256) s += part[(size_t)b * cout + c];
    __shared__ float red[256];
    red[tid] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) red[tid] += red[tid + h];
        __syncthreads();
    }
    if (tid == 0) db[c] = red[0];
}

// The ReLU-masked output gradient of a layer in BF16 (dy null: read from
// out, already masked) and, when part is set, each workgroup's sums of the
// rounded masked dY per channel in part[block][cout] (the bias gradient's
// partials for the split-K weight gradient; the tiled one sums the bias
// itself).  A work-item takes 4 channels; cout divides 1024, so its channels
// stay the same over the grid-stride loop.  The last workgroups also finish
// up to two pending sums.
__global__ __launch_bounds__(256) void mask_bf16_kernel(const float *__restrict__ dy, const uint16_t *__restrict__ y,
                                                        long long E4, int cout, uint16_t *__restrict__ out,
                                                        float *__restrict__ part, SideSums side) {
    const int tid = threadIdx.x;
    float sb[4] = {0.f, 0.f, 0.f, 0.f};
    for (long long q = blockIdx.x * 256LL + tid; q < E4; q += (long long)gridDim.x * 256) {
        uint2 r;
        if (dy) {
            r = mask4(((const float4 *)dy)[q], ((const uint2 *)y)[q]);
            ((uint2 *)out)[q] = r;
        } else {
            r = ((const uint2 *)out)[q];
        }
        if (part) {
            sb[0] += from_bf16((uint16_t)r.x);
            sb[1] += from_bf16((uint16_t)(r.x >> 16));
            sb[2] += from_bf16((uint16_t)r.y);
            sb[3] += from_bf16((uint16_t)(r.y >> 16));
        }
    }
    __shared__ float red[256][4];
    if (part) {
        for (int k = 0; k < 4; ++k) red[tid][k] = sb[k];
        __syncthreads();
    }
    const int groups = cout / 4;               // work-items of one channel set: tid, tid + groups, ...
    if (part && tid < groups)
        for (int k = 0; k < 4; ++k) {
            float s = 0.f;
            for (int j = tid; j < 256; j += groups) s += red[j][k];
            part[(size_t)blockIdx.x * cout + 4 * tid + k] = s;
        }
    // the last workgroups also finish up to two jobs of earlier launches
    for (int q = 0; q < 2; ++q)
        if (side.job[q].part && (int)blockIdx.x == (int)gridDim.x - 1 - min(q, (int)gridDim.x - 1)) {
            __syncthreads();
            fixed_order_sum(side.job[q]);
        }
}

// up to two sums left as partials by earlier launches, one workgroup each
__global__ __launch_bounds__(256) void finish_sums_kernel(SideSums side) {
    if (side.job[blockIdx.x].part) fixed_order_sum(side.job[blockIdx.x]);
}

// 2x2 average pooling, BF16 NHWC (h, w even)
__global__ void pool_fwd_kernel(const uint16_t *__restrict__ x, int n, int h, int w, int c,
                                uint16_t *__restrict__ y) {
    const int ho = h / 2, wo = w / 2;
    const long long total = (long long)n * ho * wo * c;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(e % c);
        const long long pix = e / c;
        const int xo = (int)(pix % wo), yo = (int)((pix / wo) % ho), ni = (int)(pix / ((long long)wo * ho));
        const long long b = (((long long)ni * h + 2 * yo) * w + 2 * xo) * c + ch;
        const float s = from_bf16(x[b]) + from_bf16(x[b + c]) + from_bf16(x[b + (long long)w * c]) +
                        from_bf16(x[b + (long long)w * c + c]);
        y[e] = to_bf16(0.25f * s);
    }
}

// gradient of the pooling: dx[2y+a][2x+b] (+)= dy[y][x] / 4 (FP32)
__global__ void pool_bwd_kernel(const float *__restrict__ dy, int n, int h, int w, int c, float *__restrict__ dx,
                                int accumulate) {
    const long long total = (long long)n * h * w * c;
    const int ho = h / 2, wo = w / 2;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(e % c);
        const long long pix = e / c;
        const int x = (int)(pix % w), y = (int)((pix / w) % h), ni = (int)(pix / ((long long)w * h));
        const float g = 0.25f * dy[(((long long)ni * ho + y / 2) * wo + x / 2) * c + ch];
        dx[e] = accumulate ? dx[e] + g : g;
    }
}

// Vector forms of the pooling and upsampling kernels: one work-item per 8
// BF16 or 4 FP32 channels of a pixel (16-byte accesses) and 32-bit index
// arithmetic, for channel counts that divide and groups that fit in 31 bits.
// Every element is computed by the same expression as in the scalar forms.
struct PixIdx {
    unsigned ni, y, x, g;
};
__device__ __forceinline__ PixIdx pix_idx(unsigned e, unsigned groups, unsigned w, unsigned h) {
    const unsigned pix = e / groups, t = pix / w;
    return PixIdx{t / h, t - (t / h) * h, pix - t * w, e - pix * groups};
}

__global__ void pool_fwd_vec_kernel(const uint4 *__restrict__ x, unsigned n, unsigned h, unsigned w, unsigned groups,
                                    uint4 *__restrict__ y) {
    const unsigned ho = h / 2, wo = w / 2, total = n * ho * wo * groups;
    for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const PixIdx q = pix_idx(e, groups, wo, ho);
        const size_t b = (((size_t)q.ni * h + 2 * q.y) * w + 2 * q.x) * groups + q.g;
        Pack8 A, B, C, D, o;
        A.u = x[b]; B.u = x[b + groups]; C.u = x[b + (size_t)w * groups]; D.u = x[b + (size_t)w * groups + groups];
        for (int k = 0; k < 8; ++k) {
            const float s = from_bf16(A.h[k]) + from_bf16(B.h[k]) + from_bf16(C.h[k]) + from_bf16(D.h[k]);
            o.h[k] = to_bf16(0.25f * s);
        }
        y[e] = o.u;
    }
}

// m.gm set: the sum goes out ReLU-masked in BF16 to m.gm instead of dx
__global__ void pool_bwd_vec_kernel(const float4 *__restrict__ dy, unsigned n, unsigned h, unsigned w,
                                    unsigned groups, float4 *__restrict__ dx, int accumulate, MaskOut m) {
    const unsigned ho = h / 2, wo = w / 2, total = n * h * w * groups;
    for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const PixIdx q = pix_idx(e, groups, w, h);
        const float4 d = dy[(((size_t)q.ni * ho + q.y / 2) * wo + q.x / 2) * groups + q.g];
        float4 g = make_float4(0.25f * d.x, 0.25f * d.y, 0.25f * d.z, 0.25f * d.w);
        if (accumulate) {
            const float4 a = dx[e];
            g = make_float4(a.x + g.x, a.y + g.y, a.z + g.z, a.w + g.w);
        }
        if (m.gm)
            ((uint2 *)m.gm)[e] = mask4(g, ((const uint2 *)m.y)[e]);
        else
            dx[e] = g;
    }
}

__global__ void up_fwd_vec_kernel(const uint4 *__restrict__ x, unsigned n, unsigned h, unsigned w, unsigned groups,
                                  uint4 *__restrict__ y) {
    const unsigned ho = 2 * h, wo = 2 * w, total = n * ho * wo * groups;
    for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const PixIdx q = pix_idx(e, groups, wo, ho);
        y[e] = x[(((size_t)q.ni * h + q.y / 2) * w + q.x / 2) * groups + q.g];
    }
}

// m.gm set: the sum goes out ReLU-masked in BF16 to m.gm instead of dx
__global__ void up_bwd_vec_kernel(const float4 *__restrict__ dy, unsigned n, unsigned h, unsigned w, unsigned groups,
                                  float4 *__restrict__ dx, MaskOut m) {
    const unsigned wo = 2 * w, total = n * h * w * groups;
    for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const PixIdx q = pix_idx(e, groups, w, h);
        const size_t b = (((size_t)q.ni * 2 * h + 2 * q.y) * wo + 2 * q.x) * groups + q.g;
        const float4 a = dy[b], bb = dy[b + groups], c = dy[b + (size_t)wo * groups],
                     d = dy[b + (size_t)wo * groups + groups];
        const float4 g = make_float4(a.x + bb.x + c.x + d.x, a.y + bb.y + c.y + d.y, a.z + bb.z + c.z + d.z,
                                     a.w + bb.w + c.w + d.w);
        if (m.gm)
            ((uint2 *)m.gm)[e] = mask4(g, ((const uint2 *)m.y)[e]);
        else
            dx[e] = g;
    }
}

// nearest-neighbour 2x upsampling, BF16 NHWC
__global__ void up_fwd_kernel(const uint16_t *__restrict__ x, int n, int h, int w, int c, uint16_t *__restrict__ y) {
    const int ho = 2 * h, wo = 2 * w;
    const long long total = (long long)n * ho * wo * c;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(e % c);
        const long long pix = e / c;
        const int xo = (int)(pix % wo), yo = (int)((pix / wo) % ho), ni = (int)(pix / ((long long)wo * ho));
        y[e] = x[(((long long)ni * h + yo / 2) * w + xo / 2) * c + ch];
    }
}

// its gradient: dx[y][x] = sum of the 4 dy it fed (FP32)
__global__ void up_bwd_kernel(const float *__restrict__ dy, int n, int h, int w, int c, float *__restrict__ dx) {
    const long long total = (long long)n * h * w * c;
    const int wo = 2 * w;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(e % c);
        const long long pix = e / c;
        const int x = (int)(pix % w), y = (int)((pix / w) % h), ni = (int)(pix / ((long long)w * h));
        const long long b = (((long long)ni * 2 * h + 2 * y) * wo + 2 * x) * c + ch;
        dx[e] = dy[b] + dy[b + c] + dy[b + (long long)wo * c] + dy[b + (long long)wo * c + c];
    }
}

// network input: x [n][cin][h][w] FP32 -> NHWC BF16 with `c` channels
// (input channel q in channel q, zeros above cin); with x0 != null also
// channel 0 as [n][h][w] FP32 (the residual's input)
__global__ void input_kernel(const float *__restrict__ x, long long P, int hw, int cin, int c,
                             uint16_t *__restrict__ y, float *__restrict__ x0) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < P * c;
         e += (long long)gridDim.x * blockDim.x) {
        const long long p = e / c;
        const int ch = (int)(e - p * c);
        if (ch >= cin) {
            y[e] = 0;
            continue;
        }
        const long long ni = p / hw, pix = p - ni * hw;
        const float v = x[(ni * cin + ch) * hw + pix];
        y[e] = to_bf16(v);
        if (x0 && ch == 0) x0[p] = v;
    }
}

// The same conversion, one work-item per pixel (c a multiple of 8): its cin
// values are read (coalesced across the wave), and its c channels written
// with 16-byte stores.  Bit-identical to input_kernel.
__global__ void input_pix_kernel(const float *__restrict__ x, long long P, int hw, int cin, int c,
                                 uint16_t *__restrict__ y, float *__restrict__ x0) {
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < P; p += (long long)gridDim.x * blockDim.x) {
        const long long ni = p / hw, pix = p - ni * hw;
        const float *xp = x + ni * cin * hw + pix;
        uint4 *yp = (uint4 *)(y + p * c);
        for (int g = 0; g < c; g += 8) {
            Pack8 o;
            for (int q = 0; q < 8; ++q) o.h[q] = g + q < cin ? to_bf16(xp[(long long)(g + q) * hw]) : (uint16_t)0;
            yp[g / 8] = o.u;
        }
        if (x0) x0[p] = xp[0];
    }
}

// head: out[p] = x[p] + b + sum_c wh[c] * d[p][c]   (FP32; d BF16 with kHead
// channels, wh = kHead weights then the bias)
constexpr int kHead = 32;

__global__ void head_fwd_kernel(const uint16_t *__restrict__ d, const float *__restrict__ x,
                                const float *__restrict__ wh, long long P, float *__restrict__ out) {
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < P; p += (long long)gridDim.x * blockDim.x) {
        float s = wh[kHead];
        for (int q = 0; q < kHead; ++q) s += wh[q] * from_bf16(d[p * kHead + q]);
        out[p] = x[p] + s;
    }
}

// loss += mean wgt (out - t)^2 (wgt = 1 when null); g = d loss / d out;
// head gradients dwh; dd[p][c] = g[p] wh[c], or with gm set the last
// layer's masked gradient gm[p][c] = bf16(d[p][c] > 0 ? g[p] wh[c] : 0)
// (d is that layer's ReLU output) in place of dd
__global__ __launch_bounds__(256) void head_bwd_kernel(const uint16_t *__restrict__ d, const float *__restrict__ out,
                                                       const float *__restrict__ t, const float *__restrict__ wgt,
                                                       const float *__restrict__ wh, long long P,
                                                       float *__restrict__ dd, uint16_t *__restrict__ gm,
                                                       float *__restrict__ part) {
    // four work-items per pixel, 8 of its kHead channels each (16-byte loads)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, cg = tid & 3;
    const float inv = 1.0f / (float)P;
    float v[10];                          // dwh of the 8 channels, the bias, the squared error
    for (int q = 0; q < 10; ++q) v[q] = 0.f;
    float w8[8];
    for (int q = 0; q < 8; ++q) w8[q] = wh[8 * cg + q];
    for (long long p = (blockIdx.x * 256LL + tid) >> 2; p < P; p += (long long)gridDim.x * 64) {
        const float e = out[p] - t[p];
        const float wv = wgt ? wgt[p] : 1.0f;
        const float g = 2.0f * e * wv * inv;
        if (cg == 0) {
            v[8] += g;
            v[9] += wv * e * e;
        }
        Pack8 dv;
        dv.u = *(const uint4 *)(d + p * kHead + 8 * cg);
        for (int q = 0; q < 8; ++q) v[q] += g * from_bf16(dv.h[q]);
        const float4 g0 = make_float4(g * w8[0], g * w8[1], g * w8[2], g * w8[3]);
        const float4 g1 = make_float4(g * w8[4], g * w8[5], g * w8[6], g * w8[7]);
        if (gm) {
            uint4 r;
            const uint2 lo = mask4(g0, make_uint2(dv.u.x, dv.u.y)), hi = mask4(g1, make_uint2(dv.u.z, dv.u.w));
            r.x = lo.x; r.y = lo.y; r.z = hi.x; r.w = hi.y;
            *(uint4 *)(gm + p * kHead + 8 * cg) = r;
        } else {
            float *o = dd + p * kHead + 8 * cg;
            *(float4 *)o = g0;
            *(float4 *)(o + 4) = g1;
        }
    }
    // sum over the lanes of a wave with the same cg, then over the waves
    for (int off = 4; off < 64; off <<= 1)
        for (int q = 0; q < 10; ++q) v[q] += __shfl_xor(v[q], off, 64);
    __shared__ float red[4][4][10];
    if (lane < 4)
        for (int q = 0; q < 10; ++q) red[wave][lane][q] = v[q];
    __syncthreads();
    // this workgroup's partials: the kHead weight gradients, the bias
    // gradient, the weighted squared error over P (summed by a later launch)
    constexpr int m = kHead + 2;
    static_assert(m == kHeadPartials, "head partials layout");
    if (tid < 40) {
        const int c = tid / 10, q = tid - c * 10;
        const float s = red[0][c][q] + red[1][c][q] + red[2][c][q] + red[3][c][q];
        float *pb = part + (size_t)blockIdx.x * m;
        if (q < 8)
            pb[8 * c + q] = s;
        else if (c == 0)
            pb[kHead + q - 8] = q == 9 ? s * inv : s;
    }
}

// w'[ci][tap'][co] = w[co][8 - tap][ci]: the weights of the backward-data conv
__global__ void flip_kernel(const uint16_t *__restrict__ w, int cout, int cin, uint16_t *__restrict__ wf) {
    const long long total = (long long)cout * 9 * cin;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int ci = (int)(e % cin);
        const int tap = (int)((e / cin) % 9);
        const int co = (int)(e / ((long long)cin * 9));
        wf[((size_t)ci * 9 + (8 - tap)) * cout + co] = w[e];
    }
}

// Adam (Kingma & Ba, ICLR 2015) on the FP32 master parameters; c1, c2 are the
// bias corrections 1 - b1^t, 1 - b2^t
__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, long long n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const float gi = g[e];
        const float mi = b1 * m[e] + (1.f - b1) * gi;
        const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// The BF16 copy of a conv weight and its flipped, transposed copy
// (w'[ci][8 - tap][co] = w[co][tap][ci], as flip_kernel), if parameter e is a
// conv weight.
__device__ __forceinline__ void weight_copies(long long e, float x, const WeightMap &map, uint16_t *__restrict__ wb,
                                              uint16_t *__restrict__ wf) {
    for (int l = 0; l < map.layers; ++l) {
        const long long idx = e - map.w_off[l];
        const int cout = map.cout[l], cin = map.cin[l];
        if (idx < 0 || idx >= (long long)cout * 9 * cin) continue;
        const int ci = (int)(idx % cin), t = (int)(idx / cin);
        const int tap = t % 9, co = t / 9;
        const uint16_t h = to_bf16(x);
        wb[map.wb_off[l] + idx] = h;
        wf[map.wb_off[l] + ((long long)ci * 9 + (8 - tap)) * cout + co] = h;
        return;
    }
}

// Adam as adam_kernel, then the BF16 copies of the updated conv weights.
__global__ void adam_bf16_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                                 float *__restrict__ v, long long n, float lr, float b1, float b2, float eps, float c1,
                                 float c2, WeightMap map, uint16_t *__restrict__ wb, uint16_t *__restrict__ wf) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const float gi = g[e];
        const float mi = b1 * m[e] + (1.f - b1) * gi;
        const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        const float pi = p[e] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        p[e] = pi;
        weight_copies(e, pi, map, wb, wf);
    }
}

__global__ void weights_bf16_kernel(const float *__restrict__ p, long long n, WeightMap map, uint16_t *__restrict__ wb,
                                    uint16_t *__restrict__ wf) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        weight_copies(e, p[e], map, wb, wf);
}

__global__ void to_bf16_kernel(const float *__restrict__ x, long long n, uint16_t *__restrict__ y) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        y[e] = to_bf16(x[e]);
}

__global__ void fill_f32_kernel(float *__restrict__ a, long long n, float v) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        a[e] = v;
}

bool job_ok(const PartialSums &j) { return !j.part || (j.m >= 1 && j.m <= 256 && j.nb >= 1 && j.stride >= j.m); }

int blocks_for(long long n) {
    long long b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    return (int)(b > 0 ? b : 1);
}

}  // namespace

// where no kernel writes it directly, a masked output is the FP32 result masked by a second launch
static hipError_t mask_after(const float *dx, long long E, int c, const MaskOut &m, hipStream_t s) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !m.gm) return e;
    int nb = 0;
    return launch_mask_bf16(dx, m.y, E, c, m.gm, nullptr, 0, &nb, SideSums{}, s);
}

hipError_t launch_conv3x3(const ConvArgs &a, bool fp32_input, hipStream_t s) {
    const long long P = (long long)a.n * a.h * a.w;
    const int nout = a.cy1 + a.cy2;
    if (P <= 0 || nout <= 0) return hipSuccess;
    if ((a.c1 % 32) || (a.c2 % 32) || (nout % 32) || ((a.cy1 % 32) && a.cy2)) return hipErrorInvalidValue;
    // a masked BF16 output: the input gradient of a single-input layer, stored
    if (a.mo && (a.mode != 1 || a.cy2 || a.acc1 || !a.my || !a.y1)) return hipErrorInvalidValue;
    if (a.pg && !a.mo) return hipErrorInvalidValue;
    if (a.uo && (a.mode != 1 || !a.cy2 || a.acc2 || !a.uy || !a.y2 || !a.ud || (a.h | a.w) & 1))
        return hipErrorInvalidValue;
    if (!fp32_input && !a.m1) {
        // tiled path: R rows per workgroup (R * w <= 256), halved while the
        // grid has fewer than 512 workgroups, as long as the stage fits in LDS
        // (in the training step 512 measured best: 1.25 ms at 256, 1.28 ms at 1024)
        int R = 256 / a.w < a.h ? 256 / a.w : a.h;
        if (R < 1) R = 1;
        // stage bytes with the weights in LDS (wl) or in registers, at least
        // the epilogue's tile of r w pixels x 32 channels
        auto lds = [&](int r, bool wl) {
            const size_t stage = ((size_t)(r + 2) * (a.w + 2) + (wl ? kCvWRows : 0)) * kCvPS * sizeof(uint16_t);
            const size_t out = (size_t)r * a.w * (a.mode == 0 ? kCvPS * sizeof(uint16_t) : kCvOutF * sizeof(float));
            return stage > out ? stage : out;
        };
        while (R > 1 && lds(R, false) > 64 * 1024) --R;
        while (R > 1 && (long long)a.n * ((a.h + R - 1) / R) * (nout / 32) < 512) R = (R + 1) / 2;
        if (lds(R, false) <= 64 * 1024 && R * a.w <= 256) {
            const int rowsPerImg = (a.h + R - 1) / R;
            // the upsampling's backward needs row pairs in a stage; with odd
            // R it runs as a second launch on the FP32 gradient
            ConvArgs b = a;
            const bool upLater = a.uo && (R & 1);
            if (upLater) b.uo = nullptr;
            const dim3 grid((unsigned)(a.n * rowsPerImg), (unsigned)(nout / 32));
            const int RW = R * a.w;
            if (lds(R, true) > 64 * 1024)
                conv3x3_tile_kernel<4, false><<<grid, 256, lds(R, false), s>>>(b, R, rowsPerImg);
            else if (RW <= 64)
                conv3x3_tile_kernel<1, true><<<grid, 256, lds(R, true), s>>>(b, R, rowsPerImg);
            else if (RW <= 128)
                conv3x3_tile_kernel<2, true><<<grid, 256, lds(R, true), s>>>(b, R, rowsPerImg);
            else
                conv3x3_tile_kernel<4, true><<<grid, 256, lds(R, true), s>>>(b, R, rowsPerImg);
            if (!upLater) return hipGetLastError();
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            return launch_up_bwd((const float *)a.y2, a.n, a.h / 2, a.w / 2, a.cy2, a.ud, MaskOut{a.uo, a.uy}, s);
        }
    }
    // the general kernel writes the FP32 gradients to y1 and y2; second
    // launches mask them (through the pooling's or the upsampling's backward)
    const dim3 grid((unsigned)((P + 127) / 128), (unsigned)(nout / 32));
    if (fp32_input)
        conv3x3_kernel<float><<<grid, 256, 0, s>>>(a);
    else
        conv3x3_kernel<uint16_t><<<grid, 256, 0, s>>>(a);
    if (a.uo) {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        const hipError_t e2 = launch_up_bwd((const float *)a.y2, a.n, a.h / 2, a.w / 2, a.cy2, a.ud,
                                            MaskOut{a.uo, a.uy}, s);
        if (e2 != hipSuccess || !a.mo) return e2;
    }
    if (a.pg) {                         // the pooling's backward, masked, by a second launch
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        return launch_pool_bwd((const float *)a.y1, a.n, 2 * a.h, 2 * a.w, a.cy1, const_cast<float *>(a.pg), true,
                               MaskOut{a.mo, a.my}, s);
    }
    return mask_after((const float *)a.y1, P * a.cy1, a.cy1, MaskOut{a.mo, a.my}, s);
}

// The tiled weight gradient's shape for these arguments: stage rows R,
// stages, partial slices zn and slice length E (dW, then the bias gradient);
// false when the tiled path does not apply (the split-K path runs).
struct WgradTiling {
    int R, rowsPerImg, stages, zn;
    size_t lds, E;
};
static bool wgrad_tiling(const WgradArgs &a, WgradTiling &t) {
    const int cin = a.c1 + a.c2;
    if (!a.part || !a.dyb || a.w % 8 || a.w > kWgPix) return false;  // BF16 dY, widths a multiple of 8
    t.R = kWgPix / a.w < a.h ? kWgPix / a.w : a.h;
    t.lds = (size_t)(32 * (t.R * a.w + 8) + 32 * ((t.R + 2) * (a.w + 16) + 8)) * sizeof(uint16_t);
    t.rowsPerImg = (a.h + t.R - 1) / t.R;
    t.stages = a.n * t.rowsPerImg;
    const int blocks = (a.cout / 32) * (cin / 32);
    t.zn = kWgTargetBlocks / blocks;
    if (t.zn > t.stages) t.zn = t.stages;
    if (t.zn < 1) t.zn = 1;
    t.E = (size_t)a.cout * 9 * cin + (a.db ? a.cout : 0);
    return t.lds <= 64 * 1024 && t.E * t.zn <= a.partCap;             // the stage fits, and the partials
}

bool wgrad_is_tiled(const WgradArgs &a) {
    WgradTiling t;
    return wgrad_tiling(a, t);
}

bool wgrad_plan(const WgradArgs &a, WgradJob &j, size_t *partFloats) {
    WgradArgs probe = a;
    float dummy;
    if (!probe.part) probe.part = &dummy;              // the tiling only asks that a workspace exist
    probe.partCap = (size_t)-1;
    WgradTiling t;
    if (!wgrad_tiling(probe, t)) return false;
    j.a = a;
    j.R = t.R; j.rowsPerImg = t.rowsPerImg; j.stages = t.stages; j.zn = t.zn;
    j.gx = a.cout / 32; j.gy = (a.c1 + a.c2) / 32;
    j.first = 0;
    *partFloats = t.E * t.zn;
    return true;
}

hipError_t launch_wgrad_batch(WgradJobs &b, hipStream_t s) {
    if (b.nj < 1 || b.nj > kWgradBatchMax) return hipErrorInvalidValue;
    SliceSumJobs sums{};
    sums.n = b.nj;
    int wgs = 0, sumBlocks = 0;
    size_t lds = 0;
    for (int q = 0; q < b.nj; ++q) {
        WgradJob &j = b.job[q];
        const WgradArgs &a = j.a;
        const int cin = a.c1 + a.c2;
        WgradTiling t;
        if ((cin % 32) || (a.cout % 32) || !wgrad_tiling(a, t) || t.zn != j.zn) return hipErrorInvalidValue;
        if (a.db && a.db != a.dw + (t.E - a.cout)) return hipErrorInvalidValue;
        if (((uintptr_t)a.dw | (uintptr_t)a.part) & 15) return hipErrorInvalidValue;
        j.first = wgs;
        wgs += j.gx * j.gy * j.zn;
        lds = t.lds > lds ? t.lds : lds;
        int zg = 1;                                    // slice groups: the largest power of two <= min(16, zn)
        while (zg * 2 <= 16 && zg * 2 <= t.zn) zg *= 2;
        const size_t E4 = t.E / 4, cols = 256 / zg;
        sums.s[q] = SliceSumJob{reinterpret_cast<const float4 *>(a.part), reinterpret_cast<float4 *>(a.dw), E4, t.zn, zg,
                                sumBlocks};
        sumBlocks += (int)((E4 + cols - 1) / cols);
    }
    wgrad_batch_kernel<<<wgs, 256, lds, s>>>(b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    wgrad_sum_batch_kernel<<<sumBlocks, 256, 0, s>>>(sums);
    return hipGetLastError();
}

hipError_t launch_wgrad(const WgradArgs &a, hipStream_t s) {
    const int cin = a.c1 + a.c2;
    if ((cin % 32) || (a.cout % 32) || a.splits < 1) return hipErrorInvalidValue;
    WgradTiling t;
    if (wgrad_tiling(a, t)) {
        // a slice of partials: dW, then the bias gradient (a.db must be a.dw + cout 9 cin)
        if (a.db && a.db != a.dw + (t.E - a.cout)) return hipErrorInvalidValue;
        // float4 sums: E is a multiple of 4 (cout is), dW and the partials 16-byte aligned
        if (((uintptr_t)a.dw | (uintptr_t)a.part) & 15) return hipErrorInvalidValue;
        const dim3 grid((unsigned)(a.cout / 32), (unsigned)(cin / 32), (unsigned)t.zn);
        wgrad_tile_kernel<<<grid, 256, t.lds, s>>>(a, t.R, t.rowsPerImg, t.stages);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        // slice groups: the largest power of two <= min(16, zn)
        int zg = 1;
        while (zg * 2 <= 16 && zg * 2 <= t.zn) zg *= 2;
        const size_t E4 = t.E / 4, cols = 256 / zg;
        wgrad_sum_slices_kernel<<<(unsigned)((E4 + cols - 1) / cols), 256, 0, s>>>(
            reinterpret_cast<const float4 *>(a.part), t.zn, zg, E4, reinterpret_cast<float4 *>(a.dw));
        return hipGetLastError();
    }
    // the split-K kernel adds with atomics: dW starts from zero
    hipError_t e = launch_fill_f32(a.dw, (long long)a.cout * 9 * cin, 0.f, s);
    if (e != hipSuccess) return e;
    const dim3 grid((unsigned)(a.cout / 32), (unsigned)(cin / 32), (unsigned)(9 * a.splits));
    wgrad_kernel<<<grid, 256, 0, s>>>(a);
    e = hipGetLastError();
    if (e != hipSuccess || !a.db) return e;
    if (!a.dyb || !a.bpart || a.bnb < 1) return hipErrorInvalidValue;
    bias_reduce_kernel<<<a.cout, 256, 0, s>>>(a.bpart, a.bnb, a.cout, a.db);
    return hipGetLastError();
}

// E = pixels x cout; cout a multiple of 4 that divides 1024; part holds
// blocks x cout floats (at most 16384 x cout)
int mask_bf16_blocks(long long E) {
    // at least 8 groups of 4 elements per work-item, at most 1024 workgroups:
    // the copy runs at full rate and the bias partials stay few
    long long b = (E / 4 + 2047) / 2048;
    return (int)(b > 1024 ? 1024 : (b < 1 ? 1 : b));
}

hipError_t launch_mask_bf16(const float *dy, const uint16_t *y, long long E, int cout, uint16_t *out, float *part,
                            size_t partCap, int *nblocks, const SideSums &side, hipStream_t s) {
    if (cout < 4 || cout > 256 || cout % 4 || 1024 % cout || E % cout) return hipErrorInvalidValue;
    if (!job_ok(side.job[0]) || !job_ok(side.job[1])) return hipErrorInvalidValue;
    const int b = mask_bf16_blocks(E);
    if (part && (size_t)b * cout > partCap) return hipErrorInvalidValue;
    mask_bf16_kernel<<<b, 256, 0, s>>>(dy, y, E / 4, cout, out, part, side);
    *nblocks = b;
    return hipGetLastError();
}

hipError_t launch_finish_sums(const SideSums &side, hipStream_t s) {
    if (!job_ok(side.job[0]) || !job_ok(side.job[1])) return hipErrorInvalidValue;
    if (!side.job[0].part && !side.job[1].part) return hipSuccess;
    finish_sums_kernel<<<2, 256, 0, s>>>(side);
    return hipGetLastError();
}

// the vector forms apply when the channels split into 16-byte groups and the
// group count of the larger tensor stays below 2^31
static bool vec_ok(long long groups_total, int c, int per) { return c % per == 0 && groups_total < (1LL << 31); }

hipError_t launch_pool_fwd(const uint16_t *x, int n, int h, int w, int c, uint16_t *y, hipStream_t s) {
    if (vec_ok((long long)n * h * w * (c / 8), c, 8)) {
        const long long g = (long long)n * (h / 2) * (w / 2) * (c / 8);
        pool_fwd_vec_kernel<<<blocks_for(g), 256, 0, s>>>((const uint4 *)x, n, h, w, c / 8, (uint4 *)y);
    } else {
        pool_fwd_kernel<<<blocks_for((long long)n * (h / 2) * (w / 2) * c), 256, 0, s>>>(x, n, h, w, c, y);
    }
    return hipGetLastError();
}
hipError_t launch_pool_bwd(const float *dy, int n, int h, int w, int c, float *dx, bool accumulate, const MaskOut &m,
                           hipStream_t s) {
    if (vec_ok((long long)n * h * w * (c / 4), c, 4)) {
        pool_bwd_vec_kernel<<<blocks_for((long long)n * h * w * (c / 4)), 256, 0, s>>>(
            (const float4 *)dy, n, h, w, c / 4, (float4 *)dx, accumulate ? 1 : 0, m);
        return hipGetLastError();
    }
    pool_bwd_kernel<<<blocks_for((long long)n * h * w * c), 256, 0, s>>>(dy, n, h, w, c, dx, accumulate ? 1 : 0);
    return mask_after(dx, (long long)n * h * w * c, c, m, s);
}
hipError_t launch_up_fwd(const uint16_t *x, int n, int h, int w, int c, uint16_t *y, hipStream_t s) {
    if (vec_ok((long long)n * 4 * h * w * (c / 8), c, 8))
        up_fwd_vec_kernel<<<blocks_for((long long)n * 4 * h * w * (c / 8)), 256, 0, s>>>((const uint4 *)x, n, h, w,
                                                                                     c / 8, (uint4 *)y);
    else
        up_fwd_kernel<<<blocks_for((long long)n * 4 * h * w * c), 256, 0, s>>>(x, n, h, w, c, y);
    return hipGetLastError();
}
hipError_t launch_up_bwd(const float *dy, int n, int h, int w, int c, float *dx, const MaskOut &m, hipStream_t s) {
    if (vec_ok((long long)n * 4 * h * w * (c / 4), c, 4)) {
        up_bwd_vec_kernel<<<blocks_for((long long)n * h * w * (c / 4)), 256, 0, s>>>((const float4 *)dy, n, h, w,
                                                                                 c / 4, (float4 *)dx, m);
        return hipGetLastError();
    }
    up_bwd_kernel<<<blocks_for((long long)n * h * w * c), 256, 0, s>>>(dy, n, h, w, c, dx);
    return mask_after(dx, (long long)n * h * w * c, c, m, s);
}
hipError_t launch_unet_input(const float *x, long long P, int hw, int cin, int c, uint16_t *y, float *x0,
                             hipStream_t s) {
    if (cin < 1 || cin > c) return hipErrorInvalidValue;
    if (c % 8 == 0)
        input_pix_kernel<<<blocks_for(P), 256, 0, s>>>(x, P, hw, cin, c, y, x0);
    else
        input_kernel<<<blocks_for(P * c), 256, 0, s>>>(x, P, hw, cin, c, y, x0);
    return hipGetLastError();
}
hipError_t launch_head_fwd(const uint16_t *d, const float *x, const float *wh, long long P, float *out,
                           hipStream_t s) {
    head_fwd_kernel<<<blocks_for(P), 256, 0, s>>>(d, x, wh, P, out);
    return hipGetLastError();
}
hipError_t launch_head_bwd(const uint16_t *d, const float *out, const float *t, const float *wgt, const float *wh,
                           long long P, float *dd, uint16_t *gm, float *part, size_t partCap, int *nblocks,
                           hipStream_t s) {
    int b = blocks_for(P);
    if (b > 256) b = 256;
    if ((size_t)b * (kHead + 2) > partCap) return hipErrorInvalidValue;
    head_bwd_kernel<<<b, 256, 0, s>>>(d, out, t, wgt, wh, P, dd, gm, part);
    *nblocks = b;
    return hipGetLastError();
}
hipError_t launch_flip(const uint16_t *w, int cout, int cin, uint16_t *wf, hipStream_t s) {
    flip_kernel<<<blocks_for((long long)cout * 9 * cin), 256, 0, s>>>(w, cout, cin, wf);
    return hipGetLastError();
}
hipError_t launch_adam(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2,
                       float eps, int t, hipStream_t s) {
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    adam_kernel<<<blocks_for(n), 256, 0, s>>>(p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}
hipError_t launch_adam_bf16(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2,
                            float eps, int t, const WeightMap &map, uint16_t *wb, uint16_t *wf, hipStream_t s) {
    if (map.layers < 0 || map.layers > kMaxConvLayers) return hipErrorInvalidValue;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    adam_bf16_kernel<<<blocks_for(n), 256, 0, s>>>(p, g, m, v, n, lr, b1, b2, eps, c1, c2, map, wb, wf);
    return hipGetLastError();
}
hipError_t launch_weights_bf16(const float *p, const WeightMap &map, uint16_t *wb, uint16_t *wf, hipStream_t s) {
    if (map.layers < 1 || map.layers > kMaxConvLayers) return hipErrorInvalidValue;
    const int L = map.layers - 1;
    const long long n = map.w_off[L] + (long long)map.cout[L] * 9 * map.cin[L];
    weights_bf16_kernel<<<blocks_for(n), 256, 0, s>>>(p, n, map, wb, wf);
    return hipGetLastError();
}
hipError_t launch_to_bf16(const float *x, long long n, uint16_t *y, hipStream_t s) {
    to_bf16_kernel<<<blocks_for(n), 256, 0, s>>>(x, n, y);
    return hipGetLastError();
}
hipError_t launch_fill_f32(float *a, long long n, float v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    fill_f32_kernel<<<blocks_for(n), 256, 0, s>>>(a, n, v);
    return hipGetLastError();
}

}  // namespace wc
