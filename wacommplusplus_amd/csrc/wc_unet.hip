// Kernels of the concentration-reconstruction network (a 4-level U-Net with
// 32 base channels) for gfx950: training and inference on BF16 matrix cores
// with FP32 accumulation.  No counterpart in the reference.
//
// Layout: activations NHWC in BF16 (round to nearest even), activation
// gradients NHWC in FP32, weights [cout][3][3][cin] in BF16 (FP32 master
// copies live in the optimiser).  Every channel count is a multiple of 32.
//
// conv3x3_kernel: 3x3, stride 1, zero padding 1, as an implicit GEMM
//   out[p][co] = sum_{tap, ci} in[p + tap][ci] * w[co][tap][ci]
// on v_mfma_f32_16x16x32_bf16.  One wave computes a 32-pixel x 32-channel
// tile (2 x 2 MFMA tiles); each lane loads 8 consecutive channels (16 B) of
// one pixel for A and of one output channel's weights for B, straight from
// global memory (NHWC makes both contiguous).  The same kernel serves
//   - forward: two inputs concatenated along channels (decoder skip + up),
//     bias + ReLU, BF16 out;
//   - backward data: the output gradient (FP32, masked by the ReLU of the
//     forward output) convolved with the flipped, transposed weights,
//     FP32 out split over two tensors (the concat's two inputs), stored or
//     accumulated.
// MFMA operand layout (16x16x32 bf16): lane l holds A[l%16][8*(l/16)..+7]
// and B[8*(l/16)..+7][l%16]; the accumulator holds D[4*(l/16)+r][l%16],
// r = 0..3.
//
// wgrad_kernel: dW[co][tap][ci] += sum_p dYm[p][co] * X[p + tap][ci], a GEMM
// whose reduction runs over pixels: a workgroup stages 64 pixels of dYm
// (32 co) and X (32 ci) transposed in LDS, four waves take the 2 x 2 16x16
// tiles, and each workgroup adds its pixel range's partial sums into dW with
// FP32 atomics (split-K).
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
            float g = 0.f;
            if (p < pend && co0 + c < a.cout) {
                const size_t o = (size_t)p * a.cout + co0 + c;
                g = a.dy[o];
                if (!(from_bf16(a.y[o]) > 0.f)) g = 0.f;
            }
            sdy[c][pl] = to_bf16(g);
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

// db[co] += sum_p dY[p][co] * (Y[p][co] > 0); one block per 32 channels x pixel slice
__global__ __launch_bounds__(256) void bias_grad_kernel(const float *__restrict__ dy, const uint16_t *__restrict__ y,
                                                        long long P, int cout, float *__restrict__ db) {
    const int c = blockIdx.x * 32 + (threadIdx.x & 31);
    const int lanegrp = threadIdx.x >> 5;                        // 8 pixel lanes per channel
    float s = 0.f;
    if (c < cout)
        for (long long p = (long long)blockIdx.y * 8 + lanegrp; p < P; p += (long long)gridDim.y * 8) {
            const size_t o = (size_t)p * cout + c;
            if (from_bf16(y[o]) > 0.f) s += dy[o];
        }
    __shared__ float part[8][32];
    part[lanegrp][threadIdx.x & 31] = s;
    __syncthreads();
    if (threadIdx.x < 32 && c < cout) {
        float t = 0.f;
        for (int g = 0; g < 8; ++g) t += part[g][threadIdx.x];
        if (t != 0.f) atomicAdd(&db[c], t);
    }
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

// network input: x [n][h][w] FP32 -> NHWC BF16 with `c` channels, channel 0 = x
__global__ void input_kernel(const float *__restrict__ x, long long P, int c, uint16_t *__restrict__ y) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < P * c;
         e += (long long)gridDim.x * blockDim.x) {
        const long long p = e / c;
        y[e] = (e - p * c) == 0 ? to_bf16(x[p]) : (uint16_t)0;
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

// loss += mean (out - t)^2; g = d loss / d out; head gradients dwh;
// dd[p][c] = g[p] wh[c]
__global__ __launch_bounds__(256) void head_bwd_kernel(const uint16_t *__restrict__ d, const float *__restrict__ out,
                                                       const float *__restrict__ t, const float *__restrict__ wh,
                                                       long long P, float *__restrict__ dd,
                                                       float *__restrict__ dwh, float *__restrict__ loss) {
    __shared__ float part[256];
    const float inv = 1.0f / (float)P;
    float l = 0.f;
    float gw[kHead + 1];
    for (int q = 0; q <= kHead; ++q) gw[q] = 0.f;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < P; p += (long long)gridDim.x * blockDim.x) {
        const float e = out[p] - t[p];
        l += e * e;
        const float g = 2.0f * e * inv;
        for (int q = 0; q < kHead; ++q) {
            gw[q] += g * from_bf16(d[p * kHead + q]);
            dd[p * kHead + q] = g * wh[q];
        }
        gw[kHead] += g;
    }
    part[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, part[0] * inv);
    for (int q = 0; q <= kHead; ++q) {
        __syncthreads();
        part[threadIdx.x] = gw[q];
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0 && part[0] != 0.f) atomicAdd(&dwh[q], part[0]);
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

__global__ void to_bf16_kernel(const float *__restrict__ x, long long n, uint16_t *__restrict__ y) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        y[e] = to_bf16(x[e]);
}

__global__ void fill_f32_kernel(float *__restrict__ a, long long n, float v) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        a[e] = v;
}

int blocks_for(long long n) {
    long long b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    return (int)(b > 0 ? b : 1);
}

}  // namespace

hipError_t launch_conv3x3(const ConvArgs &a, bool fp32_input, hipStream_t s) {
    const long long P = (long long)a.n * a.h * a.w;
    const int nout = a.cy1 + a.cy2;
    if (P <= 0 || nout <= 0) return hipSuccess;
    if ((a.c1 % 32) || (a.c2 % 32) || (nout % 32) || ((a.cy1 % 32) && a.cy2)) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((P + 127) / 128), (unsigned)(nout / 32));
    if (fp32_input)
        conv3x3_kernel<float><<<grid, 256, 0, s>>>(a);
    else
        conv3x3_kernel<uint16_t><<<grid, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_wgrad(const WgradArgs &a, hipStream_t s) {
    const int cin = a.c1 + a.c2;
    if ((cin % 32) || (a.cout % 32) || a.splits < 1) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(a.cout / 32), (unsigned)(cin / 32), (unsigned)(9 * a.splits));
    wgrad_kernel<<<grid, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_bias_grad(const float *dy, const uint16_t *y, long long P, int cout, float *db, hipStream_t s) {
    const dim3 grid((unsigned)((cout + 31) / 32), 256);
    bias_grad_kernel<<<grid, 256, 0, s>>>(dy, y, P, cout, db);
    return hipGetLastError();
}

hipError_t launch_pool_fwd(const uint16_t *x, int n, int h, int w, int c, uint16_t *y, hipStream_t s) {
    pool_fwd_kernel<<<blocks_for((long long)n * (h / 2) * (w / 2) * c), 256, 0, s>>>(x, n, h, w, c, y);
    return hipGetLastError();
}
hipError_t launch_pool_bwd(const float *dy, int n, int h, int w, int c, float *dx, bool accumulate, hipStream_t s) {
    pool_bwd_kernel<<<blocks_for((long long)n * h * w * c), 256, 0, s>>>(dy, n, h, w, c, dx, accumulate ? 1 : 0);
    return hipGetLastError();
}
hipError_t launch_up_fwd(const uint16_t *x, int n, int h, int w, int c, uint16_t *y, hipStream_t s) {
    up_fwd_kernel<<<blocks_for((long long)n * 4 * h * w * c), 256, 0, s>>>(x, n, h, w, c, y);
    return hipGetLastError();
}
hipError_t launch_up_bwd(const float *dy, int n, int h, int w, int c, float *dx, hipStream_t s) {
    up_bwd_kernel<<<blocks_for((long long)n * h * w * c), 256, 0, s>>>(dy, n, h, w, c, dx);
    return hipGetLastError();
}
hipError_t launch_unet_input(const float *x, long long P, int c, uint16_t *y, hipStream_t s) {
    input_kernel<<<blocks_for(P * c), 256, 0, s>>>(x, P, c, y);
    return hipGetLastError();
}
hipError_t launch_head_fwd(const uint16_t *d, const float *x, const float *wh, long long P, float *out,
                           hipStream_t s) {
    head_fwd_kernel<<<blocks_for(P), 256, 0, s>>>(d, x, wh, P, out);
    return hipGetLastError();
}
hipError_t launch_head_bwd(const uint16_t *d, const float *out, const float *t, const float *wh, long long P,
                           float *dd, float *dwh, float *loss, hipStream_t s) {
    int b = blocks_for(P);
    if (b > 1024) b = 1024;
    head_bwd_kernel<<<b, 256, 0, s>>>(d, out, t, wh, P, dd, dwh, loss);
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
