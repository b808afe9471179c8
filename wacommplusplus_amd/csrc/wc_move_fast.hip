// Mixed-precision move kernel for gfx950: the same operator as Particle::move
// (Particle.cpp:71-824: time-lerped bilinear u, v at the containing rho level,
// trilinear w and AKt over the two W levels, explicit midpoint or Euler step,
// keyed random displacement, surface/bottom/shore/open-boundary closures,
// exponential decay), computed in FP32 from FP32, FP16 or BF16 forcing.
//
// Gather layout.  The move is bound by the number of scattered load
// instructions, not by bytes, so once per interval a pack kernel interleaves
// the interval's two forcing records into per-grid-point structs that one
// naturally aligned load fetches:
//   UV[k][j][i] = {u(t0), u(t1), v(t0), v(t1)}                 rho level k
//   WA[k][j][i] = {w_k(t0), w_k(t1), w_k-1(t0), w_k-1(t1),
//                  a_k(t0), a_k(t1), a_k-1(t0), a_k-1(t1)}     W levels k, k-1
//   Z [j][i]    = {zeta(t0), zeta(t1), h, -}                   (FP32)
//   MD[j][i]    = {1/dx, 1/dy}                                   (FP32, static)
// A bilinear stage is then 4 UV + 4 WA + 4 Z loads instead of ~60.
//
// State: each work-item keeps its particle as an integer cell (jI, iI) plus
// FP32 offsets in [0,1) and an FP32 level k for a whole forcing interval in
// registers; elapsed time and age stay FP64.  The FP64 SoA state is read once
// and written once per interval (j = jI + offset is exact in FP64).  Health is
// evaluated once per interval: exp(-age/tau0) < survprob is equivalent to
// age > -tau0*log(survprob), tested every substep.
//
// Diffusion draws: Philox4x32-10 with key = seed, counter = (substep,
// interval key, id lo, id hi); one call gives the four uniforms of two
// Box-Muller pairs (three normals used).  Statistically equivalent to the
// reference's SplitMix64 draws, not draw-for-draw identical.
#include <hip/hip_fp16.h>

#include <type_traits>

#include "wc_launch.h"
#include "wc_philox.h"

// Multiply-adds are fused only within one source expression (contract(on),
// not fast): fusion across statements depends on what the substep loop is
// inlined into, and near-zero water columns (coast node lines, where the
// reference's vertical step is 0/0) would then end differently in the
// context kernel and the ensemble kernel.  With it, both give the same bits.
#pragma clang fp contract(on)

namespace wc {
namespace {

template <class T>
struct alignas(4 * sizeof(T)) UV4 {
    T v[4];
};
template <class T>
struct alignas(8 * sizeof(T)) WA8 {
    T v[8];
};
struct alignas(16) ZH {
    float z0, z1, h, pad;
};

template <class T>
__device__ __forceinline__ float tof(T x);
template <>
__device__ __forceinline__ float tof<float>(float x) { return x; }
template <>
__device__ __forceinline__ float tof<_Float16>(_Float16 x) { return (float)x; }
template <>
__device__ __forceinline__ float tof<uint16_t>(uint16_t x) { return __uint_as_float((uint32_t)x << 16); }

template <class T>
__device__ __forceinline__ T fromf(float x);
template <>
__device__ __forceinline__ float fromf<float>(float x) { return x; }
template <>
__device__ __forceinline__ _Float16 fromf<_Float16>(float x) { return (_Float16)x; }
template <>
__device__ __forceinline__ uint16_t fromf<uint16_t>(float x) {
    const uint32_t b = __float_as_uint(x);   // round to nearest even (finite inputs)
    return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float lerpf(float a, float b, float t) { return fmaf(t, b - a, a); }

// The mixed-precision path is held to the FP64 path statistically, not bit
// for bit, so its quotients and square roots use the hardware reciprocal and
// square root (v_rcp_f32, v_sqrt_f32: 1 ulp) instead of the IEEE sequences.
// A zero divisor gives inf, so 0/0 still ends as NaN, as with the division.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Corner arithmetic of the samplers (shared by the global and the LDS-tiled
// sampler, so both give the same bits).
struct W4 {
    float w00, w01, w10, w11;
};
__device__ __forceinline__ W4 weights(float fi, float fj) {
    const float gi = 1.0f - fi, gj = 1.0f - fj;
    return W4{gi * gj, fi * gj, gi * fj, fi * fj};
}
__device__ __forceinline__ float hc4(const ZH &a, const ZH &b, const ZH &d, const ZH &e, const W4 &w, float al) {
    const float zz = w.w00 * lerpf(a.z0, a.z1, al) + w.w01 * lerpf(b.z0, b.z1, al) +
                     w.w10 * lerpf(d.z0, d.z1, al) + w.w11 * lerpf(e.z0, e.z1, al);
    const float hh = w.w00 * a.h + w.w01 * b.h + w.w10 * d.h + w.w11 * e.h;
    return hh + zz;
}
template <class T>
__device__ __forceinline__ void uv4(const UV4<T> &a, const UV4<T> &b, const UV4<T> &d, const UV4<T> &e, const W4 &w,
                                    float al, float &uu, float &vv) {
    uu = w.w00 * lerpf(tof(a.v[0]), tof(a.v[1]), al) + w.w01 * lerpf(tof(b.v[0]), tof(b.v[1]), al) +
         w.w10 * lerpf(tof(d.v[0]), tof(d.v[1]), al) + w.w11 * lerpf(tof(e.v[0]), tof(e.v[1]), al);
    vv = w.w00 * lerpf(tof(a.v[2]), tof(a.v[3]), al) + w.w01 * lerpf(tof(b.v[2]), tof(b.v[3]), al) +
         w.w10 * lerpf(tof(d.v[2]), tof(d.v[3]), al) + w.w11 * lerpf(tof(e.v[2]), tof(e.v[3]), al);
}
// w (and AKt when wantA) at W levels kI (weight 1-kF) and kI-1 (weight kF)
template <class T>
__device__ __forceinline__ void wa4(const WA8<T> &a, const WA8<T> &b, const WA8<T> &d, const WA8<T> &e, float kF,
                                    const W4 &w, float alw, float ala, bool wantA, float &ww, float &aa) {
    auto corner = [&](const WA8<T> &q, float &wc, float &ac) {
        const float wt = lerpf(tof(q.v[0]), tof(q.v[1]), alw), wb = lerpf(tof(q.v[2]), tof(q.v[3]), alw);
        wc = lerpf(wt, wb, kF);
        if (wantA) {
            const float at = lerpf(tof(q.v[4]), tof(q.v[5]), ala), ab = lerpf(tof(q.v[6]), tof(q.v[7]), ala);
            ac = lerpf(at, ab, kF);
        } else {
            ac = 0.f;
        }
    };
    float w0, w1, w2, w3, a0, a1, a2, a3;
    corner(a, w0, a0);
    corner(b, w1, a1);
    corner(d, w2, a2);
    corner(e, w3, a3);
    ww = w.w00 * w0 + w.w01 * w1 + w.w10 * w2 + w.w11 * w3;
    aa = w.w00 * a0 + w.w01 * a1 + w.w10 * a2 + w.w11 * a3;
}

// The interval's gather layout in global memory.  Stencils at cell (j, i)
// read (j, i), (j, i+1), (j+1, i), (j+1, i+1).
template <class T>
struct Packed {
    const UV4<T> *uv;     // [s_w-1 levels: k = kLL..0][eta][xi]   (rho levels)
    const WA8<T> *wa;     // [s_w-1 levels: k = kLL..0][eta][xi]   (W levels k, k-1)
    const ZH *z;          // [eta][xi]
    const float2 *md;     // [eta][xi] {1/dx, 1/dy}
    const float *depthf;
    int xi;
    int nlev;             // s_w - 1
    size_t plane;

    // level index of logical level k (k = -(nlev-1) .. 0)
    __device__ __forceinline__ size_t lev(int k) const { return (size_t)(k + nlev - 1) * plane; }
    __device__ __forceinline__ size_t cell(int j, int i) const { return (size_t)j * xi + i; }

    // corners are copied whole (one 16-byte or 8-byte load each), not field by field
    __device__ __forceinline__ float hc(int j, int i, const W4 &w, float al) const {
        const size_t c = cell(j, i);
        const ZH a = z[c], b = z[c + 1], d = z[c + xi], e = z[c + xi + 1];
        return hc4(a, b, d, e, w, al);
    }
    __device__ __forceinline__ void uv_at(int kI, int j, int i, const W4 &w, float al, float &uu, float &vv) const {
        const UV4<T> *p = uv + lev(kI) + cell(j, i);
        const UV4<T> a = p[0], b = p[1], d = p[xi], e = p[xi + 1];
        uv4<T>(a, b, d, e, w, al, uu, vv);
    }
    __device__ __forceinline__ void wa_at(int kI, float kF, int j, int i, const W4 &w, float alw, float ala,
                                          bool wantA, float &ww, float &aa) const {
        const WA8<T> *p = wa + lev(kI) + cell(j, i);
        const WA8<T> a = p[0], b = p[1], d = p[xi], e = p[xi + 1];
        wa4<T>(a, b, d, e, kF, w, alw, ala, wantA, ww, aa);
    }
    __device__ __forceinline__ float2 md_at(int j, int i) const { return md[cell(j, i)]; }
    __device__ __forceinline__ float depth_at(int q) const { return depthf[q]; }
    // a corner of cell (j, i) is land (h = 0)
    __device__ __forceinline__ bool coastal(int j, int i) const {
        const size_t c = cell(j, i);
        return z[c].h == 0.f || z[c + 1].h == 0.f || z[c + xi].h == 0.f || z[c + xi + 1].h == 0.f;
    }
};

// The same layout with a box of cells [j0, j0+bh) x [i0, i0+bw) and levels
// [l0, l0+nl) (level indices) staged in LDS by the workgroup: a stencil
// inside the box reads LDS, any other reads global memory, so the values and
// the results are the same wherever a particle goes.
template <class T>
struct TilePacked {
    Packed<T> g;
    const WA8<T> *wL;     // [nl][bh][bw]
    const ZH *zL;         // [bh][bw]
    const UV4<T> *uL;     // [nl][bh][bw]
    const float2 *mL;     // [bh][bw]
    int j0, i0, bh, bw, l0, nl;

    __device__ __forceinline__ bool stencil_in(int j, int i) const {
        return (unsigned)(j - j0) < (unsigned)(bh - 1) && (unsigned)(i - i0) < (unsigned)(bw - 1);
    }
    __device__ __forceinline__ bool level_in(int kI) const { return (unsigned)(kI + g.nlev - 1 - l0) < (unsigned)nl; }
    __device__ __forceinline__ int local(int j, int i) const { return (j - j0) * bw + (i - i0); }
    __device__ __forceinline__ int local(int kI, int j, int i) const {
        return ((kI + g.nlev - 1 - l0) * bh + (j - j0)) * bw + (i - i0);
    }

    __device__ __forceinline__ float hc(int j, int i, const W4 &w, float al) const {
        if (!stencil_in(j, i)) return g.hc(j, i, w, al);
        const ZH *p = zL + local(j, i);
        const ZH a = p[0], b = p[1], d = p[bw], e = p[bw + 1];
        return hc4(a, b, d, e, w, al);
    }
    __device__ __forceinline__ void uv_at(int kI, int j, int i, const W4 &w, float al, float &uu, float &vv) const {
        if (!(stencil_in(j, i) && level_in(kI))) return g.uv_at(kI, j, i, w, al, uu, vv);
        const UV4<T> *p = uL + local(kI, j, i);
        const UV4<T> a = p[0], b = p[1], d = p[bw], e = p[bw + 1];
        uv4<T>(a, b, d, e, w, al, uu, vv);
    }
    __device__ __forceinline__ void wa_at(int kI, float kF, int j, int i, const W4 &w, float alw, float ala,
                                          bool wantA, float &ww, float &aa) const {
        if (!(stencil_in(j, i) && level_in(kI))) return g.wa_at(kI, kF, j, i, w, alw, ala, wantA, ww, aa);
        const WA8<T> *p = wL + local(kI, j, i);
        const WA8<T> a = p[0], b = p[1], d = p[bw], e = p[bw + 1];
        wa4<T>(a, b, d, e, kF, w, alw, ala, wantA, ww, aa);
    }
    __device__ __forceinline__ float2 md_at(int j, int i) const {
        if ((unsigned)(j - j0) < (unsigned)bh && (unsigned)(i - i0) < (unsigned)bw) return mL[local(j, i)];
        return g.md_at(j, i);
    }
    __device__ __forceinline__ float depth_at(int q) const { return g.depthf[q]; }
};

// Tile-mean storage (WC_STORE_BF16_TILE, WC_STORE_F16_TILE): u, v and w are
// stored as 16-bit residuals from a mean per tile of kTile x kTile cells (per
// level, field and record; see tile_pack_strip_kernel).  The cells [kTile*tj, kTile*tj + kTile) of a
// tile read nodes up to kTile*tj + kTile, so each tile keeps its own copy of
// (kTile+1)^2 nodes and a stencil reads one tile and one mean:
//   UVR[lev][tile][kTile+1][kTile+1] = residuals {u(t0), u(t1), v(t0), v(t1)}
//   WAR[lev][tile][kTile+1][kTile+1] = residual w at W levels k, k-1 of both
//                                      records, then AKt as stored (no mean)
//   TM [lev][tile] = means {u(t0), u(t1), v(t0), v(t1),
//                           w_k(t0), w_k(t1), w_k-1(t0), w_k-1(t1)}, in T
//                           (16 bytes: one load serves a stencil's u, v, w)
// A corner's value is its time-interpolated residual plus the time-
// interpolated mean.  The mean is a 16-bit value, so a zero (land, where
// u = v = w = 0, ROMSAdapter.cpp:244-264) is stored as the residual -mean
// exactly and decodes to exactly 0, as with the other storage types.
constexpr int kTile = 4;
constexpr int kTileNodes = (kTile + 1) * (kTile + 1);

// Geometry shared by the tiled samplers: the tile of a cell, the node of a
// stencil's first corner inside the tile's (kTile+1)^2 block, zeta/h, metric.
struct TileBase {
    const ZH *z;
    const float2 *md;
    const float *depthf;
    int xi;
    int nlev;
    int tilesX;           // tiles per row: ceil(xi / kTile)
    size_t tplane;        // tiles per level

    __device__ __forceinline__ size_t cell(int j, int i) const { return (size_t)j * xi + i; }
    // j, i >= 0 (stencils are read at valid cells only)
    __device__ __forceinline__ size_t tile(int kI, int j, int i) const {
        return (size_t)(kI + nlev - 1) * tplane + (size_t)((unsigned)j / kTile) * tilesX + (unsigned)i / kTile;
    }
    __device__ __forceinline__ unsigned node(int j, int i) const {
        return ((unsigned)j % kTile) * (kTile + 1) + (unsigned)i % kTile;
    }
    __device__ __forceinline__ float hc(int j, int i, const W4 &w, float al) const {
        const size_t c = cell(j, i);
        const ZH a = z[c], b = z[c + 1], d = z[c + xi], e = z[c + xi + 1];
        return hc4(a, b, d, e, w, al);
    }
    __device__ __forceinline__ float2 md_at(int j, int i) const { return md[cell(j, i)]; }
    __device__ __forceinline__ float depth_at(int q) const { return depthf[q]; }
    __device__ __forceinline__ bool coastal(int j, int i) const {
        const size_t c = cell(j, i);
        return z[c].h == 0.f || z[c + 1].h == 0.f || z[c + xi].h == 0.f || z[c + xi + 1].h == 0.f;
    }
};

template <class T>
struct TiledPacked : TileBase {
    const UV4<T> *uvr;
    const WA8<T> *war;
    const WA8<T> *tm;

    __device__ __forceinline__ void uv_at(int kI, int j, int i, const W4 &w, float al, float &uu, float &vv) const {
        const size_t t = tile(kI, j, i);
        const UV4<T> *p = uvr + t * kTileNodes + node(j, i);
        const UV4<T> a = p[0], b = p[1], d = p[kTile + 1], e = p[kTile + 2];
        const WA8<T> m = tm[t];
        const float Lu = lerpf(tof(m.v[0]), tof(m.v[1]), al), Lv = lerpf(tof(m.v[2]), tof(m.v[3]), al);
        auto cu = [&](const UV4<T> &q) { return lerpf(tof(q.v[0]), tof(q.v[1]), al) + Lu; };
        auto cv = [&](const UV4<T> &q) { return lerpf(tof(q.v[2]), tof(q.v[3]), al) + Lv; };
        uu = w.w00 * cu(a) + w.w01 * cu(b) + w.w10 * cu(d) + w.w11 * cu(e);
        vv = w.w00 * cv(a) + w.w01 * cv(b) + w.w10 * cv(d) + w.w11 * cv(e);
    }
    __device__ __forceinline__ void wa_at(int kI, float kF, int j, int i, const W4 &w, float alw, float ala,
                                          bool wantA, float &ww, float &aa) const {
        const size_t t = tile(kI, j, i);
        const WA8<T> *p = war + t * kTileNodes + node(j, i);
        const WA8<T> a = p[0], b = p[1], d = p[kTile + 1], e = p[kTile + 2];
        const WA8<T> m = tm[t];
        const float Lt = lerpf(tof(m.v[4]), tof(m.v[5]), alw), Lb = lerpf(tof(m.v[6]), tof(m.v[7]), alw);
        auto corner = [&](const WA8<T> &q, float &wc, float &ac) {
            const float wt = lerpf(tof(q.v[0]), tof(q.v[1]), alw) + Lt;
            const float wb = lerpf(tof(q.v[2]), tof(q.v[3]), alw) + Lb;
            wc = lerpf(wt, wb, kF);
            if (wantA) {
                const float at = lerpf(tof(q.v[4]), tof(q.v[5]), ala), ab = lerpf(tof(q.v[6]), tof(q.v[7]), ala);
                ac = lerpf(at, ab, kF);
            } else {
                ac = 0.f;
            }
        };
        float w0, w1, w2, w3, a0, a1, a2, a3;
        corner(a, w0, a0);
        corner(b, w1, a1);
        corner(d, w2, a2);
        corner(e, w3, a3);
        ww = w.w00 * w0 + w.w01 * w1 + w.w10 * w2 + w.w11 * w3;
        aa = w.w00 * a0 + w.w01 * a1 + w.w10 * a2 + w.w11 * a3;
    }
};

// Block-scaled storage (WC_STORE_I16_TILE): the tile layout above, each
// residual an int16 multiple q of a power-of-two scale s of its tile (one
// per field and W level, shared by both records) and each mean an integer
// multiple of s (tile_group_scale).  A corner's value is
//   s * lerp(q(t0), q(t1), al) + lerp(m(t0), m(t1), al),
// off by at most s/2 per record: about 2^-16 of the tile's largest deviation
// from its mean.  Scaling by a power of two commutes with rounding, so with
// m(t) = s * M(t) the two terms of a land node (q(t) = -M(t)) cancel exactly
// and zeros decode to exactly 0.  AKt has no mean (q = AKt / s_akt).
struct alignas(16) TileScaleUV {
    float m[4];           // u(t0), u(t1), v(t0), v(t1)
    float s[2];           // u, v
    float pad[2];
};
struct alignas(16) TileScaleWA {
    float m[4];           // w_k(t0), w_k(t1), w_k-1(t0), w_k-1(t1)
    float s[4];           // w_k, w_k-1, AKt_k, AKt_k-1
};

__device__ __forceinline__ float i16f(int16_t x) { return (float)x; }

struct TiledI16 : TileBase {
    const UV4<int16_t> *uvr;
    const WA8<int16_t> *war;
    const TileScaleUV *tuv;
    const TileScaleWA *twa;

    __device__ __forceinline__ void uv_at(int kI, int j, int i, const W4 &w, float al, float &uu, float &vv) const {
        const size_t t = tile(kI, j, i);
        const UV4<int16_t> *p = uvr + t * kTileNodes + node(j, i);
        const UV4<int16_t> a = p[0], b = p[1], d = p[kTile + 1], e = p[kTile + 2];
        const TileScaleUV m = tuv[t];
        const float Lu = lerpf(m.m[0], m.m[1], al), Lv = lerpf(m.m[2], m.m[3], al);
        auto cu = [&](const UV4<int16_t> &q) { return fmaf(m.s[0], lerpf(i16f(q.v[0]), i16f(q.v[1]), al), Lu); };
        auto cv = [&](const UV4<int16_t> &q) { return fmaf(m.s[1], lerpf(i16f(q.v[2]), i16f(q.v[3]), al), Lv); };
        uu = w.w00 * cu(a) + w.w01 * cu(b) + w.w10 * cu(d) + w.w11 * cu(e);
        vv = w.w00 * cv(a) + w.w01 * cv(b) + w.w10 * cv(d) + w.w11 * cv(e);
    }
    __device__ __forceinline__ void wa_at(int kI, float kF, int j, int i, const W4 &w, float alw, float ala,
                                          bool wantA, float &ww, float &aa) const {
        const size_t t = tile(kI, j, i);
        const WA8<int16_t> *p = war + t * kTileNodes + node(j, i);
        const WA8<int16_t> a = p[0], b = p[1], d = p[kTile + 1], e = p[kTile + 2];
        const TileScaleWA m = twa[t];
        const float Lt = lerpf(m.m[0], m.m[1], alw), Lb = lerpf(m.m[2], m.m[3], alw);
        auto corner = [&](const WA8<int16_t> &q, float &wc, float &ac) {
            const float wt = fmaf(m.s[0], lerpf(i16f(q.v[0]), i16f(q.v[1]), alw), Lt);
            const float wb = fmaf(m.s[1], lerpf(i16f(q.v[2]), i16f(q.v[3]), alw), Lb);
            wc = lerpf(wt, wb, kF);
            if (wantA) {
                const float at = m.s[2] * lerpf(i16f(q.v[4]), i16f(q.v[5]), ala);
                const float ab = m.s[3] * lerpf(i16f(q.v[6]), i16f(q.v[7]), ala);
                ac = lerpf(at, ab, kF);
            } else {
                ac = 0.f;
            }
        };
        float w0, w1, w2, w3, a0, a1, a2, a3;
        corner(a, w0, a0);
        corner(b, w1, a1);
        corner(d, w2, a2);
        corner(e, w3, a3);
        ww = w.w00 * w0 + w.w01 * w1 + w.w10 * w2 + w.w11 * w3;
        aa = w.w00 * a0 + w.w01 * a1 + w.w10 * a2 + w.w11 * a3;
    }
};

__device__ __forceinline__ bool valid_cell(int j, int i, int eta, int xi) {
    return j >= 0 && i >= 0 && j + 1 < eta && i + 1 < xi;
}

// reflectDomain (NumericalHelpers.hpp:67-74), rare path, in FP64
__device__ double reflect_domain(double c, double m) {
    if (m <= 0) return 0;
    double r = fmod(c, 2.0 * m);
    if (r < 0) r += 2.0 * m;
    if (r >= m) r = 2.0 * m - r;
    if (r >= m) r = nextafter(m, 0.0);
    return r;
}

// Split a coordinate held as cell + offset after a displacement d (in cells).
__device__ __forceinline__ void advance(int cell, float off, float d, int &ncell, float &noff) {
    const float x = off + d;
    const float fl = floorf(x);
    ncell = cell + (int)fl;
    noff = x - fl;
    if (noff >= 1.0f) {   // x just below an integer rounds up
        noff -= 1.0f;
        ncell += 1;
    }
}

// A particle of the mixed-precision path during one interval, in registers.
struct Reg {
    int jc, ic;
    float fj, fi, k;
    double health, age;
    double elapsed;       // time into the interval
    bool inside, healthFresh, changed;
    bool done;            // the substep loop has ended for this interval
    uint32_t tag;         // drift type and side (read and flipped only with the jibe on)
    uint32_t ev;          // WC_EVENT_* bits of the closures that acted (event tracking)
};

// Converts an FP64 position into cell + offset; `inside` is the reference's
// validInterpolationCell test (NumericalHelpers.hpp:12-14) and a finite k.
__device__ __forceinline__ void to_cell(double jd, double id_, double kd, int eta, int xi, Reg &q) {
    const double jfl = floor(jd), ifl = floor(id_);
    q.inside = jfl >= 0 && ifl >= 0 && jfl + 1 < eta && ifl + 1 < xi && kd == kd;
    q.jc = q.inside ? (int)jfl : 0;
    q.ic = q.inside ? (int)ifl : 0;
    q.fj = q.inside ? (float)(jd - jfl) : 0.f;
    q.fi = q.inside ? (float)(id_ - ifl) : 0.f;
    if (q.fj >= 1.0f) { q.fj -= 1.0f; ++q.jc; }
    if (q.fi >= 1.0f) { q.fi -= 1.0f; ++q.ic; }
}

// The substeps of one forcing interval for one particle (Particle.cpp:
// 178-815) on the cell + offset state; sigma is the horizontal diffusion
// scale (config sigma, or the particle's scenario sigma in an ensemble).
// Runs the substeps that start before tEnd (the interval length, or the end
// of a chunk of it: a later call continues from q.elapsed); q.done marks the
// end of the particle's loop for the interval.  Returns the live substeps
// taken.  S is a Packed<T> or a TilePacked<T>.  Jibe: apply the hourly side
// flip of a non-passive particle (q.tag) after each live substep, with the
// reference's keyed draw.
template <class T, class Smp, bool Jibe = false, bool Ev = false>
__device__ __forceinline__ unsigned run_interval(const MoveParams &P, const Smp &S, float sigma, uint64_t id,
                                                 double ptime, Reg &q, double tEnd) {
    unsigned steps = 0;
    if (q.done) return 0;
    int jc = q.jc, ic = q.ic;
    float fj = q.fj, fi = q.fi, k = q.k;
    double health = q.health, age = q.age;
    bool inside = q.inside, healthFresh = q.healthFresh, changed = q.changed, done = false;
    const float kLL = (float)P.kLowerLimit;
    const int eta = P.eta, xi = P.xi;
    const float dir = (float)P.direction;
    const float sv = (float)P.sv, shore = (float)P.shoreLimit;
    const uint32_t seedLo = (uint32_t)P.seed, seedHi = (uint32_t)(P.seed >> 32);
    const uint32_t ik = (uint32_t)P.intervalKey ^ (uint32_t)((uint64_t)P.intervalKey >> 32);
    const double L = P.intervalLength;
    const double invL = L == 0 ? 0.0 : 1.0 / L, invDti = 1.0 / P.dti;
    double elapsed = q.elapsed;

        while (elapsed < tEnd) {
            const double stepDt = fmin(P.dti, L - elapsed);
            const double physicalTime = P.intervalStart + P.dirInt * elapsed;
            if (P.dirInt > 0 && ptime > physicalTime) {           // not yet emitted (Particle.cpp:211-215)
                elapsed = fmin(L, ptime - P.intervalStart);
                continue;
            }
            if (healthFresh ? (health < P.survprob) : (age > P.ageKill)) {   // :218-223
                health = -1;
                healthFresh = true;
                changed = true;
                done = true;
                break;
            }
            // open boundary (:241-248); a level outside [kLL, 0] (reachable
            // only from a non-finite step or an unknown closure code) ends the
            // particle as the FP64 path's level guard does (kernel.cu:195)
            if (!inside || !(k <= 0.f && k >= kLL)) {
                health = -1;
                healthFresh = true;
                changed = true;
                done = true;
                break;
            }
            ++steps;
            changed = true;
            const float alpha = (float)((elapsed + 0.5 * stepDt) * invL);
            const float alphaStart = (float)(elapsed * invL);
            const float dt = (float)stepDt;
            const int kI = (int)ceilf(k);
            const float kF = (float)kI - k;
            const W4 w0 = weights(fi, fj);

            const float hc = S.hc(jc, ic, w0, alpha);
            bool kill = false;
            if (hc <= shore && P.horizontal == WC_CLOSURE_KILL) kill = true;
            if constexpr (Ev) {
                if (hc <= shore) q.ev |= WC_EVENT_SHALLOW;
                if (S.coastal(jc, ic)) q.ev |= WC_EVENT_COAST;
            }

            const float a1 = P.midpoint ? alphaStart : alpha;
            float uu, vv, ww, aa;
            S.uv_at(kI, jc, ic, w0, a1, uu, vv);
            S.wa_at(kI, kF, jc, ic, w0, a1, alpha, P.randomOn != 0, ww, aa);
            float rx = 0, ry = 0, rz = 0;
            if (P.randomOn) {
                const float sigmaDepth = sigma * (1.0f - k / kLL);
                const uint32_t substep = (uint32_t)floor(elapsed * invDti + 1e-9);
                const Philox4 r = philox4x32_10(substep, ik, (uint32_t)id, (uint32_t)(id >> 32), seedLo, seedHi);
                const float u1 = ((float)(r.v[0] >> 8) + 0.5f) * 0x1.0p-24f;
                const float u2 = (float)(r.v[1] >> 8) * 0x1.0p-24f;
                const float u3 = ((float)(r.v[2] >> 8) + 0.5f) * 0x1.0p-24f;
                const float u4 = (float)(r.v[3] >> 8) * 0x1.0p-24f;
                const float ra = fsqrt(-2.0f * __logf(u1)), rb = fsqrt(-2.0f * __logf(u3));
                float s1, c1, s2, c2;
                __sincosf(6.2831853071795865f * u2, &s1, &c1);
                __sincosf(6.2831853071795865f * u4, &s2, &c2);
                const float scale = fsqrt((float)(stepDt * invDti)) * sigmaDepth;
                rx = scale * ra * c1;
                ry = scale * ra * s1;
                rz = scale * rb * c2 * aa * (float)P.crid;
            }
            const float2 md = S.md_at(jc, ic);
            const float idx_ = md.x, idy_ = md.y;
            const float zdist = hc * S.depth_at(kI + P.s_w - 2);
            // Vertical step in levels: the leap capped at one cell height and
            // divided by it (Particle.cpp:585-594, 644-657).  Where the water
            // column has no depth (hc == 0: a particle exactly on a land node
            // line) the reference arithmetic gives 0/0; as in the FP64 path
            // the non-finite level then ends the particle at its next substep.
            auto levels = [&](float leap) -> float {
                if (fabsf(leap) > zdist) leap = copysignf(fabsf(zdist), leap);
                return leap * frcp(zdist);
            };
            float dxl = dir * uu * dt, dyl = dir * vv * dt, dzl = dir * (sv + ww) * dt;

            if (P.midpoint) {                                      // Particle.cpp:585-641
                const float kHalf = dir * (sv + ww) * 0.5f * dt;
                int jm, im;
                float fjm, fim;
                advance(jc, fj, dir * vv * 0.5f * dt * idy_, jm, fjm);
                advance(ic, fi, dir * uu * 0.5f * dt * idx_, im, fim);
                const float kMid = fminf(0.0f, fmaxf(kLL, k + levels(kHalf)));
                bool midValid = valid_cell(jm, im, eta, xi);
                float hcMid = 0;
                W4 wm{};
                if (midValid) {
                    wm = weights(fim, fjm);
                    hcMid = S.hc(jm, im, wm, alpha);
                    midValid = hcMid > shore && hcMid > 0;
                }
                if (Ev && !midValid) q.ev |= WC_EVENT_FALLBACK;
                if (midValid) {
                    const int kMI = (int)ceilf(kMid);
                    const float kMF = (float)kMI - kMid;
                    float um, vm, wmv, dummy;
                    S.uv_at(kMI, jm, im, wm, alpha, um, vm);
                    S.wa_at(kMI, kMF, jm, im, wm, alpha, alpha, false, wmv, dummy);
                    // leap in metres of the start cell: u* dt dx/dx*
                    const float2 mdm = S.md_at(jm, im);
                    dxl = dir * um * dt * (mdm.x * frcp(idx_));
                    dyl = dir * vm * dt * (mdm.y * frcp(idy_));
                    dzl = dir * (sv + wmv) * dt * (zdist * frcp(hcMid * S.depth_at(kMI + P.s_w - 2)));
                } else {
                    float ue, ve, we, dummy;
                    S.uv_at(kI, jc, ic, w0, alpha, ue, ve);
                    S.wa_at(kI, kF, jc, ic, w0, alpha, alpha, false, we, dummy);
                    dxl = dir * ue * dt;
                    dyl = dir * ve * dt;
                    dzl = dir * (sv + we) * dt;
                }
            }
            const float xleap = dxl + rx, yleap = dyl + ry;
            const float zleap = dzl + rz;                          // :644-648 (cap in levels())
            int jn, in_;
            float fjn, fin;
            advance(jc, fj, yleap * idy_, jn, fjn);
            advance(ic, fi, xleap * idx_, in_, fin);
            float kdet = k + levels(zleap);
            if (Ev && (kdet > 0.f || kdet < kLL)) q.ev |= WC_EVENT_VERTICAL;
            if (kdet > 0.f) {                                      // :665-704
                if (P.upper == WC_CLOSURE_CONSTRAINT) kdet = 0.f;
                else if (P.upper == WC_CLOSURE_KILL) kill = true;
                else if (P.upper == WC_CLOSURE_REFLECTION) kdet = -kdet;
            }
            if (kdet < kLL) {
                if (P.lower == WC_CLOSURE_CONSTRAINT) kdet = kLL;
                else if (P.lower == WC_CLOSURE_KILL) kill = true;
                else if (P.lower == WC_CLOSURE_REFLECTION) kdet = 2.0f * kLL - kdet;
            }
            if (valid_cell(jn, in_, eta, xi)) {                    // :709-776
                const float hcdet = S.hc(jn, in_, weights(fin, fjn), alpha);
                if (hcdet <= shore) {
                    if (Ev) q.ev |= WC_EVENT_SHORE;
                    if (P.horizontal == WC_CLOSURE_CONSTRAINT) {
                        jn = jc; fjn = fj; in_ = ic; fin = fi;
                    } else if (P.horizontal == WC_CLOSURE_KILL) {
                        kill = true;
                    } else if (P.horizontal == WC_CLOSURE_REFLECTION) {
                        // reflectCell (NumericalHelpers.hpp:62-65) relative to the old cell:
                        // behind -> old cell + |old - candidate|; ahead -> candidate cell - fraction
                        if (in_ < ic) advance(ic, 0.0f, fabsf(fi - ((float)(in_ - ic) + fin)), in_, fin);
                        else if (in_ > ic) advance(in_, 0.0f, -fin, in_, fin);
                        if (jn < jc) advance(jc, 0.0f, fabsf(fj - ((float)(jn - jc) + fjn)), jn, fjn);
                        else if (jn > jc) advance(jn, 0.0f, -fjn, jn, fjn);
                    }
                }
                jc = jn; fj = fjn; ic = in_; fi = fin; k = kdet;
            } else {                                               // :777-791
                if (Ev) q.ev |= WC_EVENT_DOMAIN;
                if (P.horizontal == WC_CLOSURE_CONSTRAINT) {
                    k = kdet;
                } else if (P.horizontal == WC_CLOSURE_KILL) {
                    kill = true;
                } else if (P.horizontal == WC_CLOSURE_REFLECTION) {
                    const double ri = reflect_domain((double)in_ + (double)fin, xi - 1);
                    const double rj = reflect_domain((double)jn + (double)fjn, eta - 1);
                    const double rif = floor(ri), rjf = floor(rj);
                    ic = (int)rif; fi = (float)(ri - rif);
                    jc = (int)rjf; fj = (float)(rj - rjf);
                    k = kdet;
                }
            }
            inside = valid_cell(jc, ic, eta, xi) && fj == fj && fi == fi;
            if (kill) {
                health = -1;
                healthFresh = true;
            } else if (!(healthFresh && health <= 0)) {            // :797-804
                age = age + stepDt;
                healthFresh = false;
                if (Jibe && tag_type(q.tag) != 0) {                // :805-813
                    const uint64_t prefix = mix64(mix64(mix64(P.seed) ^ id) ^ (uint64_t)P.intervalKey);
                    if (jibe_flips(P, prefix, elapsed, stepDt)) q.tag = flip_side(q.tag);
                }
            }
            elapsed += stepDt;
        }
        q.jc = jc; q.ic = ic; q.fj = fj; q.fi = fi; q.k = k;
    q.health = health; q.age = age;
    q.inside = inside; q.healthFresh = healthFresh; q.changed = changed;
    q.elapsed = elapsed;
    q.done = done;
    return steps;
}

// perm (optional): work-item t moves particle perm[t] (a cell-tile order)
template <class T, class Smp, bool Jibe, bool Ev>
__global__ __launch_bounds__(256) void move_fast_kernel(MoveParams P, Smp S, SoA st, int n,
                                                        unsigned long long *stepCounter,
                                                        const uint32_t *__restrict__ perm) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned steps = 0;
    if (t < n) {
        const int idx = perm ? (int)perm[t] : t;
        const uint64_t id = st.id[idx];
        double kd = st.k[idx];
        const double jd = st.j[idx], id_ = st.i[idx];
        double health = st.health[idx];
        const double ptime = st.time[idx];

        // pre-step closures (Particle.cpp:134-176)
        Reg q;
        q.ev = (Ev && (kd > 0 || kd < P.kLowerLimit)) ? (uint32_t)WC_EVENT_VERTICAL : 0u;
        if (kd > 0) {
            if (P.upper == WC_CLOSURE_CONSTRAINT) kd = 0;
            else if (P.upper == WC_CLOSURE_KILL) health = -1;
            else if (P.upper == WC_CLOSURE_REFLECTION) kd = -kd;
        }
        if (kd < P.kLowerLimit) {
            if (P.lower == WC_CLOSURE_CONSTRAINT) kd = P.kLowerLimit;
            else if (P.lower == WC_CLOSURE_KILL) health = -1;
            else if (P.lower == WC_CLOSURE_REFLECTION) kd = 2.0 * P.kLowerLimit - kd;
        }
        to_cell(jd, id_, kd, P.eta, P.xi, q);
        q.k = (float)kd;
        q.health = health;
        q.age = st.age[idx];
        q.healthFresh = true;   // `health` holds the current value (else derive from age)
        q.changed = false;
        q.elapsed = P.elapsed0;
        q.done = false;
        q.tag = Jibe ? st.tag[idx] : 0u;
        steps = run_interval<T, Smp, Jibe, Ev>(P, S, (float)P.sigma, id, ptime, q, P.intervalLength);
        if (!q.healthFresh) q.health = exp(-q.age / P.tau0);
        if (Jibe && tag_type(q.tag) != 0) st.tag[idx] = q.tag;
        if (Ev && q.ev) st.ev[idx] |= (uint8_t)q.ev;
        if (q.changed) {
            st.k[idx] = (double)q.k;
            st.j[idx] = (double)q.jc + (double)q.fj;
            st.i[idx] = (double)q.ic + (double)q.fi;
            st.health[idx] = q.health;
            st.age[idx] = q.age;
        } else if (kd != st.k[idx] || q.health != st.health[idx]) {  // pre-step closures only
            st.k[idx] = kd;
            st.health[idx] = q.health;
        }
    }
    if (stepCounter) wave_add_u64(stepCounter, steps);
}

// Interleaves the interval's records (FP32 ring slots) into the gather layout.
template <class T>
__device__ __forceinline__ void pack_body(const RecordPair<float> &F, const float *__restrict__ hf,
                                          UV4<T> *__restrict__ uv, WA8<T> *__restrict__ wa, ZH *__restrict__ z,
                                          int s_rho, int s_w, size_t plane) {
    const int nlev = s_w - 1;                 // particle levels k = -(nlev-1) .. 0
    const size_t total = plane * nlev;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
        const size_t lv = q / plane, c = q - lv * plane;     // lv = k + nlev - 1
        const int k = (int)lv - (nlev - 1);
        const size_t r = (size_t)(k + s_rho - 1) * plane + c;          // rho level k
        const size_t wt = (size_t)(k + s_w - 1) * plane + c;           // W level k
        const size_t wb = wt - plane;                                    // W level k-1
        UV4<T> a;
        a.v[0] = fromf<T>(F.u0[r]); a.v[1] = fromf<T>(F.u1[r]);
        a.v[2] = fromf<T>(F.v0[r]); a.v[3] = fromf<T>(F.v1[r]);
        uv[q] = a;
        WA8<T> b;
        b.v[0] = fromf<T>(F.w0[wt]); b.v[1] = fromf<T>(F.w1[wt]);
        b.v[2] = fromf<T>(F.w0[wb]); b.v[3] = fromf<T>(F.w1[wb]);
        b.v[4] = fromf<T>(F.akt0[wt]); b.v[5] = fromf<T>(F.akt1[wt]);
        b.v[6] = fromf<T>(F.akt0[wb]); b.v[7] = fromf<T>(F.akt1[wb]);
        wa[q] = b;
        if (lv == 0) z[c] = ZH{tof<T>(fromf<T>(F.zeta0[c])), tof<T>(fromf<T>(F.zeta1[c])), hf[c], 0.0f};
    }
}

template <class T>
__global__ void pack_kernel(RecordPair<float> F, const float *__restrict__ hf, UV4<T> *__restrict__ uv,
                            WA8<T> *__restrict__ wa, ZH *__restrict__ z, int s_rho, int s_w, size_t plane) {
    pack_body<T>(F, hf, uv, wa, z, s_rho, s_w, plane);
}

// Smallest power of two >= x / cap (x >= 0), kept within [2^-100, 2^100]
// (2^-100 for x = 0), so every scaled int16 stays a normal float.
__device__ __forceinline__ double pow2_at_least(double x, double cap) {
    int e = 0;
    const double f = frexp(x / cap, &e);     // x / cap = f * 2^e, f in [0.5, 1)
    if (f == 0.0) return ldexp(1.0, -100);
    if (e < -100) e = -100;
    if (e > 100) e = 100;
    return ldexp(1.0, e);
}

// (v - m) / s to the nearest integer, given inv = 1 / s (exact: s is a power
// of two, and so is the product in FP64); |v - m| <= 32766.5 s by
// construction of the scale, the clamp only guards the int16 range.
__device__ __forceinline__ int16_t quant16(float v, float m, double inv) {
    const double q = rint(((double)v - (double)m) * inv);
    return (int16_t)fmin(fmax(q, -32767.0), 32767.0);
}

// Scale and mean of one group over a tile's nodes (rows [0, nj), columns
// [c0, c0 + ni) of two staged records a, b of row pitch kI16Pitch): the mean
// (FP64, rows in order) of each record's non-zero values, a power-of-two
// scale covering the largest deviation of either record from its mean
// (zeros included) in 32766 steps and keeping |mean / scale| below 2^23, and
// the means rounded to multiples of it (exact in FP32).  amax: AKt, no mean:
// a scale covering the largest magnitude in 32767 steps.
constexpr int kI16Tiles = 64;                            // tiles per workgroup (a strip of one tile row)
constexpr int kI16Pitch = kI16Tiles * kTile + 4;         // staged nodes per row (257 used)

__device__ void tile_group_scale(const float *a, const float *b, int nj, int c0, int ni, bool amax, float &ma,
                                 float &mb, float &sc) {
    if (amax) {
        double m = 0.0;
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i)
                m = fmax(m, fmax(fabs((double)a[j * kI16Pitch + c0 + i]), fabs((double)b[j * kI16Pitch + c0 + i])));
        ma = mb = 0.f;
        sc = (float)pow2_at_least(m, 32767.0);
        return;
    }
    double sa = 0.0, sb = 0.0;
    int na = 0, nb = 0;
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) {
            const float va = a[j * kI16Pitch + c0 + i], vb = b[j * kI16Pitch + c0 + i];
            if (va != 0.0f) { sa += va; ++na; }
            if (vb != 0.0f) { sb += vb; ++nb; }
        }
    const double mda = na ? sa / na : 0.0, mdb = nb ? sb / nb : 0.0;
    double dev = 0.0;
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i)
            dev = fmax(dev, fmax(fabs((double)a[j * kI16Pitch + c0 + i] - mda),
                                 fabs((double)b[j * kI16Pitch + c0 + i] - mdb)));
    const double s = fmax(pow2_at_least(dev, 32766.0), pow2_at_least(fmax(fabs(mda), fabs(mdb)), 8388608.0));
    ma = (float)(rint(mda / s) * s);
    mb = (float)(rint(mdb / s) * s);
    sc = (float)s;
}

// Mean of each of two staged records' non-zero values over a tile's nodes
// (FP64, rows in order; 0 when all are zero), as tile_group_scale takes it.
__device__ void tile_group_mean(const float *a, const float *b, int nj, int c0, int ni, double &ma, double &mb) {
    double sa = 0.0, sb = 0.0;
    int na = 0, nb = 0;
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) {
            const float va = a[j * kI16Pitch + c0 + i], vb = b[j * kI16Pitch + c0 + i];
            if (va != 0.0f) { sa += va; ++na; }
            if (vb != 0.0f) { sb += vb; ++nb; }
        }
    ma = na ? sa / na : 0.0;
    mb = nb ? sb / nb : 0.0;
}

// The whole tiled pack of one interval in one pass, for the block-scaled
// store (T = int16_t) and the tile-mean stores (T = BF16 bits or _Float16).
// A workgroup takes a strip of up to kI16Tiles tiles of one tile row at one
// level: it stages the strip's (kTile+1) node rows of u, v (both records) in
// LDS with coalesced loads (all of a work-item's loads before its stores),
// computes each tile's means (and, block-scaled, scales: tile_group_scale),
// and writes them and the residuals of every node of every tile; then the
// same for w at W levels k and k-1 and AKt.  Tile means: the mean of the
// non-zero values rounded to T (with a mean that T represents, a zero
// node's residual -mean is exact, so zeros decode to 0), each residual
// v - mean rounded to T, AKt rounded to T.  Nodes beyond the grid hold zeros
// (no valid stencil reads them).
// Level 0 also writes zeta (FP32 block-scaled, rounded to T otherwise) and h
// for the strip's cells.
template <class T>
__global__ __launch_bounds__(256) void tile_pack_strip_kernel(RecordPair<float> F, const float *__restrict__ hf,
                                                              TileScaleUV *__restrict__ tuv,
                                                              TileScaleWA *__restrict__ twa, WA8<T> *__restrict__ tm,
                                                              UV4<T> *__restrict__ uvr, WA8<T> *__restrict__ war,
                                                              ZH *__restrict__ z, int s_rho, int s_w, int eta, int xi,
                                                              int tilesX, size_t tplane) {
    constexpr bool kScaled = std::is_same<T, int16_t>::value;
    extern __shared__ float st[];                         // [8][kTile + 1][kI16Pitch]
    __shared__ float scl[kI16Tiles][8];                   // per tile: 4 means, 2 or 4 scales
    __shared__ float sinv[kI16Tiles][4];                  // and the scales' reciprocals (exact)
    __shared__ float smean[kI16Tiles][4];                 // tile means: the u, v means (phase 0)
    const int nlev = s_w - 1;
    const int strips = (tilesX + kI16Tiles - 1) / kI16Tiles;
    const int tilesY = (int)(tplane / (size_t)tilesX);
    const int b = blockIdx.x, lv = b / (tilesY * strips), rem = b - lv * tilesY * strips;
    const int tj = rem / strips, ti0 = (rem - tj * strips) * kI16Tiles;
    const int nt = min(kI16Tiles, tilesX - ti0);
    const int k = lv - (nlev - 1);
    const size_t plane = (size_t)eta * xi;
    const size_t r0 = (size_t)(k + s_rho - 1) * plane, wt = (size_t)(k + s_w - 1) * plane, wb = wt - plane;
    const int j0 = tj * kTile, i0 = ti0 * kTile;
    const int nj = min(kTile + 1, eta - j0), ni = min(nt * kTile + 1, xi - i0);    // staged node rows, columns
    const size_t lt0 = (size_t)lv * tplane + (size_t)tj * tilesX + ti0;          // first tile of the strip
    const int tid = threadIdx.x;
    constexpr int kRows = kTile + 1, kArr = kRows * kI16Pitch;
    for (int phase = 0; phase < 2; ++phase) {
        const int na = phase == 0 ? 4 : 8;
        const float *src[8];
        if (phase == 0) {
            src[0] = F.u0 + r0; src[1] = F.u1 + r0; src[2] = F.v0 + r0; src[3] = F.v1 + r0;
            src[4] = src[5] = src[6] = src[7] = nullptr;
        } else {
            src[0] = F.w0 + wt; src[1] = F.w1 + wt; src[2] = F.w0 + wb; src[3] = F.w1 + wb;
            src[4] = F.akt0 + wt; src[5] = F.akt1 + wt; src[6] = F.akt0 + wb; src[7] = F.akt1 + wb;
        }
        __syncthreads();                                  // the previous phase's reads are done
        // column tid of every staged row: all loads issued before any store
        // (one memory latency per phase, not one per row)
        float vals[8][kRows];
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int j = 0; j < kRows; ++j)
                vals[a][j] = (a < na && j < nj && tid < ni) ? src[a][(size_t)(j0 + j) * xi + i0 + tid] : 0.f;
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int j = 0; j < kRows; ++j)
                if (a < na && tid < ni) st[a * kArr + j * kI16Pitch + tid] = vals[a][j];
        for (int e = tid; ni > 256 && e < na * kRows; e += 256) {      // the strip's last node column
            const int a = e / kRows, j = e - a * kRows;
            st[a * kArr + j * kI16Pitch + 256] = j < nj ? src[a][(size_t)(j0 + j) * xi + i0 + 256] : 0.f;
        }
        __syncthreads();
        // means (and scales): one work-item per (tile, group)
        const int groups = phase == 0 ? 2 : (kScaled ? 4 : 2);
        if (tid < nt * groups) {
            const int t = tid / groups, g = tid - t * groups;
            const int c0 = t * kTile, nic = min(kTile + 1, ni - c0);
            if constexpr (kScaled) {
                float ma, mb, sc;
                tile_group_scale(st + 2 * g * kArr, st + (2 * g + 1) * kArr, nj, c0, nic, phase == 1 && g >= 2, ma,
                                 mb, sc);
                if (phase == 0 || g < 2) {
                    scl[t][2 * g] = ma;
                    scl[t][2 * g + 1] = mb;
                }
                scl[t][4 + g] = sc;
                sinv[t][g] = 1.0f / sc;
            } else {
                double ma, mb;
                tile_group_mean(st + 2 * g * kArr, st + (2 * g + 1) * kArr, nj, c0, nic, ma, mb);
                scl[t][2 * g] = tof<T>(fromf<T>((float)ma));
                scl[t][2 * g + 1] = tof<T>(fromf<T>((float)mb));
                if (phase == 0) {
                    smean[t][2 * g] = scl[t][2 * g];
                    smean[t][2 * g + 1] = scl[t][2 * g + 1];
                }
            }
        }
        __syncthreads();
        if (tid < nt) {
            if constexpr (kScaled) {
                if (phase == 0) {
                    TileScaleUV o;
                    for (int f = 0; f < 4; ++f) o.m[f] = scl[tid][f];
                    o.s[0] = scl[tid][4];
                    o.s[1] = scl[tid][5];
                    o.pad[0] = o.pad[1] = 0.f;
                    tuv[lt0 + tid] = o;
                } else {
                    TileScaleWA o;
                    for (int f = 0; f < 4; ++f) o.m[f] = scl[tid][f];
                    for (int f = 0; f < 4; ++f) o.s[f] = scl[tid][4 + f];
                    twa[lt0 + tid] = o;
                }
            } else if (phase == 1) {
                WA8<T> o;
                for (int f = 0; f < 4; ++f) o.v[f] = fromf<T>(smean[tid][f]);
                for (int f = 0; f < 4; ++f) o.v[4 + f] = fromf<T>(scl[tid][f]);
                tm[lt0 + tid] = o;
            }
        }
        // residuals: one work-item per (tile, node), nodes of a tile consecutive
        for (int e = tid; e < nt * kTileNodes; e += 256) {
            const int t = e / kTileNodes, nd = e - t * kTileNodes;
            const int dj = nd / (kTile + 1), di = nd - dj * (kTile + 1);
            const int c = t * kTile + di;                          // staged column
            const bool in = dj < nj && c < ni;
            const int o = dj * kI16Pitch + c;
            auto res = [&](int a, int m, int sg) -> T {       // array a, mean m, scale group sg
                if constexpr (kScaled) return quant16(st[a * kArr + o], scl[t][m], sinv[t][sg]);
                else return fromf<T>(st[a * kArr + o] - scl[t][m]);
            };
            auto raw = [&](int a, int sg) -> T {               // AKt: no mean
                if constexpr (kScaled) return quant16(st[a * kArr + o], 0.f, sinv[t][sg]);
                else return fromf<T>(st[a * kArr + o]);
            };
            if (phase == 0) {
                UV4<T> q{};
                if (in) {
                    q.v[0] = res(0, 0, 0);
                    q.v[1] = res(1, 1, 0);
                    q.v[2] = res(2, 2, 1);
                    q.v[3] = res(3, 3, 1);
                }
                uvr[(lt0 + t) * kTileNodes + nd] = q;
            } else {
                WA8<T> q{};
                if (in) {
                    q.v[0] = res(0, 0, 0);
                    q.v[1] = res(1, 1, 0);
                    q.v[2] = res(2, 2, 1);
                    q.v[3] = res(3, 3, 1);
                    q.v[4] = raw(4, 2);
                    q.v[5] = raw(5, 2);
                    q.v[6] = raw(6, 3);
                    q.v[7] = raw(7, 3);
                }
                war[(lt0 + t) * kTileNodes + nd] = q;
            }
        }
        __syncthreads();                                  // scl is rewritten by the next phase
    }
    if (lv == 0)                                          // zeta and h of the strip's cells
        for (int e = tid; e < kTile * nt * kTile; e += 256) {
            const int dj = e / (nt * kTile), di = e - dj * nt * kTile;
            const int j = j0 + dj, i = i0 + di;
            if (j < eta && i < xi) {
                const size_t c = (size_t)j * xi + i;
                if constexpr (kScaled) z[c] = ZH{F.zeta0[c], F.zeta1[c], hf[c], 0.0f};
                else z[c] = ZH{tof<T>(fromf<T>(F.zeta0[c])), tof<T>(fromf<T>(F.zeta1[c])), hf[c], 0.0f};
            }
        }
}

// ---- batched ensemble ------------------------------------------------------
//
// Many release scenarios share the grid, the forcing and one population.
// Each particle carries its scenario (one byte), which selects its horizontal
// diffusion scale; release offsets are particle emission times (the
// reference's late activation, Particle.cpp:211-215).  State is compact: cell
// index + FP32 offsets and level, FP32 health, FP64 age and emission time.
// The interval's scalars and record pointers are read from a device array
// indexed by a device step counter, so one captured graph (pack -> move ->
// clear -> surface grid -> advance) replays interval after interval.

template <class T>
__global__ void ens_pack_kernel(const EnsStep *__restrict__ steps, const int *__restrict__ stepIdx,
                                const float *__restrict__ hf, UV4<T> *__restrict__ uv, WA8<T> *__restrict__ wa,
                                ZH *__restrict__ z, int s_rho, int s_w, size_t plane) {
    const RecordPair<float> F = steps[*stepIdx].F;
    pack_body<T>(F, hf, uv, wa, z, s_rho, s_w, plane);
}

// cell + offset -> FP64 coordinate -> cell + FP32 offset, as the FP64 state
// of the context path stores and reloads it (to_cell); NaN stays NaN.
__device__ __forceinline__ void requantize(int &cell, float &off) {
    const double v = (double)cell + (double)off;
    if (!(v == v)) return;
    const double fl = floor(v);
    cell = (int)fl;
    off = (float)(v - fl);
    if (off >= 1.0f) { off -= 1.0f; ++cell; }
}

// Occ = 0: no occupancy hint (the compiler's register allocation: 175-191
// VGPRs, 2 waves per SIMD); Occ = 3: at most 168 VGPRs (3 waves per SIMD);
// Occ = 4: at most 128 VGPRs (4 waves per SIMD, with spills).  The
// autotuner times them.
template <class T, int Occ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Occ > 0 ? Occ : 1)))
void move_ens_kernel(const EnsStep *__restrict__ steps, const int *__restrict__ stepIdx, Packed<T> S, EnsSoA st,
                     int n, const float *__restrict__ sigmaTab, int ppt, unsigned long long *stepCounter) {
    const MoveParams P = steps[*stepIdx].P;
    unsigned total = 0;
    const int stride = gridDim.x * blockDim.x;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    for (int r = 0; r < ppt && idx < n; ++r, idx += stride) {
        const uint32_t cell = st.cell[idx];
        Reg q;
        q.jc = (int)(cell / (uint32_t)P.xi);
        q.ic = (int)(cell - (uint32_t)q.jc * (uint32_t)P.xi);
        q.fj = st.fj[idx];
        q.fi = st.fi[idx];
        // an offset that rounded to 1.0f (the domain reflection can produce
        // one) belongs to the next cell, as the FP64 reload of the context
        // path (to_cell) makes it
        if (q.fj >= 1.0f) { q.fj -= 1.0f; ++q.jc; }
        if (q.fi >= 1.0f) { q.fi -= 1.0f; ++q.ic; }
        float k0 = st.k[idx];
        const float h0 = st.health[idx];
        double health = h0;
        const float kLL = (float)P.kLowerLimit;
        float k = k0;
        if (k > 0.f) {                                             // Particle.cpp:134-176
            if (P.upper == WC_CLOSURE_CONSTRAINT) k = 0.f;
            else if (P.upper == WC_CLOSURE_KILL) health = -1;
            else if (P.upper == WC_CLOSURE_REFLECTION) k = -k;
        }
        if (k < kLL) {
            if (P.lower == WC_CLOSURE_CONSTRAINT) k = kLL;
            else if (P.lower == WC_CLOSURE_KILL) health = -1;
            else if (P.lower == WC_CLOSURE_REFLECTION) k = 2.0f * kLL - k;
        }
        q.k = k;
        q.inside = valid_cell(q.jc, q.ic, P.eta, P.xi) && q.fj == q.fj && q.fi == q.fi && k == k;
        q.health = health;
        q.age = st.age[idx];
        q.healthFresh = true;
        q.changed = false;
        q.elapsed = P.elapsed0;
        q.done = false;
        total += run_interval<T>(P, S, sigmaTab[st.scen[idx]], st.id[idx], st.time[idx], q, P.intervalLength);
        if (!q.healthFresh) q.health = exp(-q.age / P.tau0);
        if (q.changed) {
            // store the position the context path stores (FP64 cell +
            // offset, read back into cell + FP32 offset): an offset below the
            // FP64 resolution of the coordinate becomes 0 there, and both
            // paths must continue from the same state
            requantize(q.jc, q.fj);
            requantize(q.ic, q.fi);
            st.cell[idx] = (uint32_t)q.jc * (uint32_t)P.xi + (uint32_t)q.ic;
            st.fj[idx] = q.fj;
            st.fi[idx] = q.fi;
            st.k[idx] = q.k;
            st.health[idx] = (float)q.health;
            st.age[idx] = q.age;
        } else if (k != k0 || (float)q.health != h0) {
            st.k[idx] = k;
            st.health[idx] = (float)q.health;
        }
    }
    if (stepCounter) wave_add_u64(stepCounter, total);
}

// The LDS box of one chunk of substeps, shared by the LDS-tiled moves.  Each
// work-item offers its cell and level when `act` (it steps in the chunk);
// the workgroup reduces the box of those cells widened by `halo` cells, one
// level each way and the stencil's next cell, shrinks the halo until the
// box fits in `ldsBytes`, and copies the box's packed u/v, w/AKt, zeta/h and
// metric into LDS with coalesced loads (no box when nothing steps: a 1 x 1
// box at (-2, -2) with no levels, which no cell or stencil lies in).  The
// first barrier also ends every read of the previous chunk's box.  The
// returned sampler reads LDS inside the box and global memory outside, so
// the values are those of S wherever a particle goes.
template <class T>
__device__ TilePacked<T> stage_box(const Packed<T> &S, bool act, int jc, int ic, int kI, int eta, int xi, int halo,
                                   int ldsBytes, unsigned char *lds, int (*part)[6], int *box) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nlev = S.nlev;
    int v[6] = {act ? jc : INT_MAX, act ? -jc : INT_MAX, act ? ic : INT_MAX,
                act ? -ic : INT_MAX, act ? kI : INT_MAX, act ? -kI : INT_MAX};
    for (int off = 32; off > 0; off >>= 1)
        for (int r = 0; r < 6; ++r) v[r] = min(v[r], __shfl_xor(v[r], off, 64));
    if (lane == 0)
        for (int r = 0; r < 6; ++r) part[wave][r] = v[r];
    __syncthreads();
    if (tid == 0) {
        int m[6];
        for (int r = 0; r < 6; ++r) m[r] = min(min(part[0][r], part[1][r]), min(part[2][r], part[3][r]));
        int b[6] = {-2, 1, -2, 1, 0, 0};                  // j0, bh, i0, bw, l0, nl
        const int lLo = max(0, m[4] - 2 + nlev - 1), lHi = min(nlev - 1, -m[5] + 1 + nlev - 1);
        const int nl = lHi - lLo + 1;
        if (m[0] != INT_MAX && nl >= 1) {
            for (int h = halo; h >= 0; --h) {
                const int j0 = max(0, m[0] - h), j1 = min(eta - 1, -m[1] + 1 + h);
                const int i0 = max(0, m[2] - h), i1 = min(xi - 1, -m[3] + 1 + h);
                const int bh = j1 - j0 + 1, bw = i1 - i0 + 1, cells = bh * bw;
                const long long need = (long long)nl * cells * (long long)(sizeof(WA8<T>) + sizeof(UV4<T>)) +
                                       (long long)cells * (long long)(sizeof(ZH) + sizeof(float2));
                if (need <= ldsBytes) {
                    b[0] = j0; b[1] = bh; b[2] = i0; b[3] = bw; b[4] = lLo; b[5] = nl;
                    break;
                }
            }
        }
        for (int r = 0; r < 6; ++r) box[r] = b[r];
    }
    __syncthreads();
    TilePacked<T> TS;
    TS.g = S;
    TS.j0 = box[0]; TS.bh = box[1]; TS.i0 = box[2]; TS.bw = box[3]; TS.l0 = box[4]; TS.nl = box[5];
    const int cells = TS.bh * TS.bw;
    WA8<T> *wL = reinterpret_cast<WA8<T> *>(lds);
    ZH *zL = reinterpret_cast<ZH *>(lds + (size_t)TS.nl * cells * sizeof(WA8<T>));
    UV4<T> *uL = reinterpret_cast<UV4<T> *>(reinterpret_cast<unsigned char *>(zL) + (size_t)cells * sizeof(ZH));
    float2 *mL = reinterpret_cast<float2 *>(reinterpret_cast<unsigned char *>(uL) +
                                            (size_t)TS.nl * cells * sizeof(UV4<T>));
    TS.wL = wL; TS.zL = zL; TS.uL = uL; TS.mL = mL;
    if (TS.nl > 0) {
        for (int c = tid; c < cells; c += 256) {
            const int jj = c / TS.bw, ii = c - jj * TS.bw;
            const size_t g = (size_t)(TS.j0 + jj) * S.xi + (TS.i0 + ii);
            zL[c] = S.z[g];
            mL[c] = S.md[g];
        }
        const int items = TS.nl * cells;
        for (int c = tid; c < items; c += 256) {
            const int l = c / cells, r = c - l * cells;
            const int jj = r / TS.bw, ii = r - jj * TS.bw;
            const size_t g = (size_t)(TS.l0 + l) * S.plane + (size_t)(TS.j0 + jj) * S.xi + (TS.i0 + ii);
            wL[c] = S.wa[g];
            uL[c] = S.uv[g];
        }
    }
    __syncthreads();
    return TS;
}

// The ensemble move with the forcing staged in LDS.  One particle per work-
// item, 256 per workgroup.  The interval runs in chunks of `chunk` substeps;
// before each chunk the workgroup reduces the cells and levels of its
// particles that will step in it, widens that box by `halo` cells (and one
// level each way plus the stencil's next cell), and, when the box fits in
// `ldsBytes`, copies its packed u/v, w/AKt, zeta/h and metric into LDS with
// coalesced loads.  The substeps then sample through TilePacked (LDS inside
// the box, global memory outside), so the result is the one of
// move_ens_kernel for any box.  With the population sorted by cell tile the
// box stays small.
template <class T, int Occ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Occ > 0 ? Occ : 1)))
void move_ens_lds_kernel(const EnsStep *__restrict__ steps, const int *__restrict__ stepIdx, Packed<T> S, EnsSoA st,
                         int n, const float *__restrict__ sigmaTab, int chunk, int halo, int ldsBytes,
                         unsigned long long *stepCounter) {
    extern __shared__ __align__(32) unsigned char lds[];
    __shared__ int part[4][6];
    __shared__ int box[6];
    const MoveParams P = steps[*stepIdx].P;
    const int idx = blockIdx.x * 256 + threadIdx.x;
    const bool have = idx < n;
    Reg q{};
    float k0 = 0.f, h0 = 0.f, kStart = 0.f, sigma = 0.f;
    uint64_t id = 0;
    double ptime = 0;
    q.done = true;
    if (have) {
        const uint32_t cell = st.cell[idx];
        q.jc = (int)(cell / (uint32_t)P.xi);
        q.ic = (int)(cell - (uint32_t)q.jc * (uint32_t)P.xi);
        q.fj = st.fj[idx];
        q.fi = st.fi[idx];
        if (q.fj >= 1.0f) { q.fj -= 1.0f; ++q.jc; }
        if (q.fi >= 1.0f) { q.fi -= 1.0f; ++q.ic; }
        k0 = st.k[idx];
        h0 = st.health[idx];
        double health = h0;
        const float kLL = (float)P.kLowerLimit;
        float k = k0;
        if (k > 0.f) {                                             // Particle.cpp:134-176
            if (P.upper == WC_CLOSURE_CONSTRAINT) k = 0.f;
            else if (P.upper == WC_CLOSURE_KILL) health = -1;
            else if (P.upper == WC_CLOSURE_REFLECTION) k = -k;
        }
        if (k < kLL) {
            if (P.lower == WC_CLOSURE_CONSTRAINT) k = kLL;
            else if (P.lower == WC_CLOSURE_KILL) health = -1;
            else if (P.lower == WC_CLOSURE_REFLECTION) k = 2.0f * kLL - k;
        }
        kStart = k;
        q.k = k;
        q.inside = valid_cell(q.jc, q.ic, P.eta, P.xi) && q.fj == q.fj && q.fi == q.fi && k == k;
        q.health = health;
        q.age = st.age[idx];
        q.healthFresh = true;
        q.changed = false;
        q.elapsed = P.elapsed0;
        q.done = false;
        sigma = sigmaTab[st.scen[idx]];
        id = st.id[idx];
        ptime = st.time[idx];
    }
    const double L = P.intervalLength, span = (double)chunk * P.dti;
    const int nChunks = L > P.elapsed0 ? (int)ceil((L - P.elapsed0) / span) : 0;
    unsigned total = 0;
    for (int ch = 0; ch < nChunks; ++ch) {
        const double tEnd = ch + 1 == nChunks ? L : fmin(L, P.elapsed0 + (double)(ch + 1) * span);
        // box of the particles that step in this chunk (forward: emitted before its end)
        const bool act = !q.done && q.inside && q.elapsed < tEnd && q.k >= -1e6f && q.k <= 1e6f &&
                         (P.dirInt < 0 || ptime - P.intervalStart < tEnd);
        const TilePacked<T> TS = stage_box<T>(S, act, q.jc, q.ic, act ? (int)ceilf(q.k) : 0, P.eta, P.xi, halo,
                                              ldsBytes, lds, part, box);
        if (have) total += run_interval<T>(P, TS, sigma, id, ptime, q, tEnd);
    }
    if (have) {
        if (!q.healthFresh) q.health = exp(-q.age / P.tau0);
        if (q.changed) {
            requantize(q.jc, q.fj);
            requantize(q.ic, q.fi);
            st.cell[idx] = (uint32_t)q.jc * (uint32_t)P.xi + (uint32_t)q.ic;
            st.fj[idx] = q.fj;
            st.fi[idx] = q.fi;
            st.k[idx] = q.k;
            st.health[idx] = (float)q.health;
            st.age[idx] = q.age;
        } else if (kStart != k0 || (float)q.health != h0) {
            st.k[idx] = kStart;
            st.health[idx] = (float)q.health;
        }
    }
    if (stepCounter) wave_add_u64(stepCounter, total);
}

// The context's mixed-precision move (move_fast_kernel, plain packed
// storage, no jibe, no event tracking) with the forcing staged in LDS as in
// move_ens_lds_kernel: the interval runs in chunks of `chunk` substeps, and
// before each chunk the workgroup stages the box of its particles that step
// in it (stage_box).  Work-item t moves particle perm[t] (a cell-tile
// order), so with the population sorted the box stays small.  Every value a
// stencil reads is the one the plain kernel reads, and run_interval resumes
// a particle exactly at a chunk boundary, so the results are identical.
// Three waves per SIMD: three workgroups of kMoveLdsBytes share a CU.
template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
void move_fast_lds_kernel(MoveParams P, Packed<T> S, SoA st, int n, int chunk,
                                                            int halo, int ldsBytes, unsigned long long *stepCounter,
                                                            const uint32_t *__restrict__ perm) {
    extern __shared__ __align__(32) unsigned char lds[];
    __shared__ int part[4][6];
    __shared__ int box[6];
    const int t = blockIdx.x * 256 + threadIdx.x;
    const bool have = t < n;
    const int idx = have ? (perm ? (int)perm[t] : t) : 0;
    Reg q{};
    q.done = true;
    uint64_t id = 0;
    double ptime = 0, kd = 0;
    if (have) {
        id = st.id[idx];
        kd = st.k[idx];
        double health = st.health[idx];
        ptime = st.time[idx];
        if (kd > 0) {                                              // Particle.cpp:134-176
            if (P.upper == WC_CLOSURE_CONSTRAINT) kd = 0;
            else if (P.upper == WC_CLOSURE_KILL) health = -1;
            else if (P.upper == WC_CLOSURE_REFLECTION) kd = -kd;
        }
        if (kd < P.kLowerLimit) {
            if (P.lower == WC_CLOSURE_CONSTRAINT) kd = P.kLowerLimit;
            else if (P.lower == WC_CLOSURE_KILL) health = -1;
            else if (P.lower == WC_CLOSURE_REFLECTION) kd = 2.0 * P.kLowerLimit - kd;
        }
        to_cell(st.j[idx], st.i[idx], kd, P.eta, P.xi, q);
        q.k = (float)kd;
        q.health = health;
        q.age = st.age[idx];
        q.healthFresh = true;
        q.changed = false;
        q.elapsed = P.elapsed0;
        q.done = false;
    }
    const double L = P.intervalLength, span = (double)chunk * P.dti;
    const int nChunks = L > P.elapsed0 ? (int)ceil((L - P.elapsed0) / span) : 0;
    unsigned total = 0;
    for (int ch = 0; ch < nChunks; ++ch) {
        const double tEnd = ch + 1 == nChunks ? L : fmin(L, P.elapsed0 + (double)(ch + 1) * span);
        const bool act = !q.done && q.inside && q.elapsed < tEnd && q.k >= -1e6f && q.k <= 1e6f &&
                         (P.dirInt < 0 || ptime - P.intervalStart < tEnd);
        const TilePacked<T> TS = stage_box<T>(S, act, q.jc, q.ic, act ? (int)ceilf(q.k) : 0, P.eta, P.xi, halo,
                                              ldsBytes, lds, part, box);
        if (have) total += run_interval<T>(P, TS, (float)P.sigma, id, ptime, q, tEnd);
    }
    if (have) {
        if (!q.healthFresh) q.health = exp(-q.age / P.tau0);
        if (q.changed) {
            st.k[idx] = (double)q.k;
            st.j[idx] = (double)q.jc + (double)q.fj;
            st.i[idx] = (double)q.ic + (double)q.fi;
            st.health[idx] = q.health;
            st.age[idx] = q.age;
        } else if (kd != st.k[idx] || q.health != st.health[idx]) {  // pre-step closures only
            st.k[idx] = kd;
            st.health[idx] = q.health;
        }
    }
    if (stepCounter) wave_add_u64(stepCounter, total);
}

// Per-scenario surface-layer counts [scenario][eta][xi] at the end of the
// step's interval: particles released by then (emission time not after the
// interval end in forward tracking: the reference creates a particle in the
// interval of its emission), alive, with rounded level 0 (k > -0.5,
// Concentration.hpp:31-36 rounds half away from zero), at their rounded
// (j, i); wave-ballot grouping of equal cells, one atomic per group.
__global__ __launch_bounds__(256) void ens_surface_kernel(const EnsStep *__restrict__ steps,
                                                          const int *__restrict__ stepIdx, EnsSoA st, int n,
                                                          int eta, int xi, uint32_t *__restrict__ counts) {
    const MoveParams &P = steps[*stepIdx].P;
    const double tEnd = P.intervalStart + P.dirInt * P.intervalLength;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = false;
    uint32_t key = 0;
    if (idx < n && st.health[idx] > 0.f && st.k[idx] > -0.5f && (P.dirInt < 0 || st.time[idx] <= tEnd)) {
        const uint32_t cell = st.cell[idx];
        const int jc = (int)(cell / (uint32_t)xi), ic = (int)(cell - (uint32_t)jc * (uint32_t)xi);
        const int jr = jc + (st.fj[idx] >= 0.5f ? 1 : 0), ir = ic + (st.fi[idx] >= 0.5f ? 1 : 0);
        if (jr < eta && ir < xi) {
            key = ((uint32_t)st.scen[idx] * (uint32_t)eta + (uint32_t)jr) * (uint32_t)xi + (uint32_t)ir;
            active = true;
        }
    }
    for (;;) {
        const unsigned long long m = __ballot(active);
        if (m == 0) break;
        const int leader = __ffsll((long long)m) - 1;
        const uint32_t lk = __shfl(key, leader, 64);
        const bool same = active && key == lk;
        const unsigned long long sm = __ballot(same);
        if ((int)(threadIdx.x & 63) == leader) atomicAdd(&counts[lk], (uint32_t)__popcll(sm));
        if (same) active = false;
    }
}

__global__ void ens_advance_kernel(int *stepIdx) { stepIdx[0] += 1; }

// ---- cell-tile sort of the ensemble population -----------------------------
//
// Every `period`-th step (read from the device step counter, so the captured
// step stays the same) the population is reordered by the tile of its cell
// (tile x tile cells, row-major tiles; a particle outside the grid goes
// last): a histogram, an exclusive scan and a scatter into scratch, then a
// copy back.  Lanes of a wave then gather from neighbouring cells.  Every
// result is per particle (draws keyed by id, integer counts), so the order
// changes no result; the order inside a tile is not fixed.

__device__ __forceinline__ bool sort_step(const int *stepIdx, int period) {
    return period > 0 && (*stepIdx % period) == 0;
}

__device__ __forceinline__ uint32_t tile_key(uint32_t cell, int xi, int tile, int tilesX, uint32_t nTiles) {
    if (cell == 0xFFFFFFFFu) return nTiles;
    const uint32_t j = cell / (uint32_t)xi, i = cell - j * (uint32_t)xi;
    return (j / (uint32_t)tile) * (uint32_t)tilesX + i / (uint32_t)tile;
}

__global__ void ens_sort_clear_kernel(const int *stepIdx, int period, uint32_t *bins, uint32_t nBins) {
    if (!sort_step(stepIdx, period)) return;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nBins; q += gridDim.x * blockDim.x) bins[q] = 0;
}

// histogram with equal keys of a wave added by one atomic
__global__ __launch_bounds__(256) void ens_sort_hist_kernel(const int *stepIdx, int period, const uint32_t *cell,
                                                            int n, int xi, int tile, int tilesX, uint32_t nTiles,
                                                            uint32_t *bins) {
    if (!sort_step(stepIdx, period)) return;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = idx < n;
    const uint32_t key = active ? tile_key(cell[idx], xi, tile, tilesX, nTiles) : 0;
    for (;;) {
        const unsigned long long m = __ballot(active);
        if (m == 0) break;
        const int leader = __ffsll((long long)m) - 1;
        const uint32_t lk = __shfl(key, leader, 64);
        const bool same = active && key == lk;
        const unsigned long long sm = __ballot(same);
        if ((int)(threadIdx.x & 63) == leader) atomicAdd(&bins[lk], (uint32_t)__popcll(sm));
        if (same) active = false;
    }
}

// exclusive scan of nBins counts in place, one block of 1024
__global__ __launch_bounds__(1024) void ens_sort_scan_kernel(const int *stepIdx, int period, uint32_t *bins,
                                                             uint32_t nBins) {
    if (!sort_step(stepIdx, period)) return;
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x, per = (nBins + 1023) / 1024, b0 = t * per;
    uint32_t s = 0;
    for (uint32_t q = 0; q < per && b0 + q < nBins; ++q) s += bins[b0 + q];
    part[t] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = t == 0 ? 0 : part[t - 1];
    for (uint32_t q = 0; q < per && b0 + q < nBins; ++q) {
        const uint32_t c = bins[b0 + q];
        bins[b0 + q] = run;
        run += c;
    }
}

// scatter into dst at bins[key] (advanced per group of equal keys in a wave)
__global__ __launch_bounds__(256) void ens_sort_scatter_kernel(const int *stepIdx, int period, EnsSoA src,
                                                               EnsSoA dst, int n, int xi, int tile, int tilesX,
                                                               uint32_t nTiles, uint32_t *bins) {
    if (!sort_step(stepIdx, period)) return;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    bool active = idx < n;
    const uint32_t key = active ? tile_key(src.cell[idx], xi, tile, tilesX, nTiles) : 0;
    uint32_t pos = 0;
    for (;;) {
        const unsigned long long m = __ballot(active);
        if (m == 0) break;
        const int leader = __ffsll((long long)m) - 1;
        const uint32_t lk = __shfl(key, leader, 64);
        const bool same = active && key == lk;
        const unsigned long long sm = __ballot(same);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&bins[lk], (uint32_t)__popcll(sm));
        base = __shfl(base, leader, 64);
        if (same) {
            pos = base + (uint32_t)__popcll(sm & ((1ull << lane) - 1ull));
            active = false;
        }
    }
    if (idx < n) {
        dst.id[pos] = src.id[idx]; dst.cell[pos] = src.cell[idx]; dst.fj[pos] = src.fj[idx];
        dst.fi[pos] = src.fi[idx]; dst.k[pos] = src.k[idx]; dst.health[pos] = src.health[idx];
        dst.age[pos] = src.age[idx]; dst.time[pos] = src.time[idx]; dst.scen[pos] = src.scen[idx];
    }
}

__global__ void ens_sort_copy_kernel(const int *stepIdx, int period, EnsSoA src, EnsSoA dst, int n) {
    if (!sort_step(stepIdx, period)) return;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        dst.id[q] = src.id[q]; dst.cell[q] = src.cell[q]; dst.fj[q] = src.fj[q]; dst.fi[q] = src.fi[q];
        dst.k[q] = src.k[q]; dst.health[q] = src.health[q]; dst.age[q] = src.age[q]; dst.time[q] = src.time[q];
        dst.scen[q] = src.scen[q];
    }
}

// Device release of an ensemble population: particle q (global id q) belongs
// to the source s with first[s] <= q < first[s+1]; its j and i are offset by
// spread * N(0, 1) draws of Philox4x32-10 keyed on (seed; q, 0xE5, 0, 0), so a
// population is reproducible and independent of the launch; every
// sub_every-th particle of a source (by its index in the source) goes to the
// source's sub-scenario.
__global__ void ens_release_kernel(const EnsSource *__restrict__ src, const unsigned long long *__restrict__ first,
                                   int nsrc, long long n, int eta, int xi, uint64_t seed, EnsSoA st) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
         q += (long long)gridDim.x * blockDim.x) {
        int lo = 0, hi = nsrc - 1;                         // last s with first[s] <= q
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (first[mid] <= (unsigned long long)q) lo = mid; else hi = mid - 1;
        }
        const EnsSource &S = src[lo];
        const unsigned long long r = (unsigned long long)q - first[lo];
        const Philox4 z = philox4x32_10((uint32_t)q, (uint32_t)(q >> 32), 0xE5u, 0u, (uint32_t)seed,
                                        (uint32_t)(seed >> 32));
        const double u1 = ((double)(z.v[0] >> 8) + 0.5) * 0x1.0p-24, u2 = (double)(z.v[1] >> 8) * 0x1.0p-24;
        const double rad = sqrt(-2.0 * log(u1));
        const double j = S.j + S.spread * rad * cos(6.283185307179586 * u2);
        const double i = S.i + S.spread * rad * sin(6.283185307179586 * u2);
        const double jfl = floor(j), ifl = floor(i);
        const bool inside = jfl >= 0 && ifl >= 0 && jfl + 1 < eta && ifl + 1 < xi;
        int jc = (int)jfl, ic = (int)ifl;
        float fj = inside ? (float)(j - jfl) : 0.f, fi = inside ? (float)(i - ifl) : 0.f;
        if (fj >= 1.0f) { fj -= 1.0f; ++jc; }
        if (fi >= 1.0f) { fi -= 1.0f; ++ic; }
        st.id[q] = (uint64_t)q;
        st.cell[q] = inside ? (uint32_t)jc * (uint32_t)xi + (uint32_t)ic : 0xFFFFFFFFu;
        st.fj[q] = fj;
        st.fi[q] = fi;
        st.k[q] = (float)S.k;
        st.health[q] = 1.0f;
        st.age[q] = 0.0;
        st.time[q] = S.time;
        st.scen[q] = (uint8_t)((S.subEvery && S.subScenario >= 0 && r % S.subEvery == 0) ? S.subScenario : S.scenario);
    }
}

int grid_blocks(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return (int)(b ? b : 1);
}

template <class T, class Smp>
hipError_t launch_move(const MoveParams &P, const Smp &S, const SoA &s, int n, unsigned long long *ctr, bool events,
                       const uint32_t *perm, hipStream_t stream) {
    const int block = 256, blocks = (n + block - 1) / block;
    if (P.jibe) {
        if (events) move_fast_kernel<T, Smp, true, true><<<blocks, block, 0, stream>>>(P, S, s, n, ctr, perm);
        else move_fast_kernel<T, Smp, true, false><<<blocks, block, 0, stream>>>(P, S, s, n, ctr, perm);
    } else {
        if (events) move_fast_kernel<T, Smp, false, true><<<blocks, block, 0, stream>>>(P, S, s, n, ctr, perm);
        else move_fast_kernel<T, Smp, false, false><<<blocks, block, 0, stream>>>(P, S, s, n, ctr, perm);
    }
    return hipGetLastError();
}

// Byte offsets of the tile-mean layout in the pack buffer (each section
// 256-byte aligned).
struct TileLayout {
    size_t uvr, war, tm, z, total;
    int tilesX;
    size_t tplane;
};
TileLayout tile_layout(size_t tbytes, int s_w, int eta, int xi) {
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    TileLayout L;
    const size_t nlev = (size_t)(s_w - 1);
    L.tilesX = (xi + kTile - 1) / kTile;
    L.tplane = (size_t)L.tilesX * (size_t)((eta + kTile - 1) / kTile);
    const size_t blocks = nlev * L.tplane;
    L.uvr = 0;
    L.war = up(L.uvr + blocks * kTileNodes * 4 * tbytes);
    L.tm = up(L.war + blocks * kTileNodes * 8 * tbytes);
    L.z = up(L.tm + blocks * 8 * tbytes);
    L.total = up(L.z + (size_t)eta * xi * sizeof(ZH));
    return L;
}

// Byte offsets of the block-scaled layout: residuals as the tile-mean layout
// with T = int16, then the two scale arrays in place of the means.
struct TileLayoutI16 {
    size_t uvr, war, tuv, twa, z, total;
    int tilesX;
    size_t tplane;
};
TileLayoutI16 tile_layout_i16(int s_w, int eta, int xi) {
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const TileLayout B = tile_layout(2, s_w, eta, xi);
    const size_t blocks = (size_t)(s_w - 1) * B.tplane;
    TileLayoutI16 L;
    L.tilesX = B.tilesX;
    L.tplane = B.tplane;
    L.uvr = B.uvr;
    L.war = B.war;
    L.tuv = B.tm;
    L.twa = up(L.tuv + blocks * sizeof(TileScaleUV));
    L.z = up(L.twa + blocks * sizeof(TileScaleWA));
    L.total = up(L.z + (size_t)eta * xi * sizeof(ZH));
    return L;
}

hipError_t launch_tiled_i16(const MoveParams &P, const RecordPair<float> &F, void *packBuf, const GridPtrs &G,
                            const SoA &s, int n, unsigned long long *ctr, bool events, const uint32_t *perm,
                            int stages, hipStream_t stream) {
    const TileLayoutI16 L = tile_layout_i16(P.s_w, P.eta, P.xi);
    char *base = (char *)packBuf;
    TiledI16 S;
    S.uvr = (const UV4<int16_t> *)(base + L.uvr);
    S.war = (const WA8<int16_t> *)(base + L.war);
    S.tuv = (const TileScaleUV *)(base + L.tuv);
    S.twa = (const TileScaleWA *)(base + L.twa);
    S.z = (const ZH *)(base + L.z);
    S.md = (const float2 *)G.idxf;
    S.depthf = G.depthf;
    S.xi = P.xi;
    S.nlev = P.s_w - 1;
    S.tilesX = L.tilesX;
    S.tplane = L.tplane;
    if (stages & kStagePack) {
        const int strips = (L.tilesX + kI16Tiles - 1) / kI16Tiles;
        const size_t blocks = (size_t)S.nlev * (L.tplane / (size_t)L.tilesX) * strips;
        const size_t lds = 8 * (kTile + 1) * kI16Pitch * sizeof(float);
        tile_pack_strip_kernel<int16_t><<<(unsigned)blocks, 256, lds, stream>>>(
            F, G.hf, (TileScaleUV *)S.tuv, (TileScaleWA *)S.twa, nullptr, (UV4<int16_t> *)S.uvr,
            (WA8<int16_t> *)S.war, (ZH *)S.z, P.s_rho, P.s_w, P.eta, P.xi, L.tilesX, L.tplane);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (!(stages & kStageMove) || n <= 0) return hipSuccess;
    return launch_move<int16_t>(P, S, s, n, ctr, events, perm, stream);
}

template <class T>
hipError_t launch_tiled_t(const MoveParams &P, const RecordPair<float> &F, void *packBuf, const GridPtrs &G,
                          const SoA &s, int n, unsigned long long *ctr, bool events, const uint32_t *perm,
                          int stages, hipStream_t stream) {
    const TileLayout L = tile_layout(sizeof(T), P.s_w, P.eta, P.xi);
    char *base = (char *)packBuf;
    TiledPacked<T> S;
    S.uvr = (const UV4<T> *)(base + L.uvr);
    S.war = (const WA8<T> *)(base + L.war);
    S.tm = (const WA8<T> *)(base + L.tm);
    S.z = (const ZH *)(base + L.z);
    S.md = (const float2 *)G.idxf;
    S.depthf = G.depthf;
    S.xi = P.xi;
    S.nlev = P.s_w - 1;
    S.tilesX = L.tilesX;
    S.tplane = L.tplane;
    if (stages & kStagePack) {
        const int strips = (L.tilesX + kI16Tiles - 1) / kI16Tiles;
        const size_t blocks = (size_t)S.nlev * (L.tplane / (size_t)L.tilesX) * strips;
        const size_t lds = 8 * (kTile + 1) * kI16Pitch * sizeof(float);
        tile_pack_strip_kernel<T><<<(unsigned)blocks, 256, lds, stream>>>(
            F, G.hf, nullptr, nullptr, (WA8<T> *)S.tm, (UV4<T> *)S.uvr, (WA8<T> *)S.war, (ZH *)S.z, P.s_rho, P.s_w,
            P.eta, P.xi, L.tilesX, L.tplane);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (!(stages & kStageMove) || n <= 0) return hipSuccess;
    return launch_move<T>(P, S, s, n, ctr, events, perm, stream);
}

template <class T>
hipError_t launch_t(const MoveParams &P, const RecordPair<float> &F, void *packBuf, const GridPtrs &G, const SoA &s,
                    int n, unsigned long long *ctr, bool events, const uint32_t *perm, int stages, const LdsTiling &lt,
                    hipStream_t stream) {
    const size_t plane = (size_t)P.eta * P.xi;
    const int nlev = P.s_w - 1;
    Packed<T> S;
    char *base = (char *)packBuf;
    S.uv = (const UV4<T> *)base;
    base += sizeof(UV4<T>) * plane * nlev;
    S.wa = (const WA8<T> *)base;
    base += sizeof(WA8<T>) * plane * nlev;
    S.z = (const ZH *)base;
    S.md = (const float2 *)G.idxf;
    S.depthf = G.depthf;
    S.xi = P.xi;
    S.nlev = nlev;
    S.plane = plane;
    if (stages & kStagePack) {
        pack_kernel<T><<<grid_blocks(plane * nlev), 256, 0, stream>>>(F, G.hf, (UV4<T> *)S.uv, (WA8<T> *)S.wa,
                                                                      (ZH *)S.z, P.s_rho, P.s_w, plane);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (!(stages & kStageMove) || n <= 0) return hipSuccess;
    if (lt.chunk > 0 && !P.jibe && !events) {
        move_fast_lds_kernel<T><<<(n + 255) / 256, 256, kMoveLdsBytes, stream>>>(P, S, s, n, lt.chunk, lt.halo,
                                                                                 kMoveLdsBytes, ctr, perm);
        return hipGetLastError();
    }
    return launch_move<T>(P, S, s, n, ctr, events, perm, stream);
}

template <class T>
Packed<T> packed_view(void *packBuf, const GridPtrs &G, int s_w, int xi, size_t plane) {
    const int nlev = s_w - 1;
    Packed<T> S;
    char *base = (char *)packBuf;
    S.uv = (const UV4<T> *)base;
    base += sizeof(UV4<T>) * plane * nlev;
    S.wa = (const WA8<T> *)base;
    base += sizeof(WA8<T>) * plane * nlev;
    S.z = (const ZH *)base;
    S.md = (const float2 *)G.idxf;
    S.depthf = G.depthf;
    S.xi = xi;
    S.nlev = nlev;
    S.plane = plane;
    return S;
}

hipError_t ens_sort(const EnsLaunch &L, hipStream_t stream) {
    if (L.sortTile <= 0 || L.sortPeriod <= 0 || L.n <= 0) return hipSuccess;
    const int tilesX = (L.xi + L.sortTile - 1) / L.sortTile, tilesY = (L.eta + L.sortTile - 1) / L.sortTile;
    const uint32_t nTiles = (uint32_t)tilesX * (uint32_t)tilesY, nBins = nTiles + 1;
    const int nb = (L.n + 255) / 256;
    ens_sort_clear_kernel<<<(nBins + 255) / 256, 256, 0, stream>>>(L.stepIdx, L.sortPeriod, L.bins, nBins);
    ens_sort_hist_kernel<<<nb, 256, 0, stream>>>(L.stepIdx, L.sortPeriod, L.s.cell, L.n, L.xi, L.sortTile, tilesX,
                                                 nTiles, L.bins);
    ens_sort_scan_kernel<<<1, 1024, 0, stream>>>(L.stepIdx, L.sortPeriod, L.bins, nBins);
    ens_sort_scatter_kernel<<<nb, 256, 0, stream>>>(L.stepIdx, L.sortPeriod, L.s, L.tmp, L.n, L.xi, L.sortTile,
                                                    tilesX, nTiles, L.bins);
    ens_sort_copy_kernel<<<nb < 16384 ? nb : 16384, 256, 0, stream>>>(L.stepIdx, L.sortPeriod, L.tmp, L.s, L.n);
    return hipGetLastError();
}

template <class T>
hipError_t ens_launch_t(const EnsLaunch &L, hipStream_t stream) {
    const size_t plane = (size_t)L.eta * L.xi;
    const Packed<T> S = packed_view<T>(L.packBuf, L.G, L.s_w, L.xi, plane);
    hipError_t se = ens_sort(L, stream);
    if (se != hipSuccess) return se;
    ens_pack_kernel<T><<<grid_blocks(plane * (L.s_w - 1)), 256, 0, stream>>>(
        L.steps, L.stepIdx, L.G.hf, (UV4<T> *)S.uv, (WA8<T> *)S.wa, (ZH *)S.z, L.s_rho, L.s_w, plane);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (L.n > 0 && L.ldsChunk > 0) {
        const int blocks = (L.n + 255) / 256;
        if (L.occ == 3)
            move_ens_lds_kernel<T, 3><<<blocks, 256, kEnsLdsBytes3, stream>>>(
                L.steps, L.stepIdx, S, L.s, L.n, L.sigma, L.ldsChunk, L.ldsHalo, kEnsLdsBytes3, L.stepCounter);
        else
            move_ens_lds_kernel<T, 0><<<blocks, 256, kEnsLdsBytes2, stream>>>(
                L.steps, L.stepIdx, S, L.s, L.n, L.sigma, L.ldsChunk, L.ldsHalo, kEnsLdsBytes2, L.stepCounter);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    } else if (L.n > 0) {
        const long long per = (long long)L.block * L.ppt;
        const int blocks = (int)((L.n + per - 1) / per);
        if (L.occ == 4)
            move_ens_kernel<T, 4><<<blocks, L.block, 0, stream>>>(L.steps, L.stepIdx, S, L.s, L.n, L.sigma, L.ppt,
                                                                   L.stepCounter);
        else if (L.occ == 3)
            move_ens_kernel<T, 3><<<blocks, L.block, 0, stream>>>(L.steps, L.stepIdx, S, L.s, L.n, L.sigma, L.ppt,
                                                                   L.stepCounter);
        else
            move_ens_kernel<T, 0><<<blocks, L.block, 0, stream>>>(L.steps, L.stepIdx, S, L.s, L.n, L.sigma, L.ppt,
                                                                   L.stepCounter);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (L.surface) {
        const size_t cells = (size_t)L.nScen * plane;
        e = launch_fill_u32(L.surface, cells, 0u, stream);
        if (e != hipSuccess) return e;
        if (L.n > 0) {
            ens_surface_kernel<<<(L.n + 255) / 256, 256, 0, stream>>>(L.steps, L.stepIdx, L.s, L.n, L.eta, L.xi,
                                                                      L.surface);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
    }
    ens_advance_kernel<<<1, 1, 0, stream>>>(L.stepIdx);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_ens_release(const EnsSource *src, const unsigned long long *first, int nsrc, long long n, int eta,
                              int xi, uint64_t seed, const EnsSoA &st, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    long long b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    ens_release_kernel<<<(int)b, 256, 0, stream>>>(src, first, nsrc, n, eta, xi, seed, st);
    return hipGetLastError();
}

hipError_t launch_ens_step(int store, const EnsLaunch &L, hipStream_t stream) {
    if (store != WC_STORE_F32 && store != WC_STORE_F16 && store != WC_STORE_BF16) return hipErrorInvalidValue;
    if (L.block != 64 && L.block != 128 && L.block != 256) return hipErrorInvalidValue;
    if (L.ppt < 1 || (L.occ != 0 && L.occ != 3 && L.occ != 4)) return hipErrorInvalidValue;
    if (L.ldsChunk < 0 || L.ldsHalo < 0 || (L.ldsChunk > 0 && L.occ == 4)) return hipErrorInvalidValue;
    if (store == WC_STORE_F16) return ens_launch_t<_Float16>(L, stream);
    if (store == WC_STORE_BF16) return ens_launch_t<uint16_t>(L, stream);
    return ens_launch_t<float>(L, stream);
}

size_t fast_pack_bytes(int store, int s_w, int eta, int xi) {
    if (store == WC_STORE_BF16_TILE || store == WC_STORE_F16_TILE) return tile_layout(2, s_w, eta, xi).total;
    if (store == WC_STORE_I16_TILE) return tile_layout_i16(s_w, eta, xi).total;
    const size_t nlev = (size_t)(s_w - 1), plane = (size_t)eta * xi;
    const size_t t = store == WC_STORE_F32 ? 4 : 2;
    return plane * nlev * (4 * t + 8 * t) + plane * sizeof(ZH);
}

hipError_t launch_move_fast_soa(int store, const MoveParams &P, const RecordPair<float> &F, void *packBuf,
                                const GridPtrs &G, const SoA &s, int n, unsigned long long *ctr, bool events,
                                const uint32_t *perm, int stages, const LdsTiling &lt, hipStream_t stream) {
    if (store == WC_STORE_F16)
        return launch_t<_Float16>(P, F, packBuf, G, s, n, ctr, events, perm, stages, lt, stream);
    if (store == WC_STORE_BF16)
        return launch_t<uint16_t>(P, F, packBuf, G, s, n, ctr, events, perm, stages, lt, stream);
    if (store == WC_STORE_F16_TILE)
        return launch_tiled_t<_Float16>(P, F, packBuf, G, s, n, ctr, events, perm, stages, stream);
    if (store == WC_STORE_BF16_TILE)
        return launch_tiled_t<uint16_t>(P, F, packBuf, G, s, n, ctr, events, perm, stages, stream);
    if (store == WC_STORE_I16_TILE) return launch_tiled_i16(P, F, packBuf, G, s, n, ctr, events, perm, stages, stream);
    return launch_t<float>(P, F, packBuf, G, s, n, ctr, events, perm, stages, lt, stream);
}

}  // namespace wc
