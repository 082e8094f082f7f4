// ws_kernels.hip -- the learned warm-start network's kernels for gfx950:
// 3x3 convolutions as implicit GEMMs on bfloat16 MFMA (v_mfma_f32_32x32x16_bf16)
// accumulating in FP32, their weight gradients, the masked MSE loss and Adam.
//
// No counterpart in the reference; see include/hystrath_ws.h for the layouts.
//
// MFMA 32x32x16 operand layout (wave of 64 lanes, lane l, half h = l / 32):
//   A (32 x 16, rows m):  lane l holds A[m = l % 32][k = 8h .. 8h+7]
//   B (16 x 32, cols n):  lane l holds B[k = 8h .. 8h+7][n = l % 32]
//   D (32 x 32):          lane l holds D[m = 8(v/4) + 4h + v%4][n = l % 32], v = 0..15
// (checked on the device by hs_ws_mfma_tile in tests/test_ws_kernels.py).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "../../include/hystrath_dsmc.h"
#include "../../include/hystrath_ws.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kThreads = 256;  // four waves

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    const uint4 u = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ bf16x8 zero8() {
    const uint4 u = { 0u, 0u, 0u, 0u };
    return __builtin_bit_cast(bf16x8, u);
}

// One wave: 32 consecutive pixels of one image row times NB blocks of 32
// output channels.  The K loop runs over the 9 taps and the input channels
// in steps of 16.  Taps are dil pixels apart (a dilated 3x3 convolution,
// zero padding dil).  Waves are independent (no barriers): a wave past the
// end returns.
template <int NB>
__global__ __launch_bounds__(kThreads) void k_conv3x3(const uint16_t* __restrict__ x,
                                                     const uint16_t* __restrict__ wp,
                                                     const float* __restrict__ bias,
                                                     const uint16_t* __restrict__ mask,
                                                     uint16_t* __restrict__ y, float* __restrict__ y32, int n,
                                                     int h, int w, int cin, int cout, int relu, int dil) {
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int m = l & 31, hh = l >> 5;
    const int segs = w >> 5;
    const int64_t seg = (int64_t)blockIdx.x * (kThreads / 64) + wave;
    if (seg >= (int64_t)n * h * segs) return;
    const int xs = (int)(seg % segs);
    const int64_t row = seg / segs;
    const int yy0 = (int)(row % h);
    const int64_t img = row / h;
    const int x0 = xs << 5;
    const int cout_pad = NB * 32;
    f32x16 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[nb][v] = 0.0f;
    for (int tap = 0; tap < 9; ++tap) {
        const int ky = tap / 3, kx = tap - 3 * ky;
        const int yy = yy0 + (ky - 1) * dil, xx = x0 + m + (kx - 1) * dil;
        const bool ok = yy >= 0 && yy < h && xx >= 0 && xx < w;
        const uint16_t* px = x + ((img * h + (ok ? yy : 0)) * w + (ok ? xx : 0)) * cin + 8 * hh;
        const uint16_t* pw = wp + ((int64_t)tap * cout_pad + m) * cin + 8 * hh;
        for (int c0 = 0; c0 < cin; c0 += 16) {
            const bf16x8 a = ok ? load8(px + c0) : zero8();
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const bf16x8 b = load8(pw + (int64_t)nb * 32 * cin + c0);
                acc[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[nb], 0, 0, 0);
            }
        }
    }
    const int64_t pix0 = (img * h + yy0) * w + x0;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const int co = nb * 32 + m;
        if (co >= cout) continue;
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int64_t pix = pix0 + 8 * (v >> 2) + 4 * hh + (v & 3);
            float val = acc[nb][v] + bv;
            if (relu) val = fmaxf(val, 0.0f);
            if (mask && !(bf2f(mask[pix * cout + co]) > 0.0f)) val = 0.0f;
            if (y) y[pix * cout + co] = f2bf(val);
            if (y32) y32[pix * cout + co] = val;
        }
    }
}

// Weight gradient: dw[tap][ci][co] += sum_p x(p + dil * tap) dy(p).  A workgroup
// takes whole image rows; per chunk of 16 pixels it stages the nine shifted
// input patches and the output-gradient patch in LDS, transposed so each
// lane's 8 consecutive pixels are one 16-byte read, and the four waves share
// the 9 x MB x NB output tiles (K = pixels).
constexpr int kMaxC = 64;
constexpr int kChunk = 16;

template <int MB, int NB>
__global__ __launch_bounds__(kThreads) void k_wgrad(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy,
                                                   float* __restrict__ dw, float* __restrict__ db, int n, int h,
                                                   int w, int cin, int cout, int rows_per_block, int dil) {
    __shared__ __attribute__((aligned(16))) uint16_t xs[9][kMaxC][kChunk];   // [tap][ci][pixel]
    __shared__ __attribute__((aligned(16))) uint16_t ds[kMaxC][kChunk];      // [co][pixel]
    constexpr int T = 9 * MB * NB;
    constexpr int S = (T + 3) / 4;
    const int t = threadIdx.x, wave = t >> 6, l = t & 63, m = l & 31, hh = l >> 5;
    f32x16 acc[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[s][v] = 0.0f;
    float bsum = 0.0f;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
    const int64_t r1 = min((int64_t)n * h, r0 + rows_per_block);
    for (int64_t row = r0; row < r1; ++row) {
        const int yy = (int)(row % h);
        const int64_t img = row / h;
        for (int p0 = 0; p0 < w; p0 += kChunk) {
            for (int e = t; e < 9 * cin * kChunk; e += kThreads) {
                const int ci = e % cin, j = (e / cin) % kChunk, tap = e / (cin * kChunk);
                const int ky = tap / 3, kx = tap - 3 * ky;
                const int sy = yy + (ky - 1) * dil, sx = p0 + j + (kx - 1) * dil;
                uint16_t v = 0;
                if (sy >= 0 && sy < h && sx >= 0 && sx < w) v = x[((img * h + sy) * w + sx) * cin + ci];
                xs[tap][ci][j] = v;
            }
            for (int e = t; e < cout * kChunk; e += kThreads) {
                const int co = e % cout, j = e / cout;
                ds[co][j] = dy[((img * h + yy) * w + p0 + j) * cout + co];
            }
            __syncthreads();
            if (t < cout) {
#pragma unroll
                for (int j = 0; j < kChunk; ++j) bsum += bf2f(ds[t][j]);
            }
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int tile = wave + 4 * s;
                if (tile < T) {
                    const int tap = tile / (MB * NB), mb = (tile / NB) % MB, nb = tile % NB;
                    const int ci = mb * 32 + m, co = nb * 32 + m;
                    const bf16x8 a = ci < cin ? load8(&xs[tap][ci][8 * hh]) : zero8();
                    const bf16x8 b = co < cout ? load8(&ds[co][8 * hh]) : zero8();
                    acc[s] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[s], 0, 0, 0);
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int tile = wave + 4 * s;
        if (tile >= T) continue;
        const int tap = tile / (MB * NB), mb = (tile / NB) % MB, nb = tile % NB;
        const int co = nb * 32 + m;
        if (co >= cout) continue;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int ci = mb * 32 + 8 * (v >> 2) + 4 * hh + (v & 3);
            if (ci < cin) unsafeAtomicAdd(&dw[((int64_t)tap * cin + ci) * cout + co], acc[s][v]);
        }
    }
    if (t < cout && db) unsafeAtomicAdd(&db[t], bsum);
}

__global__ void k_pack(const float* __restrict__ wm, int cin, int cout, int cout_pad, int cin_pad,
                       uint16_t* __restrict__ wf, uint16_t* __restrict__ wb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // forward [9][cout_pad][cin]
    if (i < 9LL * cout_pad * cin) {
        const int ci = (int)(i % cin), co = (int)((i / cin) % cout_pad), tap = (int)(i / ((int64_t)cin * cout_pad));
        wf[i] = co < cout ? f2bf(wm[((int64_t)tap * cin + ci) * cout + co]) : (uint16_t)0;
    }
    // input gradient [9][cin_pad][cout], taps flipped
    if (i < 9LL * cin_pad * cout) {
        const int co = (int)(i % cout), ci = (int)((i / cout) % cin_pad), tp = (int)(i / ((int64_t)cout * cin_pad));
        const int tap = 8 - tp;
        wb[i] = ci < cin ? f2bf(wm[((int64_t)tap * cin + ci) * cout + co]) : (uint16_t)0;
    }
}

__global__ __launch_bounds__(kThreads) void k_mse(const float* __restrict__ pred, const float* __restrict__ tgt,
                                                 const float* __restrict__ mask, uint16_t* __restrict__ dy,
                                                 float* __restrict__ loss, int64_t npix, int c, int nf,
                                                 float scale) {
    __shared__ float red[kThreads / 64];
    const int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    float s = 0.0f;
    if (p < npix) {
        const float mk = mask[p];
        for (int f = 0; f < c; ++f) {
            float g = 0.0f;
            if (f < nf && mk != 0.0f) {
                const float d = pred[p * c + f] - tgt[p * c + f];
                s += mk * d * d;
                g = 2.0f * mk * d * scale;
            }
            dy[p * c + f] = f2bf(g);
        }
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) unsafeAtomicAdd(loss, (red[0] + red[1] + red[2] + red[3]) * scale);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_mfma_tile(int kind, const uint16_t* a, const uint16_t* b, float* d) {
    const int l = threadIdx.x;
    const bf16x8 av = load8(a + 8 * l), bv = load8(b + 8 * l);
    if (kind == 0) {
        f32x16 c;
        for (int v = 0; v < 16; ++v) c[v] = 0.0f;
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, c, 0, 0, 0);
        for (int v = 0; v < 16; ++v) d[16 * l + v] = c[v];
    } else {
        f32x4 c;
        for (int v = 0; v < 4; ++v) c[v] = 0.0f;
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, c, 0, 0, 0);
        for (int v = 0; v < 4; ++v) d[4 * l + v] = c[v];
    }
}

bool channels_ok(int c) { return c == 16 || c == 32 || c == 64; }

int launched() { return hipGetLastError() == hipSuccess ? 0 : HS_E_HIP; }

}  // namespace

extern "C" int hs_ws_pack(const float* w, int32_t cin, int32_t cout, uint16_t* w_fwd, uint16_t* w_bwd,
                          void* stream) {
    if (!w || !w_fwd || !w_bwd || cin < 1 || cout < 1 || cin > kMaxC || cout > kMaxC) return HS_E_INVALID;
    const int cout_pad = (cout + 31) / 32 * 32, cin_pad = (cin + 31) / 32 * 32;
    const int64_t nmax = 9LL * (cout_pad * (int64_t)cin > cin_pad * (int64_t)cout ? cout_pad * (int64_t)cin
                                                                                  : cin_pad * (int64_t)cout);
    k_pack<<<(unsigned)((nmax + 255) / 256), 256, 0, (hipStream_t)stream>>>(w, cin, cout, cout_pad, cin_pad, w_fwd,
                                                                             w_bwd);
    return launched();
}

extern "C" int hs_ws_conv3x3(const uint16_t* x, const uint16_t* w_packed, const float* bias, const uint16_t* mask,
                             uint16_t* y, float* y32, int32_t n, int32_t h, int32_t w, int32_t cin, int32_t cout,
                             int32_t relu, int32_t dilation, void* stream) {
    if (!x || !w_packed || (!y && !y32) || n < 1 || h < 1 || w < 32 || (w & 31) || !channels_ok(cin) ||
        cout < 1 || cout > kMaxC || dilation < 1 || dilation > 64)
        return HS_E_INVALID;
    const int64_t waves = (int64_t)n * h * (w / 32);
    const unsigned grid = (unsigned)((waves + 3) / 4);
    hipStream_t st = (hipStream_t)stream;
    if (cout <= 32)
        k_conv3x3<1><<<grid, kThreads, 0, st>>>(x, w_packed, bias, mask, y, y32, n, h, w, cin, cout, relu, dilation);
    else
        k_conv3x3<2><<<grid, kThreads, 0, st>>>(x, w_packed, bias, mask, y, y32, n, h, w, cin, cout, relu, dilation);
    return launched();
}

extern "C" int hs_ws_conv3x3_wgrad(const uint16_t* x, const uint16_t* dy, float* dw, float* db, int32_t n, int32_t h,
                                   int32_t w, int32_t cin, int32_t cout, int32_t dilation, void* stream) {
    if (!x || !dy || !dw || n < 1 || h < 1 || w < kChunk || (w % kChunk) || !channels_ok(cin) || !channels_ok(cout) ||
        dilation < 1 || dilation > 64)
        return HS_E_INVALID;
    const int64_t rows = (int64_t)n * h;
    const int64_t blocks = rows < 2048 ? rows : 2048;
    const int rpb = (int)((rows + blocks - 1) / blocks);
    const unsigned grid = (unsigned)((rows + rpb - 1) / rpb);
    hipStream_t st = (hipStream_t)stream;
    const int mb = cin > 32 ? 2 : 1, nb = cout > 32 ? 2 : 1;
    if (mb == 1 && nb == 1) k_wgrad<1, 1><<<grid, kThreads, 0, st>>>(x, dy, dw, db, n, h, w, cin, cout, rpb, dilation);
    else if (mb == 1) k_wgrad<1, 2><<<grid, kThreads, 0, st>>>(x, dy, dw, db, n, h, w, cin, cout, rpb, dilation);
    else if (nb == 1) k_wgrad<2, 1><<<grid, kThreads, 0, st>>>(x, dy, dw, db, n, h, w, cin, cout, rpb, dilation);
    else k_wgrad<2, 2><<<grid, kThreads, 0, st>>>(x, dy, dw, db, n, h, w, cin, cout, rpb, dilation);
    return launched();
}

extern "C" int hs_ws_mse(const float* pred, const float* target, const float* mask, uint16_t* dy, float* loss,
                         int32_t n, int32_t h, int32_t w, int32_t c, int32_t nf, double count, void* stream) {
    if (!pred || !target || !mask || !dy || !loss || n < 1 || h < 1 || w < 1 || c < 1 || nf < 1 || nf > c ||
        !(count > 0.0))
        return HS_E_INVALID;
    const int64_t npix = (int64_t)n * h * w;
    const float scale = (float)(1.0 / (nf * count));
    k_mse<<<(unsigned)((npix + kThreads - 1) / kThreads), kThreads, 0, (hipStream_t)stream>>>(
        pred, target, mask, dy, loss, npix, c, nf, scale);
    return launched();
}

extern "C" int hs_ws_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float beta1,
                          float beta2, float eps, int32_t t, void* stream) {
    if (!p || !g || !m || !v || n < 1 || t < 1) return HS_E_INVALID;
    const float c1 = 1.0f - powf(beta1, (float)t), c2 = 1.0f - powf(beta2, (float)t);
    k_adam<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(p, g, m, v, n, lr, beta1, beta2, eps, c1, c2);
    return launched();
}

extern "C" int hs_ws_mfma_tile(int32_t kind, const uint16_t* a, const uint16_t* b, float* d) {
    if (!a || !b || !d || (kind != 0 && kind != 1)) return HS_E_INVALID;
    uint16_t *da = nullptr, *dbp = nullptr;
    float* dd = nullptr;
    const size_t nd = kind == 0 ? 64 * 16 : 64 * 4;
    int rc = 0;
    if (hipMalloc(&da, 64 * 8 * 2) != hipSuccess || hipMalloc(&dbp, 64 * 8 * 2) != hipSuccess ||
        hipMalloc(&dd, nd * 4) != hipSuccess) {
        rc = HS_E_HIP;
    } else if (hipMemcpy(da, a, 64 * 8 * 2, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(dbp, b, 64 * 8 * 2, hipMemcpyHostToDevice) != hipSuccess) {
        rc = HS_E_HIP;
    } else {
        k_mfma_tile<<<1, 64>>>(kind, da, dbp, dd);
        if (hipGetLastError() != hipSuccess || hipMemcpy(d, dd, nd * 4, hipMemcpyDeviceToHost) != hipSuccess)
            rc = HS_E_HIP;
    }
    if (da) (void)hipFree(da);
    if (dbp) (void)hipFree(dbp);
    if (dd) (void)hipFree(dd);
    return rc;
}
