// BF16 MFMA kernels of the learned warm start (a small U-Net that maps
// channel-wise backprojections to a TGV estimate used as the initial u).
//
// Tensors are NHWC bf16 with a pixel stride `ld` (channels allocated per
// pixel) and a channel offset, so that a tensor can be a channel slice of a
// wider buffer (U-Net skip concatenation without copies).  Weights are
// W[co][tap][ci] (ci fastest), tap = 3*ky + kx for 3x3 (padding 1) or a
// single tap for 1x1.  All products run on v_mfma_f32_32x32x16_bf16 with
// FP32 accumulation.
//
// MFMA 32x32x16 operand layout (lane l = 0..63, h = l / 32):
//   A[m][k]: lane holds m = l % 32, k = 8h .. 8h+7
//   B[k][n]: lane holds n = l % 32, k = 8h .. 8h+7
//   D[m][n]: lane holds n = l % 32, rows m = 8(r/4) + 4h + r%4, r = 0..15
#include "common.h"

#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>

namespace gr {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

struct alignas(16) U16x8 {
    uint16_t v[8];
};

__device__ __forceinline__ bf16x8 as_bf16x8(const U16x8& u)
{
    bf16x8 r;
    __builtin_memcpy(&r, &u, 16);
    return r;
}

__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ---------------------------------------------------------------------------
// Convolution forward, LDS-tiled: a work-group of 4 waves computes MP
// consecutive pixels of one output row x NC output channels.  Per chunk of
// KC input channels it stages the R input rows the taps touch (MP + 2
// pixels, zero outside the image and beyond cin) and the NC x taps x KC
// weights in LDS; each wave then runs TM x TN MFMA tiles per 16-wide k step,
// reusing every A fragment TN times and every B fragment TM times.  Strides
// in LDS are padded (KC + 8 halves per pixel, taps * KC + 8 per channel) so
// that the 16-byte fragment reads of 16 consecutive lanes hit distinct banks.
// ---------------------------------------------------------------------------
template <int WAVES_M_, int TM_, int TN_, int KC_, int TAPS_>
struct ConvTile {
    static constexpr int WAVES_M = WAVES_M_, TM = TM_, TN = TN_, KC = KC_, TAPS = TAPS_;
    static constexpr int WAVES_N = 4 / WAVES_M;
    static constexpr int MP = WAVES_M * TM * 32;  // pixels per work-group
    static constexpr int NC = WAVES_N * TN * 32;  // output channels per work-group
    static constexpr int AS = KC + 8;             // LDS halves per staged pixel
    static constexpr int R = TAPS == 9 ? 3 : 1;   // staged input rows
    static constexpr int P = TAPS == 9 ? MP + 2 : MP;  // staged pixels per row
    static constexpr int BS = TAPS * KC + 8;      // LDS halves per staged output channel
    static constexpr int G = KC / 8;              // 16-byte groups per pixel / tap
    static constexpr int NA = (R * P * G + 255) / 256;     // A groups per thread
    static constexpr int NB = (NC * TAPS * G + 255) / 256;  // B groups per thread
    static constexpr size_t lds_bytes = 2 * ((size_t)R * P * AS + (size_t)NC * BS);
};

// Staging of one channel chunk: global -> registers (conv_stage_load) and
// registers -> LDS (conv_stage_store); thread tid owns groups tid + 256 k.
// 16-byte groups travel as 4 x u32 vectors (kept in registers).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <class T>
__device__ __forceinline__ void conv_stage_load(u32x4 (&ra)[T::NA], u32x4 (&rb)[T::NB],
                                                const uint16_t* __restrict__ x, int ldx, int cx0,
                                                const uint16_t* __restrict__ w, int c0, int tid,
                                                int row, int img, int x0, int co_base, int H,
                                                int W, int cin, int cout)
{
    constexpr int G = T::G, P = T::P, R = T::R, TAPS = T::TAPS, NC = T::NC;
    constexpr int HALO = TAPS == 9 ? 1 : 0;
#pragma unroll
    for (int k = 0; k < T::NA; ++k) {
        const int e = tid + 256 * k;
        const int g = e % G, q = e / G, pp = q % P, r = q / P;
        const int yy = row + r - HALO, xx = x0 + pp - HALO, c = c0 + 8 * g;
        const bool ok = e < R * P * G && yy >= 0 && yy < H && xx >= 0 && xx < W && c < cin;
        ra[k] = ok ? *reinterpret_cast<const u32x4*>(x + ((size_t)(img * H + yy) * W + xx) * ldx +
                                                      cx0 + c)
                   : u32x4{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int k = 0; k < T::NB; ++k) {
        const int e = tid + 256 * k;
        const int g = e % G, q = e / G, t = q % TAPS, co = q / TAPS;
        const int cg = co_base + co, c = c0 + 8 * g;
        const bool ok = e < NC * TAPS * G && cg < cout && c < cin;
        rb[k] = ok ? *reinterpret_cast<const u32x4*>(w + ((size_t)cg * TAPS + t) * cin + c)
                   : u32x4{0u, 0u, 0u, 0u};
    }
}

template <class T>
__device__ __forceinline__ void conv_stage_store(const u32x4 (&ra)[T::NA],
                                                 const u32x4 (&rb)[T::NB], uint16_t* As,
                                                 uint16_t* Bs, int tid)
{
    constexpr int G = T::G, P = T::P, R = T::R, TAPS = T::TAPS, NC = T::NC;
#pragma unroll
    for (int k = 0; k < T::NA; ++k) {
        const int e = tid + 256 * k;
        if (e < R * P * G)
            *reinterpret_cast<u32x4*>(As + (size_t)(e / G) * T::AS + 8 * (e % G)) = ra[k];
    }
#pragma unroll
    for (int k = 0; k < T::NB; ++k) {
        const int e = tid + 256 * k;
        if (e < NC * TAPS * G) {
            const int g = e % G, q = e / G, t = q % TAPS, co = q / TAPS;
            *reinterpret_cast<u32x4*>(Bs + (size_t)co * T::BS + t * T::KC + 8 * g) = rb[k];
        }
    }
}

// The next chunk's input and weight groups are loaded into registers while
// the current chunk computes, then written to LDS after a barrier.
template <int WAVES_M, int TM, int TN, int KC, int TAPS>
__global__ __launch_bounds__(256) void conv_tile_kernel(
    const uint16_t* __restrict__ x, int ldx, int cx0, const uint16_t* __restrict__ w,
    const float* __restrict__ bias, void* __restrict__ y, int ldy, int cy0, int H, int W,
    int cin, int cout, int relu, int out_f32)
{
    using T = ConvTile<WAVES_M, TM, TN, KC, TAPS>;
    constexpr int MP = T::MP, NC = T::NC, AS = T::AS, R = T::R, P = T::P, BS = T::BS, G = T::G;
    constexpr int NA = T::NA, NB = T::NB, HALO = TAPS == 9 ? 1 : 0;
    extern __shared__ __attribute__((aligned(16))) uint16_t lds[];
    uint16_t* As = lds;                  // [R][P][AS]
    uint16_t* Bs = lds + (size_t)R * P * AS;  // [NC][BS]

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, l32 = lane & 31;
    const int segs = (W + MP - 1) / MP, co_blocks = (cout + NC - 1) / NC;
    int b = blockIdx.x;
    const int cob = b % co_blocks;
    b /= co_blocks;
    const int seg = b % segs;
    b /= segs;
    const int row = b % H, img = b / H;
    const int x0 = seg * MP, co_base = cob * NC;
    const int wm = wave % WAVES_M, wn = wave / WAVES_M;

    floatx16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    u32x4 ra[NA], rb[NB];
    conv_stage_load<T>(ra, rb, x, ldx, cx0, w, 0, tid, row, img, x0, co_base, H, W, cin, cout);
    for (int c0 = 0; c0 < cin; c0 += KC) {
        __syncthreads();  // the previous chunk is no longer read
        conv_stage_store<T>(ra, rb, As, Bs, tid);
        __syncthreads();
        if (c0 + KC < cin)  // in flight during this chunk's MFMAs
            conv_stage_load<T>(ra, rb, x, ldx, cx0, w, c0 + KC, tid, row, img, x0, co_base, H, W,
                               cin, cout);
#pragma unroll
        for (int t = 0; t < TAPS; ++t) {
            const int dy = TAPS == 9 ? t / 3 : 0, dx = TAPS == 9 ? t % 3 : 0;
            const uint16_t* arow = As + (size_t)dy * P * AS;
#pragma unroll
            for (int k = 0; k < KC; k += 16) {
                U16x8 a[TM], bb[TN];
#pragma unroll
                for (int i = 0; i < TM; ++i) {
                    const int pl = (wm * TM + i) * 32 + l32 + dx;  // staged pixel index
                    a[i] = *reinterpret_cast<const U16x8*>(arow + (size_t)pl * AS + k + 8 * h);
                }
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    const int cl = (wn * TN + j) * 32 + l32;
                    bb[j] = *reinterpret_cast<const U16x8*>(Bs + (size_t)cl * BS + t * KC + k +
                                                            8 * h);
                }
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            as_bf16x8(a[i]), as_bf16x8(bb[j]), acc[i][j], 0, 0, 0);
            }
        }
    }
    // epilogue: lane holds column co, rows m = 8(r/4) + 4h + r%4 of each tile
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int co = co_base + (wn * TN + j) * 32 + l32;
        if (co >= cout) continue;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < TM; ++i) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = 8 * (r >> 2) + 4 * h + (r & 3);
                const int xo = x0 + (wm * TM + i) * 32 + m;
                if (xo >= W) continue;
                float v = acc[i][j][r] + bv;
                if (relu) v = v > 0.f ? v : 0.f;
                const size_t pix = (size_t)(img * H + row) * W + xo;
                if (out_f32)
                    reinterpret_cast<float*>(y)[pix * ldy + cy0 + co] = v;
                else
                    reinterpret_cast<uint16_t*>(y)[pix * ldy + cy0 + co] = f2bf(v);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[co][tap][ci] += sum over pixels of
// dy[p][co] * x[p shifted by tap][ci].  M = co (32 per wave), N = ci (32 per
// wave), K = pixels.  A work-group covers 64 co x 64 ci and one row of taps
// (the three taps of one vertical offset, or the single 1x1 tap) over a chunk
// of rows; per 64-pixel step it stages dy as [co][px] and the input row as
// [shift][ci][px] (one copy per horizontal shift, so every MFMA operand is
// an aligned 16-byte LDS read), then runs 4 k steps x the row's taps.
// Partial sums go to the FP32 accumulator with atomics (split-K over the
// chunks).
// ---------------------------------------------------------------------------
constexpr int WG_PX = 64;            // pixels per staged step (4 MFMA k steps)
constexpr int WG_LD = WG_PX + 8;     // LDS row stride (halves): conflict-free fragment reads

template <int TAPS>
__global__ __launch_bounds__(256) void conv_wgrad_kernel(
    const uint16_t* __restrict__ x, int ldx, int cx0, const uint16_t* __restrict__ dy, int ldd,
    int cd0, float* __restrict__ dw, int n, int H, int W, int cin, int cout, int rows_per_chunk)
{
    constexpr int NS = TAPS == 9 ? 3 : 1;  // horizontal shifts = taps per work-group
    __shared__ __attribute__((aligned(16))) uint16_t sdy[64][WG_LD];      // [co][px]
    __shared__ __attribute__((aligned(16))) uint16_t sx[NS][64][WG_LD];   // [shift][ci][px]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, l32 = lane & 31;
    const int co_blocks = (cout + 63) / 64, ci_blocks = (cin + 63) / 64;
    int b = blockIdx.x;
    const int cob = b % co_blocks; b /= co_blocks;
    const int cib = b % ci_blocks; b /= ci_blocks;
    const int trow = b % NS; b /= NS;  // vertical offset index (3x3) or 0 (1x1)
    const int chunk = b;               // chunk of rows over all images
    const int total_rows = n * H;
    const int r0 = chunk * rows_per_chunk;
    const int r1 = min(total_rows, r0 + rows_per_chunk);
    const int dyo = TAPS == 9 ? trow - 1 : 0;
    const int wm = wave & 1, wn = wave >> 1;

    floatx16 acc[NS];
#pragma unroll
    for (int d = 0; d < NS; ++d)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[d][r] = 0.f;

    for (int gr_ = r0; gr_ < r1; ++gr_) {
        const int img = gr_ / H, row = gr_ - img * H;
        const int yy = row + dyo;
        for (int x0 = 0; x0 < W; x0 += WG_PX) {
            __syncthreads();
            // dy: 64 px x 8 channel groups (two per thread)
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int e = tid + 256 * k, p = e & 63, g = e >> 6;
                const int xo = x0 + p, co = cob * 64 + 8 * g;
                U16x8 v = {{0, 0, 0, 0, 0, 0, 0, 0}};
                if (co < cout && xo < W)
                    v = *reinterpret_cast<const U16x8*>(
                        dy + ((size_t)(img * H + row) * W + xo) * ldd + cd0 + co);
#pragma unroll
                for (int j = 0; j < 8; ++j) sdy[8 * g + j][p] = v.v[j];
            }
            // x: pixels x0 - 1 .. x0 + 64 of row yy (3x3) or x0 .. x0 + 63 (1x1),
            // written once per horizontal shift they serve
            constexpr int QP = TAPS == 9 ? WG_PX + 2 : WG_PX;
            for (int e = tid; e < QP * 8; e += 256) {
                const int q = e % QP, g = e / QP;
                const int xs = x0 + q - (TAPS == 9 ? 1 : 0), ci = cib * 64 + 8 * g;
                U16x8 v = {{0, 0, 0, 0, 0, 0, 0, 0}};
                if (ci < cin && yy >= 0 && yy < H && xs >= 0 && xs < W)
                    v = *reinterpret_cast<const U16x8*>(
                        x + ((size_t)(img * H + yy) * W + xs) * ldx + cx0 + ci);
#pragma unroll
                for (int d = 0; d < NS; ++d) {
                    const int p = q - d;  // shift d reads pixel x0 + p + d - 1
                    if (p >= 0 && p < WG_PX) {
#pragma unroll
                        for (int j = 0; j < 8; ++j) sx[d][8 * g + j][p] = v.v[j];
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int ks = 0; ks < WG_PX / 16; ++ks) {
                const U16x8 a =
                    *reinterpret_cast<const U16x8*>(&sdy[wm * 32 + l32][ks * 16 + 8 * h]);
#pragma unroll
                for (int d = 0; d < NS; ++d) {
                    const U16x8 bb =
                        *reinterpret_cast<const U16x8*>(&sx[d][wn * 32 + l32][ks * 16 + 8 * h]);
                    acc[d] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(a), as_bf16x8(bb),
                                                                     acc[d], 0, 0, 0);
                }
            }
        }
    }
    const int ci = cib * 64 + wn * 32 + l32;
#pragma unroll
    for (int d = 0; d < NS; ++d) {
        const int t = TAPS == 9 ? 3 * trow + d : 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int co = cob * 64 + wm * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
            if (co < cout && ci < cin) atomicAdd(&dw[((size_t)co * TAPS + t) * cin + ci], acc[d][r]);
        }
    }
}

// db[co] += sum over pixels of dy[p][co] (FP32 accumulation).  A block of
// 256 threads covers 256 / cout pixel lanes x cout channels (cout divides
// 256); lanes are summed in LDS and each block adds one partial per channel.
__global__ __launch_bounds__(256) void bias_grad_kernel(const uint16_t* __restrict__ dy, int ldd,
                                                        int cd0, size_t npix, int cout,
                                                        float* __restrict__ db)
{
    __shared__ float red[256];
    const int co = threadIdx.x % cout, lanes = 256 / cout, lane = threadIdx.x / cout;
    float acc = 0.f;
    for (size_t p = (size_t)blockIdx.x * lanes + lane; p < npix; p += (size_t)gridDim.x * lanes)
        acc += bf2f(dy[p * ldd + cd0 + co]);
    red[threadIdx.x] = acc;
    __syncthreads();
    if (lane == 0) {
        for (int l = 1; l < lanes; ++l) acc += red[l * cout + co];
        atomicAdd(&db[co], acc);
    }
}

// The same sum with 16-byte reads: a thread owns 8 consecutive channels of
// one pixel lane (cout / 8 groups, 256 * 8 / cout pixel lanes per block).
__global__ __launch_bounds__(256) void bias_grad8_kernel(const uint16_t* __restrict__ dy,
                                                         int ldd, int cd0, size_t npix, int cout,
                                                         float* __restrict__ db)
{
    __shared__ float red[8][256];
    const int groups = cout / 8, lanes = 256 / groups;
    const int g = threadIdx.x % groups, lane = threadIdx.x / groups;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const size_t stride = (size_t)gridDim.x * lanes;
    // four independent loads in flight per thread
    for (size_t p = (size_t)blockIdx.x * lanes + lane; p < npix; p += 4 * stride) {
        U16x8 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const size_t q = p + u * stride;
            v[u] = q < npix ? *reinterpret_cast<const U16x8*>(dy + q * ldd + cd0 + 8 * g)
                            : U16x8{{0, 0, 0, 0, 0, 0, 0, 0}};
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[j] += bf2f(v[u].v[j]);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) red[j][threadIdx.x] = acc[j];
    __syncthreads();
    if (threadIdx.x < cout) {  // one thread per channel sums the pixel lanes
        const int c = threadIdx.x, gg = c / 8, j = c % 8;
        float s = 0.f;
        for (int l = 0; l < lanes; ++l) s += red[j][l * groups + gg];
        atomicAdd(&db[c], s);
    }
}

// Wt[ci][8 - tap][co] = W[co][tap][ci]: weights of the data-gradient
// convolution (transpose of channels, flip of taps).
__global__ void weight_flip_kernel(const uint16_t* __restrict__ w, uint16_t* __restrict__ wt,
                                   int cin, int cout, int taps)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)cin * cout * taps;
    if (i >= total) return;
    const int ci = (int)(i % cin);
    const int t = (int)((i / cin) % taps);
    const int co = (int)(i / ((size_t)cin * taps));
    wt[((size_t)ci * taps + (taps - 1 - t)) * cout + co] = w[i];
}

// dy <- dy * (y > 0) (ReLU backward, y = forward output), in place
__global__ void relu_bwd_kernel(uint16_t* __restrict__ dy, int ldd, int cd0,
                                const uint16_t* __restrict__ y, int ldy, int cy0, size_t npix,
                                int c)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix * c) return;
    const size_t p = i / c;
    const int ch = (int)(i % c);
    if (!(bf2f(y[p * ldy + cy0 + ch]) > 0.f)) dy[p * ldd + cd0 + ch] = 0;
}

// 2x2 average pool (fwd) and its adjoint (bwd) on NHWC bf16
__global__ void pool_fwd_kernel(const uint16_t* __restrict__ x, int ldx, int cx0,
                                uint16_t* __restrict__ y, int ldy, int cy0, int n, int H, int W,
                                int c)
{
    const int Ho = H / 2, Wo = W / 2;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * Ho * Wo * c) return;
    const int ch = (int)(i % c);
    size_t q = i / c;
    const int xo = (int)(q % Wo); q /= Wo;
    const int yo = (int)(q % Ho);
    const int im = (int)(q / Ho);
    float s = 0.f;
    for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx)
            s += bf2f(x[((size_t)(im * H + 2 * yo + dy) * W + 2 * xo + dx) * ldx + cx0 + ch]);
    y[((size_t)(im * Ho + yo) * Wo + xo) * ldy + cy0 + ch] = f2bf(0.25f * s);
}

// dx (fine) = 1/4 dy (coarse) broadcast; accumulate = add into dx
__global__ void pool_bwd_kernel(const uint16_t* __restrict__ dy, int ldd, int cd0,
                                uint16_t* __restrict__ dx, int ldx, int cx0, int n, int H, int W,
                                int c, int accumulate)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * H * W * c) return;
    const int ch = (int)(i % c);
    size_t q = i / c;
    const int xf = (int)(q % W); q /= W;
    const int yf = (int)(q % H);
    const int im = (int)(q / H);
    const int Ho = H / 2, Wo = W / 2;
    float g = 0.25f * bf2f(dy[((size_t)(im * Ho + yf / 2) * Wo + xf / 2) * ldd + cd0 + ch]);
    uint16_t* o = dx + ((size_t)(im * H + yf) * W + xf) * ldx + cx0 + ch;
    if (accumulate) g += bf2f(*o);
    *o = f2bf(g);
}

// nearest 2x upsample (fwd) and its adjoint (bwd: sum of the 2x2 block)
__global__ void up_fwd_kernel(const uint16_t* __restrict__ x, int ldx, int cx0,
                              uint16_t* __restrict__ y, int ldy, int cy0, int n, int H, int W,
                              int c)  // H, W of the coarse input
{
    const int Ho = 2 * H, Wo = 2 * W;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * Ho * Wo * c) return;
    const int ch = (int)(i % c);
    size_t q = i / c;
    const int xo = (int)(q % Wo); q /= Wo;
    const int yo = (int)(q % Ho);
    const int im = (int)(q / Ho);
    y[((size_t)(im * Ho + yo) * Wo + xo) * ldy + cy0 + ch] =
        x[((size_t)(im * H + yo / 2) * W + xo / 2) * ldx + cx0 + ch];
}

__global__ void up_bwd_kernel(const uint16_t* __restrict__ dy, int ldd, int cd0,
                              uint16_t* __restrict__ dx, int ldx, int cx0, int n, int H, int W,
                              int c)  // H, W of the coarse gradient
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * H * W * c) return;
    const int ch = (int)(i % c);
    size_t q = i / c;
    const int xc = (int)(q % W); q /= W;
    const int yc = (int)(q % H);
    const int im = (int)(q / H);
    const int Wf = 2 * W, Hf = 2 * H;
    float s = 0.f;
    for (int a = 0; a < 2; ++a)
        for (int bb = 0; bb < 2; ++bb)
            s += bf2f(dy[((size_t)(im * Hf + 2 * yc + a) * Wf + 2 * xc + bb) * ldd + cd0 + ch]);
    dx[((size_t)(im * H + yc) * W + xc) * ldx + cx0 + ch] = f2bf(s);
}

// ---- 8-channel (16-byte) forms of the elementwise kernels, used when every
// channel count, stride and offset is a multiple of 8 ----------------------
__device__ __forceinline__ U16x8 ld8(const uint16_t* p) { return *reinterpret_cast<const U16x8*>(p); }
__device__ __forceinline__ void st8(uint16_t* p, const U16x8& v) { *reinterpret_cast<U16x8*>(p) = v; }

__global__ void relu_bwd8_kernel(uint16_t* __restrict__ dy, int ldd, int cd0,
                                 const uint16_t* __restrict__ y, int ldy, int cy0, size_t npix,
                                 int c8)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix * c8) return;
    const size_t p = i / c8;
    const int ch = (int)(i % c8) * 8;
    uint16_t* d = dy + p * ldd + cd0 + ch;
    U16x8 g = ld8(d);
    const U16x8 v = ld8(y + p * ldy + cy0 + ch);
#pragma unroll
    for (int j = 0; j < 8; ++j)
        if (!(bf2f(v.v[j]) > 0.f)) g.v[j] = 0;
    st8(d, g);
}

__global__ void pool_fwd8_kernel(const uint16_t* __restrict__ x, int ldx, int cx0,
                                 uint16_t* __restrict__ y, int ldy, int cy0, int n, int H, int W,
                                 int c8)
{
    const int Ho = H / 2, Wo = W / 2;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * Ho * Wo * c8) return;
    const int ch = (int)(i % c8) * 8;
    size_t q = i / c8;
    const int xo = (int)(q % Wo); q /= Wo;
    const int yo = (int)(q % Ho);
    const int im = (int)(q / Ho);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
            const U16x8 v =
                ld8(x + ((size_t)(im * H + 2 * yo + dy) * W + 2 * xo + dx) * ldx + cx0 + ch);
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += bf2f(v.v[j]);
        }
    U16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o.v[j] = f2bf(0.25f * s[j]);
    st8(y + ((size_t)(im * Ho + yo) * Wo + xo) * ldy + cy0 + ch, o);
}

__global__ void pool_bwd8_kernel(const uint16_t* __restrict__ dy, int ldd, int cd0,
                                 uint16_t* __restrict__ dx, int ldx, int cx0, int n, int H, int W,
                                 int c8, int accumulate)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * H * W * c8) return;
    const int ch = (int)(i % c8) * 8;
    size_t q = i / c8;
    const int xf = (int)(q % W); q /= W;
    const int yf = (int)(q % H);
    const int im = (int)(q / H);
    const int Ho = H / 2, Wo = W / 2;
    const U16x8 g = ld8(dy + ((size_t)(im * Ho + yf / 2) * Wo + xf / 2) * ldd + cd0 + ch);
    uint16_t* o = dx + ((size_t)(im * H + yf) * W + xf) * ldx + cx0 + ch;
    U16x8 r;
    if (accumulate) {
        const U16x8 a = ld8(o);
#pragma unroll
        for (int j = 0; j < 8; ++j) r.v[j] = f2bf(0.25f * bf2f(g.v[j]) + bf2f(a.v[j]));
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) r.v[j] = f2bf(0.25f * bf2f(g.v[j]));
    }
    st8(o, r);
}

__global__ void up_fwd8_kernel(const uint16_t* __restrict__ x, int ldx, int cx0,
                               uint16_t* __restrict__ y, int ldy, int cy0, int n, int H, int W,
                               int c8)
{
    const int Ho = 2 * H, Wo = 2 * W;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * Ho * Wo * c8) return;
    const int ch = (int)(i % c8) * 8;
    size_t q = i / c8;
    const int xo = (int)(q % Wo); q /= Wo;
    const int yo = (int)(q % Ho);
    const int im = (int)(q / Ho);
    st8(y + ((size_t)(im * Ho + yo) * Wo + xo) * ldy + cy0 + ch,
        ld8(x + ((size_t)(im * H + yo / 2) * W + xo / 2) * ldx + cx0 + ch));
}

__global__ void up_bwd8_kernel(const uint16_t* __restrict__ dy, int ldd, int cd0,
                               uint16_t* __restrict__ dx, int ldx, int cx0, int n, int H, int W,
                               int c8)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * H * W * c8) return;
    const int ch = (int)(i % c8) * 8;
    size_t q = i / c8;
    const int xc = (int)(q % W); q /= W;
    const int yc = (int)(q % H);
    const int im = (int)(q / H);
    const int Wf = 2 * W, Hf = 2 * H;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int a = 0; a < 2; ++a)
        for (int bb = 0; bb < 2; ++bb) {
            const U16x8 v =
                ld8(dy + ((size_t)(im * Hf + 2 * yc + a) * Wf + 2 * xc + bb) * ldd + cd0 + ch);
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += bf2f(v.v[j]);
        }
    U16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o.v[j] = f2bf(s[j]);
    st8(dx + ((size_t)(im * H + yc) * W + xc) * ldx + cx0 + ch, o);
}

// y = bf16(x) for a dense float array (with output stride / offset per pixel)
__global__ void pack_bf16_kernel(const float* __restrict__ x, int c_in, uint16_t* __restrict__ y,
                                 int ldy, int cy0, size_t npix, float scale)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix * c_in) return;
    const size_t p = i / c_in;
    const int ch = (int)(i % c_in);
    y[p * ldy + cy0 + ch] = f2bf(x[i] * scale);
}

// MSE loss gradient: g = (pred - target) * gscale as bf16 into an ld-strided
// buffer (first c channels); when loss is given, the sum of squares is added
// to *loss (one atomic per block)
__global__ __launch_bounds__(256) void mse_grad_kernel(const float* __restrict__ pred, int ldp,
                                                       const float* __restrict__ target, int c,
                                                       size_t npix, uint16_t* __restrict__ g,
                                                       int ldg, float gscale,
                                                       float* __restrict__ loss)
{
    __shared__ float red[4];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float d = 0.f;
    if (i < npix * c) {
        const size_t p = i / c;
        const int ch = (int)(i % c);
        d = pred[p * ldp + ch] - target[i];
        g[p * ldg + ch] = f2bf(d * gscale);
    }
    if (!loss) return;  // uniform over the grid
    float s = d * d;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, red[0] + red[1] + red[2] + red[3]);
}

// Adam on FP32 master weights; also refreshes the bf16 copy
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v,
                            uint16_t* __restrict__ pb, size_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2, float wd)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * p[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / c1, vh = vi / c2;
    const float np = p[i] - lr * mh / (sqrtf(vh) + eps);
    p[i] = np;
    if (pb) pb[i] = f2bf(np);
}

}  // namespace

}  // namespace gr

using namespace gr;

static hipStream_t cnn_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

static unsigned nblk(size_t n, int tb) { return (unsigned)((n + tb - 1) / tb); }

// a channel slice [c0, c0 + c) of a buffer with ld channels per pixel
static bool slice_ok(int ld, int c0, int c) { return c > 0 && c0 >= 0 && c0 + c <= ld; }
// 16-byte (8-channel) vector access is aligned for this slice
static bool vec8(int c, int ld, int c0, const void* base)
{
    return c % 8 == 0 && ld % 8 == 0 && c0 % 8 == 0 && ((uintptr_t)base & 15) == 0;
}

template <int WAVES_M, int TM, int TN, int KC>
static void conv_tile_launch(const uint16_t* x, int ldx, int cx0, const uint16_t* w,
                             const float* bias, void* y, int ldy, int cy0, int n, int H, int W,
                             int cin, int cout, int taps, int relu, int out_f32, hipStream_t st)
{
    using T9 = ConvTile<WAVES_M, TM, TN, KC, 9>;
    using T1 = ConvTile<WAVES_M, TM, TN, KC, 1>;
    const long long blocks =
        (long long)n * H * ((W + T9::MP - 1) / T9::MP) * ((cout + T9::NC - 1) / T9::NC);
    if (taps == 9)
        hipLaunchKernelGGL((conv_tile_kernel<WAVES_M, TM, TN, KC, 9>), dim3((unsigned)blocks),
                           dim3(256), T9::lds_bytes, st, x, ldx, cx0, w, bias, y, ldy, cy0, H, W,
                           cin, cout, relu, out_f32);
    else
        hipLaunchKernelGGL((conv_tile_kernel<WAVES_M, TM, TN, KC, 1>), dim3((unsigned)blocks),
                           dim3(256), T1::lds_bytes, st, x, ldx, cx0, w, bias, y, ldy, cy0, H, W,
                           cin, cout, relu, out_f32);
}

extern "C" {

int gr_cnn_conv_fwd(const void* x, int ldx, int cx0, const void* w, const float* bias, void* y,
                    int ldy, int cy0, int n, int H, int W, int cin, int cout, int taps, int relu,
                    int out_f32, void* stream)
{
    if (!x || !w || !y || n <= 0 || H <= 0 || W <= 0 || cin % 16 || cout % 32 ||
        (taps != 1 && taps != 9) || ldx % 8 || cx0 % 8 || cx0 + cin > ldx || cy0 + cout > ldy) {
        set_error("cnn_conv_fwd: unsupported shape (cin %% 16, cout %% 32, 8-aligned ld)");
        return GR_ERR_INVALID;
    }
    // tile shape by row width and output channels: wide rows take 128-pixel
    // tiles, 64- and 32-pixel rows take tiles with more output channels
    const hipStream_t st = cnn_stream(stream);
    const uint16_t* xp = (const uint16_t*)x;
    const uint16_t* wp = (const uint16_t*)w;
    if (W <= 32 && cout % 256 == 0)
        conv_tile_launch<1, 1, 2, 16>(xp, ldx, cx0, wp, bias, y, ldy, cy0, n, H, W, cin, cout,
                                      taps, relu, out_f32, st);
    else if (W <= 32 && cout % 128 == 0)
        conv_tile_launch<1, 1, 1, 16>(xp, ldx, cx0, wp, bias, y, ldy, cy0, n, H, W, cin, cout,
                                      taps, relu, out_f32, st);
    else if (W <= 64 && cout % 128 == 0)
        conv_tile_launch<2, 1, 2, 16>(xp, ldx, cx0, wp, bias, y, ldy, cy0, n, H, W, cin, cout,
                                      taps, relu, out_f32, st);
    else if (W <= 64 && cout % 64 == 0)
        conv_tile_launch<2, 1, 1, 32>(xp, ldx, cx0, wp, bias, y, ldy, cy0, n, H, W, cin, cout,
                                      taps, relu, out_f32, st);
    else if (cout % 64 == 0)
        conv_tile_launch<2, 2, 1, 32>(xp, ldx, cx0, wp, bias, y, ldy, cy0, n, H, W, cin, cout,
                                      taps, relu, out_f32, st);
    else
        conv_tile_launch<4, 1, 1, 32>(xp, ldx, cx0, wp, bias, y, ldy, cy0, n, H, W, cin, cout,
                                      taps, relu, out_f32, st);
    GR_LAUNCH_CHECK("conv_tile_kernel");
    return GR_OK;
}

int gr_cnn_conv_wgrad(const void* x, int ldx, int cx0, const void* dy, int ldd, int cd0,
                      float* dw, int n, int H, int W, int cin, int cout, int taps,
                      int rows_per_chunk, void* stream)
{
    if (!x || !dy || !dw || n <= 0 || H <= 0 || W <= 0 || cin % 8 || cout % 8 || ldx % 8 ||
        cx0 % 8 || ldd % 8 || cd0 % 8 || cx0 + cin > ldx || cd0 + cout > ldd ||
        (taps != 1 && taps != 9) || rows_per_chunk <= 0) {
        set_error("cnn_conv_wgrad: invalid arguments");
        return GR_ERR_INVALID;
    }
    const int chunks = (n * H + rows_per_chunk - 1) / rows_per_chunk;
    const long long blocks = (long long)((cout + 63) / 64) * ((cin + 63) / 64) *
                             (taps == 9 ? 3 : 1) * chunks;
    if (taps == 9)
        hipLaunchKernelGGL(conv_wgrad_kernel<9>, dim3((unsigned)blocks), dim3(256), 0,
                           cnn_stream(stream), (const uint16_t*)x, ldx, cx0, (const uint16_t*)dy,
                           ldd, cd0, dw, n, H, W, cin, cout, rows_per_chunk);
    else
        hipLaunchKernelGGL(conv_wgrad_kernel<1>, dim3((unsigned)blocks), dim3(256), 0,
                           cnn_stream(stream), (const uint16_t*)x, ldx, cx0, (const uint16_t*)dy,
                           ldd, cd0, dw, n, H, W, cin, cout, rows_per_chunk);
    GR_LAUNCH_CHECK("conv_wgrad_kernel");
    return GR_OK;
}

int gr_cnn_bias_grad(const void* dy, int ldd, int cd0, size_t npix, int cout, float* db,
                     void* stream)
{
    if (!dy || !db || npix == 0 || cout <= 0 || cout > 256 || 256 % cout ||
        !slice_ok(ldd, cd0, cout)) {
        set_error("cnn_bias_grad: invalid arguments (cout must divide 256)");
        return GR_ERR_INVALID;
    }
    if (cout % 8 == 0 && vec8(cout, ldd, cd0, dy)) {
        // one block per CU: every block ends in cout atomics on the same cout
        // addresses, so fewer, longer blocks
        const size_t lanes = 256 / (cout / 8);
        const unsigned blocks = (unsigned)std::min<size_t>(256, (npix + lanes - 1) / lanes);
        hipLaunchKernelGGL(bias_grad8_kernel, dim3(blocks), dim3(256), 0, cnn_stream(stream),
                           (const uint16_t*)dy, ldd, cd0, npix, cout, db);
        GR_LAUNCH_CHECK("bias_grad8_kernel");
        return GR_OK;
    }
    const size_t lanes = 256 / cout;
    const unsigned blocks = (unsigned)std::min<size_t>(2048, (npix + lanes - 1) / lanes);
    hipLaunchKernelGGL(bias_grad_kernel, dim3(blocks), dim3(256), 0, cnn_stream(stream),
                       (const uint16_t*)dy, ldd, cd0, npix, cout, db);
    GR_LAUNCH_CHECK("bias_grad_kernel");
    return GR_OK;
}

int gr_cnn_weight_flip(const void* w, void* wt, int cin, int cout, int taps, void* stream)
{
    if (!w || !wt || cin <= 0 || cout <= 0 || (taps != 1 && taps != 9)) {
        set_error("cnn_weight_flip: invalid arguments");
        return GR_ERR_INVALID;
    }
    const size_t total = (size_t)cin * cout * taps;
    hipLaunchKernelGGL(weight_flip_kernel, dim3(nblk(total, 256)), dim3(256), 0,
                       cnn_stream(stream), (const uint16_t*)w, (uint16_t*)wt, cin, cout, taps);
    GR_LAUNCH_CHECK("weight_flip_kernel");
    return GR_OK;
}

int gr_cnn_relu_bwd(void* dy, int ldd, int cd0, const void* y, int ldy, int cy0, size_t npix,
                    int c, void* stream)
{
    if (!dy || !y || npix == 0 || !slice_ok(ldd, cd0, c) || !slice_ok(ldy, cy0, c)) {
        set_error("cnn_relu_bwd: invalid arguments");
        return GR_ERR_INVALID;
    }
    if (vec8(c, ldd, cd0, dy) && vec8(c, ldy, cy0, y))
        hipLaunchKernelGGL(relu_bwd8_kernel, dim3(nblk(npix * (c / 8), 256)), dim3(256), 0,
                           cnn_stream(stream), (uint16_t*)dy, ldd, cd0, (const uint16_t*)y, ldy,
                           cy0, npix, c / 8);
    else
        hipLaunchKernelGGL(relu_bwd_kernel, dim3(nblk(npix * c, 256)), dim3(256), 0,
                           cnn_stream(stream), (uint16_t*)dy, ldd, cd0, (const uint16_t*)y, ldy,
                           cy0, npix, c);
    GR_LAUNCH_CHECK("relu_bwd_kernel");
    return GR_OK;
}

int gr_cnn_pool(const void* x, int ldx, int cx0, void* y, int ldy, int cy0, int n, int H, int W,
                int c, int backward, int accumulate, void* stream)
{
    if (!x || !y || n <= 0 || H <= 0 || W <= 0 || H % 2 || W % 2 || !slice_ok(ldx, cx0, c) ||
        !slice_ok(ldy, cy0, c)) {
        set_error("cnn_pool: invalid arguments (even sizes, channel slices within ld)");
        return GR_ERR_INVALID;
    }
    const bool v8 = vec8(c, ldx, cx0, x) && vec8(c, ldy, cy0, y);
    if (!backward && v8)
        hipLaunchKernelGGL(pool_fwd8_kernel, dim3(nblk((size_t)n * H / 2 * W / 2 * (c / 8), 256)),
                           dim3(256), 0, cnn_stream(stream), (const uint16_t*)x, ldx, cx0,
                           (uint16_t*)y, ldy, cy0, n, H, W, c / 8);
    else if (!backward)
        hipLaunchKernelGGL(pool_fwd_kernel, dim3(nblk((size_t)n * H / 2 * W / 2 * c, 256)),
                           dim3(256), 0, cnn_stream(stream), (const uint16_t*)x, ldx, cx0,
                           (uint16_t*)y, ldy, cy0, n, H, W, c);
    else if (v8)  // x = coarse gradient, y = fine gradient (H, W fine)
        hipLaunchKernelGGL(pool_bwd8_kernel, dim3(nblk((size_t)n * H * W * (c / 8), 256)),
                           dim3(256), 0, cnn_stream(stream), (const uint16_t*)x, ldx, cx0,
                           (uint16_t*)y, ldy, cy0, n, H, W, c / 8, accumulate);
    else
        hipLaunchKernelGGL(pool_bwd_kernel, dim3(nblk((size_t)n * H * W * c, 256)), dim3(256), 0,
                           cnn_stream(stream), (const uint16_t*)x, ldx, cx0, (uint16_t*)y, ldy,
                           cy0, n, H, W, c, accumulate);
    GR_LAUNCH_CHECK("pool");
    return GR_OK;
}

int gr_cnn_upsample(const void* x, int ldx, int cx0, void* y, int ldy, int cy0, int n, int H,
                    int W, int c, int backward, void* stream)
{
    if (!x || !y || n <= 0 || H <= 0 || W <= 0 || !slice_ok(ldx, cx0, c) ||
        !slice_ok(ldy, cy0, c)) {
        set_error("cnn_upsample: invalid arguments");
        return GR_ERR_INVALID;
    }
    const bool v8 = vec8(c, ldx, cx0, x) && vec8(c, ldy, cy0, y);
    if (!backward && v8)
        hipLaunchKernelGGL(up_fwd8_kernel, dim3(nblk((size_t)n * 4 * H * W * (c / 8), 256)),
                           dim3(256), 0, cnn_stream(stream), (const uint16_t*)x, ldx, cx0,
                           (uint16_t*)y, ldy, cy0, n, H, W, c / 8);
    else if (!backward)
        hipLaunchKernelGGL(up_fwd_kernel, dim3(nblk((size_t)n * 4 * H * W * c, 256)), dim3(256),
                           0, cnn_stream(stream), (const uint16_t*)x, ldx, cx0, (uint16_t*)y, ldy,
                           cy0, n, H, W, c);
    else if (v8)  // x = fine gradient, y = coarse gradient (H, W coarse)
        hipLaunchKernelGGL(up_bwd8_kernel, dim3(nblk((size_t)n * H * W * (c / 8), 256)),
                           dim3(256), 0, cnn_stream(stream), (const uint16_t*)x, ldx, cx0,
                           (uint16_t*)y, ldy, cy0, n, H, W, c / 8);
    else
        hipLaunchKernelGGL(up_bwd_kernel, dim3(nblk((size_t)n * H * W * c, 256)), dim3(256), 0,
                           cnn_stream(stream), (const uint16_t*)x, ldx, cx0, (uint16_t*)y, ldy,
                           cy0, n, H, W, c);
    GR_LAUNCH_CHECK("upsample");
    return GR_OK;
}

int gr_cnn_pack_bf16(const float* x, int c_in, void* y, int ldy, int cy0, size_t npix,
                     float scale, void* stream)
{
    if (!x || !y || npix == 0 || !slice_ok(ldy, cy0, c_in)) {
        set_error("cnn_pack_bf16: invalid arguments");
        return GR_ERR_INVALID;
    }
    hipLaunchKernelGGL(pack_bf16_kernel, dim3(nblk(npix * c_in, 256)), dim3(256), 0,
                       cnn_stream(stream), x, c_in, (uint16_t*)y, ldy, cy0, npix, scale);
    GR_LAUNCH_CHECK("pack_bf16_kernel");
    return GR_OK;
}

int gr_cnn_mse_grad(const float* pred, int ldp, const float* target, int c, size_t npix, void* g,
                    int ldg, float gscale, float* loss, void* stream)
{
    if (!pred || !target || !g || npix == 0 || !slice_ok(ldp, 0, c) ||
        !slice_ok(ldg, 0, c)) {
        set_error("cnn_mse_grad: invalid arguments");
        return GR_ERR_INVALID;
    }
    hipLaunchKernelGGL(mse_grad_kernel, dim3(nblk(npix * c, 256)), dim3(256), 0,
                       cnn_stream(stream), pred, ldp, target, c, npix, (uint16_t*)g, ldg, gscale,
                       loss);
    GR_LAUNCH_CHECK("mse_grad_kernel");
    return GR_OK;
}

int gr_cnn_adam(float* p, const float* g, float* m, float* v, void* pb, size_t n, float lr,
                float b1, float b2, float eps, int step, float wd, void* stream)
{
    if (!p || !g || !m || !v || n == 0 || step <= 0) {
        set_error("cnn_adam: invalid arguments");
        return GR_ERR_INVALID;
    }
    const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(nblk(n, 256)), dim3(256), 0, cnn_stream(stream), p, g, m,
                       v, (uint16_t*)pb, n, lr, b1, b2, eps, c1, c2, wd);
    GR_LAUNCH_CHECK("adam_kernel");
    return GR_OK;
}

}  // extern "C"
