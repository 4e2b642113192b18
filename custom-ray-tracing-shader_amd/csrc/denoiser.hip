// denoiser.hip - U-Net denoiser trained and run by hand-written gfx950 BF16 MFMA kernels.
//
// Network (all convolutions 3x3, stride 1, zero padding, ReLU except the last):
//   level 0 (H x W)      e0a 10->32, e0b 32->32           ... d0a 96->32, d0b 32->32, out 32->3
//   level 1 (H/2)        e1a 32->64, e1b 64->64           ... d1a 128->64, d1b 64->64
//   level 2 (H/4)        e2a 64->64, e2b 64->64           ... d2a 160->64, d2b 64->64
//   level 3 (H/8)        e3a 64->96, e3b 96->96           ... d3a 224->96, d3b 96->96
//   level 4 (H/16)       ba 96->128, bb 128->128
// 2x2 average pooling down, nearest upsampling up, skip connections by concatenation.
// Input features: log(1 + radiance / (albedo + e)) 3 (albedo-demodulated), albedo 3, normal 3,
// log(1+depth) 1, zero-padded to 32 channels (e = 0.01).  The output is added to the
// demodulated log radiance (residual), mapped back with exp(x)-1 and re-modulated by
// (albedo + e); training minimises the L1 distance in that demodulated log space to the
// high-sample target.
//
// Tensors are NHWC BF16.  A concatenation is a buffer holding [upsampled | skip] channels:
// producers write their slice through a (pointer, channel stride, channel offset) view, so no
// copies happen.  Convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation:
//   forward      M = pixels, N = output channels, K = 9 taps x input channels
//   input grad   the same kernel over the output gradient with the weights transposed and
//                the taps mirrored
//   weight grad  M = output channels, N = input channels, K = pixels; operands are staged in
//                LDS transposed so each lane reads 8 consecutive pixels
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include "denoiser.h"

namespace rt {

namespace {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kB = 256;

struct View {
    bf16* p;
    int cs;   // channels per pixel in the buffer
    int co;   // channel offset of the view
    int c;    // channels of the view
};

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int e = 0; e < 8; e++) z[e] = (bf16)0.0f;
    return z;
}

// ------------------------------------------------------------------ convolution kernels
struct ConvArgs {
    View in;
    const bf16* wt;       // [9][coutPad][cinPad]
    const float* bias;    // [coutPad] or null
    View out;             // optional BF16 output (p may be null)
    float* outF;          // optional FP32 output, [pixel][outFStride], channels < outFStride
    int outFStride;
    int n, h, w;
    int cinPad, coutPad;
    int relu;
    // Optional: input channels [0, lowC) are the nearest-neighbour upsampling of `low`, a
    // (h / 2) x (w / 2) buffer (read in place, so the upsampled tensor is never written), and
    // `in` holds channels lowC.. (in channel ci - lowC of the view).  Tile kernel only.
    View low;
    int lowC;
    // Optional: the output is the input gradient of a ReLU layer's output; it is kept where
    // `omask` (that output: the same pixels and channels as `out`) is positive and zero
    // elsewhere, so the layer's backward reads it as its pre-activation gradient.
    View omask;
};

template <int NT>
__global__ __launch_bounds__(kB) void convFwdKernel(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r = lane & 15;
    const int npix = a.n * a.h * a.w;
    const int base = blockIdx.x * kB + wave * 64;
    const int cob = blockIdx.y * 16 * NT;
    int py[4], px[4], pn[4];
    bool pv[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int p = base + t * 16 + r;
        pv[t] = p < npix;
        const int pp = pv[t] ? p : 0;
        px[t] = pp % a.w;
        const int q = pp / a.w;
        py[t] = q % a.h;
        pn[t] = q / a.h;
    }
    f32x4 acc[4][NT];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int j = 0; j < NT; j++) acc[t][j] = f32x4{0, 0, 0, 0};
    for (int tap = 0; tap < 9; tap++) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const bf16* src[4];
        bool ok[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int yy = py[t] + dy, xx = px[t] + dx;
            ok[t] = pv[t] && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w;
            const int q = ok[t] ? (pn[t] * a.h + yy) * a.w + xx : 0;
            src[t] = a.in.p + (size_t)q * a.in.cs + a.in.co + 8 * g;
        }
        const bf16* wrow = a.wt + ((size_t)tap * a.coutPad + cob + r) * a.cinPad + 8 * g;
        for (int ci0 = 0; ci0 < a.cinPad; ci0 += 32) {
            bf16x8 av[4];
#pragma unroll
            for (int t = 0; t < 4; t++) av[t] = ok[t] ? *reinterpret_cast<const bf16x8*>(src[t] + ci0) : zero8();
#pragma unroll
            for (int j = 0; j < NT; j++) {
                const bf16x8 bv = *reinterpret_cast<const bf16x8*>(wrow + (size_t)16 * j * a.cinPad + ci0);
#pragma unroll
                for (int t = 0; t < 4; t++) acc[t][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[t], bv, acc[t][j], 0, 0, 0);
            }
        }
    }
    // D layout of 16x16 MFMA: lane (g, r) holds rows 4g..4g+3 (pixels), column r (channel).
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int j = 0; j < NT; j++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int p = base + t * 16 + 4 * g + i;
                const int c = cob + 16 * j + r;
                if (p >= npix) continue;
                float v = acc[t][j][i] + (a.bias ? a.bias[c] : 0.0f);
                if (a.relu) v = fmaxf(v, 0.0f);
                if (a.out.p && c < a.out.c) {
                    bf16* dst = a.out.p + (size_t)p * a.out.cs + a.out.co + c;
                    if (a.omask.p && !((float)a.omask.p[(size_t)p * a.omask.cs + a.omask.co + c] > 0.0f)) v = 0.0f;
                    *dst = (bf16)v;
                }
                if (a.outF && c < a.outFStride) a.outF[(size_t)p * a.outFStride + c] = v;
            }
}

// The same convolution for images at least 16 pixels wide (or 16384 pixels), as a
// persistent kernel: a block computes ROWS-row x 64-column output tiles of one group of
// 16 NT output channels (blockIdx.y; one row per wave), looping over tiles and 32-channel
// input slices ("steps").
// Each step's input slice plus its one-pixel halo ((ROWS + 2) x 66 pixels) is staged in LDS
// with the weights of that slice (for a layer of one slice, once before the first step) and
// read there by all nine taps.  The next step's halo (and weights) are loaded into registers
// while this step's MFMAs run and written to LDS after it, so the loads' latency hides
// behind the MFMAs.  Loads and stores are raw buffer accesses whose out-of-image offsets lie
// past the buffer's end, so the hardware returns the zero padding and drops the stores
// without branches.  The MFMA takes the weights as its A operand, so a lane's results are 4
// consecutive output channels of one pixel, stored as 8 bytes straight from registers.
// Output rows per tile (one per wavefront of the block): 8, or 4 for layers with too few
// 8-row tiles to give every compute unit a block (the 4-row block's smaller LDS lets two or
// three share one).
constexpr int kTileW = 64, kHaloW = kTileW + 2;
template <int ROWS> struct TileShape {
    static constexpr int kH = ROWS, kHaloH = ROWS + 2, kThreads = 64 * ROWS;
    static constexpr int kHaloChunks = kHaloH * kHaloW * 4;   // 16-byte chunks of one slice
    static constexpr int kHaloIters = (kHaloChunks + kThreads - 1) / kThreads;
};
constexpr unsigned kOutside = 0x7FFFFFF0u;                        // a buffer offset past any view

// LDS position (in bf16) of 16-byte chunk c of staged row `row` (a halo pixel, or one output
// channel's weights): 64 bytes per row, the chunk XOR-swizzled by bits 1-2 of `key`, which
// is the row's position modulo 8 in 128-byte LDS lines (the halo column; a halo row is 66
// pixels, 33 lines, so the halo row does not change it).  A wave's 16-byte reads then hit
// every bank once per eight lanes (eight consecutive pixels or channels), where the plain
// layout put two lanes on each bank.
__device__ __forceinline__ int swz(int row, int c, int key) { return row * 32 + 8 * (c ^ ((key >> 1) & 3)); }

__device__ __forceinline__ __amdgpu_buffer_rsrc_t bufferOf(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ bf16x8 bufferLoad16(__amdgpu_buffer_rsrc_t r, unsigned offset) {
    return __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(r, (int)offset, 0, 0));
}

typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Buffer resources of a tile kernel's outputs (raw buffer stores drop offsets past the end,
// which is how lanes outside the image skip their stores without branches).
struct TileOut {
    __amdgpu_buffer_rsrc_t out, mask, outF;
};

// Bias, ReLU and the stores of a finished tile row: lane (g, r) holds output channels
// cBase + 16 j + 4 g .. + 3 of pixel 16 q + r of its wave's row (biasOf[j]: their biases).
// Each lane's byte offsets are formed once per row; pixel group q adds a fixed stride and
// channel group j a constant.
template <int NT>
__device__ __forceinline__ void tileEpilogue(const ConvArgs& a, const TileOut& o, const f32x4 (&acc)[4][NT],
                                             const float4 (&biasOf)[NT], int cBase, int x0, int y, int nn, int g, int r) {
    if (y >= a.h) return;
    const int pix = (nn * a.h + y) * a.w + x0 + r;   // this lane's pixel in group q = 0
    const int left = a.w - x0 - r;                    // group q holds an image pixel while 16 q < left
    const int c0 = cBase + 4 * g;                     // this lane's first channel in group j = 0
    const unsigned outBase = (unsigned)(pix * a.out.cs + a.out.co + c0) * 2;
    const unsigned maskBase = (unsigned)(pix * a.omask.cs + a.omask.co + c0) * 2;
    const unsigned fBase = (unsigned)(pix * a.outFStride + c0) * 4;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const bool in = 16 * q < left;
#pragma unroll
        for (int j = 0; j < NT; j++) {
            const unsigned outOff = in && c0 + 16 * j < a.out.c ? outBase + (unsigned)(32 * q * a.out.cs + 32 * j) : kOutside;
            const unsigned maskOff = maskBase + (unsigned)(32 * q * a.omask.cs + 32 * j);
            const unsigned fOff = in && c0 + 16 * j < a.outFStride ? fBase + (unsigned)(64 * q * a.outFStride + 64 * j) : kOutside;
            const float4 bias = biasOf[j];
            float v[4] = {acc[q][j][0] + bias.x, acc[q][j][1] + bias.y, acc[q][j][2] + bias.z, acc[q][j][3] + bias.w};
            if (a.relu)
#pragma unroll
                for (int i = 0; i < 4; i++) v[i] = fmaxf(v[i], 0.0f);
            if (a.out.p) {
                if (a.omask.p) {
                    const bf16x4 m = __builtin_bit_cast(bf16x4, __builtin_amdgcn_raw_buffer_load_b64(o.mask, (int)maskOff, 0, 0));
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        if (!((float)m[i] > 0.0f)) v[i] = 0.0f;
                }
                bf16x4 h;
#pragma unroll
                for (int i = 0; i < 4; i++) h[i] = (bf16)v[i];
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, h), o.out, (int)outOff, 0, 0);
            }
            if (a.outF)
                __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, f32x4{v[0], v[1], v[2], v[3]}), o.outF,
                                                       (int)fOff, 0, 0);
        }
    }
}

// RT_CONV_WAVES_PER_EU asks the compiler for that many resident waves per SIMD (it then caps
// the registers); RT_CONV_TAP_UNROLL limits how many taps' LDS reads it schedules together.
#ifndef RT_CONV_TAP_UNROLL
#define RT_CONV_TAP_UNROLL 9
#endif
#ifdef RT_CONV_WAVES_PER_EU
#define RT_CONV_OCCUPANCY __attribute__((amdgpu_waves_per_eu(RT_CONV_WAVES_PER_EU)))
#else
#define RT_CONV_OCCUPANCY
#endif
// Resident blocks per compute unit a tile kernel is built for: 4-row blocks of up to 32
// output channels fit three (their LDS), 64-channel ones two; 8-row blocks one.
// (The multi-slice 32-channel kernel holds a slice of weights in registers too: two blocks.)
// Whether a tile kernel loads its halo straight into two LDS buffers (the 8-row kernels with
// resident weights and 32 or 64 output channels; the 16-channel output layer is faster
// staging its halo through registers).
template <int NT, int ROWS, bool MULTI> constexpr bool tileDmaHalo() { return !MULTI && ROWS == 8 && NT >= 2; }
template <int NT, int ROWS, bool MULTI> constexpr int tileBlocksPerCu() {
    return ROWS == 4 ? (NT == 4 || (NT == 2 && MULTI) ? 2 : 3) : 1;
}
// MULTI: each step stages its input slice's weights as well (layers whose weights do not fit
// in LDS beside the halo); without it all of the block's weights (every slice: dynamic LDS of
// 9 x 16 NT x cinPad BF16) are staged once, before the first step.
template <int NT, int ROWS, bool MULTI>
__global__ __launch_bounds__(64 * ROWS, (tileBlocksPerCu<NT, ROWS, MULTI>())) RT_CONV_OCCUPANCY void convFwdTileKernel(ConvArgs a, int tilesX, int tilesY, int tileCount) {
    using T = TileShape<ROWS>;
    constexpr int kTileH = T::kH, kHaloH = T::kHaloH, kTileThreads = T::kThreads;
    constexpr int kHaloChunks = T::kHaloChunks, kHaloIters = T::kHaloIters;
    constexpr int kCout = 16 * NT, kWChunks = 9 * kCout * 4, kWIters = (kWChunks + kTileThreads - 1) / kTileThreads;
    // 8-row kernels with resident weights and 32 or 64 output channels load each step's halo
    // straight into LDS (64 lanes x 16 bytes per instruction) into one of two buffers, the
    // next step's while this step's taps read the other.
    constexpr bool kDma = tileDmaHalo<NT, ROWS, MULTI>();
    constexpr int kDmaInstr = (kHaloChunks + 63) / 64, kDmaPerWave = (kDmaInstr + ROWS - 1) / ROWS;
    __shared__ __attribute__((aligned(16))) bf16 tile[kDma ? kDmaInstr * 512 : kHaloH * kHaloW * 32];
    __shared__ __attribute__((aligned(16))) bf16 tile2[kDma ? kDmaInstr * 512 : 8];
    __shared__ __attribute__((aligned(16))) bf16 sWStep[MULTI ? 9 * kCout * 32 : 8];
    extern __shared__ __attribute__((aligned(16))) bf16 sWAll[];
    __shared__ __attribute__((aligned(16))) float sBias[kCout];
    bf16* const sW = MULTI ? sWStep : sWAll;   // [slice][tap][output channel][32 input channels]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r = lane & 15;
    const int slices = a.cinPad / 32, cBase = blockIdx.y * kCout;   // this block's output channels
    const int steps = ((tileCount - (int)blockIdx.x + (int)gridDim.x - 1) / (int)gridDim.x) * slices;
    if (steps <= 0) return;
    const __amdgpu_buffer_rsrc_t inBuf = bufferOf(a.in.p, (unsigned)((size_t)a.n * a.h * a.w * a.in.cs * 2));
    const __amdgpu_buffer_rsrc_t wBuf = bufferOf(a.wt, (unsigned)((size_t)9 * a.coutPad * a.cinPad * 2));
    const __amdgpu_buffer_rsrc_t lowBuf =
        bufferOf(a.low.p ? a.low.p : a.in.p, (unsigned)((size_t)a.n * (a.h / 2) * (a.w / 2) * a.low.cs * 2));
    const size_t pixels = (size_t)a.n * a.h * a.w;
    const TileOut o{bufferOf(a.out.p, (unsigned)(pixels * a.out.cs * 2)), bufferOf(a.omask.p, (unsigned)(pixels * a.omask.cs * 2)),
                    bufferOf(a.outF, (unsigned)(pixels * a.outFStride * 4))};
    // this thread's halo chunk `it`: pixel row and column in the halo tile
    auto hRow = [&](int it) __attribute__((always_inline)) {
        const int v = threadIdx.x + it * kTileThreads;
        return v < kHaloChunks ? (v >> 2) / kHaloW : -1000;   // -1000: no chunk (always outside)
    };
    auto hCol = [&](int it) __attribute__((always_inline)) { return ((int)threadIdx.x + it * kTileThreads) / 4 % kHaloW; };
    const int part = threadIdx.x & 3;   // every chunk index of a thread has the same low two bits
    // byte offset of chunk `it` from the tile's first pixel (channel 0 of the slice), or past any
    // buffer for a thread without that chunk (24-bit products: a halo row is under 2^24 bytes)
    const unsigned rowBytes = (unsigned)(a.w * a.in.cs * 2), pixBytes = (unsigned)(a.in.cs * 2);
    auto rel = [&](int it) __attribute__((always_inline)) {
        return threadIdx.x + it * kTileThreads < kHaloChunks
                   ? __umul24((unsigned)hRow(it), rowBytes) + __umul24((unsigned)hCol(it), pixBytes) - rowBytes - pixBytes + 16u * part
                   : kOutside;
    };
    bf16x8 rin[kHaloIters], rw[kWIters];
    // loads the weights of input slice ci0 into dst
    auto fetchW = [&](int ci0, bf16x8 (&dst)[kWIters]) {
#pragma unroll
        for (int it = 0; it < kWIters; it++) {
            const int v = threadIdx.x + it * kTileThreads;
            const int row = v >> 2, tap = row / kCout, co = row % kCout;
            const unsigned off =
                v < kWChunks ? (unsigned)(((tap * a.coutPad + cBase + co) * a.cinPad) + ci0 + 8 * part) * 2 : kOutside;
            dst[it] = bufferLoad16(wBuf, off);
        }
    };
    // stores src as the weights of LDS slice `slice`
    auto storeW = [&](int slice, const bf16x8 (&src)[kWIters]) {
        bf16* const dst = sW + slice * (9 * kCout * 32);
#pragma unroll
        for (int it = 0; it < kWIters; it++) {
            const int v = threadIdx.x + it * kTileThreads;
            if (v < kWChunks) *reinterpret_cast<bf16x8*>(dst + swz(v >> 2, v & 3, v >> 2)) = src[it];
        }
    };
    // loads the halo (and for MULTI the weights) of step k into rin / rw
    auto fetch = [&](int k) __attribute__((always_inline)) {
        const int t = blockIdx.x + (k / slices) * gridDim.x, ci0 = (k % slices) * 32;
        const int x0 = (t % tilesX) * kTileW, y0 = ((t / tilesX) % tilesY) * kTileH, nn = t / (tilesX * tilesY);
        if (ci0 < a.lowC) {   // upsampled channels: the half-resolution pixel under each halo pixel
            const unsigned chan = (unsigned)(a.low.co + ci0 + 8 * part) * 2;
            const int hl = a.h / 2, wl = a.w / 2;
#pragma unroll
            for (int it = 0; it < kHaloIters; it++) {
                const int yy = y0 + hRow(it) - 1, xx = x0 + hCol(it) - 1;
                const bool in = (unsigned)yy < (unsigned)a.h && (unsigned)xx < (unsigned)a.w;
                const unsigned off =
                    in ? (unsigned)(((nn * hl + (yy >> 1)) * wl + (xx >> 1)) * a.low.cs) * 2 + chan : kOutside;
                rin[it] = bufferLoad16(lowBuf, off);
            }
        } else {
            // Tiles whose halo lies inside the image skip the per-chunk bounds test (their
            // threads without a chunk still land past the buffer: base + kOutside).
            const unsigned base = (unsigned)(((nn * a.h + y0) * a.w + x0) * a.in.cs + a.in.co + ci0 - a.lowC) * 2;
            const bool inside = x0 >= 1 && x0 + kTileW < a.w && y0 >= 1 && y0 + kTileH < a.h;
#pragma unroll
            for (int it = 0; it < kHaloIters; it++) {
                const int yy = y0 + hRow(it) - 1, xx = x0 + hCol(it) - 1;
                const bool in = inside || ((unsigned)yy < (unsigned)a.h && (unsigned)xx < (unsigned)a.w);
                rin[it] = bufferLoad16(inBuf, in ? base + rel(it) : kOutside);
            }
        }
        if (MULTI) fetchW(ci0, rw);
    };
    // kDma: the halo of step k into LDS buffer dst.  Lane l of the wave's instruction i fills
    // 16-byte slot s = 64 i + l, i.e. chunk slot s & 3 of halo pixel s / 4, so it loads the
    // channel chunk that the swizzled layout keeps there; out-of-image and surplus lanes read
    // past the buffer's end, which writes zeros.
    auto dma = [&](int k, bf16* dst) __attribute__((always_inline)) {
        const int t = blockIdx.x + (k / slices) * gridDim.x, ci0 = (k % slices) * 32;
        const int x0 = (t % tilesX) * kTileW, y0 = ((t / tilesX) % tilesY) * kTileH, nn = t / (tilesX * tilesY);
        // Byte offset of the tile's first pixel and the strides of one halo row and pixel; the
        // upsampled channels step through the half-resolution tensor (tiles start on even
        // rows and columns, so halo offset d maps to d >> 1 there).
        const bool low = ci0 < a.lowC;
        unsigned base, rowB, pixB;
        __amdgpu_buffer_rsrc_t src;
        if (low) {
            const int wl = a.w / 2, cs = a.low.cs;
            base = (unsigned)(((nn * (a.h / 2) + y0 / 2) * wl + x0 / 2) * cs + a.low.co + ci0) * 2;
            rowB = (unsigned)(wl * cs * 2); pixB = (unsigned)(cs * 2);
            src = lowBuf;
        } else {
            const int cs = a.in.cs;
            base = (unsigned)(((nn * a.h + y0) * a.w + x0) * cs + a.in.co + ci0 - a.lowC) * 2;
            rowB = (unsigned)(a.w * cs * 2); pixB = (unsigned)(cs * 2);
            src = inBuf;
        }
        const int shift = low ? 1 : 0;
#pragma unroll
        for (int n = 0; n < kDmaPerWave; n++) {
            const int i = wave + n * ROWS;
            if (i < kDmaInstr) {
                const int sl = 64 * i + lane, px = sl >> 2, hr = px / kHaloW, hc = px - hr * kHaloW;
                const int c = (sl & 3) ^ ((hc >> 1) & 3);
                const int yy = y0 + hr - 1, xx = x0 + hc - 1;
                const bool in = sl < kHaloChunks && (unsigned)yy < (unsigned)a.h && (unsigned)xx < (unsigned)a.w;
                const unsigned off = base + (unsigned)((hr - 1) >> shift) * rowB + (unsigned)((hc - 1) >> shift) * pixB + 16u * c;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(src, (__attribute__((address_space(3))) void*)(dst + i * 512), 16,
                                                         (int)(in ? off : kOutside), 0, 0, 0);
            }
        }
    };
    if (!MULTI) {   // every slice's weights, staged once (published by the first step's barrier)
        bf16x8 wa[kWIters], wb[kWIters];
        fetchW(0, wa);
        for (int slice = 0; slice < slices; slice += 2) {   // the next slice's loads overlap this slice's stores
            if (slice + 1 < slices) fetchW(32 * (slice + 1), wb);
            storeW(slice, wa);
            if (slice + 1 < slices) {
                if (slice + 2 < slices) fetchW(32 * (slice + 2), wa);
                storeW(slice + 1, wb);
            }
        }
    }
    if ((int)threadIdx.x < kCout) sBias[threadIdx.x] = a.bias ? a.bias[cBase + threadIdx.x] : 0.0f;   // read after a barrier
    f32x4 acc[4][NT];
    int px0 = 0, py0 = 0, pnn = 0;   // the tile whose results acc holds
    // The biases of this lane's output channels: registers for the DMA kernels (an LDS read
    // there would wait for the loads in flight into the other buffer), else read from LDS.
    float4 biasReg[NT];
    if (kDma)
#pragma unroll
        for (int j = 0; j < NT; j++)
            biasReg[j] = a.bias ? *reinterpret_cast<const float4*>(a.bias + cBase + 16 * j + 4 * g) : make_float4(0, 0, 0, 0);
    auto epilogue = [&]() __attribute__((always_inline)) {
        float4 biasOf[NT];
#pragma unroll
        for (int j = 0; j < NT; j++) biasOf[j] = kDma ? biasReg[j] : *reinterpret_cast<const float4*>(sBias + 16 * j + 4 * g);
        tileEpilogue<NT>(a, o, acc, biasOf, cBase, px0, py0 + wave, pnn, g, r);
    };
    // at a tile's first slice: the previous tile's results out, the accumulators cleared
    auto beginStep = [&](int k) __attribute__((always_inline)) {
        if (k % slices) return;
        if (k > 0) epilogue();
        const int t = blockIdx.x + (k / slices) * gridDim.x;
        px0 = (t % tilesX) * kTileW; py0 = ((t / tilesX) % tilesY) * kTileH; pnn = t / (tilesX * tilesY);
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int j = 0; j < NT; j++) acc[q][j] = f32x4{0, 0, 0, 0};
    };
    // the nine taps of one step over its halo in LDS
    auto taps = [&](const bf16* halo, int slice) __attribute__((always_inline)) {
        const bf16* const wSlice = sW + (MULTI ? 0 : slice * (9 * kCout * 32));
#pragma unroll RT_CONV_TAP_UNROLL
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap % 3;
            bf16x8 av[4];
#pragma unroll
            for (int q = 0; q < 4; q++)
                av[q] = *reinterpret_cast<const bf16x8*>(halo + swz((wave + dy) * kHaloW + 16 * q + r + dx, g, r + dx));
#pragma unroll
            for (int j = 0; j < NT; j++) {
                const bf16x8 wv = *reinterpret_cast<const bf16x8*>(wSlice + swz(tap * kCout + 16 * j + r, g, r));
#pragma unroll
                for (int q = 0; q < 4; q++) acc[q][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wv, av[q], acc[q][j], 0, 0, 0);
            }
        }
    };
    if constexpr (kDma) {
        // vmcnt(0) with the other counters unconstrained: this wave's loads into LDS (and its
        // stores) are complete
        constexpr int kWaitLoads = 0x0F70;
        dma(0, tile);
        __builtin_amdgcn_s_waitcnt(kWaitLoads);
        __syncthreads();
        // one barrier per step: the next step's halo goes into the other buffer, whose last
        // readers finished before the previous step's barrier
        auto step = [&](int k, const bf16* cur, bf16* nxt) __attribute__((always_inline)) {
            beginStep(k);
            if (k + 1 < steps) dma(k + 1, nxt);
            taps(cur, k % slices);
            __builtin_amdgcn_s_waitcnt(kWaitLoads);
            __syncthreads();
        };
        for (int k = 0; k < steps; k += 2) {
            step(k, tile, tile2);
            if (k + 1 < steps) step(k + 1, tile2, tile);
        }
    } else {
        fetch(0);
        for (int k = 0; k < steps; k++) {
            __syncthreads();   // the previous step's reads of tile / sW are done
#pragma unroll
            for (int it = 0; it < kHaloIters; it++) {
                const int v = threadIdx.x + it * kTileThreads;
                if (v < kHaloChunks) *reinterpret_cast<bf16x8*>(tile + swz(v >> 2, v & 3, hCol(it))) = rin[it];
            }
            if (MULTI) storeW(0, rw);
            __syncthreads();
            // The previous tile's stores go out here, after this step's wait for its loads and
            // before the next step's loads: the wait at the next step (which counts stores too)
            // then finds both done behind this step's MFMAs.
            beginStep(k);
            if (k + 1 < steps) fetch(k + 1);
            taps(tile, k % slices);
        }
    }
    epilogue();
}

struct BwdWArgs {
    View dy;       // gradient w.r.t. the pre-activation output, coutPad channels
    View in;       // layer input, cinPad channels
    float* part;   // [9][coutPad][cinPad] weight gradient, accumulated with atomics
    int n, h, w;
    int cinPad, coutPad;
    int pixPerPart;
};

template <int TCO, int TCI>
__global__ __launch_bounds__(kB) void convBwdWKernel(BwdWArgs a) {
    __shared__ __attribute__((aligned(16))) bf16 sA[4][16 * TCO * 32];
    __shared__ __attribute__((aligned(16))) bf16 sB[4][16 * TCI * 32];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r = lane & 15;
    const int tap = blockIdx.y, dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int ncib = a.cinPad / (16 * TCI);
    const int cob = (blockIdx.z / ncib) * 16 * TCO, cib = (blockIdx.z % ncib) * 16 * TCI;
    const int npix = a.n * a.h * a.w;
    const int begin = blockIdx.x * a.pixPerPart;
    const int end = min(npix, begin + a.pixPerPart);
    const int steps = (a.pixPerPart + 127) / 128;   // each step: 4 waves x 32 pixels
    f32x4 acc[TCO][TCI];
#pragma unroll
    for (int i = 0; i < TCO; i++)
#pragma unroll
        for (int j = 0; j < TCI; j++) acc[i][j] = f32x4{0, 0, 0, 0};
    // A wave's 32 pixels per step: this lane's 16-byte chunks of dy (TCO of them) and of the
    // shifted input (TCI), loaded one step ahead into registers.
    auto fetch = [&](int s, bf16x8 (&va)[TCO], bf16x8 (&vb)[TCI]) {
        const int p0 = begin + (s * 4 + wave) * 32;
#pragma unroll
        for (int k = 0; k < TCO; k++) {
            const int idx = lane + 64 * k, pxl = idx / (2 * TCO), cg = idx % (2 * TCO), p = p0 + pxl;
            va[k] = p < end ? *reinterpret_cast<const bf16x8*>(a.dy.p + (size_t)p * a.dy.cs + a.dy.co + cob + 8 * cg)
                            : zero8();
        }
#pragma unroll
        for (int k = 0; k < TCI; k++) {
            const int idx = lane + 64 * k, pxl = idx / (2 * TCI), cg = idx % (2 * TCI), p = p0 + pxl;
            bf16x8 v = zero8();
            if (p < end) {
                const int x = p % a.w, q = p / a.w, y = q % a.h, nn = q / a.h;
                const int yy = y + dy, xx = x + dx;
                if (yy >= 0 && yy < a.h && xx >= 0 && xx < a.w)
                    v = *reinterpret_cast<const bf16x8*>(a.in.p + (size_t)((nn * a.h + yy) * a.w + xx) * a.in.cs + a.in.co +
                                                         cib + 8 * cg);
            }
            vb[k] = v;
        }
    };
    bf16x8 va[TCO], vb[TCI];
    fetch(0, va, vb);
    // Each wave stages and reads only its own LDS region, so the waves synchronise with
    // themselves (ordering their LDS writes and reads), not with the workgroup.
    for (int s = 0; s < steps; s++) {
#pragma unroll
        for (int k = 0; k < TCO; k++) {
            const int idx = lane + 64 * k, pxl = idx / (2 * TCO), cg = idx % (2 * TCO);
#pragma unroll
            for (int e = 0; e < 8; e++) sA[wave][(8 * cg + e) * 32 + pxl] = va[k][e];
        }
#pragma unroll
        for (int k = 0; k < TCI; k++) {
            const int idx = lane + 64 * k, pxl = idx / (2 * TCI), cg = idx % (2 * TCI);
#pragma unroll
            for (int e = 0; e < 8; e++) sB[wave][(8 * cg + e) * 32 + pxl] = vb[k][e];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (s + 1 < steps) fetch(s + 1, va, vb);
#pragma unroll
        for (int i = 0; i < TCO; i++) {
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(&sA[wave][(16 * i + r) * 32 + 8 * g]);
#pragma unroll
            for (int j = 0; j < TCI; j++) {
                const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&sB[wave][(16 * j + r) * 32 + 8 * g]);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[i][j], 0, 0, 0);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    float* part = a.part + (size_t)tap * a.coutPad * a.cinPad;   // the layer's gradient, zeroed before backward
#pragma unroll
    for (int i = 0; i < TCO; i++)
#pragma unroll
        for (int j = 0; j < TCI; j++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int co = cob + 16 * i + 4 * g + e, ci = cib + 16 * j + r;
                atomicAdd(&part[(size_t)co * a.cinPad + ci], acc[i][j][e]);
            }
}

// Weight and bias gradients as one persistent kernel over tiles of pixels, all nine taps per
// block (the per-tap kernel above stages every operand nine times).
//
// A block of 4 waves handles one pair of 32-channel groups (32 output x 32 input channels;
// blockIdx.y) and loops over tiles of TH x TW dy pixels (8 x 64 for S = 2 strips of 32
// columns, 16 x 32 for S = 1), accumulating in registers.  Each tile's dy pixels and the input
// pixels under them with a one-pixel halo are staged in LDS in NHWC order (64 bytes per pixel,
// chunk order swizzled by wgByte).  The MFMA operands need 8 consecutive pixels of one channel
// per lane; ds_read_b64_tr_b16 gathers them from that pixel-major layout (lane i of a 16-lane
// group receives element i % 4 of the 8-byte words that lanes 4e + i / 4 name, e = 0..3), so
// a tap's pixel shift is only an address offset.
// A wave owns a 32-column strip of 4 dy rows: it walks the 6 input rows under them, loads each
// input row once per column shift and uses it for the (up to) three dy rows that reach it,
// so one dy row's operands serve all nine taps.  The bias gradient is one more MFMA per dy row
// against ones.  At the end the 4 waves' sums are added through LDS and the block writes them
// to its slot of `part`; wgReduceKernel (one layer) or wgReduceAllKernel (every layer of a
// backward pass) sums the slots in runs and adds each run's sum to the gradient with one
// atomic.
constexpr int kWgThreads = 256, kWgSlot = 9 * 32 * 32 + 32;   // floats per block slot: weights + bias

struct BwdWTileArgs {
    View dy, in;
    float* part;   // [pairs][gridDim.x][kWgSlot]
    int n, h, w;
    int nci;       // input-channel groups of 32
    int tilesX, tilesY, tileCount;
};

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) s16x4 LdsS16x4;

// Position of 16-byte chunk c of pixel p in its LDS tile: pixels 2 and 3 of every 4 swap the
// 32-byte halves of their 64 bytes, so four consecutive pixels put a 16-channel range in four
// different 32-byte bank groups of a 128-byte window.  Tile rows are padded to a multiple of 4
// pixels, so p & 3 is the column & 3.
__device__ __forceinline__ int wgByte(int p, int chunk) { return p * 64 + ((chunk ^ (p & 2)) << 4); }

typedef __attribute__((address_space(3))) char LdsChar;
// A transposing read of 4 pixels x 16 channels: lane jj of a 16-lane group names pixel
// col0 + shift + (jj >> 2) and channels n0 + 4 (jj & 3) .. + 3, and gets channel n0 + jj of
// the four pixels.  Returns the byte offset of the lane's 8 bytes from the start of the tile
// row (col0 a multiple of 4; the next four pixels are 256 bytes further).
__device__ __forceinline__ int wgLaneOffset(int col0, int shift, int n0, int jj) {
    const int col = col0 + shift + (jj >> 2);
    return wgByte(col, (n0 >> 3) + ((jj >> 1) & 1)) + ((jj & 1) << 3);
}
__device__ __forceinline__ s16x4 wgTr(const char* p) {
    return __builtin_amdgcn_ds_read_tr16_b64_v4i16((LdsS16x4*)(LdsChar*)p);
}
__device__ __forceinline__ bf16x8 wgPack(s16x4 lo, s16x4 hi) {
    typedef short s16x8 __attribute__((ext_vector_type(8)));
    return __builtin_bit_cast(bf16x8, (s16x8)__builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
}

// WG_IN_PARTS splits the input tile's loads into that many batches (fewer registers in
// flight while staging).
#ifndef WG_IN_PARTS
#define WG_IN_PARTS 1
#endif

template <int S>
__global__ __launch_bounds__(kWgThreads, 2) void convBwdWTileKernel(BwdWTileArgs a) {
    constexpr int RG = 4 / S, TH = 4 * RG, TW = 32 * S, HH = TH + 2, HW = TW + 2, HWP = (HW + 3) / 4 * 4;
    constexpr int kDyBytes = TH * TW * 64, kInPix = HH * HWP;
    constexpr int kDyIters = TH * TW * 4 / kWgThreads, kInIters = (kInPix * 4 + kWgThreads - 1) / kWgThreads;
    constexpr int kInParts = WG_IN_PARTS;   // input-tile loads in flight at once: kInIters / kInParts
    static_assert(kDyBytes + kInPix * 64 >= kWgSlot * 4, "the block's sums are added in the tile buffer");
    __shared__ __attribute__((aligned(16))) char lds[kDyBytes + kInPix * 64];
    char* const sDy = lds;
    char* const sIn = lds + kDyBytes;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, jj = lane & 15;
    const int strip = wave % S, rg = wave / S;
    const int pair = blockIdx.y, cob = (pair / a.nci) * 32, cib = (pair % a.nci) * 32;
    const bool bias = pair % a.nci == 0;
    const __amdgpu_buffer_rsrc_t dyBuf = bufferOf(a.dy.p, (unsigned)((size_t)a.n * a.h * a.w * a.dy.cs * 2));
    const __amdgpu_buffer_rsrc_t inBuf = bufferOf(a.in.p, (unsigned)((size_t)a.n * a.h * a.w * a.in.cs * 2));
    // This lane's operand offsets within a tile row: dy (channel group i, pixels +0 / +4) and
    // input (column shift dx, channel group j, pixels +0 / +4).
    const int col0 = 32 * strip + 8 * g;
    const char* baseA = sDy + 4 * rg * (TW * 64);
    const char* baseB = sIn + 4 * rg * (HWP * 64);
    const char* opA[2];
    const char* opB[3][2];
#pragma unroll
    for (int i = 0; i < 2; i++) opA[i] = baseA + wgLaneOffset(col0, 0, 16 * i, jj);
#pragma unroll
    for (int dx = 0; dx < 3; dx++)
#pragma unroll
        for (int j = 0; j < 2; j++) opB[dx][j] = baseB + wgLaneOffset(col0, dx, 16 * j, jj);
    f32x4 acc[9][2][2], accB[2];
#pragma unroll
    for (int t = 0; t < 9; t++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[t][i][j] = f32x4{0, 0, 0, 0};
    accB[0] = accB[1] = f32x4{0, 0, 0, 0};
    bf16x8 ones;
#pragma unroll
    for (int e = 0; e < 8; e++) ones[e] = (bf16)1.0f;
    for (int t = blockIdx.x; t < a.tileCount; t += gridDim.x) {
        const int x0 = (t % a.tilesX) * TW, y0 = ((t / a.tilesX) % a.tilesY) * TH, nn = t / (a.tilesX * a.tilesY);
        __syncthreads();   // the previous tile's reads are done
        {
            bf16x8 v[kDyIters];
#pragma unroll
            for (int k = 0; k < kDyIters; k++) {
                const int q = threadIdx.x + k * kWgThreads, p = q >> 2, c = q & 3;
                const int y = y0 + p / TW, x = x0 + p % TW;
                const bool in = y < a.h && x < a.w;
                v[k] = bufferLoad16(dyBuf, in ? (unsigned)(((nn * a.h + y) * a.w + x) * a.dy.cs + a.dy.co + cob + 8 * c) * 2
                                              : kOutside);
            }
#pragma unroll
            for (int k = 0; k < kDyIters; k++) {
                const int q = threadIdx.x + k * kWgThreads;
                *reinterpret_cast<bf16x8*>(sDy + wgByte(q >> 2, q & 3)) = v[k];
            }
        }
#pragma unroll
        for (int half = 0; half < kInParts; half++) {
            constexpr int kHalf = (kInIters + kInParts - 1) / kInParts;
            bf16x8 v[kHalf];
#pragma unroll
            for (int k = 0; k < kHalf; k++) {
                const int q = threadIdx.x + (half * kHalf + k) * kWgThreads, p = q >> 2, c = q & 3;
                const int row = p / HWP, col = p % HWP, y = y0 - 1 + row, x = x0 - 1 + col;
                const bool in = q < kInPix * 4 && col < HW && (unsigned)y < (unsigned)a.h && (unsigned)x < (unsigned)a.w;
                v[k] = bufferLoad16(inBuf, in ? (unsigned)(((nn * a.h + y) * a.w + x) * a.in.cs + a.in.co + cib + 8 * c) * 2
                                              : kOutside);
            }
#pragma unroll
            for (int k = 0; k < kHalf; k++) {
                const int q = threadIdx.x + (half * kHalf + k) * kWgThreads;
                if (q < kInPix * 4) *reinterpret_cast<bf16x8*>(sIn + wgByte(q >> 2, q & 3)) = v[k];
            }
        }
        __syncthreads();
        // dy row 4 rg + delta of the tile meets its tap row ddy (0: the row above) in input
        // (halo) row 4 rg + delta + ddy = 4 rg + rho.
        bf16x8 A[4][2];
#pragma unroll
        for (int rho = 0; rho < 6; rho++) {
            if (rho < 4) {
#pragma unroll
                for (int i = 0; i < 2; i++)
                    A[rho][i] = wgPack(wgTr(opA[i] + rho * (TW * 64)), wgTr(opA[i] + rho * (TW * 64) + 256));
                if (bias) {
#pragma unroll
                    for (int i = 0; i < 2; i++) accB[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[rho][i], ones, accB[i], 0, 0, 0);
                }
            }
#pragma unroll
            for (int dx = 0; dx < 3; dx++) {
                bf16x8 B[2];
#pragma unroll
                for (int j = 0; j < 2; j++)
                    B[j] = wgPack(wgTr(opB[dx][j] + rho * (HWP * 64)), wgTr(opB[dx][j] + rho * (HWP * 64) + 256));
#pragma unroll
                for (int ddy = 0; ddy < 3; ddy++) {
                    const int delta = rho - ddy;
                    if (delta < 0 || delta > 3) continue;
#pragma unroll
                    for (int i = 0; i < 2; i++)
#pragma unroll
                        for (int j = 0; j < 2; j++)
                            acc[ddy * 3 + dx][i][j] =
                                __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[delta][i], B[j], acc[ddy * 3 + dx][i][j], 0, 0, 0);
                }
            }
        }
    }
    // Sum the four waves' partial gradients: waves 2 and 3 park theirs in LDS for waves 0 and 1,
    // then wave 1 parks the pair's sum for wave 0, which writes the block's slot.  (Lane-major
    // parking keeps the LDS stores and loads free of bank conflicts.)
    __syncthreads();
    float* park = reinterpret_cast<float*>(lds);          // [2 regions][144 values][64 lanes]
    float* biasPark = park + 2 * 144 * 64;                // [4 waves][32 channels]
    static_assert((2 * 144 * 64 + 4 * 32) * 4 <= kDyBytes + kInPix * 64, "parking fits in the tile buffer");
    auto parkAcc = [&](int region) {
#pragma unroll
        for (int t = 0; t < 9; t++)
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
#pragma unroll
                    for (int e = 0; e < 4; e++) park[(region * 144 + ((t * 2 + i) * 2 + j) * 4 + e) * 64 + lane] = acc[t][i][j][e];
    };
    auto takeAcc = [&](int region) {
#pragma unroll
        for (int t = 0; t < 9; t++)
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
#pragma unroll
                    for (int e = 0; e < 4; e++) acc[t][i][j][e] += park[(region * 144 + ((t * 2 + i) * 2 + j) * 4 + e) * 64 + lane];
    };
    if (jj == 0)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int e = 0; e < 4; e++) biasPark[wave * 32 + 16 * i + 4 * g + e] = accB[i][e];
    if (wave >= 2) parkAcc(wave - 2);
    __syncthreads();
    if (wave < 2) takeAcc(wave);
    __syncthreads();
    if (wave == 1) parkAcc(0);
    __syncthreads();
    if (wave != 0) return;
    takeAcc(0);
    float* slot = a.part + ((size_t)pair * gridDim.x + blockIdx.x) * kWgSlot;
#pragma unroll
    for (int t = 0; t < 9; t++)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int e = 0; e < 4; e++) slot[(t * 32 + 16 * i + 4 * g + e) * 32 + 16 * j + jj] = acc[t][i][j][e];
    if (bias && lane < 32)
        slot[9 * 1024 + lane] = biasPark[lane] + biasPark[32 + lane] + biasPark[64 + lane] + biasPark[96 + lane];
}

// dW[tap][co][ci] += the sum over a pair's blocks of their slots (and dB for the pairs of
// input-channel group 0).  Run `run` takes kWgReduceRun consecutive blocks of every pair, and
// their sum is added with one atomic.
constexpr int kWgReduceRun = 16;
__device__ __forceinline__ void wgReduceRun(const float* part, int blocksPerPair, int pairs, int nci, int cinPad,
                                            int coutPad, float* dW, float* dB, long long idx, int run) {
    if (idx >= (long long)pairs * kWgSlot) return;
    const int pair = (int)(idx / kWgSlot), k = (int)(idx % kWgSlot);
    const int cob = (pair / nci) * 32, cib = (pair % nci) * 32;
    if (k >= 9 * 1024 && cib != 0) return;
    const int b0 = run * kWgReduceRun, b1 = min(blocksPerPair, b0 + kWgReduceRun);
    const float* src = part + (size_t)pair * blocksPerPair * kWgSlot + k;
    float sum = 0.0f;
    for (int b = b0; b < b1; b++) sum += src[(size_t)b * kWgSlot];
    if (k < 9 * 1024) {
        const int tap = k >> 10, co = (k >> 5) & 31, ci = k & 31;
        atomicAdd(&dW[((size_t)tap * coutPad + cob + co) * cinPad + cib + ci], sum);
    } else {
        atomicAdd(&dB[cob + k - 9 * 1024], sum);
    }
}
__global__ void wgReduceKernel(const float* part, int blocksPerPair, int pairs, int nci, int cinPad, int coutPad,
                               float* dW, float* dB) {
    wgReduceRun(part, blocksPerPair, pairs, nci, cinPad, coutPad, dW, dB, blockIdx.x * (long long)kB + threadIdx.x,
                blockIdx.y);
}
// The reductions of every layer of a backward pass in one launch: blocks firstBlock.. of the
// grid belong to a layer, xBlocks of them per run.
constexpr int kMaxWgLayers = 24;
struct WgReduceLayer {
    const float* part;
    float *dW, *dB;
    int per, pairs, nci, cinPad, coutPad, xBlocks, firstBlock;
};
struct WgReduceArgs {
    WgReduceLayer layer[kMaxWgLayers];
    int count;
};
__global__ void wgReduceAllKernel(WgReduceArgs a) {
    int l = 0;
    while (l + 1 < a.count && (int)blockIdx.x >= a.layer[l + 1].firstBlock) l++;
    const WgReduceLayer& L = a.layer[l];
    const int local = blockIdx.x - L.firstBlock;
    wgReduceRun(L.part, L.per, L.pairs, L.nci, L.cinPad, L.coutPad, L.dW, L.dB,
                (local % L.xBlocks) * (long long)kB + threadIdx.x, local / L.xBlocks);
}

// ------------------------------------------------------------------ elementwise kernels
// Pooling and upsampling move 8 channels (16 bytes) per thread: every view's channel
// stride, offset and width are multiples of 32, so each 8-channel group is 16-byte aligned.
__device__ __forceinline__ bf16x8 load8(const View& v, size_t pixel, int c) {
    return *reinterpret_cast<const bf16x8*>(v.p + pixel * v.cs + v.co + c);
}
__device__ __forceinline__ void store8(const View& v, size_t pixel, int c, bf16x8 x) {
    *reinterpret_cast<bf16x8*>(v.p + pixel * v.cs + v.co + c) = x;
}

// dBias[c] += sum over a part of the pixels of dy[p][c] (channels a multiple of 8, at most
// 8 kB): each thread sums 8 channels over a strided share of the part's pixels, the block's
// threads are combined in LDS, and each block adds once per channel.
__global__ __launch_bounds__(kB) void biasGradKernel(View dy, int npix, int channels, int pixPerPart, float* db) {
    __shared__ float red[kB * 8];
    const int groups = channels / 8, lanesPer = kB / groups;
    const int cg = threadIdx.x % groups, k = threadIdx.x / groups;
    const int begin = blockIdx.x * pixPerPart, end = min(npix, begin + pixPerPart);
    float sum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (k < lanesPer)
        for (int p = begin + k; p < end; p += lanesPer) {
            const bf16x8 v = load8(dy, (size_t)p, 8 * cg);
#pragma unroll
            for (int e = 0; e < 8; e++) sum[e] += (float)v[e];
        }
#pragma unroll
    for (int e = 0; e < 8; e++) red[threadIdx.x * 8 + e] = sum[e];
    __syncthreads();
    if ((int)threadIdx.x < channels) {
        const int c = threadIdx.x;
        float t = 0;
        for (int j = 0; j < lanesPer; j++) t += red[(j * groups + c / 8) * 8 + c % 8];
        atomicAdd(&db[c], t);
    }
}

__global__ void avgPoolKernel(View in, View out, int n, int h, int w, int channels) {
    const int ho = h / 2, wo = w / 2, groups = channels / 8;
    const long long i = blockIdx.x * (long long)kB + threadIdx.x;
    if (i >= (long long)n * ho * wo * groups) return;
    const int c = 8 * (int)(i % groups);
    const long long po = i / groups;
    const int x = (int)(po % wo), y = (int)((po / wo) % ho), nn = (int)(po / ((long long)wo * ho));
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            const bf16x8 v = load8(in, ((size_t)nn * h + 2 * y + a) * w + 2 * x + b, c);
#pragma unroll
            for (int e = 0; e < 8; e++) s[e] += (float)v[e];
        }
    bf16x8 o;
#pragma unroll
    for (int e = 0; e < 8; e++) o[e] = (bf16)(0.25f * s[e]);
    store8(out, (size_t)po, c, o);
}

// Adds the pooling gradient to the input-side gradient view.
// Adds the pooling gradient to the input-side gradient view where the pooled layer's output
// (`mask`: a ReLU output, the same pixels and channels as gin) is positive.
__global__ void avgPoolBackKernel(View gout, View gin, View mask, int n, int h, int w, int channels) {
    const int groups = channels / 8;
    const long long i = blockIdx.x * (long long)kB + threadIdx.x;
    if (i >= (long long)n * h * w * groups) return;
    const int c = 8 * (int)(i % groups);
    const long long p = i / groups;
    const int x = (int)(p % w), y = (int)((p / w) % h), nn = (int)(p / ((long long)w * h));
    const size_t po = ((size_t)nn * (h / 2) + y / 2) * (w / 2) + x / 2;
    const bf16x8 g = load8(gout, po, c), m = load8(mask, (size_t)p, c);
    bf16x8 d = load8(gin, (size_t)p, c);
#pragma unroll
    for (int e = 0; e < 8; e++) d[e] = (float)m[e] > 0.0f ? (bf16)((float)d[e] + 0.25f * (float)g[e]) : d[e];
    store8(gin, (size_t)p, c, d);
}

__global__ void upsampleKernel(View in, View out, int n, int h, int w, int channels) {   // h, w: output size
    const int groups = channels / 8;
    const long long i = blockIdx.x * (long long)kB + threadIdx.x;
    if (i >= (long long)n * h * w * groups) return;
    const int c = 8 * (int)(i % groups);
    const long long p = i / groups;
    const int x = (int)(p % w), y = (int)((p / w) % h), nn = (int)(p / ((long long)w * h));
    store8(out, (size_t)p, c, load8(in, ((size_t)nn * (h / 2) + y / 2) * (w / 2) + x / 2, c));
}

__global__ void upsampleBackKernel(View gout, View gin, int n, int h, int w, int channels) {   // h, w: input size
    const int groups = channels / 8;
    const long long i = blockIdx.x * (long long)kB + threadIdx.x;
    if (i >= (long long)n * h * w * groups) return;
    const int c = 8 * (int)(i % groups);
    const long long p = i / groups;
    const int x = (int)(p % w), y = (int)((p / w) % h), nn = (int)(p / ((long long)w * h));
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            const bf16x8 v = load8(gout, ((size_t)nn * 2 * h + 2 * y + a) * (2 * w) + 2 * x + b, c);
#pragma unroll
            for (int e = 0; e < 8; e++) s[e] += (float)v[e];
        }
    bf16x8 o;
#pragma unroll
    for (int e = 0; e < 8; e++) o[e] = (bf16)s[e];
    store8(gin, (size_t)p, c, o);
}

constexpr float kDemodEps = 0.01f;

// Raw planes [p][10] = radiance, albedo, normal, depth -> BF16 features [p][32], the
// demodulated log radiance the residual is added to, and the modulation factors.
// Each pixel's outputs go out as whole 16-byte vectors.
__device__ __forceinline__ float4 writeFeatures(const float* rawp, bf16* x0, float* noisyLog, float* demod, size_t p) {
    float raw[kDenoiseInputChannels], f[kDenoiseInputChannels], m[3];
    const float2* r2 = reinterpret_cast<const float2*>(rawp);   // 40 bytes per pixel: five 8-byte loads
#pragma unroll
    for (int k = 0; k < kDenoiseInputChannels / 2; k++) {
        const float2 v = r2[k];
        raw[2 * k] = v.x;
        raw[2 * k + 1] = v.y;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        m[c] = fmaxf(raw[3 + c], 0.0f) + kDemodEps;
        f[c] = log1pf(fmaxf(raw[c], 0.0f) / m[c]);
    }
#pragma unroll
    for (int c = 3; c < 9; c++) f[c] = raw[c];
    f[9] = log1pf(fmaxf(raw[9], 0.0f));
    bf16x8 lo, hi = zero8();
#pragma unroll
    for (int c = 0; c < 8; c++) lo[c] = (bf16)f[c];
    hi[0] = (bf16)f[8];
    hi[1] = (bf16)f[9];
    bf16x8* dst = reinterpret_cast<bf16x8*>(x0 + p * 32);
    dst[0] = lo; dst[1] = hi; dst[2] = zero8(); dst[3] = zero8();
    const float4 dm = make_float4(m[0], m[1], m[2], 0.0f);
    *reinterpret_cast<float4*>(demod + p * 4) = dm;
    *reinterpret_cast<float4*>(noisyLog + p * 4) = make_float4(f[0], f[1], f[2], 0.0f);
    return dm;
}

__global__ void featuresKernel(const float* raw, int npix, bf16* x0, float* noisyLog, float* demod) {
    const int p = blockIdx.x * kB + threadIdx.x;
    if (p >= npix) return;
    writeFeatures(raw + (size_t)p * kDenoiseInputChannels, x0, noisyLog, demod, p);
}

struct Crop { int image, ox, oy, flip; };

__global__ void cropKernel(const float* dsIn, const float* dsTgt, int dsH, int dsW, const Crop* crops, int crop,
                           int batch, bf16* x0, float* noisyLog, float* demod, float* targetLog) {
    const long long i = blockIdx.x * (long long)kB + threadIdx.x;
    if (i >= (long long)batch * crop * crop) return;
    const int x = (int)(i % crop), y = (int)((i / crop) % crop), b = (int)(i / ((long long)crop * crop));
    const Crop cr = crops[b];
    const int sx = cr.ox + ((cr.flip & 1) ? crop - 1 - x : x);
    const int sy = cr.oy + ((cr.flip & 2) ? crop - 1 - y : y);
    const size_t sp = ((size_t)cr.image * dsH + sy) * dsW + sx;
    const float4 m = writeFeatures(dsIn + sp * kDenoiseInputChannels, x0, noisyLog, demod, (size_t)i);
    const float* t = dsTgt + sp * 3;
    *reinterpret_cast<float4*>(targetLog + i * 4) = make_float4(
        log1pf(fmaxf(t[0], 0.0f) / m.x), log1pf(fmaxf(t[1], 0.0f) / m.y), log1pf(fmaxf(t[2], 0.0f) / m.z), 0.0f);
}

// Display transform of Native/renderer.cu:86-90 at exposure 1 without the 8-bit rounding:
// t = 1 - exp(-L), then the sRGB curve; and its derivative with respect to L.
__device__ __forceinline__ float displayValue(float L, float& slope) {
    const float e = expf(-L), t = 1.0f - e;
    if (t <= 0.0031308f) { slope = 12.92f * e; return 12.92f * t; }
    const float pw = powf(t, 1.0f / 2.4f);
    slope = 1.055f / 2.4f * pw / t * e;
    return 1.055f * pw - 0.055f;
}

// Loss on the residual prediction y = noisyLog + out (demodulated log radiance):
// mean over pixels and RGB of |y - targetLog| + displayWeight * |D(L(y)) - D(L(target))|,
// where L(y) = max(expm1(y), 0) * demod is the radiance the network outputs and D the
// display transform.  Writes the gradient w.r.t. out (BF16, 32 channels).
__global__ void lossKernel(const float* out, const float* noisyLog, const float* targetLog, const float* demod,
                           float displayWeight, int npix, bf16* gout, float* loss) {
    __shared__ float red[kB];
    const int p = blockIdx.x * kB + threadIdx.x;
    float l = 0;
    if (p < npix) {
        const float scale = 1.0f / (3.0f * npix);
        const float4 n4 = *reinterpret_cast<const float4*>(noisyLog + (size_t)p * 4);
        const float4 o4 = *reinterpret_cast<const float4*>(out + (size_t)p * 4);
        const float4 t4 = *reinterpret_cast<const float4*>(targetLog + (size_t)p * 4);
        const float4 m4 = *reinterpret_cast<const float4*>(demod + (size_t)p * 4);
        const float ny[3] = {n4.x, n4.y, n4.z}, oy[3] = {o4.x, o4.y, o4.z}, ty[3] = {t4.x, t4.y, t4.z};
        const float my[3] = {m4.x, m4.y, m4.z};
        bf16x8 g = zero8();
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float y = ny[c] + oy[c];
            const float tl = ty[c];
            const float d = y - tl;
            l += fabsf(d);
            float gr = d > 0 ? scale : (d < 0 ? -scale : 0.0f);
            if (displayWeight > 0) {
                const float m = my[c];
                const float em = expm1f(y);
                float sp, st;
                const float dp = displayValue(fmaxf(em, 0.0f) * m, sp);
                const float dt = displayValue(fmaxf(expm1f(tl), 0.0f) * m, st);
                const float dd = dp - dt;
                l += displayWeight * fabsf(dd);
                if (em > 0 && dd != 0) gr += (dd > 0 ? scale : -scale) * displayWeight * sp * (em + 1.0f) * m;
            }
            g[c] = (bf16)gr;
        }
        bf16x8* dst = reinterpret_cast<bf16x8*>(gout + (size_t)p * 32);
        dst[0] = g; dst[1] = zero8(); dst[2] = zero8(); dst[3] = zero8();
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] / (3.0f * npix));
}

__global__ void finishOutputKernel(const float* out, const float* noisyLog, const float* demod, int npix,
                                   float* radiance) {
    const int p = blockIdx.x * kB + threadIdx.x;
    if (p >= npix) return;
    for (int c = 0; c < 3; c++)
        radiance[(size_t)p * 4 + c] =
            fmaxf(expm1f(noisyLog[(size_t)p * 4 + c] + out[(size_t)p * 4 + c]), 0.0f) * demod[(size_t)p * 4 + c];
    radiance[(size_t)p * 4 + 3] = 0.0f;
}

__global__ void adamKernel(float* w, const float* g, float* m, float* v, long long count, float lr, float b1, float b2,
                           float eps, float c1, float c2) {
    const long long i = blockIdx.x * (long long)kB + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1 - b1) * gi, vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi; v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// FP32 master [9][co][ci] -> BF16 forward copy (same layout) and BF16 input-gradient copy
// [9][ci][co] with mirrored taps, for every layer in one launch (blocks firstBlock.. of the
// grid belong to a layer).
constexpr int kMaxPackLayers = 24;
struct PackLayer {
    const float* w;
    bf16 *wf, *wb;
    int coPad, ciPad, firstBlock;
};
struct PackArgs {
    PackLayer layer[kMaxPackLayers];
    int count;
};
__global__ void packWeightsKernel(PackArgs a) {
    int l = 0;
    while (l + 1 < a.count && (int)blockIdx.x >= a.layer[l + 1].firstBlock) l++;
    const PackLayer& L = a.layer[l];
    const long long i = (blockIdx.x - L.firstBlock) * (long long)kB + threadIdx.x;
    if (i >= 9LL * L.coPad * L.ciPad) return;
    const int ci = (int)(i % L.ciPad), co = (int)((i / L.ciPad) % L.coPad), tap = (int)(i / ((long long)L.ciPad * L.coPad));
    const bf16 v = (bf16)L.w[i];
    L.wf[i] = v;
    L.wb[((size_t)(8 - tap) * L.ciPad + ci) * L.coPad + co] = v;
}

inline int blocks(long long n) { return (int)((n + kB - 1) / kB); }

}  // namespace

// ====================================================================== network
struct DnLayer {
    int cin, cout, cinPad, coutPad;
    bool relu;
    long long wOff, bOff;   // offsets in the FP32 parameter vector
    long long bfOff;        // offset of this layer's BF16 copies (forward then input-gradient)
};

struct DenoiserNet {
    std::vector<DnLayer> layers;
    long long nParams = 0, nBf = 0;
    float *params = nullptr, *grads = nullptr, *adamM = nullptr, *adamV = nullptr;
    bf16* wbf = nullptr;
    int adamStep = 0;
    // activations for the current (n, h, w)
    int n = 0, h = 0, w = 0;
    std::vector<void*> bufs;
    bf16 *x0 = nullptr, *e0a, *D0, *p0, *e1a, *D1, *p1, *e2a, *D2, *p2, *e3a, *D3, *p3, *ba, *bb, *d3a, *d3b, *d2a, *d2b,
         *d1a, *d1b, *d0a, *d0b;
    bf16 *gD0, *gD1, *gD2, *gD3, *ge0a, *gp0, *ge1a, *gp1, *ge2a, *gp2, *ge3a, *gp3, *gba, *gbb, *gd3a, *gd3b, *gd2a,
         *gd2b, *gd1a, *gd1b, *gd0a, *gd0b, *gout;
    float *outF = nullptr, *noisyLog = nullptr, *demod = nullptr, *targetLog = nullptr, *loss = nullptr;
    // dataset
    float *dsIn = nullptr, *dsTgt = nullptr;
    int dsN = 0, dsH = 0, dsW = 0;
    Crop* dCrops = nullptr;
    int cropCap = 0;
    // per-block slots of the weight-gradient kernel (summed by wgReduceKernel); a backward
    // pass gives every layer its own slots and sums them all in one launch at the end
    float* wgPart = nullptr;
    size_t wgPartFloats = 0;
    struct WgPending {
        size_t off;
        int per, pairs, nci, cinPad, coutPad;
        float *dW, *dB;
    };
    std::vector<WgPending> wgPending;
    bool wgDefer = false;
    size_t wgUsed = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float lastFwdMs = 0, lastTrainMs = 0;
    float displayWeight = 0;   // weight of the display-space L1 term of the training loss
};

namespace {

struct LayerSpec { int cin, cout; bool relu; };
// e0a e0b e1a e1b e2a e2b e3a e3b ba bb d3a d3b d2a d2b d1a d1b d0a d0b out
const LayerSpec kSpec[] = {{10, 32, true},  {32, 32, true},   {32, 64, true},  {64, 64, true}, {64, 64, true},
                           {64, 64, true},  {64, 96, true},   {96, 96, true},  {96, 128, true}, {128, 128, true},
                           {224, 96, true}, {96, 96, true},   {160, 64, true}, {64, 64, true}, {128, 64, true},
                           {64, 64, true},  {96, 32, true},   {32, 32, true},  {32, 3, false}};
constexpr int kLayers = sizeof(kSpec) / sizeof(kSpec[0]);
enum { E0A, E0B, E1A, E1B, E2A, E2B, E3A, E3B, BA, BB, D3A, D3B, D2A, D2B, D1A, D1B, D0A, D0B, OUT };

inline int pad32(int c) { return (c + 31) / 32 * 32; }

View V(bf16* p, int cs, int co = 0, int c = -1) { return View{p, cs, co, c < 0 ? cs : c}; }

bf16* wfOf(DenoiserNet* net, int l) { return net->wbf + net->layers[l].bfOff; }
bf16* wbOf(DenoiserNet* net, int l) {
    const DnLayer& L = net->layers[l];
    return net->wbf + L.bfOff + 9LL * L.coutPad * L.cinPad;
}

// Whether launchConvFwd runs a convolution on the tile kernel.
bool tileConv(const ConvArgs& a) {
    return a.coutPad % 32 == 0 && (a.w >= 16 || (long long)a.w * a.h >= 16384) &&
           (long long)a.n * a.h * a.w * a.in.cs * 2 < (long long)kOutside && (long long)a.w * a.in.cs * 2 < (1LL << 24) &&
           (long long)9 * a.coutPad * a.cinPad * 2 < kOutside && (a.in.cs | a.in.co) % 8 == 0 &&
           (!a.out.p || ((a.out.cs | a.out.co) % 4 == 0 && (long long)a.n * a.h * a.w * a.out.cs * 2 < (long long)kOutside)) &&
           (!a.outF || (a.outFStride % 4 == 0 && (long long)a.n * a.h * a.w * a.outFStride * 4 < (long long)kOutside)) &&
           (!a.low.p || (a.lowC % 32 == 0 && a.h % 2 == 0 && a.w % 2 == 0 && (a.low.cs | a.low.co) % 8 == 0)) &&
           (!a.omask.p || ((a.omask.cs | a.omask.co) % 4 == 0 && (long long)a.n * a.h * a.w * a.omask.cs * 2 < (long long)kOutside));
}

// A layer runs 4-row tiles when its 8-row work items (tiles x 32-channel groups) number fewer
// than RT_CONV_SMALL percent of the compute units (the 1/16- and 1/8-resolution layers of a
// 1080p frame), 8-row tiles otherwise (their halo rows and weights serve twice the pixels).
#ifndef RT_CONV_SMALL
#define RT_CONV_SMALL 50
#endif

// Launches one tile kernel as a persistent grid: as many blocks of each channel group as the
// GPU holds at once (its occupancy with `dynLds` bytes of weights, found once per size),
// fewer for small images.
template <int NT, int ROWS, bool MULTI>
hipError_t launchTileKernel(const ConvArgs& a, int groups, int tilesX, int tilesY, int tileCount, int dynLds, hipStream_t s) {
    constexpr int threads = 64 * ROWS;
    static std::mutex lock;
    static int sizes[8], resident[8], known = 0;
    int blocks = 0;
    {
        std::lock_guard<std::mutex> guard(lock);
        for (int i = 0; i < known && !blocks; i++)
            if (sizes[i] == dynLds) blocks = resident[i];
        if (!blocks) {
            int dev = 0, cus = 0, perCu = 0;
            hipError_t e = hipGetDevice(&dev);
            if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            // a request to allow that much dynamic LDS (the launch itself reports a size the
            // device cannot give; a runtime that needs no such request may refuse it)
            if (e == hipSuccess && dynLds > 0 &&
                hipFuncSetAttribute(reinterpret_cast<const void*>(convFwdTileKernel<NT, ROWS, MULTI>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, dynLds) != hipSuccess)
                (void)hipGetLastError();
            if (e == hipSuccess)
                e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, convFwdTileKernel<NT, ROWS, MULTI>, threads, dynLds);
            if (e != hipSuccess) return e;
            blocks = std::max(1, cus) * std::max(1, perCu);
            if (known < 8) {
                sizes[known] = dynLds;
                resident[known++] = blocks;
            }
        }
    }
    const dim3 grid(std::min(tileCount, std::max(1, blocks / groups)), groups);
    hipLaunchKernelGGL((convFwdTileKernel<NT, ROWS, MULTI>), grid, dim3(threads), dynLds, s, a, tilesX, tilesY, tileCount);
    return hipGetLastError();
}

// LDS of a resident-weight tile block besides its weights: the halo slice (two buffers of
// whole 1 KiB load instructions where the halo is loaded straight into LDS) and the biases.
template <int NT, int ROWS> constexpr int tileStaticLds() {
    return (tileDmaHalo<NT, ROWS, false>() ? 2 * ((TileShape<ROWS>::kHaloChunks + 63) / 64) * 1024
                                           : TileShape<ROWS>::kHaloH * kHaloW * 64) +
           64 * NT + 16;
}

// Whether a block's weights (all input slices) fit in LDS beside its halo at the resident-weight
// kernel's blocks per compute unit.
template <int NT, int ROWS> bool weightsResident(const ConvArgs& a, int ldsPerCu) {
    return a.cinPad == 32 ||
           (tileStaticLds<NT, ROWS>() + 9 * 16 * NT * a.cinPad * 2) * tileBlocksPerCu<NT, ROWS, false>() <= ldsPerCu;
}

// The persistent tile kernel for a convolution that tileConv accepts, with ROWS-row tiles: all
// weights of a block resident in LDS when they fit, else one input slice's weights per step.
template <int NT, int ROWS>
hipError_t launchTileWeights(const ConvArgs& a, int groups, int tilesX, int tilesY, int tileCount, int ldsPerCu, hipStream_t s) {
    if (weightsResident<NT, ROWS>(a, ldsPerCu))
        return launchTileKernel<NT, ROWS, false>(a, groups, tilesX, tilesY, tileCount, 9 * 16 * NT * a.cinPad * 2, s);
    return launchTileKernel<NT, ROWS, true>(a, groups, tilesX, tilesY, tileCount, 0, s);
}

template <int ROWS>
hipError_t launchConvTileRows(const ConvArgs& a, int cus, int ldsPerCu, hipStream_t s) {
    const int tilesX = (a.w + kTileW - 1) / kTileW, tilesY = (a.h + ROWS - 1) / ROWS;
    const int tileCount = tilesX * tilesY * a.n;
    // Only an FP32 output of at most 16 channels (the network's last layer): the first
    // 16-channel group suffices.
    if (!a.out.p && a.outF && a.outFStride <= 16) return launchTileWeights<1, ROWS>(a, 1, tilesX, tilesY, tileCount, ldsPerCu, s);
    // 64-channel groups read each halo once for twice the channels; 32-channel groups make
    // twice the blocks, for small images that would leave compute units idle, and keep their
    // weights in LDS where 64-channel groups could not
    const bool wide = a.coutPad % 64 == 0 && (long long)tileCount * (a.coutPad / 64) >= 2LL * cus &&
                      (weightsResident<4, ROWS>(a, ldsPerCu) || !weightsResident<2, ROWS>(a, ldsPerCu));
    const int groups = a.coutPad / (wide ? 64 : 32);
    return wide ? launchTileWeights<4, ROWS>(a, groups, tilesX, tilesY, tileCount, ldsPerCu, s)
                : launchTileWeights<2, ROWS>(a, groups, tilesX, tilesY, tileCount, ldsPerCu, s);
}

hipError_t launchConvTile(const ConvArgs& a, hipStream_t s) {
    static int cus = 0, ldsPerCu = 0;
    if (!cus) {
        int dev = 0, c = 0, l = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&l, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
        if (e != hipSuccess) return e;
        ldsPerCu = l;
        cus = c;
    }
    const long long tiles8 = (long long)((a.w + kTileW - 1) / kTileW) * ((a.h + 7) / 8) * a.n;
    const long long items8 = tiles8 * (a.coutPad / 32);
    return items8 * 100 < (long long)RT_CONV_SMALL * cus ? launchConvTileRows<4>(a, cus, ldsPerCu, s)
                                                          : launchConvTileRows<8>(a, cus, ldsPerCu, s);
}

hipError_t launchConvFwd(const ConvArgs& a, hipStream_t s) {
    const int npix = a.n * a.h * a.w;
    const int tiles = a.coutPad / 16;
    if (a.cinPad % 32 || a.coutPad % 32 || a.in.c != a.cinPad - a.lowC) return hipErrorInvalidValue;
    if (a.low.p && !tileConv(a)) return hipErrorInvalidValue;   // only the tile kernel reads upsampled channels
    if (tileConv(a)) return launchConvTile(a, s);
    if (tiles % 4 == 0) hipLaunchKernelGGL(convFwdKernel<4>, dim3(blocks(npix), tiles / 4), dim3(kB), 0, s, a);
    else hipLaunchKernelGGL(convFwdKernel<2>, dim3(blocks(npix), tiles / 2), dim3(kB), 0, s, a);
    return hipGetLastError();
}

// Weight and bias gradients of one layer from the pre-activation gradient dy, added with
// atomics into dW and dB (zeroed once per backward pass).
#define DN_TRY(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return e_; } while (0)

// Launch shape of the tiled weight-gradient kernel for one layer.
struct WgPlan {
    int S, tilesX, tilesY, tileCount, nci, pairs, per;
    size_t slots() const { return (size_t)pairs * per * kWgSlot; }
};
hipError_t wgPlan(int n, int h, int w, int cinPad, int coutPad, WgPlan& p) {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
    }
    p.S = w > 32 ? 2 : 1;
    const int TH = 16 / p.S, TW = 32 * p.S;
    p.tilesX = (w + TW - 1) / TW;
    p.tilesY = (h + TH - 1) / TH;
    p.tileCount = p.tilesX * p.tilesY * n;
    p.nci = cinPad / 32;
    p.pairs = (coutPad / 32) * p.nci;
    p.per = std::max(1, std::min(p.tileCount, 2 * cus / p.pairs));   // two blocks per compute unit
    return hipSuccess;
}

hipError_t ensureWgSlots(DenoiserNet* net, size_t need) {
    if (need <= net->wgPartFloats) return hipSuccess;
    if (net->wgPart) (void)hipFree(net->wgPart);
    net->wgPart = nullptr;
    net->wgPartFloats = 0;
    const hipError_t e = hipMalloc(&net->wgPart, need * sizeof(float));
    if (e == hipSuccess) net->wgPartFloats = need;
    return e;
}

hipError_t convWeightGrad(DenoiserNet* net, const View& dy, const View& in, int n, int h, int w, int cinPad,
                          int coutPad, float* dW, float* dB, hipStream_t s) {
    if ((dy.cs | dy.co | in.cs | in.co) % 8 == 0 && (long long)n * h * w * std::max(dy.cs, in.cs) * 2 < (long long)kOutside) {
        WgPlan p;
        DN_TRY(wgPlan(n, h, w, cinPad, coutPad, p));
        size_t off = 0;
        if (net->wgDefer) {   // the backward pass sized the slots for every layer
            off = net->wgUsed;
            if (off + p.slots() > net->wgPartFloats) return hipErrorInvalidValue;
            net->wgUsed += p.slots();
        } else {
            DN_TRY(ensureWgSlots(net, p.slots()));
        }
        BwdWTileArgs a{dy, in, net->wgPart + off, n, h, w, p.nci, p.tilesX, p.tilesY, p.tileCount};
        if (p.S == 2) hipLaunchKernelGGL(convBwdWTileKernel<2>, dim3(p.per, p.pairs), dim3(kWgThreads), 0, s, a);
        else hipLaunchKernelGGL(convBwdWTileKernel<1>, dim3(p.per, p.pairs), dim3(kWgThreads), 0, s, a);
        DN_TRY(hipGetLastError());
        if (net->wgDefer) {
            net->wgPending.push_back(DenoiserNet::WgPending{off, p.per, p.pairs, p.nci, cinPad, coutPad, dW, dB});
            return hipSuccess;
        }
        hipLaunchKernelGGL(wgReduceKernel, dim3(blocks((long long)p.pairs * kWgSlot), (p.per + kWgReduceRun - 1) / kWgReduceRun),
                           dim3(kB), 0, s, net->wgPart, p.per, p.pairs, p.nci, cinPad, coutPad, dW, dB);
        return hipGetLastError();
    }
    const int npix = n * h * w;
    int parts = 64;
    int per = (npix + parts - 1) / parts;
    per = (per + 127) / 128 * 128;
    parts = (npix + per - 1) / per;
    BwdWArgs a{dy, in, dW, n, h, w, cinPad, coutPad, per};
    const int tco = (coutPad / 16) % 4 == 0 ? 4 : 2, tci = (cinPad / 16) % 4 == 0 ? 4 : 2;
    dim3 grid(parts, 9, (coutPad / (16 * tco)) * (cinPad / (16 * tci)));
    if (tco == 4 && tci == 4) hipLaunchKernelGGL((convBwdWKernel<4, 4>), grid, dim3(kB), 0, s, a);
    else if (tco == 4) hipLaunchKernelGGL((convBwdWKernel<4, 2>), grid, dim3(kB), 0, s, a);
    else if (tci == 4) hipLaunchKernelGGL((convBwdWKernel<2, 4>), grid, dim3(kB), 0, s, a);
    else hipLaunchKernelGGL((convBwdWKernel<2, 2>), grid, dim3(kB), 0, s, a);
    hipError_t e;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // Bias: many short pixel ranges so the reduction spreads over the whole GPU.
    const int biasBlocks = std::max(1, std::min(2048, npix / 64));
    const int biasPer = (npix + biasBlocks - 1) / biasBlocks;
    hipLaunchKernelGGL(biasGradKernel, dim3(biasBlocks), dim3(kB), 0, s, dy, npix, coutPad, biasPer, dB);
    return hipGetLastError();
}

void freeActivations(DenoiserNet* net) {
    for (void* p : net->bufs) (void)hipFree(p);
    net->bufs.clear();
    net->x0 = nullptr;
    net->n = net->h = net->w = 0;
}

hipError_t ensureActivations(DenoiserNet* net, int n, int h, int w, bool training) {
    if (net->x0 && net->n == n && net->h == h && net->w == w && (!training || net->gout)) return hipSuccess;
    freeActivations(net);
    net->gout = nullptr;
    hipError_t err = hipSuccess;
    const size_t P[5] = {(size_t)n * h * w, (size_t)n * (h / 2) * (w / 2), (size_t)n * (h / 4) * (w / 4),
                         (size_t)n * (h / 8) * (w / 8), (size_t)n * (h / 16) * (w / 16)};
    auto alloc = [&](bf16*& p, int level, int ch) {
        if (err != hipSuccess) return;
        void* q = nullptr;
        err = hipMalloc(&q, P[level] * ch * sizeof(bf16));
        if (err == hipSuccess) { net->bufs.push_back(q); p = static_cast<bf16*>(q); }
    };
    auto allocF = [&](float*& p, size_t count) {
        if (err != hipSuccess) return;
        void* q = nullptr;
        err = hipMalloc(&q, count * sizeof(float));
        if (err == hipSuccess) { net->bufs.push_back(q); p = static_cast<float*>(q); }
    };
    alloc(net->x0, 0, 32); alloc(net->e0a, 0, 32); alloc(net->D0, 0, 96); alloc(net->p0, 1, 32);
    alloc(net->e1a, 1, 64); alloc(net->D1, 1, 128); alloc(net->p1, 2, 64); alloc(net->e2a, 2, 64);
    alloc(net->D2, 2, 160); alloc(net->p2, 3, 64); alloc(net->e3a, 3, 96); alloc(net->D3, 3, 224);
    alloc(net->p3, 4, 96); alloc(net->ba, 4, 128); alloc(net->bb, 4, 128); alloc(net->d3a, 3, 96);
    alloc(net->d3b, 3, 96); alloc(net->d2a, 2, 64); alloc(net->d2b, 2, 64); alloc(net->d1a, 1, 64);
    alloc(net->d1b, 1, 64); alloc(net->d0a, 0, 32); alloc(net->d0b, 0, 32);
    allocF(net->outF, P[0] * 4); allocF(net->noisyLog, P[0] * 4); allocF(net->targetLog, P[0] * 4);
    allocF(net->demod, P[0] * 4);
    allocF(net->loss, 4);
    if (training) {
        alloc(net->gD0, 0, 96); alloc(net->gD1, 1, 128); alloc(net->gD2, 2, 160); alloc(net->gD3, 3, 224);
        alloc(net->ge0a, 0, 32); alloc(net->gp0, 1, 32); alloc(net->ge1a, 1, 64); alloc(net->gp1, 2, 64);
        alloc(net->ge2a, 2, 64); alloc(net->gp2, 3, 64); alloc(net->ge3a, 3, 96); alloc(net->gp3, 4, 96);
        alloc(net->gba, 4, 128); alloc(net->gbb, 4, 128); alloc(net->gd3a, 3, 96); alloc(net->gd3b, 3, 96);
        alloc(net->gd2a, 2, 64); alloc(net->gd2b, 2, 64); alloc(net->gd1a, 1, 64); alloc(net->gd1b, 1, 64);
        alloc(net->gd0a, 0, 32); alloc(net->gd0b, 0, 32);
        bf16* go = nullptr;
        alloc(go, 0, 32);
        if (err == hipSuccess) net->gout = go;
    }
    if (err != hipSuccess) { freeActivations(net); return err; }
    net->n = n; net->h = h; net->w = w;
    return hipSuccess;
}

ConvArgs convArgs(DenoiserNet* net, int l, const View& in, const View& out, int n, int h, int w, float* outF = nullptr) {
    const DnLayer& L = net->layers[l];
    ConvArgs a{};
    a.in = in; a.wt = wfOf(net, l); a.bias = net->params + L.bOff; a.out = out;
    a.outF = outF; a.outFStride = 4;
    a.n = n; a.h = h; a.w = w; a.cinPad = L.cinPad; a.coutPad = L.coutPad; a.relu = L.relu;
    return a;
}

hipError_t conv(DenoiserNet* net, int l, const View& in, const View& out, int n, int h, int w, hipStream_t s,
                float* outF = nullptr) {
    return launchConvFwd(convArgs(net, l, in, out, n, h, w, outF), s);
}

hipError_t pool(const View& in, const View& out, int n, int h, int w, int c, hipStream_t s) {
    hipLaunchKernelGGL(avgPoolKernel, dim3(blocks((long long)n * (h / 2) * (w / 2) * (c / 8))), dim3(kB), 0, s, in, out, n, h, w, c);
    return hipGetLastError();
}
hipError_t upsample(const View& in, const View& out, int n, int h, int w, int c, hipStream_t s) {
    hipLaunchKernelGGL(upsampleKernel, dim3(blocks((long long)n * h * w * (c / 8))), dim3(kB), 0, s, in, out, n, h, w, c);
    return hipGetLastError();
}


// Forward pass from net->x0 to net->outF.
//
// Training (fuseUp false) keeps each level's concatenation buffer D = [upsampled | skip] whole,
// since the weight gradients read it: the encoder writes its skip output into D, pooling reads
// it there and the upsampling writes the other channels.  At inference (fuseUp), on levels whose
// first decoder convolution runs on the tile kernel, nothing is concatenated: the encoder writes
// its skip output to the buffer of the decoder's second convolution (free until that
// convolution overwrites it), pooling reads it there, and the decoder convolution reads the
// upsampled channels from the half-resolution tensor in place and the skip channels from that
// buffer, each contiguous.
hipError_t forward(DenoiserNet* net, hipStream_t s, bool fuseUp = false) {
    const int n = net->n;
    int H[5], W[5];
    for (int l = 0; l < 5; l++) { H[l] = net->h >> l; W[l] = net->w >> l; }
    // per level: concatenation buffer, its width, the upsampled channels, skip channels, the
    // encoder / decoder buffers, decoder layers
    struct Level { bf16* D; int dc; int upC; int skipC; bf16* enc; int encC; bf16* skipBuf; int ea, eb, da; };
    const Level L[4] = {
        {net->D0, 96, 64, 32, net->e0a, 32, net->d0b, E0A, E0B, D0A},
        {net->D1, 128, 64, 64, net->e1a, 64, net->d1b, E1A, E1B, D1A},
        {net->D2, 160, 96, 64, net->e2a, 64, net->d2b, E2A, E2B, D2A},
        {net->D3, 224, 128, 96, net->e3a, 96, net->d3b, E3A, E3B, D3A}};
    const View up[4] = {V(net->d1b, 64), V(net->d2b, 64), V(net->d3b, 96), V(net->bb, 128)};
    const View dOut[4] = {V(net->d0a, 32), V(net->d1a, 64), V(net->d2a, 64), V(net->d3a, 96)};
    bool fused[4];
    View skip[4];
    for (int l = 0; l < 4; l++) {
        ConvArgs a = convArgs(net, L[l].da, V(L[l].skipBuf, L[l].skipC), dOut[l], n, H[l], W[l]);
        a.low = up[l]; a.lowC = L[l].upC;
        fused[l] = fuseUp && tileConv(a);
        skip[l] = fused[l] ? V(L[l].skipBuf, L[l].skipC) : V(L[l].D, L[l].dc, L[l].upC, L[l].skipC);
    }
    const View pooled[4] = {V(net->p0, 32), V(net->p1, 64), V(net->p2, 64), V(net->p3, 96)};
    View x = V(net->x0, 32);
    for (int l = 0; l < 4; l++) {   // encoder
        DN_TRY(conv(net, L[l].ea, x, V(L[l].enc, L[l].encC), n, H[l], W[l], s));
        DN_TRY(conv(net, L[l].eb, V(L[l].enc, L[l].encC), skip[l], n, H[l], W[l], s));
        DN_TRY(pool(skip[l], pooled[l], n, H[l], W[l], L[l].skipC, s));
        x = pooled[l];
    }
    DN_TRY(conv(net, BA, V(net->p3, 96), V(net->ba, 128), n, H[4], W[4], s));
    DN_TRY(conv(net, BB, V(net->ba, 128), V(net->bb, 128), n, H[4], W[4], s));
    const int db[4] = {D0B, D1B, D2B, D3B};
    for (int l = 3; l >= 0; l--) {   // decoder
        if (fused[l]) {
            ConvArgs a = convArgs(net, L[l].da, skip[l], dOut[l], n, H[l], W[l]);
            a.low = up[l]; a.lowC = L[l].upC;
            DN_TRY(launchConvFwd(a, s));
        } else {
            DN_TRY(upsample(up[l], V(L[l].D, L[l].dc, 0, L[l].upC), n, H[l], W[l], L[l].upC, s));
            DN_TRY(conv(net, L[l].da, V(L[l].D, L[l].dc), dOut[l], n, H[l], W[l], s));
        }
        DN_TRY(conv(net, db[l], dOut[l], V(L[l].skipBuf, L[l].skipC), n, H[l], W[l], s));
    }
    DN_TRY(conv(net, OUT, V(net->d0b, 32), View{nullptr, 32, 0, 32}, n, H[0], W[0], s, net->outF));
    return hipSuccess;
}

// Backward through one convolution: weight/bias gradients and (if gin.p) the input
// gradient.  g is the gradient w.r.t. the layer's pre-activation output:
// every producer of a ReLU layer's output gradient zeroes it where that output is not
// positive (the input-gradient convolution's `omask`, the pooling backward's mask; nearest
// upsampling's backward sums pixels that share one mask value).  gmask: the ReLU output under
// gin, or none when gin is a pooling output's gradient (no activation).
hipError_t convBack(DenoiserNet* net, int l, const View& g, const View& in, const View& gin, const View& gmask,
                    int n, int h, int w, hipStream_t s) {
    const DnLayer& L = net->layers[l];
    DN_TRY(convWeightGrad(net, g, in, n, h, w, L.cinPad, L.coutPad, net->grads + L.wOff, net->grads + L.bOff, s));
    if (gin.p) {
        ConvArgs a{};
        a.in = g; a.wt = wbOf(net, l); a.bias = nullptr; a.out = gin; a.outF = nullptr; a.outFStride = 0;
        a.n = n; a.h = h; a.w = w; a.cinPad = L.coutPad; a.coutPad = L.cinPad; a.relu = 0;
        a.omask = gmask;
        DN_TRY(launchConvFwd(a, s));
    }
    return hipSuccess;
}

hipError_t backwardLayers(DenoiserNet* net, hipStream_t s) {
    const int n = net->n, h0 = net->h, w0 = net->w;
    const int h1 = h0 / 2, w1 = w0 / 2, h2 = h0 / 4, w2 = w0 / 4, h3 = h0 / 8, w3 = w0 / 8, h4 = h0 / 16, w4 = w0 / 16;
    const View none{nullptr, 0, 0, 0};
    // convBack(layer, output gradient, input, input gradient, ReLU output under it, ...)
    DN_TRY(convBack(net, OUT, V(net->gout, 32), V(net->d0b, 32), V(net->gd0b, 32), V(net->d0b, 32), n, h0, w0, s));
    DN_TRY(convBack(net, D0B, V(net->gd0b, 32), V(net->d0a, 32), V(net->gd0a, 32), V(net->d0a, 32), n, h0, w0, s));
    DN_TRY(convBack(net, D0A, V(net->gd0a, 32), V(net->D0, 96), V(net->gD0, 96), V(net->D0, 96), n, h0, w0, s));
    hipLaunchKernelGGL(upsampleBackKernel, dim3(blocks((long long)n * h1 * w1 * 8)), dim3(kB), 0, s, V(net->gD0, 96, 0, 64),
                       V(net->gd1b, 64), n, h1, w1, 64);
    DN_TRY(convBack(net, D1B, V(net->gd1b, 64), V(net->d1a, 64), V(net->gd1a, 64), V(net->d1a, 64), n, h1, w1, s));
    DN_TRY(convBack(net, D1A, V(net->gd1a, 64), V(net->D1, 128), V(net->gD1, 128), V(net->D1, 128), n, h1, w1, s));
    hipLaunchKernelGGL(upsampleBackKernel, dim3(blocks((long long)n * h2 * w2 * 8)), dim3(kB), 0, s, V(net->gD1, 128, 0, 64),
                       V(net->gd2b, 64), n, h2, w2, 64);
    DN_TRY(convBack(net, D2B, V(net->gd2b, 64), V(net->d2a, 64), V(net->gd2a, 64), V(net->d2a, 64), n, h2, w2, s));
    DN_TRY(convBack(net, D2A, V(net->gd2a, 64), V(net->D2, 160), V(net->gD2, 160), V(net->D2, 160), n, h2, w2, s));
    hipLaunchKernelGGL(upsampleBackKernel, dim3(blocks((long long)n * h3 * w3 * 12)), dim3(kB), 0, s, V(net->gD2, 160, 0, 96),
                       V(net->gd3b, 96), n, h3, w3, 96);
    DN_TRY(convBack(net, D3B, V(net->gd3b, 96), V(net->d3a, 96), V(net->gd3a, 96), V(net->d3a, 96), n, h3, w3, s));
    DN_TRY(convBack(net, D3A, V(net->gd3a, 96), V(net->D3, 224), V(net->gD3, 224), V(net->D3, 224), n, h3, w3, s));
    hipLaunchKernelGGL(upsampleBackKernel, dim3(blocks((long long)n * h4 * w4 * 16)), dim3(kB), 0, s, V(net->gD3, 224, 0, 128),
                       V(net->gbb, 128), n, h4, w4, 128);
    DN_TRY(convBack(net, BB, V(net->gbb, 128), V(net->ba, 128), V(net->gba, 128), V(net->ba, 128), n, h4, w4, s));
    DN_TRY(convBack(net, BA, V(net->gba, 128), V(net->p3, 96), V(net->gp3, 96), none, n, h4, w4, s));
    hipLaunchKernelGGL(avgPoolBackKernel, dim3(blocks((long long)n * h3 * w3 * 12)), dim3(kB), 0, s, V(net->gp3, 96),
                       V(net->gD3, 224, 128, 96), V(net->D3, 224, 128, 96), n, h3, w3, 96);
    DN_TRY(convBack(net, E3B, V(net->gD3, 224, 128, 96), V(net->e3a, 96), V(net->ge3a, 96), V(net->e3a, 96), n, h3, w3, s));
    DN_TRY(convBack(net, E3A, V(net->ge3a, 96), V(net->p2, 64), V(net->gp2, 64), none, n, h3, w3, s));
    hipLaunchKernelGGL(avgPoolBackKernel, dim3(blocks((long long)n * h2 * w2 * 8)), dim3(kB), 0, s, V(net->gp2, 64),
                       V(net->gD2, 160, 96, 64), V(net->D2, 160, 96, 64), n, h2, w2, 64);
    DN_TRY(convBack(net, E2B, V(net->gD2, 160, 96, 64), V(net->e2a, 64), V(net->ge2a, 64), V(net->e2a, 64), n, h2, w2, s));
    DN_TRY(convBack(net, E2A, V(net->ge2a, 64), V(net->p1, 64), V(net->gp1, 64), none, n, h2, w2, s));
    hipLaunchKernelGGL(avgPoolBackKernel, dim3(blocks((long long)n * h1 * w1 * 8)), dim3(kB), 0, s, V(net->gp1, 64),
                       V(net->gD1, 128, 64, 64), V(net->D1, 128, 64, 64), n, h1, w1, 64);
    DN_TRY(convBack(net, E1B, V(net->gD1, 128, 64, 64), V(net->e1a, 64), V(net->ge1a, 64), V(net->e1a, 64), n, h1, w1, s));
    DN_TRY(convBack(net, E1A, V(net->ge1a, 64), V(net->p0, 32), V(net->gp0, 32), none, n, h1, w1, s));
    hipLaunchKernelGGL(avgPoolBackKernel, dim3(blocks((long long)n * h0 * w0 * 4)), dim3(kB), 0, s, V(net->gp0, 32),
                       V(net->gD0, 96, 64, 32), V(net->D0, 96, 64, 32), n, h0, w0, 32);
    DN_TRY(convBack(net, E0B, V(net->gD0, 96, 64, 32), V(net->e0a, 32), V(net->ge0a, 32), V(net->e0a, 32), n, h0, w0, s));
    DN_TRY(convBack(net, E0A, V(net->ge0a, 32), V(net->x0, 32), none, none, n, h0, w0, s));
    return hipGetLastError();
}

// Level (pooling depth) of each layer, in kSpec order.
constexpr int kLayerLevel[] = {0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 3, 3, 2, 2, 1, 1, 0, 0, 0};
static_assert(sizeof(kLayerLevel) / sizeof(kLayerLevel[0]) == kLayers, "one level per layer");

// Gradients of every parameter: the layers' weight-gradient kernels write their block slots
// into one workspace, and a single launch at the end sums them into the gradient.
hipError_t backward(DenoiserNet* net, hipStream_t s) {
    DN_TRY(hipMemsetAsync(net->grads, 0, net->nParams * sizeof(float), s));
    size_t total = 0;
    for (int l = 0; l < kLayers; l++) {
        const DnLayer& L = net->layers[l];
        WgPlan p;
        DN_TRY(wgPlan(net->n, net->h >> kLayerLevel[l], net->w >> kLayerLevel[l], L.cinPad, L.coutPad, p));
        total += p.slots();
    }
    DN_TRY(ensureWgSlots(net, total));
    net->wgPending.clear();
    net->wgUsed = 0;
    net->wgDefer = true;
    const hipError_t e = backwardLayers(net, s);
    net->wgDefer = false;
    if (e != hipSuccess) return e;
    if (net->wgPending.empty()) return hipSuccess;
    if (net->wgPending.size() > (size_t)kMaxWgLayers) return hipErrorInvalidValue;
    WgReduceArgs a{};
    int first = 0;
    for (size_t i = 0; i < net->wgPending.size(); i++) {
        const DenoiserNet::WgPending& q = net->wgPending[i];
        const int xBlocks = blocks((long long)q.pairs * kWgSlot), runs = (q.per + kWgReduceRun - 1) / kWgReduceRun;
        a.layer[i] = WgReduceLayer{net->wgPart + q.off, q.dW, q.dB, q.per, q.pairs, q.nci, q.cinPad, q.coutPad, xBlocks, first};
        first += xBlocks * runs;
    }
    a.count = (int)net->wgPending.size();
    hipLaunchKernelGGL(wgReduceAllKernel, dim3(first), dim3(kB), 0, s, a);
    return hipGetLastError();
}

hipError_t packAll(DenoiserNet* net, hipStream_t s) {
    static_assert(kLayers <= kMaxPackLayers, "one packing launch covers every layer");
    PackArgs a{};
    int total = 0;
    for (int l = 0; l < kLayers; l++) {
        const DnLayer& L = net->layers[l];
        a.layer[l] = PackLayer{net->params + L.wOff, wfOf(net, l), wbOf(net, l), L.coutPad, L.cinPad, total};
        total += blocks(9LL * L.coutPad * L.cinPad);
    }
    a.count = kLayers;
    hipLaunchKernelGGL(packWeightsKernel, dim3(total), dim3(kB), 0, s, a);
    return hipGetLastError();
}

}  // namespace

// ====================================================================== public functions
DenoiserNet* denoiserCreate() {
    DenoiserNet* net = new DenoiserNet();
    long long off = 0, bfOff = 0;
    for (const LayerSpec& sp : kSpec) {
        DnLayer L;
        L.cin = sp.cin; L.cout = sp.cout; L.cinPad = pad32(sp.cin); L.coutPad = pad32(sp.cout); L.relu = sp.relu;
        L.wOff = off; off += 9LL * L.coutPad * L.cinPad;
        L.bOff = off; off += L.coutPad;
        L.bfOff = bfOff; bfOff += 2 * 9LL * L.coutPad * L.cinPad;
        net->layers.push_back(L);
    }
    net->nParams = off; net->nBf = bfOff;
    if (hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&net->ev0) != hipSuccess || hipEventCreate(&net->ev1) != hipSuccess ||
        hipMalloc(&net->params, off * 4) != hipSuccess || hipMalloc(&net->grads, off * 4) != hipSuccess ||
        hipMalloc(&net->adamM, off * 4) != hipSuccess || hipMalloc(&net->adamV, off * 4) != hipSuccess ||
        hipMalloc(&net->wbf, bfOff * sizeof(bf16)) != hipSuccess) {
        denoiserDestroy(net);
        return nullptr;
    }
    return net;
}

void denoiserDestroy(DenoiserNet* net) {
    if (!net) return;
    freeActivations(net);
    for (void* p : {(void*)net->params, (void*)net->grads, (void*)net->adamM, (void*)net->adamV, (void*)net->wbf,
                    (void*)net->dsIn, (void*)net->dsTgt, (void*)net->dCrops, (void*)net->wgPart})
        if (p) (void)hipFree(p);
    if (net->ev0) (void)hipEventDestroy(net->ev0);
    if (net->ev1) (void)hipEventDestroy(net->ev1);
    if (net->stream) (void)hipStreamDestroy(net->stream);
    delete net;
}

long long denoiserParamCount(const DenoiserNet* net) { return net ? net->nParams : 0; }

// He-normal weights (over the real fan-in), zero biases; the output layer starts small so
// the initial network is close to the identity (residual formulation).
hipError_t denoiserInit(DenoiserNet* net, unsigned seed) {
    std::vector<float> host(net->nParams, 0.0f);
    std::mt19937 rng(seed);
    for (int l = 0; l < kLayers; l++) {
        const DnLayer& L = net->layers[l];
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (9.0f * L.cin)) * (l == OUT ? 0.1f : 1.0f));
        for (int tap = 0; tap < 9; tap++)
            for (int co = 0; co < L.cout; co++)
                for (int ci = 0; ci < L.cin; ci++) host[L.wOff + ((size_t)tap * L.coutPad + co) * L.cinPad + ci] = nd(rng);
    }
    return denoiserSetParams(net, host.data());
}

hipError_t denoiserGetParams(DenoiserNet* net, float* host) {
    return hipMemcpy(host, net->params, net->nParams * 4, hipMemcpyDeviceToHost);
}

hipError_t denoiserSetParams(DenoiserNet* net, const float* host) {
    DN_TRY(hipMemcpy(net->params, host, net->nParams * 4, hipMemcpyHostToDevice));
    DN_TRY(hipMemset(net->adamM, 0, net->nParams * 4));
    DN_TRY(hipMemset(net->adamV, 0, net->nParams * 4));
    net->adamStep = 0;
    DN_TRY(packAll(net, net->stream));
    return hipStreamSynchronize(net->stream);
}

bool denoiserSave(DenoiserNet* net, const char* path, std::string& error) {
    std::vector<float> host(net->nParams);
    if (denoiserGetParams(net, host.data()) != hipSuccess) { error = "cannot read denoiser weights"; return false; }
    FILE* f = std::fopen(path, "wb");
    if (!f) { error = std::string("cannot write ") + path; return false; }
    const unsigned hdr[4] = {kWeightFileMagic, kWeightFileVersion, (unsigned)kLayers, (unsigned)net->nParams};
    bool ok = std::fwrite(hdr, sizeof(hdr), 1, f) == 1 && std::fwrite(host.data(), 4, host.size(), f) == host.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) error = std::string("short write to ") + path;
    return ok;
}

bool denoiserLoad(DenoiserNet* net, const char* path, std::string& error) {
    FILE* f = std::fopen(path, "rb");
    if (!f) { error = std::string("denoiser weights not found: ") + path; return false; }
    unsigned hdr[4] = {};
    std::vector<float> host(net->nParams);
    bool ok = std::fread(hdr, sizeof(hdr), 1, f) == 1;
    if (ok && (hdr[0] != kWeightFileMagic || hdr[1] != kWeightFileVersion || hdr[2] != (unsigned)kLayers ||
               hdr[3] != (unsigned)net->nParams)) {
        error = std::string("denoiser weight file has another format: ") + path;
        std::fclose(f);
        return false;
    }
    ok = ok && std::fread(host.data(), 4, host.size(), f) == host.size();
    std::fclose(f);
    if (!ok) { error = std::string("truncated denoiser weight file: ") + path; return false; }
    if (denoiserSetParams(net, host.data()) != hipSuccess) { error = "cannot upload denoiser weights"; return false; }
    return true;
}

hipError_t denoiserInferDevice(DenoiserNet* net, const float* dInput, int n, int h, int w, float* dOutput, hipStream_t s) {
    if (h % 16 || w % 16 || n < 1) return hipErrorInvalidValue;
    DN_TRY(ensureActivations(net, n, h, w, false));
    const int npix = n * h * w;
    hipLaunchKernelGGL(featuresKernel, dim3(blocks(npix)), dim3(kB), 0, s, dInput, npix, net->x0, net->noisyLog,
                       net->demod);
    DN_TRY(hipGetLastError());
    DN_TRY(hipEventRecord(net->ev0, s));
    DN_TRY(forward(net, s, true));
    DN_TRY(hipEventRecord(net->ev1, s));
    hipLaunchKernelGGL(finishOutputKernel, dim3(blocks(npix)), dim3(kB), 0, s, net->outF, net->noisyLog, net->demod, npix,
                       dOutput);
    DN_TRY(hipGetLastError());
    DN_TRY(hipEventSynchronize(net->ev1));
    return hipEventElapsedTime(&net->lastFwdMs, net->ev0, net->ev1);
}

hipError_t denoiserInferHost(DenoiserNet* net, const float* input, int n, int h, int w, float* output) {
    const size_t npix = (size_t)n * h * w;
    float *di = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc(&di, npix * kDenoiseInputChannels * 4);
    if (e == hipSuccess) e = hipMalloc(&dout, npix * 16);
    if (e == hipSuccess) e = hipMemcpy(di, input, npix * kDenoiseInputChannels * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = denoiserInferDevice(net, di, n, h, w, dout, net->stream);
    if (e == hipSuccess) {
        std::vector<float> o(npix * 4);
        e = hipMemcpy(o.data(), dout, npix * 16, hipMemcpyDeviceToHost);
        for (size_t p = 0; p < npix && e == hipSuccess; p++)
            for (int c = 0; c < 3; c++) output[p * 3 + c] = o[p * 4 + c];
    }
    if (di) (void)hipFree(di);
    if (dout) (void)hipFree(dout);
    return e;
}

hipError_t denoiserGradients(DenoiserNet* net, const float* input, const float* target, int n, int h, int w, float* loss,
                             float* grads) {
    if (h % 16 || w % 16 || n < 1) return hipErrorInvalidValue;
    hipStream_t s = net->stream;
    DN_TRY(ensureActivations(net, n, h, w, true));
    const size_t npix = (size_t)n * h * w;
    // Reuse the dataset path: the images are the dataset and each "crop" is a whole image.
    DN_TRY(denoiserSetDataset(net, input, target, n, h, w));
    std::vector<Crop> crops(n);
    for (int b = 0; b < n; b++) crops[b] = Crop{b, 0, 0, 0};
    if (net->cropCap < n) {
        if (net->dCrops) (void)hipFree(net->dCrops);
        net->dCrops = nullptr;
        DN_TRY(hipMalloc(&net->dCrops, n * sizeof(Crop)));
        net->cropCap = n;
    }
    if (h != w) return hipErrorInvalidValue;   // square images only (crop geometry)
    DN_TRY(hipMemcpy(net->dCrops, crops.data(), n * sizeof(Crop), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(cropKernel, dim3(blocks((long long)npix)), dim3(kB), 0, s, net->dsIn, net->dsTgt, h, w, net->dCrops, h, n,
                       net->x0, net->noisyLog, net->demod, net->targetLog);
    DN_TRY(hipGetLastError());
    DN_TRY(forward(net, s));
    DN_TRY(hipMemsetAsync(net->loss, 0, 4, s));
    hipLaunchKernelGGL(lossKernel, dim3(blocks((long long)npix)), dim3(kB), 0, s, net->outF, net->noisyLog, net->targetLog,
                       net->demod, net->displayWeight, (int)npix, net->gout, net->loss);
    DN_TRY(hipGetLastError());
    DN_TRY(backward(net, s));
    DN_TRY(hipStreamSynchronize(s));
    DN_TRY(hipMemcpy(loss, net->loss, 4, hipMemcpyDeviceToHost));
    return hipMemcpy(grads, net->grads, net->nParams * 4, hipMemcpyDeviceToHost);
}

hipError_t denoiserSetDataset(DenoiserNet* net, const float* inputs, const float* targets, int count, int h, int w) {
    if (net->dsIn) (void)hipFree(net->dsIn);
    if (net->dsTgt) (void)hipFree(net->dsTgt);
    net->dsIn = net->dsTgt = nullptr;
    const size_t npix = (size_t)count * h * w;
    DN_TRY(hipMalloc(&net->dsIn, npix * kDenoiseInputChannels * 4));
    DN_TRY(hipMalloc(&net->dsTgt, npix * 3 * 4));
    DN_TRY(hipMemcpy(net->dsIn, inputs, npix * kDenoiseInputChannels * 4, hipMemcpyHostToDevice));
    DN_TRY(hipMemcpy(net->dsTgt, targets, npix * 3 * 4, hipMemcpyHostToDevice));
    net->dsN = count; net->dsH = h; net->dsW = w;
    return hipSuccess;
}

hipError_t denoiserTrain(DenoiserNet* net, int steps, int batch, int crop, float lr, unsigned seed, float* losses) {
    if (!net->dsIn || crop % 16 || crop > net->dsH || crop > net->dsW || batch < 1 || steps < 1) return hipErrorInvalidValue;
    hipStream_t s = net->stream;
    DN_TRY(ensureActivations(net, batch, crop, crop, true));
    std::mt19937 rng(seed);
    std::vector<Crop> crops((size_t)steps * batch);
    for (auto& c : crops) {
        c.image = (int)(rng() % net->dsN);
        c.ox = (int)(rng() % (net->dsW - crop + 1));
        c.oy = (int)(rng() % (net->dsH - crop + 1));
        c.flip = (int)(rng() % 4);
    }
    if (net->cropCap < (int)crops.size()) {
        if (net->dCrops) (void)hipFree(net->dCrops);
        net->dCrops = nullptr;
        DN_TRY(hipMalloc(&net->dCrops, crops.size() * sizeof(Crop)));
        net->cropCap = (int)crops.size();
    }
    DN_TRY(hipMemcpy(net->dCrops, crops.data(), crops.size() * sizeof(Crop), hipMemcpyHostToDevice));
    float* dLoss = nullptr;
    DN_TRY(hipMalloc(&dLoss, steps * sizeof(float)));
    DN_TRY(hipMemsetAsync(dLoss, 0, steps * sizeof(float), s));
    const int npix = batch * crop * crop;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    hipError_t err = hipSuccess;
    for (int it = 0; it < steps && err == hipSuccess; it++) {
        if (it == steps - 1) err = hipEventRecord(net->ev0, s);
        hipLaunchKernelGGL(cropKernel, dim3(blocks((long long)npix)), dim3(kB), 0, s, net->dsIn, net->dsTgt, net->dsH, net->dsW,
                           net->dCrops + (size_t)it * batch, crop, batch, net->x0, net->noisyLog, net->demod,
                           net->targetLog);
        if (err == hipSuccess) err = forward(net, s);
        hipLaunchKernelGGL(lossKernel, dim3(blocks((long long)npix)), dim3(kB), 0, s, net->outF, net->noisyLog,
                           net->targetLog, net->demod, net->displayWeight, npix, net->gout, dLoss + it);
        if (err == hipSuccess) err = backward(net, s);
        if (it == steps - 1 && err == hipSuccess) err = hipEventRecord(net->ev1, s);
        net->adamStep++;
        const float c1 = 1 - std::pow(b1, (float)net->adamStep), c2 = 1 - std::pow(b2, (float)net->adamStep);
        hipLaunchKernelGGL(adamKernel, dim3(blocks(net->nParams)), dim3(kB), 0, s, net->params, net->grads, net->adamM,
                           net->adamV, net->nParams, lr, b1, b2, eps, c1, c2);
        if (err == hipSuccess) err = hipGetLastError();
        if (err == hipSuccess) err = packAll(net, s);
    }
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    if (err == hipSuccess) err = hipEventElapsedTime(&net->lastTrainMs, net->ev0, net->ev1);
    if (err == hipSuccess && losses) err = hipMemcpy(losses, dLoss, steps * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(dLoss);
    return err;
}

double denoiserForwardFlops(int n, int h, int w) {
    const int level[kLayers] = {0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 3, 3, 2, 2, 1, 1, 0, 0, 0};
    double f = 0;
    for (int l = 0; l < kLayers; l++) {
        const double pix = (double)n * (h >> level[l]) * (w >> level[l]);
        f += 2.0 * 9.0 * pad32(kSpec[l].cin) * pad32(kSpec[l].cout) * pix;
    }
    return f;
}
float denoiserLastForwardMs(const DenoiserNet* net) { return net ? net->lastFwdMs : 0; }
bool denoiserSetDisplayWeight(DenoiserNet* net, float weight) {
    if (!net || !(weight >= 0) || weight > 1e6f) return false;
    net->displayWeight = weight;
    return true;
}
float denoiserLastTrainStepMs(const DenoiserNet* net) { return net ? net->lastTrainMs : 0; }

// ---------------------------------------------------------------- single-layer checks
hipError_t convForwardHost(const float* in, int n, int h, int w, int cin, const float* weight, const float* bias, int cout,
                           int relu, float* out) {
    const int ciPad = pad32(cin), coPad = pad32(cout);
    const size_t npix = (size_t)n * h * w;
    std::vector<bf16> hin(npix * ciPad, (bf16)0.0f), hw(9ULL * coPad * ciPad, (bf16)0.0f);
    std::vector<float> hb(coPad, 0.0f);
    for (size_t p = 0; p < npix; p++)
        for (int c = 0; c < cin; c++) hin[p * ciPad + c] = (bf16)in[p * cin + c];
    for (int t = 0; t < 9; t++)
        for (int co = 0; co < cout; co++)
            for (int ci = 0; ci < cin; ci++) hw[((size_t)t * coPad + co) * ciPad + ci] = (bf16)weight[((size_t)t * cout + co) * cin + ci];
    for (int co = 0; co < cout; co++) hb[co] = bias ? bias[co] : 0.0f;
    bf16 *din = nullptr, *dw = nullptr;
    float *db = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc(&din, hin.size() * 2);
    if (e == hipSuccess) e = hipMalloc(&dw, hw.size() * 2);
    if (e == hipSuccess) e = hipMalloc(&db, coPad * 4);
    if (e == hipSuccess) e = hipMalloc(&dout, npix * coPad * 4);
    if (e == hipSuccess) e = hipMemcpy(din, hin.data(), hin.size() * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dw, hw.data(), hw.size() * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, hb.data(), coPad * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        ConvArgs a{};
        a.in = View{din, ciPad, 0, ciPad}; a.wt = dw; a.bias = db; a.out = View{nullptr, coPad, 0, coPad};
        a.outF = dout; a.outFStride = coPad; a.n = n; a.h = h; a.w = w; a.cinPad = ciPad; a.coutPad = coPad;
        a.relu = relu;
        e = launchConvFwd(a, nullptr);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) {
        std::vector<float> o(npix * coPad);
        e = hipMemcpy(o.data(), dout, o.size() * 4, hipMemcpyDeviceToHost);
        for (size_t p = 0; p < npix && e == hipSuccess; p++)
            for (int c = 0; c < cout; c++) out[p * cout + c] = o[p * coPad + c];
    }
    for (void* p : {(void*)din, (void*)dw, (void*)db, (void*)dout})
        if (p) (void)hipFree(p);
    return e;
}

hipError_t convBackwardHost(const float* in, int n, int h, int w, int cin, const float* weight, const float* gradOut,
                            int cout, float* gradIn, float* gradWeight, float* gradBias) {
    DenoiserNet* net = denoiserCreate();   // only for its scratch buffers
    if (!net) return hipErrorOutOfMemory;
    const int ciPad = pad32(cin), coPad = pad32(cout);
    const size_t npix = (size_t)n * h * w;
    std::vector<bf16> hin(npix * ciPad, (bf16)0.0f), hg(npix * coPad, (bf16)0.0f), hwb(9ULL * ciPad * coPad, (bf16)0.0f);
    for (size_t p = 0; p < npix; p++) {
        for (int c = 0; c < cin; c++) hin[p * ciPad + c] = (bf16)in[p * cin + c];
        for (int c = 0; c < cout; c++) hg[p * coPad + c] = (bf16)gradOut[p * cout + c];
    }
    for (int t = 0; t < 9; t++)
        for (int co = 0; co < cout; co++)
            for (int ci = 0; ci < cin; ci++)
                hwb[((size_t)(8 - t) * ciPad + ci) * coPad + co] = (bf16)weight[((size_t)t * cout + co) * cin + ci];
    bf16 *din = nullptr, *dg = nullptr, *dwb = nullptr;
    float *dgi = nullptr, *dgw = nullptr, *dgb = nullptr;
    hipError_t e = hipMalloc(&din, hin.size() * 2);
    if (e == hipSuccess) e = hipMalloc(&dg, hg.size() * 2);
    if (e == hipSuccess) e = hipMalloc(&dwb, hwb.size() * 2);
    if (e == hipSuccess) e = hipMalloc(&dgi, npix * ciPad * 4);
    if (e == hipSuccess) e = hipMalloc(&dgw, 9ULL * coPad * ciPad * 4);
    if (e == hipSuccess) e = hipMalloc(&dgb, coPad * 4);
    if (e == hipSuccess) e = hipMemcpy(din, hin.data(), hin.size() * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dg, hg.data(), hg.size() * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dwb, hwb.data(), hwb.size() * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        ConvArgs a{};
        a.in = View{dg, coPad, 0, coPad}; a.wt = dwb; a.bias = nullptr; a.out = View{nullptr, ciPad, 0, ciPad};
        a.outF = dgi; a.outFStride = ciPad; a.n = n; a.h = h; a.w = w; a.cinPad = coPad; a.coutPad = ciPad;
        a.relu = 0;
        e = launchConvFwd(a, nullptr);
    }
    if (e == hipSuccess) e = hipMemset(dgw, 0, 9ULL * coPad * ciPad * 4);
    if (e == hipSuccess) e = hipMemset(dgb, 0, coPad * 4);
    if (e == hipSuccess)
        e = convWeightGrad(net, View{dg, coPad, 0, coPad}, View{din, ciPad, 0, ciPad}, n, h, w, ciPad, coPad, dgw, dgb,
                           nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) {
        std::vector<float> gi(npix * ciPad), gw(9ULL * coPad * ciPad), gb(coPad);
        e = hipMemcpy(gi.data(), dgi, gi.size() * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(gw.data(), dgw, gw.size() * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(gb.data(), dgb, gb.size() * 4, hipMemcpyDeviceToHost);
        if (e == hipSuccess) {
            for (size_t p = 0; p < npix; p++)
                for (int c = 0; c < cin; c++) gradIn[p * cin + c] = gi[p * ciPad + c];
            for (int t = 0; t < 9; t++)
                for (int co = 0; co < cout; co++)
                    for (int ci = 0; ci < cin; ci++)
                        gradWeight[((size_t)t * cout + co) * cin + ci] = gw[((size_t)t * coPad + co) * ciPad + ci];
            for (int co = 0; co < cout; co++) gradBias[co] = gb[co];
        }
    }
    for (void* p : {(void*)din, (void*)dg, (void*)dwb, (void*)dgi, (void*)dgw, (void*)dgb})
        if (p) (void)hipFree(p);
    denoiserDestroy(net);
    return e;
}

}  // namespace rt
