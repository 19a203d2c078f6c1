// Learned warm start: a small residual CNN, trained and run on the GPU with hand-written
// kernels.  No counterpart in the reference (it is a start image for the reference's loop).
//
// Network: five 3x3 convolutions with zero padding, widths 2 -> 32 -> 32 -> 32 -> 32 -> 1,
// ReLU after the first four, the output added to input feature 0.  Features of a plane:
// f0 = x0 / 64 (the decoded, box-projected plane) and f1 = q(y mod 8, x mod 8) / 64, the
// plane's quantisation table tiled over its 8x8 blocks (a per-pixel quantisation-step map, which
// tells the network each pixel's position in its block as well as the quantisation strength);
// the prediction is 64 * (f0 + net(f)).  Activations are NHWC bf16 with
// the channel dimension padded to 16 or 32; weights are kept in f32 (Adam's master copy) and
// repacked to bf16 operand layouts after every update.
//
// Convolution = implicit GEMM on v_mfma_f32_16x16x16_bf16 (A = weights [cout x cin] per tap,
// B = input [cin x 16 pixels], C = output [cout x 16 pixels], f32 accumulation).  A wave owns
// 16 consecutive pixels of a row and all output channels; its weights for all 9 taps sit in
// registers.  The backward pass over the data is the same kernel on weights transposed
// (cin <-> cout) and flipped (tap (dy, dx) -> (2 - dy, 2 - dx)).  Weight gradients: the same
// MFMA with the pixels as the contraction dimension, per-workgroup partials reduced in a
// fixed order.  Operand lane layouts (16x16x16, 64 lanes): A[m][k] at lane m + 16*(k/4) item
// k%4; B[k][n] at lane n + 16*(k/4) item k%4; C[m][n] at lane n + 16*(m/4) item m%4.
//
// Inference (k_cnn_stream) runs all five layers in one kernel that streams down 64-column
// bands, with activations kept in LDS rows and v_mfma_f32_16x16x32_bf16 (K = 32: lane l holds
// k = 8*(l/16) .. +7 of A row / B column l%16; C as above).  The layered kernels stay for
// training and as the ARTEFACT_CNN_LAYERED=1 inference path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "../../include/artefact_hip.h"
#include "cnn_internal.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NLAYER = 5;
constexpr int WIDTH[NLAYER + 1] = {2, 32, 32, 32, 32, 1};
constexpr int CINP[NLAYER] = {16, 32, 32, 32, 32};   // padded input channels
constexpr int COUTP[NLAYER] = {32, 32, 32, 32, 16};  // padded output channels
constexpr float FEAT = 1.0f / 64.0f;
constexpr int CT = 256;  // threads per workgroup (4 waves)

__device__ __forceinline__ uint16_t f2bf(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Offsets of layer l's weights and biases in the flat parameter vector (w[cout][cin][3][3]
// then b[cout], layer after layer).
struct ParamLayout {
    int w[NLAYER], b[NLAYER], total;
};
constexpr ParamLayout param_layout()
{
    ParamLayout p{};
    int o = 0;
    for (int l = 0; l < NLAYER; l++) {
        p.w[l] = o;
        o += WIDTH[l + 1] * WIDTH[l] * 9;
        p.b[l] = o;
        o += WIDTH[l + 1];
    }
    p.total = o;
    return p;
}
constexpr ParamLayout PL = param_layout();

// MODE_RELU: + bias, ReLU, bf16 out.  MODE_LINEAR: no bias, bf16 out (backward data).
// MODE_FINAL: + bias, channel 0 only, f32 out = residual feature + conv (the prediction).
// MODE_FINAL_IMAGE: MODE_FINAL with the residual formed from the decoded plane (x0 / 64, what
// k_pack_input stores) and the prediction written in image units (x 64), for inference.
enum { MODE_RELU = 0, MODE_LINEAR = 1, MODE_FINAL = 2, MODE_FINAL_IMAGE = 3 };

// A plane's quantisation table (natural order), passed by value to the kernels that form the
// features; pixel (y, x) of the plane takes entry (y mod 8) * 8 + (x mod 8).
struct QuantTable {
    float q[64];
};
__device__ __forceinline__ int qmap_index(int y, int x) { return ((y & 7) << 3) | (x & 7); }

// Output tile of one workgroup: 64 pixels (wave w: pixels 16w .. 16w+15) x RT rows.  The input
// tile (RT + 2 rows x 66 pixels x CI channels, zero outside the plane) is staged in LDS once;
// each wave walks its input rows and feeds every B operand it reads to the (up to three)
// output rows that use it.  Every output still accumulates its taps in the order (dy, dx, kc),
// so the result does not depend on the tiling.
constexpr int RT = 8;
constexpr int TILE_PX = 66;

template <int CI>
constexpr int lds_px_bytes() { return CI * 2 + 8; }  // padded pixel stride (bank spread), a multiple of 8

// FEAT_IN: the input is the decoded plane itself (f32, `in` cast back) and the kernel forms the
// 16 feature channels while staging its tile: bf16(x0 / 64), bf16(q(y mod 8, x mod 8) / 64), then
// zeros, as k_pack_input does.
template <int CI, int CO, int MODE, bool FEAT_IN = false>
__global__ __launch_bounds__(CT) void k_conv(const uint16_t *__restrict__ in, const bf16x4 *__restrict__ wpk,
                                             const float *__restrict__ bias, uint16_t *__restrict__ out,
                                             float *__restrict__ pred, const float *__restrict__ resid, int H,
                                             int W, const QuantTable qt)
{
    static_assert(!FEAT_IN || CI == 16, "features fill one 16-channel chunk");
    constexpr int KC = CI / 16, MH = CO / 16, PXB = lds_px_bytes<CI>();
    __shared__ __attribute__((aligned(16))) uint8_t tile[(RT + 2) * TILE_PX * PXB];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b = blockIdx.z, y0 = blockIdx.y * RT, x0 = blockIdx.x * 64;
    // Stage the input tile: 16-byte units (8 channels), consecutive threads on consecutive units.
    // Every load of the thread is issued before the first store, so the tile costs one memory
    // round trip; a unit past the tile reloads the last one and is not stored.
    constexpr int UPX = CI / 8, UNITS = (RT + 2) * TILE_PX * UPX, PER = (UNITS + CT - 1) / CT;
    uint4 tv[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int u = min((int)threadIdx.x + i * CT, UNITS - 1);
        const int part = u % UPX, px = (u / UPX) % TILE_PX, r = u / (UPX * TILE_PX);
        const int gy = y0 - 1 + r, gx = x0 - 1 + px;
        const bool ok = gy >= 0 && gy < H && gx >= 0 && gx < W;
        const size_t pix = ((size_t)b * H + (ok ? gy : 0)) * W + (ok ? gx : 0);
        uint4 v;
        if constexpr (FEAT_IN) {
            const float x0 = reinterpret_cast<const float *>(in)[pix];
            const uint32_t f01 =
                (uint32_t)f2bf(x0 * FEAT) | ((uint32_t)f2bf(qt.q[qmap_index(gy, gx)] * FEAT) << 16);
            v = make_uint4(part == 0 ? f01 : 0u, 0u, 0u, 0u);
        } else {
            v = *reinterpret_cast<const uint4 *>(in + pix * CI + part * 8);
        }
        tv[i] = ok ? v : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int u = (int)threadIdx.x + i * CT;
        if (u >= UNITS) break;
        const int part = u % UPX, px = (u / UPX) % TILE_PX, r = u / (UPX * TILE_PX);
        // The padded pixel stride is a multiple of 8 bytes, not 16: two 8-byte stores.
        uint2 *dst = reinterpret_cast<uint2 *>(tile + (r * TILE_PX + px) * PXB + part * 16);
        dst[0] = make_uint2(tv[i].x, tv[i].y);
        dst[1] = make_uint2(tv[i].z, tv[i].w);
    }
    // The weights (all 9 taps in registers) are loaded after the tile's loads are consumed.
    bf16x4 wr[9][KC][MH];
#pragma unroll
    for (int t = 0; t < 9; t++)
#pragma unroll
        for (int kc = 0; kc < KC; kc++)
#pragma unroll
            for (int mh = 0; mh < MH; mh++) wr[t][kc][mh] = wpk[((t * KC + kc) * MH + mh) * 64 + lane];
    __syncthreads();
    const int n = lane & 15, kq = (lane >> 4) * 4;
    f32x4 acc[RT][MH];
#pragma unroll
    for (int ro = 0; ro < RT; ro++)
#pragma unroll
        for (int mh = 0; mh < MH; mh++) acc[ro][mh] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ir = 0; ir < RT + 2; ir++) {
        bf16x4 bv[3][KC];
#pragma unroll
        for (int dx = 0; dx < 3; dx++)
#pragma unroll
            for (int kc = 0; kc < KC; kc++)
                bv[dx][kc] = *reinterpret_cast<const bf16x4 *>(tile + (ir * TILE_PX + wave * 16 + n + dx) * PXB +
                                                               (kc * 16 + kq) * 2);
#pragma unroll
        for (int dx = 0; dx < 3; dx++)
#pragma unroll
            for (int kc = 0; kc < KC; kc++)
#pragma unroll
                for (int dy = 2; dy >= 0; dy--) {  // output row ro = ir - dy uses this input row as tap dy
                    const int ro = ir - dy;
                    if (ro < 0 || ro >= RT) continue;
#pragma unroll
                    for (int mh = 0; mh < MH; mh++)
                        acc[ro][mh] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wr[dy * 3 + dx][kc][mh], bv[dx][kc],
                                                                                acc[ro][mh], 0, 0, 0);
                }
    }
    const int px = x0 + wave * 16 + n;
    if (px >= W) return;
#pragma unroll
    for (int ro = 0; ro < RT; ro++) {
        const int y = y0 + ro;
        if (y >= H) break;
        const size_t p = ((size_t)b * H + y) * W + px;
        if (MODE == MODE_FINAL) {
            if (kq == 0) pred[p] = resid[p] + (acc[ro][0][0] + bias[0]);
            continue;
        }
        if (MODE == MODE_FINAL_IMAGE) {
            if (kq == 0) pred[p] = (resid[p] * FEAT + (acc[ro][0][0] + bias[0])) * 64.0f;
            continue;
        }
#pragma unroll
        for (int mh = 0; mh < MH; mh++) {
            uint16_t v[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                float r = acc[ro][mh][t];
                if (MODE == MODE_RELU) {
                    const int co = mh * 16 + kq + t;
                    r += bias[co];
                    r = r > 0.0f ? r : 0.0f;
                }
                v[t] = f2bf(r);
            }
            uint2 packed;
            packed.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
            packed.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
            *reinterpret_cast<uint2 *>(out + p * CO + mh * 16 + kq) = packed;
        }
    }
}

// Weight and bias gradients of one layer over a range of 16-pixel row segments.  Wave w
// handles taps w, w+4, w+8 of the ten "taps" (0..8 the 3x3 taps, 9 the bias, whose B operand
// is 1 in column 0).  part: [workgroup][10][CO][CI] f32.
template <int CI, int CO>
__global__ __launch_bounds__(CT) void k_wgrad(const uint16_t *__restrict__ act, const uint16_t *__restrict__ gout,
                                              float *__restrict__ part, int B, int H, int W)
{
    constexpr int KC = CI / 16, MH = CO / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, kq = (lane >> 4) * 4;
    const long segs = (long)B * H * (W / 16);
    const long s0 = segs * blockIdx.x / gridDim.x, s1 = segs * (blockIdx.x + 1) / gridDim.x;
    f32x4 acc[3][MH][KC];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int mh = 0; mh < MH; mh++)
#pragma unroll
            for (int kc = 0; kc < KC; kc++) acc[i][mh][kc] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (long s = s0; s < s1; s++) {
        const int xs = (int)(s % (W / 16)) * 16;
        const long row = s / (W / 16);
        const int y = (int)(row % H), b = (int)(row / H);
        // A[m = co][k = pixel]: gout of pixels xs + kq + t.
        bf16x4 av[MH];
#pragma unroll
        for (int mh = 0; mh < MH; mh++)
#pragma unroll
            for (int t = 0; t < 4; t++)
                av[mh][t] = (short)gout[(((size_t)b * H + y) * W + xs + kq + t) * CO + mh * 16 + n];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int tap = wave + 4 * i;
            if (tap > 9) break;
            bf16x4 bv[KC];
#pragma unroll
            for (int kc = 0; kc < KC; kc++) {
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    short v = 0;
                    if (tap == 9) {
                        v = (kc == 0 && n == 0) ? (short)0x3F80 : (short)0;  // bf16 1.0 in column 0
                    } else {
                        const int yy = y + tap / 3 - 1, px = xs + kq + t + tap % 3 - 1;
                        if (yy >= 0 && yy < H && px >= 0 && px < W)
                            v = (short)act[(((size_t)b * H + yy) * W + px) * CI + kc * 16 + n];
                    }
                    bv[kc][t] = v;
                }
            }
#pragma unroll
            for (int mh = 0; mh < MH; mh++)
#pragma unroll
                for (int kc = 0; kc < KC; kc++)
                    acc[i][mh][kc] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[mh], bv[kc], acc[i][mh][kc], 0, 0, 0);
        }
    }
    // C[m = co][n = ci] at lane: co = mh*16 + kq + t, ci = kc*16 + n.
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int tap = wave + 4 * i;
        if (tap > 9) break;
        float *dst = part + ((size_t)blockIdx.x * 10 + tap) * CO * CI;
#pragma unroll
        for (int mh = 0; mh < MH; mh++)
#pragma unroll
            for (int kc = 0; kc < KC; kc++)
#pragma unroll
                for (int t = 0; t < 4; t++) dst[(size_t)(mh * 16 + kq + t) * CI + kc * 16 + n] = acc[i][mh][kc][t];
    }
}

// Sums the workgroup partials of layer l in a fixed order into the flat gradient vector.
__global__ __launch_bounds__(CT) void k_wgrad_reduce(const float *part, int groups, int CI, int CO, int cin,
                                                     int cout, float *gw, float *gb)
{
    const int i = blockIdx.x * CT + threadIdx.x;  // over cout * cin * 9 + cout
    const int nw = cout * cin * 9;
    if (i >= nw + cout) return;
    int tap, co, ci;
    if (i < nw) {
        co = i / (cin * 9);
        ci = (i / 9) % cin;
        tap = i % 9;
    } else {
        co = i - nw;
        ci = 0;
        tap = 9;
    }
    double s = 0.0;
    for (int g = 0; g < groups; g++) s += part[(((size_t)g * 10 + tap) * CO + co) * CI + ci];
    if (i < nw) gw[i] = (float)s;
    else gb[co] = (float)s;
}

// Weight packing into the A-operand layout [tap][kc][mh][lane] (4 bf16 per lane).  With
// transpose set, packs the backward-data operator: A[m = ci][k = co] at tap (2-dy, 2-dx).
__global__ __launch_bounds__(CT) void k_pack_weights(const float *w, int cin, int cout, int CI, int CO, int transpose,
                                                     bf16x4 *dst)
{
    const int KCin = (transpose ? CO : CI) / 16, MHo = (transpose ? CI : CO) / 16;
    const int total = 9 * KCin * MHo * 64;
    const int i = blockIdx.x * CT + threadIdx.x;
    if (i >= total) return;
    const int lane = i & 63, mh = (i >> 6) % MHo, kc = (i >> 6) / MHo % KCin, tap = (i >> 6) / MHo / KCin;
    bf16x4 v;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int m = mh * 16 + (lane & 15), k = kc * 16 + (lane >> 4) * 4 + t;
        float x = 0.0f;
        if (!transpose) {
            if (m < cout && k < cin) x = w[((size_t)m * cin + k) * 9 + tap];
        } else {
            const int ftap = 8 - tap;  // (dy, dx) -> (2 - dy, 2 - dx)
            if (k < cout && m < cin) x = w[((size_t)k * cin + m) * 9 + ftap];
        }
        v[t] = (short)f2bf(x);
    }
    dst[i] = v;
}

// Features of B planes of width w: channel 0 = x0 / 64, channel 1 = q(y mod 8, x mod 8) / 64 with
// plane b's table quant[b * 64 ..], channels 2..15 = 0.
__global__ __launch_bounds__(CT) void k_pack_input(const float *x0, const float *quant, size_t plane, int w, int B,
                                                   uint16_t *a0, float *resid)
{
    const size_t i = (size_t)blockIdx.x * CT + threadIdx.x;
    if (i >= plane * B) return;
    const int b = (int)(i / plane);
    const int p = (int)(i - (size_t)b * plane);
    const float f0 = x0[i] * FEAT;
    resid[i] = f0;
    uint16_t *o = a0 + i * 16;
    o[0] = f2bf(f0);
    o[1] = f2bf(quant[b * 64 + qmap_index(p / w, p % w)] * FEAT);
#pragma unroll
    for (int c = 2; c < 16; c++) o[c] = 0;
}

// Prediction back to image units.

// MSE loss gradient in feature units: g = 2 (pred - target / 64) / n into channel 0 of a
// 16-channel bf16 gradient; per-workgroup partial sums of the squared error (f64).
__global__ __launch_bounds__(CT) void k_loss(const float *pred, const float *target, size_t n, uint16_t *g,
                                             double *partial)
{
    __shared__ double red[CT];
    const size_t i = (size_t)blockIdx.x * CT + threadIdx.x;
    double e2 = 0.0;
    if (i < n) {
        const float d = pred[i] - target[i] * FEAT;
        e2 = (double)d * d;
        uint16_t *o = g + i * 16;
        o[0] = f2bf(2.0f * d / (float)n);
#pragma unroll
        for (int c = 1; c < 16; c++) o[c] = 0;
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = CT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// g *= (act > 0): the ReLU's backward pass, in place.
__global__ __launch_bounds__(CT) void k_relu_bwd(uint16_t *g, const uint16_t *act, size_t n)
{
    const size_t i = (size_t)blockIdx.x * CT + threadIdx.x;
    if (i < n && !(bf2f(act[i]) > 0.0f)) g[i] = 0;
}

// Adam (Kingma & Ba 2015, Algorithm 1) on the f32 master parameters.
__global__ __launch_bounds__(CT) void k_adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1,
                                             float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * CT + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

unsigned blocks(size_t n) { return (unsigned)((n + CT - 1) / CT); }

}  // namespace

namespace {

// ---------------------------------------------------------------------------------------------
// Inference in one kernel: the five layers stream down a column band.
//
// A workgroup owns SB = 64 output columns and RH output rows of one plane.  Wave 0 stages the
// input features and runs layer 1; waves 1-6 run layers 2-4, one wave per half of a layer's
// output channels; wave 7 runs layer 5 and writes the prediction.  Each step, every layer
// produces one row, over 80 columns starting 5 - l columns left of the band (the next layers
// read 64 + 2 (5 - l) of them).  Rows pass from layer to layer through LDS rings of two rows:
// what a layer writes in step t the next one reads in step t + 1, with one barrier per step.
// Layers 2-5 keep three output rows of accumulators: taking in input row j adds the dy = 0 taps
// of output row j + 1, the dy = 1 taps of row j and the dy = 2 taps of row j - 1, which is then
// complete.  Every output therefore sums its taps in the order (dy, dx), each tap one
// v_mfma_f32_16x16x32_bf16 over all 32 input channels.  Layer 1 packs its 9 taps x 2 features
// into one K = 32 operand.  Outputs outside the plane are stored as zeros: the next layer's
// zero padding.  Rows computed above or below the ones the band needs (pipeline fill and
// drain) only feed rows that are not needed either.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <int V>
using IC = std::integral_constant<int, V>;

constexpr int SB = 64;                            // output columns per band
constexpr int ST = 5;                             // 16-pixel tiles per layer row (80 columns)
constexpr int SPX = 82;                           // ring row: 80 columns + the 2 a 3-tap read reaches past
constexpr int SSTR = 80;                          // bytes per ring pixel: 32 bf16 channels + 16 padding
constexpr int SROW = SPX * SSTR;                  // one ring row
constexpr int SRING = 2 * SROW;                   // one activation ring (rows alternate by step)
constexpr int SFEAT = 4 * SPX * 4;                // feature ring: 4 rows of (f0, q) bf16 pairs
constexpr int SQMAP = 64 * 4;                     // the plane's q map entries, bf16 in the high half
constexpr int SLDS = 4 * SRING + SFEAT + SQMAP;   // rings of layers 1-4, the features, the q map
constexpr int SFILL = 12;                         // steps from a band's first input row to its first output row
// Packed weights, one u32x4 (8 bf16) per lane: layer 1 [mh][lane]; layers 2-4 [tap][mh][lane];
// layer 5 [tap][lane].
constexpr int SW_L2 = 128, SW_MID = 9 * 2 * 64, SW_L5 = SW_L2 + 3 * SW_MID, SW_TOTAL = SW_L5 + 9 * 64;
static_assert(SLDS % 16 == 0, "rings are cleared in 16-byte units");

__device__ __forceinline__ f32x4 mfma32(u32x4 a, u32x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                    0, 0);
}

// bias + ReLU + bf16 of one accumulator (channels 4g .. 4g+3 of a 16-channel half), packed.
__device__ __forceinline__ uint2 relu_pack(f32x4 acc, const float b[4], bool keep)
{
    uint16_t v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float r = acc[k] + b[k];
        r = r > 0.0f ? r : 0.0f;
        v[k] = f2bf(r);
    }
    const uint2 p = make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16), (uint32_t)v[2] | ((uint32_t)v[3] << 16));
    return keep ? p : make_uint2(0u, 0u);
}

// Layers 2-4 (l = 2, 3, 4): one wave per half (mh) of the 32 output channels, 5 tiles.
__device__ __forceinline__ void stream_mid(uint8_t *lds, const u32x4 *__restrict__ wts, const float *__restrict__ params,
                                           int l, int mh, int bx0, int y0, int H, int W, int T, int lane)
{
    const int n = lane & 15, g = lane >> 4;
    u32x4 w[9];
#pragma unroll
    for (int t = 0; t < 9; t++) w[t] = wts[SW_L2 + (l - 2) * SW_MID + (t * 2 + mh) * 64 + lane];
    float bias[4];
#pragma unroll
    for (int k = 0; k < 4; k++) bias[k] = params[PL.b[l - 1] + mh * 16 + 4 * g + k];
    const uint8_t *in = lds + (l - 2) * SRING;
    uint8_t *outr = lds + (l - 1) * SRING + (mh * 16 + 4 * g) * 2;
    const int xo = bx0 - (5 - l);   // plane column of ring pixel 0 of this layer's output
    const int j0 = y0 - 1 - 2 * l;  // input row taken in at step 0
    f32x4 acc[3][ST];
#pragma unroll
    for (int s = 0; s < 3; s++)
#pragma unroll
        for (int i = 0; i < ST; i++) acc[s][i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    auto step = [&](int t, auto PC) {
        constexpr int P = decltype(PC)::value;
        constexpr int S0 = (P + 1) % 3, S1 = P, S2 = (P + 2) % 3;  // rows j + 1, j, j - 1
        const uint8_t *rin = in + ((t + 1) & 1) * SROW + g * 16;
        uint8_t *rout = outr + (t & 1) * SROW;
        const int y = j0 + t - 1;
        const bool rowok = y >= 0 && y < H;
#pragma unroll
        for (int i = 0; i < ST; i++) {
#pragma unroll
            for (int dx = 0; dx < 3; dx++) {
                const u32x4 bv = *reinterpret_cast<const u32x4 *>(rin + (i * 16 + n + dx) * SSTR);
                acc[S0][i] = mfma32(w[dx], bv, acc[S0][i]);
                acc[S1][i] = mfma32(w[3 + dx], bv, acc[S1][i]);
                acc[S2][i] = mfma32(w[6 + dx], bv, acc[S2][i]);
            }
            // Row j - 1 of this tile is complete: out while the next tiles' products run.
            const int x = xo + i * 16 + n;
            *reinterpret_cast<uint2 *>(rout + (i * 16 + n) * SSTR) =
                relu_pack(acc[S2][i], bias, rowok && x >= 0 && x < W);
            acc[S2][i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
        __syncthreads();
    };
    for (int t = 0; t < T; t += 3) {
        step(t, IC<0>{});
        step(t + 1, IC<1>{});
        step(t + 2, IC<2>{});
    }
}

// Layer 1 and the feature staging: one wave.
__device__ __forceinline__ void stream_first(uint8_t *lds, const float *__restrict__ x0, const u32x4 *__restrict__ wts,
                                             const float *__restrict__ params, int bx0, int y0, int H, int W, int T,
                                             const float *__restrict__ quant, int lane)
{
    const int n = lane & 15, g = lane >> 4;
    u32x4 w1[2];
#pragma unroll
    for (int mh = 0; mh < 2; mh++) w1[mh] = wts[mh * 64 + lane];
    float b1[2][4];
#pragma unroll
    for (int mh = 0; mh < 2; mh++)
#pragma unroll
        for (int k = 0; k < 4; k++) b1[mh][k] = params[PL.b[0] + mh * 16 + 4 * g + k];
    uint32_t *feat = reinterpret_cast<uint32_t *>(lds + 4 * SRING);  // [4][SPX]
    uint32_t *qmap = reinterpret_cast<uint32_t *>(lds + 4 * SRING + SFEAT);
    uint8_t *r1 = lds;                                               // layer 1's output ring
    qmap[lane] = (uint32_t)f2bf(quant[lane] * FEAT) << 16;  // read only by this wave

    // Feature row f is plane row y0 - 5 + f; ring pixel p is plane column bx0 - 5 + p.
    auto load_row = [&](int f, float v[2]) {
        const int y = min(max(y0 - 5 + f, 0), H - 1);
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int x = min(max(bx0 - 5 + lane + 64 * k, 0), W - 1);
            v[k] = x0[(size_t)y * W + x];
        }
    };
    auto store_row = [&](int f, const float v[2]) {
        const int y = y0 - 5 + f;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int p = lane + 64 * k;
            const int x = bx0 - 5 + p;
            const bool ok = y >= 0 && y < H && x >= 0 && x < W;
            if (p < SPX) feat[(f & 3) * SPX + p] = ok ? ((uint32_t)f2bf(v[k] * FEAT) | qmap[qmap_index(y, x)]) : 0u;
        }
    };
    float pf[2];
#pragma unroll
    for (int f = 0; f < 3; f++) {
        load_row(f, pf);
        store_row(f, pf);
    }
    load_row(3, pf);

    // B operand: lane (n, g) holds k = 8g .. 8g+7, i.e. taps 4g .. 4g+3 with their two features
    // (k = 2 tap + feature); taps past 8 are zero.
    int tdy[4], tdx[4];
    bool tok[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int tap = 4 * g + q;
        tok[q] = tap < 9;
        tdy[q] = tap / 3;
        tdx[q] = tap % 3;
    }
    for (int t = 0; t < T; t++) {
        // Output row y0 - 4 + t from feature rows t .. t + 2.
        const int y = y0 - 4 + t;
        const bool rowok = y >= 0 && y < H;
        uint8_t *rout = r1 + (t & 1) * SROW;
#pragma unroll
        for (int i = 0; i < ST; i++) {
            u32x4 bv;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t v = feat[((t + tdy[q]) & 3) * SPX + i * 16 + n + tdx[q]];
                bv[q] = tok[q] ? v : 0u;
            }
            const int x = bx0 - 4 + i * 16 + n;
            const bool ok = rowok && x >= 0 && x < W;
#pragma unroll
            for (int mh = 0; mh < 2; mh++) {
                const f32x4 a = mfma32(w1[mh], bv, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
                *reinterpret_cast<uint2 *>(rout + (i * 16 + n) * SSTR + (mh * 16 + 4 * g) * 2) =
                    relu_pack(a, b1[mh], ok);
            }
        }
        // Next feature row in (loaded one step ahead), the one after that requested.
        store_row(t + 3, pf);
        load_row(t + 4, pf);
        __syncthreads();
    }
}

// Layer 5 and the output: one wave.  Takes in layer 4's row y0 - 11 + t at step t; output row
// y0 - 12 + t is then complete.
__device__ __forceinline__ void stream_last(uint8_t *lds, const float *__restrict__ x0, const u32x4 *__restrict__ wts,
                                            const float *__restrict__ params, float *__restrict__ out, int bx0, int y0,
                                            int RH, int H, int W, int T, int lane)
{
    const int n = lane & 15, g = lane >> 4;
    u32x4 w5[9];
#pragma unroll
    for (int t = 0; t < 9; t++) w5[t] = wts[SW_L5 + t * 64 + lane];
    const float b5 = params[PL.b[4]];
    const uint8_t *r4 = lds + 3 * SRING + g * 16;  // layer 4's output ring
    f32x4 acc5[3][ST - 1];
#pragma unroll
    for (int s = 0; s < 3; s++)
#pragma unroll
        for (int i = 0; i < ST - 1; i++) acc5[s][i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    auto step = [&](int t, auto PC) {
        constexpr int P = decltype(PC)::value;
        constexpr int S0 = (P + 1) % 3, S1 = P, S2 = (P + 2) % 3;
        const uint8_t *rin = r4 + ((t + 1) & 1) * SROW;
        const int y = y0 - 12 + t;
        const bool rowok = y >= y0 && y < y0 + RH && y < H;
#pragma unroll
        for (int i = 0; i < ST - 1; i++) {
#pragma unroll
            for (int dx = 0; dx < 3; dx++) {
                const u32x4 bv = *reinterpret_cast<const u32x4 *>(rin + (i * 16 + n + dx) * SSTR);
                acc5[S0][i] = mfma32(w5[dx], bv, acc5[S0][i]);
                acc5[S1][i] = mfma32(w5[3 + dx], bv, acc5[S1][i]);
                acc5[S2][i] = mfma32(w5[6 + dx], bv, acc5[S2][i]);
            }
            const int x = bx0 + i * 16 + n;
            if (g == 0 && rowok && x < W) {
                const size_t p = (size_t)y * W + x;
                out[p] = (x0[p] * FEAT + (acc5[S2][i][0] + b5)) * 64.0f;
            }
            acc5[S2][i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
        __syncthreads();
    };
    for (int t = 0; t < T; t += 3) {
        step(t, IC<0>{});
        step(t + 1, IC<1>{});
        step(t + 2, IC<2>{});
    }
}

// Grid: (bands of SB columns, segments of RH rows); RH a multiple of 3.  Eight waves: wave 0
// layer 1, waves 1-6 layers 2-4 (two channel halves each), wave 7 layer 5.
constexpr int SCT = 512;
// Up to SPLANES planes per launch (blockIdx.z), so one plane's last workgroups overlap the next
// plane's first.
constexpr int SPLANES = 4;
struct StreamPlanes {
    const float *x0[SPLANES];
    float *out[SPLANES];
    float quant[SPLANES][64];
};

__global__ __launch_bounds__(SCT, 4) void k_cnn_stream(const StreamPlanes pp, const u32x4 *__restrict__ wts,
                                                       const float *__restrict__ params, int H, int W, int RH)
{
    const float *__restrict__ x0 = pp.x0[blockIdx.z];
    float *__restrict__ out = pp.out[blockIdx.z];
    const float *quant = pp.quant[blockIdx.z];
    __shared__ __attribute__((aligned(16))) uint8_t lds[SLDS];
    for (int i = threadIdx.x; i < SLDS / 16; i += SCT) reinterpret_cast<u32x4 *>(lds)[i] = u32x4{0u, 0u, 0u, 0u};
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int bx0 = blockIdx.x * SB, y0 = blockIdx.y * RH, T = RH + SFILL;
    if (wave == 0) stream_first(lds, x0, wts, params, bx0, y0, H, W, T, quant, lane);
    else if (wave == 7) stream_last(lds, x0, wts, params, out, bx0, y0, RH, H, W, T, lane);
    else stream_mid(lds, wts, params, 2 + (wave - 1) / 2, (wave - 1) & 1, bx0, y0, H, W, T, lane);
}

// Weights of all five layers into the streaming kernel's operand layouts.
__global__ __launch_bounds__(CT) void k_pack_stream(const float *__restrict__ params, u32x4 *__restrict__ dst)
{
    const int i = blockIdx.x * CT + threadIdx.x;
    if (i >= SW_TOTAL) return;
    const int lane = i & 63, m16 = lane & 15, g = lane >> 4;
    float v[8];
    if (i < SW_L2) {  // layer 1, [mh][lane]: k = 2 tap + feature
        const int m = (i >> 6) * 16 + m16;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int k = 8 * g + e, tap = k >> 1, ch = k & 1;
            v[e] = tap < 9 ? params[PL.w[0] + (m * 2 + ch) * 9 + tap] : 0.0f;
        }
    } else if (i < SW_L5) {  // layers 2-4, [tap][mh][lane]
        const int r = i - SW_L2, l = 1 + r / SW_MID, tap = (r % SW_MID) / 128, m = ((r % 128) >> 6) * 16 + m16;
#pragma unroll
        for (int e = 0; e < 8; e++) v[e] = params[PL.w[l] + (m * 32 + 8 * g + e) * 9 + tap];
    } else {  // layer 5, [tap][lane]: output channel 0
        const int tap = (i - SW_L5) >> 6;
#pragma unroll
        for (int e = 0; e < 8; e++) v[e] = m16 == 0 ? params[PL.w[4] + (8 * g + e) * 9 + tap] : 0.0f;
    }
    u32x4 o;
#pragma unroll
    for (int j = 0; j < 4; j++) o[j] = (uint32_t)f2bf(v[2 * j]) | ((uint32_t)f2bf(v[2 * j + 1]) << 16);
    dst[i] = o;
}

}  // namespace

struct artefact_cnn {
    int device = 0;
    hipStream_t stream = nullptr;
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;  // PL.total each
    bf16x4 *wf[NLAYER] = {}, *wb[NLAYER] = {};  // forward / backward-data operands
    u32x4 *ws = nullptr;                        // operands of the streaming inference kernel
    bool ws_stale = true;                       // params changed since ws was packed
    int cus = 0;                                // compute units of the device
    int adam_t = 0;
    // workspace for a batch of B planes of H x W
    size_t cap_px = 0;
    uint16_t *act[NLAYER] = {};  // act[0] = features (16 ch), act[l] = output of layer l-1 (32 ch)
    uint16_t *g32a = nullptr, *g32b = nullptr, *g16 = nullptr;
    float *resid = nullptr, *pred = nullptr, *x0 = nullptr, *target = nullptr;
    float *quant = nullptr;  // [64 planes][64]: the tables of a batch
    float *part = nullptr;
    double *lpart = nullptr;
    int wgroups = 0;
    // resident training set
    float *ds_x0 = nullptr, *ds_target = nullptr, *ds_q = nullptr;
    uint32_t ds_count = 0, ds_h = 0, ds_w = 0;
};

namespace {

bool ensure_workspace(artefact_cnn *net, size_t px)
{
    if (px <= net->cap_px) return true;
    // Every pointer is cleared as it is released, and the capacity is 0 until the whole set is
    // allocated again: after a failed allocation nothing holds a released address, so a later
    // call (or artefact_cnn_destroy) never frees one twice.
    net->cap_px = 0;
    for (int l = 0; l < NLAYER; l++) {
        if (net->act[l]) (void)hipFree(net->act[l]);
        net->act[l] = nullptr;
    }
    void **bufs[] = {(void **)&net->g32a, (void **)&net->g32b, (void **)&net->g16, (void **)&net->resid,
                     (void **)&net->pred, (void **)&net->x0, (void **)&net->target, (void **)&net->lpart};
    for (void **p : bufs) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    bool ok = true;
    ok = ok && hipMalloc(&net->act[0], px * 16 * 2) == hipSuccess;
    for (int l = 1; l < NLAYER; l++) ok = ok && hipMalloc(&net->act[l], px * 32 * 2) == hipSuccess;
    ok = ok && hipMalloc(&net->g32a, px * 32 * 2) == hipSuccess;
    ok = ok && hipMalloc(&net->g32b, px * 32 * 2) == hipSuccess;
    ok = ok && hipMalloc(&net->g16, px * 16 * 2) == hipSuccess;
    ok = ok && hipMalloc(&net->resid, px * 4) == hipSuccess;
    ok = ok && hipMalloc(&net->pred, px * 4) == hipSuccess;
    ok = ok && hipMalloc(&net->x0, px * 4) == hipSuccess;
    ok = ok && hipMalloc(&net->target, px * 4) == hipSuccess;
    ok = ok && hipMalloc(&net->lpart, blocks(px) * sizeof(double)) == hipSuccess;
    net->cap_px = ok ? px : 0;
    // A failed allocation leaves its error as the thread's last HIP error, which the next launch
    // check (hipGetLastError) would report as its own: clear it here, where it is handled.
    if (!ok) (void)hipGetLastError();
    return ok;
}

void repack(artefact_cnn *net)
{
    net->ws_stale = true;
    for (int l = 0; l < NLAYER; l++) {
        const int nf = 9 * (CINP[l] / 16) * (COUTP[l] / 16) * 64;
        hipLaunchKernelGGL(k_pack_weights, dim3(blocks(nf)), dim3(CT), 0, net->stream, net->params + PL.w[l],
                           WIDTH[l], WIDTH[l + 1], CINP[l], COUTP[l], 0, net->wf[l]);
        hipLaunchKernelGGL(k_pack_weights, dim3(blocks(nf)), dim3(CT), 0, net->stream, net->params + PL.w[l],
                           WIDTH[l], WIDTH[l + 1], CINP[l], COUTP[l], 1, net->wb[l]);
    }
}

template <int CI, int CO, int MODE, bool FEAT_IN = false>
void conv(const uint16_t *in, const bf16x4 *w, const float *bias, uint16_t *out, float *pred,
          const float *resid, int B, int H, int W, hipStream_t s, const QuantTable &qt = QuantTable{})
{
    hipLaunchKernelGGL((k_conv<CI, CO, MODE, FEAT_IN>), dim3((W + 63) / 64, (H + RT - 1) / RT, B), dim3(CT), 0, s,
                       in, w, bias, out, pred, resid, H, W, qt);
}

// Forward pass on the features in act[0] / resid; the prediction (feature units) in pred.
void forward(artefact_cnn *net, int B, int H, int W, hipStream_t s)
{
    conv<16, 32, MODE_RELU>(net->act[0], net->wf[0], net->params + PL.b[0], net->act[1], nullptr, nullptr, B, H,
                            W, s);
    conv<32, 32, MODE_RELU>(net->act[1], net->wf[1], net->params + PL.b[1], net->act[2], nullptr, nullptr, B, H,
                            W, s);
    conv<32, 32, MODE_RELU>(net->act[2], net->wf[2], net->params + PL.b[2], net->act[3], nullptr, nullptr, B, H,
                            W, s);
    conv<32, 32, MODE_RELU>(net->act[3], net->wf[3], net->params + PL.b[3], net->act[4], nullptr, nullptr, B, H,
                            W, s);
    conv<32, 16, MODE_FINAL>(net->act[4], net->wf[4], net->params + PL.b[4], nullptr, net->pred, net->resid, B,
                             H, W, s);
}

template <int CI, int CO>
void wgrad(artefact_cnn *net, int l, const uint16_t *act, const uint16_t *gout, int B, int H, int W)
{
    hipLaunchKernelGGL((k_wgrad<CI, CO>), dim3(net->wgroups), dim3(CT), 0, net->stream, act, gout, net->part, B, H, W);
    const int n = WIDTH[l + 1] * WIDTH[l] * 9 + WIDTH[l + 1];
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(blocks(n)), dim3(CT), 0, net->stream, (const float *)net->part,
                       net->wgroups, CI, CO, WIDTH[l], WIDTH[l + 1], net->grads + PL.w[l], net->grads + PL.b[l]);
}

// Loss and gradients for the batch whose features, residual and target are on the device.
bool backward(artefact_cnn *net, int B, int H, int W, double *loss)
{
    const size_t px = (size_t)B * H * W;
    const unsigned nb = blocks(px);
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(CT), 0, net->stream, (const float *)net->pred,
                       (const float *)net->target, px, net->g16, net->lpart);
    // layer 4 (32 -> 16 padded): weights from act[4]; data gradient into g32a
    wgrad<32, 16>(net, 4, net->act[4], net->g16, B, H, W);
    conv<16, 32, MODE_LINEAR>(net->g16, net->wb[4], nullptr, net->g32a, nullptr, nullptr, B, H, W,
                              net->stream);
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks(px * 32)), dim3(CT), 0, net->stream, net->g32a,
                       (const uint16_t *)net->act[4], px * 32);
    uint16_t *g = net->g32a, *gn = net->g32b;
    for (int l = 3; l >= 1; l--) {
        wgrad<32, 32>(net, l, net->act[l], g, B, H, W);
        conv<32, 32, MODE_LINEAR>(g, net->wb[l], nullptr, gn, nullptr, nullptr, B, H, W, net->stream);
        hipLaunchKernelGGL(k_relu_bwd, dim3(blocks(px * 32)), dim3(CT), 0, net->stream, gn,
                           (const uint16_t *)net->act[l], px * 32);
        std::swap(g, gn);
    }
    wgrad<16, 32>(net, 0, net->act[0], g, B, H, W);
    std::vector<double> lp(nb);
    bool ok = hipMemcpyAsync(lp.data(), net->lpart, nb * sizeof(double), hipMemcpyDeviceToHost, net->stream) ==
              hipSuccess;
    ok = ok && hipStreamSynchronize(net->stream) == hipSuccess;
    double s = 0.0;
    for (double v : lp) s += v;
    if (loss) *loss = s / (double)px;
    return ok;
}

bool upload_batch(artefact_cnn *net, const float *x0, const float *q, const float *target, int B, int H, int W,
                  bool device)
{
    const size_t px = (size_t)B * H * W;
    if (!ensure_workspace(net, px)) return false;
    const hipMemcpyKind kind = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    bool ok = hipMemcpyAsync(net->x0, x0, px * 4, kind, net->stream) == hipSuccess;
    ok = ok && hipMemcpyAsync(net->quant, q, (size_t)B * 64 * 4, kind, net->stream) == hipSuccess;
    if (target) ok = ok && hipMemcpyAsync(net->target, target, px * 4, kind, net->stream) == hipSuccess;
    hipLaunchKernelGGL(k_pack_input, dim3(blocks(px)), dim3(CT), 0, net->stream, (const float *)net->x0,
                       (const float *)net->quant, (size_t)H * W, W, B, net->act[0], net->resid);
    return ok && hipGetLastError() == hipSuccess;
}

// The layered inference path (one kernel per layer, activations through HBM), kept for
// comparison: selected by ARTEFACT_CNN_LAYERED=1.
bool layered_requested()
{
    const char *e = getenv("ARTEFACT_CNN_LAYERED");
    return e && e[0] == '1';
}

}  // namespace

int cnn_predict_planes(artefact_cnn *net, const float *const *planes, const float *quant, uint32_t nplanes,
                       uint32_t h, uint32_t w, float *const *out, hipStream_t s)
{
    if (!net || !planes || !quant || !out || !nplanes || nplanes > 64 || !h || !w) return ARTEFACT_E_ARGUMENT;
    // A training step returns with its Adam update and weight repacking still queued on the
    // net's stream; the kernels below read those weights on stream s.
    if (hipStreamSynchronize(net->stream) != hipSuccess) return ARTEFACT_E_HIP;
    bool ok = true;
    if (!layered_requested()) {
        if (net->ws_stale) {
            hipLaunchKernelGGL(k_pack_stream, dim3(blocks(SW_TOTAL)), dim3(CT), 0, s, (const float *)net->params,
                               net->ws);
            net->ws_stale = false;
        }
        // Bands of SB columns; row segments sized for about two workgroups per compute unit over
        // all the launch's planes (a multiple of 3 rows, at least 12, so that the pipeline fill
        // stays a modest share).
        const int nbx = ((int)w + SB - 1) / SB;
        for (uint32_t i0 = 0; i0 < nplanes && ok; i0 += SPLANES) {
            StreamPlanes pp{};
            const uint32_t n = std::min<uint32_t>(SPLANES, nplanes - i0);
            int nby = std::max(1, (2 * std::max(net->cus, 1) + nbx * (int)n - 1) / (nbx * (int)n));
            int rh = ((int)h + nby - 1) / nby;
            rh = std::max(12, (rh + 2) / 3 * 3);
            nby = ((int)h + rh - 1) / rh;
            for (uint32_t k = 0; k < n; k++) {
                pp.x0[k] = planes[i0 + k];
                pp.out[k] = out[i0 + k];
                memcpy(pp.quant[k], quant + (size_t)(i0 + k) * 64, 64 * sizeof(float));
            }
            hipLaunchKernelGGL(k_cnn_stream, dim3(nbx, nby, n), dim3(SCT), 0, s, pp, (const u32x4 *)net->ws,
                               (const float *)net->params, (int)h, (int)w, rh);
            ok = hipGetLastError() == hipSuccess;
        }
        return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
    }
    const size_t plane = (size_t)h * w;
    if (!ensure_workspace(net, plane)) return ARTEFACT_E_HIP;
    // The first layer forms the features from the plane while staging its tiles and the last
    // writes the prediction in image units, so no packing kernels run (same arithmetic).
    for (uint32_t i = 0; i < nplanes && ok; i++) {
        QuantTable qt;
        memcpy(qt.q, quant + (size_t)i * 64, sizeof qt.q);
        conv<16, 32, MODE_RELU, true>(reinterpret_cast<const uint16_t *>(planes[i]), net->wf[0],
                                      net->params + PL.b[0], net->act[1], nullptr, nullptr, 1, (int)h, (int)w, s,
                                      qt);
        conv<32, 32, MODE_RELU>(net->act[1], net->wf[1], net->params + PL.b[1], net->act[2], nullptr, nullptr, 1,
                                (int)h, (int)w, s);
        conv<32, 32, MODE_RELU>(net->act[2], net->wf[2], net->params + PL.b[2], net->act[3], nullptr, nullptr, 1,
                                (int)h, (int)w, s);
        conv<32, 32, MODE_RELU>(net->act[3], net->wf[3], net->params + PL.b[3], net->act[4], nullptr, nullptr, 1,
                                (int)h, (int)w, s);
        conv<32, 16, MODE_FINAL_IMAGE>(net->act[4], net->wf[4], net->params + PL.b[4], nullptr, out[i], planes[i], 1,
                                       (int)h, (int)w, s);
        ok = hipGetLastError() == hipSuccess;
    }
    return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

extern "C" {

uint32_t artefact_cnn_param_count(void) { return (uint32_t)PL.total; }

int artefact_cnn_create(const float *params, artefact_cnn **out)
{
    if (!out || !params) return ARTEFACT_E_ARGUMENT;
    *out = nullptr;
    artefact_cnn *net = new (std::nothrow) artefact_cnn;
    if (!net) return ARTEFACT_E_HIP;
    bool ok = hipGetDevice(&net->device) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking) == hipSuccess;
    const size_t pb = (size_t)PL.total * 4;
    ok = ok && hipMalloc(&net->params, pb) == hipSuccess && hipMalloc(&net->grads, pb) == hipSuccess;
    ok = ok && hipMalloc(&net->m, pb) == hipSuccess && hipMalloc(&net->v, pb) == hipSuccess;
    ok = ok && hipMalloc(&net->quant, 64 * 64 * 4) == hipSuccess;
    for (int l = 0; l < NLAYER && ok; l++) {
        const size_t nf = (size_t)9 * (CINP[l] / 16) * (COUTP[l] / 16) * 64 * sizeof(bf16x4);
        ok = hipMalloc(&net->wf[l], nf) == hipSuccess && hipMalloc(&net->wb[l], nf) == hipSuccess;
    }
    ok = ok && hipMalloc(&net->ws, (size_t)SW_TOTAL * sizeof(u32x4)) == hipSuccess;
    ok = ok && hipDeviceGetAttribute(&net->cus, hipDeviceAttributeMultiprocessorCount, net->device) == hipSuccess;
    net->wgroups = 512;
    ok = ok && hipMalloc(&net->part, (size_t)net->wgroups * 10 * 32 * 32 * 4) == hipSuccess;
    ok = ok && hipMemcpyAsync(net->params, params, pb, hipMemcpyHostToDevice, net->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(net->m, 0, pb, net->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(net->v, 0, pb, net->stream) == hipSuccess;
    if (ok) repack(net);
    ok = ok && hipStreamSynchronize(net->stream) == hipSuccess;
    if (!ok) {
        artefact_cnn_destroy(net);
        (void)hipGetLastError();  // handled here: not the next call's launch error
        return ARTEFACT_E_HIP;
    }
    *out = net;
    return ARTEFACT_OK;
}

int artefact_cnn_get_params(artefact_cnn *net, float *out)
{
    if (!net || !out) return ARTEFACT_E_ARGUMENT;
    bool ok = hipMemcpyAsync(out, net->params, (size_t)PL.total * 4, hipMemcpyDeviceToHost, net->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(net->stream) == hipSuccess;
    return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

int artefact_cnn_forward(artefact_cnn *net, const float *x0, const float *quant, uint32_t batch, uint32_t h,
                         uint32_t w, float *pred)
{
    if (!net || !x0 || !quant || !pred || !batch || batch > 64 || !h || !w) return ARTEFACT_E_ARGUMENT;
    const size_t plane = (size_t)h * w, px = batch * plane;
    if (!ensure_workspace(net, px)) return ARTEFACT_E_HIP;
    // The inference path warm_init runs, plane by plane.
    const float *planes[64];
    float *outs[64];
    for (uint32_t i = 0; i < batch; i++) {
        planes[i] = net->x0 + i * plane;
        outs[i] = net->pred + i * plane;
    }
    bool ok = hipMemcpyAsync(net->x0, x0, px * 4, hipMemcpyHostToDevice, net->stream) == hipSuccess;
    ok = ok && cnn_predict_planes(net, planes, quant, batch, h, w, outs, net->stream) == ARTEFACT_OK;
    ok = ok && hipMemcpyAsync(pred, net->pred, px * 4, hipMemcpyDeviceToHost, net->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(net->stream) == hipSuccess;
    return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

int artefact_cnn_gradients(artefact_cnn *net, const float *x0, const float *quant, const float *target,
                           uint32_t batch, uint32_t h, uint32_t w, double *loss, float *grads)
{
    if (!net || !x0 || !quant || !target || !batch || batch > 64 || !h || !w || w % 16) return ARTEFACT_E_ARGUMENT;
    if (!upload_batch(net, x0, quant, target, (int)batch, (int)h, (int)w, false)) return ARTEFACT_E_HIP;
    forward(net, (int)batch, (int)h, (int)w, net->stream);
    if (!backward(net, (int)batch, (int)h, (int)w, loss)) return ARTEFACT_E_HIP;
    if (grads) {
        bool ok = hipMemcpyAsync(grads, net->grads, (size_t)PL.total * 4, hipMemcpyDeviceToHost, net->stream) ==
                  hipSuccess;
        ok = ok && hipStreamSynchronize(net->stream) == hipSuccess;
        if (!ok) return ARTEFACT_E_HIP;
    }
    return ARTEFACT_OK;
}

int artefact_cnn_set_dataset(artefact_cnn *net, const float *x0, const float *quant, const float *target,
                             uint32_t count, uint32_t h, uint32_t w)
{
    if (!net || !x0 || !quant || !target || !count || !h || !w || w % 16) return ARTEFACT_E_ARGUMENT;
    for (float *p : {net->ds_x0, net->ds_target, net->ds_q})
        if (p) (void)hipFree(p);
    net->ds_x0 = net->ds_target = net->ds_q = nullptr;
    const size_t px = (size_t)count * h * w;
    bool ok = hipMalloc(&net->ds_x0, px * 4) == hipSuccess && hipMalloc(&net->ds_target, px * 4) == hipSuccess &&
              hipMalloc(&net->ds_q, (size_t)count * 64 * 4) == hipSuccess;
    ok = ok && hipMemcpyAsync(net->ds_x0, x0, px * 4, hipMemcpyHostToDevice, net->stream) == hipSuccess;
    ok = ok && hipMemcpyAsync(net->ds_target, target, px * 4, hipMemcpyHostToDevice, net->stream) == hipSuccess;
    ok = ok && hipMemcpyAsync(net->ds_q, quant, (size_t)count * 64 * 4, hipMemcpyHostToDevice, net->stream) ==
                   hipSuccess;
    ok = ok && hipStreamSynchronize(net->stream) == hipSuccess;
    if (!ok) (void)hipGetLastError();  // handled here: not the next call's launch error
    net->ds_count = ok ? count : 0;
    net->ds_h = h;
    net->ds_w = w;
    return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

int artefact_cnn_train_step(artefact_cnn *net, const uint32_t *indices, uint32_t batch, float lr, double *loss)
{
    if (!net || !indices || !batch || batch > 64 || !net->ds_count) return ARTEFACT_E_ARGUMENT;
    for (uint32_t i = 0; i < batch; i++)
        if (indices[i] >= net->ds_count) return ARTEFACT_E_ARGUMENT;
    const int H = (int)net->ds_h, W = (int)net->ds_w;
    const size_t plane = (size_t)H * W;
    if (!ensure_workspace(net, plane * batch)) return ARTEFACT_E_HIP;
    bool ok = true;
    for (uint32_t i = 0; i < batch && ok; i++) {
        ok = hipMemcpyAsync(net->x0 + i * plane, net->ds_x0 + indices[i] * plane, plane * 4,
                            hipMemcpyDeviceToDevice, net->stream) == hipSuccess;
        ok = ok && hipMemcpyAsync(net->target + i * plane, net->ds_target + indices[i] * plane, plane * 4,
                                  hipMemcpyDeviceToDevice, net->stream) == hipSuccess;
        ok = ok && hipMemcpyAsync(net->quant + i * 64, net->ds_q + (size_t)indices[i] * 64, 64 * 4,
                                  hipMemcpyDeviceToDevice, net->stream) == hipSuccess;
    }
    // A failed copy leaves the batch incomplete: no step is taken on it.
    if (!ok) return ARTEFACT_E_HIP;
    hipLaunchKernelGGL(k_pack_input, dim3(blocks(plane * batch)), dim3(CT), 0, net->stream, (const float *)net->x0,
                       (const float *)net->quant, plane, W, (int)batch, net->act[0], net->resid);
    forward(net, (int)batch, H, W, net->stream);
    ok = ok && backward(net, (int)batch, H, W, loss);
    net->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)net->adam_t), c2 = 1.0f - powf(b2, (float)net->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(blocks(PL.total)), dim3(CT), 0, net->stream, net->params,
                       (const float *)net->grads, net->m, net->v, PL.total, lr, b1, b2, 1e-8f, c1, c2);
    repack(net);
    ok = ok && hipGetLastError() == hipSuccess;
    return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

void artefact_cnn_destroy(artefact_cnn *net)
{
    if (!net) return;
    if (net->stream) (void)hipStreamSynchronize(net->stream);
    std::vector<void *> bufs{net->params, net->grads, net->m, net->v, net->quant, net->part, net->g32a, net->g32b,
                             net->g16, net->resid, net->pred, net->x0, net->target, net->lpart, net->ds_x0,
                             net->ds_target, net->ds_q, net->ws};
    for (int l = 0; l < NLAYER; l++) {
        bufs.push_back(net->wf[l]);
        bufs.push_back(net->wb[l]);
        bufs.push_back(net->act[l]);
    }
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (net->stream) (void)hipStreamDestroy(net->stream);
    delete net;
}

}  // extern "C"
