// artefact_hip: the reconstruction loop of artefact-lib as hand-written gfx950 kernels.
//
// One iteration of backend/artefact-lib/pipeline_scalar/mod.rs:49-75 is two kernels:
//
//   k_gradient  FISTA extrapolation (mod.rs:54-59) + prob gradient (compute_step_prob.rs)
//               + TV gradient (compute_step_tv.rs) + second-order gradient
//               (compute_step_tv2.rs), in gather form over LDS-staged 16x64 tiles with a
//               2-pixel halo; per-tile partial sums of g^2 in double-double, reduced to
//               the per-channel gradient norm by the last tile to finish.
//   k_project   FISTA extrapolation again (recomputed, never stored), the normalised
//               descent step (compute_step.rs:61-65) and the box projection
//               (compute_projection.rs): chroma box mean, 8x8 DCT, clamp to the
//               quantisation interval, coefficient write-back, 8x8 IDCT, mean re-add.
//
// Both kernels reproduce the reference's f32 operation order (no contraction: built
// with -ffp-contract=off; IEEE division and square root), so their output equals the
// CPU restatement bit for bit once the gradient norm is computed the same way.  The
// gradient norm is the one place where the order differs: the reference folds
// N squares sequentially in f32 (compute_step.rs:54-58); here the squares are summed
// exactly enough in double-double and rounded once, an order-independent value.
//
// Gather form of the scatter loops.  The reference scatters from every source pixel p
// into up to 3 (TV) and 7 (second-order) neighbours, in raster order of p.  Every
// destination q therefore receives its terms in a fixed order:
//   prob(q); TV from q-W, q-1, q; second-order from q-W, q-W+1, q-1, q, q+1, q+W-1, q+W.
// The accumulator starts at +0.0 and can never become -0.0 (x + y == -0.0 needs both
// operands -0.0), so adding a zero term is a no-op: sources outside the image and
// sources whose norm is 0 contribute an explicit 0 instead of a branch, and every
// border rule of compute_step_tv2.rs:96-135 reduces to "the source pixel exists".
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/artefact_hip.h"

namespace {

constexpr int TW = 64;          // tile width (full-resolution pixels)
constexpr int TH = 16;          // tile height
constexpr int NT = 256;         // threads per workgroup
constexpr int HX = TW + 4;      // extrapolated-image tile with a 2-pixel halo
constexpr int HY = TH + 4;
constexpr int RX = TW + 2;      // gradient source positions: tile + 1-pixel ring
constexpr int RY = TH + 2;
constexpr int BSTR = 72;        // LDS floats per 8x8 block: 8 rows of 9 (bank spread)
constexpr int MAXBLK = 16;      // 8x8 blocks per tile at most (luma: 2 x 8)

// Ooura shrtdct 8-point constants, backend/artefact-lib/utils/dct.rs:20-27.
constexpr float C8_1R = 0.49039264020161522456f;
constexpr float C8_1I = 0.09754516100806413392f;
constexpr float C8_2R = 0.46193976625564337806f;
constexpr float C8_2I = 0.19134171618254488586f;
constexpr float C8_3R = 0.41573480615127261854f;
constexpr float C8_3I = 0.27778511650980111237f;
constexpr float C8_4R = 0.35355339059327376220f;
constexpr float W8_4R = 0.70710678118654752440f;

// One 8-point pass of idct8x8s (utils/dct.rs:30-98); v[k] is element k of the column
// (first pass) or row (second pass).
__device__ __forceinline__ void idct8(float v[8])
{
    float x1r = C8_1R * v[1] + C8_1I * v[7];
    float x1i = C8_1R * v[7] - C8_1I * v[1];
    float x3r = C8_3R * v[3] + C8_3I * v[5];
    float x3i = C8_3R * v[5] - C8_3I * v[3];
    float xr = x1r - x3r;
    float xi = x1i + x3i;
    x1r += x3r;
    x3i -= x1i;
    x1i = W8_4R * (xr + xi);
    x3r = W8_4R * (xr - xi);
    xr = C8_2R * v[2] + C8_2I * v[6];
    xi = C8_2R * v[6] - C8_2I * v[2];
    float x0r = C8_4R * (v[0] + v[4]);
    float x0i = C8_4R * (v[0] - v[4]);
    float x2r = x0r - xr;
    float x2i = x0i - xi;
    x0r += xr;
    x0i += xi;
    v[0] = x0r + x1r;
    v[7] = x0r - x1r;
    v[2] = x0i + x1i;
    v[5] = x0i - x1i;
    v[4] = x2r - x3i;
    v[3] = x2r + x3i;
    v[6] = x2i - x3r;
    v[1] = x2i + x3r;
}

// One 8-point pass of dct8x8s (utils/dct.rs:100-169).
__device__ __forceinline__ void dct8(float v[8])
{
    float x0r = v[0] + v[7];
    float x1r = v[0] - v[7];
    float x0i = v[2] + v[5];
    float x1i = v[2] - v[5];
    float x2r = v[4] + v[3];
    float x3r = v[4] - v[3];
    float x2i = v[6] + v[1];
    float x3i = v[6] - v[1];
    float xr = x0r + x2r;
    float xi = x0i + x2i;
    v[0] = C8_4R * (xr + xi);
    v[4] = C8_4R * (xr - xi);
    xr = x0r - x2r;
    xi = x0i - x2i;
    v[2] = C8_2R * xr - C8_2I * xi;
    v[6] = C8_2R * xi + C8_2I * xr;
    xr = W8_4R * (x1i - x3i);
    x1i = W8_4R * (x1i + x3i);
    x3i = x1i - x3r;
    x1i += x3r;
    x3r = x1r - xr;
    x1r += xr;
    v[1] = C8_1R * x1r - C8_1I * x1i;
    v[7] = C8_1R * x1i + C8_1I * x1r;
    v[3] = C8_3R * x3r - C8_3I * x3i;
    v[5] = C8_3R * x3i + C8_3I * x3r;
}

// FISTA extrapolation, pipeline_scalar/mod.rs:56: factor*(x - x_prev) + x.
__device__ __forceinline__ float fista(float x, float xp, float f) { return f * (x - xp) + x; }

// f32::midpoint as Rust std computes it on x86_64: widened to f64, halved, rounded once.
__device__ __forceinline__ float midpoint(float a, float b)
{
    return (float)(((double)a + (double)b) * 0.5);
}

// Rust's f32::clamp(lo, hi): lower bound first, NaN passes through.
__device__ __forceinline__ float clamp_ref(float v, float lo, float hi)
{
    v = v < lo ? lo : v;
    v = v > hi ? hi : v;
    return v;
}

// Double-double accumulation (Knuth two-sum), for an order-insensitive sum of squares.
struct DD {
    double hi, lo;
};
__device__ __forceinline__ void dd_add(DD &a, double x)
{
    double s = a.hi + x;
    double bb = s - a.hi;
    double e = (a.hi - (s - bb)) + (x - bb);
    a.hi = s;
    a.lo += e;
}
__device__ __forceinline__ DD dd_merge(DD a, DD b)
{
    double s = a.hi + b.hi;
    double bb = s - a.hi;
    double e = (a.hi - (s - bb)) + (b.hi - bb);
    e += a.lo + b.lo;
    DD r;
    r.hi = s + e;
    r.lo = e - (r.hi - s);
    return r;
}
__device__ __forceinline__ DD dd_shfl_xor(DD a, int m)
{
    DD r;
    r.hi = __shfl_xor(a.hi, m, 64);
    r.lo = __shfl_xor(a.lo, m, 64);
    return r;
}

struct ChanDev {
    const int16_t *coef;  // M_c, block order
    float *cos;           // M_c, block order
    const float *quant;   // 64, natural order, as f32
    const float *qq;      // 64, quant*quant rounded to f32 (compute_step_prob.rs:29)
    uint32_t rw, rh, bw, bh, hs, vs, resample;
    float alpha_p;        // pweight*2*255*sqrt(2) (compute_step.rs:30); 0 disables the term
};

struct IterArgs {
    uint32_t W, H, C;
    float factor, step, alpha_tv, alpha_tgv;
    const float *x[3];
    const float *xp[3];
    float *g[3];
    float *xnew[3];
    ChanDev ch[3];
    double *partials;        // [workgroups][3][2]
    unsigned int *counter;   // tiles finished; reset by the last one
    float *norms;            // [3]
    double *norm_sq;         // [3]
};

enum { TVS = 0, TVR, TVD, TGS, TGX, TGY, TGD };

// Forward differences at an LDS position of the extrapolated tile, with the
// reference's edge rule (compute_step_tv.rs:48-60): 0 on the last column / row.
__device__ __forceinline__ void fwd_diff(const float (*ys)[HX], int sy, int sx, int py, int px, int W, int H,
                                         float &gx, float &gy)
{
    float v = ys[sy][sx];
    gx = px < W - 1 ? ys[sy][sx + 1] - v : 0.0f;
    gy = py < H - 1 ? ys[sy + 1][sx] - v : 0.0f;
}

// Second-order terms at a source pixel (compute_step_tv2.rs:53-72).
__device__ __forceinline__ void tgv_terms(const float (*ys)[HX], int sy, int sx, int py, int px, int W, int H,
                                          float gx, float gy, float &gxx, float &gyy, float &sym)
{
    float gyx = 0.0f, gxy = 0.0f;
    gxx = 0.0f;
    gyy = 0.0f;
    if (px != 0) {
        float lx, ly;
        fwd_diff(ys, sy, sx - 1, py, px - 1, W, H, lx, ly);
        gxx = gx - lx;
        gyx = gy - ly;
    }
    if (py != 0) {
        float ux, uy;
        fwd_diff(ys, sy - 1, sx, py - 1, px, W, H, ux, uy);
        gyy = gy - uy;
        gxy = gx - ux;
    }
    sym = midpoint(gxy, gyx);
}

template <bool TGV>
__global__ __launch_bounds__(NT) void k_gradient(const IterArgs a)
{
    __shared__ float ys[3][HY][HX];
    __shared__ float nrm[RY][RX];
    __shared__ float nrm2[RY][RX];
    __shared__ float ctb[7][RY][RX];
    __shared__ float pblk[MAXBLK * BSTR];
    __shared__ float probv[TH][TW];
    __shared__ DD red[NT / 64][3];
    __shared__ int is_last;

    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    const int W = a.W, H = a.H, C = a.C;
    const float f = a.factor;

    // 1. Extrapolated image y = FISTA(x, x_prev) on the tile + 2-pixel halo.
    for (int i = tid; i < HY * HX; i += NT) {
        const int ry = i / HX, rx = i - ry * HX;
        const int gy = y0 - 2 + ry, gx = x0 - 2 + rx;
        const bool in = gy >= 0 && gy < H && gx >= 0 && gx < W;
        const size_t p = in ? (size_t)gy * W + gx : 0;
#pragma unroll
        for (int c = 0; c < 3; c++)
            if (c < C) ys[c][ry][rx] = in ? fista(a.x[c][p], a.xp[c][p], f) : 0.0f;
    }
    __syncthreads();

    // 2. Joint norms n(p) (compute_step_tv.rs:62-68) and n2(p) (compute_step_tv2.rs:76-86)
    //    at every source position that reaches the tile; 0 outside the image.
    for (int i = tid; i < RY * RX; i += NT) {
        const int ry = i / RX, rx = i - ry * RX;
        const int py = y0 - 1 + ry, px = x0 - 1 + rx;
        float n = 0.0f, n2 = 0.0f;
        if (py >= 0 && py < H && px >= 0 && px < W) {
            const int sy = ry + 1, sx = rx + 1;
            float nacc = 0.0f, n2acc = 0.0f;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                if (c < C) {
                    float gx, gy;
                    fwd_diff(ys[c], sy, sx, py, px, W, H, gx, gy);
                    nacc += gx * gx;
                    nacc += gy * gy;
                    if (TGV) {
                        float gxx, gyy, sym;
                        tgv_terms(ys[c], sy, sx, py, px, W, H, gx, gy, gxx, gyy, sym);
                        n2acc += gyy * gyy + (2.0f * (sym * sym) + gxx * gxx);
                    }
                }
            }
            n = sqrtf(nacc);
            if (TGV) n2 = sqrtf(n2acc);
        }
        nrm[ry][rx] = n;
        nrm2[ry][rx] = n2;
    }
    __syncthreads();

    DD acc[3];
#pragma unroll
    for (int c = 0; c < 3; c++) acc[c] = DD{0.0, 0.0};

    const int tr = tid >> 4, tc = (tid & 15) * 4;
    const float atv = a.alpha_tv, atg = a.alpha_tgv;

    for (int c = 0; c < C; c++) {
        const ChanDev &ch = a.ch[c];
        // 3a. Per-source contributions of channel c (compute_step_tv.rs:74-86,
        //     compute_step_tv2.rs:96-135), zero where the norm is 0.
        for (int i = tid; i < RY * RX; i += NT) {
            const int ry = i / RX, rx = i - ry * RX;
            const int py = y0 - 1 + ry, px = x0 - 1 + rx;
            float v[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            if (py >= 0 && py < H && px >= 0 && px < W) {
                const int sy = ry + 1, sx = rx + 1;
                float gx, gy;
                fwd_diff(ys[c], sy, sx, py, px, W, H, gx, gy);
                const float n = nrm[ry][rx];
                if (n != 0.0f) {
                    v[TVS] = (atv * -(gx + gy)) / n;
                    v[TVR] = (atv * gx) / n;
                    v[TVD] = (atv * gy) / n;
                }
                if (TGV) {
                    const float n2 = nrm2[ry][rx];
                    if (n2 != 0.0f) {
                        float gxx, gyy, s;
                        tgv_terms(ys[c], sy, sx, py, px, W, H, gx, gy, gxx, gyy, s);
                        v[TGS] = atg * (-(2.0f * gyy + (2.0f * gxx + 2.0f * s)) / n2);
                        v[TGX] = atg * ((s + gxx) / n2);
                        v[TGY] = atg * ((gyy + s) / n2);
                        v[TGD] = atg * ((-s) / n2);
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 7; k++) ctb[k][ry][rx] = v[k];
        }
        // 3b. Prob term (compute_step_prob.rs:14-58): IDCT of (cos - c*q)/q^2 for the
        //     component blocks under this tile, replicated onto the full grid.
        for (int i = tid; i < TH * TW; i += NT) probv[i / TW][i % TW] = 0.0f;
        const int hs = ch.hs, vs = ch.vs;
        const int nbx = TW / hs / 8, nby = TH / vs / 8;
        const int b = tid >> 3, j = tid & 7;
        const int bx = x0 / hs / 8 + b % nbx, by = y0 / vs / 8 + b / nbx;
        const bool pvalid = ch.alpha_p != 0.0f && b < nbx * nby && bx < (int)ch.bw && by < (int)ch.bh;
        float pv[8];
        if (pvalid) {
            const size_t base = ((size_t)by * ch.bw + bx) * 64;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                float d = ch.cos[base + k * 8 + j];
                d -= (float)ch.coef[base + k * 8 + j] * ch.quant[k * 8 + j];
                d /= ch.qq[k * 8 + j];
                pv[k] = d;
            }
            idct8(pv);
#pragma unroll
            for (int k = 0; k < 8; k++) pblk[b * BSTR + k * 9 + j] = pv[k];
        }
        __syncthreads();
        if (pvalid) {
#pragma unroll
            for (int k = 0; k < 8; k++) pv[k] = pblk[b * BSTR + j * 9 + k];
            idct8(pv);
            const int ty = (by * 8 + j) * vs - y0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int tx = (bx * 8 + k) * hs - x0;
                const float val = ch.alpha_p * pv[k];
                for (int sy = 0; sy < vs; sy++)
                    for (int sx = 0; sx < hs; sx++) probv[ty + sy][tx + sx] = val;
            }
        }
        __syncthreads();

        // 3c. Gather in the reference's accumulation order.
        float out[4];
        const int qy = y0 + tr;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int ry = tr + 1, rx = tc + k + 1;
            float gv = 0.0f;
            gv += probv[tr][tc + k];
            gv += ctb[TVD][ry - 1][rx];
            gv += ctb[TVR][ry][rx - 1];
            gv += ctb[TVS][ry][rx];
            if (TGV) {
                gv += ctb[TGY][ry - 1][rx];
                gv += ctb[TGD][ry - 1][rx + 1];
                gv += ctb[TGX][ry][rx - 1];
                gv += ctb[TGS][ry][rx];
                gv += ctb[TGX][ry][rx + 1];
                gv += ctb[TGD][ry + 1][rx - 1];
                gv += ctb[TGY][ry + 1][rx];
            }
            out[k] = gv;
        }
        if (qy < H && x0 + tc < W) {
            *reinterpret_cast<float4 *>(a.g[c] + (size_t)qy * W + x0 + tc) =
                make_float4(out[0], out[1], out[2], out[3]);
#pragma unroll
            for (int k = 0; k < 4; k++) dd_add(acc[c], (double)out[k] * (double)out[k]);
        }
        __syncthreads();
    }

    // 4. Per-tile partial sums, then the last tile reduces all partials in a fixed order.
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        DD v = acc[c];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) v = dd_merge(v, dd_shfl_xor(v, m));
        if (lane == 0) red[wave][c] = v;
    }
    __syncthreads();
    const unsigned int nwg = gridDim.x * gridDim.y;
    const unsigned int wg = blockIdx.y * gridDim.x + blockIdx.x;
    if (tid == 0) {
        for (int c = 0; c < 3; c++) {
            DD v = red[0][c];
            for (int w = 1; w < NT / 64; w++) v = dd_merge(v, red[w][c]);
            __hip_atomic_store(&a.partials[(size_t)wg * 6 + c * 2 + 0], v.hi, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&a.partials[(size_t)wg * 6 + c * 2 + 1], v.lo, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        __threadfence();
        const unsigned int prev = atomicAdd(a.counter, 1u);
        is_last = prev == nwg - 1;
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    DD s[3];
#pragma unroll
    for (int c = 0; c < 3; c++) s[c] = DD{0.0, 0.0};
    for (unsigned int w = tid; w < nwg; w += NT) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            DD v;
            v.hi = __hip_atomic_load(&a.partials[(size_t)w * 6 + c * 2 + 0], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
            v.lo = __hip_atomic_load(&a.partials[(size_t)w * 6 + c * 2 + 1], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
            s[c] = dd_merge(s[c], v);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        DD v = s[c];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) v = dd_merge(v, dd_shfl_xor(v, m));
        if (lane == 0) red[wave][c] = v;
    }
    __syncthreads();
    if (tid == 0) {
        for (int c = 0; c < C; c++) {
            DD v = red[0][c];
            for (int w = 1; w < NT / 64; w++) v = dd_merge(v, red[w][c]);
            const double sum = v.hi + v.lo;
            a.norm_sq[c] = sum;
            a.norms[c] = (float)sqrt(sum);
        }
        atomicExch(a.counter, 0u);
    }
}

__global__ __launch_bounds__(NT) void k_project(const IterArgs a)
{
    __shared__ float yv[TH][TW + 1];
    __shared__ float mv[TH][TW + 1];
    __shared__ float blk[MAXBLK * BSTR];

    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    const int W = a.W, H = a.H, C = a.C;
    const int tr = tid >> 4, tc = (tid & 15) * 4;
    const int qy = y0 + tr, qx = x0 + tc;
    const bool live = qy < H && qx < W;
    const float f = a.factor, step = a.step;

    for (int c = 0; c < C; c++) {
        const ChanDev &ch = a.ch[c];
        const float norm = a.norms[c];
        // Descent step on the extrapolated image (compute_step.rs:61-65).
        float yd[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (live) {
            const size_t p = (size_t)qy * W + qx;
            const float4 x4 = *reinterpret_cast<const float4 *>(a.x[c] + p);
            const float4 p4 = *reinterpret_cast<const float4 *>(a.xp[c] + p);
            const float4 g4 = *reinterpret_cast<const float4 *>(a.g[c] + p);
            const float xs[4] = {x4.x, x4.y, x4.z, x4.w};
            const float ps[4] = {p4.x, p4.y, p4.z, p4.w};
            const float gs[4] = {g4.x, g4.y, g4.z, g4.w};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float y = fista(xs[k], ps[k], f);
                if (norm != 0.0f) y -= step * (gs[k] / norm);
                yd[k] = y;
            }
        }
#pragma unroll
        for (int k = 0; k < 4; k++) yv[tr][tc + k] = yd[k];
        __syncthreads();

        // Component-resolution view of the tile (compute_projection.rs:18-61).
        const int hs = ch.hs, vs = ch.vs;
        const int cw = TW / hs, chh = TH / vs;
        const int cx0 = x0 / hs, cy0 = y0 / vs;
        const int nbx = cw / 8, nby = chh / 8;
        for (int i = tid; i < cw * chh; i += NT) {
            const int ly = i / cw, lx = i - ly * cw;
            if (cx0 + lx < (int)ch.rw && cy0 + ly < (int)ch.rh) {
                float val;
                if (ch.resample) {
                    float m = 0.0f;
                    for (int sy = 0; sy < vs; sy++)
                        for (int sx = 0; sx < hs; sx++) m += yv[ly * vs + sy][lx * hs + sx];
                    m /= (float)(hs * vs);
                    mv[ly][lx] = m;
                    val = m;
                } else {
                    val = yv[ly][lx];
                }
                blk[((ly >> 3) * nbx + (lx >> 3)) * BSTR + (ly & 7) * 9 + (lx & 7)] = val;
            }
        }
        __syncthreads();

        // DCT, clamp to [(c - 1/2) q, (c + 1/2) q], keep cos, IDCT
        // (compute_projection.rs:63-92), 8 threads per block, column pass first.
        const int b = tid >> 3, j = tid & 7;
        const int bx = cx0 / 8 + b % nbx, by = cy0 / 8 + b / nbx;
        const bool bvalid = b < nbx * nby && bx < (int)ch.bw && by < (int)ch.bh;
        float v[8];
        float *bp = blk + b * BSTR;
        if (bvalid) {
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bp[k * 9 + j];
            dct8(v);
#pragma unroll
            for (int k = 0; k < 8; k++) bp[k * 9 + j] = v[k];
        }
        __syncthreads();
        if (bvalid) {
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bp[j * 9 + k];
            dct8(v);
            const size_t base = ((size_t)by * ch.bw + bx) * 64 + j * 8;
            float o[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const float cf = (float)ch.coef[base + k];
                const float q = ch.quant[j * 8 + k];
                o[k] = clamp_ref(v[k], (cf - 0.5f) * q, (cf + 0.5f) * q);
            }
            *reinterpret_cast<float4 *>(ch.cos + base) = make_float4(o[0], o[1], o[2], o[3]);
            *reinterpret_cast<float4 *>(ch.cos + base + 4) = make_float4(o[4], o[5], o[6], o[7]);
#pragma unroll
            for (int k = 0; k < 8; k++) bp[j * 9 + k] = o[k];
        }
        __syncthreads();
        if (bvalid) {
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bp[k * 9 + j];
            idct8(v);
#pragma unroll
            for (int k = 0; k < 8; k++) bp[k * 9 + j] = v[k];
        }
        __syncthreads();
        if (bvalid) {
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bp[j * 9 + k];
            idct8(v);
#pragma unroll
            for (int k = 0; k < 8; k++) bp[j * 9 + k] = v[k];
        }
        __syncthreads();

        // Unbox and re-add the projected mean (compute_projection.rs:94-121).
        if (live) {
            float o[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int ly = tr / vs, lx = (tc + k) / hs;
                float r = yd[k];
                if (cx0 + lx < (int)ch.rw && cy0 + ly < (int)ch.rh) {
                    const float pr = blk[((ly >> 3) * nbx + (lx >> 3)) * BSTR + (ly & 7) * 9 + (lx & 7)];
                    r = ch.resample ? (yd[k] - mv[ly][lx]) + pr : pr;
                }
                o[k] = r;
            }
            *reinterpret_cast<float4 *>(a.xnew[c] + (size_t)qy * W + qx) = make_float4(o[0], o[1], o[2], o[3]);
        }
        __syncthreads();
    }
}

// Start image, step 1: dequantise and IDCT every block into the component plane
// (pipeline_simd_8/coef.rs:88-118: all blocks, then one unboxing).
__global__ __launch_bounds__(NT) void k_init_blocks(const ChanDev ch, float *plane)
{
    __shared__ float blk[(NT / 8) * BSTR];
    const int tid = threadIdx.x;
    const int b = tid >> 3, j = tid & 7;
    const size_t bi = (size_t)blockIdx.x * (NT / 8) + b;
    const bool valid = bi < (size_t)ch.bw * ch.bh;
    float v[8];
    float *bp = blk + b * BSTR;
    if (valid) {
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = (float)ch.coef[bi * 64 + k * 8 + j] * ch.quant[k * 8 + j];
        idct8(v);
#pragma unroll
        for (int k = 0; k < 8; k++) bp[k * 9 + j] = v[k];
    }
    __syncthreads();
    if (valid) {
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = bp[j * 9 + k];
        idct8(v);
        const size_t bx = bi % ch.bw, by = bi / ch.bw;
        float *row = plane + (by * 8 + j) * ch.rw + bx * 8;
#pragma unroll
        for (int k = 0; k < 8; k++) row[k] = v[k];
    }
}

// Start image, step 2: nearest-neighbour upsample onto the grid (coef.rs:69-92),
// x = x_prev = that image (aux.rs:58-62).
__global__ __launch_bounds__(NT) void k_init_upsample(const ChanDev ch, const float *plane, float *x, float *xp,
                                                       uint32_t W, uint32_t H)
{
    const size_t i = (size_t)blockIdx.x * NT + threadIdx.x;
    if (i >= (size_t)W * H) return;
    const uint32_t y = i / W, xx = i % W;
    uint32_t cy = y / ch.vs, cx = xx / ch.hs;
    cy = cy < ch.rh - 1 ? cy : ch.rh - 1;
    cx = cx < ch.rw - 1 ? cx : ch.rw - 1;
    const float v = plane[(size_t)cy * ch.rw + cx];
    x[i] = v;
    xp[i] = v;
}

// cos = c * q (coef.rs:94-104).
__global__ __launch_bounds__(NT) void k_init_cos(const ChanDev ch)
{
    const size_t i = (size_t)blockIdx.x * NT + threadIdx.x;
    if (i >= (size_t)ch.bw * ch.bh * 64) return;
    ch.cos[i] = (float)ch.coef[i] * ch.quant[i & 63];
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace

struct artefact_hip_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t C = 0, W = 0, H = 0, iterations = 0;
    size_t N = 0;
    size_t M[3] = {0, 0, 0};
    float weight = 0.0f, pweight[3] = {0, 0, 0};
    bool tgv = true;
    float alpha_tv = 0.0f, alpha_tgv = 0.0f, step = 0.0f;
    std::vector<float> factors;
    ChanDev ch[3];
    float *buf[2] = {nullptr, nullptr};  // each C*N: x and x_prev, swapped every iteration
    float *g = nullptr;                  // C*N
    float *cos = nullptr;                // sum M_c
    int16_t *coef = nullptr;             // sum M_c
    float *tables = nullptr;             // [3][2][64]: quant, quant^2
    double *partials = nullptr;
    unsigned int *counter = nullptr;
    float *norms = nullptr;
    double *norm_sq = nullptr;
    std::vector<std::vector<float>> init_override;
    dim3 grid;
    int cur = 0;            // buf[cur] holds x
    bool initialized = false;
    bool use_graph = true;
    hipGraphExec_t graph = nullptr;
};

namespace {

int check_inputs(const artefact_params *p, const artefact_channel *ch)
{
    if (!p || !ch) return ARTEFACT_E_ARGUMENT;
    if (p->nchannel < 1 || p->nchannel > 3) return ARTEFACT_E_ARGUMENT;
    if (p->mode != ARTEFACT_MODE_TGV && p->mode != ARTEFACT_MODE_TV) return ARTEFACT_E_ARGUMENT;
    if (p->max_w < 8 || p->max_h < 8 || p->max_w % 8 || p->max_h % 8) return ARTEFACT_E_GEOMETRY;
    for (uint32_t c = 0; c < p->nchannel; c++) {
        const artefact_channel &k = ch[c];
        if (!k.coefs || !k.quant) return ARTEFACT_E_ARGUMENT;
        if (!k.rounded_w || !k.rounded_h || k.rounded_w % 8 || k.rounded_h % 8) return ARTEFACT_E_GEOMETRY;
        if (k.block_w * 8 != k.rounded_w || k.block_h * 8 != k.rounded_h) return ARTEFACT_E_GEOMETRY;
        if ((k.h_samp != 1 && k.h_samp != 2) || (k.v_samp != 1 && k.v_samp != 2)) return ARTEFACT_E_GEOMETRY;
        if ((uint64_t)k.rounded_w * k.h_samp > p->max_w || (uint64_t)k.rounded_h * k.v_samp > p->max_h)
            return ARTEFACT_E_GEOMETRY;
        for (int j = 0; j < 64; j++)
            if (k.quant[j] == 0) return ARTEFACT_E_QUANT;
    }
    return ARTEFACT_OK;
}

void free_plan(artefact_hip_plan *pl)
{
    if (!pl) return;
    DeviceGuard guard(pl->device);
    if (pl->graph) (void)hipGraphExecDestroy(pl->graph);
    (void)hipFree(pl->buf[0]);
    (void)hipFree(pl->buf[1]);
    (void)hipFree(pl->g);
    (void)hipFree(pl->cos);
    (void)hipFree(pl->coef);
    (void)hipFree(pl->tables);
    (void)hipFree(pl->partials);
    (void)hipFree(pl->counter);
    (void)hipFree(pl->norms);
    (void)hipFree(pl->norm_sq);
    if (pl->stream) (void)hipStreamDestroy(pl->stream);
    delete pl;
}

IterArgs make_args(const artefact_hip_plan *pl, int cur, float factor)
{
    IterArgs a;
    memset(&a, 0, sizeof(a));
    a.W = pl->W;
    a.H = pl->H;
    a.C = pl->C;
    a.factor = factor;
    a.step = pl->step;
    a.alpha_tv = pl->alpha_tv;
    a.alpha_tgv = pl->alpha_tgv;
    for (uint32_t c = 0; c < 3; c++) {
        const size_t off = (size_t)(c < pl->C ? c : 0) * pl->N;
        a.x[c] = pl->buf[cur] + off;
        a.xp[c] = pl->buf[1 - cur] + off;
        a.xnew[c] = pl->buf[1 - cur] + off;
        a.g[c] = pl->g + off;
        a.ch[c] = pl->ch[c < pl->C ? c : 0];
    }
    a.partials = pl->partials;
    a.counter = pl->counter;
    a.norms = pl->norms;
    a.norm_sq = pl->norm_sq;
    return a;
}

int launch_iterations(artefact_hip_plan *pl)
{
    int cur = pl->cur;
    for (uint32_t it = 0; it < pl->iterations; it++) {
        const IterArgs a = make_args(pl, cur, pl->factors[it]);
        if (pl->tgv)
            hipLaunchKernelGGL(k_gradient<true>, pl->grid, dim3(NT), 0, pl->stream, a);
        else
            hipLaunchKernelGGL(k_gradient<false>, pl->grid, dim3(NT), 0, pl->stream, a);
        hipLaunchKernelGGL(k_project, pl->grid, dim3(NT), 0, pl->stream, a);
        cur = 1 - cur;
    }
    return hipGetLastError() == hipSuccess ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

}  // namespace

extern "C" {

const char *artefact_hip_strerror(int code)
{
    switch (code) {
    case ARTEFACT_OK: return "ok";
    case ARTEFACT_E_ARGUMENT: return "invalid argument";
    case ARTEFACT_E_GEOMETRY: return "invalid geometry";
    case ARTEFACT_E_QUANT: return "zero entry in a quantisation table";
    case ARTEFACT_E_HIP: return "HIP runtime failure";
    case ARTEFACT_E_STATE: return "plan used out of order";
    default: return "unknown error";
    }
}

int artefact_hip_plan_create(const artefact_params *p, const artefact_channel *chs, artefact_hip_plan **out)
{
    if (!out) return ARTEFACT_E_ARGUMENT;
    *out = nullptr;
    int rc = check_inputs(p, chs);
    if (rc) return rc;
    artefact_hip_plan *pl = new (std::nothrow) artefact_hip_plan();
    if (!pl) return ARTEFACT_E_HIP;
    if (hipGetDevice(&pl->device) != hipSuccess) {
        delete pl;
        return ARTEFACT_E_HIP;
    }
    pl->C = p->nchannel;
    pl->W = p->max_w;
    pl->H = p->max_h;
    pl->N = (size_t)pl->W * pl->H;
    pl->iterations = p->iterations;
    pl->weight = p->mode == ARTEFACT_MODE_TV ? 0.0f : p->weight;
    for (int c = 0; c < 3; c++) pl->pweight[c] = p->pweight[c];
    // With weight 0 every second-order contribution is +-0 and adding it is a no-op
    // on the (never -0) accumulator, so the TV-only kernel is bit-identical.
    pl->tgv = pl->weight != 0.0f;

    // f32 constants exactly as the reference derives them (compute_step_tv.rs:70,
    // compute_step.rs:30,47 with compute_step_tv2.rs:88-89, mod.rs:45-52,71).
    const float Cf = (float)pl->C;
    pl->alpha_tv = 1.0f / sqrtf(Cf);
    pl->alpha_tgv = (pl->weight / sqrtf(2.0f)) / sqrtf(Cf);
    const float radius = sqrtf((float)pl->N) / 2.0f;
    pl->step = radius / sqrtf(1.0f + (float)pl->iterations);
    pl->factors.resize(pl->iterations);
    float term = 1.0f;
    for (uint32_t i = 0; i < pl->iterations; i++) {
        const float nt = (float)(((double)1.0f + (double)sqrtf(4.0f * (term * term) + 1.0f)) / 2.0);
        pl->factors[i] = (term - 1.0f) / nt;
        term = nt;
    }

    if (p->init_override) {
        pl->init_override.resize(pl->C);
        for (uint32_t c = 0; c < pl->C; c++) {
            if (!p->init_override[c]) {
                delete pl;
                return ARTEFACT_E_ARGUMENT;
            }
            pl->init_override[c].assign(p->init_override[c], p->init_override[c] + pl->N);
        }
    }

    DeviceGuard guard(pl->device);
    size_t mtot = 0;
    for (uint32_t c = 0; c < pl->C; c++) {
        pl->M[c] = (size_t)chs[c].rounded_w * chs[c].rounded_h;
        mtot += pl->M[c];
    }
    pl->grid = dim3((pl->W + TW - 1) / TW, (pl->H + TH - 1) / TH);
    const size_t nwg = (size_t)pl->grid.x * pl->grid.y;
    bool ok = hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&pl->buf[0], pl->C * pl->N * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&pl->buf[1], pl->C * pl->N * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&pl->g, pl->C * pl->N * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&pl->cos, mtot * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&pl->coef, mtot * sizeof(int16_t)) == hipSuccess;
    ok = ok && hipMalloc(&pl->tables, 3 * 128 * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&pl->partials, nwg * 6 * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&pl->counter, sizeof(unsigned int)) == hipSuccess;
    ok = ok && hipMalloc(&pl->norms, 3 * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(&pl->norm_sq, 3 * sizeof(double)) == hipSuccess;
    if (!ok) {
        free_plan(pl);
        return ARTEFACT_E_HIP;
    }
    std::vector<float> tables(3 * 128, 1.0f);
    size_t moff = 0;
    for (uint32_t c = 0; c < pl->C; c++) {
        const artefact_channel &k = chs[c];
        for (int j = 0; j < 64; j++) {
            const float q = (float)k.quant[j];
            tables[c * 128 + j] = q;
            tables[c * 128 + 64 + j] = q * q;
        }
        ok = ok && hipMemcpyAsync(pl->coef + moff, k.coefs, pl->M[c] * sizeof(int16_t), hipMemcpyHostToDevice,
                                  pl->stream) == hipSuccess;
        ChanDev &d = pl->ch[c];
        d.coef = pl->coef + moff;
        d.cos = pl->cos + moff;
        d.quant = pl->tables + c * 128;
        d.qq = pl->tables + c * 128 + 64;
        d.rw = k.rounded_w;
        d.rh = k.rounded_h;
        d.bw = k.block_w;
        d.bh = k.block_h;
        d.hs = k.h_samp;
        d.vs = k.v_samp;
        d.resample = (k.rounded_w != pl->W || k.rounded_h != pl->H) ? 1u : 0u;
        d.alpha_p = pl->pweight[c] != 0.0f ? pl->pweight[c] * 2.0f * 255.0f * sqrtf(2.0f) : 0.0f;
        moff += pl->M[c];
    }
    for (uint32_t c = pl->C; c < 3; c++) pl->ch[c] = pl->ch[0];
    ok = ok && hipMemcpyAsync(pl->tables, tables.data(), tables.size() * sizeof(float), hipMemcpyHostToDevice,
                              pl->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(pl->counter, 0, sizeof(unsigned int), pl->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(pl->norms, 0, 3 * sizeof(float), pl->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(pl->norm_sq, 0, 3 * sizeof(double), pl->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(pl->stream) == hipSuccess;
    if (!ok) {
        free_plan(pl);
        return ARTEFACT_E_HIP;
    }
    *out = pl;
    return ARTEFACT_OK;
}

int artefact_hip_plan_init(artefact_hip_plan *pl)
{
    if (!pl) return ARTEFACT_E_ARGUMENT;
    DeviceGuard guard(pl->device);
    bool ok = true;
    for (uint32_t c = 0; c < pl->C; c++) {
        const ChanDev &d = pl->ch[c];
        float *x = pl->buf[0] + c * pl->N, *xp = pl->buf[1] + c * pl->N;
        if (!pl->init_override.empty()) {
            ok = ok && hipMemcpyAsync(x, pl->init_override[c].data(), pl->N * sizeof(float),
                                      hipMemcpyHostToDevice, pl->stream) == hipSuccess;
            ok = ok && hipMemcpyAsync(xp, pl->init_override[c].data(), pl->N * sizeof(float),
                                      hipMemcpyHostToDevice, pl->stream) == hipSuccess;
        } else {
            float *plane = pl->g + c * pl->N;  // scratch: M_c <= N
            const size_t nb = (size_t)d.bw * d.bh;
            hipLaunchKernelGGL(k_init_blocks, dim3((unsigned)((nb + NT / 8 - 1) / (NT / 8))), dim3(NT), 0,
                               pl->stream, d, plane);
            hipLaunchKernelGGL(k_init_upsample, dim3((unsigned)((pl->N + NT - 1) / NT)), dim3(NT), 0, pl->stream,
                               d, (const float *)plane, x, xp, pl->W, pl->H);
        }
        hipLaunchKernelGGL(k_init_cos, dim3((unsigned)((pl->M[c] + NT - 1) / NT)), dim3(NT), 0, pl->stream, d);
    }
    ok = ok && hipGetLastError() == hipSuccess;
    pl->cur = 0;
    pl->initialized = ok;
    return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

int artefact_hip_plan_iterate(artefact_hip_plan *pl)
{
    if (!pl) return ARTEFACT_E_ARGUMENT;
    if (!pl->initialized) return ARTEFACT_E_STATE;
    DeviceGuard guard(pl->device);
    if (pl->iterations == 0) return ARTEFACT_OK;
    int rc = ARTEFACT_OK;
    if (pl->use_graph) {
        if (!pl->graph) {
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(pl->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return ARTEFACT_E_HIP;
            rc = launch_iterations(pl);
            const hipError_t e = hipStreamEndCapture(pl->stream, &g);
            if (rc || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return ARTEFACT_E_HIP;
            }
            const hipError_t ie = hipGraphInstantiate(&pl->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                pl->graph = nullptr;
                return ARTEFACT_E_HIP;
            }
        }
        if (hipGraphLaunch(pl->graph, pl->stream) != hipSuccess) return ARTEFACT_E_HIP;
    } else {
        rc = launch_iterations(pl);
        if (rc) return rc;
    }
    if (pl->iterations % 2) pl->cur = 1 - pl->cur;
    // The iterate leaves x in buf[cur]; a second iterate would restart the FISTA
    // sequence from the current image, which the reference never does.
    pl->initialized = false;
    return ARTEFACT_OK;
}

int artefact_hip_plan_download(artefact_hip_plan *pl, float *const *out)
{
    if (!pl || !out) return ARTEFACT_E_ARGUMENT;
    DeviceGuard guard(pl->device);
    bool ok = true;
    for (uint32_t c = 0; c < pl->C; c++) {
        if (!out[c]) return ARTEFACT_E_ARGUMENT;
        ok = ok && hipMemcpyAsync(out[c], pl->buf[pl->cur] + c * pl->N, pl->N * sizeof(float),
                                  hipMemcpyDeviceToHost, pl->stream) == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(pl->stream) == hipSuccess;
    return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

int artefact_hip_plan_download_cos(artefact_hip_plan *pl, float *const *out)
{
    if (!pl || !out) return ARTEFACT_E_ARGUMENT;
    DeviceGuard guard(pl->device);
    bool ok = true;
    for (uint32_t c = 0; c < pl->C; c++) {
        if (!out[c]) return ARTEFACT_E_ARGUMENT;
        ok = ok && hipMemcpyAsync(out[c], pl->ch[c].cos, pl->M[c] * sizeof(float), hipMemcpyDeviceToHost,
                                  pl->stream) == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(pl->stream) == hipSuccess;
    return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

int artefact_hip_plan_last_norms(artefact_hip_plan *pl, float *out)
{
    if (!pl || !out) return ARTEFACT_E_ARGUMENT;
    DeviceGuard guard(pl->device);
    bool ok = hipMemcpyAsync(out, pl->norms, pl->C * sizeof(float), hipMemcpyDeviceToHost, pl->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(pl->stream) == hipSuccess;
    return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

int artefact_hip_plan_first_gradient(artefact_hip_plan *pl, float *const *out)
{
    if (!pl || !out) return ARTEFACT_E_ARGUMENT;
    if (!pl->initialized) return ARTEFACT_E_STATE;
    DeviceGuard guard(pl->device);
    const IterArgs a = make_args(pl, pl->cur, 0.0f);
    if (pl->tgv)
        hipLaunchKernelGGL(k_gradient<true>, pl->grid, dim3(NT), 0, pl->stream, a);
    else
        hipLaunchKernelGGL(k_gradient<false>, pl->grid, dim3(NT), 0, pl->stream, a);
    bool ok = hipGetLastError() == hipSuccess;
    for (uint32_t c = 0; c < pl->C; c++) {
        if (!out[c]) return ARTEFACT_E_ARGUMENT;
        ok = ok && hipMemcpyAsync(out[c], pl->g + c * pl->N, pl->N * sizeof(float), hipMemcpyDeviceToHost,
                                  pl->stream) == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(pl->stream) == hipSuccess;
    return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

int artefact_hip_plan_profile(artefact_hip_plan *pl, float *avg_ms)
{
    if (!pl || !avg_ms) return ARTEFACT_E_ARGUMENT;
    if (!pl->initialized) return ARTEFACT_E_STATE;
    DeviceGuard guard(pl->device);
    const uint32_t n = pl->iterations;
    std::vector<hipEvent_t> ev(2 * (size_t)n + 1, nullptr);
    bool ok = true;
    for (auto &e : ev) ok = ok && hipEventCreate(&e) == hipSuccess;
    int cur = pl->cur;
    if (ok && n) {
        ok = hipEventRecord(ev[0], pl->stream) == hipSuccess;
        for (uint32_t it = 0; ok && it < n; it++) {
            const IterArgs a = make_args(pl, cur, pl->factors[it]);
            if (pl->tgv)
                hipLaunchKernelGGL(k_gradient<true>, pl->grid, dim3(NT), 0, pl->stream, a);
            else
                hipLaunchKernelGGL(k_gradient<false>, pl->grid, dim3(NT), 0, pl->stream, a);
            ok = ok && hipEventRecord(ev[2 * it + 1], pl->stream) == hipSuccess;
            hipLaunchKernelGGL(k_project, pl->grid, dim3(NT), 0, pl->stream, a);
            ok = ok && hipEventRecord(ev[2 * it + 2], pl->stream) == hipSuccess;
            cur = 1 - cur;
        }
        ok = ok && hipGetLastError() == hipSuccess && hipEventSynchronize(ev[2 * n]) == hipSuccess;
    }
    double tg = 0.0, tp = 0.0;
    for (uint32_t it = 0; ok && it < n; it++) {
        float a = 0.0f, b = 0.0f;
        ok = ok && hipEventElapsedTime(&a, ev[2 * it], ev[2 * it + 1]) == hipSuccess;
        ok = ok && hipEventElapsedTime(&b, ev[2 * it + 1], ev[2 * it + 2]) == hipSuccess;
        tg += a;
        tp += b;
    }
    float total = 0.0f;
    if (ok && n) ok = hipEventElapsedTime(&total, ev[0], ev[2 * n]) == hipSuccess;
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (!ok) return ARTEFACT_E_HIP;
    avg_ms[0] = n ? (float)(tg / n) : 0.0f;
    avg_ms[1] = n ? (float)(tp / n) : 0.0f;
    avg_ms[2] = total;
    pl->cur = cur;
    pl->initialized = false;
    return ARTEFACT_OK;
}

void *artefact_hip_plan_stream(artefact_hip_plan *pl) { return pl ? (void *)pl->stream : nullptr; }

int artefact_hip_plan_set_graph(artefact_hip_plan *pl, int enable)
{
    if (!pl) return ARTEFACT_E_ARGUMENT;
    pl->use_graph = enable != 0;
    return ARTEFACT_OK;
}

void artefact_hip_plan_destroy(artefact_hip_plan *pl)
{
    if (!pl) return;
    {
        DeviceGuard guard(pl->device);
        if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    }
    free_plan(pl);
}

int artefact_hip_compute(const artefact_params *p, const artefact_channel *ch, float *const *out)
{
    if (!out) return ARTEFACT_E_ARGUMENT;
    artefact_hip_plan *pl = nullptr;
    int rc = artefact_hip_plan_create(p, ch, &pl);
    if (rc) return rc;
    pl->use_graph = false;
    rc = artefact_hip_plan_init(pl);
    if (!rc) rc = artefact_hip_plan_iterate(pl);
    if (!rc) rc = artefact_hip_plan_download(pl, out);
    artefact_hip_plan_destroy(pl);
    return rc;
}

}  // extern "C"
