// artefact_hip: the reconstruction loop of artefact-lib as hand-written gfx950 kernels.
//
// One iteration of backend/artefact-lib/pipeline_scalar/mod.rs:49-75 is two kernels:
//
//   k_gradient  FISTA extrapolation (mod.rs:54-59) + TV gradient (compute_step_tv.rs) +
//               second-order gradient (compute_step_tv2.rs) + the prob gradient
//               (compute_step_prob.rs, read from the plane k_project left behind).
//               A register-streaming stencil: each lane owns one image column and walks
//               down a strip of rows; horizontal neighbours come from DPP lane shifts,
//               so there is no LDS staging and no barrier in the stencil.  Each
//               workgroup's sum of g^2 per channel (double-double) is added, exactly, into
//               a fixed-point integer accumulator with 64-bit atomics.
//   k_project   Every workgroup first forms the gradient norm (compute_step.rs:54-58) from
//               the accumulator, then the normalised
//               descent step (compute_step.rs:61-65) and the box projection
//               (compute_projection.rs): chroma box mean, 8x8 DCT, clamp to the
//               quantisation interval, 8x8 IDCT, mean re-add.  It also evaluates the next
//               iteration's prob term, alpha_p * IDCT((cos - c q) / q^2), while the
//               clamped coefficients are still in LDS, so the coefficient estimate itself
//               only goes to HBM on the last iteration.
//
// Both kernels reproduce the reference's f32 operation order (no contraction: built with
// -ffp-contract=off), so their output equals the CPU restatement bit for bit once the
// gradient norm is computed the same way.  The gradient norm is the one place where the
// order differs: the reference folds N squares sequentially in f32 (compute_step.rs:54-58);
// here the squares are summed in f64 per lane, merged in double-double per workgroup, and the
// workgroups' sums are added exactly (fixed point) and rounded once.
//
// Gather form of the scatter loops.  The reference scatters from every source pixel p
// into up to 3 (TV) and 7 (second-order) neighbours, in raster order of p.  Every
// destination q therefore receives its terms in a fixed order:
//   prob(q); TV from q-W, q-1, q; second-order from q-W, q-W+1, q-1, q, q+1, q+W-1, q+W.
// The accumulator starts at +0.0 and can never become -0.0 (x + y == -0.0 needs both
// operands -0.0), so adding a zero term is a no-op: sources outside the image and
// sources whose norm is 0 contribute an explicit 0 instead of a branch, and every
// border rule of compute_step_tv2.rs:96-135 reduces to "the source pixel exists".
//
// Correctly rounded f32 division by a shared denominator.  The reference divides each
// numerator by the joint norm (compute_step_tv.rs:136-143, compute_step_tv2.rs:256-286),
// by q^2 and by the gradient norm.  Here a/b is evaluated as (float)((double)a * r) with
// r a reciprocal of b accurate to ~2^-52.  That equals the IEEE f32 quotient: for f32
// a, b the exact quotient is never an f32 rounding midpoint and lies at least 2^-49
// (relative) away from every midpoint (a/b - m = (a - m b)/b with m = M 2^mu, M an odd
// 25-bit integer, and |a - m b| >= 2^e gives a relative gap >= 1/(M B) > 2^-49 for the
// 24-bit significand B of b), while the f64 product is within ~2^-52 of a/b, so rounding
// it to f32 picks the same neighbour.  All quotients here are finite: the iterates stay
// far below 2^60 in magnitude (int16 coefficients, u16 tables, bounded steps).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <new>
#include <type_traits>
#include <vector>

#include "../../include/artefact_hip.h"
#include "cnn_internal.h"

namespace {

constexpr int NT = 256;         // threads per workgroup
constexpr int WAVES = NT / 64;  // waves per workgroup
constexpr int KG_COLS = 60;     // k_gradient: output columns per wave (64 lanes, 2+2 halo)
constexpr int TW = 64;          // k_project tile width (full-resolution pixels)
constexpr int TH = 16;          // k_project tile height
constexpr int BSTR = 72;        // LDS floats per 8x8 block: 8 rows of 9 (bank spread)
constexpr int MAXBLK = 16;      // 8x8 blocks per k_project tile at most (2 x 8)
constexpr int NXCD = 8;         // accelerator complex dies on gfx950; workgroups are dealt round-robin

// Logical tile of this workgroup such that the tiles one XCD runs form a contiguous band
// of rows: the dispatcher sends workgroup i to XCD i % NXCD, and every XCD has its own L2,
// so neighbouring strips (which share halo rows and cache lines) should share an XCD.
__device__ __forceinline__ void xcd_tile(unsigned int &tx, unsigned int &ty)
{
    const unsigned int n = gridDim.x * gridDim.y;
    const unsigned int h = blockIdx.y * gridDim.x + blockIdx.x;
    const unsigned int xcd = h % NXCD, j = h / NXCD;
    // XCD k receives ceil((n - k) / NXCD) workgroups; those before it hold the earlier tiles.
    const unsigned int q = n / NXCD, r = n % NXCD;
    const unsigned int base = xcd * q + (xcd < r ? xcd : r);
    const unsigned int l = base + j;
    ty = l / gridDim.x;
    tx = l % gridDim.x;
}

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
// (first pass) or row (second pass).  a + -(b) of the reference is a - b in IEEE.
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

// f32::midpoint, which Rust std computes on x86_64 widened to f64, halved and rounded once:
// (float)(((double)a + (double)b) * 0.5).  In f32 arithmetic, fl32(fl32(a + b) * 0.5) equals
// that widened form for every pair whose sum stays below 2^128.  If a + b is exact in f32, both
// round the exact half once (the f64 sum and halving are exact).  If it is not, |a + b| >=
// 2^-125 (below that both operands are multiples of 2^-149, or Sterbenz applies, and the sum is
// exact), so the f32 half is exact and both forms equal RN32(a + b) / 2: the f64 sum is exact
// unless |b| < 2^-29 |a|, and then both round to a.  The iterates here stay far below 2^128 (see
// the header comment).
__device__ __forceinline__ float midpoint_f32(float a, float b) { return (a + b) * 0.5f; }

// Rust's f32::clamp(lo, hi): lower bound first, NaN passes through.
__device__ __forceinline__ float clamp_ref(float v, float lo, float hi)
{
    v = v < lo ? lo : v;
    v = v > hi ? hi : v;
    return v;
}

// Reciprocal of a positive finite f32 in f64, accurate to ~2^-52: the hardware estimate r
// (v_rcp_f64, relative error e = 1 - b r of order 2^-23) refined by one third-order step
// r (1 + e + e^2), whose truncation error e^3 (~2^-69) is far below the final rounding.
__device__ __forceinline__ double recip64(float b)
{
    const double d = (double)b;
    const double r = __builtin_amdgcn_rcp(d);
    const double e = __builtin_fma(-d, r, 1.0);
    const double t = __builtin_fma(e, e, e);
    return __builtin_fma(r, t, r);
}

// a / b for f32 a and b, given r ~ 1/b (see the header comment): bit-identical to IEEE.
__device__ __forceinline__ float qdiv(float a, double r) { return (float)((double)a * r); }

// Value of the same register in lane - 1 (wave_shr:1) / lane + 1 (wave_shl:1); the edge
// lane reads 0.  Both are DPP modifiers on a v_mov, not LDS traffic.
__device__ __forceinline__ float from_left(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float from_right(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
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

enum ProjMode : uint32_t { PM_FULL = 0, PM_SUB11 = 1, PM_SUB22 = 2, PM_SUB21 = 3, PM_SUB12 = 4 };
// Channel layouts with a fused projection kernel; anything else runs channel by channel.
enum Layout : int { LAYOUT_GENERIC = 0, LAYOUT_Y = 1, LAYOUT_444 = 2, LAYOUT_420 = 3 };

struct ChanDev {
    const int16_t *coef;  // M_c, block order
    float *cos;           // M_c, block order (written on the last iteration)
    float *prob;          // M_c, component raster: alpha_p * IDCT((cos - c q) / q^2)
    const float *quant;   // 64, natural order, as f32
    const double *rqq;    // 64, 1 / (quant*quant rounded to f32), correctly rounded f64
    uint32_t rw, rh, bw, bh, hs, vs, resample, mode;
    float alpha_p;        // pweight*2*255*sqrt(2) (compute_step.rs:30); 0 disables the term
};

struct IterArgs {
    uint32_t W, H, C, write_cos;
    uint32_t prob_bytes;     // bytes of the prob arena (all channels)
    float factor, step, alpha_tv, alpha_tgv;
    const float *x[3];
    const float *xp[3];
    float *y[3];             // extrapolated image y = FISTA(x, x_prev), written by k_gradient
    float *g[3];
    float *xnew[3];
    ChanDev ch[3];
    long long *acc;          // ACC_WORDS: this iteration's norm accumulator (k_gradient adds)
    long long *acc_next;     // ACC_WORDS: the next iteration's, cleared by the projection
    float *norms;            // [3]
    double *rnorm;           // [3] 1/norm, correctly rounded f64
    double *norm_sq;         // [3]
    uint32_t kg_rows;        // k_gradient output rows per wave
    uint32_t xcd_bands;      // projection workgroups on the XCD that swept their rows (see k_project_lane)
};

// ---- Norm accumulator ----
// The sum of g^2 per channel over all k_gradient workgroups, kept exactly as a fixed-point
// integer whose unit is 2^-352: ACC_DIGITS digits of 32 bits, digit d of weight 2^(32 d - 352),
// each held in an int64 that takes signed 32-bit chunks by atomic addition (carries are
// resolved by the reader), plus a flag word counting values that do not fit.  Every value a
// workgroup adds is a multiple of 2^-298 (g is an f32, a multiple of 2^-149, so g^2 and every
// rounded sum or double-double error term of such squares is too) and below 2^288 (|g| <
// 2^128 and at most 2^29 pixels), so it fits the 640 bits exactly.  Integer addition is
// associative: the total does not depend on the order in which the workgroups finish.
// ACC_SLOTS copies spread the workgroups' atomics over more addresses.
constexpr int ACC_DIGITS = 20;
constexpr int ACC_CH = ACC_DIGITS + 1;  // digits, then the flag
constexpr int ACC_SLOTS = 8;
constexpr int ACC_WORDS = ACC_SLOTS * 3 * ACC_CH;  // one accumulator (4032 bytes)

// Adds chunk j (0..2) of the finite double x to the accumulator of one channel (ACC_CH words).
__device__ __forceinline__ void acc_add(long long *acc, double x, int j)
{
    if (x == 0.0) return;
    const unsigned long long bits = (unsigned long long)__double_as_longlong(x);
    const int ex = (int)((bits >> 52) & 0x7ff);
    // The value range above: 2^-298 <= |x| < 2^288 (biased exponents 725 .. 1310).  Anything
    // else (a non-finite x included) is counted in the flag, which makes the norm NaN.
    if (ex < 723 || ex > 1310) {
        if (j == 0) __hip_atomic_fetch_add(acc + ACC_DIGITS, 1LL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    const unsigned long long m = (bits & ((1ull << 52) - 1)) | (1ull << 52);
    const int pos = ex - 1075 + 352;  // bit position of m's lowest bit, >= 0
    const int q = pos >> 5, sh = pos & 31;
    const unsigned long long lo = m << sh, hi = sh ? m >> (64 - sh) : 0ull;
    const unsigned long long chunk = j == 0 ? (lo & 0xffffffffull) : (j == 1 ? lo >> 32 : hi);
    const int d = q + j;
    if (chunk == 0 || d >= ACC_DIGITS) return;  // a chunk above the top digit is 0 (|x| < 2^288)
    const long long v = (long long)chunk;
    __hip_atomic_fetch_add(acc + d, (long long)(bits >> 63) ? -v : v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The accumulated sum of one channel, rounded once to f64 (to nearest, ties to even); NaN if a
// value was out of range.  Called by a whole wave: lane d < ACC_DIGITS holds digit d with the
// ACC_SLOTS copies added, lane ACC_DIGITS the flag, the other lanes 0.  The digits' carries are
// resolved in parallel: each round every digit keeps its low 32 bits and hands the signed rest
// to the digit above.  The first round leaves carries of -1, 0 or +1 only, and such a carry
// moves up one digit per round, so at most ACC_DIGITS + 1 rounds have any.  Every lane
// returns the same value.
__device__ __forceinline__ double acc_value(long long v, int lane)
{
    const bool bad = __ballot(lane == ACC_DIGITS && v != 0) != 0;
    long long x = lane < ACC_DIGITS ? v : 0;
    long long top = 0;  // carries out of the top digit, summed
#pragma unroll 1
    for (int r = 0; r <= ACC_DIGITS + 1; r++) {
        const long long carry = x >> 32;  // arithmetic shift: a negative digit borrows
        if (__ballot(carry != 0) == 0) break;
        top += __shfl(carry, ACC_DIGITS - 1, 64);
        const long long cin = __shfl_up(carry, 1, 64);
        x = lane < ACC_DIGITS ? (x & 0xffffffffll) + (lane > 0 ? cin : 0) : 0;
    }
    // A negative total (impossible for sums of squares) leaves a borrow out of the top digit.
    if (bad || top != 0) return __builtin_nan("");
    const unsigned int d32 = (unsigned int)x;
    const unsigned long long nz = __ballot(d32 != 0);
    if (nz == 0) return 0.0;
    const int k = 63 - __builtin_clzll(nz);
    // The top three digits as a 96-bit integer T:R, the digits below as a sticky bit.
    const unsigned int dk = (unsigned int)__builtin_amdgcn_readlane((int)d32, k);
    const unsigned int dk1 = k >= 1 ? (unsigned int)__builtin_amdgcn_readlane((int)d32, k - 1) : 0u;
    unsigned int R = k >= 2 ? (unsigned int)__builtin_amdgcn_readlane((int)d32, k - 2) : 0u;
    const bool sticky = k >= 3 && (nz & ((1ull << (k - 2)) - 1)) != 0;
    unsigned long long T = ((unsigned long long)dk << 32) | dk1;
    const int lz = __builtin_clzll(T);  // < 32: the top digit is non-zero
    if (lz) {
        T = (T << lz) | ((unsigned long long)R >> (32 - lz));
        R <<= lz;
    }
    unsigned long long mant = T >> 11;
    const unsigned int rest = (unsigned int)(T & 0x7ff);
    const bool round = (rest >> 10) & 1u;
    const bool below = (rest & 0x3ffu) != 0 || R != 0 || sticky;
    if (round && (below || (mant & 1ull))) mant++;
    // value = mant * 2^(32 (k - 2) + 32 + 11 - lz) in units of 2^-352
    return ldexp((double)mant, 32 * (k - 2) + 43 - lz - 352);
}

// Raw buffer resource over `bytes` bytes: loads outside return 0, stores outside are dropped.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void *p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}

template <int N>
using IC = std::integral_constant<int, N>;

// k_gradient: C channels, TGV on/off.  Every wave owns a strip of a.kg_rows output rows of
// its 64 columns (2 halo columns each side) and walks rows r0-2 .. r0+kg_rows+1 once: a
// four-step prologue, then a loop unrolled by three so that the three-row ring of
// contributions sits in fixed registers.  The host sizes the strips so that the whole grid
// is resident at once (one round of waves), which keeps the 2-row halo cost per strip low.
// FAST (ARTEFACT_PRECISION_F32_FAST): the same gather, with the per-source arithmetic relaxed:
// v_sqrt_f32 for the joint norms, one v_rcp_f32 per norm and products (alpha / n) * term in
// place of the correctly rounded (alpha * term) / n, and a plain f64 sum of g^2.  Each term is
// then within a few ulp of the exact one instead of equal to it.
// WIDE: the same sweep and arithmetic for grids whose C planes reach 2^31 bytes, where the
// 32-bit buffer offsets below (lane offset plus row and channel offset) could wrap.  Every
// load and store then goes through a buffer resource over one row of one channel's plane,
// based at that row's 64-bit address, with the lane offset alone: the same clamped loads,
// zero loads and dropped stores, with offsets below one row's bytes.
template <int C, bool TGV, bool FAST = false, bool WIDE = false>
__global__ __launch_bounds__(NT, 5) void k_gradient(const IterArgs a)
{
    __shared__ DD lred[WAVES][3][64];
    // Rows shared by the strips of consecutive waves: the 4 rows around each of the workgroup's
    // 3 inner strip boundaries, x and x_prev per channel, handed from the lower strip's wave
    // (its first rows) to the upper one's (its last rows).
    __shared__ float hal[WAVES - 1][4][2 * C][64];

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int W = a.W, H = a.H, RW = (int)a.kg_rows;
    unsigned int tx, ty;
    xcd_tile(tx, ty);
    const int x = (int)tx * KG_COLS - 2 + lane;
    const int r0 = ((int)ty * WAVES + wave) * RW;
    const bool xin = x >= 0 && x < W;
    const bool out_lane = lane >= 2 && lane < 62 && xin;
    const bool xl = x > 0;
    const float f = a.factor, atv = a.alpha_tv, atg = a.alpha_tgv;
    const uint32_t pbytes = (uint32_t)((size_t)W * H * 4);

    // One buffer resource per arena (the C planes of x, x_prev, g and y are contiguous, the
    // prob planes too); channel c sits c * pbytes (prob: poff[c]) bytes in.
    const __amdgpu_buffer_rsrc_t rx = plane_rsrc(a.x[0], pbytes * C);
    const __amdgpu_buffer_rsrc_t rxp = plane_rsrc(a.xp[0], pbytes * C);
    const __amdgpu_buffer_rsrc_t rg = plane_rsrc(a.g[0], pbytes * C);
    const __amdgpu_buffer_rsrc_t ry = plane_rsrc(a.y[0], pbytes * C);
    const __amdgpu_buffer_rsrc_t rpr = plane_rsrc(a.ch[0].prob, a.prob_bytes);
    uint32_t poff[C];
    int hsh[C], vsh[C], crw[C], crh[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        poff[c] = (uint32_t)((const char *)a.ch[c].prob - (const char *)a.ch[0].prob);
        hsh[c] = a.ch[c].hs - 1;
        vsh[c] = a.ch[c].vs - 1;
        crw[c] = a.ch[c].rw;
        crh[c] = a.ch[c].rh;
    }

    // Row state in rings of three, indexed by k mod 3 so that no register moves between
    // steps: y of rows t, t-1, t-2; x-differences of rows t, t-1, t-2; y-differences of
    // source rows s = t-1 and s-1.
    float yring[3][C], dxring[3][C], dyring[3][C], pring[3][C], nx[C], nxp[C];
    // Running sums of the output rows in flight (see the gather below): head[c] holds row
    // s + 1's first two terms, body[c] row s's sum up to its last two terms, and tgy_up[c] /
    // tgd_up[c] the second-order terms of source row s, which row s + 1 takes next step.
    float head[C], body[C], tgy_up[C], tgd_up[C];
    DD acc[C];

    // Addressing: the lane part of every offset is constant and the row and channel part is
    // wave-uniform (the scalar offset).  Loads clamp the column to [0, W-1] and the row to
    // [0, H-1]: the replicated edge makes the reference's zero differences at x = W-1 and
    // y = H-1 come out of the subtractions themselves (y - y = +0), and every value a lane
    // outside the image holds is finite and feeds only contributions that are zeroed.  Stores
    // and prob loads use the arena size as the lane offset outside the image, which the
    // buffer range check (on the lane offset) turns into a dropped store or a zero load.
    const uint32_t arena = pbytes * C;
    const uint32_t vin = (uint32_t)min(max(x, 0), W - 1) * 4u;
    const uint32_t rowb = (uint32_t)W * 4u;  // WIDE: bytes of a row, its out-of-range lane offset
    const uint32_t vout = out_lane ? (uint32_t)x * 4u : (WIDE ? rowb : arena);
    // WIDE: vpr is the offset in a component row, or the row's size (a zero load).
    uint32_t vpr[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        const int cx = x >> hsh[c];
        if constexpr (WIDE)
            vpr[c] = (out_lane && cx < crw[c]) ? (uint32_t)cx * 4u : (uint32_t)crw[c] * 4u;
        else
            vpr[c] = (out_lane && cx < crw[c]) ? poff[c] + (uint32_t)cx * 4u : a.prob_bytes;
    }
    // Halo sharing needs the hand-over (steps 0-3 of the lower wave, then a barrier) to come
    // before the first row the upper wave takes (step RW - 1).
    const bool share = RW >= 5;
    auto rowload = [&](int t) {
        const int kk = t - (r0 - 2);  // this strip's row index, 0 .. RW + 3
        if (share && wave < WAVES - 1 && kk >= RW) {
#pragma unroll
            for (int c = 0; c < C; c++) {
                nx[c] = hal[wave][kk - RW][2 * c][lane];
                nxp[c] = hal[wave][kk - RW][2 * c + 1][lane];
            }
            return;
        }
        const uint32_t tr = (uint32_t)min(max(t, 0), H - 1);
#pragma unroll
        for (int c = 0; c < C; c++) {
            if constexpr (WIDE) {
                const size_t ro = (size_t)tr * (size_t)W;
                const __amdgpu_buffer_rsrc_t r1 = plane_rsrc(a.x[c] + ro, rowb);
                const __amdgpu_buffer_rsrc_t r2 = plane_rsrc(a.xp[c] + ro, rowb);
                nx[c] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r1, (int)vin, 0, 0));
                nxp[c] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r2, (int)vin, 0, 0));
            } else {
                const int so = (int)(tr * (uint32_t)W * 4u + (uint32_t)c * pbytes);
                nx[c] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rx, (int)vin, so, 0));
                nxp[c] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rxp, (int)vin, so, 0));
            }
        }
    };
    // Prob term of output row q, one step ahead of its use.  A row outside the component moves
    // the lane offset out of range (a zero load) instead of branching, so the wait for it can
    // sit after the next row's loads.
    auto probload = [&](int q, float *dst) {
#pragma unroll
        for (int c = 0; c < C; c++) {
            const int cy = q >> vsh[c];
            const bool cin = cy < crh[c];
            // The row offset is clamped rather than selected, so no branch splits the step.
            if constexpr (WIDE) {
                const uint32_t cb = (uint32_t)crw[c] * 4u;
                const __amdgpu_buffer_rsrc_t rp =
                    plane_rsrc(a.ch[c].prob + (size_t)min(cy, crh[c] - 1) * (size_t)crw[c], cb);
                dst[c] = __builtin_bit_cast(float,
                                            __builtin_amdgcn_raw_buffer_load_b32(rp, (int)(cin ? vpr[c] : cb), 0, 0));
            } else {
                const int so = min(cy, crh[c] - 1) * crw[c] * 4;
                dst[c] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                       rpr, (int)(cin ? vpr[c] : a.prob_bytes), so, 0));
            }
        }
    };

#pragma unroll
    for (int c = 0; c < C; c++) {
#pragma unroll
        for (int j = 0; j < 3; j++) yring[j][c] = dxring[j][c] = dyring[j][c] = pring[j][c] = 0.0f;
        head[c] = body[c] = tgy_up[c] = tgd_up[c] = 0.0f;
        acc[c] = DD{0.0, 0.0};
    }
    rowload(r0 - 2);
    const int kend = RW + 4;  // steps k = 0 .. RW + 3; step k reads row r0 - 2 + k

    // Step k: extrapolate row t = r0 - 2 + k, form its x-differences and the y-differences of
    // row t - 1, evaluate the contributions of source row s = t - 1 (k >= 2) and finish output
    // row t - 2 (k >= 4).  KM = k mod 3 picks the ring slots; PH = min(k, 4) the prologue phase.
    // From k = 4 on the source row is >= 1, so the top-border tests drop out.
    auto step = [&](const int k, auto KM, auto PH) {
        constexpr int km = decltype(KM)::value, ph = decltype(PH)::value;
        constexpr int k1 = (km + 2) % 3, k2 = (km + 1) % 3;  // rings: one and two rows back
        const int t = r0 - 2 + k;
        float *yt = yring[km], *dxt = dxring[km];
        if (k < 4 && share && wave > 0) {
            // The lower wave's first rows are the upper wave's last: hand them over (here, where
            // the loads have arrived anyway).
#pragma unroll
            for (int c = 0; c < C; c++) {
                hal[wave - 1][k][2 * c][lane] = nx[c];
                hal[wave - 1][k][2 * c + 1][lane] = nxp[c];
            }
        }
#pragma unroll
        for (int c = 0; c < C; c++) yt[c] = fista(nx[c], nxp[c], f);
        {
            // Store y of row t at once if the row is one of this strip's outputs (otherwise the
            // lane offset moves out of range: a dropped store), so no row of y stays in registers
            // for the two steps until its gradient is done.
            const bool tin = t >= r0 && t < r0 + RW && t < H;
#pragma unroll
            for (int c = 0; c < C; c++) {
                if constexpr (WIDE) {
                    const size_t ro = tin ? (size_t)t * (size_t)W : 0;
                    const __amdgpu_buffer_rsrc_t r2 = plane_rsrc(a.y[c] + ro, rowb);
                    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, yt[c]), r2,
                                                          (int)(tin ? vout : rowb), 0, 0);
                } else {
                    const uint32_t toff = tin ? (uint32_t)t * (uint32_t)W * 4u : 0u;
                    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, yt[c]), ry,
                                                          (int)(tin ? vout : arena),
                                                          (int)(toff + (uint32_t)c * pbytes), 0);
                }
            }
        }
        if (k + 1 < kend) {
            rowload(t + 1);
            if constexpr (ph >= 1) probload(t + 1, pring[k2]);  // the head of step k + 1 (row t + 1)
        }
#pragma unroll
        for (int c = 0; c < C; c++) dxt[c] = from_right(yt[c]) - yt[c];
        if constexpr (ph >= 1) {
            const int s = t - 1;  // source row whose dy is now known
            const float *yprev = yring[k1], *dxprev = dxring[k1], *dxs1 = dxring[k2], *dys1 = dyring[k1];
            float *dys = dyring[km];
#pragma unroll
            for (int c = 0; c < C; c++) dys[c] = yt[c] - yprev[c];
            if constexpr (ph >= 2) {
                const bool top = ph < 4 && s <= 0;  // s == 0 has no row above; s < 0 is outside
                const bool valid = xin && s < H && (ph >= 4 || s >= 0);
                // Joint norms at source (s, x): compute_step_tv.rs:62-68, compute_step_tv2.rs:53-86.
                float nacc = 0.0f, n2acc = 0.0f;
                float gxx[C], gyy[C], sym[C];
#pragma unroll
                for (int c = 0; c < C; c++) {
                    const float gx = dxprev[c], gy = dys[c];
                    nacc += gx * gx;
                    nacc += gy * gy;
                    gxx[c] = gyy[c] = sym[c] = 0.0f;
                    if (TGV) {
                        const float lgx = from_left(gx), lgy = from_left(gy);
                        const float xx = xl ? gx - lgx : 0.0f;
                        // No select: at x = 0 the lane to the left (column -1) loaded column 0, so
                        // its y-difference equals this one bit for bit and yx is +0, the value the
                        // reference's x > 0 test gives (compute_step_tv2.rs:53-58).
                        const float yx = gy - lgy;
                        const float yy = top ? 0.0f : gy - dys1[c];
                        const float xy = top ? 0.0f : gx - dxs1[c];
                        const float sm = midpoint_f32(xy, yx);
                        n2acc += yy * yy + (2.0f * (sm * sm) + xx * xx);
                        gxx[c] = xx;
                        gyy[c] = yy;
                        sym[c] = sm;
                    }
                }
                // A zero norm or a source outside the image contributes nothing: its
                // reciprocal is set to 0, so every contribution becomes +-0, which adds
                // nothing to an accumulator that is never -0.  (An infinite norm would give
                // a NaN reciprocal where IEEE division gives a / inf = 0; it needs squared
                // differences above 3.4e38, which int16 coefficients and u16 tables cannot
                // produce.)
                float tvs[C], tvr[C], tvd[C], tgs[C], tgx[C], tgy[C], tgd[C];
                if constexpr (FAST) {
                    const float n = __builtin_amdgcn_sqrtf(nacc);
                    const float ar = (valid && n != 0.0f) ? atv * __builtin_amdgcn_rcpf(n) : 0.0f;
                    float gr = 0.0f;
                    if (TGV) {
                        const float n2 = __builtin_amdgcn_sqrtf(n2acc);
                        gr = (valid && n2 != 0.0f) ? atg * __builtin_amdgcn_rcpf(n2) : 0.0f;
                    }
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        const float gx = dxprev[c], gy = dys[c];
                        tvs[c] = ar * -(gx + gy);
                        tvr[c] = ar * gx;
                        tvd[c] = ar * gy;
                        tgs[c] = tgx[c] = tgy[c] = tgd[c] = 0.0f;
                        if (TGV) {
                            const float xx = gxx[c], yy = gyy[c], sm = sym[c];
                            const float sx = sm + xx;
                            tgs[c] = gr * (-2.0f * (yy + sx));
                            tgx[c] = gr * sx;
                            tgy[c] = gr * (yy + sm);
                            tgd[c] = gr * -sm;
                        }
                    }
                } else {
                    const float n = sqrtf(nacc);
                    const double rn = (valid && n != 0.0f) ? recip64(n) : 0.0;
                    double rn2 = 0.0;
                    if (TGV) {
                        const float n2 = sqrtf(n2acc);
                        rn2 = (valid && n2 != 0.0f) ? recip64(n2) : 0.0;
                    }
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        const float gx = dxprev[c], gy = dys[c];
                        tvs[c] = qdiv(atv * -(gx + gy), rn);
                        tvr[c] = qdiv(atv * gx, rn);
                        tvd[c] = qdiv(atv * gy, rn);
                        tgs[c] = tgx[c] = tgy[c] = tgd[c] = 0.0f;
                        if (TGV) {
                            const float xx = gxx[c], yy = gyy[c], sm = sym[c];
                            const float sx = sm + xx;
                            // -(2 yy + (2 xx + 2 sm)) == -2 (yy + (sm + xx)) bit for bit: doubling
                            // is exact and commutes with rounding (a subnormal sum is exact).
                            tgs[c] = atg * qdiv(-2.0f * (yy + sx), rn2);
                            tgx[c] = atg * qdiv(sx, rn2);
                            tgy[c] = atg * qdiv(yy + sm, rn2);
                            tgd[c] = atg * qdiv(-sm, rn2);
                        }
                    }
                }
                // Gather in the reference's order for destination q:
                //   prob(q); TV from q-W, q-1, q; second-order from q-W, q-W+1, q-1, q, q+1,
                //   q+W-1, q+W.
                // A row's sum is built as its terms become known, so only four values per channel
                // stay live between steps: it starts one step early (prob, then tvd of the row
                // above: the head), takes its own row's terms with the row above's tgy and tgd
                // (the body), and ends with the row below's tgd (from the left) and tgy.
                if constexpr (ph >= 4) {
                    // Finish output row q = s - 1 and store it.
                    const int q = s - 1;
                    const bool qin = q < H;
                    const bool acc_on = out_lane && qin;
                    // Branch-free: a row outside the image moves the lane offset out of range
                    // (a dropped store) instead, so no wait is forced at a join.
                    const uint32_t vst = qin ? vout : arena;
                    const uint32_t qoff = qin ? (uint32_t)q * (uint32_t)W * 4u : 0u;
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        float gv = body[c];
                        if (TGV) {
                            gv += from_left(tgd[c]);
                            gv += tgy[c];
                        }
                        if constexpr (WIDE) {
                            // A row outside the image keeps row 0's base with the out-of-range offset.
                            const size_t ro = qin ? (size_t)q * (size_t)W : 0;
                            const uint32_t vw = qin ? vout : rowb;
                            const __amdgpu_buffer_rsrc_t r1 = plane_rsrc(a.g[c] + ro, rowb);
                            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, gv), r1, (int)vw, 0, 0);
                        } else {
                            const int so = (int)(qoff + (uint32_t)c * pbytes);
                            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, gv), rg, (int)vst, so,
                                                                  0);
                        }
                        // A lane's sum of g^2 over its strip rows is a plain f64 sum (g^2 is exact in
                        // f64); lanes, waves and workgroups are merged in double-double (the
                        // exact mode) or f64 (the fast mode).
                        const float ga = acc_on ? gv : 0.0f;
                        acc[c].hi = __builtin_fma((double)ga, (double)ga, acc[c].hi);
                    }
                }
                if constexpr (ph >= 3) {
                    // Body of row s: its own terms and the row above's second-order terms.
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        float gv = head[c];
                        gv += from_left(tvr[c]);
                        gv += tvs[c];
                        if (TGV) {
                            gv += tgy_up[c];
                            gv += from_right(tgd_up[c]);
                            gv += from_left(tgx[c]);
                            gv += tgs[c];
                            gv += from_right(tgx[c]);
                        }
                        body[c] = gv;
                    }
                }
                // Head of row s + 1 (its prob term, prefetched last step, then row s's tvd).
#pragma unroll
                for (int c = 0; c < C; c++) {
                    float gv = 0.0f;
                    gv += pring[km][c];
                    gv += tvd[c];
                    head[c] = gv;
                    tgy_up[c] = tgy[c];
                    tgd_up[c] = tgd[c];
                }
            }
        }
    };

    step(0, IC<0>{}, IC<0>{});
    step(1, IC<1>{}, IC<1>{});
    step(2, IC<2>{}, IC<2>{});
    step(3, IC<0>{}, IC<3>{});
    if (share) __syncthreads();  // the shared rows are in LDS before any upper wave takes them
    // Whole groups of three steps (no exit inside a group, so values that one step hands to the
    // next stay in one basic block), then the remaining 0-2 steps.
    int k = 4;
    for (; k + 3 <= kend; k += 3) {
        step(k, IC<1>{}, IC<4>{});
        step(k + 1, IC<2>{}, IC<4>{});
        step(k + 2, IC<0>{}, IC<4>{});
    }
    if (k < kend) step(k, IC<1>{}, IC<4>{});
    if (k + 1 < kend) step(k + 1, IC<2>{}, IC<4>{});

    // Per-workgroup sums: wave c merges channel c across the workgroup's waves lane by lane,
    // then across its lanes, and adds lane 0's double-double result to the accumulator (lanes
    // 0-2 the chunks of its high part, lanes 3-5 those of its low part).
#pragma unroll
    for (int c = 0; c < C; c++) lred[wave][c][lane] = acc[c];
    __syncthreads();
    if (wave < C) {
        const int c = wave;
        DD v = lred[0][c][lane];
#pragma unroll
        for (int w = 1; w < WAVES; w++) v = dd_merge(v, lred[w][c][lane]);
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) v = dd_merge(v, dd_shfl_xor(v, m));
        const double hi = __shfl(v.hi, 0, 64), lo = __shfl(v.lo, 0, 64);
        if (lane < 6) {
            const unsigned int wg = ty * gridDim.x + tx;
            long long *acc = a.acc + (size_t)(wg % ACC_SLOTS) * 3 * ACC_CH + (size_t)c * ACC_CH;
            acc_add(acc, lane < 3 ? hi : lo, lane < 3 ? lane : lane - 3);
        }
    }
}

// Gradient norm per channel (compute_step.rs:54-58) from the accumulator k_gradient filled,
// so every workgroup that forms it gets the same bits, whatever order k_gradient's workgroups
// ran in.  Called by all threads of a workgroup (it has a barrier); on return out.norm[c] and
// out.rnorm[c] hold the results in LDS.  If `writer` is set, the workgroup also stores them,
// with the squared norm, to the plan's buffers, and clears the next iteration's accumulator
// (which no kernel of this iteration touches).  k_gradient adds with device-scope atomics;
// the reads and the clearing are plain accesses in the kernels before and after it, ordered
// by the kernel boundaries like every other buffer the kernels hand on.
struct NormShared {
    float norm[3];
    double rnorm[3];
};
template <int C>
__device__ __forceinline__ void accumulated_norms(const IterArgs &a, bool writer, NormShared &out)
{
    const int tid = threadIdx.x, lane = tid & 63, c = tid >> 6;  // wave c forms channel c
    if (c < C) {
        long long v = 0;
        if (lane < ACC_CH) {
#pragma unroll
            for (int sl = 0; sl < ACC_SLOTS; sl++) v += a.acc[(size_t)sl * 3 * ACC_CH + c * ACC_CH + lane];
        }
        const double sum = acc_value(v, lane);
        if (lane == 0) {
            const float norm = (float)sqrt(sum);
            const double rn = norm != 0.0f ? 1.0 / (double)norm : 0.0;
            out.norm[c] = norm;
            out.rnorm[c] = rn;
            if (writer) {
                a.norm_sq[c] = sum;
                a.norms[c] = norm;
                a.rnorm[c] = rn;
            }
        }
    }
    if (writer)
        for (int i = tid; i < ACC_WORDS; i += blockDim.x) a.acc_next[i] = 0;
    __syncthreads();
}

// The norms a projection kernel descends with, formed by each of its workgroups; the first
// workgroup is the writer.
template <int C>
__device__ __forceinline__ void projection_norms(const IterArgs &a, NormShared &ns)
{
    accumulated_norms<C>(a, (blockIdx.x | blockIdx.y) == 0, ns);
}

// Clears both accumulators of a plan (2 * ACC_WORDS words): before a loop starts, so that the
// first iteration's sum starts from zero whatever ran before.
__global__ __launch_bounds__(NT) void k_zero_acc(long long *acc)
{
    for (int i = threadIdx.x; i < 2 * ACC_WORDS; i += NT) acc[i] = 0;
}

// The norms alone (one workgroup), for the calls that run k_gradient without a projection.
template <int C>
__global__ __launch_bounds__(NT) void k_norms(const IterArgs a)
{
    __shared__ NormShared ns;
    accumulated_norms<C>(a, true, ns);
}

struct ProjShared {
    float yv[TH][TW + 1];        // descended values, full resolution
    float mv[TH][TW + 1];        // component means (resampled channels)
    float blk[MAXBLK * BSTR];    // projection blocks
    float pbk[MAXBLK * BSTR];    // prob-term blocks
};

// Box projection of one channel over one 16x64 tile (compute_projection.rs:8-122), plus
// the prob plane of the next iteration (compute_step_prob.rs:14-33).
template <int HS, int VS, bool RES>
__device__ __forceinline__ void project_channel(const IterArgs &a, const ChanDev &ch, float *xnew, ProjShared &sh,
                                                const float yd[4], int tr, int tc, int x0, int y0, bool live)
{
    constexpr int CW = TW / HS, CT = TH / VS, NBX = CW / 8, NBY = CT / 8, NB = NBX * NBY;
    const int tid = threadIdx.x;
    const int cx0 = x0 / HS, cy0 = y0 / VS;
    const int rw = ch.rw, rh = ch.rh;

    if (RES) {
#pragma unroll
        for (int k = 0; k < 4; k++) sh.yv[tr][tc + k] = yd[k];
        __syncthreads();
        for (int i = tid; i < CW * CT; i += NT) {
            const int ly = i / CW, lx = i % CW;
            if (cx0 + lx < rw && cy0 + ly < rh) {
                float m = 0.0f;
#pragma unroll
                for (int sy = 0; sy < VS; sy++)
#pragma unroll
                    for (int sx = 0; sx < HS; sx++) m += sh.yv[ly * VS + sy][lx * HS + sx];
                m /= (float)(HS * VS);
                sh.mv[ly][lx] = m;
                sh.blk[((ly >> 3) * NBX + (lx >> 3)) * BSTR + (ly & 7) * 9 + (lx & 7)] = m;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            sh.blk[((tr >> 3) * NBX + ((tc + k) >> 3)) * BSTR + (tr & 7) * 9 + ((tc + k) & 7)] = yd[k];
    }
    __syncthreads();

    // Forward DCT (column pass, row pass), clamp, prob difference.
    {
        const int b = tid >> 3, j = tid & 7;
        const int bx = cx0 / 8 + b % NBX, by = cy0 / 8 + b / NBX;
        const bool bv = b < NB && bx < (int)ch.bw && by < (int)ch.bh;
        float *bp = sh.blk + b * BSTR;
        float v[8];
        if (bv) {
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bp[k * 9 + j];
            dct8(v);
#pragma unroll
            for (int k = 0; k < 8; k++) bp[k * 9 + j] = v[k];
        }
        __syncthreads();
        if (bv) {
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bp[j * 9 + k];
            dct8(v);
            const size_t base = ((size_t)by * ch.bw + bx) * 64 + j * 8;
            const short4 c03 = *reinterpret_cast<const short4 *>(ch.coef + base);
            const short4 c47 = *reinterpret_cast<const short4 *>(ch.coef + base + 4);
            const float cf[8] = {(float)c03.x, (float)c03.y, (float)c03.z, (float)c03.w,
                                 (float)c47.x, (float)c47.y, (float)c47.z, (float)c47.w};
            float o[8], d[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const float q = ch.quant[j * 8 + k];
                o[k] = clamp_ref(v[k], (cf[k] - 0.5f) * q, (cf[k] + 0.5f) * q);
                d[k] = qdiv(o[k] - cf[k] * q, ch.rqq[j * 8 + k]);
            }
            if (a.write_cos) {
                *reinterpret_cast<float4 *>(ch.cos + base) = make_float4(o[0], o[1], o[2], o[3]);
                *reinterpret_cast<float4 *>(ch.cos + base + 4) = make_float4(o[4], o[5], o[6], o[7]);
            }
            float *pp = sh.pbk + b * BSTR;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                bp[j * 9 + k] = o[k];
                pp[j * 9 + k] = d[k];
            }
        }
        __syncthreads();
    }

    // Two IDCTs at once: threads 0..127 the projected block, 128..255 the prob block.
    {
        const int half = tid >> 7, t2 = tid & 127;
        const int b = t2 >> 3, j = t2 & 7;
        const int bx = cx0 / 8 + b % NBX, by = cy0 / 8 + b / NBX;
        const bool bv = b < NB && bx < (int)ch.bw && by < (int)ch.bh && (half == 0 || ch.alpha_p != 0.0f);
        float *bp = (half ? sh.pbk : sh.blk) + b * BSTR;
        float v[8];
        if (bv) {
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bp[k * 9 + j];
            idct8(v);
#pragma unroll
            for (int k = 0; k < 8; k++) bp[k * 9 + j] = v[k];
        }
        __syncthreads();
        if (bv) {
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bp[j * 9 + k];
            idct8(v);
            if (half == 0) {
#pragma unroll
                for (int k = 0; k < 8; k++) bp[j * 9 + k] = v[k];
            } else {
                float *row = ch.prob + (size_t)(by * 8 + j) * rw + bx * 8;
                const float ap = ch.alpha_p;
                *reinterpret_cast<float4 *>(row) = make_float4(ap * v[0], ap * v[1], ap * v[2], ap * v[3]);
                *reinterpret_cast<float4 *>(row + 4) = make_float4(ap * v[4], ap * v[5], ap * v[6], ap * v[7]);
            }
        }
        __syncthreads();
    }

    // Unbox and re-add the projected mean (compute_projection.rs:94-121).
    if (live) {
        float o[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int ly = tr / VS, lx = (tc + k) / HS;
            float r = yd[k];
            if (cx0 + lx < rw && cy0 + ly < rh) {
                const float pr = sh.blk[((ly >> 3) * NBX + (lx >> 3)) * BSTR + (ly & 7) * 9 + (lx & 7)];
                r = RES ? (yd[k] - sh.mv[ly][lx]) + pr : pr;
            }
            o[k] = r;
        }
        *reinterpret_cast<float4 *>(xnew + (size_t)(y0 + tr) * a.W + x0 + tc) = make_float4(o[0], o[1], o[2], o[3]);
    }
    __syncthreads();
}

template <int C>
__global__ __launch_bounds__(NT) void k_project(const IterArgs a)
{
    __shared__ ProjShared sh;
    __shared__ NormShared ns;
    const int tid = threadIdx.x;
    unsigned int tx = blockIdx.x, ty = blockIdx.y;
    if (a.xcd_bands) xcd_tile(tx, ty);  // the XCD's band of rows, as in k_gradient
    const int x0 = (int)tx * TW, y0 = (int)ty * TH;
    const int W = a.W, H = a.H;
    const int tr = tid >> 4, tc = (tid & 15) * 4;
    const int qy = y0 + tr, qx = x0 + tc;
    const bool live = qy < H && qx < W;
    const float step = a.step;
    const size_t p = live ? (size_t)qy * W + qx : 0;

    // Every channel's loads first, so they are all in flight together.
    float4 yv[C], gv[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        yv[c] = *reinterpret_cast<const float4 *>(a.y[c] + p);
        gv[c] = *reinterpret_cast<const float4 *>(a.g[c] + p);
    }
    projection_norms<C>(a, ns);
#pragma unroll
    for (int c = 0; c < C; c++) {
        const ChanDev &ch = a.ch[c];
        const float norm = ns.norm[c];
        const double rn = ns.rnorm[c];
        // Descent step on the extrapolated image (compute_step.rs:61-65).
        const float ys[4] = {yv[c].x, yv[c].y, yv[c].z, yv[c].w};
        const float gs[4] = {gv[c].x, gv[c].y, gv[c].z, gv[c].w};
        float yd[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float y = ys[k];
            if (norm != 0.0f) y -= step * qdiv(gs[k], rn);
            yd[k] = y;
        }
        switch (ch.mode) {
        case PM_FULL: project_channel<1, 1, false>(a, ch, a.xnew[c], sh, yd, tr, tc, x0, y0, live); break;
        case PM_SUB11: project_channel<1, 1, true>(a, ch, a.xnew[c], sh, yd, tr, tc, x0, y0, live); break;
        case PM_SUB22: project_channel<2, 2, true>(a, ch, a.xnew[c], sh, yd, tr, tc, x0, y0, live); break;
        case PM_SUB21: project_channel<2, 1, true>(a, ch, a.xnew[c], sh, yd, tr, tc, x0, y0, live); break;
        default: project_channel<1, 2, true>(a, ch, a.xnew[c], sh, yd, tr, tc, x0, y0, live); break;
        }
    }
}

// Compile-time geometry of a channel inside a k_project tile.
template <int MODE> struct ModeGeo {
    static constexpr int HS = (MODE == PM_SUB22 || MODE == PM_SUB21) ? 2 : 1;
    static constexpr int VS = (MODE == PM_SUB22 || MODE == PM_SUB12) ? 2 : 1;
    static constexpr bool RES = MODE != PM_FULL;
};
template <int MODE> constexpr int mode_nbx() { return TW / ModeGeo<MODE>::HS / 8; }
template <int MODE> constexpr int mode_nb() { return mode_nbx<MODE>() * (TH / ModeGeo<MODE>::VS / 8); }

template <int C, int M0, int M1, int M2> struct FusedLayout {
    static constexpr int NB0 = mode_nb<M0>();
    static constexpr int NB1 = C > 1 ? mode_nb<M1>() : 0;
    static constexpr int NB2 = C > 2 ? mode_nb<M2>() : 0;
    static constexpr int TOTB = NB0 + NB1 + NB2;
    // LDS slots for the resampled channels' full-resolution tile and box means.
    static constexpr int R0 = ModeGeo<M0>::RES ? 1 : 0;
    static constexpr int R1 = C > 1 && ModeGeo<M1>::RES ? 1 : 0;
    static constexpr int R2 = C > 2 && ModeGeo<M2>::RES ? 1 : 0;
    static constexpr int NRES = R0 + R1 + R2 > 0 ? R0 + R1 + R2 : 1;
    static constexpr int SLOT1 = R0, SLOT2 = R0 + R1;
};

template <int C, int M0, int M1, int M2> struct FusedShared {
    float blk[FusedLayout<C, M0, M1, M2>::TOTB * BSTR];   // projection blocks of every channel
    float pbk[FusedLayout<C, M0, M1, M2>::TOTB * BSTR];   // prob-term blocks
    float yv[FusedLayout<C, M0, M1, M2>::NRES][TH][TW + 1];  // descended values (resampled channels)
    float mv[FusedLayout<C, M0, M1, M2>::NRES][TH][TW + 1];  // their box means
};

// Stage one channel's descended tile into LDS: straight into 8x8 blocks at full
// resolution, or into yv for the box means (compute_projection.rs:18-61).
template <int SLOT, int MODE, int BOFF, typename SH>
__device__ __forceinline__ void fused_stage(SH &sh, const float yd[4], int tr, int tc)
{
    if (ModeGeo<MODE>::RES) {
#pragma unroll
        for (int k = 0; k < 4; k++) sh.yv[SLOT][tr][tc + k] = yd[k];
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            sh.blk[(BOFF + (tr >> 3) * 8 + ((tc + k) >> 3)) * BSTR + (tr & 7) * 9 + ((tc + k) & 7)] = yd[k];
    }
}

template <int SLOT, int MODE, int BOFF, typename SH>
__device__ __forceinline__ void fused_means(SH &sh, const ChanDev &ch, int x0, int y0)
{
    if (!ModeGeo<MODE>::RES) return;
    constexpr int HS = ModeGeo<MODE>::HS, VS = ModeGeo<MODE>::VS;
    constexpr int CW = TW / HS, CT = TH / VS, NBX = CW / 8;
    const int cx0 = x0 / HS, cy0 = y0 / VS;
    for (int i = threadIdx.x; i < CW * CT; i += NT) {
        const int ly = i / CW, lx = i % CW;
        if (cx0 + lx < (int)ch.rw && cy0 + ly < (int)ch.rh) {
            float m = 0.0f;
#pragma unroll
            for (int sy = 0; sy < VS; sy++)
#pragma unroll
                for (int sx = 0; sx < HS; sx++) m += sh.yv[SLOT][ly * VS + sy][lx * HS + sx];
            m /= (float)(HS * VS);
            sh.mv[SLOT][ly][lx] = m;
            sh.blk[(BOFF + (ly >> 3) * NBX + (lx >> 3)) * BSTR + (ly & 7) * 9 + (lx & 7)] = m;
        }
    }
}

template <int SLOT, int MODE, int BOFF, typename SH>
__device__ __forceinline__ void fused_emit(SH &sh, const ChanDev &ch, float *xnew, int W, const float yd[4], int tr,
                                           int tc, int x0, int y0, bool live)
{
    constexpr int HS = ModeGeo<MODE>::HS, VS = ModeGeo<MODE>::VS, NBX = mode_nbx<MODE>();
    if (!live) return;
    const int cx0 = x0 / HS, cy0 = y0 / VS;
    float o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int ly = tr / VS, lx = (tc + k) / HS;
        float r = yd[k];
        if (cx0 + lx < (int)ch.rw && cy0 + ly < (int)ch.rh) {
            const float pr = sh.blk[(BOFF + (ly >> 3) * NBX + (lx >> 3)) * BSTR + (ly & 7) * 9 + (lx & 7)];
            r = ModeGeo<MODE>::RES ? (yd[k] - sh.mv[SLOT][ly][lx]) + pr : pr;
        }
        o[k] = r;
    }
    *reinterpret_cast<float4 *>(xnew + (size_t)(y0 + tr) * W + x0 + tc) = make_float4(o[0], o[1], o[2], o[3]);
}

// Which channel block b of the tile belongs to, and its block coordinates in that
// channel's component plane.
template <int C, int M0, int M1, int M2>
__device__ __forceinline__ int fused_block(int b, int x0, int y0, int &bx, int &by)
{
    using L = FusedLayout<C, M0, M1, M2>;
    if (b < L::NB0) {
        bx = x0 / ModeGeo<M0>::HS / 8 + b % mode_nbx<M0>();
        by = y0 / ModeGeo<M0>::VS / 8 + b / mode_nbx<M0>();
        return 0;
    }
    if (C > 1 && b < L::NB0 + L::NB1) {
        const int lb = b - L::NB0;
        bx = x0 / ModeGeo<M1>::HS / 8 + lb % mode_nbx<M1>();
        by = y0 / ModeGeo<M1>::VS / 8 + lb / mode_nbx<M1>();
        return 1;
    }
    const int lb = b - L::NB0 - L::NB1;
    bx = x0 / ModeGeo<M2>::HS / 8 + lb % mode_nbx<M2>();
    by = y0 / ModeGeo<M2>::VS / 8 + lb / mode_nbx<M2>();
    return 2;
}

// k_project for the common layouts (channel modes known at compile time): every
// channel's blocks go through one DCT pass and one double IDCT pass together.
template <int C, int M0, int M1, int M2>
__global__ __launch_bounds__(NT) void k_project_fused(const IterArgs a)
{
    using L = FusedLayout<C, M0, M1, M2>;
    constexpr int TOTB = L::TOTB;
    __shared__ FusedShared<C, M0, M1, M2> sh;
    __shared__ NormShared ns;
    const int tid = threadIdx.x;
    unsigned int tx = blockIdx.x, ty = blockIdx.y;
    if (a.xcd_bands) xcd_tile(tx, ty);  // the XCD's band of rows, as in k_gradient
    const int x0 = (int)tx * TW, y0 = (int)ty * TH;
    const int W = a.W, H = a.H;
    const int tr = tid >> 4, tc = (tid & 15) * 4;
    const int qy = y0 + tr, qx = x0 + tc;
    const bool live = qy < H && qx < W;
    const float step = a.step;
    const size_t p = live ? (size_t)qy * W + qx : 0;

    float4 yv[C], gv[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        yv[c] = *reinterpret_cast<const float4 *>(a.y[c] + p);
        gv[c] = *reinterpret_cast<const float4 *>(a.g[c] + p);
    }
    projection_norms<C>(a, ns);
    // The coefficients of this thread's row-pass items, loaded now so that their latency
    // overlaps the tile loads (the row pass runs after three barriers).
    constexpr int ITEMS = (TOTB * 8 + NT - 1) / NT;
    short4 cpre[ITEMS][2];
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const int jb = tid + i * NT;
        cpre[i][0] = cpre[i][1] = make_short4(0, 0, 0, 0);
        if (jb < TOTB * 8) {
            int bx, by;
            const int c = fused_block<C, M0, M1, M2>(jb >> 3, x0, y0, bx, by);
            const ChanDev &ch = a.ch[c];
            if (bx < (int)ch.bw && by < (int)ch.bh) {
                const size_t base = ((size_t)by * ch.bw + bx) * 64 + (jb & 7) * 8;
                cpre[i][0] = *reinterpret_cast<const short4 *>(ch.coef + base);
                cpre[i][1] = *reinterpret_cast<const short4 *>(ch.coef + base + 4);
            }
        }
    }
    // Descent step on the extrapolated image (compute_step.rs:61-65).
    float yd[C][4];
#pragma unroll
    for (int c = 0; c < C; c++) {
        const float norm = ns.norm[c];
        const double rn = ns.rnorm[c];
        const float ys[4] = {yv[c].x, yv[c].y, yv[c].z, yv[c].w};
        const float gs[4] = {gv[c].x, gv[c].y, gv[c].z, gv[c].w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float y = ys[k];
            if (norm != 0.0f) y -= step * qdiv(gs[k], rn);
            yd[c][k] = y;
        }
    }
    fused_stage<0, M0, 0>(sh, yd[0], tr, tc);
    if constexpr (C > 1) fused_stage<L::SLOT1, M1, L::NB0>(sh, yd[1], tr, tc);
    if constexpr (C > 2) fused_stage<L::SLOT2, M2, L::NB0 + L::NB1>(sh, yd[2], tr, tc);
    __syncthreads();
    if (ModeGeo<M0>::RES || (C > 1 && ModeGeo<M1>::RES) || (C > 2 && ModeGeo<M2>::RES)) {
        fused_means<0, M0, 0>(sh, a.ch[0], x0, y0);
        if constexpr (C > 1) fused_means<L::SLOT1, M1, L::NB0>(sh, a.ch[1], x0, y0);
        if constexpr (C > 2) fused_means<L::SLOT2, M2, L::NB0 + L::NB1>(sh, a.ch[2], x0, y0);
        __syncthreads();
    }

    // Forward DCT, column pass (utils/dct.rs:100-134).
    for (int jb = tid; jb < TOTB * 8; jb += NT) {
        const int b = jb >> 3, j = jb & 7;
        int bx, by;
        const int c = fused_block<C, M0, M1, M2>(b, x0, y0, bx, by);
        const ChanDev &ch = a.ch[c];
        if (bx < (int)ch.bw && by < (int)ch.bh) {
            float *bp = sh.blk + b * BSTR;
            float v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bp[k * 9 + j];
            dct8(v);
#pragma unroll
            for (int k = 0; k < 8; k++) bp[k * 9 + j] = v[k];
        }
    }
    __syncthreads();
    // Row pass, clamp to [(c - 1/2) q, (c + 1/2) q] (compute_projection.rs:63-82) and the
    // prob difference (cos - c q) / q^2 (compute_step_prob.rs:24-30).
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
        const int jb = tid + i * NT, b = jb >> 3, j = jb & 7;
        if (jb >= TOTB * 8) break;
        int bx, by;
        const int c = fused_block<C, M0, M1, M2>(b, x0, y0, bx, by);
        const ChanDev &ch = a.ch[c];
        if (bx < (int)ch.bw && by < (int)ch.bh) {
            float *bp = sh.blk + b * BSTR;
            float v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bp[j * 9 + k];
            dct8(v);
            const size_t base = ((size_t)by * ch.bw + bx) * 64 + j * 8;
            const short4 c03 = cpre[i][0], c47 = cpre[i][1];
            const float cf[8] = {(float)c03.x, (float)c03.y, (float)c03.z, (float)c03.w,
                                 (float)c47.x, (float)c47.y, (float)c47.z, (float)c47.w};
            float o[8], d[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const float q = ch.quant[j * 8 + k];
                o[k] = clamp_ref(v[k], (cf[k] - 0.5f) * q, (cf[k] + 0.5f) * q);
                d[k] = qdiv(o[k] - cf[k] * q, ch.rqq[j * 8 + k]);
            }
            if (a.write_cos) {
                *reinterpret_cast<float4 *>(ch.cos + base) = make_float4(o[0], o[1], o[2], o[3]);
                *reinterpret_cast<float4 *>(ch.cos + base + 4) = make_float4(o[4], o[5], o[6], o[7]);
            }
            float *pp = sh.pbk + b * BSTR;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                bp[j * 9 + k] = o[k];
                pp[j * 9 + k] = d[k];
            }
        }
    }
    __syncthreads();
    // Two IDCTs (utils/dct.rs:30-98): the projected blocks and the prob blocks.
    for (int jb = tid; jb < 2 * TOTB * 8; jb += NT) {
        const int half = jb >= TOTB * 8, jr = jb - half * TOTB * 8;
        const int b = jr >> 3, j = jr & 7;
        int bx, by;
        const int c = fused_block<C, M0, M1, M2>(b, x0, y0, bx, by);
        const ChanDev &ch = a.ch[c];
        if (bx < (int)ch.bw && by < (int)ch.bh && (!half || ch.alpha_p != 0.0f)) {
            float *bp = (half ? sh.pbk : sh.blk) + b * BSTR;
            float v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bp[k * 9 + j];
            idct8(v);
#pragma unroll
            for (int k = 0; k < 8; k++) bp[k * 9 + j] = v[k];
        }
    }
    __syncthreads();
    for (int jb = tid; jb < 2 * TOTB * 8; jb += NT) {
        const int half = jb >= TOTB * 8, jr = jb - half * TOTB * 8;
        const int b = jr >> 3, j = jr & 7;
        int bx, by;
        const int c = fused_block<C, M0, M1, M2>(b, x0, y0, bx, by);
        const ChanDev &ch = a.ch[c];
        if (bx < (int)ch.bw && by < (int)ch.bh && (!half || ch.alpha_p != 0.0f)) {
            float *bp = (half ? sh.pbk : sh.blk) + b * BSTR;
            float v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bp[j * 9 + k];
            idct8(v);
            if (!half) {
#pragma unroll
                for (int k = 0; k < 8; k++) bp[j * 9 + k] = v[k];
            } else {
                float *row = ch.prob + (size_t)(by * 8 + j) * ch.rw + bx * 8;
                const float ap = ch.alpha_p;
                *reinterpret_cast<float4 *>(row) = make_float4(ap * v[0], ap * v[1], ap * v[2], ap * v[3]);
                *reinterpret_cast<float4 *>(row + 4) = make_float4(ap * v[4], ap * v[5], ap * v[6], ap * v[7]);
            }
        }
    }
    __syncthreads();
    // Unbox and re-add the projected mean (compute_projection.rs:94-121).
    fused_emit<0, M0, 0>(sh, a.ch[0], a.xnew[0], W, yd[0], tr, tc, x0, y0, live);
    if constexpr (C > 1)
        fused_emit<L::SLOT1, M1, L::NB0>(sh, a.ch[1], a.xnew[1], W, yd[1], tr, tc, x0, y0, live);
    if constexpr (C > 2)
        fused_emit<L::SLOT2, M2, L::NB0 + L::NB1>(sh, a.ch[2], a.xnew[2], W, yd[2], tr, tc, x0, y0, live);
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

// Bilinear upsampling of component-resolution planes onto the grid, for the learned start
// image: component sample i sits at grid coordinate (i + 0.5) * s - 0.5, and the edges clamp.
// Without subsampling (s = 1) every weight is 0 or 1 and the plane is copied.  Replication
// (k_init_upsample) would leave a TV jump every s pixels in a smooth prediction.  One launch
// for all channels (blockIdx.z); a thread writes 4 adjacent pixels (one 16-byte store; W is a
// multiple of 8) of every gridDim.y-th row.
struct UpsampleArgs {
    const float *plane[3];
    float *out[3];
    uint32_t rw[3], rh[3], hs[3], vs[3];
    uint32_t W, H;
};
__global__ __launch_bounds__(NT) void k_upsample_linear(const UpsampleArgs u)
{
    const uint32_t c = blockIdx.z, xb = (blockIdx.x * NT + threadIdx.x) * 4;
    if (xb >= u.W) return;
    const uint32_t rw = u.rw[c], rh = u.rh[c];
    auto clamp = [](float t, uint32_t n) {
        const int k = (int)t;
        return (uint32_t)(k < 0 ? 0 : (k > (int)n - 1 ? (int)n - 1 : k));
    };
    uint32_t x0[4], x1[4];
    float au[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float uu = ((float)(xb + i) + 0.5f) / (float)u.hs[c] - 0.5f;
        const float fu = floorf(uu);
        au[i] = uu - fu;
        x0[i] = clamp(fu, rw);
        x1[i] = clamp(fu + 1.0f, rw);
    }
    const float *plane = u.plane[c];
    float *out = u.out[c];
    for (uint32_t y = blockIdx.y; y < u.H; y += gridDim.y) {
        const float v = ((float)y + 0.5f) / (float)u.vs[c] - 0.5f;
        const float fv = floorf(v), av = v - fv;
        const uint32_t y0 = clamp(fv, rh), y1 = clamp(fv + 1.0f, rh);
        const float *r0 = plane + (size_t)y0 * rw, *r1 = plane + (size_t)y1 * rw;
        float o[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float top = au[i] == 0.0f ? r0[x0[i]] : (1.0f - au[i]) * r0[x0[i]] + au[i] * r0[x1[i]];
            const float bot = au[i] == 0.0f ? r1[x0[i]] : (1.0f - au[i]) * r1[x0[i]] + au[i] * r1[x1[i]];
            o[i] = av == 0.0f ? top : (1.0f - av) * top + av * bot;
        }
        *reinterpret_cast<float4 *>(out + (size_t)y * u.W + xb) = make_float4(o[0], o[1], o[2], o[3]);
    }
}

// cos = c * q (coef.rs:94-104); the prob term of the first iteration is 0 (cos == c q).
__global__ __launch_bounds__(NT) void k_init_cos(const ChanDev ch)
{
    const size_t i = (size_t)blockIdx.x * NT + threadIdx.x;
    if (i >= (size_t)ch.bw * ch.bh * 64) return;
    ch.cos[i] = (float)ch.coef[i] * ch.quant[i & 63];
    ch.prob[i] = 0.0f;
}

// ---- Objective of an iterate (build-defined; the reference only forms the discarded
// partial sums of compute_step_tv.rs:70-71 and compute_step_tv2.rs:88-90):
//   F(x) = sum_p alpha n(p) + sum_p a2 n2(p)
//        + sum_c (alpha_p,c / 2) sum_blocks sum_j ((DCT(down_c x_c)_j - c_j q_j) / q_j)^2
// with the f32 constants of the loop and every term in f64 (n, n2 by the T1/T2 forms).

// Ooura 8-point DCT pass in f64 (the oracle's double build of utils/dct.rs:100-169).
__device__ __forceinline__ void dct8d(double v[8])
{
    const double c1r = 0.49039264020161522456, c1i = 0.09754516100806413392, c2r = 0.46193976625564337806,
                 c2i = 0.19134171618254488586, c3r = 0.41573480615127261854, c3i = 0.27778511650980111237,
                 c4r = 0.35355339059327376220, w4r = 0.70710678118654752440;
    double x0r = v[0] + v[7], x1r = v[0] - v[7], x0i = v[2] + v[5], x1i = v[2] - v[5];
    double x2r = v[4] + v[3], x3r = v[4] - v[3], x2i = v[6] + v[1], x3i = v[6] - v[1];
    double xr = x0r + x2r, xi = x0i + x2i;
    v[0] = c4r * (xr + xi);
    v[4] = c4r * (xr - xi);
    xr = x0r - x2r;
    xi = x0i - x2i;
    v[2] = c2r * xr - c2i * xi;
    v[6] = c2r * xi + c2i * xr;
    xr = w4r * (x1i - x3i);
    x1i = w4r * (x1i + x3i);
    x3i = x1i - x3r;
    x1i += x3r;
    x3r = x1r - xr;
    x1r += xr;
    v[1] = c1r * x1r - c1i * x1i;
    v[7] = c1r * x1i + c1i * x1r;
    v[3] = c3r * x3r - c3i * x3i;
    v[5] = c3r * x3i + c3i * x3r;
}

struct ObjArgs {
    uint32_t W, H, C;
    const float *x[3];
    ChanDev ch[3];
    double alpha, a2, ap[3];
    double *partials;  // [workgroups][3 terms][2]
    unsigned int pixel_wgs;
};

__device__ __forceinline__ void block_reduce_dd(DD v, DD (*red)[3], int slot, double *out)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = dd_merge(v, dd_shfl_xor(v, m));
    if (lane == 0) red[wave][slot] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        DD t = red[0][slot];
        for (int w = 1; w < NT / 64; w++) t = dd_merge(t, red[w][slot]);
        out[0] = t.hi;
        out[1] = t.lo;
    }
}

// TV and second-order terms: one thread per grid pixel.
template <int C>
__global__ __launch_bounds__(NT) void k_obj_pixels(const ObjArgs a)
{
    __shared__ DD red[NT / 64][3];
    const int W = a.W, H = a.H;
    const size_t i = (size_t)blockIdx.x * NT + threadIdx.x;
    DD tv{0.0, 0.0}, tg{0.0, 0.0};
    if (i < (size_t)W * H) {
        const int y = (int)(i / W), x = (int)(i % W);
        auto px = [&](int c, int yy, int xx) -> double { return (double)a.x[c][(size_t)yy * W + xx]; };
        auto dxy = [&](int c, int yy, int xx, double &gx, double &gy) {
            const double v = px(c, yy, xx);
            gx = xx < W - 1 ? px(c, yy, xx + 1) - v : 0.0;
            gy = yy < H - 1 ? px(c, yy + 1, xx) - v : 0.0;
        };
        double n = 0.0, n2 = 0.0;
#pragma unroll
        for (int c = 0; c < C; c++) {
            double gx, gy;
            dxy(c, y, x, gx, gy);
            n += gx * gx + gy * gy;
            double gxx = 0, gyx = 0, gyy = 0, gxy = 0;
            if (x > 0) {
                double lx, ly;
                dxy(c, y, x - 1, lx, ly);
                gxx = gx - lx;
                gyx = gy - ly;
            }
            if (y > 0) {
                double ux, uy;
                dxy(c, y - 1, x, ux, uy);
                gyy = gy - uy;
                gxy = gx - ux;
            }
            const double sm = 0.5 * (gxy + gyx);
            n2 += gyy * gyy + 2.0 * sm * sm + gxx * gxx;
        }
        tv.hi = a.alpha * sqrt(n);
        tg.hi = a.a2 * sqrt(n2);
    }
    block_reduce_dd(tv, red, 0, a.partials + (size_t)blockIdx.x * 6 + 0);
    block_reduce_dd(tg, red, 1, a.partials + (size_t)blockIdx.x * 6 + 2);
    if (threadIdx.x == 0) {
        a.partials[(size_t)blockIdx.x * 6 + 4] = 0.0;
        a.partials[(size_t)blockIdx.x * 6 + 5] = 0.0;
    }
}

// Prob term: one thread per 8x8 block of every channel (box means for resampled channels).
template <int C>
__global__ __launch_bounds__(NT) void k_obj_blocks(const ObjArgs a)
{
    __shared__ DD red[NT / 64][3];
    const size_t b = (size_t)blockIdx.x * NT + threadIdx.x;  // block b of every channel
    DD acc{0.0, 0.0};
#pragma unroll
    for (int c = 0; c < C; c++) {
        const ChanDev &ch = a.ch[c];
        const size_t nb = (size_t)ch.bw * ch.bh;
        if (b < nb) {
            if (a.ap[c] != 0.0) {
                const int bx = (int)(b % ch.bw), by = (int)(b / ch.bw);
                const int hs = ch.hs, vs = ch.vs, W = a.W;
                double blk[64];
                for (int iy = 0; iy < 8; iy++)
                    for (int ix = 0; ix < 8; ix++) {
                        const int cy = by * 8 + iy, cx = bx * 8 + ix;
                        double v;
                        if (ch.resample) {
                            v = 0.0;
                            for (int sy = 0; sy < vs; sy++)
                                for (int sx = 0; sx < hs; sx++)
                                    v += (double)a.x[c][(size_t)(cy * vs + sy) * W + cx * hs + sx];
                            v /= (double)(hs * vs);
                        } else {
                            v = (double)a.x[c][(size_t)cy * W + cx];
                        }
                        blk[iy * 8 + ix] = v;
                    }
                for (int j = 0; j < 8; j++) {
                    double v[8];
                    for (int k = 0; k < 8; k++) v[k] = blk[k * 8 + j];
                    dct8d(v);
                    for (int k = 0; k < 8; k++) blk[k * 8 + j] = v[k];
                }
                double s = 0.0;
                for (int j = 0; j < 8; j++) {
                    double v[8];
                    for (int k = 0; k < 8; k++) v[k] = blk[j * 8 + k];
                    dct8d(v);
                    for (int k = 0; k < 8; k++) {
                        const double q = (double)ch.quant[j * 8 + k];
                        const double d = (v[k] - (double)ch.coef[b * 64 + j * 8 + k] * q) / q;
                        s += d * d;
                    }
                }
                dd_add(acc, 0.5 * a.ap[c] * s);
            }
        }
    }
    double out[2] = {0.0, 0.0};
    block_reduce_dd(acc, red, 2, out);
    if (threadIdx.x == 0) {
        double *pp = a.partials + ((size_t)a.pixel_wgs + blockIdx.x) * 6;
        pp[0] = pp[1] = pp[2] = pp[3] = 0.0;
        pp[4] = out[0];
        pp[5] = out[1];
    }
}

// Fixed-order reduction of the objective partials: out[0..2] = TV, second-order, prob.
__global__ __launch_bounds__(NT) void k_obj_final(const double *partials, unsigned int n, double *out)
{
    __shared__ DD red[NT / 64][3];
    DD t[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (unsigned int w = threadIdx.x; w < n; w += NT)
#pragma unroll
        for (int k = 0; k < 3; k++) t[k] = dd_merge(t[k], DD{partials[(size_t)w * 6 + 2 * k], partials[(size_t)w * 6 + 2 * k + 1]});
    double o[3][2];
#pragma unroll
    for (int k = 0; k < 3; k++) block_reduce_dd(t[k], red, k, o[k]);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) out[k] = o[k][0] + o[k][1];
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

// ---------------------------------------------------------------------------------------
// Output stage (lib.rs:159-211).  Each thread converts 4 consecutive pixels of one row: the
// planes are row-major with a stride that is a multiple of 8, so the float4 loads at a
// column that is a multiple of 4 stay inside the row.  Products and sums are rounded one at
// a time (-ffp-contract=off), as the reference's default mul_add! (utils/macros.rs:10-15).
struct ColourArgs {
    const float *y, *cb, *cr;
    uint32_t stride, real_w, real_h, colour;
    uint8_t *rgb;
};

// f32::clamp(0, 255) then `as u8`: NaN survives the clamp and the saturating cast maps it to 0.
__device__ __forceinline__ uint32_t rust_u8(float v)
{
    v = v < 0.0f ? 0.0f : v;
    v = v > 255.0f ? 255.0f : v;
    return v == v ? (uint32_t)v : 0u;
}

__global__ __launch_bounds__(NT) void k_colour(ColourArgs a)
{
    const uint32_t j0 = (blockIdx.x * NT + threadIdx.x) * 4;
    if (j0 >= a.real_w) return;
    for (uint32_t i = blockIdx.y; i < a.real_h; i += gridDim.y) {
        const size_t row = (size_t)i * a.stride + j0;
        const float4 yv = *reinterpret_cast<const float4 *>(a.y + row);
        float4 cbv = make_float4(0.f, 0.f, 0.f, 0.f), crv = cbv;
        if (a.colour) {
            cbv = *reinterpret_cast<const float4 *>(a.cb + row);
            crv = *reinterpret_cast<const float4 *>(a.cr + row);
        }
        const float ys[4] = {yv.x, yv.y, yv.z, yv.w};
        const float cbs[4] = {cbv.x, cbv.y, cbv.z, cbv.w};
        const float crs[4] = {crv.x, crv.y, crv.z, crv.w};
        uint32_t bytes[12];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float y = ys[k] + 128.0f;
            if (a.colour) {
                const float r = 1.402f * crs[k] + y;
                const float gi = 0.34414f * -cbs[k] + y;
                const float g = 0.71414f * -crs[k] + gi;
                const float b = 1.772f * cbs[k] + y;
                bytes[3 * k + 0] = rust_u8(r);
                bytes[3 * k + 1] = rust_u8(g);
                bytes[3 * k + 2] = rust_u8(b);
            } else {
                const uint32_t v = rust_u8(y);
                bytes[3 * k + 0] = bytes[3 * k + 1] = bytes[3 * k + 2] = v;
            }
        }
        const size_t o = ((size_t)i * a.real_w + j0) * 3;
        if (j0 + 4 <= a.real_w && (o & 3) == 0) {
            uint32_t *dst = reinterpret_cast<uint32_t *>(a.rgb + o);
#pragma unroll
            for (int w = 0; w < 3; w++)
                dst[w] = bytes[4 * w] | bytes[4 * w + 1] << 8 | bytes[4 * w + 2] << 16 | bytes[4 * w + 3] << 24;
        } else {
            const uint32_t n = min(4u, a.real_w - j0) * 3;
            for (uint32_t k = 0; k < n; k++) a.rgb[o + k] = (uint8_t)bytes[k];
        }
    }
}

// ---------------------------------------------------------------------------------------
// Precision form of the projection: the 8x8 DCT and IDCTs as BF16 MFMA contractions with
// split hi/lo operands and f32 accumulation (v_mfma_f32_16x16x16_bf16), for layouts whose
// channels all cover the grid at full resolution (4:4:4, luma only).  Not bit-identical to
// the reference: each product keeps about 16 significant bits.
//
// A wave owns a 16x16 super-tile (2x2 blocks).  The image enters as the A operand: lane l,
// item t holds X[l%16][4*(l/16)+t], four consecutive pixels of one row, so loads and stores
// are one float4 per lane and plane.  With Db = blockdiag(D, D) (D the orthonormal 8-point
// DCT-II matrix) one stage computes C = A * B with a constant B.  The accumulator layout
// (lane l, item t holds C[4*(l/16)+t][l%16]) read as the next stage's A operand is C^T, so
// two chained stages give
//   forward:  (X Db^T)^T Db^T = Db X^T Db^T = Y^T      with Y = Db X Db^T,
//   inverse:  (Y Db)^T Db = Db^T Y^T Db = X'^T         with X' = Db^T Y Db,
// and the accumulator then holds Y (resp. X') in the A operand layout again: lane l has
// coefficients (u = l%16 mod 8, v = 4*(l/16) mod 8 + t) and, after the inverse, pixels of the
// same row and columns it loaded.  No LDS, no barriers.  Each stage is three MFMAs:
// a_lo*b_hi + a_hi*b_lo + a_hi*b_hi (the a_lo*b_lo term, ~2^-16 relative, is dropped).
// Constants bypass the contractions, exactly.  Each forward stage transforms every block row
// minus its first element and adds K0 times that element (K0 = the sum of the f32 basis row 0)
// to the row's frequency 0; the inverse runs on the AC coefficients and adds DC / 8 to every
// pixel.  The Ooura passes give exact zeros wherever a block row is constant (their
// butterflies subtract equal values), and TV's g / |g| turns any non-zero difference, however
// small, into a unit-size gradient term; so flat blocks, and blocks whose rows (or columns) are
// all equal, come out exact here too, where the split operands alone leave ~1e-7.  Elsewhere
// the ~2^-16 rounding of the products still moves the trajectory off the f32 one: one
// iteration agrees with the f32 loop to 3e-5 RMS, the next differs by 0.27, and after 100 the
// result is about as far from the FP64 loop as the f32 loop's (RMS 0.027 against 0.026 at
// 4096x3072 q30; 0.045 without the exact constants).
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_rne(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_to_f32(short h) { return __uint_as_float((uint32_t)(uint16_t)h << 16); }

// A constant B operand as two bf16 parts: hi, and lo = the bf16 rounding of what hi leaves.
struct SplitB {
    bf16x4 p[2];
};
typedef __bf16 bf16x2_hw __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Two f32 to bf16 with round-to-nearest-even (v_cvt_pk_bf16_f32); the same bits as bf16_rne
// for finite inputs.  Element 0 in the low half.
__device__ __forceinline__ uint32_t cvt_pk_bf16(float x0, float x1)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){x0, x1}, bf16x2_hw));
}

__device__ __forceinline__ f32x4 mfma_stage(const f32x4 a, const SplitB &b)
{
    // a = hi + lo with hi = bf16(a) and lo = bf16(a - hi) (a - hi is exact in f32).
    const uint32_t h01 = cvt_pk_bf16(a[0], a[1]), h23 = cvt_pk_bf16(a[2], a[3]);
    const uint32_t l01 = cvt_pk_bf16(a[0] - __uint_as_float(h01 << 16), a[1] - __uint_as_float(h01 & 0xFFFF0000u));
    const uint32_t l23 = cvt_pk_bf16(a[2] - __uint_as_float(h23 << 16), a[3] - __uint_as_float(h23 & 0xFFFF0000u));
    const bf16x4 ah = __builtin_bit_cast(bf16x4, (u32x2){h01, h23});
    const bf16x4 al = __builtin_bit_cast(bf16x4, (u32x2){l01, l23});
    f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, b.p[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, b.p[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, b.p[0], c, 0, 0, 0);
    return c;
}

// One forward stage (C = A Db^T per 8x8 block) with constant rows taken out exactly: lane l
// holds A[m][k] for m = l % 16, k = 4 (l / 16) + t, so A[m][8 b] (the first element of row m of
// block column b) is item 0 of lane m + 32 b; the output C[m'][n] sits in lane n + 16 (m' / 4),
// item m' % 4, and frequency 0 of a block is n % 8 == 0.
__device__ __forceinline__ f32x4 mfma_forward_stage(const f32x4 a, const SplitB &b, int lane)
{
    constexpr float K0 = 0x1.6a09e6p+1f;  // 8 * D[0][k] (f32), exact
    const float first = __shfl(a[0], lane & 47, 64);
    f32x4 d;
#pragma unroll
    for (int t = 0; t < 4; t++) d[t] = a[t] - first;
    f32x4 c = mfma_stage(d, b);
    // The first elements of the output rows m' = 4 (l / 16) + t of this lane's block column.
    const int bc = (lane & 15) >> 3;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const float f = __shfl(a[0], ((lane >> 4) * 4 + t) + 32 * bc, 64);
        if ((lane & 7) == 0) c[t] += K0 * f;
    }
    return c;
}

// Orthonormal 8-point DCT-II basis D[u][x] = c(u) cos((2x + 1) u pi / 16), c(0) = 1/sqrt(8),
// c(u > 0) = 1/2, rounded to f32 (the B operands of the precision form).
__constant__ float kDct8[64] = {
    0x1.6a09e6p-2f, 0x1.6a09e6p-2f, 0x1.6a09e6p-2f, 0x1.6a09e6p-2f, 0x1.6a09e6p-2f, 0x1.6a09e6p-2f, 0x1.6a09e6p-2f, 0x1.6a09e6p-2f,
    0x1.f6297cp-2f, 0x1.a9b662p-2f, 0x1.1c73b4p-2f, 0x1.8f8b84p-4f, -0x1.8f8b84p-4f, -0x1.1c73b4p-2f, -0x1.a9b662p-2f, -0x1.f6297cp-2f,
    0x1.d906bcp-2f, 0x1.87de2ap-3f, -0x1.87de2ap-3f, -0x1.d906bcp-2f, -0x1.d906bcp-2f, -0x1.87de2ap-3f, 0x1.87de2ap-3f, 0x1.d906bcp-2f,
    0x1.a9b662p-2f, -0x1.8f8b84p-4f, -0x1.f6297cp-2f, -0x1.1c73b4p-2f, 0x1.1c73b4p-2f, 0x1.f6297cp-2f, 0x1.8f8b84p-4f, -0x1.a9b662p-2f,
    0x1.6a09e6p-2f, -0x1.6a09e6p-2f, -0x1.6a09e6p-2f, 0x1.6a09e6p-2f, 0x1.6a09e6p-2f, -0x1.6a09e6p-2f, -0x1.6a09e6p-2f, 0x1.6a09e6p-2f,
    0x1.1c73b4p-2f, -0x1.f6297cp-2f, 0x1.8f8b84p-4f, 0x1.a9b662p-2f, -0x1.a9b662p-2f, -0x1.8f8b84p-4f, 0x1.f6297cp-2f, -0x1.1c73b4p-2f,
    0x1.87de2ap-3f, -0x1.d906bcp-2f, 0x1.d906bcp-2f, -0x1.87de2ap-3f, -0x1.87de2ap-3f, 0x1.d906bcp-2f, -0x1.d906bcp-2f, 0x1.87de2ap-3f,
    0x1.8f8b84p-4f, -0x1.1c73b4p-2f, 0x1.a9b662p-2f, -0x1.f6297cp-2f, 0x1.f6297cp-2f, -0x1.a9b662p-2f, 0x1.1c73b4p-2f, -0x1.8f8b84p-4f,
};

// SUB: channels 1 and 2 are 4:2:0 chroma (C = 3).  A workgroup's TW x TH tile then holds one
// super-tile of each chroma component (4 component blocks side by side, 32 x 8 component
// pixels): every wave forms the 2x2 box means of its pixels into an LDS tile, waves 0 and 1 run
// the transforms of Cb and Cr on it, and every wave adds the projected means back.
template <int C, bool SUB = false>
__global__ __launch_bounds__(NT) void k_project_mfma(const IterArgs a)
{
    static_assert(!SUB || C == 3, "k_project_mfma: 4:2:0 chroma needs three channels");
    __shared__ NormShared ns;
    __shared__ __align__(16) float ctile[SUB ? 2 : 1][8][36];  // SUB: each chroma component's 8 x 32 tile
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int W = a.W, H = a.H;
    // The wave's four 8x8 blocks sit side by side in one 8-row band: super-tile blocks (0,0),
    // (0,1), (1,0), (1,1) are image blocks 0, 1, 2, 3 from the left; a workgroup covers 2 x 2
    // such bands of 32 columns (its TW x TH tile).
    unsigned int tx = blockIdx.x, ty = blockIdx.y;
    if (a.xcd_bands) xcd_tile(tx, ty);  // the XCD's band of rows, as in k_gradient
    const int x0 = (int)tx * TW + (wave & 1) * 32, y0 = (int)ty * TH + (wave >> 1) * 8;
    // X enters in the A operand layout (lane l: row l % 16, columns 4 (l / 16) .. + 3), so each
    // lane loads and stores one float4 per plane.  B operands: lane l holds element
    // (k = 4 (l / 16) + t, n = l % 16).  Forward stages use Db^T, inverse stages Db.
    const int r = lane & 15, cq = (lane >> 4) * 4;
    SplitB fwd, inv;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int k = cq + t, n = r;
        const bool same = (k >> 3) == (n >> 3);
        const int ku = k & 7, nu = n & 7;
        float vf = same ? kDct8[nu * 8 + ku] : 0.0f;  // Db^T[k][n] = D[n][k]
        float vi = same ? kDct8[ku * 8 + nu] : 0.0f;  // Db[k][n] = D[k][n]
#pragma unroll
        for (int i = 0; i < 2; i++) {
            fwd.p[i][t] = bf16_rne(vf);
            inv.p[i][t] = bf16_rne(vi);
            vf -= bf16_to_f32(fwd.p[i][t]);
            vi -= bf16_to_f32(inv.p[i][t]);
        }
    }
    // This lane's four pixels: super-tile row r, columns cq .. cq + 3, i.e. image row
    // y0 + r % 8, columns x0 + 16 (r / 8) + cq .. + 3 (W and H are multiples of 8, so the four are
    // all inside the grid or all outside).  Every load of every channel is issued here, without
    // branches: a lane outside reads pixel 0 and coefficient 0 and stores nothing.
    const int py = y0 + (r & 7), px = x0 + 16 * (r >> 3) + cq;
    const bool in = px < W && py < H;
    const size_t pix = in ? (size_t)py * W + px : 0;
    const int kq = (r & 7) * 8 + (cq & 7);  // natural-order index of the first coefficient
    auto coef_index = [&](int c) -> uint32_t {  // block order, < 2^31 by the plan's size limit
        return in ? ((uint32_t)(py >> 3) * a.ch[c].bw + (uint32_t)(px >> 3)) * 64u + (uint32_t)kq : 0u;
    };
    const float step = a.step;
    float4 yl[C], gl[C];
    int2 cl[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        yl[c] = *reinterpret_cast<const float4 *>(a.y[c] + pix);
        gl[c] = *reinterpret_cast<const float4 *>(a.g[c] + pix);
        if (!SUB || c == 0) cl[c] = *reinterpret_cast<const int2 *>(a.ch[c].coef + coef_index(c));
    }
    projection_norms<C>(a, ns);
#pragma unroll
    for (int c = 0; c < (SUB ? 1 : C); c++) {
        const ChanDev &ch = a.ch[c];
        const float norm = ns.norm[c];
        const double rn = ns.rnorm[c];
        const float ys[4] = {yl[c].x, yl[c].y, yl[c].z, yl[c].w};
        const float gs[4] = {gl[c].x, gl[c].y, gl[c].z, gl[c].w};
        f32x4 xv;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            float v = ys[t];
            // Descent step on the extrapolated image (compute_step.rs:61-65).
            if (norm != 0.0f) v -= step * qdiv(gs[t], rn);
            xv[t] = in ? v : 0.0f;
        }
        // Forward DCT (row pass, then column pass): this lane then holds coefficients
        // (u = r & 7, v = (cq & 7) + t) of its block.  Clamp to [(c - 1/2) q, (c + 1/2) q] and
        // form the prob difference (cos - c q) / q^2.
        const f32x4 yv = mfma_forward_stage(mfma_forward_stage(xv, fwd, lane), fwd, lane);
        // Lane (lane & 40) holds the DC coefficient of this lane's block (item 0).
        const bool dc_lane = (lane & 40) == lane;
        const float4 q4 = *reinterpret_cast<const float4 *>(ch.quant + kq);
        const float qs[4] = {q4.x, q4.y, q4.z, q4.w};
        const int cw[2] = {cl[c].x, cl[c].y};
        f32x4 ov, dv;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const float q = qs[t];
            const float cf = (float)(short)((uint32_t)cw[t >> 1] >> (16 * (t & 1)));
            float o = clamp_ref(yv[t], (cf - 0.5f) * q, (cf + 0.5f) * q);
            float d = qdiv(o - cf * q, ch.rqq[kq + t]);
            ov[t] = in ? o : 0.0f;
            dv[t] = in ? d : 0.0f;
        }
        if (in && a.write_cos)
            *reinterpret_cast<float4 *>(ch.cos + coef_index(c)) = make_float4(ov[0], ov[1], ov[2], ov[3]);
        // Inverse DCT (row pass, then column pass): this lane then holds row py, columns px ..
        // px + 3 of the projected image and of the prob plane.
        const float odc = __shfl(ov[0], lane & 40, 64) * 0.125f;
        f32x4 oac = ov;
        if (dc_lane) oac[0] = 0.0f;
        f32x4 xn = mfma_stage(mfma_stage(oac, inv), inv);
#pragma unroll
        for (int t = 0; t < 4; t++) xn[t] += odc;
        const bool prob_on = ch.alpha_p != 0.0f;
        f32x4 pn = {0.0f, 0.0f, 0.0f, 0.0f};
        if (prob_on) {
            const float ddc = __shfl(dv[0], lane & 40, 64) * 0.125f;
            f32x4 dac = dv;
            if (dc_lane) dac[0] = 0.0f;
            pn = mfma_stage(mfma_stage(dac, inv), inv);
#pragma unroll
            for (int t = 0; t < 4; t++) pn[t] += ddc;
        }
        if (in) {
            *reinterpret_cast<float4 *>(a.xnew[c] + pix) = make_float4(xn[0], xn[1], xn[2], xn[3]);
            if (prob_on) {
                const float ap = ch.alpha_p;
                *reinterpret_cast<float4 *>(ch.prob + (size_t)py * ch.rw + px) =
                    make_float4(ap * pn[0], ap * pn[1], ap * pn[2], ap * pn[3]);
            }
        }
    }
    if constexpr (SUB) {
        // This lane's pixels (row py, columns px .. px + 3) are component row py / 2, columns
        // px / 2 and px / 2 + 1: in the workgroup's component tile, row 4 (wave / 2) + (r & 7) / 2
        // and columns 16 (wave & 1) + 8 (r / 8) + cq / 2 + e.  The other row of each 2x2 box is
        // in lane ^ 1 (r ^ 1).
        const int trow = 4 * (wave >> 1) + ((r & 7) >> 1), tcol = 16 * (wave & 1) + 8 * (r >> 3) + (cq >> 1);
        const bool lower = (r & 1) != 0;
        float res[2][4];
#pragma unroll
        for (int cc = 0; cc < 2; cc++) {
            const int c = cc + 1;
            const float norm = ns.norm[c];
            const double rn = ns.rnorm[c];
            const float ys[4] = {yl[c].x, yl[c].y, yl[c].z, yl[c].w};
            const float gs[4] = {gl[c].x, gl[c].y, gl[c].z, gl[c].w};
            float v[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                float y = ys[t];
                if (norm != 0.0f) y -= step * qdiv(gs[t], rn);  // compute_step.rs:61-65
                v[t] = in ? y : 0.0f;
            }
            // Box means in the reference's order (compute_projection.rs:21-31): the upper row's
            // pair, then the lower row's; both lanes of the pair form the same sums.
            float m[2];
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const float a0 = v[2 * e], a1 = v[2 * e + 1];
                const float b0 = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(a0), 0xB1, 0xF, 0xF, false));
                const float b1 = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(a1), 0xB1, 0xF, 0xF, false));
                const float u0 = lower ? b0 : a0, u1 = lower ? b1 : a1;
                const float l0 = lower ? a0 : b0, l1 = lower ? a1 : b1;
                float mm = 0.0f;
                mm += u0;
                mm += u1;
                mm += l0;
                mm += l1;
                mm /= 4.0f;
                m[e] = mm;
                res[cc][2 * e] = v[2 * e] - mm;
                res[cc][2 * e + 1] = v[2 * e + 1] - mm;
            }
            *reinterpret_cast<float2 *>(&ctile[cc][trow][tcol]) = make_float2(m[0], m[1]);
        }
        __syncthreads();
        if (wave < 2) {
            // Wave cc projects chroma component cc + 1's four blocks: lane (r, cq) holds
            // component row cy, columns cx .. cx + 3, in the same A layout as the luma tiles.
            const int cc = wave, c = cc + 1;
            const ChanDev &ch = a.ch[c];
            const int cy = (int)ty * (TH / 2) + (r & 7), cx = (int)tx * (TW / 2) + 16 * (r >> 3) + cq;
            const bool cin = cx < (int)ch.rw && cy < (int)ch.rh;
            const uint32_t ci = cin ? ((uint32_t)(cy >> 3) * ch.bw + (uint32_t)(cx >> 3)) * 64u + (uint32_t)kq : 0u;
            const int2 craw = *reinterpret_cast<const int2 *>(ch.coef + ci);
            const float4 tv = *reinterpret_cast<const float4 *>(&ctile[cc][r & 7][16 * (r >> 3) + cq]);
            const f32x4 xv = {tv.x, tv.y, tv.z, tv.w};
            const f32x4 yv = mfma_forward_stage(mfma_forward_stage(xv, fwd, lane), fwd, lane);
            const bool dc_lane = (lane & 40) == lane;
            const float4 q4 = *reinterpret_cast<const float4 *>(ch.quant + kq);
            const float qs[4] = {q4.x, q4.y, q4.z, q4.w};
            const int cw[2] = {craw.x, craw.y};
            f32x4 ov, dv;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const float q = qs[t];
                const float cf = (float)(short)((uint32_t)cw[t >> 1] >> (16 * (t & 1)));
                float o = clamp_ref(yv[t], (cf - 0.5f) * q, (cf + 0.5f) * q);
                float d = qdiv(o - cf * q, ch.rqq[kq + t]);
                ov[t] = cin ? o : 0.0f;
                dv[t] = cin ? d : 0.0f;
            }
            if (cin && a.write_cos)
                *reinterpret_cast<float4 *>(ch.cos + ci) = make_float4(ov[0], ov[1], ov[2], ov[3]);
            const float odc = __shfl(ov[0], lane & 40, 64) * 0.125f;
            f32x4 oac = ov;
            if (dc_lane) oac[0] = 0.0f;
            f32x4 xn = mfma_stage(mfma_stage(oac, inv), inv);
#pragma unroll
            for (int t = 0; t < 4; t++) xn[t] += odc;
            if (ch.alpha_p != 0.0f) {
                const float ddc = __shfl(dv[0], lane & 40, 64) * 0.125f;
                f32x4 dac = dv;
                if (dc_lane) dac[0] = 0.0f;
                f32x4 pn = mfma_stage(mfma_stage(dac, inv), inv);
                const float ap = ch.alpha_p;
                if (cin)
                    *reinterpret_cast<float4 *>(ch.prob + (size_t)cy * ch.rw + cx) =
                        make_float4(ap * (pn[0] + ddc), ap * (pn[1] + ddc), ap * (pn[2] + ddc), ap * (pn[3] + ddc));
            }
            // The projected means replace the box means in the tile (this lane's own four).
            *reinterpret_cast<float4 *>(&ctile[cc][r & 7][16 * (r >> 3) + cq]) = make_float4(xn[0], xn[1], xn[2], xn[3]);
        }
        __syncthreads();
        // compute_projection.rs:108-121: each pixel's difference from its box mean on top of
        // the projected mean.
#pragma unroll
        for (int cc = 0; cc < 2; cc++) {
            const float2 pm = *reinterpret_cast<const float2 *>(&ctile[cc][trow][tcol]);
            if (in)
                *reinterpret_cast<float4 *>(a.xnew[cc + 1] + pix) =
                    make_float4(res[cc][0] + pm.x, res[cc][1] + pm.x, res[cc][2] + pm.y, res[cc][3] + pm.y);
        }
    }
}

// ---------------------------------------------------------------------------------------
// k_project_lane: the same projection as k_project_fused (same Ooura passes on the same
// vectors, so the same bits) without workgroup barriers past the norms.  Eight lanes own one
// 8x8 block, one block row each; the column passes run on columns obtained by an 8x8 transpose
// across those lanes.  The block-row index j sits in lane bits 3-5, which makes two of the
// transpose's three exchange stages plain register swaps between lane halves
// (v_permlane32_swap, j bit 2) and between 16-lane rows (v_permlane16_swap, j bit 1); only the
// j bit 0 stage (lane ^ 8, DPP row_ror:8) selects.  A wave takes 8 consecutive blocks of one
// block row of a full-resolution channel (lane = 8 j + block: rows load and store coalesced),
// or 4 blocks of 4:2:0 chroma, whose 16-row band it loads row-contiguously and hands to the
// transform layout (lane = 8 j + 4 sy + block, both sy lanes carrying component row j) through
// a wave-private LDS tile (lane_unit_sub22).

// Value of the same register in lane ^ 8 (DPP row_ror:8).
__device__ __forceinline__ float lane_xor8(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xF, 0xF, false));
}

// Lane with block row j (lane bits 3-5) holds a[k] = M[j][k]; afterwards a[k] = M[k][j].
// Stage S exchanges a[k + S] of the lanes with j & S == 0 against a[k] of their partners
// (j ^ S) for every k with k & S == 0.
__device__ __forceinline__ void transpose8(float a[8], int j)
{
    // j bit 2 = lane bit 5: the upper 32 lanes of a[k] trade with the lower 32 of a[k + 4].
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a[k]), __float_as_uint(a[k + 4]), false, false);
        a[k] = __uint_as_float(r[0]);
        a[k + 4] = __uint_as_float(r[1]);
    }
    // j bit 1 = lane bit 4: the odd 16-lane rows of a[k] trade with the even rows of a[k + 2].
#pragma unroll
    for (int k = 0; k < 8; k++) {
        if (k & 2) continue;
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a[k]), __float_as_uint(a[k + 2]), false, false);
        a[k] = __uint_as_float(r[0]);
        a[k + 2] = __uint_as_float(r[1]);
    }
    // j bit 0 = lane bit 3: lane ^ 8 with selects (on values read out first, so that they stay
    // register selects and never become a lane-dependent index into the array).
    const bool hi = (j & 1) != 0;
#pragma unroll
    for (int k = 0; k < 8; k += 2) {
        const float lo_v = a[k], hi_v = a[k + 1];
        const float recv = lane_xor8(hi ? lo_v : hi_v);
        a[k] = hi ? recv : lo_v;
        a[k + 1] = hi ? hi_v : recv;
    }
}

// DCT of the block whose row j this lane holds (column pass, then row pass: dct.rs:100-169),
// the clamp and the prob difference for coefficient row j, then the IDCTs of both (column
// pass, then row pass: dct.rs:30-98).  raw: the block's int16 coefficients of row j, loaded by
// the caller.  On return v holds row j of the projected block and pr row j of the prob block
// (before alpha_p).  `store` selects the lane that writes cos.
__device__ __forceinline__ void lane_project_block(float v[8], int j, const ChanDev &ch, size_t blk, int4 raw,
                                                   const float q8[8], const double rq8[8], bool valid, bool store,
                                                   bool write_cos, float pr[8])
{
    transpose8(v, j);
    dct8(v);  // column j
    transpose8(v, j);
    dct8(v);  // row j: coefficients (u = j, v = k)
    float cf[8];
    const int w4[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        cf[2 * k] = (float)(short)(w4[k] & 0xFFFF);
        cf[2 * k + 1] = (float)(short)((uint32_t)w4[k] >> 16);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const float q = q8[k];
        const float o = clamp_ref(v[k], (cf[k] - 0.5f) * q, (cf[k] + 0.5f) * q);
        pr[k] = qdiv(o - cf[k] * q, rq8[k]);
        v[k] = o;
    }
    if (write_cos && valid && store) {
        float *c = ch.cos + blk * 64 + j * 8;
        *reinterpret_cast<float4 *>(c) = make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4 *>(c + 4) = make_float4(v[4], v[5], v[6], v[7]);
    }
    transpose8(v, j);
    transpose8(pr, j);
    idct8(v);
    idct8(pr);
    transpose8(v, j);
    transpose8(pr, j);
    idct8(v);
    idct8(pr);
}

__device__ __forceinline__ void lane_descend(float norm, double rn, float step, float4 yv, float4 gv, float out[4])
{
    const float ys[4] = {yv.x, yv.y, yv.z, yv.w};
    const float gs[4] = {gv.x, gv.y, gv.z, gv.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float y = ys[k];
        if (norm != 0.0f) y -= step * qdiv(gs[k], rn);  // compute_step.rs:61-65
        out[k] = y;
    }
}

__device__ __forceinline__ void lane_store_prob(const ChanDev &ch, int by, int bx, int j, const float pr[8])
{
    // Next iteration's prob plane, component raster (compute_step_prob.rs:24-55 for x_new).
    if (ch.alpha_p == 0.0f) return;
    const float ap = ch.alpha_p;
    float *row = ch.prob + (size_t)(by * 8 + j) * ch.rw + bx * 8;
    *reinterpret_cast<float4 *>(row) = make_float4(ap * pr[0], ap * pr[1], ap * pr[2], ap * pr[3]);
    *reinterpret_cast<float4 *>(row + 4) = make_float4(ap * pr[4], ap * pr[5], ap * pr[6], ap * pr[7]);
}

// Every load of a unit is issued up front and unconditionally: lanes of blocks past the
// channel's width (valid false) read block 0 / pixel 0, run the transforms on those values
// within their own lane group, and store nothing.

// A unit's y, g and coefficient loads.  Full resolution: one block row of 8 pixels (y[0..1],
// g[0..1]).  4:2:0 chroma: the unit's band of 16 grid rows by 64 columns row-contiguously, load
// q covering rows 4q .. 4q + 3 (lane 16 h + s: pixels 4s .. 4s + 3 of row 4q + h, in y[q] and
// g[q]), so that each load instruction reads four whole 256-byte row segments.  A lane of a
// block past the channel's width loads from offset 0.
struct LaneLoad {
    float4 y[4], g[4];
    int4 raw;
};
template <int MODE>
__device__ __forceinline__ void lane_load(const IterArgs &a, int c, int by, int bx, int r, bool valid, LaneLoad &L)
{
    const ChanDev &ch = a.ch[c];
    const size_t blk = valid ? (size_t)by * ch.bw + bx : 0;
    if constexpr (MODE == PM_SUB22) {
        const int lane = threadIdx.x & 63, h = lane >> 4, s = lane & 15;
        const int bx0 = bx - (lane & 3);  // the unit's first block
        const bool lvalid = bx0 + (s >> 2) < (int)ch.bw;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const size_t p = lvalid ? (size_t)(by * 16 + 4 * q + h) * a.W + (size_t)bx0 * 16 + 4 * s : 0;
            L.y[q] = *reinterpret_cast<const float4 *>(a.y[c] + p);
            L.g[q] = *reinterpret_cast<const float4 *>(a.g[c] + p);
        }
        L.raw = *reinterpret_cast<const int4 *>(ch.coef + blk * 64 + (r >> 1) * 8);
    } else {
        const size_t p = valid ? (size_t)(by * 8 + r) * a.W + bx * 8 : 0;
        const float4 *yp = reinterpret_cast<const float4 *>(a.y[c] + p);
        const float4 *gp = reinterpret_cast<const float4 *>(a.g[c] + p);
#pragma unroll
        for (int q4 = 0; q4 < 2; q4++) {
            L.y[q4] = yp[q4];
            L.g[q4] = gp[q4];
        }
        L.raw = *reinterpret_cast<const int4 *>(ch.coef + blk * 64 + r * 8);
    }
}

// Full-resolution channel: 8 lanes per block, block row j = lane >> 3.
__device__ __forceinline__ void lane_unit_full(const IterArgs &a, const NormShared &ns, int c, int by, int bx, int j,
                                               bool valid, const LaneLoad &L)
{
    const ChanDev &ch = a.ch[c];
    const float norm = ns.norm[c];
    const double rn = ns.rnorm[c];
    const size_t blk = valid ? (size_t)by * ch.bw + bx : 0;
    const size_t p = valid ? (size_t)(by * 8 + j) * a.W + bx * 8 : 0;
    const float4 y0 = L.y[0], y1 = L.y[1], g0 = L.g[0], g1 = L.g[1];
    const int4 raw = L.raw;
    float q8[8];
    double rq8[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        q8[k] = ch.quant[j * 8 + k];
        rq8[k] = ch.rqq[j * 8 + k];
    }
    float v[8], pr[8];
    lane_descend(norm, rn, a.step, y0, g0, v);
    lane_descend(norm, rn, a.step, y1, g1, v + 4);
    lane_project_block(v, j, ch, blk, raw, q8, rq8, valid, true, a.write_cos != 0, pr);
    if (valid) {
        *reinterpret_cast<float4 *>(a.xnew[c] + p) = make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4 *>(a.xnew[c] + p + 4) = make_float4(v[4], v[5], v[6], v[7]);
        lane_store_prob(ch, by, bx, j, pr);
    }
}

// Orders a wave's LDS accesses across its lanes: the wavefront-scope release/acquire pair makes
// the stores before it visible to the loads after it, and wave_barrier keeps the compiler from
// moving memory accesses across it.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Value of v in lane ^ 16: v_permlane16_swap of v with itself gives the odd 16-lane rows the
// even rows' values (first result) and the even rows the odd rows' (second).
__device__ __forceinline__ float lane_xor16(float v, bool odd_row)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(odd_row ? r[0] : r[1]);
}

// 4:2:0 chroma.  The band arrives row-contiguously (lane_load: lane 16 h + s holds pixels
// 4s .. 4s + 3 of rows 4q + h).  The rows of a 2x2 box, 4q + h and 4q + (h ^ 1), sit in lanes 16
// apart.  A wave-private LDS tile of the unit's 8 x 32 component pixels carries the box means
// into the transform layout (lane 8 j + 4 sy + b: component row j of block b, held by both sy
// lanes) and the projected means back, with wave_sync() between the stores and the loads (no
// workgroup barrier).  Every value goes through the same operations as in k_project_fused.
__device__ __forceinline__ void lane_unit_sub22(const IterArgs &a, const NormShared &ns, int c, int by, int bx, int r,
                                                bool valid, const LaneLoad &L, float (*tile)[36])
{
    const ChanDev &ch = a.ch[c];
    const float norm = ns.norm[c];
    const double rn = ns.rnorm[c];
    const int lane = threadIdx.x & 63, h = lane >> 4, s = lane & 15, b = lane & 3;
    const int j = r >> 1, sy = r & 1, bx0 = bx - b;
    const bool lvalid = bx0 + (s >> 2) < (int)ch.bw, lower = (h & 1) != 0;
    const size_t blk = valid ? (size_t)by * ch.bw + bx : 0;
    float q8[8];
    double rq8[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        q8[k] = ch.quant[j * 8 + k];
        rq8[k] = ch.rqq[j * 8 + k];
    }
    float yd[16];  // yd[4q + t]: pixel 4s + t of band row 4q + h
#pragma unroll
    for (int q = 0; q < 4; q++) lane_descend(norm, rn, a.step, L.y[q], L.g[q], &yd[4 * q]);
    // Box means in the reference's order (compute_projection.rs:21-31): the upper row's pair,
    // then the lower row's; both lanes of a row pair form the same sums.
    float mq[8];  // mq[2q + e]: component row 2q + (h >> 1), column 2s + e
#pragma unroll
    for (int q = 0; q < 4; q++) {
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const float a0 = yd[4 * q + 2 * e], a1 = yd[4 * q + 2 * e + 1];
            const float b0 = lane_xor16(a0, lower), b1 = lane_xor16(a1, lower);
            const float u0 = lower ? b0 : a0, u1 = lower ? b1 : a1;
            const float l0 = lower ? a0 : b0, l1 = lower ? a1 : b1;
            float m = 0.0f;
            m += u0;
            m += u1;
            m += l0;
            m += l1;
            m /= 4.0f;
            mq[2 * q + e] = m;
            yd[4 * q + 2 * e] -= m;
            yd[4 * q + 2 * e + 1] -= m;
        }
    }
    // Both lanes of a pair store the same means (a lane-dependent choice of registers would
    // send the array to memory).
#pragma unroll
    for (int q = 0; q < 4; q++)
        *reinterpret_cast<float2 *>(&tile[2 * q + (h >> 1)][2 * s]) = make_float2(mq[2 * q], mq[2 * q + 1]);
    wave_sync();
    const float4 t0 = *reinterpret_cast<const float4 *>(&tile[j][8 * b]);
    const float4 t1 = *reinterpret_cast<const float4 *>(&tile[j][8 * b + 4]);
    wave_sync();
    float v[8] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w}, pr[8];
    lane_project_block(v, j, ch, blk, L.raw, q8, rq8, valid, sy == 0, a.write_cos != 0, pr);
    // The projected means back; the two sy lanes of a block row hold and store the same values.
    *reinterpret_cast<float4 *>(&tile[j][8 * b]) = make_float4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<float4 *>(&tile[j][8 * b + 4]) = make_float4(v[4], v[5], v[6], v[7]);
    wave_sync();
    if (lvalid) {
        float *dst = a.xnew[c] + (size_t)(by * 16 + h) * a.W + (size_t)bx0 * 16 + 4 * s;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float2 pm = *reinterpret_cast<const float2 *>(&tile[2 * q + (h >> 1)][2 * s]);
            *reinterpret_cast<float4 *>(dst + (size_t)4 * q * a.W) =
                make_float4(yd[4 * q] + pm.x, yd[4 * q + 1] + pm.x, yd[4 * q + 2] + pm.y, yd[4 * q + 3] + pm.y);
        }
    }
    wave_sync();
    if (valid && sy == 0) lane_store_prob(ch, by, bx, j, pr);
}

// Blocks per wave: 8 for full-resolution channels, 4 for 4:2:0 chroma.
template <int MODE>
constexpr int lane_blocks_per_wave() { return MODE == PM_SUB22 ? 4 : 8; }

// One wave per unit = 8 (full resolution) or 4 (4:2:0 chroma) consecutive blocks of a block row;
// units of channel 0 first.  u0, u1, u2: cumulative unit counts after channels 0, 1, 2.  Each
// workgroup forms the norms once and its waves then take several units, so a grid smaller than
// the unit count spreads that cost.
//
// XCD bands (a.xcd_bands).  The workgroups an XCD runs (workgroup i runs on XCD i % NXCD) take
// the units of the block rows that start in its eighth of the grid's rows, channel after
// channel: the rows k_gradient's workgroups on the same XCD sweep (xcd_tile), so the y and g
// this kernel reads and the x_new and prob it writes for the next sweep stay in that XCD's L2,
// which keeps a kernel's lines for the next one (measured: a 32 MiB buffer read back in 3.4 us
// on the XCDs that wrote it, 18.5 us on the others).  That pays where an iteration's planes fit
// the L2s (512^2 luma: 12.1 -> 10.4 us per iteration); on large grids the bands cost 1-2 %, so
// the plan decides (set_layout, the autotuner).
template <int C, int M0, int M1, int M2>
__global__ __launch_bounds__(NT) void k_project_lane(const IterArgs a, uint32_t u0, uint32_t u1, uint32_t u2)
{
    __shared__ NormShared ns;
    __shared__ __align__(16) float ctile[NT / 64][8][36];  // 4:2:0 chroma units: each wave's component tile
    const int lane = threadIdx.x & 63;
    projection_norms<C>(a, ns);  // before any wave leaves: it has a barrier
    // A grid of fewer than NXCD workgroups leaves XCDs without one: then every workgroup takes
    // from the whole grid.
    const bool bands = a.xcd_bands && gridDim.x >= NXCD;
    // The units this workgroup's waves take, as an index k into a list: unit k itself, or with
    // bands the k-th unit of this XCD's band (its block rows of channel 0, then 1, then 2).
    uint32_t first[3] = {0, u0, u1}, count[3] = {u0, u1 - u0, u2 - u1};
    uint32_t wg = blockIdx.x, nwg = gridDim.x;
    if (bands) {
        const uint32_t xcd = blockIdx.x % NXCD;
        wg = blockIdx.x / NXCD;
        nwg = (gridDim.x - xcd + NXCD - 1) / NXCD;
        const uint32_t r_lo = xcd * a.H / NXCD, r_hi = (xcd + 1) * a.H / NXCD;
        auto band = [&](auto MODE_C, auto C_C) {
            constexpr int MODE = decltype(MODE_C)::value, c = decltype(C_C)::value;
            constexpr uint32_t BPW = lane_blocks_per_wave<MODE>(), ROWS = MODE == PM_SUB22 ? 16 : 8;  // grid rows per block row
            const ChanDev &ch = a.ch[c];
            const uint32_t per_row = (ch.bw + BPW - 1) / BPW;
            const uint32_t by_lo = min((r_lo + ROWS - 1) / ROWS, ch.bh), by_hi = min((r_hi + ROWS - 1) / ROWS, ch.bh);
            first[c] += by_lo * per_row;
            count[c] = (by_hi - by_lo) * per_row;
        };
        band(IC<M0>{}, IC<0>{});
        if constexpr (C > 1) band(IC<M1>{}, IC<1>{});
        if constexpr (C > 2) band(IC<M2>{}, IC<2>{});
    }
    const uint32_t n = count[0] + (C > 1 ? count[1] : 0) + (C > 2 ? count[2] : 0), stride = nwg * (NT / 64);
    for (uint32_t k = __builtin_amdgcn_readfirstlane(wg * (NT / 64) + (threadIdx.x >> 6)); k < n; k += stride) {
        const uint32_t unit = k < count[0] ? first[0] + k
                              : k < count[0] + count[1] ? first[1] + (k - count[0])
                                                        : first[2] + (k - count[0] - count[1]);
        auto run = [&](auto MODE_C, auto C_C, uint32_t local) {
            constexpr int MODE = decltype(MODE_C)::value, c = decltype(C_C)::value;
            constexpr int BPW = lane_blocks_per_wave<MODE>();
            const ChanDev &ch = a.ch[c];
            const uint32_t per_row = (ch.bw + BPW - 1) / BPW;
            const int by = (int)(local / per_row), bx = (int)(local % per_row) * BPW + lane % BPW;
            const bool valid = bx < (int)ch.bw;
            LaneLoad L;
            lane_load<MODE>(a, c, by, bx, lane / BPW, valid, L);
            if constexpr (MODE == PM_SUB22) lane_unit_sub22(a, ns, c, by, bx, lane / BPW, valid, L, ctile[threadIdx.x >> 6]);
            else lane_unit_full(a, ns, c, by, bx, lane / BPW, valid, L);
        };
        if (unit < u0) run(IC<M0>{}, IC<0>{}, unit);
        else if constexpr (C > 2 && M1 == M2 && M1 == PM_SUB22) {
            // Both chroma components through one copy of the code, the channel a (wave-uniform)
            // index: a smaller kernel for the instruction cache.
            const int c = unit < u1 ? 1 : 2;
            const uint32_t local = unit - (c == 1 ? u0 : u1);
            const ChanDev &ch = a.ch[c];
            const uint32_t per_row = (ch.bw + 3) / 4;
            const int by = (int)(local / per_row), bx = (int)(local % per_row) * 4 + lane % 4;
            const bool valid = bx < (int)ch.bw;
            LaneLoad L;
            lane_load<PM_SUB22>(a, c, by, bx, lane / 4, valid, L);
            lane_unit_sub22(a, ns, c, by, bx, lane / 4, valid, L, ctile[threadIdx.x >> 6]);
        } else if (C > 1 && unit < u1) run(IC<M1>{}, IC<1>{}, unit - u0);
        else if (C > 2) run(IC<M2>{}, IC<2>{}, unit - u1);
    }
}

}  // namespace

struct artefact_hip_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t C = 0, W = 0, H = 0, iterations = 0;
    size_t N = 0;
    size_t M[3] = {0, 0, 0};
    float weight = 0.0f, pweight[3] = {0, 0, 0};
    bool tgv = true;
    int rows = 8;  // k_gradient rows per wave
    int layout = 0;  // LAYOUT_*: which k_project variant runs
    int layout_fused = 0;  // the fused layout this plan's channels allow (LAYOUT_GENERIC if none)
    int lane_project = 0;  // fused layouts: 1 = k_project_lane (no LDS), 0 = k_project_fused
    uint32_t lane_units[3] = {0, 0, 0};  // k_project_lane: cumulative units (one wave's blocks) per channel
    uint32_t lane_wgs = 0;               // k_project_lane: workgroup cap (0: one unit per wave)
    bool xcd_bands = false;              // projection workgroups by XCD row band (IterArgs::xcd_bands)
    uint32_t precision = ARTEFACT_PRECISION_F32;  // ARTEFACT_PRECISION_BF16X3: MFMA projection
    float qtab[3][64] = {};                       // quantisation tables per channel, f32 (warm-start features)
    float alpha_tv = 0.0f, alpha_tgv = 0.0f, step = 0.0f;
    std::vector<float> factors;
    std::vector<float> steps;  // per-iteration steps (artefact_hip_plan_set_step_schedule); empty: `step`
    ChanDev ch[3];
    void *arena = nullptr;               // the one device allocation every buffer below lives in
    float *buf[2] = {nullptr, nullptr};  // each C*N: x and x_prev, swapped every iteration
    float *g = nullptr;                  // C*N
    float *ybuf = nullptr;               // C*N: extrapolated image of the current iteration
    float *cos = nullptr;                // sum M_c
    float *prob = nullptr;               // sum M_c
    int16_t *coef = nullptr;             // sum M_c
    float *quant = nullptr;              // [3][64]
    double *rqq = nullptr;               // [3][64]
    long long *acc = nullptr;  // [2][ACC_WORDS]: norm accumulators, by x's buffer index
    float *norms = nullptr;
    double *rnorm = nullptr;
    double *norm_sq = nullptr;
    std::vector<std::vector<float>> init_override;
    dim3 grid_g, grid_p;
    int cur = 0;            // buf[cur] holds x
    bool initialized = false;
    bool use_graph = true;
    // k_gradient with 64-bit addressing: the C planes of the grid reach 2^31 bytes (or
    // ARTEFACT_WIDE_ADDRESSING=1).
    bool wide = false;
    hipGraphExec_t graph = nullptr;
    // Optional event pairs around iterate()'s launches (artefact_hip_plan_set_timing).
    std::vector<hipEvent_t> timing_events;
    uint32_t timing_used = 0;  // pairs recorded
};

namespace {

int check_inputs(const artefact_params *p, const artefact_channel *ch)
{
    if (!p || !ch) return ARTEFACT_E_ARGUMENT;
    if (p->nchannel < 1 || p->nchannel > 3) return ARTEFACT_E_ARGUMENT;
    if (p->mode != ARTEFACT_MODE_TGV && p->mode != ARTEFACT_MODE_TV) return ARTEFACT_E_ARGUMENT;
    if (p->max_w < 8 || p->max_h < 8 || p->max_w % 8 || p->max_h % 8) return ARTEFACT_E_GEOMETRY;
    // int16 coefficients, u16 tables and a bounded grid keep every iterate far below
    // 2^60 in magnitude, so every quotient the kernels form is finite.
    // The cap also keeps every element index within a plane below 2^29 and every byte offset
    // within one plane below 2^31 (grids whose C planes together reach 2^31 bytes use the
    // 64-bit-addressed gradient sweep).
    if ((uint64_t)p->max_w * p->max_h > (1ull << 29)) return ARTEFACT_E_GEOMETRY;
    if (!(p->weight >= 0.0f && p->weight < 1e6f)) return ARTEFACT_E_ARGUMENT;
    for (int c = 0; c < 3; c++)
        if (!(p->pweight[c] >= 0.0f && p->pweight[c] < 1e6f)) return ARTEFACT_E_ARGUMENT;
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
    for (hipEvent_t e : pl->timing_events) (void)hipEventDestroy(e);
    if (pl->arena) (void)hipFree(pl->arena);
    if (pl->stream) (void)hipStreamDestroy(pl->stream);
    delete pl;
}

// The step of iteration `it`: the plan's constant step, or its decaying schedule.
float step_at(const artefact_hip_plan *pl, uint32_t it)
{
    return pl->steps.empty() ? pl->step : pl->steps[it];
}

IterArgs make_args(const artefact_hip_plan *pl, int cur, float factor, bool write_cos)
{
    IterArgs a;
    memset(&a, 0, sizeof(a));
    a.W = pl->W;
    a.H = pl->H;
    a.C = pl->C;
    a.write_cos = write_cos ? 1u : 0u;
    size_t mtot = 0;
    for (uint32_t c = 0; c < pl->C; c++) mtot += pl->M[c];
    a.prob_bytes = (uint32_t)(mtot * sizeof(float));
    a.factor = factor;
    a.step = pl->step;
    a.alpha_tv = pl->alpha_tv;
    a.alpha_tgv = pl->alpha_tgv;
    for (uint32_t c = 0; c < 3; c++) {
        const size_t off = (size_t)(c < pl->C ? c : 0) * pl->N;
        a.x[c] = pl->buf[cur] + off;
        a.xp[c] = pl->buf[1 - cur] + off;
        a.y[c] = pl->ybuf + off;
        a.xnew[c] = pl->buf[1 - cur] + off;
        a.g[c] = pl->g + off;
        a.ch[c] = pl->ch[c < pl->C ? c : 0];
    }
    a.acc = pl->acc + (size_t)cur * ACC_WORDS;  // x's buffer index alternates as the iterations do
    a.acc_next = pl->acc + (size_t)(1 - cur) * ACC_WORDS;
    a.norms = pl->norms;
    a.rnorm = pl->rnorm;
    a.norm_sq = pl->norm_sq;
    a.kg_rows = (uint32_t)pl->rows;
    a.xcd_bands = pl->xcd_bands ? 1u : 0u;
    return a;
}

template <int C, bool TGV>
void launch_gradient_c(const artefact_hip_plan *pl, const IterArgs &a, hipStream_t s)
{
    const bool fast = pl->precision == ARTEFACT_PRECISION_F32_FAST;
    if (pl->wide && fast)
        hipLaunchKernelGGL((k_gradient<C, TGV, true, true>), pl->grid_g, dim3(NT), 0, s, a);
    else if (pl->wide)
        hipLaunchKernelGGL((k_gradient<C, TGV, false, true>), pl->grid_g, dim3(NT), 0, s, a);
    else if (fast)
        hipLaunchKernelGGL((k_gradient<C, TGV, true>), pl->grid_g, dim3(NT), 0, s, a);
    else
        hipLaunchKernelGGL((k_gradient<C, TGV>), pl->grid_g, dim3(NT), 0, s, a);
}

// The gradient sweep (its workgroups' sums of g^2 go to the accumulator a.acc).
void launch_gradient(const artefact_hip_plan *pl, const IterArgs &a, hipStream_t s)
{
    switch (pl->C * 2 + (pl->tgv ? 1 : 0)) {
    case 2: launch_gradient_c<1, false>(pl, a, s); break;
    case 3: launch_gradient_c<1, true>(pl, a, s); break;
    case 4: launch_gradient_c<2, false>(pl, a, s); break;
    case 5: launch_gradient_c<2, true>(pl, a, s); break;
    case 6: launch_gradient_c<3, false>(pl, a, s); break;
    default: launch_gradient_c<3, true>(pl, a, s); break;
    }
}

// The projection kernel for the plan's layout and precision.
void launch_project(const artefact_hip_plan *pl, const IterArgs &a, hipStream_t s)
{
    if (pl->precision == ARTEFACT_PRECISION_BF16X3) {
        if (pl->C == 1) hipLaunchKernelGGL((k_project_mfma<1>), pl->grid_p, dim3(NT), 0, s, a);
        else if (pl->C == 2) hipLaunchKernelGGL((k_project_mfma<2>), pl->grid_p, dim3(NT), 0, s, a);
        else if (pl->ch[1].mode == PM_SUB22) hipLaunchKernelGGL((k_project_mfma<3, true>), pl->grid_p, dim3(NT), 0, s, a);
        else hipLaunchKernelGGL((k_project_mfma<3>), pl->grid_p, dim3(NT), 0, s, a);
        return;
    }
    if (pl->lane_project && pl->layout != LAYOUT_GENERIC) {
        uint32_t wgs = (pl->lane_units[2] + NT / 64 - 1) / (NT / 64);
        if (pl->lane_wgs && pl->lane_wgs < wgs) wgs = pl->lane_wgs;
        const dim3 grid(wgs);
        const uint32_t u0 = pl->lane_units[0], u1 = pl->lane_units[1], u2 = pl->lane_units[2];
        switch (pl->layout) {
        case LAYOUT_Y:
            hipLaunchKernelGGL((k_project_lane<1, PM_FULL, PM_FULL, PM_FULL>), grid, dim3(NT), 0, s, a, u0, u1, u2);
            return;
        case LAYOUT_444:
            hipLaunchKernelGGL((k_project_lane<3, PM_FULL, PM_FULL, PM_FULL>), grid, dim3(NT), 0, s, a, u0, u1, u2);
            return;
        default:
            hipLaunchKernelGGL((k_project_lane<3, PM_FULL, PM_SUB22, PM_SUB22>), grid, dim3(NT), 0, s, a, u0, u1,
                               u2);
            return;
        }
    }
    switch (pl->layout) {
    case LAYOUT_Y:
        hipLaunchKernelGGL((k_project_fused<1, PM_FULL, PM_FULL, PM_FULL>), pl->grid_p, dim3(NT), 0, s, a);
        return;
    case LAYOUT_444:
        hipLaunchKernelGGL((k_project_fused<3, PM_FULL, PM_FULL, PM_FULL>), pl->grid_p, dim3(NT), 0, s, a);
        return;
    case LAYOUT_420:
        hipLaunchKernelGGL((k_project_fused<3, PM_FULL, PM_SUB22, PM_SUB22>), pl->grid_p, dim3(NT), 0, s, a);
        return;
    default: break;
    }
    if (pl->C == 1)
        hipLaunchKernelGGL((k_project<1>), pl->grid_p, dim3(NT), 0, s, a);
    else if (pl->C == 2)
        hipLaunchKernelGGL((k_project<2>), pl->grid_p, dim3(NT), 0, s, a);
    else
        hipLaunchKernelGGL((k_project<3>), pl->grid_p, dim3(NT), 0, s, a);
}

int launch_iterations(artefact_hip_plan *pl, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_acc, dim3(1), dim3(NT), 0, s, pl->acc);
    int cur = pl->cur;
    for (uint32_t it = 0; it < pl->iterations; it++) {
        IterArgs a = make_args(pl, cur, pl->factors[it], it + 1 == pl->iterations);
        a.step = step_at(pl, it);
        launch_gradient(pl, a, s);
        launch_project(pl, a, s);
        cur = 1 - cur;
    }
    return hipGetLastError() == hipSuccess ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

// Start image from the coefficients (no override): the kernels of artefact_hip_plan_init.
// upsample = false (the learned start): only the component planes and cos / prob; x and
// x_prev are left for the caller to write.
void launch_init(artefact_hip_plan *pl, hipStream_t s, bool upsample = true)
{
    hipLaunchKernelGGL(k_zero_acc, dim3(1), dim3(NT), 0, s, pl->acc);
    for (uint32_t c = 0; c < pl->C; c++) {
        const ChanDev &d = pl->ch[c];
        float *x = pl->buf[0] + c * pl->N, *xp = pl->buf[1] + c * pl->N;
        float *plane = pl->g + c * pl->N;  // scratch: M_c <= N
        const size_t nb = (size_t)d.bw * d.bh;
        hipLaunchKernelGGL(k_init_blocks, dim3((unsigned)((nb + NT / 8 - 1) / (NT / 8))), dim3(NT), 0, s, d, plane);
        if (upsample)
            hipLaunchKernelGGL(k_init_upsample, dim3((unsigned)((pl->N + NT - 1) / NT)), dim3(NT), 0, s, d,
                               (const float *)plane, x, xp, pl->W, pl->H);
        hipLaunchKernelGGL(k_init_cos, dim3((unsigned)((pl->M[c] + NT - 1) / NT)), dim3(NT), 0, s, d);
    }
}

// The k_gradient instance for (C, TGV, FAST, WIDE), for occupancy queries.
const void *gradient_kernel(uint32_t C, bool tgv, bool fast, bool wide)
{
    if (wide) {
        switch (C * 4 + (tgv ? 2 : 0) + (fast ? 1 : 0)) {
        case 4: return (const void *)k_gradient<1, false, false, true>;
        case 5: return (const void *)k_gradient<1, false, true, true>;
        case 6: return (const void *)k_gradient<1, true, false, true>;
        case 7: return (const void *)k_gradient<1, true, true, true>;
        case 8: return (const void *)k_gradient<2, false, false, true>;
        case 9: return (const void *)k_gradient<2, false, true, true>;
        case 10: return (const void *)k_gradient<2, true, false, true>;
        case 11: return (const void *)k_gradient<2, true, true, true>;
        case 12: return (const void *)k_gradient<3, false, false, true>;
        case 13: return (const void *)k_gradient<3, false, true, true>;
        case 14: return (const void *)k_gradient<3, true, false, true>;
        default: return (const void *)k_gradient<3, true, true, true>;
        }
    }
    switch (C * 4 + (tgv ? 2 : 0) + (fast ? 1 : 0)) {
    case 4: return (const void *)k_gradient<1, false>;
    case 5: return (const void *)k_gradient<1, false, true>;
    case 6: return (const void *)k_gradient<1, true>;
    case 7: return (const void *)k_gradient<1, true, true>;
    case 8: return (const void *)k_gradient<2, false>;
    case 9: return (const void *)k_gradient<2, false, true>;
    case 10: return (const void *)k_gradient<2, true>;
    case 11: return (const void *)k_gradient<2, true, true>;
    case 12: return (const void *)k_gradient<3, false>;
    case 13: return (const void *)k_gradient<3, false, true>;
    case 14: return (const void *)k_gradient<3, true>;
    default: return (const void *)k_gradient<3, true, true>;
    }
}

// k_gradient strip height RW (output rows per wave).  A wave reads RW + 4 rows and evaluates
// RW + 2 rows of sources, so taller strips waste less, but a grid larger than the resident
// capacity runs a partial second round.  Pick the smallest RW whose grid (4 strips per
// workgroup, stacked) fits the device's resident workgroups at once.  An override
// (ARTEFACT_KG_ROWS) is honoured.
// Resident k_gradient workgroups per CU and the CU count (4 and 256 if the queries fail).
void gradient_residency(int device, uint32_t C, bool tgv, bool fast, bool wide, int &per_cu, int &cus)
{
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, gradient_kernel(C, tgv, fast, wide), NT, 0) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || per_cu < 1 ||
        cus < 1) {
        per_cu = 4;
        cus = 256;
    }
}

// Smallest strip height (>= 2: a 1-row strip spends 5 row loads per output row) whose grid has
// at most `wgs` workgroups.
int rows_for_workgroups(uint32_t W, uint32_t H, uint64_t wgs)
{
    const uint64_t cols = (W + KG_COLS - 1) / KG_COLS;
    for (uint32_t r = 2; r < H; r++)
        if (cols * ((H + WAVES * r - 1) / (WAVES * r)) <= wgs) return (int)r;
    return (int)std::max(H, 1u);
}

int pick_rows(int device, uint32_t W, uint32_t H, uint32_t C, bool tgv, bool fast, bool wide)
{
    // An override below 2 rows is ignored (2 is the tuner's shortest strip).
    if (const char *e = getenv("ARTEFACT_KG_ROWS")) {
        const int r = atoi(e);
        if (r >= 2 && r <= 4096) return r;
    }
    int per_cu = 0, cus = 0;
    gradient_residency(device, C, tgv, fast, wide, per_cu, cus);
    return rows_for_workgroups(W, H, (uint64_t)per_cu * cus);
}

// One step of the FISTA scalar recurrence (pipeline_scalar/mod.rs:50-52): t' = midpoint(1,
// sqrt(4 t^2 + 1)) (f32::midpoint widened to f64), factor = (t - 1) / t'.  Returns the factor
// and advances term.
float fista_factor(float &term)
{
    const float nt = (float)(((double)1.0f + (double)sqrtf(4.0f * (term * term) + 1.0f)) / 2.0);
    const float f = (term - 1.0f) / nt;
    term = nt;
    return f;
}

// The plan's scalar loop parameters, as the reference derives them in f32 (compute_step_tv.rs:70,
// compute_step.rs:30,47 with compute_step_tv2.rs:88-89, mod.rs:45-52,71).
void set_loop_scalars(artefact_hip_plan *pl, const artefact_params *p)
{
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
    const float Cf = (float)pl->C;
    pl->alpha_tv = 1.0f / sqrtf(Cf);
    pl->alpha_tgv = (pl->weight / sqrtf(2.0f)) / sqrtf(Cf);
    const float radius = sqrtf((float)pl->N) / 2.0f;
    pl->step = radius / sqrtf(1.0f + (float)pl->iterations);
}

// Geometry fields of channel c (pointers excluded) and its quantisation tables: q as f32 and the
// correctly rounded f64 reciprocal of q*q, the product rounded to f32 first (compute_step_prob.rs:29).
void set_channel(artefact_hip_plan *pl, uint32_t c, const artefact_channel &k, float *quant, double *rqq)
{
    for (int j = 0; j < 64; j++) {
        const float q = (float)k.quant[j];
        quant[j] = q;
        rqq[j] = 1.0 / (double)(q * q);
        pl->qtab[c][j] = q;
    }
    ChanDev &d = pl->ch[c];
    d.rw = k.rounded_w;
    d.rh = k.rounded_h;
    d.bw = k.block_w;
    d.bh = k.block_h;
    d.hs = k.h_samp;
    d.vs = k.v_samp;
    d.resample = (k.rounded_w != pl->W || k.rounded_h != pl->H) ? 1u : 0u;
    if (!d.resample) d.mode = PM_FULL;
    else if (d.hs == 1 && d.vs == 1) d.mode = PM_SUB11;
    else if (d.hs == 2 && d.vs == 2) d.mode = PM_SUB22;
    else if (d.hs == 2) d.mode = PM_SUB21;
    else d.mode = PM_SUB12;
    d.alpha_p = pl->pweight[c] != 0.0f ? pl->pweight[c] * 2.0f * 255.0f * sqrtf(2.0f) : 0.0f;
}

// The projection kernel the channels allow (fused layouts, lane-kernel units) and the untuned
// defaults with their environment overrides.
void set_layout(artefact_hip_plan *pl)
{
    for (uint32_t c = pl->C; c < 3; c++) pl->ch[c] = pl->ch[0];
    if (pl->C == 1 && pl->ch[0].mode == PM_FULL) pl->layout = LAYOUT_Y;
    else if (pl->C == 3 && pl->ch[0].mode == PM_FULL && pl->ch[1].mode == PM_FULL && pl->ch[2].mode == PM_FULL)
        pl->layout = LAYOUT_444;
    else if (pl->C == 3 && pl->ch[0].mode == PM_FULL && pl->ch[1].mode == PM_SUB22 && pl->ch[2].mode == PM_SUB22)
        pl->layout = LAYOUT_420;
    else pl->layout = LAYOUT_GENERIC;
    pl->layout_fused = pl->layout;
    uint32_t acc = 0;
    for (uint32_t c = 0; c < 3; c++) {
        if (c < pl->C) {
            const uint32_t bpw = pl->ch[c].mode == PM_SUB22 ? 4 : 8;  // blocks per wave, k_project_lane
            acc += ((pl->ch[c].bw + bpw - 1) / bpw) * pl->ch[c].bh;
        }
        pl->lane_units[c] = acc;
    }
    // Untuned default: the LDS-free lane kernel, measured on MI355X faster than the LDS kernel
    // (k_project_fused) on every fused layout (4096x3072 4:4:4: 134 vs 194 us).
    pl->lane_project = 1;
    if (const char *e = getenv("ARTEFACT_PROJECT_KERNEL")) pl->lane_project = strcmp(e, "lds") != 0;
    // k_project_lane's grid: three workgroups per CU (each forms the norms once, then its
    // waves take several units), measured on MI355X faster than one unit per wave (2048x1536
    // 4:2:0: 29.5 against 35 us, the difference being the per-workgroup norm evaluation).
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, pl->device) != hipSuccess || cus < 1)
        cus = 256;
    pl->lane_wgs = 3u * (uint32_t)cus;
    if (const char *e = getenv("ARTEFACT_LANE_WGS")) pl->lane_wgs = (uint32_t)atoi(e);
    // Projection workgroups by XCD row band where the iteration's planes (x, x_prev, y, g: 16 C N
    // bytes) fit the 8 XCDs' L2s (4 MiB each); the autotuner times the other choice too.
    pl->xcd_bands = 16.0 * (double)pl->C * (double)pl->N <= 32.0 * 1048576.0;
    if (const char *e = getenv("ARTEFACT_XCD_BANDS")) pl->xcd_bands = atoi(e) != 0;
    if (const char *e = getenv("ARTEFACT_GENERIC_PROJECT"))
        if (atoi(e)) pl->layout = LAYOUT_GENERIC;
    // Lets callers of the one-shot entry points (artefact_hip_compute / _process, batches) opt
    // into the fast gradient arithmetic without an API change.
    if (const char *e = getenv("ARTEFACT_PRECISION"))
        if (strcmp(e, "f32_fast") == 0) pl->precision = ARTEFACT_PRECISION_F32_FAST;
}

// Output stage on `stream` from device planes (y, cb, cr; cb/cr ignored unless colour) into
// host memory, staging the bytes in `staging` (at least real_w*real_h*3 bytes of device
// memory).  Synchronises the stream.
int run_colour(hipStream_t stream, const float *y, const float *cb, const float *cr, uint32_t stride,
               uint32_t real_w, uint32_t real_h, bool colour, uint8_t *staging, uint8_t *out)
{
    ColourArgs a{y, cb, cr, stride, real_w, real_h, colour ? 1u : 0u, staging};
    const dim3 grid((real_w + 4 * NT - 1) / (4 * NT), std::min(real_h, 65535u));
    hipLaunchKernelGGL(k_colour, grid, dim3(NT), 0, stream, a);
    bool ok = hipGetLastError() == hipSuccess;
    ok = ok && hipMemcpyAsync(out, staging, (size_t)real_w * real_h * 3, hipMemcpyDeviceToHost, stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(stream) == hipSuccess;
    return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

// ---- Autotuner: per geometry, the k_gradient strip height and the projection kernel ----
// Candidates are timed on the plan itself (a few iterations after a fresh init, replayed
// from a graph: time_loop) and the winner is cached for the process, per device and geometry,
// under g_tune_mu.
struct TuneChoice {
    int rows, layout, lane;
    uint32_t wgs;  // k_project_lane's workgroup cap
    float ms;
    bool bands;    // projection workgroups by XCD row band
};
std::mutex g_tune_mu;
std::map<std::string, TuneChoice> g_tune_cache;

std::string tune_key(const artefact_hip_plan *pl)
{
    std::string k = "d" + std::to_string(pl->device) + " " + std::to_string(pl->W) + "x" + std::to_string(pl->H) +
                    " C" + std::to_string(pl->C) + (pl->tgv ? " tgv" : " tv") + " p" + std::to_string(pl->precision) +
                    (pl->wide ? " w1" : " w0");  // the two k_gradient instances differ in registers and occupancy
    for (uint32_t c = 0; c < pl->C; c++)
        k += " " + std::to_string(pl->ch[c].rw) + "x" + std::to_string(pl->ch[c].rh) + "/" +
             std::to_string(pl->ch[c].hs) + std::to_string(pl->ch[c].vs);
    // A choice made under environment pins is not one to reuse without them, nor the reverse.
    for (const char *pin : {"ARTEFACT_KG_ROWS", "ARTEFACT_LANE_WGS", "ARTEFACT_PROJECT_KERNEL",
                            "ARTEFACT_GENERIC_PROJECT", "ARTEFACT_XCD_BANDS"})
        if (const char *v = getenv(pin)) k += std::string(" ") + pin + "=" + v;
    return k;
}

void set_rows(artefact_hip_plan *pl, int rows)
{
    pl->rows = rows;
    pl->grid_g = dim3((pl->W + KG_COLS - 1) / KG_COLS, (pl->H + WAVES * rows - 1) / (WAVES * rows));
}

// Mean milliseconds per iteration of `nt` iterations (nt <= the plan's iterations) after a
// fresh init, replayed from a graph as the plan runs them: timed eagerly, launch gaps would
// favour the variants with fewer kernels per iteration.  The first replay warms caches and
// code; the faster of the next two is taken.  < 0 on failure.
float time_loop(artefact_hip_plan *pl, uint32_t nt)
{
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    bool ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
              hipEventCreate(&e2) == hipSuccess;
    float ms = -1.0f, ms2 = -1.0f;
    if (ok) {
        launch_init(pl, pl->stream);
        ok = hipStreamBeginCapture(pl->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            // Each replay starts from cleared accumulators, as launch_iterations does.
            hipLaunchKernelGGL(k_zero_acc, dim3(1), dim3(NT), 0, pl->stream, pl->acc);
            int cur = 0;
            for (uint32_t it = 0; it < nt; it++) {
                IterArgs a = make_args(pl, cur, pl->factors[it], false);
                a.step = step_at(pl, it);
                launch_gradient(pl, a, pl->stream);
                launch_project(pl, a, pl->stream);
                cur = 1 - cur;
            }
            const bool launched = hipGetLastError() == hipSuccess;
            ok = hipStreamEndCapture(pl->stream, &g) == hipSuccess && g && launched;
        }
        ok = ok && hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) == hipSuccess;
        ok = ok && hipGraphLaunch(ge, pl->stream) == hipSuccess && hipEventRecord(e0, pl->stream) == hipSuccess &&
             hipGraphLaunch(ge, pl->stream) == hipSuccess && hipEventRecord(e1, pl->stream) == hipSuccess &&
             hipGraphLaunch(ge, pl->stream) == hipSuccess && hipEventRecord(e2, pl->stream) == hipSuccess &&
             hipEventSynchronize(e2) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
             hipEventElapsedTime(&ms2, e1, e2) == hipSuccess;
    }
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    for (hipEvent_t e : {e0, e1, e2})
        if (e) (void)hipEventDestroy(e);
    return ok ? std::min(ms, ms2) / (float)nt : -1.0f;
}

void drop_graph(artefact_hip_plan *pl)
{
    if (pl->graph) (void)hipGraphExecDestroy(pl->graph);
    pl->graph = nullptr;
}

// Picks (rows, layout) for the plan, from the cache or by timing; leaves the plan
// uninitialised.  Environment overrides (ARTEFACT_KG_ROWS, ARTEFACT_GENERIC_PROJECT) pin
// their parameter.
int tune_plan(artefact_hip_plan *pl, TuneChoice *out)
{
    const std::string key = tune_key(pl);
    TuneChoice best{pl->rows, pl->layout, pl->lane_project, pl->lane_wgs, -1.0f, pl->xcd_bands};
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) best = it->second;
    }
    if (best.ms < 0.0f && pl->iterations > 0) {
        const TuneChoice dflt = best;  // the plan's untuned configuration
        auto apply = [pl](const TuneChoice &c) {
            pl->layout = c.layout;
            pl->lane_project = c.lane;
            pl->lane_wgs = c.wgs;
            pl->xcd_bands = c.bands;
            set_rows(pl, c.rows);
        };
        const uint32_t nt = std::min<uint32_t>(pl->iterations, 16);
        // Warm the GPU (clocks, caches, code) on the default configuration before timing
        // anything: the first loops of a cold process run slower and would favour whichever
        // candidate comes later.
        for (int w = 0; w < 2; w++)
            if (time_loop(pl, nt) < 0.0f) return ARTEFACT_E_HIP;
        const int r0 = pl->rows;
        std::vector<int> rows{r0};
        if (!getenv("ARTEFACT_KG_ROWS")) {
            // Around the single-wave height, and the heights that give every CU exactly
            // per_cu - 1 or per_cu - 2 workgroups: the sweep is VALU-bound per SIMD, so its time
            // follows the most loaded CU's (workgroups x strip steps).
            int per_cu = 0, cus = 0;
            gradient_residency(pl->device, pl->C, pl->tgv, pl->precision == ARTEFACT_PRECISION_F32_FAST, pl->wide,
                               per_cu, cus);
            // Shorter strips run several waves of workgroups at full occupancy: on large images,
            // where the sweep waits on memory, that beats a single wave of workgroups with few
            // waves per SIMD (4096x3072 4:4:4: 337 us per iteration at RW = 18 against 351 us
            // at the single-wave 70).
            std::vector<int> cand{r0 + 1, r0 + 2, r0 - 1, r0 - 2, 2 * r0, r0 / 2, r0 / 3, r0 / 4};
            for (int k = per_cu - 1; k >= std::max(1, per_cu - 2); k--) {
                const int rk = rows_for_workgroups(pl->W, pl->H, (uint64_t)k * cus);
                cand.push_back(rk);
                cand.push_back(rk + 1);
            }
            for (int r : cand)
                if (r >= 2 && r <= (int)pl->H && std::find(rows.begin(), rows.end(), r) == rows.end()) rows.push_back(r);
        }
        // Projection candidates: (layout, lane kernel) -- the fused layout with the LDS-free and
        // the LDS kernel, and the per-channel kernel.
        std::vector<std::pair<int, int>> projs{{pl->layout, pl->lane_project}};
        if (pl->layout != LAYOUT_GENERIC && !getenv("ARTEFACT_PROJECT_KERNEL")) projs.push_back({pl->layout, 1 - pl->lane_project});
        if (pl->layout != LAYOUT_GENERIC) projs.push_back({LAYOUT_GENERIC, 0});
        for (const auto &pj : projs)
            for (int r : rows) {
                pl->layout = pj.first;
                pl->lane_project = pj.second;
                set_rows(pl, r);
                const float t = time_loop(pl, nt);
                if (t < 0.0f) return ARTEFACT_E_HIP;
                if (best.ms < 0.0f || t < best.ms) best = TuneChoice{r, pj.first, pj.second, pl->lane_wgs, t, pl->xcd_bands};
            }
        // The lane kernel's workgroup count on the winner: 2-4 per CU, or one unit per wave.
        if (best.lane && best.layout != LAYOUT_GENERIC && !getenv("ARTEFACT_LANE_WGS")) {
            int per_cu = 0, cus = 0;
            gradient_residency(pl->device, pl->C, pl->tgv, false, false, per_cu, cus);
            pl->layout = best.layout;
            pl->lane_project = best.lane;
            set_rows(pl, best.rows);
            for (uint32_t w : {2u * cus, 4u * cus, 0u}) {
                if (w == best.wgs) continue;
                pl->lane_wgs = w;
                const float t = time_loop(pl, nt);
                if (t < 0.0f) return ARTEFACT_E_HIP;
                if (t < best.ms) best = TuneChoice{best.rows, best.layout, best.lane, w, t, best.bands};
            }
        }
        // XCD bands on the winner: the other choice.
        if (!getenv("ARTEFACT_XCD_BANDS")) {
            TuneChoice other = best;
            other.bands = !best.bands;
            apply(other);
            const float t = time_loop(pl, nt);
            if (t < 0.0f) return ARTEFACT_E_HIP;
            if (t < best.ms) {
                best = other;
                best.ms = t;
            }
        }
        // Guard against a noisy pick: time the winner and the default alternately, twice each,
        // and keep the winner only if it is faster both times.
        if (best.rows != dflt.rows || best.layout != dflt.layout || best.lane != dflt.lane || best.wgs != dflt.wgs ||
            best.bands != dflt.bands) {
            int wins = 0;
            float tb = 0.0f, td = 0.0f;
            for (int round = 0; round < 2; round++) {
                apply(best);
                tb = time_loop(pl, nt);
                apply(dflt);
                td = time_loop(pl, nt);
                if (tb < 0.0f || td < 0.0f) return ARTEFACT_E_HIP;
                wins += tb < td ? 1 : 0;
            }
            if (wins < 2) {
                best = dflt;
                best.ms = td;
            }
        }
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tune_cache[key] = best;
    }
    pl->layout = best.layout;
    pl->lane_project = best.lane;
    pl->lane_wgs = best.wgs;
    pl->xcd_bands = best.bands;
    set_rows(pl, best.rows);
    drop_graph(pl);
    pl->initialized = false;
    if (out) *out = best;
    return ARTEFACT_OK;
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
    case ARTEFACT_E_JPEG: return "malformed or unsupported JPEG data";
    case ARTEFACT_E_SCRATCH: return "grid too large for the no-allocation entry point's static scratch";
    case ARTEFACT_BENCHMARK: return "benchmark run: no image written";
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
    set_loop_scalars(pl, p);
    // Host vectors sized by the caller's iteration count and grid: an allocation failure is
    // an error code, never an exception through the C interface.
    try {
        pl->factors.resize(pl->iterations);
        float term = 1.0f;
        for (uint32_t i = 0; i < pl->iterations; i++) pl->factors[i] = fista_factor(term);

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
    } catch (const std::bad_alloc &) {
        delete pl;
        return ARTEFACT_E_HIP;
    }

    DeviceGuard guard(pl->device);
    size_t mtot = 0;
    for (uint32_t c = 0; c < pl->C; c++) {
        pl->M[c] = (size_t)chs[c].rounded_w * chs[c].rounded_h;
        mtot += pl->M[c];
    }
    const char *wide_env = getenv("ARTEFACT_WIDE_ADDRESSING");
    pl->wide = (uint64_t)pl->C * pl->N * 4 >= (1ull << 31) || (wide_env && wide_env[0] == '1');
    pl->rows = pick_rows(pl->device, pl->W, pl->H, pl->C, pl->tgv, false, pl->wide);
    pl->grid_g = dim3((pl->W + KG_COLS - 1) / KG_COLS, (pl->H + WAVES * pl->rows - 1) / (WAVES * pl->rows));
    pl->grid_p = dim3((pl->W + TW - 1) / TW, (pl->H + TH - 1) / TH);
    bool ok = hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking) == hipSuccess;
    // One device allocation holds every buffer, each at a 256-byte boundary: a one-shot call
    // (artefact_hip_compute / _process) pays one hipMalloc and one hipFree, not 13 of each.
    const size_t cn = (size_t)pl->C * pl->N;
    size_t arena_bytes = 0;
    auto take = [&arena_bytes](size_t bytes) {
        const size_t at = arena_bytes;
        arena_bytes += (bytes + 255) & ~(size_t)255;
        return at;
    };
    const size_t at_x = take(cn * 4), at_xp = take(cn * 4), at_g = take(cn * 4), at_y = take(cn * 4);
    const size_t at_cos = take(mtot * 4), at_prob = take(mtot * 4), at_coef = take(mtot * 2);
    const size_t at_quant = take(3 * 64 * 4), at_rqq = take(3 * 64 * 8), at_acc = take(2 * ACC_WORDS * 8);
    const size_t at_norms = take(3 * 4), at_rnorm = take(3 * 8), at_nsq = take(3 * 8);
    ok = ok && hipMalloc(&pl->arena, arena_bytes) == hipSuccess;
    // A failed allocation stays the thread's last HIP error until read; it is handled here, so it
    // must not surface as a later launch's error.
    if (!ok) (void)hipGetLastError();
    if (ok) {
        char *base = static_cast<char *>(pl->arena);
        pl->buf[0] = reinterpret_cast<float *>(base + at_x);
        pl->buf[1] = reinterpret_cast<float *>(base + at_xp);
        pl->g = reinterpret_cast<float *>(base + at_g);
        pl->ybuf = reinterpret_cast<float *>(base + at_y);
        pl->cos = reinterpret_cast<float *>(base + at_cos);
        pl->prob = reinterpret_cast<float *>(base + at_prob);
        pl->coef = reinterpret_cast<int16_t *>(base + at_coef);
        pl->quant = reinterpret_cast<float *>(base + at_quant);
        pl->rqq = reinterpret_cast<double *>(base + at_rqq);
        pl->acc = reinterpret_cast<long long *>(base + at_acc);
        pl->norms = reinterpret_cast<float *>(base + at_norms);
        pl->rnorm = reinterpret_cast<double *>(base + at_rnorm);
        pl->norm_sq = reinterpret_cast<double *>(base + at_nsq);
    }
    if (!ok) {
        free_plan(pl);
        return ARTEFACT_E_HIP;
    }
    std::vector<float> quant(3 * 64, 1.0f);
    std::vector<double> rqq(3 * 64, 1.0);
    size_t moff = 0;
    for (uint32_t c = 0; c < pl->C; c++) {
        const artefact_channel &k = chs[c];
        set_channel(pl, c, k, &quant[c * 64], &rqq[c * 64]);
        ok = ok && hipMemcpyAsync(pl->coef + moff, k.coefs, pl->M[c] * sizeof(int16_t), hipMemcpyHostToDevice,
                                  pl->stream) == hipSuccess;
        ChanDev &d = pl->ch[c];
        d.coef = pl->coef + moff;
        d.cos = pl->cos + moff;
        d.prob = pl->prob + moff;
        d.quant = pl->quant + c * 64;
        d.rqq = pl->rqq + c * 64;
        moff += pl->M[c];
    }
    set_layout(pl);
    ok = ok && hipMemcpyAsync(pl->quant, quant.data(), quant.size() * sizeof(float), hipMemcpyHostToDevice,
                              pl->stream) == hipSuccess;
    ok = ok && hipMemcpyAsync(pl->rqq, rqq.data(), rqq.size() * sizeof(double), hipMemcpyHostToDevice,
                              pl->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(pl->norms, 0, 3 * sizeof(float), pl->stream) == hipSuccess;
    ok = ok && hipMemsetAsync(pl->rnorm, 0, 3 * sizeof(double), pl->stream) == hipSuccess;
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
    if (pl->init_override.empty()) {
        launch_init(pl, pl->stream);
    } else {
        for (uint32_t c = 0; c < pl->C; c++) {
            const ChanDev &d = pl->ch[c];
            float *x = pl->buf[0] + c * pl->N, *xp = pl->buf[1] + c * pl->N;
            ok = ok && hipMemcpyAsync(x, pl->init_override[c].data(), pl->N * sizeof(float),
                                      hipMemcpyHostToDevice, pl->stream) == hipSuccess;
            ok = ok && hipMemcpyAsync(xp, pl->init_override[c].data(), pl->N * sizeof(float),
                                      hipMemcpyHostToDevice, pl->stream) == hipSuccess;
            hipLaunchKernelGGL(k_init_cos, dim3((unsigned)((pl->M[c] + NT - 1) / NT)), dim3(NT), 0, pl->stream, d);
        }
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
            rc = launch_iterations(pl, pl->stream);
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
    }
    const bool timed = pl->timing_used < pl->timing_events.size() / 2;
    if (timed && hipEventRecord(pl->timing_events[2 * pl->timing_used], pl->stream) != hipSuccess)
        return ARTEFACT_E_HIP;
    if (pl->use_graph) {
        if (hipGraphLaunch(pl->graph, pl->stream) != hipSuccess) return ARTEFACT_E_HIP;
    } else {
        rc = launch_iterations(pl, pl->stream);
        if (rc) return rc;
    }
    if (timed) {
        if (hipEventRecord(pl->timing_events[2 * pl->timing_used + 1], pl->stream) != hipSuccess)
            return ARTEFACT_E_HIP;
        pl->timing_used++;
    }
    if (pl->iterations % 2) pl->cur = 1 - pl->cur;
    // The loop consumed the start image; a new solve needs a new init (the reference
    // never restarts the FISTA sequence from an iterate).
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
    const IterArgs ga = make_args(pl, pl->cur, 0.0f, false);
    // The gradient and its norms (last_norms), from a cleared accumulator.
    hipLaunchKernelGGL(k_zero_acc, dim3(1), dim3(NT), 0, pl->stream, pl->acc);
    launch_gradient(pl, ga, pl->stream);
    if (pl->C == 1) hipLaunchKernelGGL(k_norms<1>, dim3(1), dim3(NT), 0, pl->stream, ga);
    else if (pl->C == 2) hipLaunchKernelGGL(k_norms<2>, dim3(1), dim3(NT), 0, pl->stream, ga);
    else hipLaunchKernelGGL(k_norms<3>, dim3(1), dim3(NT), 0, pl->stream, ga);
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
        hipLaunchKernelGGL(k_zero_acc, dim3(1), dim3(NT), 0, pl->stream, pl->acc);
        ok = hipEventRecord(ev[0], pl->stream) == hipSuccess;
        for (uint32_t it = 0; ok && it < n; it++) {
            IterArgs a = make_args(pl, cur, pl->factors[it], it + 1 == n);
            a.step = step_at(pl, it);
            launch_gradient(pl, a, pl->stream);
            ok = ok && hipEventRecord(ev[2 * it + 1], pl->stream) == hipSuccess;
            launch_project(pl, a, pl->stream);
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

int artefact_hip_plan_set_timing(artefact_hip_plan *pl, uint32_t launches)
{
    if (!pl) return ARTEFACT_E_ARGUMENT;
    DeviceGuard guard(pl->device);
    if (pl->stream && hipStreamSynchronize(pl->stream) != hipSuccess) return ARTEFACT_E_HIP;
    for (hipEvent_t e : pl->timing_events) (void)hipEventDestroy(e);
    pl->timing_events.clear();
    pl->timing_used = 0;
    for (uint32_t i = 0; i < 2 * launches; i++) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return ARTEFACT_E_HIP;
        pl->timing_events.push_back(e);
    }
    return ARTEFACT_OK;
}

int artefact_hip_plan_timing(artefact_hip_plan *pl, float *total_ms, uint32_t *launches)
{
    if (!pl || !total_ms || !launches) return ARTEFACT_E_ARGUMENT;
    DeviceGuard guard(pl->device);
    double total = 0.0;
    for (uint32_t i = 0; i < pl->timing_used; i++) {
        float ms = 0.0f;
        if (hipEventSynchronize(pl->timing_events[2 * i + 1]) != hipSuccess ||
            hipEventElapsedTime(&ms, pl->timing_events[2 * i], pl->timing_events[2 * i + 1]) != hipSuccess)
            return ARTEFACT_E_HIP;
        total += ms;
    }
    *total_ms = (float)total;
    *launches = pl->timing_used;
    return ARTEFACT_OK;
}

int artefact_hip_plan_objective(artefact_hip_plan *pl, double *terms)
{
    if (!pl || !terms) return ARTEFACT_E_ARGUMENT;
    DeviceGuard guard(pl->device);
    ObjArgs a;
    memset(&a, 0, sizeof(a));
    a.W = pl->W;
    a.H = pl->H;
    a.C = pl->C;
    for (uint32_t c = 0; c < 3; c++) {
        a.x[c] = pl->buf[pl->cur] + (size_t)(c < pl->C ? c : 0) * pl->N;
        a.ch[c] = pl->ch[c < pl->C ? c : 0];
        a.ap[c] = c < pl->C ? (double)pl->ch[c].alpha_p : 0.0;
    }
    const float Cf = (float)pl->C;
    a.alpha = (double)(1.0f / sqrtf(Cf));
    a.a2 = (double)((pl->weight / sqrtf(2.0f)) / sqrtf(Cf));
    const unsigned int pix_wgs = (unsigned int)((pl->N + NT - 1) / NT);
    size_t maxb = 0;
    for (uint32_t c = 0; c < pl->C; c++) maxb = std::max(maxb, (size_t)pl->ch[c].bw * pl->ch[c].bh);
    const unsigned int blk_wgs = (unsigned int)((maxb + NT - 1) / NT);
    double *part = nullptr, *out = nullptr;
    bool ok = hipMalloc(&part, ((size_t)pix_wgs + blk_wgs) * 6 * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&out, 3 * sizeof(double)) == hipSuccess;
    if (!ok) (void)hipGetLastError();  // handled here: not a later launch's error
    if (ok) {
        a.partials = part;
        a.pixel_wgs = pix_wgs;
        if (pl->C == 1) {
            hipLaunchKernelGGL(k_obj_pixels<1>, dim3(pix_wgs), dim3(NT), 0, pl->stream, a);
            hipLaunchKernelGGL(k_obj_blocks<1>, dim3(blk_wgs), dim3(NT), 0, pl->stream, a);
        } else if (pl->C == 2) {
            hipLaunchKernelGGL(k_obj_pixels<2>, dim3(pix_wgs), dim3(NT), 0, pl->stream, a);
            hipLaunchKernelGGL(k_obj_blocks<2>, dim3(blk_wgs), dim3(NT), 0, pl->stream, a);
        } else {
            hipLaunchKernelGGL(k_obj_pixels<3>, dim3(pix_wgs), dim3(NT), 0, pl->stream, a);
            hipLaunchKernelGGL(k_obj_blocks<3>, dim3(blk_wgs), dim3(NT), 0, pl->stream, a);
        }
        hipLaunchKernelGGL(k_obj_final, dim3(1), dim3(NT), 0, pl->stream, (const double *)part, pix_wgs + blk_wgs,
                           out);
        ok = hipGetLastError() == hipSuccess;
        ok = ok && hipMemcpyAsync(terms, out, 3 * sizeof(double), hipMemcpyDeviceToHost, pl->stream) == hipSuccess;
        ok = ok && hipStreamSynchronize(pl->stream) == hipSuccess;
    }
    if (part) (void)hipFree(part);
    if (out) (void)hipFree(out);
    return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
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

int artefact_hip_plan_download_rgb(artefact_hip_plan *pl, uint32_t real_w, uint32_t real_h, uint8_t *out)
{
    if (!pl || !out) return ARTEFACT_E_ARGUMENT;
    if (!real_w || !real_h || real_w > pl->W || real_h > pl->H) return ARTEFACT_E_GEOMETRY;
    DeviceGuard guard(pl->device);
    const float *x = pl->buf[pl->cur];
    // The gradient buffer (C*N floats >= 4N bytes) stages the 3*real_w*real_h bytes.
    return run_colour(pl->stream, x, x + (pl->C > 1 ? pl->N : 0), x + (pl->C > 2 ? 2 * pl->N : 0), pl->W, real_w,
                      real_h, pl->C == 3, reinterpret_cast<uint8_t *>(pl->g), out);
}

int artefact_hip_process(const artefact_process_params *p, const artefact_channel *ch, uint8_t *out)
{
    if (!p || !ch) return ARTEFACT_E_ARGUMENT;
    if (p->nchannel < 1 || p->nchannel > 3) return ARTEFACT_E_ARGUMENT;
    if (!p->benchmark && !out) return ARTEFACT_E_ARGUMENT;
    uint32_t W = 0, H = 0;
    for (uint32_t c = 0; c < p->nchannel; c++) {
        W = std::max(W, ch[c].rounded_w);
        H = std::max(H, ch[c].rounded_h);
    }
    if (!p->real_w || !p->real_h || p->real_w > W || p->real_h > H) return ARTEFACT_E_GEOMETRY;
    const bool joint = p->nchannel == 3 && !p->separate_components;
    const uint32_t nplan = joint ? 1 : p->nchannel;
    artefact_hip_plan *plans[3] = {nullptr, nullptr, nullptr};
    int rc = ARTEFACT_OK;
    for (uint32_t k = 0; k < nplan && !rc; k++) {
        artefact_params q{};
        q.nchannel = joint ? 3 : 1;
        q.max_w = W;
        q.max_h = H;
        q.weight = p->weight[k];
        for (int c = 0; c < 3; c++) q.pweight[c] = p->pweight[c];
        q.iterations = p->iterations[k];
        q.mode = ARTEFACT_MODE_TGV;
        q.init_override = nullptr;
        rc = artefact_hip_plan_create(&q, ch + k, &plans[k]);
        if (!rc) plans[k]->use_graph = false;
    }
    // Every loop is queued before any is awaited, so separate components overlap on the GPU.
    for (uint32_t k = 0; k < nplan && !rc; k++) rc = artefact_hip_plan_init(plans[k]);
    for (uint32_t k = 0; k < nplan && !rc; k++) rc = artefact_hip_plan_iterate(plans[k]);
    for (uint32_t k = 1; k < nplan && !rc; k++) {
        DeviceGuard guard(plans[k]->device);
        if (hipStreamSynchronize(plans[k]->stream) != hipSuccess) rc = ARTEFACT_E_HIP;
    }
    if (!rc && p->benchmark) {
        DeviceGuard guard(plans[0]->device);
        rc = hipStreamSynchronize(plans[0]->stream) == hipSuccess ? ARTEFACT_BENCHMARK : ARTEFACT_E_HIP;
    } else if (!rc) {
        artefact_hip_plan *p0 = plans[0];
        DeviceGuard guard(p0->device);
        const float *planes[3];
        for (uint32_t c = 0; c < p->nchannel; c++)
            planes[c] = joint ? p0->buf[p0->cur] + (size_t)c * p0->N : plans[c]->buf[plans[c]->cur];
        for (uint32_t c = p->nchannel; c < 3; c++) planes[c] = planes[0];
        rc = run_colour(p0->stream, planes[0], planes[1], planes[2], W, p->real_w, p->real_h, p->nchannel == 3,
                        reinterpret_cast<uint8_t *>(p0->g), out);
    }
    for (uint32_t k = 0; k < nplan; k++) artefact_hip_plan_destroy(plans[k]);
    return rc;
}

int artefact_hip_plan_set_precision(artefact_hip_plan *pl, uint32_t precision)
{
    if (!pl) return ARTEFACT_E_ARGUMENT;
    if (precision != ARTEFACT_PRECISION_F32 && precision != ARTEFACT_PRECISION_BF16X3 &&
        precision != ARTEFACT_PRECISION_F32_FAST)
        return ARTEFACT_E_ARGUMENT;
    if (precision == ARTEFACT_PRECISION_BF16X3) {
        // Every channel at full resolution, or 4:2:0 colour (full-resolution luma, both chroma
        // components 2 x 2 subsampled).
        const bool sub420 = pl->C == 3 && pl->ch[0].mode == PM_FULL && pl->ch[1].mode == PM_SUB22 &&
                            pl->ch[2].mode == PM_SUB22;
        if (!sub420)
            for (uint32_t c = 0; c < pl->C; c++)
                if (pl->ch[c].mode != PM_FULL) return ARTEFACT_E_GEOMETRY;
    }
    if (pl->precision != precision) {
        DeviceGuard guard(pl->device);
        if (pl->stream) (void)hipStreamSynchronize(pl->stream);
        drop_graph(pl);
    }
    pl->precision = precision;
    return ARTEFACT_OK;
}

int artefact_hip_plan_set_momentum_offset(artefact_hip_plan *pl, uint32_t k)
{
    if (!pl) return ARTEFACT_E_ARGUMENT;
    if (k > (1u << 20)) return ARTEFACT_E_ARGUMENT;
    float term = 1.0f;
    for (uint32_t i = 0; i < k; i++) (void)fista_factor(term);
    for (uint32_t i = 0; i < pl->iterations; i++) pl->factors[i] = fista_factor(term);
    DeviceGuard guard(pl->device);
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    drop_graph(pl);
    return ARTEFACT_OK;
}

int artefact_hip_plan_set_step_schedule(artefact_hip_plan *pl, uint32_t first, uint32_t last)
{
    const int rc = artefact_hip_plan_set_step_iterations(pl, last);  // the constant step; clears a schedule
    if (rc != ARTEFACT_OK || first == 0) return rc;
    const float radius = sqrtf((float)pl->N) / 2.0f;
    const float s1 = radius / sqrtf(1.0f + (float)first), s2 = pl->step;
    try {
        pl->steps.resize(pl->iterations);
    } catch (...) {
        return ARTEFACT_E_HIP;
    }
    for (uint32_t i = 0; i < pl->iterations; i++)
        pl->steps[i] = pl->iterations > 1 ? s1 * powf(s2 / s1, (float)i / (float)(pl->iterations - 1)) : s2;
    return ARTEFACT_OK;
}

int artefact_hip_plan_set_step_iterations(artefact_hip_plan *pl, uint32_t n)
{
    if (!pl) return ARTEFACT_E_ARGUMENT;
    const float radius = sqrtf((float)pl->N) / 2.0f;
    const uint32_t k = n ? n : pl->iterations;
    pl->step = radius / sqrtf(1.0f + (float)k);  // pipeline_scalar/mod.rs:45,64-74 with k in place of iterations
    pl->steps.clear();
    DeviceGuard guard(pl->device);
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    drop_graph(pl);
    return ARTEFACT_OK;
}

int artefact_hip_plan_warm_init(artefact_hip_plan *pl, artefact_cnn *net)
{
    if (!pl || !net) return ARTEFACT_E_ARGUMENT;
    DeviceGuard guard(pl->device);
    // The buffers are overwritten from here on: a failure below leaves no start image.
    pl->initialized = false;
    // The decode: each channel's component plane (rw x rh) in the gradient buffer, and cos /
    // prob; x and x_prev are written below, from the prediction.
    launch_init(pl, pl->stream, false);
    // The CNN runs on every channel at its component resolution, as the decoder produced it
    // (subsampled chroma at a quarter of the grid's pixels), channels of equal size in one call.
    // Its prediction goes to the channel's prob plane (M_c floats, free until the projection
    // below writes it) and is upsampled bilinearly onto the grid, into the y buffer the
    // projection reads.
    bool done[3] = {false, false, false};
    for (uint32_t c0 = 0; c0 < pl->C; c0++) {
        if (done[c0]) continue;
        const float *planes[3];
        float *pred[3];
        float quant[3 * 64];
        uint32_t n = 0;
        for (uint32_t c = c0; c < pl->C; c++) {
            if (done[c] || pl->ch[c].rw != pl->ch[c0].rw || pl->ch[c].rh != pl->ch[c0].rh) continue;
            planes[n] = pl->g + (size_t)c * pl->N;
            pred[n] = pl->ch[c].prob;
            memcpy(quant + n * 64, pl->qtab[c], 64 * sizeof(float));
            done[c] = true;
            n++;
        }
        const int rc = cnn_predict_planes(net, planes, quant, n, pl->ch[c0].rh, pl->ch[c0].rw, pred, pl->stream);
        if (rc) return rc;
    }
    UpsampleArgs ua{};
    for (uint32_t c = 0; c < pl->C; c++) {
        ua.plane[c] = pl->ch[c].prob;
        ua.out[c] = pl->ybuf + (size_t)c * pl->N;
        ua.rw[c] = pl->ch[c].rw;
        ua.rh[c] = pl->ch[c].rh;
        ua.hs[c] = pl->ch[c].hs;
        ua.vs[c] = pl->ch[c].vs;
    }
    ua.W = pl->W;
    ua.H = pl->H;
    hipLaunchKernelGGL(k_upsample_linear, dim3((pl->W / 4 + NT - 1) / NT, std::min<uint32_t>(pl->H, 128u), pl->C),
                       dim3(NT), 0, pl->stream, ua);
    // Projection of the prediction: the projection kernel with zero norms (the accumulator
    // launch_init cleared: no descent step) reads y = prediction and writes the projected
    // image into buf[1] and the prob plane of its clamped coefficients, the state the loop
    // carries into its next gradient.
    const IterArgs pa = make_args(pl, 0, 0.0f, false);
    launch_project(pl, pa, pl->stream);
    bool ok = hipMemcpyAsync(pl->buf[0], pl->buf[1], (size_t)pl->C * pl->N * sizeof(float),
                              hipMemcpyDeviceToDevice, pl->stream) == hipSuccess;
    ok = ok && hipGetLastError() == hipSuccess;
    pl->cur = 0;
    pl->initialized = ok;
    return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

int artefact_hip_plan_tune(artefact_hip_plan *pl, artefact_tuning *out)
{
    if (!pl) return ARTEFACT_E_ARGUMENT;
    DeviceGuard guard(pl->device);
    TuneChoice t{pl->rows, pl->layout, pl->lane_project, pl->lane_wgs, -1.0f};
    const int rc = tune_plan(pl, &t);
    if (rc) return rc;
    if (out) {
        out->kg_rows = (uint32_t)t.rows;
        out->project_kernel = t.layout == LAYOUT_GENERIC ? 0u : (t.lane ? 2u : 1u);
        out->ms_per_iteration = t.ms;
        out->project_workgroups = t.wgs;
        out->xcd_bands = t.bands ? 1u : 0u;
    }
    return ARTEFACT_OK;
}

}  // extern "C"

struct artefact_hip_batch {
    int device = 0;
    std::vector<artefact_hip_plan *> plans;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;  // [0] fork, [k] join of stream k
    hipGraphExec_t graph = nullptr;
    uint32_t streams_used = 0;       // streams the graph spreads the images over
};

namespace {

void free_batch(artefact_hip_batch *b)
{
    if (!b) return;
    DeviceGuard guard(b->device);
    for (hipStream_t s : b->streams)
        if (s) (void)hipStreamSynchronize(s);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (hipEvent_t e : b->events)
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t s : b->streams)
        if (s) (void)hipStreamDestroy(s);
    for (artefact_hip_plan *pl : b->plans) artefact_hip_plan_destroy(pl);
    delete b;
}

}  // namespace

extern "C" {

int artefact_hip_batch_create(const artefact_batch_image *images, uint32_t count, uint32_t nstreams, uint32_t tune,
                              artefact_hip_batch **out)
{
    if (!out) return ARTEFACT_E_ARGUMENT;
    *out = nullptr;
    if (!images || !count) return ARTEFACT_E_ARGUMENT;
    for (uint32_t i = 0; i < count; i++)
        if (!images[i].params || !images[i].channels || images[i].params->init_override) return ARTEFACT_E_ARGUMENT;
    artefact_hip_batch *b = new (std::nothrow) artefact_hip_batch;
    if (!b) return ARTEFACT_E_HIP;
    if (hipGetDevice(&b->device) != hipSuccess) {
        delete b;
        return ARTEFACT_E_HIP;
    }
    int rc = ARTEFACT_OK;
    for (uint32_t i = 0; i < count && !rc; i++) {
        artefact_hip_plan *pl = nullptr;
        rc = artefact_hip_plan_create(images[i].params, images[i].channels, &pl);
        if (!rc) {
            pl->use_graph = false;
            b->plans.push_back(pl);
            if (tune) rc = tune_plan(pl, nullptr);
        }
    }
    if (rc) {
        free_batch(b);
        return rc;
    }
    // Streams: the caller's count, else (tune) the fastest of 1-4, else 2.  On the 64-image
    // mixed workload 2 streams measured fastest (219 ms against 229 with 4 and 249 with 1): two
    // large images at once already fill the GPU, and more concurrent working sets compete for
    // the Infinity Cache.
    const uint32_t most = std::max(1u, std::min(nstreams ? nstreams : 4u, count));
    bool ok = true;
    b->streams.assign(most, nullptr);
    b->events.assign(most, nullptr);
    for (uint32_t k = 0; k < most && ok; k++) {
        ok = hipStreamCreateWithFlags(&b->streams[k], hipStreamNonBlocking) == hipSuccess;
        ok = ok && hipEventCreateWithFlags(&b->events[k], hipEventDisableTiming) == hipSuccess;
    }
    auto cost = [&](uint32_t i) {
        const artefact_hip_plan *pl = b->plans[i];
        return (double)pl->N * pl->C * (1.0 + pl->iterations);
    };
    std::vector<uint32_t> order(count);
    for (uint32_t i = 0; i < count; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return cost(x) > cost(y); });
    // One graph over ns streams: largest solves first, each to the stream with the least work so
    // far; the streams fork off stream 0, every image's init and loop go on its stream, and they
    // join back into stream 0.
    auto build = [&](uint32_t ns, hipGraphExec_t *exec) {
        std::vector<uint32_t> stream_of(count, 0);
        std::vector<double> load(ns, 0.0);
        for (uint32_t i : order) {
            const uint32_t k = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
            stream_of[i] = k;
            load[k] += cost(i);
        }
        hipGraph_t g = nullptr;
        bool cap = hipStreamBeginCapture(b->streams[0], hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (cap) {
            cap = hipEventRecord(b->events[0], b->streams[0]) == hipSuccess;
            for (uint32_t k = 1; k < ns; k++)
                cap = cap && hipStreamWaitEvent(b->streams[k], b->events[0], 0) == hipSuccess;
            for (uint32_t i : order) {
                artefact_hip_plan *pl = b->plans[i];
                const hipStream_t s = b->streams[stream_of[i]];
                pl->cur = 0;
                launch_init(pl, s);
                cap = cap && launch_iterations(pl, s) == ARTEFACT_OK;
            }
            for (uint32_t k = 1; k < ns; k++) {
                cap = cap && hipEventRecord(b->events[k], b->streams[k]) == hipSuccess;
                cap = cap && hipStreamWaitEvent(b->streams[0], b->events[k], 0) == hipSuccess;
            }
            const hipError_t e = hipStreamEndCapture(b->streams[0], &g);
            cap = cap && e == hipSuccess && g;
        }
        cap = cap && hipGraphInstantiate(exec, g, nullptr, nullptr, 0) == hipSuccess;
        if (g) (void)hipGraphDestroy(g);
        return cap;
    };
    if (ok && !nstreams && tune && count > 1) {
        // Each candidate: one replay to warm up, one timed with events on stream 0.
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
        float best = -1.0f;
        for (uint32_t ns = 1; ns <= most && ok; ns++) {
            hipGraphExec_t ge = nullptr;
            float ms = 0.0f;
            ok = build(ns, &ge) && hipGraphLaunch(ge, b->streams[0]) == hipSuccess &&
                 hipEventRecord(e0, b->streams[0]) == hipSuccess && hipGraphLaunch(ge, b->streams[0]) == hipSuccess &&
                 hipEventRecord(e1, b->streams[0]) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
                 hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
            if (ok && (best < 0.0f || ms < best)) {
                best = ms;
                if (b->graph) (void)hipGraphExecDestroy(b->graph);
                b->graph = ge;
                b->streams_used = ns;
            } else if (ge) {
                (void)hipGraphExecDestroy(ge);
            }
        }
        for (hipEvent_t e : {e0, e1})
            if (e) (void)hipEventDestroy(e);
    } else if (ok) {
        b->streams_used = nstreams ? most : std::min(2u, most);
        ok = build(b->streams_used, &b->graph);
    }
    if (!ok) {
        free_batch(b);
        return ARTEFACT_E_HIP;
    }
    *out = b;
    return ARTEFACT_OK;
}

uint32_t artefact_hip_batch_streams(const artefact_hip_batch *b)
{
    return b ? b->streams_used : 0u;
}

int artefact_hip_batch_run(artefact_hip_batch *b)
{
    if (!b) return ARTEFACT_E_ARGUMENT;
    DeviceGuard guard(b->device);
    if (hipGraphLaunch(b->graph, b->streams[0]) != hipSuccess) return ARTEFACT_E_HIP;
    for (artefact_hip_plan *pl : b->plans) {
        pl->cur = (int)(pl->iterations % 2);
        pl->initialized = false;
    }
    return ARTEFACT_OK;
}

int artefact_hip_batch_synchronize(artefact_hip_batch *b)
{
    if (!b) return ARTEFACT_E_ARGUMENT;
    DeviceGuard guard(b->device);
    return hipStreamSynchronize(b->streams[0]) == hipSuccess ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

int artefact_hip_batch_plan(artefact_hip_batch *b, uint32_t index, artefact_hip_plan **out)
{
    if (!b || !out || index >= b->plans.size()) return ARTEFACT_E_ARGUMENT;
    *out = b->plans[index];
    return ARTEFACT_OK;
}

int artefact_hip_batch_download(artefact_hip_batch *b, uint32_t index, float *const *out_planes)
{
    if (!b || index >= b->plans.size()) return ARTEFACT_E_ARGUMENT;
    const int rc = artefact_hip_batch_synchronize(b);
    return rc ? rc : artefact_hip_plan_download(b->plans[index], out_planes);
}

void artefact_hip_batch_destroy(artefact_hip_batch *b) { free_batch(b); }

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

// ---------------------------------------------------------------------------------------
// The loop without any allocation (artefact_hip_compute_noalloc).  Every buffer beyond the
// caller's coefficients and output planes -- x and x_prev, the extrapolated image and the
// gradient, the coefficient estimate and the prob plane, the quantisation tables, the norm
// accumulators and norms -- is carved from one static device array, which each device that runs
// the library holds from the first call until the process ends.  Host code never passes the
// array's name to a kernel (that is a host-side stand-in, not a device address): the device
// address comes from hipGetSymbolAddress once per device, under g_scratch_mu.
static __device__ __align__(256) unsigned char scratch[ARTEFACT_SCRATCH_BYTES];

namespace {

constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[SCRATCH_DEVICES] = {};

// Device address of this device's copy of `scratch` (the current device), or nullptr.
void *scratch_address(int device)
{
    if (device < 0 || device >= SCRATCH_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

// The quantisation tables as kernel arguments, written into the scratch tables by k_tables.
struct TableArgs {
    float quant[3][64];
    double rqq[3][64];
};
__global__ __launch_bounds__(64) void k_tables(const TableArgs t, float *quant, double *rqq)
{
    const int j = threadIdx.x;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        quant[c * 64 + j] = t.quant[c][j];
        rqq[c * 64 + j] = t.rqq[c][j];
    }
}

}  // namespace

extern "C" {

int artefact_hip_compute_noalloc(const artefact_params *p, const artefact_channel *ch, float *const *out,
                                 void *stream)
{
    int rc = check_inputs(p, ch);
    if (rc) return rc;
    if (!out || p->init_override) return ARTEFACT_E_ARGUMENT;
    for (uint32_t c = 0; c < p->nchannel; c++) {
        if (!out[c]) return ARTEFACT_E_ARGUMENT;
        // The projection kernels read a block row of coefficients as one 16-byte load.
        if (reinterpret_cast<uintptr_t>(ch[c].coefs) % 16) return ARTEFACT_E_ARGUMENT;
    }
    // The plan lives on the stack: its host vectors stay empty, so nothing is allocated.
    artefact_hip_plan pl;
    pl.stream = static_cast<hipStream_t>(stream);
    set_loop_scalars(&pl, p);
    size_t mtot = 0;
    for (uint32_t c = 0; c < pl.C; c++) {
        pl.M[c] = (size_t)ch[c].rounded_w * ch[c].rounded_h;
        mtot += pl.M[c];
    }
    // Scratch layout, each buffer at a 256-byte boundary, as the plan arena.
    const size_t cn = (size_t)pl.C * pl.N;
    size_t bytes = 0;
    auto take = [&bytes](size_t n) {
        const size_t at = bytes;
        bytes += (n + 255) & ~(size_t)255;
        return at;
    };
    const size_t at_x = take(cn * 4), at_xp = take(cn * 4), at_g = take(cn * 4), at_y = take(cn * 4);
    const size_t at_cos = take(mtot * 4), at_prob = take(mtot * 4);
    const size_t at_quant = take(3 * 64 * 4), at_rqq = take(3 * 64 * 8), at_acc = take(2 * ACC_WORDS * 8);
    const size_t at_norms = take(3 * 4), at_rnorm = take(3 * 8), at_nsq = take(3 * 8);
    if (bytes > ARTEFACT_SCRATCH_BYTES) return ARTEFACT_E_SCRATCH;
    if (hipGetDevice(&pl.device) != hipSuccess) return ARTEFACT_E_HIP;
    const char *wide_env = getenv("ARTEFACT_WIDE_ADDRESSING");
    pl.wide = wide_env && wide_env[0] == '1';
    pl.rows = pick_rows(pl.device, pl.W, pl.H, pl.C, pl.tgv, false, pl.wide);
    pl.grid_g = dim3((pl.W + KG_COLS - 1) / KG_COLS, (pl.H + WAVES * pl.rows - 1) / (WAVES * pl.rows));
    pl.grid_p = dim3((pl.W + TW - 1) / TW, (pl.H + TH - 1) / TH);
    char *base = static_cast<char *>(scratch_address(pl.device));
    if (!base) return ARTEFACT_E_HIP;
    pl.buf[0] = reinterpret_cast<float *>(base + at_x);
    pl.buf[1] = reinterpret_cast<float *>(base + at_xp);
    pl.g = reinterpret_cast<float *>(base + at_g);
    pl.ybuf = reinterpret_cast<float *>(base + at_y);
    pl.cos = reinterpret_cast<float *>(base + at_cos);
    pl.prob = reinterpret_cast<float *>(base + at_prob);
    pl.quant = reinterpret_cast<float *>(base + at_quant);
    pl.rqq = reinterpret_cast<double *>(base + at_rqq);
    pl.acc = reinterpret_cast<long long *>(base + at_acc);
    pl.norms = reinterpret_cast<float *>(base + at_norms);
    pl.rnorm = reinterpret_cast<double *>(base + at_rnorm);
    pl.norm_sq = reinterpret_cast<double *>(base + at_nsq);

    TableArgs tables;
    for (int c = 0; c < 3; c++)
        for (int j = 0; j < 64; j++) {
            tables.quant[c][j] = 1.0f;
            tables.rqq[c][j] = 1.0;
        }
    size_t moff = 0;
    for (uint32_t c = 0; c < pl.C; c++) {
        set_channel(&pl, c, ch[c], tables.quant[c], tables.rqq[c]);
        ChanDev &d = pl.ch[c];
        d.coef = ch[c].coefs;
        d.cos = pl.cos + moff;
        d.prob = pl.prob + moff;
        d.quant = pl.quant + c * 64;
        d.rqq = pl.rqq + c * 64;
        moff += pl.M[c];
    }
    set_layout(&pl);

    // Every buffer the loop reads is written first (tables, start image, cos, prob = 0) or
    // cleared by a kernel (the norm accumulators, in launch_init), so the previous call's
    // contents do not matter.
    hipLaunchKernelGGL(k_tables, dim3(1), dim3(64), 0, pl.stream, tables, pl.quant, pl.rqq);
    launch_init(&pl, pl.stream);
    float term = 1.0f;
    int cur = 0;
    for (uint32_t it = 0; it < pl.iterations; it++) {
        const IterArgs a = make_args(&pl, cur, fista_factor(term), it + 1 == pl.iterations);
        launch_gradient(&pl, a, pl.stream);
        launch_project(&pl, a, pl.stream);
        cur = 1 - cur;
    }
    bool ok = hipGetLastError() == hipSuccess;
    for (uint32_t c = 0; c < pl.C && ok; c++)
        ok = hipMemcpyAsync(out[c], pl.buf[cur] + c * pl.N, pl.N * sizeof(float), hipMemcpyDeviceToDevice,
                            pl.stream) == hipSuccess;
    return ok ? ARTEFACT_OK : ARTEFACT_E_HIP;
}

}  // extern "C"
