/*
 * CPU oracle for artefact-lib's reconstruction loop -- TEST INFRASTRUCTURE ONLY.
 *
 * A restatement of backend/artefact-lib/pipeline_scalar (the reference's default CPU
 * path, selected at backend/artefact-lib/lib.rs:27-28), used by tests/, the smoke
 * check and bench.py's CPU-baseline leg as the checker.  Never linked into the
 * product.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math).
 *
 * Parity status of each function (how it is pinned; see tests/test_oracle.py):
 *   dct8x8 / idct8x8        closed-form orthonormal DCT-II (D X D^T, D^T X D) + round trip
 *   boxing / unboxing       brute-force permutation + inverse pair
 *   image_data (corrected)  closed form: dequantize -> D^T X D -> raster
 *   image_data (quirk)      brute-force simulation of the reference's in-loop unboxing
 *   upsample                closed-form nearest-neighbour replicate
 *   fista scalars           closed-form t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2 sequence
 *   step_prob               zero on the first iteration; equals grad of the prob term of the
 *                           objective (finite differences, luma)
 *   step_tv                 finite differences of sum(alpha*n) away from n ~ 0
 *   step_tgv                interior: finite differences of sum(a2*n2); borders: brute
 *                           force against an independent numpy scatter restatement
 *   projection              idempotence, re-quantisation |DCT(down x)/q - c| <= 1/2,
 *                           mean-preserving chroma decomposition
 *   grad_norm               norm_mode 1 vs exact rational sum (fractions) on small inputs
 *   compute                 special cases: iterations=0 returns the init, weight=0 is the
 *                           TV-only loop, uniform-DC image never moves
 *   objective               closed-form values on constant / ramp planes
 *   output_rgb              inverse of the BT.601 full-range forward transform (round trip
 *                           within the truncation), clamp / NaN / truncation special cases
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* Threads for the loops the reference runs over channels with rayon (pipeline_scalar/mod.rs:
 * 27-30 coefficient conversion, 54-59 FISTA update, 77 collection); everything inside
 * compute_step stays serial, as there.  1 unless oracle_set_threads says otherwise. */
static int g_channel_threads = 1;

void oracle_set_threads(int n) { g_channel_threads = n < 1 ? 1 : n; }

static double now_seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

#define R float
#define FN(x) x##_f32
#include "oracle_impl.h"
#undef R
#undef FN

#define R double
#define FN(x) x##_f64
#include "oracle_impl.h"
#undef R
#undef FN

static int valid_channel(const oracle_channel *c, uint32_t W, uint32_t H)
{
    if (c->rounded_w % 8 || c->rounded_h % 8 || !c->rounded_w || !c->rounded_h) return 0;
    if (c->block_w * 8 != c->rounded_w || c->block_h * 8 != c->rounded_h) return 0;
    if ((c->h_samp != 1 && c->h_samp != 2) || (c->v_samp != 1 && c->v_samp != 2)) return 0;
    if (c->rounded_w * c->h_samp > W || c->rounded_h * c->v_samp > H) return 0;
    return 1;
}

int oracle_compute_f32(uint32_t C, const oracle_channel *ch, float weight, const float *pweight,
                       uint32_t iterations, uint32_t W, uint32_t H, int init_mode, int norm_mode,
                       const float *const *init_override, float *const *out, double *loop_seconds,
                       float *norm_trace)
{
    if (C < 1 || C > 3) return -1;
    for (uint32_t c = 0; c < C; c++)
        if (!valid_channel(&ch[c], W, H)) return -2;
    return compute_f32(C, ch, weight, pweight, iterations, W, H, init_mode, norm_mode, init_override,
                       out, loop_seconds, norm_trace);
}

int oracle_compute_f64(uint32_t C, const oracle_channel *ch, float weight, const float *pweight,
                       uint32_t iterations, uint32_t W, uint32_t H, int init_mode,
                       const double *const *init_override, double *const *out, double *loop_seconds)
{
    if (C < 1 || C > 3) return -1;
    for (uint32_t c = 0; c < C; c++)
        if (!valid_channel(&ch[c], W, H)) return -2;
    return compute_f64(C, ch, weight, pweight, iterations, W, H, init_mode, 1, init_override, out,
                       loop_seconds, NULL);
}

int oracle_init_f32(const oracle_channel *ch, uint32_t W, uint32_t H, int init_mode, float *fdata)
{
    if (!valid_channel(ch, W, H)) return -2;
    aux_t_f32 a;
    memset(&a, 0, sizeof(a));
    if (aux_alloc_f32(&a, ch, (size_t)W * H)) return -1;
    aux_init_f32(&a, ch, W, H, init_mode, NULL);
    memcpy(fdata, a.fdata, (size_t)W * H * sizeof(float));
    aux_free_f32(&a);
    return 0;
}

int oracle_gradient_f32(uint32_t C, const oracle_channel *ch, float weight, const float *pweight,
                        uint32_t W, uint32_t H, const float *const *fdata, const float *const *cos,
                        float *const *grad, double *tv, double *tv2)
{
    if (C < 1 || C > 3) return -1;
    size_t N = (size_t)W * H;
    aux_t_f32 aux[3];
    memset(aux, 0, sizeof(aux));
    int rc = 0;
    for (uint32_t c = 0; c < C; c++) {
        if (!valid_channel(&ch[c], W, H)) { rc = -2; break; }
        if (aux_alloc_f32(&aux[c], &ch[c], N)) { rc = -1; break; }
        memcpy(aux[c].fdata, fdata[c], N * sizeof(float));
        memcpy(aux[c].cos, cos[c], (size_t)ch[c].rounded_w * ch[c].rounded_h * sizeof(float));
    }
    if (rc == 0) {
        for (uint32_t c = 0; c < C; c++) {
            if (pweight[c] != 0.0f)
                step_prob_f32(W, pweight[c] * 2.0f * 255.0f * sqrtf(2.0f), &ch[c], &aux[c], aux[c].grad);
        }
        double a = step_tv_f32(W, H, C, aux);
        double b = step_tgv_f32(W, H, C, aux, weight / sqrtf(2.0f));
        if (tv) *tv = a;
        if (tv2) *tv2 = b;
        for (uint32_t c = 0; c < C; c++) memcpy(grad[c], aux[c].grad, N * sizeof(float));
    }
    for (uint32_t c = 0; c < C; c++) aux_free_f32(&aux[c]);
    return rc;
}

int oracle_projection_f32(const oracle_channel *ch, uint32_t W, uint32_t H, float *fdata, float *cos)
{
    if (!valid_channel(ch, W, H)) return -2;
    size_t N = (size_t)W * H;
    aux_t_f32 a;
    memset(&a, 0, sizeof(a));
    if (aux_alloc_f32(&a, ch, N)) return -1;
    memcpy(a.fdata, fdata, N * sizeof(float));
    projection_f32(W, H, ch, &a);
    memcpy(fdata, a.fdata, N * sizeof(float));
    memcpy(cos, a.cos, (size_t)ch->rounded_w * ch->rounded_h * sizeof(float));
    aux_free_f32(&a);
    return 0;
}

void oracle_dct8x8_f32(float *b, uint64_t n) { for (uint64_t i = 0; i < n; i++) dct8x8_f32(b + 64 * i); }
void oracle_idct8x8_f32(float *b, uint64_t n) { for (uint64_t i = 0; i < n; i++) idct8x8_f32(b + 64 * i); }
void oracle_dct8x8_f64(double *b, uint64_t n) { for (uint64_t i = 0; i < n; i++) dct8x8_f64(b + 64 * i); }
void oracle_idct8x8_f64(double *b, uint64_t n) { for (uint64_t i = 0; i < n; i++) idct8x8_f64(b + 64 * i); }
void oracle_boxing_f32(const float *in, float *out, uint32_t rw, uint32_t bw, uint32_t bh)
{
    boxing_f32(in, out, rw, bw, bh);
}
void oracle_unboxing_f32(const float *in, float *out, uint32_t rw, uint32_t bw, uint32_t bh)
{
    unboxing_f32(in, out, rw, bw, bh);
}
float oracle_midpoint_f32(float a, float b) { return midpoint_f32(a, b); }

/* FISTA scalars of backend/artefact-lib/pipeline_scalar/mod.rs:45-71. */
void oracle_fista_scalars(uint32_t iterations, uint32_t N, float *next_terms, float *factors, float *step)
{
    float radius = sqrtf((float)N) / 2.0f, term = 1.0f;
    *step = radius / sqrtf(1.0f + (float)iterations);
    for (uint32_t i = 0; i < iterations; i++) {
        float nt = midpoint_f32(1.0f, sqrtf(4.0f * (term * term) + 1.0f));
        factors[i] = (term - 1.0f) / nt;
        next_terms[i] = nt;
        term = nt;
    }
}

/* Objective of the loop (build-defined; the reference computes only the discarded
 * partial sums of compute_step_tv.rs:70-71 and compute_step_tv2.rs:88-90):
 *   F(x) = sum_p alpha*n(p)                      [T1 form, compute_step_tv.rs:48-71]
 *        + sum_p a2*n2(p)                        [T2 form, compute_step_tv2.rs:53-90]
 *        + sum_c (alpha_p,c / 2) * sum_blocks sum_j ((DCT(down_c x_c)_j - c_j q_j) / q_j)^2
 * with alpha, a2, alpha_p the f32 constants the loop uses, every pixel term in f64.
 * terms[0..2] = TV, TGV, prob parts. */
static double objective_impl(uint32_t C, const oracle_channel *ch, float weight, const float *pweight,
                             uint32_t W, uint32_t H, const double *const *x, double *terms)
{
    double alpha = (double)(1.0f / sqrtf((float)C));
    double a2 = (double)((weight / sqrtf(2.0f)) / sqrtf((float)C));
    double tv = 0.0, tgv = 0.0, prob = 0.0;
    size_t N = (size_t)W * H;
    double *dx = (double *)malloc(3 * N * sizeof(double));
    double *dy = (double *)malloc(3 * N * sizeof(double));
    for (uint32_t c = 0; c < C; c++)
        for (uint32_t y = 0; y < H; y++)
            for (uint32_t xx = 0; xx < W; xx++) {
                size_t p = (size_t)y * W + xx;
                dx[c * N + p] = xx < W - 1 ? x[c][p + 1] - x[c][p] : 0.0;
                dy[c * N + p] = y < H - 1 ? x[c][p + W] - x[c][p] : 0.0;
            }
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t xx = 0; xx < W; xx++) {
            size_t p = (size_t)y * W + xx;
            double n = 0.0, n2 = 0.0;
            for (uint32_t c = 0; c < C; c++) {
                double gx = dx[c * N + p], gy = dy[c * N + p];
                n += gx * gx + gy * gy;
                double gxx = 0, gyx = 0, gyy = 0, gxy = 0;
                if (xx > 0) { gxx = gx - dx[c * N + p - 1]; gyx = gy - dy[c * N + p - 1]; }
                if (y > 0) { gyy = gy - dy[c * N + p - W]; gxy = gx - dx[c * N + p - W]; }
                double s = 0.5 * (gxy + gyx);
                n2 += gyy * gyy + 2.0 * s * s + gxx * gxx;
            }
            tv += alpha * sqrt(n);
            tgv += a2 * sqrt(n2);
        }
    free(dx);
    free(dy);
    for (uint32_t c = 0; c < C; c++) {
        if (pweight[c] == 0.0f) continue;
        const oracle_channel *k = &ch[c];
        double ap = (double)(pweight[c] * 2.0f * 255.0f * sqrtf(2.0f));
        int resample = k->rounded_w != W || k->rounded_h != H;
        double blk[64], s = 0.0;
        for (uint32_t by = 0; by < k->block_h; by++)
            for (uint32_t bx = 0; bx < k->block_w; bx++) {
                for (uint32_t iy = 0; iy < 8; iy++)
                    for (uint32_t ix = 0; ix < 8; ix++) {
                        uint32_t cy = by * 8 + iy, cx = bx * 8 + ix;
                        double v;
                        if (resample) {
                            v = 0.0;
                            for (uint32_t sy = 0; sy < k->v_samp; sy++)
                                for (uint32_t sx = 0; sx < k->h_samp; sx++)
                                    v += x[c][(size_t)(cy * k->v_samp + sy) * W + cx * k->h_samp + sx];
                            v /= (double)(k->h_samp * k->v_samp);
                        } else {
                            v = x[c][(size_t)cy * W + cx];
                        }
                        blk[iy * 8 + ix] = v;
                    }
                dct8x8_f64(blk);
                size_t i = (size_t)by * k->block_w + bx;
                for (int j = 0; j < 64; j++) {
                    double q = (double)k->quant[j];
                    double d = (blk[j] - (double)k->coefs[i * 64 + j] * q) / q;
                    s += d * d;
                }
            }
        prob += 0.5 * ap * s;
    }
    if (terms) { terms[0] = tv; terms[1] = tgv; terms[2] = prob; }
    return tv + tgv + prob;
}

double oracle_objective_f64planes(uint32_t C, const oracle_channel *ch, float weight, const float *pweight,
                                  uint32_t W, uint32_t H, const double *const *planes, double *terms)
{
    return objective_impl(C, ch, weight, pweight, W, H, planes, terms);
}

double oracle_objective(uint32_t C, const oracle_channel *ch, float weight, const float *pweight,
                        uint32_t W, uint32_t H, const float *const *planes, double *terms)
{
    size_t N = (size_t)W * H;
    double *x[3] = {NULL, NULL, NULL};
    for (uint32_t c = 0; c < C; c++) {
        x[c] = (double *)malloc(N * sizeof(double));
        for (size_t i = 0; i < N; i++) x[c][i] = (double)planes[c][i];
    }
    double r = objective_impl(C, ch, weight, pweight, W, H, (const double *const *)x, terms);
    for (uint32_t c = 0; c < C; c++) free(x[c]);
    return r;
}

/* Output stage, backend/artefact-lib/lib.rs:159-211: luma gets +128 in f32 (:160-162); for three
 * channels each pixel becomes R = 1.402*cr + y, G = 0.71414*(-cr) + (0.34414*(-cb) + y),
 * B = 1.772*cb + y, each product and sum rounded to f32 on its own (utils/macros.rs:10-15, the
 * default without the `native` feature), clamped to [0, 255] (f32::clamp keeps NaN) and cast with
 * Rust's saturating `as u8` (truncation, NaN -> 0).  Other channel counts give grey R = G = B from
 * channel 0 (:188-210).  Rows i < real_h, columns j < real_w of the max grid (stride max_w);
 * rgb is real_w*real_h*3 bytes, row-major, interleaved. */
static uint8_t rust_u8(float v)
{
    if (v < 0.0f) v = 0.0f;
    if (v > 255.0f) v = 255.0f;
    if (v != v) return 0;
    return (uint8_t)v;
}

void oracle_output_rgb(uint32_t nchannel, const float *const *planes, uint32_t max_w, uint32_t real_w,
                       uint32_t real_h, uint8_t *rgb)
{
    for (uint32_t i = 0; i < real_h; i++)
        for (uint32_t j = 0; j < real_w; j++) {
            const size_t idx = (size_t)i * max_w + j;
            uint8_t *o = rgb + ((size_t)i * real_w + j) * 3;
            const float y = planes[0][idx] + 128.0f;
            if (nchannel == 3) {
                const float cb = planes[1][idx], cr = planes[2][idx];
                const float r = 1.402f * cr + y;
                const float gi = 0.34414f * -cb + y;
                const float g = 0.71414f * -cr + gi;
                const float b = 1.772f * cb + y;
                o[0] = rust_u8(r);
                o[1] = rust_u8(g);
                o[2] = rust_u8(b);
            } else {
                o[0] = o[1] = o[2] = rust_u8(y);
            }
        }
}
