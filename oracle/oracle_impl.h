/*
 * CPU restatement of artefact-lib's scalar reconstruction loop, generic over the
 * element type.  TEST INFRASTRUCTURE ONLY: this file is compiled into the oracle
 * library that the tests, the smoke check and the CPU-baseline leg of the
 * benchmark use as the checker.  The product (artefact_amd) never links it.
 *
 * Included twice from oracle.c with
 *   R      the element type (float / double)
 *   FN(x)  the name mangler (x##_f32 / x##_f64)
 * All float arithmetic is written in the reference's operation order and the
 * file is compiled with -ffp-contract=off (the reference's default build has
 * mul_add!(a,b,c) == a*b+c with two roundings, backend/artefact-lib/utils/macros.rs:8-13).
 *
 * Citations are relative to the reference checkout.
 */

/* Ooura 8-point constants, backend/artefact-lib/utils/dct.rs:20-27 (values of
 * sqrt(2/8)*cos(k*pi/16), sqrt(2/8)*sin(k*pi/16) and cos(pi/4)). */
#define C8_1R ((R)0.49039264020161522456)
#define C8_1I ((R)0.09754516100806413392)
#define C8_2R ((R)0.46193976625564337806)
#define C8_2I ((R)0.19134171618254488586)
#define C8_3R ((R)0.41573480615127261854)
#define C8_3I ((R)0.27778511650980111237)
#define C8_4R ((R)0.35355339059327376220)
#define W8_4R ((R)0.70710678118654752440)

/* idct8x8s, backend/artefact-lib/utils/dct.rs:30-98: column pass then row pass.
 * Equals the orthonormal D^T X D (checked against the closed form in the tests). */
static void FN(idct8x8)(R *a)
{
    R x0r, x0i, x1r, x1i, x2r, x2i, x3r, x3i, xr, xi;
    for (int j = 0; j < 8; j++) {
        x1r = C8_1R * a[1 * 8 + j] + C8_1I * a[7 * 8 + j];
        x1i = C8_1R * a[7 * 8 + j] + -(C8_1I * a[1 * 8 + j]);
        x3r = C8_3R * a[3 * 8 + j] + C8_3I * a[5 * 8 + j];
        x3i = C8_3R * a[5 * 8 + j] + -(C8_3I * a[3 * 8 + j]);
        xr = x1r - x3r;
        xi = x1i + x3i;
        x1r += x3r;
        x3i -= x1i;
        x1i = W8_4R * (xr + xi);
        x3r = W8_4R * (xr - xi);
        xr = C8_2R * a[2 * 8 + j] + C8_2I * a[6 * 8 + j];
        xi = C8_2R * a[6 * 8 + j] + -(C8_2I * a[2 * 8 + j]);
        x0r = C8_4R * (a[0 * 8 + j] + a[4 * 8 + j]);
        x0i = C8_4R * (a[0 * 8 + j] - a[4 * 8 + j]);
        x2r = x0r - xr;
        x2i = x0i - xi;
        x0r += xr;
        x0i += xi;
        a[0 * 8 + j] = x0r + x1r;
        a[7 * 8 + j] = x0r - x1r;
        a[2 * 8 + j] = x0i + x1i;
        a[5 * 8 + j] = x0i - x1i;
        a[4 * 8 + j] = x2r - x3i;
        a[3 * 8 + j] = x2r + x3i;
        a[6 * 8 + j] = x2i - x3r;
        a[1 * 8 + j] = x2i + x3r;
    }
    for (int j = 0; j < 8; j++) {
        R *r = a + j * 8;
        x1r = C8_1R * r[1] + C8_1I * r[7];
        x1i = C8_1R * r[7] + -(C8_1I * r[1]);
        x3r = C8_3R * r[3] + C8_3I * r[5];
        x3i = C8_3R * r[5] + -(C8_3I * r[3]);
        xr = x1r - x3r;
        xi = x1i + x3i;
        x1r += x3r;
        x3i -= x1i;
        x1i = W8_4R * (xr + xi);
        x3r = W8_4R * (xr - xi);
        xr = C8_2R * r[2] + C8_2I * r[6];
        xi = C8_2R * r[6] + -(C8_2I * r[2]);
        x0r = C8_4R * (r[0] + r[4]);
        x0i = C8_4R * (r[0] - r[4]);
        x2r = x0r - xr;
        x2i = x0i - xi;
        x0r += xr;
        x0i += xi;
        r[0] = x0r + x1r;
        r[7] = x0r - x1r;
        r[2] = x0i + x1i;
        r[5] = x0i - x1i;
        r[4] = x2r - x3i;
        r[3] = x2r + x3i;
        r[6] = x2i - x3r;
        r[1] = x2i + x3r;
    }
}

/* dct8x8s, backend/artefact-lib/utils/dct.rs:100-169: column pass then row pass.
 * Equals the orthonormal D X D^T. */
static void FN(dct8x8)(R *a)
{
    R x0r, x0i, x1r, x1i, x2r, x2i, x3r, x3i, xr, xi;
    for (int j = 0; j < 8; j++) {
        x0r = a[0 * 8 + j] + a[7 * 8 + j];
        x1r = a[0 * 8 + j] - a[7 * 8 + j];
        x0i = a[2 * 8 + j] + a[5 * 8 + j];
        x1i = a[2 * 8 + j] - a[5 * 8 + j];
        x2r = a[4 * 8 + j] + a[3 * 8 + j];
        x3r = a[4 * 8 + j] - a[3 * 8 + j];
        x2i = a[6 * 8 + j] + a[1 * 8 + j];
        x3i = a[6 * 8 + j] - a[1 * 8 + j];
        xr = x0r + x2r;
        xi = x0i + x2i;
        a[0 * 8 + j] = C8_4R * (xr + xi);
        a[4 * 8 + j] = C8_4R * (xr - xi);
        xr = x0r - x2r;
        xi = x0i - x2i;
        a[2 * 8 + j] = C8_2R * xr + -(C8_2I * xi);
        a[6 * 8 + j] = C8_2R * xi + C8_2I * xr;
        xr = W8_4R * (x1i - x3i);
        x1i = W8_4R * (x1i + x3i);
        x3i = x1i - x3r;
        x1i += x3r;
        x3r = x1r - xr;
        x1r += xr;
        a[1 * 8 + j] = C8_1R * x1r + -(C8_1I * x1i);
        a[7 * 8 + j] = C8_1R * x1i + C8_1I * x1r;
        a[3 * 8 + j] = C8_3R * x3r + -(C8_3I * x3i);
        a[5 * 8 + j] = C8_3R * x3i + C8_3I * x3r;
    }
    for (int j = 0; j < 8; j++) {
        R *r = a + j * 8;
        x0r = r[0] + r[7];
        x1r = r[0] - r[7];
        x0i = r[2] + r[5];
        x1i = r[2] - r[5];
        x2r = r[4] + r[3];
        x3r = r[4] - r[3];
        x2i = r[6] + r[1];
        x3i = r[6] - r[1];
        xr = x0r + x2r;
        xi = x0i + x2i;
        r[0] = C8_4R * (xr + xi);
        r[4] = C8_4R * (xr - xi);
        xr = x0r - x2r;
        xi = x0i - x2i;
        r[2] = C8_2R * xr + -(C8_2I * xi);
        r[6] = C8_2R * xi + C8_2I * xr;
        xr = W8_4R * (x1i - x3i);
        x1i = W8_4R * (x1i + x3i);
        x3i = x1i - x3r;
        x1i += x3r;
        x3r = x1r - xr;
        x1r += xr;
        r[1] = C8_1R * x1r + -(C8_1I * x1i);
        r[7] = C8_1R * x1i + C8_1I * x1r;
        r[3] = C8_3R * x3r + -(C8_3I * x3i);
        r[5] = C8_3R * x3i + C8_3I * x3r;
    }
}

/* unboxing, backend/artefact-lib/utils/boxing.rs:8-45 (non-simd branch :34-42):
 * block order (64 contiguous values per block, raster block order) -> raster. */
static void FN(unboxing)(const R *in, R *out, uint32_t rw, uint32_t bw, uint32_t bh)
{
    size_t index = 0;
    for (uint32_t by = 0; by < bh; by++)
        for (uint32_t bx = 0; bx < bw; bx++)
            for (uint32_t iy = 0; iy < 8; iy++)
                for (uint32_t ix = 0; ix < 8; ix++)
                    out[(size_t)(by * 8 + iy) * rw + (bx * 8 + ix)] = in[index++];
}

/* boxing, backend/artefact-lib/utils/boxing.rs:48-84 (non-simd branch :74-81):
 * raster -> block order. */
static void FN(boxing)(const R *in, R *out, uint32_t rw, uint32_t bw, uint32_t bh)
{
    size_t index = 0;
    for (uint32_t by = 0; by < bh; by++)
        for (uint32_t bx = 0; bx < bw; bx++)
            for (uint32_t iy = 0; iy < 8; iy++)
                for (uint32_t ix = 0; ix < 8; ix++)
                    out[index++] = in[(size_t)(by * 8 + iy) * rw + (bx * 8 + ix)];
}

/* f32::midpoint / f64::midpoint.  Rust std on x86_64 computes the f32 form as
 * ((a as f64 + b as f64) / 2.0) as f32; the f64 form is (a + b) / 2 for finite
 * operands far from overflow, which is all this loop produces. */
static inline R FN(midpoint)(R a, R b)
{
    return (R)(((double)a + (double)b) / 2.0);
}

/* Per-channel solver state, backend/artefact-lib/utils/aux.rs:9-23. */
typedef struct {
    R *cos;          /* M_c, block order */
    R *grad;         /* N */
    R *diff_x;       /* N */
    R *diff_y;       /* N */
    R *fdata;        /* N */
    R *fista;        /* N */
    R *dct_coefs;    /* M_c, block order, the f32 coefficient values widened to R */
    R quant[64];
} FN(aux_t);

/* Image data of ScalarCoef::from.  mode 0 ("corrected"): dequantize, IDCT every
 * block, unbox once -- backend/artefact-lib/pipeline_simd_8/coef.rs:88-118.
 * mode 1 ("scalar_quirk"): the default scalar pipeline's order, which unboxes the
 * whole buffer after every block -- backend/artefact-lib/pipeline_scalar/coef.rs:24-50.
 * O(blocks * pixels): small inputs only. */
static void FN(image_data)(const oracle_channel *c, const R *coefs, const R *q,
                           int mode, R *image)
{
    size_t M = (size_t)c->rounded_w * c->rounded_h;
    size_t nb = (size_t)c->block_w * c->block_h;
    if (mode == 0) {
        R *tmp = (R *)malloc(M * sizeof(R));
        for (size_t i = 0; i < nb; i++) {
            for (int j = 0; j < 64; j++)
                tmp[i * 64 + j] = coefs[i * 64 + j] * q[j];
            FN(idct8x8)(tmp + i * 64);
        }
        FN(unboxing)(tmp, image, c->rounded_w, c->block_w, c->block_h);
        free(tmp);
    } else {
        R *clone = (R *)malloc(M * sizeof(R));
        for (size_t i = 0; i < M; i++) image[i] = 0;
        for (size_t i = 0; i < nb; i++) {
            for (int j = 0; j < 64; j++)
                image[i * 64 + j] = coefs[i * 64 + j] * q[j];
            FN(idct8x8)(image + i * 64);
            memcpy(clone, image, M * sizeof(R));
            FN(unboxing)(clone, image, c->rounded_w, c->block_w, c->block_h);
        }
        free(clone);
    }
}

/* get_fdata, backend/artefact-lib/pipeline_scalar/coef.rs:69-92: nearest-neighbour
 * upsample of the component plane onto the max grid, clamped at the last row/column. */
static void FN(upsample)(const oracle_channel *c, const R *image, uint32_t W, uint32_t H, R *fdata)
{
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            uint32_t cy = y / c->v_samp; if (cy > c->rounded_h - 1) cy = c->rounded_h - 1;
            uint32_t cx = x / c->h_samp; if (cx > c->rounded_w - 1) cx = c->rounded_w - 1;
            fdata[(size_t)y * W + x] = image[(size_t)cy * c->rounded_w + cx];
        }
}

/* compute_step_prob, backend/artefact-lib/pipeline_scalar/compute_step_prob.rs:5-60. */
static void FN(step_prob)(uint32_t W, R alpha, const oracle_channel *c, const FN(aux_t) *aux, R *grad)
{
    R cosb[64];
    for (uint32_t by = 0; by < c->block_h; by++)
        for (uint32_t bx = 0; bx < c->block_w; bx++) {
            size_t i = (size_t)by * c->block_w + bx;
            for (int j = 0; j < 64; j++) {
                cosb[j] = aux->cos[i * 64 + j];
                cosb[j] -= aux->dct_coefs[i * 64 + j] * aux->quant[j];
                cosb[j] /= aux->quant[j] * aux->quant[j];
            }
            FN(idct8x8)(cosb);
            for (uint32_t iy = 0; iy < 8; iy++)
                for (uint32_t ix = 0; ix < 8; ix++) {
                    uint32_t cx = bx * 8 + ix, cy = by * 8 + iy;
                    for (uint32_t sy = 0; sy < c->v_samp; sy++)
                        for (uint32_t sx = 0; sx < c->h_samp; sx++) {
                            uint32_t y = cy * c->v_samp + sy, x = cx * c->h_samp + sx;
                            grad[(size_t)y * W + x] += alpha * cosb[iy * 8 + ix];
                        }
                }
        }
}

/* compute_step_tv + compute_step_tv_inner,
 * backend/artefact-lib/pipeline_scalar/compute_step_tv.rs:4-93.  Scatter form, raster
 * order.  Returns the (discarded-by-the-reference) f64 TV sum. */
static double FN(step_tv)(uint32_t W, uint32_t H, uint32_t C, FN(aux_t) *aux)
{
    double tv = 0.0;
    R alpha = (R)1.0 / (R)sqrt((double)(R)C);
    if (sizeof(R) == sizeof(float)) alpha = (R)(1.0f / sqrtf((float)C));
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            R gx[3] = {0, 0, 0}, gy[3] = {0, 0, 0};
            size_t p = (size_t)y * W + x;
            int right = x >= W - 1, bottom = y >= H - 1;
            for (uint32_t c = 0; c < C; c++) {
                if (!right) gx[c] = aux[c].fdata[p + 1] - aux[c].fdata[p];
                if (!bottom) gy[c] = aux[c].fdata[p + W] - aux[c].fdata[p];
            }
            R n = 0;
            for (uint32_t c = 0; c < C; c++) {
                n += gx[c] * gx[c];
                n += gy[c] * gy[c];
            }
            if (sizeof(R) == sizeof(float)) n = (R)sqrtf((float)n);
            else n = (R)sqrt((double)n);
            tv += (double)alpha * (double)n;
            if (n != 0) {
                for (uint32_t c = 0; c < C; c++) {
                    aux[c].grad[p] += alpha * -(gx[c] + gy[c]) / n;
                    if (!right) aux[c].grad[p + 1] += alpha * gx[c] / n;
                    if (!bottom) aux[c].grad[p + W] += alpha * gy[c] / n;
                }
            }
            for (uint32_t c = 0; c < C; c++) {
                aux[c].diff_x[p] = gx[c];
                aux[c].diff_y[p] = gy[c];
            }
        }
    return tv;
}

/* compute_step_tv2 + compute_step_tv2_inner,
 * backend/artefact-lib/pipeline_scalar/compute_step_tv2.rs:4-136.  Scatter form,
 * raster order, 7 target sites with the reference's own border rules. */
static double FN(step_tgv)(uint32_t W, uint32_t H, uint32_t C, FN(aux_t) *aux, R alpha_in)
{
    double tv2 = 0.0;
    R alpha = alpha_in;
    if (sizeof(R) == sizeof(float)) alpha = (R)((float)alpha / sqrtf((float)C));
    else alpha /= (R)sqrt((double)C);
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            R gxx[3] = {0, 0, 0}, sym[3] = {0, 0, 0}, gyy[3] = {0, 0, 0};
            size_t a = (size_t)y * W + x;
            for (uint32_t c = 0; c < C; c++) {
                R gxy = 0, gyx = 0;
                if (x != 0) {
                    gxx[c] = aux[c].diff_x[a] - aux[c].diff_x[a - 1];
                    gyx = aux[c].diff_y[a] - aux[c].diff_y[a - 1];
                }
                if (y != 0) {
                    gyy[c] = aux[c].diff_y[a] - aux[c].diff_y[a - W];
                    gxy = aux[c].diff_x[a] - aux[c].diff_x[a - W];
                }
                sym[c] = FN(midpoint)(gxy, gyx);
            }
            R n2 = 0;
            for (uint32_t c = 0; c < C; c++)
                n2 += gyy[c] * gyy[c] + ((R)2 * (sym[c] * sym[c]) + gxx[c] * gxx[c]);
            if (sizeof(R) == sizeof(float)) n2 = (R)sqrtf((float)n2);
            else n2 = (R)sqrt((double)n2);
            tv2 += (double)(alpha * n2);
            if (n2 == 0) continue;
            for (uint32_t c = 0; c < C; c++) {
                R xx = gxx[c], yy = gyy[c], s = sym[c];
                R *g = aux[c].grad;
                g[a] += alpha * (-((R)2 * yy + ((R)2 * xx + (R)2 * s)) / n2);
                if (x > 0) g[a - 1] += alpha * ((s + xx) / n2);
                if (x < W - 1) g[a + 1] += alpha * ((s + xx) / n2);
                if (y > 0) g[a - W] += alpha * ((yy + s) / n2);
                if (y < H - 1) g[a + W] += alpha * ((yy + s) / n2);
                if (x < W - 1 && y > 0) g[a - W + 1] += alpha * ((-s) / n2);
                if (x > 0 && y < H - 1) g[a + W - 1] += alpha * ((-s) / n2);
            }
        }
    return tv2;
}

/* Rust's f32::clamp: if x < lo then lo; if x > hi then hi; NaN passes through. */
static inline R FN(clamp)(R v, R lo, R hi)
{
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}

/* compute_projection, backend/artefact-lib/pipeline_scalar/compute_projection.rs:8-122. */
static void FN(projection)(uint32_t W, uint32_t H, const oracle_channel *c, FN(aux_t) *aux)
{
    int resample = c->rounded_w != W || c->rounded_h != H;
    uint32_t rw = c->rounded_w, rh = c->rounded_h, hs = c->h_samp, vs = c->v_samp;
    size_t nb = (size_t)c->block_w * c->block_h;
    (void)H;
    if (resample) {
        for (uint32_t cy = 0; cy < rh; cy++)
            for (uint32_t cx = 0; cx < rw; cx++) {
                R mean = 0;
                for (uint32_t sy = 0; sy < vs; sy++)
                    for (uint32_t sx = 0; sx < hs; sx++)
                        mean += aux->fdata[(size_t)(cy * vs + sy) * W + (cx * hs + sx)];
                mean /= (R)(hs * vs);
                aux->diff_y[(size_t)cy * rw + cx] = mean;
                for (uint32_t sy = 0; sy < vs; sy++)
                    for (uint32_t sx = 0; sx < hs; sx++)
                        aux->fdata[(size_t)(cy * vs + sy) * W + (cx * hs + sx)] -= mean;
            }
    }
    FN(boxing)(resample ? aux->diff_y : aux->fdata, aux->diff_x, rw, c->block_w, c->block_h);
    for (size_t i = 0; i < nb; i++) FN(dct8x8)(aux->diff_x + i * 64);
    for (size_t i = 0; i < nb; i++)
        for (int j = 0; j < 64; j++) {
            R lo = (aux->dct_coefs[i * 64 + j] - (R)0.5) * aux->quant[j];
            R hi = (aux->dct_coefs[i * 64 + j] + (R)0.5) * aux->quant[j];
            aux->diff_x[i * 64 + j] = FN(clamp)(aux->diff_x[i * 64 + j], lo, hi);
        }
    memcpy(aux->cos, aux->diff_x, nb * 64 * sizeof(R));
    for (size_t i = 0; i < nb; i++) FN(idct8x8)(aux->diff_x + i * 64);
    FN(unboxing)(aux->diff_x, resample ? aux->diff_y : aux->fdata, rw, c->block_w, c->block_h);
    if (resample) {
        for (uint32_t cy = 0; cy < rh; cy++)
            for (uint32_t cx = 0; cx < rw; cx++) {
                R mean = aux->diff_y[(size_t)cy * rw + cx];
                for (uint32_t sy = 0; sy < vs; sy++)
                    for (uint32_t sx = 0; sx < hs; sx++)
                        aux->fdata[(size_t)(cy * vs + sy) * W + (cx * hs + sx)] += mean;
            }
    }
}

/* Gradient norm of compute_step, backend/artefact-lib/pipeline_scalar/compute_step.rs:54-58.
 * norm_mode 0: the reference's sequential fold acc = x*x + acc in R, then sqrt.
 * norm_mode 1: squares widened to f64 and summed with compensation (Neumaier),
 *              sqrt in f64, rounded once to R: the order-independent form the GPU
 *              path computes, so both agree bit-for-bit. */
static R FN(grad_norm)(const R *g, size_t n, int norm_mode)
{
    if (norm_mode == 0) {
        R acc = 0;
        for (size_t i = 0; i < n; i++) acc = g[i] * g[i] + acc;
        if (sizeof(R) == sizeof(float)) return (R)sqrtf((float)acc);
        return (R)sqrt((double)acc);
    }
    double s = 0.0, comp = 0.0;
    for (size_t i = 0; i < n; i++) {
        double v = (double)g[i] * (double)g[i];
        double t = s + v;
        if (fabs(s) >= fabs(v)) comp += (s - t) + v;
        else comp += (v - t) + s;
        s = t;
    }
    return (R)sqrt(s + comp);
}

/* compute_step, backend/artefact-lib/pipeline_scalar/compute_step.rs:8-72. */
static void FN(step)(uint32_t W, uint32_t H, uint32_t C, const oracle_channel *ch, FN(aux_t) *aux,
                     R step_size, R weight, const float *pweight, int norm_mode,
                     double *tv_out, double *tv2_out, R *norms_out)
{
    size_t N = (size_t)W * H;
    for (uint32_t c = 0; c < C; c++) {
        for (size_t i = 0; i < N; i++) aux[c].grad[i] = 0;
        if (pweight[c] != 0.0f) {
            R alpha;
            if (sizeof(R) == sizeof(float)) alpha = (R)(pweight[c] * 2.0f * 255.0f * sqrtf(2.0f));
            else alpha = (R)pweight[c] * (R)2 * (R)255 * (R)sqrt(2.0);
            FN(step_prob)(W, alpha, &ch[c], &aux[c], aux[c].grad);
        }
    }
    double tv = FN(step_tv)(W, H, C, aux);
    R a2;
    if (sizeof(R) == sizeof(float)) a2 = (R)((float)weight / sqrtf(2.0f));
    else a2 = weight / (R)sqrt(2.0);
    double tv2 = FN(step_tgv)(W, H, C, aux, a2);
    if (tv_out) *tv_out = tv;
    if (tv2_out) *tv2_out = tv2;
    for (uint32_t c = 0; c < C; c++) {
        R norm = FN(grad_norm)(aux[c].grad, N, norm_mode);
        if (norms_out) norms_out[c] = norm;
        if (norm != 0)
            for (size_t i = 0; i < N; i++) aux[c].fdata[i] -= step_size * (aux[c].grad[i] / norm);
    }
    for (uint32_t c = 0; c < C; c++) FN(projection)(W, H, &ch[c], &aux[c]);
}

static int FN(aux_alloc)(FN(aux_t) *a, const oracle_channel *c, size_t N)
{
    size_t M = (size_t)c->rounded_w * c->rounded_h;
    a->cos = (R *)calloc(M ? M : 1, sizeof(R));
    a->dct_coefs = (R *)calloc(M ? M : 1, sizeof(R));
    a->grad = (R *)calloc(N, sizeof(R));
    a->diff_x = (R *)calloc(N, sizeof(R));
    a->diff_y = (R *)calloc(N, sizeof(R));
    a->fdata = (R *)calloc(N, sizeof(R));
    a->fista = (R *)calloc(N, sizeof(R));
    if (!a->cos || !a->dct_coefs || !a->grad || !a->diff_x || !a->diff_y || !a->fdata || !a->fista)
        return -1;
    for (size_t i = 0; i < M; i++) a->dct_coefs[i] = (R)c->coefs[i];
    for (int j = 0; j < 64; j++) a->quant[j] = (R)c->quant[j];
    return 0;
}

static void FN(aux_free)(FN(aux_t) *a)
{
    free(a->cos); free(a->dct_coefs); free(a->grad); free(a->diff_x); free(a->diff_y);
    free(a->fdata); free(a->fista);
}

/* Aux::init, backend/artefact-lib/utils/aux.rs:45-65, with get_fdata/get_cos
 * (pipeline_scalar/coef.rs:69-104).  init_override, when given, replaces fdata. */
static void FN(aux_init)(FN(aux_t) *a, const oracle_channel *c, uint32_t W, uint32_t H,
                         int init_mode, const R *init_override)
{
    size_t N = (size_t)W * H, M = (size_t)c->rounded_w * c->rounded_h;
    if (init_override) {
        memcpy(a->fdata, init_override, N * sizeof(R));
    } else {
        R *image = (R *)calloc(M, sizeof(R));
        FN(image_data)(c, a->dct_coefs, a->quant, init_mode, image);
        FN(upsample)(c, image, W, H, a->fdata);
        free(image);
    }
    memcpy(a->fista, a->fdata, N * sizeof(R));
    for (size_t i = 0; i < M; i++) a->cos[i] = a->dct_coefs[i] * a->quant[i % 64];
}

/* compute, backend/artefact-lib/pipeline_scalar/mod.rs:17-78. */
static int FN(compute)(uint32_t C, const oracle_channel *ch, float weight, const float *pweight,
                       uint32_t iterations, uint32_t W, uint32_t H, int init_mode, int norm_mode,
                       const R *const *init_override, R *const *out, double *loop_seconds,
                       R *norm_trace)
{
    size_t N = (size_t)W * H;
    FN(aux_t) aux[3];
    memset(aux, 0, sizeof(aux));
    int rc = 0;
    for (uint32_t c = 0; c < C; c++)
        if (FN(aux_alloc)(&aux[c], &ch[c], N)) rc = -1;
    if (rc == 0) {
        for (uint32_t c = 0; c < C; c++)
            FN(aux_init)(&aux[c], &ch[c], W, H, init_mode, init_override ? init_override[c] : NULL);
        R radius, term = 1;
        if (sizeof(R) == sizeof(float)) radius = (R)(sqrtf((float)N) / 2.0f);
        else radius = (R)(sqrt((double)N) / 2.0);
        double t0 = now_seconds();
        for (uint32_t it = 0; it < iterations; it++) {
            R next_term, factor, step_size;
            if (sizeof(R) == sizeof(float)) {
                float t = (float)term;
                float nt = (float)(((double)1.0f + (double)sqrtf(4.0f * (t * t) + 1.0f)) / 2.0);
                next_term = (R)nt;
                factor = (R)((t - 1.0f) / nt);
                step_size = (R)((float)radius / sqrtf(1.0f + (float)iterations));
            } else {
                next_term = FN(midpoint)((R)1, (R)sqrt((double)((R)4 * (term * term) + (R)1)));
                factor = (term - (R)1) / next_term;
                step_size = radius / (R)sqrt(1.0 + (double)iterations);
            }
            /* One thread per channel when asked for, as the reference's par_iter_mut (mod.rs:54-59). */
#pragma omp parallel for num_threads(g_channel_threads) if (g_channel_threads > 1) schedule(static, 1)
            for (uint32_t c = 0; c < C; c++) {
                for (size_t i = 0; i < N; i++)
                    aux[c].fista[i] = factor * (aux[c].fdata[i] - aux[c].fista[i]) + aux[c].fdata[i];
                R *t = aux[c].fdata; aux[c].fdata = aux[c].fista; aux[c].fista = t;
            }
            term = next_term;
            FN(step)(W, H, C, ch, aux, step_size, (R)weight, pweight, norm_mode, NULL, NULL,
                     norm_trace ? norm_trace + (size_t)it * C : NULL);
        }
        if (loop_seconds) *loop_seconds = now_seconds() - t0;
        for (uint32_t c = 0; c < C; c++) memcpy(out[c], aux[c].fdata, N * sizeof(R));
    }
    for (uint32_t c = 0; c < C; c++) FN(aux_free)(&aux[c]);
    return rc;
}

#undef C8_1R
#undef C8_1I
#undef C8_2R
#undef C8_2I
#undef C8_3R
#undef C8_3I
#undef C8_4R
#undef W8_4R
