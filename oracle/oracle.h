/*
 * CPU oracle for artefact-lib's reconstruction loop -- TEST INFRASTRUCTURE ONLY.
 *
 * Used by tests/, the smoke check and the CPU-baseline leg of bench.py as the
 * checker.  The shipped library (artefact_amd/libartefact_hip.so) does not link,
 * load or call anything declared here.
 */
#ifndef ARTEFACT_ORACLE_H
#define ARTEFACT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One colour component, as backend/artefact-lib/jpeg/mod.rs:8-27 (Coefficient). */
typedef struct {
    uint32_t rounded_w, rounded_h;   /* component plane, multiples of 8 */
    uint32_t block_w, block_h;       /* rounded_w / 8, rounded_h / 8 */
    uint32_t h_samp, v_samp;         /* upsampling factor onto the max grid: 1 or 2 */
    const float *coefs;              /* block_w*block_h*64, raster block order, natural in-block order */
    const float *quant;              /* 64, natural order */
} oracle_channel;

/* init_mode: 0 corrected (pipeline_simd_8/coef.rs:88-118), 1 scalar quirk (pipeline_scalar/coef.rs:24-50)
 * norm_mode: 0 sequential f32 fold (compute_step.rs:54-58), 1 f64 compensated sum rounded once */
int oracle_compute_f32(uint32_t nchannel, const oracle_channel *ch, float weight, const float *pweight,
                       uint32_t iterations, uint32_t max_w, uint32_t max_h, int init_mode, int norm_mode,
                       const float *const *init_override, float *const *out, double *loop_seconds,
                       float *norm_trace);
int oracle_compute_f64(uint32_t nchannel, const oracle_channel *ch, float weight, const float *pweight,
                       uint32_t iterations, uint32_t max_w, uint32_t max_h, int init_mode,
                       const double *const *init_override, double *const *out, double *loop_seconds);

int oracle_init_f32(const oracle_channel *ch, uint32_t max_w, uint32_t max_h, int init_mode, float *fdata);
int oracle_gradient_f32(uint32_t nchannel, const oracle_channel *ch, float weight, const float *pweight,
                        uint32_t max_w, uint32_t max_h, const float *const *fdata, const float *const *cos,
                        float *const *grad, double *tv, double *tv2);
int oracle_projection_f32(const oracle_channel *ch, uint32_t max_w, uint32_t max_h, float *fdata, float *cos);
void oracle_dct8x8_f32(float *blocks, uint64_t nblocks);
void oracle_idct8x8_f32(float *blocks, uint64_t nblocks);
void oracle_dct8x8_f64(double *blocks, uint64_t nblocks);
void oracle_idct8x8_f64(double *blocks, uint64_t nblocks);
void oracle_boxing_f32(const float *in, float *out, uint32_t rw, uint32_t bw, uint32_t bh);
void oracle_unboxing_f32(const float *in, float *out, uint32_t rw, uint32_t bw, uint32_t bh);
float oracle_midpoint_f32(float a, float b);
void oracle_fista_scalars(uint32_t iterations, uint32_t pixel_count, float *next_terms, float *factors,
                          float *step_size);
/* Output stage of backend/artefact-lib/lib.rs:159-211 (+128 on luma, YCbCr -> RGB, clamp,
 * truncating u8); nchannel != 3 gives grey.  rgb: real_w*real_h*3 bytes. */
void oracle_output_rgb(uint32_t nchannel, const float *const *planes, uint32_t max_w, uint32_t real_w,
                       uint32_t real_h, uint8_t *rgb);
double oracle_objective(uint32_t nchannel, const oracle_channel *ch, float weight, const float *pweight,
                        uint32_t max_w, uint32_t max_h, const float *const *planes, double *terms);
double oracle_objective_f64planes(uint32_t nchannel, const oracle_channel *ch, float weight,
                                  const float *pweight, uint32_t max_w, uint32_t max_h,
                                  const double *const *planes, double *terms);
/* Threads for the loops the reference runs over channels (rayon); 1 by default. */
void oracle_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
