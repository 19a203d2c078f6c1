/*
 * artefact_hip -- MI355X (gfx950) reconstruction loop of artefact-lib behind a C ABI.
 *
 * The entry points replace the crate-private compute function of
 * backend/artefact-lib/pipeline_scalar/mod.rs:17-26 (identical signatures at
 * pipeline_simd_8/mod.rs:22-31 and pipeline_simd_adaptive/mod.rs:19-28), which
 * backend/artefact-lib/lib.rs:122-153 calls in joint (nchannel = 3) or
 * separate (nchannel = 1, one call per component) mode:
 *
 *   fn compute(nchannel: usize, coefs: Vec<Coefficient>, weight: f32, pweight: [f32; 3],
 *              iterations: usize, max_rounded_px_w: u32, max_rounded_px_h: u32,
 *              max_rounded_px_count: usize) -> Vec<Vec<f32>>
 *
 * Plain pointers and sizes only.  Every call is re-entrant: a plan owns its own
 * HIP stream and device buffers, and the library's only shared state is the
 * autotuner's per-geometry cache (behind a mutex), so concurrent calls from several
 * host threads (the reference's separate-components mode fans out over rayon
 * workers) are safe.  The exception is artefact_hip_compute_noalloc, whose calls
 * share one static scratch array per device (see there).
 */
#ifndef ARTEFACT_HIP_H
#define ARTEFACT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes. */
#define ARTEFACT_OK              0
#define ARTEFACT_E_ARGUMENT     -1  /* null pointer, nchannel not in 1..3, bad mode */
#define ARTEFACT_E_GEOMETRY     -2  /* plane sizes not multiples of 8, inconsistent block counts, samp factor not 1/2, component larger than the grid, max_w*max_h > 2^29 */
#define ARTEFACT_E_QUANT        -3  /* a quantisation table entry is 0 (the reference divides by q^2, compute_step_prob.rs:29) */
#define ARTEFACT_E_HIP          -4  /* HIP runtime failure (allocation, launch, copy) */
#define ARTEFACT_E_STATE        -5  /* plan used out of order (iterate before init) */
#define ARTEFACT_E_JPEG         -6  /* malformed or unsupported JPEG data (artefact_jpeg_decode) */
#define ARTEFACT_E_SCRATCH      -7  /* artefact_hip_compute_noalloc: the grid's buffers exceed ARTEFACT_SCRATCH_BYTES */
#define ARTEFACT_BENCHMARK       1  /* artefact_hip_process with benchmark set: the loop ran, no image was
                                       written (the reference returns Err("BENCHMARK"), lib.rs:155-157) */

#define ARTEFACT_MODE_TGV 0  /* TV + second-order term with strength weight (compute_step.rs:42-48) */
#define ARTEFACT_MODE_TV  1  /* TV only: the same loop with weight forced to 0 */

#define ARTEFACT_PRECISION_F32    0  /* the reference's f32 arithmetic, bit for bit (default) */
#define ARTEFACT_PRECISION_BF16X3 1  /* projection DCT/IDCTs on BF16 MFMA, hi/lo split, f32 accumulation */
#define ARTEFACT_PRECISION_F32_FAST 2 /* f32 gradient with hardware sqrt/reciprocal, plain f64 norm sum */

/* One colour component: backend/artefact-lib/jpeg/mod.rs:8-27 (Coefficient).
 * The reference carries coefficients and the quantisation table as f32 holding
 * i16 / u16 values (backend/artefact-lib/jpeg/zune.rs:33-45); they pass through
 * this boundary as the integers themselves, bit-exact. */
typedef struct {
    uint32_t rounded_w, rounded_h;  /* rounded_px_w / rounded_px_h: multiples of 8 */
    uint32_t block_w, block_h;      /* rounded_w / 8, rounded_h / 8 */
    uint32_t h_samp, v_samp;        /* horizontal/vertical_samp_factor: upsampling onto the max grid, 1 or 2 */
    const int16_t *coefs;           /* dct_coefs: block_w*block_h*64, raster block order, natural order in a block */
    const uint16_t *quant;          /* quant_table: 64, natural order */
} artefact_channel;

/* The scalar arguments of compute().  pweight is indexed by channel as in
 * compute_step.rs:26,30 (so a separate-mode call with nchannel = 1 uses pweight[0]). */
typedef struct {
    uint32_t nchannel;              /* 1..3 */
    uint32_t max_w, max_h;          /* max_rounded_px_w / max_rounded_px_h (lib.rs:108-116) */
    float weight;                   /* second-order weight (ignored in ARTEFACT_MODE_TV) */
    float pweight[3];               /* coefficient-distance weights */
    uint32_t iterations;
    uint32_t mode;                  /* ARTEFACT_MODE_TGV or ARTEFACT_MODE_TV */
    const float *const *init_override; /* NULL, or nchannel host planes of max_w*max_h replacing the decoded start image */
} artefact_params;

/* One-shot drop-in for compute(): copies the coefficients to the GPU, runs the
 * init and `iterations` iterations, and writes nchannel planes of max_w*max_h f32
 * into caller-allocated host buffers out_planes[c].  Channel 0 stays level-shifted
 * (the caller adds 128, lib.rs:160-162), chroma stays centred at 0.
 * Replaces pipeline_scalar/mod.rs:17-78. */
int artefact_hip_compute(const artefact_params *params, const artefact_channel *channels,
                         float *const *out_planes);

/* Resident form of the same call, for callers that keep the image in HBM across
 * calls (and for timing the iteration loop alone). */
typedef struct artefact_hip_plan artefact_hip_plan;

/* Validates the geometry, allocates device buffers on the current device and
 * copies coefficients and quantisation tables (the Coefficient conversion of
 * pipeline_scalar/mod.rs:27-30). */
int artefact_hip_plan_create(const artefact_params *params, const artefact_channel *channels,
                             artefact_hip_plan **out_plan);
/* Start image: ScalarCoef::from + Aux::init (pipeline_simd_8/coef.rs:88-118 order,
 * utils/aux.rs:45-65), or params->init_override when it was given. */
int artefact_hip_plan_init(artefact_hip_plan *plan);
/* The iteration loop, pipeline_scalar/mod.rs:45-75.  Asynchronous on the plan's stream. */
int artefact_hip_plan_iterate(artefact_hip_plan *plan);
/* Synchronises the plan's stream and copies the planes out (pipeline_scalar/mod.rs:77). */
int artefact_hip_plan_download(artefact_hip_plan *plan, float *const *out_planes);
/* Copies the current coefficient estimate `cos` (block order, M_c floats per
 * channel), the last projection's clamped DCT values (compute_projection.rs:82). */
int artefact_hip_plan_download_cos(artefact_hip_plan *plan, float *const *out_cos);
/* Per-channel gradient norms of the last iteration run (compute_step.rs:54-58). */
int artefact_hip_plan_last_norms(artefact_hip_plan *plan, float *out_norms);
/* Gradient of the first step from the start image, before the norm
 * (compute_step.rs:21-48): prob + TV + second-order term, nchannel planes of max_w*max_h. */
int artefact_hip_plan_first_gradient(artefact_hip_plan *plan, float *const *out_grad);
/* Runs the iteration loop without a graph, bracketing every kernel with HIP events
 * on the plan's stream: avg_ms[0] = mean k_gradient duration, avg_ms[1] = mean
 * k_project duration, avg_ms[2] = the whole loop, all in milliseconds.  Like
 * iterate, it consumes the initialised state.  No reference counterpart: the
 * reference's only timing hook is its benchmark exit (lib.rs:155-157). */
int artefact_hip_plan_profile(artefact_hip_plan *plan, float *avg_ms);
/* Timing of the product path itself: after set_timing(plan, n), the next n calls of
 * artefact_hip_plan_iterate record a HIP event pair on the plan's stream around their
 * launch (the captured graph, or the eager loop), without synchronising.  timing() waits
 * for the last pair and returns the summed duration and the number of timed launches.
 * set_timing(plan, 0) switches it off.  No reference counterpart (see profile above). */
int artefact_hip_plan_set_timing(artefact_hip_plan *plan, uint32_t launches);
int artefact_hip_plan_timing(artefact_hip_plan *plan, float *total_ms, uint32_t *launches);
/* Objective of the current iterate in f64: terms[0] = sum alpha*n(p) (TV form of
 * compute_step_tv.rs:62-71), terms[1] = sum a2*n2(p) (second-order form of
 * compute_step_tv2.rs:53-90), terms[2] = sum_c alpha_p/2 * sum ((DCT(down x)_j - c_j q_j)/q_j)^2.
 * The reference forms only the first two as discarded partial sums (compute_step.rs:39,42);
 * no reference counterpart for the total. */
int artefact_hip_plan_objective(artefact_hip_plan *plan, double *terms);
/* The plan's HIP stream (hipStream_t), for callers that record events on it. */
void *artefact_hip_plan_stream(artefact_hip_plan *plan);
/* Enables (1, default) or disables (0) capturing the iteration loop in a hipGraph. */
int artefact_hip_plan_set_graph(artefact_hip_plan *plan, int enable);
void artefact_hip_plan_destroy(artefact_hip_plan *plan);

/* Selects the loop's arithmetic: ARTEFACT_PRECISION_F32 (default, bit-identical to the
 * reference's f32 operations), ARTEFACT_PRECISION_BF16X3 (the precision form: every 8x8 DCT
 * and IDCT of the projection as BF16 MFMA contractions with hi/lo operands and f32
 * accumulation; the stencil and descent stay f32) or ARTEFACT_PRECISION_F32_FAST (the
 * gradient sweep with v_sqrt_f32 / v_rcp_f32 and (alpha / n) * term in place of the correctly
 * rounded (alpha * term) / n, and a plain f64 sum for the norm; projection unchanged; within a
 * few ulp per term, not bit-identical).  BF16X3 needs every channel at full grid resolution
 * (4:4:4 or one channel) or 4:2:0 colour, else ARTEFACT_E_GEOMETRY.  No reference counterpart. */
int artefact_hip_plan_set_precision(artefact_hip_plan *plan, uint32_t precision);

/* Autotuner result: the k_gradient strip height (output rows per wave) and the projection
 * kernel (2 = the fused multi-channel kernel without LDS, 1 = the fused kernel with LDS
 * transposes, 0 = the per-channel kernel) chosen for the plan's geometry, the measured
 * milliseconds per iteration, and the workgroup count of the kernel without LDS (each of its
 * workgroups forms the gradient norms once and then projects several block-row units; 0 =
 * one unit per wave). */
typedef struct {
    uint32_t kg_rows;
    uint32_t project_kernel;
    float ms_per_iteration;
    uint32_t project_workgroups;
    uint32_t xcd_bands;  /* 1: projection workgroups on the XCD whose k_gradient workgroups swept their rows */
} artefact_tuning;

/* Times candidate strip heights, projection kernels and projection workgroup counts on the
 * plan (a few iterations each) and keeps the fastest; the choice is cached per geometry for the
 * process.  Leaves the plan uninitialised (call artefact_hip_plan_init next).  ARTEFACT_KG_ROWS,
 * ARTEFACT_GENERIC_PROJECT, ARTEFACT_PROJECT_KERNEL, ARTEFACT_LANE_WGS and ARTEFACT_XCD_BANDS in
 * the environment pin their parameter.  No reference counterpart: the reference has one fixed schedule. */
int artefact_hip_plan_tune(artefact_hip_plan *plan, artefact_tuning *out);

/* Many images in one hipGraph (the batched form).  Each image is one compute() call's
 * arguments (init_override must be NULL).  The graph holds every image's start image and
 * iteration loop; images are spread over `streams` HIP streams, largest first onto the least
 * loaded, so independent solves overlap on the GPU.  With `tune` set every geometry is
 * autotuned first, and `streams` = 0 times the whole graph with 1 to 4 streams and keeps the
 * fastest; without `tune`, 0 means 2.  No reference counterpart: the reference processes one
 * image per call (lib.rs:99-157). */
typedef struct {
    const artefact_params *params;
    const artefact_channel *channels;
} artefact_batch_image;
typedef struct artefact_hip_batch artefact_hip_batch;

int artefact_hip_batch_create(const artefact_batch_image *images, uint32_t count, uint32_t streams, uint32_t tune,
                              artefact_hip_batch **out_batch);
/* The number of streams the batch's graph spreads its images over (0 for a null batch). */
uint32_t artefact_hip_batch_streams(const artefact_hip_batch *batch);
/* Launches the graph (asynchronous): every image is solved from its coefficients. */
int artefact_hip_batch_run(artefact_hip_batch *batch);
int artefact_hip_batch_synchronize(artefact_hip_batch *batch);
/* Synchronises and copies image `index`'s planes out, as artefact_hip_plan_download. */
int artefact_hip_batch_download(artefact_hip_batch *batch, uint32_t index, float *const *out_planes);
/* The plan of image `index` (owned by the batch), for the plan calls that read results
 * (download_rgb, objective, last_norms); synchronise the batch first. */
int artefact_hip_batch_plan(artefact_hip_batch *batch, uint32_t index, artefact_hip_plan **out_plan);
void artefact_hip_batch_destroy(artefact_hip_batch *batch);

/* Output stage of Artefact::process (lib.rs:159-211), run on the GPU from the plan's
 * current planes: luma + 128, then for three channels R = 1.402*cr + y,
 * G = 0.71414*(-cr) + (0.34414*(-cb) + y), B = 1.772*cb + y with every product and sum
 * rounded on its own (utils/macros.rs:10-15), clamp to [0, 255] and truncation to u8 (NaN
 * gives 0, as Rust's `as u8`); one or two channels give grey R = G = B from channel 0.
 * Writes real_w*real_h*3 bytes, row-major, interleaved RGB, for the top-left
 * real_w x real_h of the grid (lib.rs:168-172).  Uses the plan's gradient buffer as staging,
 * so it must not run while an iterate on the same plan is in flight (it synchronises the
 * plan's stream before returning). */
int artefact_hip_plan_download_rgb(artefact_hip_plan *plan, uint32_t real_w, uint32_t real_h, uint8_t *out_rgb);

/* The arguments of Artefact::process once the JPEG is decoded (lib.rs:49-57, 99-157):
 * ValueCollection::to_slice already applied (lib.rs:41-46). */
typedef struct {
    uint32_t nchannel;              /* Jpeg::nchannel, 1..3 */
    uint32_t real_w, real_h;        /* Jpeg::real_px_w / real_px_h */
    float weight[3];
    float pweight[3];
    uint32_t iterations[3];
    uint32_t separate_components;   /* nonzero: one single-channel loop per component */
    uint32_t benchmark;             /* nonzero: run the loops, write nothing, return ARTEFACT_BENCHMARK */
} artefact_process_params;

/* Artefact::process from decoded coefficients to the RGB image (lib.rs:105-211).  The max
 * grid is the largest component (lib.rs:108-116).  Three channels, not separate: one joint
 * loop with weight[0], pweight, iterations[0] (lib.rs:121-132).  Otherwise one loop per
 * component c with nchannel = 1, weight[c], iterations[c] and the whole pweight array,
 * so every component uses pweight[0] (lib.rs:134-149, compute_step.rs:26,30); the
 * per-component loops run concurrently, each on its own stream.  Then the output stage of
 * artefact_hip_plan_download_rgb into out_rgb (real_w*real_h*3 bytes; may be NULL when
 * benchmark is set). */
int artefact_hip_process(const artefact_process_params *params, const artefact_channel *channels,
                         uint8_t *out_rgb);

/* ---- Learned warm start (no reference counterpart) ----
 * A residual CNN (five 3x3 convolutions, widths 2 -> 32 -> 32 -> 32 -> 32 -> 1, ReLU, output
 * added to the decoded plane) that predicts the converged plane from two features per pixel:
 * the decoded plane and its quantisation-step map, the plane's 64-entry table (natural order)
 * tiled over its 8x8 blocks, q[(y mod 8) * 8 + (x mod 8)].  Parameters are a flat f32 vector:
 * per layer w[cout][cin][3][3] then b[cout].  Forward, backward and Adam run as HIP kernels
 * (BF16 MFMA, f32 accumulation).  `quant` arguments hold 64 floats per plane. */
typedef struct artefact_cnn artefact_cnn;
uint32_t artefact_cnn_param_count(void);
int artefact_cnn_create(const float *params, artefact_cnn **out_net);
int artefact_cnn_get_params(artefact_cnn *net, float *out_params);
/* Prediction (image units) for `batch` host planes of h x w with their quantisation tables,
 * through the inference path of artefact_hip_plan_warm_init: one streaming kernel for all five
 * layers (ARTEFACT_CNN_LAYERED=1 in the environment selects one kernel per layer instead). */
int artefact_cnn_forward(artefact_cnn *net, const float *x0, const float *quant, uint32_t batch, uint32_t h,
                         uint32_t w, float *out_pred);
/* Mean squared error (in units of 1/64) and its parameter gradients for a host batch; w must
 * be a multiple of 16.  grads may be NULL. */
int artefact_cnn_gradients(artefact_cnn *net, const float *x0, const float *quant, const float *target,
                           uint32_t batch, uint32_t h, uint32_t w, double *loss, float *grads);
/* Keeps a training set (count planes of h x w, their tables, targets) resident in HBM. */
int artefact_cnn_set_dataset(artefact_cnn *net, const float *x0, const float *quant, const float *target,
                             uint32_t count, uint32_t h, uint32_t w);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the samples `indices` of the set. */
int artefact_cnn_train_step(artefact_cnn *net, const uint32_t *indices, uint32_t batch, float lr, double *loss);
void artefact_cnn_destroy(artefact_cnn *net);

/* Warm start of a plan: the CNN's prediction for every channel's decoded plane at its
 * component resolution (subsampled chroma at its own size), upsampled bilinearly onto the grid
 * (the decoded start image replicates, pipeline_scalar/coef.rs:69-92), one projection into the
 * quantisation box (compute_projection.rs:8-122), and that image as both x and x_prev.
 * artefact_hip_plan_iterate then runs the plan's iterations from there. */
int artefact_hip_plan_warm_init(artefact_hip_plan *plan, artefact_cnn *net);
/* The iteration count the constant step is derived from: step = sqrt(N)/2/sqrt(1 + n)
 * (pipeline_scalar/mod.rs:45,64-74 uses the plan's own iteration count; 0 restores that).
 * A warm-started refinement of K iterations can take the step of a longer schedule. */
int artefact_hip_plan_set_step_iterations(artefact_hip_plan *plan, uint32_t n);

/* Starts the plan's FISTA recurrence (pipeline_scalar/mod.rs:50-52) k steps in: the loop's
 * first iteration takes the extrapolation factor of the recurrence's step k + 1, as if k
 * iterations had already run (0 restores the reference's start).  For a refinement from a start
 * already near the result, such as the warm start's; k <= 2^20.  No reference counterpart. */
int artefact_hip_plan_set_momentum_offset(artefact_hip_plan *plan, uint32_t k);

/* Lets the step decay over the loop: iteration i of n takes s_first^(1 - i/(n-1)) *
 * s_last^(i/(n-1)), where s_m = sqrt(N)/2/sqrt(1 + m) is the step of an m-iteration schedule
 * (pipeline_scalar/mod.rs:45,64-74; m = 0: the plan's iteration count).  first = 0 gives the
 * constant step of artefact_hip_plan_set_step_iterations(plan, last), which also clears a
 * schedule.  For a refinement from a start already near the result, such as the warm start's.
 * No reference counterpart. */
int artefact_hip_plan_set_step_schedule(artefact_hip_plan *plan, uint32_t first, uint32_t last);

/* ---- JPEG input: Jpeg::from (backend/artefact-lib/jpeg/zune.rs:5-62) on the host ----
 * Decodes baseline/extended sequential and progressive Huffman JPEGs (8-bit, 1-4 components,
 * sampling factors 1 or 2) to the reference's Coefficient layout: per component the plane
 * rounded to the MCU, int16 coefficients in raster block order and natural order inside a
 * block, not dequantised, and the component's quantisation table in natural order.  The
 * zune-jpeg fork's placement rules are kept (see jpeg_decode.cpp).  On failure returns
 * ARTEFACT_E_JPEG and writes a message into err (err_len bytes, may be NULL). */
typedef struct artefact_jpeg artefact_jpeg;
int artefact_jpeg_decode(const uint8_t *data, size_t size, artefact_jpeg **out_jpeg, char *err, size_t err_len);
/* Jpeg::nchannel, real_px_w, real_px_h (jpeg/mod.rs:29-36). */
uint32_t artefact_jpeg_nchannel(const artefact_jpeg *jpeg);
int artefact_jpeg_size(const artefact_jpeg *jpeg, uint32_t *real_w, uint32_t *real_h);
/* Fills nchannel artefact_channel entries whose pointers stay valid until artefact_jpeg_free;
 * they can go straight to artefact_hip_process or artefact_hip_compute. */
int artefact_jpeg_channels(const artefact_jpeg *jpeg, artefact_channel *out_channels);
void artefact_jpeg_free(artefact_jpeg *jpeg);

/* ---- Image files: the RGB image of artefact_hip_process to file bytes ----
 * The reference's front ends save Artefact::process's RgbImage with the `image` crate
 * (backend/artefact-cli/main.rs:156 img.save(output); backend/artefact-wasm/lib.rs:41
 * write_to(.., ImageFormat)), in the formats its features enable (png, bmp, tiff, webp;
 * backend/artefact-lib/Cargo.toml:7-12).  Built here: PNG (8-bit RGB, adaptive filters,
 * deflate), BMP (24-bit), TIFF (baseline RGB, uncompressed) and WebP (lossless VP8L, as the
 * image crate writes it; sides up to 16384).  Host code. */
#define ARTEFACT_IMAGE_PNG  0
#define ARTEFACT_IMAGE_BMP  1
#define ARTEFACT_IMAGE_TIFF 2
#define ARTEFACT_IMAGE_WEBP 3
/* Upper bound of the encoded size of a w x h RGB image, 0 for an unknown format or empty image. */
size_t artefact_image_encode_bound(uint32_t w, uint32_t h, uint32_t format);
/* Encodes w*h*3 bytes of row-major interleaved RGB into out (cap bytes, at least the bound)
 * and writes the file's size to *out_len.  ARTEFACT_E_ARGUMENT for a null pointer, an empty
 * image, an unknown format, a file past the format's 4 GiB offsets or too small a buffer. */
int artefact_image_encode(const uint8_t *rgb, uint32_t w, uint32_t h, uint32_t format, uint8_t *out, size_t cap,
                          size_t *out_len);

/* ---- No-allocation form of artefact_hip_compute ----
 * No counterpart in the reference.  For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, hosts that forbid allocation after start-up.  The
 * same loop as artefact_hip_compute (the same kernels, bit-identical results), but
 *   - channels[c].coefs are DEVICE pointers (16-byte aligned int16 coefficients, layout as
 *     artefact_channel), channels[c].quant stays a host pointer (read during the call);
 *   - out_planes[c] are DEVICE pointers to nchannel planes of max_w*max_h f32;
 *   - every other buffer is carved from one static device array of ARTEFACT_SCRATCH_BYTES that
 *     each device running the library holds until the process ends; no hipMalloc, hipFree or
 *     other allocating routine is called (the first call on a device looks up the array's
 *     address);
 *   - the kernels and a device-to-device copy into out_planes are queued on `stream` (a
 *     hipStream_t; NULL is the null stream) and the call returns without waiting;
 *   - init_override must be NULL; the strip height and projection kernel are the untuned
 *     defaults (the results do not depend on them).
 * Grids whose buffers (16 * nchannel * max_w * max_h bytes plus 8 bytes per coefficient and
 * small tables) exceed ARTEFACT_SCRATCH_BYTES return ARTEFACT_E_SCRATCH: 1024x1024 4:2:0 colour
 * or 1664x1664 luma fit.  Calls on one device share its array and keep the last call's
 * contents: two calls whose work overlaps on one device (other streams or threads) race. */
#define ARTEFACT_SCRATCH_BYTES (64u << 20)
int artefact_hip_compute_noalloc(const artefact_params *params, const artefact_channel *channels,
                                 float *const *out_planes, void *stream);

const char *artefact_hip_strerror(int code);

#ifdef __cplusplus
}
#endif
#endif
