/*
 * skewengine.h -- C ABI of the MI355X-native skewEngine (libskewengine.so).
 *
 * Plain C: pointers, sizes and int status codes only.  Every entry point
 * names the reference interface it replaces (paths relative to the
 * reference repository root).  All grids are row-major float32, dimx
 * columns by dimy rows, values in the reference's units (DEM heights in
 * cell units, i.e. metres / cell size).
 *
 * Status: 0 on success, a negative SKE_E* code otherwise; the message of the
 * last failure on the calling thread is available from ske_last_error().
 * Engine handles are not thread-safe; distinct handles may be used from
 * distinct threads (each owns one HIP stream).
 */
#ifndef SKEWENGINE_H
#define SKEWENGINE_H

#ifdef __cplusplus
extern "C" {
#endif

#define SKE_OK 0
#define SKE_EINVAL (-1)      /* bad argument (null pointer, size, sector, operator id) */
#define SKE_EHIP (-2)        /* HIP runtime error */
#define SKE_ENOMEM (-3)      /* device or host allocation failed */
#define SKE_ESTATE (-4)      /* call out of order (e.g. kernel() before skew()) */
#define SKE_EUNSUPPORTED (-5) /* valid request this build does not support */

/* Operator ids: enum KERNELS (src/global.h:77). */
#define SKE_KERNEL_SDEM 0        /* total viewshed, kernelV3 (src/kernelSDEM.h:168) */
#define SKE_KERNEL_IDENTITYDEM 1 /* identity on a DEM (src/kernelUNIT.h:16) */
#define SKE_KERNEL_RADON 2       /* global Radon / FBP (src/kernelRADN.h:213) */
#define SKE_KERNEL_IDENTITYIMG 3 /* identity on an image (src/kernelUNIT.h:16) */
#define SKE_KERNEL_CEPSTRUM 4    /* cepstral blur (src/kernelBLUR.h:267) */

#define SKE_MAX_SECTOR 180 /* maxSector (src/global.cpp:69) */

/* Per-sector public fields after skew(angle) (src/skewEngine.cuh:356-367). */
typedef struct ske_sector_params {
    int a, newAngle, sectorType;
    int dim_i, dim_o, offset, skewHeight, skewWidth;
    double scale, skewness;
} ske_sector_params;

/* Reference device-operator form (the kernelcuda pointer,
 * src/skewEngine.cuh:369): a host function that launches the user's HIP
 * kernel on `stream` (a hipStream_t) with device pointers.  val is the
 * observer height the reference passes as POVh, angle is newAngle. */
typedef void (*ske_device_operator)(float *d_skewOutput, float *d_skewInput, int dim_i, int skewHeight,
                                    unsigned short *d_last, unsigned short *d_first, float val, int angle,
                                    void *stream, void *user);

typedef struct ske_engine ske_engine;

const char *ske_last_error(void);
int ske_version(void);
int ske_device_count(int *count);

/* DLL entry point: extern "C" int run(int argc, char *argv[])
 * (src/main.cu:199-202).  argv: [--kernel ID] [--input FILE(.flt|.bil|.png)]
 * [--dims NxN (crop)] [--output FILE.flt] [--sectors A:B] [--povh H]
 * [--device D] [--local] [--png-dir DIR] [--sample-dir DIR [--samples
 * x,y;x,y;...]] (--sample-dir: the reference's saveSampleData dump for the
 * Radon and identity-image operators; default points those of
 * src/global.cpp:54-56).  With no arguments it runs the total viewshed on a
 * built-in 500x500 synthetic DEM.  Returns 0 on success. */
int run(int argc, char **argv);

/* Whole-image path: allocate + prepare + the sector farm + reduce
 * (src/main.cu:71-172,186; src/skewEngine.cuh:388-411,861-914) for sectors
 * [sector_begin, sector_end) on device 0.  input/output are host buffers of
 * dimx*dimy floats; output is overwritten. */
int ske_execute(int kernel_id, int dimx, int dimy, const float *input, float *output, float povh,
                int sector_begin, int sector_end);

/* ske_execute without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture elsewhere in their program, tight serving loops, harnesses
 * that forbid hipMalloc).  One n x n grid (2 <= n <= SKE_NOALLOC_MAX_SIDE),
 * d_input and d_output device buffers of n*n floats on the current device,
 * sectors [sector_begin, sector_end), radon_local = the LRT flag of
 * SKE_OPT_RADON_LOCAL; stream is a hipStream_t (NULL: the null stream).
 * Every other buffer lives in one static device array of
 * SKE_NOALLOC_SCRATCH_MB MiB per GPU, held until the process ends.  Calls
 * no hipMalloc/hipFree; host-side planning uses ordinary host memory.
 * Returns when the result is in d_output.  Calls on one GPU share its array:
 * calls that overlap on one GPU race.  Results equal ske_execute's. */
#define SKE_NOALLOC_SCRATCH_MB 96
#define SKE_NOALLOC_MAX_SIDE 1024
int ske_execute_noalloc(int kernel_id, int n, const float *d_input, float *d_output, float povh,
                        int sector_begin, int sector_end, int radon_local, void *stream);

/* skewEngine<float>(dimx, dimy, inData, isGPU=true, deviceId)
 * (src/skewEngine.cuh:281,461) together with allocate/prepare (:388-411,
 * :962-987) and skewAlloc (:418-456).  `input` is a host grid; the engine
 * keeps the four oriented copies in device memory. */
int ske_create(int dimx, int dimy, const float *input, int device_id, ske_engine **out);
/* Same, with the grid already resident in device memory (copied). */
int ske_create_device(int dimx, int dimy, const float *d_input, int device_id, ske_engine **out);
/* ~skewEngine (src/skewEngine.cuh:484-501) + deallocate (:378-386). */
int ske_destroy(ske_engine *e);

/* Operator selection: the kernelcpu/kernelcuda assignment of
 * src/main.cu:107-136.  povh is POVh (observer height in cell units,
 * src/kernelSDEM.h:17,159), used by SKE_KERNEL_SDEM. */
int ske_set_kernel(ske_engine *e, int kernel_id, float povh);
/* Replace the operator by a user device operator (kernelcuda form). */
int ske_set_device_operator(ske_engine *e, ske_device_operator op, void *user);

/* skew(angle) (src/skewEngine.cuh:605-682): computeParams, skewShape,
 * resetSkew and the skewed copy of the oriented input. */
int ske_skew(ske_engine *e, int angle);
/* kernel() (src/skewEngine.cuh:694-746): run the operator on the skewed rows. */
int ske_kernel(ske_engine *e);
/* deskew(isMax) (src/skewEngine.cuh:755-793) into the engine's orientation
 * accumulator of the current sector type. */
int ske_deskew(ske_engine *e, int is_max);
/* reduce(output, isMax) (src/skewEngine.cuh:861-914): output (host, dimx*dimy)
 * is accumulated into (+ or max) from the orientation accumulators. */
int ske_reduce(ske_engine *e, float *output, int is_max);

/* Inspection of the current sector (public fields of skewEngine<T>). */
int ske_get_params(ske_engine *e, ske_sector_params *p);
/* target/weight: dim_i entries; first/last: skewHeight entries. */
int ske_get_tables(ske_engine *e, int *target, float *weight, unsigned short *first, unsigned short *last);
/* Sample points (saveSampleData and samplePoints, src/global.cpp:47-56,
 * src/main.cu:79-81): record, for `count` image cells (xy: x0, y0, x1, y1,
 * ... of grid 0; at most 16) and every sector, the four 64-float vectors the
 * local Radon transform (SKE_OPT_RADON_LOCAL) forms at that cell
 * (pointRadon, src/kernelRADN.h:164-170): the Gaussian-weighted window sums,
 * the weighted row through the cell, the filtered spectrum and the real part
 * of its transform.  count 0 stops recording.  Setting points clears the
 * record; each run that records overwrites it. */
int ske_set_sample_points(ske_engine *e, int count, const int *xy);
/* The record: 180 x count x 4 x 64 floats (sector, point, vector, index).
 * Zeros where nothing was recorded: sectors not run, operators other than
 * the local transform, cells outside their skewed row (as in the reference,
 * which visits only cells inside [first, last)). */
int ske_get_sample_data(ske_engine *e, float *out);
/* The sample dump of showResultsRADN (src/kernelRADN.h:263-382): files
 * punto<p>o.txt, punto<p>i.txt, punto<p>l.txt, punto<p>c.txt (p = 1..count)
 * in directory dir from a record laid out as ske_get_sample_data returns
 * it, in the reference's text layout (two blocks of 180 lines of 32 values). */
int ske_write_sample_files(const char *dir, int count, const float *data);

/* computeParams(angle) + skewShape() (src/skewEngine.cuh:503-539, 572-596)
 * without an engine or a device: the public fields and the index maps of
 * one sector of a dimx x dimy grid (same sizes as ske_get_tables; any
 * pointer may be NULL).  Host only. */
int ske_plan_sector(int dimx, int dimy, int angle, ske_sector_params *p, int *target, float *weight,
                    unsigned short *first, unsigned short *last);
/* resetSkew() (src/skewEngine.cuh:541-570): zero skewInput and skewOutput
 * over skewHeight*dim_i cells of the current sector (ske_skew does this
 * itself before skewing). */
int ske_reset_skew(ske_engine *e);
/* skewInput / skewOutput rows (skewHeight*dim_i floats each; either may be NULL). */
int ske_get_skew(ske_engine *e, float *skewInput, float *skewOutput);
/* Host-side operators (the kernelcpu form, src/skewEngine.cuh:370,696): after
 * ske_get_skew, the caller's function fills skewOutput on the host and this
 * uploads its skewHeight*dim_i floats as the current sector's skewOutput. */
int ske_set_skew_output(ske_engine *e, const float *skewOutput);
/* Device pointers of the current sector's buffers (valid until the next skew()). */
int ske_device_pointers(ske_engine *e, float **d_skewInput, float **d_skewOutput, unsigned short **d_first,
                        unsigned short **d_last);

/* All sectors [sector_begin, sector_end) of the engine's grid through
 * skew -> operator -> deskew -> reduce in one call (the OpenMP farm of
 * src/main.cu:145-164 with a deterministic ascending sector order).  The
 * deskew mode follows the operator (max for SKE_KERNEL_CEPSTRUM, as
 * main.cu:156).  The result stays on the device (ske_output_device) and is
 * copied to `output` (host, dimx*dimy floats) when it is not NULL. */
int ske_run(ske_engine *e, int kernel_id, int sector_begin, int sector_end, float povh, float *output);
int ske_output_device(ske_engine *e, const float **d_output);
int ske_synchronize(ske_engine *e);

/* Options (no reference counterpart). */
#define SKE_OPT_SDEM_VARIANT 1  /* viewshed inner loop: 0 bit-block (default; rows <= 4096 cells,
                                   longer rows use 2), 1 IEEE divide per step, 2 exact fma division
                                   per step.  All three give identical results. */
#define SKE_OPT_ARENA_BYTES 2   /* device budget for one chunk of skewOutput rows */
#define SKE_OPT_PROFILE 3       /* 1: time every launch with HIP events on the engine stream */
#define SKE_OPT_GRAPH 4         /* 1: capture the all-sector run once in a hipGraph and replay it */
#define SKE_OPT_SDEM_WAVES 5    /* viewshed: waves per workgroup (2, 4, 8); unset: 4, or 2 for rows of
                                 * fewer than four chunk groups (or chunks) per scan direction.
                                 * ske_get_option reports the set value (4 while unset), not
                                 * the per-chunk automatic choice */
#define SKE_OPT_SDEM_PERSISTENT 6 /* viewshed: 0 one workgroup per row, k>0 persistent grid of k per CU */
#define SKE_OPT_SDEM_OCCUPANCY 7  /* viewshed persistent grid: 0 default, 8 compiled for 8 waves/SIMD */
#define SKE_OPT_RADON_LOCAL 8     /* 1: SKE_KERNEL_RADON computes the local transform (pointRadon per
                                   * cell), the reference's LRT=true (src/kernelRADN.h:212-220) */
#define SKE_OPT_COUNT_STEPS 9     /* 1: the all-sector bit-block viewshed counts the ray steps its
                                   * lanes execute (counter "op_sdem_lane_steps"; the scans stop
                                   * early once no later target can be visible) */
#define SKE_OPT_SDEM_GROUP 10      /* viewshed: 4 (default) groups of four 64-observer chunks per
                                   * wave whose finished scans are replaced by the group's remaining
                                   * ones (lane compaction); 1 one chunk per wave */
int ske_set_option(ske_engine *e, int option, long long value);
int ske_get_option(ske_engine *e, int option, long long *value);
/* Accumulated HIP-event time of the launches named `name` since the last
 * reset (profiling option on): total milliseconds and launch count. */
int ske_profile_get(ske_engine *e, const char *name, double *total_ms, int *count);
/* Accumulated value of a work counter since the last profile reset
 * ("op_sdem_lane_steps" with SKE_OPT_COUNT_STEPS on); 0 if never counted. */
int ske_counter_get(ske_engine *e, const char *name, double *value);
int ske_profile_reset(ske_engine *e);

/* Work counts of the exact maps for sectors [s0,s1) of a dimx x dimy grid:
 * cells = sum of row lengths, steps = sum L(L-1) (viewshed ray steps). */
int ske_work(int dimx, int dimy, int s0, int s1, double *cells, double *steps);

/* ---- Batches: many square grids of uneven sizes in one call.  No reference
 * counterpart (the reference runs one grid per process); each grid's result
 * equals what ske_run gives for it alone. ---- */
typedef struct ske_batch ske_batch;
/* sides[i] = dimx = dimy of grid i; inputs[i] = host (or, for the _device
 * variant, device) pointer to sides[i]^2 floats.  Inputs are copied. */
int ske_batch_create(int count, const int *sides, const float *const *inputs, int device_id, ske_batch **out);
int ske_batch_create_device(int count, const int *sides, const float *const *d_inputs, int device_id,
                            ske_batch **out);
int ske_batch_destroy(ske_batch *b);
/* Sectors [sector_begin, sector_end) of every grid: one operator launch per
 * arena chunk of (grid, sector) pairs (a persistent grid for the viewshed),
 * one deskew per grid and chunk, one reduce per grid; with SKE_OPT_GRAPH the
 * sequence is captured once and replayed.  Asynchronous: results are ready
 * after ske_batch_synchronize / ske_batch_output. */
int ske_batch_run(ske_batch *b, int kernel_id, int sector_begin, int sector_end, float povh);
int ske_batch_output(ske_batch *b, int index, float *output);
int ske_batch_output_device(ske_batch *b, int index, const float **d_output);
int ske_batch_synchronize(ske_batch *b);
int ske_batch_set_option(ske_batch *b, int option, long long value);
int ske_batch_get_option(ske_batch *b, int option, long long *value);
int ske_batch_profile_get(ske_batch *b, const char *name, double *total_ms, int *count);
int ske_batch_counter_get(ske_batch *b, const char *name, double *value);
int ske_batch_profile_reset(ske_batch *b);
/* Autotuner of the viewshed launch shape: times candidate (waves per
 * workgroup, persistent workgroups per CU, occupancy hint, chunk group)
 * shapes on a sample of the batch (one grid per distinct side, every 12th
 * sector; a warm-up, then the best of three timed launches each), keeps the
 * first candidate within 2% of the fastest and returns it. */
int ske_batch_autotune(ske_batch *b, int kernel_id, float povh, int *waves, int *wgs_per_cu, int *occupancy,
                       double *best_ms);

/* ---- Standard parallel-beam Radon / FBP.  No reference counterpart: the
 * reference's Radon operator (src/kernelRADN.h:212-259) works in its own
 * skewed geometry; this is the 725-bin geometry of the benchmark
 * configurations.  Image n x n on [-1,1]^2 (row 0 at y=+1), bins
 * s_k = (k-(n_det-1)/2)*2/n, ray samples on the same pitch with bilinear
 * interpolation, Ram-Lak filter as a (2*n_det-1)-tap convolution,
 * back-projection (pi/n_angles) * sum_a q(a, x cos + y sin), linear in s. */
typedef struct ske_fbp ske_fbp;
#define SKE_FBP_IMAGE 0    /* batch x n x n */
#define SKE_FBP_SINO 1     /* batch x n_angles x n_det */
#define SKE_FBP_FILTERED 2 /* batch x n_angles x n_det */
#define SKE_FBP_RECON 3    /* batch x n x n */
/* angles: n_angles radians, or NULL for a*pi/n_angles, a = 0..n_angles-1.
 * Device memory: the four buffers above for max_batch slices, plus the
 * projector's zero-bordered image copy, max_batch x (n+8)^2 floats. */
int ske_fbp_create(int n, int n_angles, int n_det, const float *angles, int max_batch, int device_id,
                   ske_fbp **out);
int ske_fbp_destroy(ske_fbp *f);
int ske_fbp_upload(ske_fbp *f, int buffer, int batch, const float *host);
int ske_fbp_download(ske_fbp *f, int buffer, int batch, float *host);
int ske_fbp_buffer(ske_fbp *f, int buffer, float **d_ptr);
int ske_fbp_project(ske_fbp *f, int batch);     /* IMAGE -> SINO */
int ske_fbp_filter(ske_fbp *f, int batch);      /* SINO -> FILTERED with the current taps */
int ske_fbp_backproject(ske_fbp *f, int batch); /* FILTERED -> RECON */
int ske_fbp_reconstruct(ske_fbp *f, int batch); /* filter + backproject */
/* 2*n_det-1 filter taps h(m), m = -(n_det-1)..n_det-1 (the bin pitch folded
 * in); NULL restores Ram-Lak. */
int ske_fbp_set_taps(ske_fbp *f, const float *taps);
int ske_fbp_get_taps(ske_fbp *f, float *taps);
int ske_fbp_synchronize(ske_fbp *f);
int ske_fbp_set_option(ske_fbp *f, int option, long long value); /* SKE_OPT_PROFILE */
int ske_fbp_profile_get(ske_fbp *f, const char *name, double *total_ms, int *count);
int ske_fbp_profile_reset(ske_fbp *f);

/* ---- PNG input and result images (src/helper.h:167-186 read_png via
 * lodepng; src/main.h:124-141 and src/kernelRADN.h:383-401 showResults).
 * ske_png_read: width/height of a non-interlaced PNG and, when pixels is
 * not NULL, its grey level R+G+B (0..765) per pixel of the 8-bit RGBA
 * decode, row-major, as the reference's image operators take their input. */
int ske_png_read(const char *path, int *width, int *height, float *pixels);
/* 8-bit grey (channels 1) or RGB (channels 3) PNG, rows top to bottom. */
int ske_png_write(const char *path, int width, int height, int channels, const unsigned char *data);
/* salida1.png (input grey level / 3) and salida2.png (the output mapped to
 * grey: Radon auto-scaled (out - min) * 255/(max - min), identity image
 * out/540) in directory dir; kernel_id SKE_KERNEL_RADON or
 * SKE_KERNEL_IDENTITYIMG. */
int ske_result_pngs(const char *dir, int kernel_id, int width, int height, const float *pixels,
                    const float *output);

/* ---- Learned sparse-view FBP.  No reference counterpart: it extends the
 * Radon/FBP operator (src/kernelRADN.h:212-259) with a learnable filter and
 * a residual CNN, in the geometry of an ske_fbp whose stream, timer and
 * buffers it shares (SINO in, RECON out, IMAGE = training target):
 *   r0  = BP(h * p), h[m] = taps[|m|] (n_det symmetric taps; the DCT of the
 *         taps is the filter's per-frequency response)
 *   out = r0 + conv(relu(conv(relu(conv(relu(conv(r0))))))) with 3x3 convs
 *         1->C->C->C->1 (C = 16 or 32 hidden channels), zero padding, bf16
 *         weights and activations, fp32 accumulation.
 * Parameters (fp32, taps offset + n_det of them): W1[C][9] b1[C]
 * W2[C][9][C] b2[C] W3[C][9][C] b3[C] W4[9][C] b4 taps[n_det]; tap
 * t = 3*ky + kx reads pixel (i+ky-1, j+kx-1).  The taps offset is
 * SKE_LEARNED_TAPS for C = 16 and SKE_LEARNED_TAPS_C32 for C = 32.  The
 * image side must be a multiple of 32. */
typedef struct ske_learned ske_learned;
#define SKE_LEARNED_TAPS 4945  /* offset of the filter taps in the parameters (16 channels) */
#define SKE_LEARNED_TAPS_C32 19105  /* the same for 32 hidden channels */
#define SKE_LEARNED_R0 0       /* fp32 batch x n x n: learned-filter FBP */
#define SKE_LEARNED_A1 1       /* bf16 (uint16) batch x n x n x C: hidden activations of the last
                                * ske_learned_gradient / train_step (reconstruct keeps them in LDS) */
#define SKE_LEARNED_A2 2
#define SKE_LEARNED_A3 3
#define SKE_LEARNED_DR0 4      /* fp32 batch x n x n: d loss / d r0 (after a gradient) */
#define SKE_LEARNED_FILTERED_GRAD 5 /* fp32 batch x n_angles x n_det: d loss / d filtered sinogram */
int ske_learned_create(ske_fbp *f, ske_learned **out); /* 16 hidden channels */
int ske_learned_create_channels(ske_fbp *f, int channels, ske_learned **out); /* channels: 16 or 32 */
int ske_learned_destroy(ske_learned *l);
int ske_learned_param_count(ske_learned *l, int *count);
/* hidden channels and the offset of the filter taps in the parameters */
int ske_learned_layout(ske_learned *l, int *channels, int *taps_offset);
/* He-normal hidden weights from seed, zero biases, zero output layer and
 * Ram-Lak taps: the output equals plain FBP until trained. */
int ske_learned_init(ske_learned *l, unsigned long long seed);
int ske_learned_set_params(ske_learned *l, const float *params); /* also resets the optimiser */
int ske_learned_get_params(ske_learned *l, float *params);
/* SINO -> RECON.  The CNN runs fused (one kernel, activations in LDS);
 * the result is bit-identical to the per-layer training forward. */
int ske_learned_reconstruct(ske_learned *l, int batch);
/* loss = mean((out - IMAGE)^2) over the batch and its gradient with respect
 * to every parameter (grad: param_count floats, or NULL). */
int ske_learned_gradient(ske_learned *l, int batch, double *loss, float *grad);
/* One Adam step (0.9, 0.999, 1e-8): lr for the CNN, lr_taps for the taps. */
int ske_learned_train_step(ske_learned *l, int batch, float lr, float lr_taps, double *loss);
int ske_learned_download(ske_learned *l, int buffer, int batch, void *host);
/* Random ellipses (x0, y0, a, b, phi, rho): centres uniform in the unit disk,
 * semi-axes U[0.05,0.4], rotation U[0,pi), intensity U[-0.5,1.0]. */
int ske_random_ellipses(unsigned long long seed, int batch, int n_ellipses, float *out);
/* batch x n_ellipses ellipses summed at pixel centres into IMAGE. */
int ske_learned_render(ske_learned *l, int batch, int n_ellipses, const float *ellipses);
/* Training pairs: random phantoms (seed) into IMAGE, projected into SINO, then
 * Poisson transmission noise: counts ~ Poisson(i0 exp(-mu p)),
 * p = -ln(max(counts, 1) / i0) / mu (i0 <= 0: no noise). */
int ske_learned_make_batch(ske_learned *l, int batch, int n_ellipses, unsigned long long seed, float i0, float mu);

#ifdef __cplusplus
}
#endif
#endif /* SKEWENGINE_H */
