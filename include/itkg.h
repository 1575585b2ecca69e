/*
 * itkg.h -- C ABI of the MI355X (gfx950) separable Gaussian for the itk-rs
 * image pipeline.
 *
 * The library replaces the body of itk-rs's GaussianFilter::generate_region
 * (src/filters/gaussian.rs:59-90) and the pass it calls for every axis,
 * convolve_axis (src/filters/conv.rs:11-41), with HIP kernels.  The region
 * planning it needs (kernel_radius, build_kernel, input_region_for_output,
 * split_region) is exported as well, so every caller -- a Rust ImageSource
 * node, the C++ host library, Python -- shares one definition.
 *
 * Conventions
 *   - Plain C types only; no unwinding crosses the boundary.
 *   - Status codes: 0 = ITKG_OK, negative = error; itkg_last_error() returns
 *     a thread-local message for the last failing call of this thread.
 *   - The caller owns every buffer; the library never frees caller memory.
 *     Pixel buffers passed as itkg_view are device memory unless the entry
 *     point's name ends in _host.
 *   - Pixels: dtype (f32 or f64) times ncomp components, components
 *     interleaved (VecPixel<T, N> is [T; N], array of structures).
 *   - Images are axis-0-fastest (src/image.rs:111-120).  A view's stride[d]
 *     is the distance in PIXELS between neighbours along axis d.
 *   - Every entry point is reentrant; calls on different streams may run
 *     concurrently.  Work is enqueued on `stream` (a hipStream_t; NULL = the
 *     default stream) and is complete when that stream is synchronised.
 */
#ifndef ITKG_H
#define ITKG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ITKG_MAX_DIM 4
/* Largest per-axis radius the kernels accept: 2r+1 taps travel in the
 * kernel-argument block. */
#define ITKG_MAX_RADIUS 127

typedef enum {
    ITKG_OK = 0,
    ITKG_ERR_INVALID_ARG = -1,
    ITKG_ERR_OOM = -2,
    ITKG_ERR_HIP = -3,
    ITKG_ERR_UNSUPPORTED = -4,
    ITKG_ERR_NO_DEVICE = -5
} itkg_status;

typedef enum { ITKG_F32 = 0, ITKG_F64 = 1 } itkg_dtype;

/* Element types of the pixel conversions (itkg_convert_pixels). */
typedef enum {
    ITKG_PT_U8 = 0,
    ITKG_PT_U16 = 1,
    ITKG_PT_I16 = 2,
    ITKG_PT_I32 = 3,
    ITKG_PT_F32 = 4,
    ITKG_PT_F64 = 5
} itkg_ptype;

/* ITKG_CONVERT_AS: Rust's `as` (integer -> integer truncates or extends;
 * to float rounds to nearest even; float -> integer truncates toward zero,
 * saturates, NaN -> 0).  ITKG_CONVERT_ROUND: a float is first rounded half
 * away from zero (f32::round), then `as`. */
typedef enum { ITKG_CONVERT_AS = 0, ITKG_CONVERT_ROUND = 1 } itkg_convert_mode;

/* Arithmetic mode.
 *   ITKG_EXACT: the reference's arithmetic -- every tap fl(x * (T)w_k) added
 *     to an accumulator that starts at +0, taps k = 0..2r in order, no FMA,
 *     each pass rounded to T.  Bit-identical to itk-rs.
 *   ITKG_FAST: the same weights, truncation, clamp and pass order; each tap
 *     is a fused multiply-add.  Within 1e-5 (max abs error / max abs value)
 *     of ITKG_EXACT for f32; bit-identical between full-image and streamed
 *     runs. */
typedef enum { ITKG_EXACT = 0, ITKG_FAST = 1 } itkg_arith;

/* Kernel selection (testing and tuning).  AUTO picks the fused one-round-trip
 * kernel where it applies (D = 3 or D = 2, radii within its template set) and
 * the per-axis kernel otherwise; both give bit-identical results. */
typedef enum { ITKG_PATH_AUTO = 0, ITKG_PATH_PER_AXIS = 1, ITKG_PATH_FUSED = 2 } itkg_path;

/* Region<D> (src/image.rs:9-15): index may be negative; size >= 0.
 * Entries at d >= ndim are ignored. */
typedef struct {
    int64_t index[ITKG_MAX_DIM];
    int64_t size[ITKG_MAX_DIM];
} itkg_region;

/* Strided window onto the pixels of `region`: `data` points at the pixel
 * whose coordinates are region.index. */
typedef struct {
    void* data;
    itkg_region region;
    int64_t stride[ITKG_MAX_DIM];
} itkg_view;

/* ---- region planning (host only, no device work) ----------------------- */

/* GaussianFilter::kernel_radius, src/filters/gaussian.rs:24-27:
 * ceil(3 * (sigma / spacing)) cast with Rust's saturating float->usize rule
 * (NaN and negatives give 0). */
uint64_t itkg_kernel_radius(double sigma, double spacing);

/* GaussianFilter::build_kernel, src/filters/gaussian.rs:29-40: writes the
 * 2*radius+1 normalised f64 weights to `out` (capacity `cap`). */
int itkg_build_kernel(double sigma_px, uint64_t radius, double* out, uint64_t cap);

/* GaussianFilter::input_region_for_output, src/filters/gaussian.rs:52-57:
 * output.padded_per_axis(radii).clipped_to(largest). */
int itkg_input_region_for_output(int32_t ndim, double sigma, const double* spacing,
                                 const itkg_region* largest, const itkg_region* output,
                                 itkg_region* input);

/* split_region, src/region_splitter.rs:8-39: slowest-axis slabs, the first
 * (size mod n) pieces one row larger.  Writes min(n, size[ndim-1]) pieces
 * (0 for an empty region) to `pieces` (capacity `cap`), count to `*count`. */
int itkg_split_region(int32_t ndim, const itkg_region* region, uint64_t num_pieces,
                      itkg_region* pieces, uint64_t cap, uint64_t* count);

/* ---- forward ----------------------------------------------------------- */

/* GaussianFilter::generate_region, src/filters/gaussian.rs:59-90, with each
 * pass the convolve_axis of src/filters/conv.rs:11-41.
 *
 *   largest : the source's largest_region() (the clamp bounds)
 *   input   : device view holding at least input_region_for_output(
 *             output->region); it may be larger (e.g. the whole image)
 *   output  : device view whose region is the requested region; it must lie
 *             inside `largest`.  Output and input must not overlap.
 *   sigma   : physical sigma; spacing[d] per axis (sigma_px = sigma/spacing)
 *
 * Streaming (src/source.rs:81-97) is this call once per piece with the
 * piece as output region; every voxel is computed by the same taps as in the
 * full-image call, so streamed and full outputs are bit-identical. */
int itkg_gaussian_generate_region(int32_t ndim, itkg_dtype dtype, int32_t ncomp, double sigma,
                                  const double* spacing, const itkg_region* largest,
                                  const itkg_view* input, const itkg_view* output,
                                  itkg_arith arith, void* stream);

/* Several requests of one GaussianFilter over one input view: output k
 * receives generate_region(outputs[k].region), exactly as n calls of
 * itkg_gaussian_generate_region with the same input would give it (input
 * must cover every request's input_region_for_output).  Where the fused
 * kernel serves them all (3-D, f32 scalar pixels, isotropic radii) the
 * pieces run as one persistent launch over all their tiles; otherwise one
 * call per piece.  For stream_full's pieces over a resident volume (each
 * piece still reads its own halo).  No counterpart in the reference, whose
 * stream_full calls generate_region piece by piece. */
int itkg_gaussian_generate_regions(int32_t ndim, itkg_dtype dtype, int32_t ncomp, double sigma,
                                   const double* spacing, const itkg_region* largest,
                                   const itkg_view* input, int32_t n, const itkg_view* outputs,
                                   itkg_arith arith, void* stream);

/* No counterpart in the reference: itkg_gaussian_generate_region for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results, bit for bit; it calls no hipMalloc, hipFree or allocating library
 * routine.  The fused kernel needs no scratch; the per-axis path takes its two
 * intermediates from a static device array of 64 MiB per GPU (held until the
 * process ends), and a call whose intermediates would exceed it fails with
 * ITKG_ERR_UNSUPPORTED.  All calls on one GPU share that array: calls that
 * overlap in time on one GPU race (serialise them, e.g. on one stream). */
int itkg_gaussian_generate_region_noalloc(int32_t ndim, itkg_dtype dtype, int32_t ncomp,
                                          double sigma, const double* spacing,
                                          const itkg_region* largest, const itkg_view* input,
                                          const itkg_view* output, itkg_arith arith,
                                          void* stream);

/* Same contract on dense HOST buffers (what an ImageSource whose Image owns
 * a host Vec<P> hands over): `input` holds input_region_for_output(requested)
 * densely, `output` receives `requested` densely.  Returns when the output is
 * written.  Requests over 32 MiB run as pieces along the slowest axis, as
 * stream_full does (src/source.rs:81-97): two helper threads copy input
 * chunks in and finished pieces out while the GPU computes, so the inbound
 * and outbound transfers overlap; the output is bit-identical to one
 * piece. */
int itkg_gaussian_generate_region_host(int32_t ndim, itkg_dtype dtype, int32_t ncomp,
                                       double sigma, const double* spacing,
                                       const itkg_region* largest, const void* input,
                                       const itkg_region* requested, void* output,
                                       itkg_arith arith);

/* ---- the other convolve_axis filters ----------------------------------- */

/* box_kernel, src/filters/mean.rs:15-18: 2*radius+1 weights of 1/(2r+1)
 * (f64) written to `out` (capacity `cap`). */
int itkg_box_kernel(uint64_t radius, double* out, uint64_t cap);

/* MeanFilter / BoxMeanFilterN::input_region_for_output,
 * src/filters/mean.rs:49-57 and 122-124 (and BinomialBlurFilter's,
 * src/filters/binomial_blur.rs:46-55, with radii = repetitions):
 * output.padded_per_axis(radii).clipped_to(largest). */
int itkg_separable_input_region_for_output(int32_t ndim, const uint64_t* radii,
                                           const itkg_region* largest, const itkg_region* output,
                                           itkg_region* input);

/* generate_region of a separable clamp-to-edge filter with caller-given
 * weights: MeanFilter::generate_region (src/filters/mean.rs:59-83) with
 * radii[d] = radius and box_kernel weights on every axis,
 * BoxMeanFilterN::generate_region (mean.rs:126-144) with per-axis radii, and
 * one repetition of BinomialBlurFilter::generate_region
 * (src/filters/binomial_blur.rs:57-89: radius 1, weights 1/4, 1/2, 1/4, and
 * the requested region of that repetition).  Each pass is the convolve_axis
 * of src/filters/conv.rs:11-41 and the pass regions are those of
 * itkg_gaussian_generate_region.
 *
 *   radii   : radius per axis (at most ITKG_MAX_RADIUS)
 *   weights : axis 0's 2*radii[0]+1 f64 weights, then axis 1's, ...; each
 *             axis symmetric (w_k == w_{2r-k}), else ITKG_ERR_INVALID_ARG;
 *             cast to the pixel type at each use, as the reference does
 *   input, output, largest, arith, stream: as itkg_gaussian_generate_region */
int itkg_separable_generate_region(int32_t ndim, itkg_dtype dtype, int32_t ncomp,
                                   const uint64_t* radii, const double* weights,
                                   const itkg_region* largest, const itkg_view* input,
                                   const itkg_view* output, itkg_arith arith, void* stream);

/* The same on dense host buffers, as itkg_gaussian_generate_region_host. */
int itkg_separable_generate_region_host(int32_t ndim, itkg_dtype dtype, int32_t ncomp,
                                        const uint64_t* radii, const double* weights,
                                        const itkg_region* largest, const void* input,
                                        const itkg_region* requested, void* output,
                                        itkg_arith arith);

/* ---- Gaussian consumers with an index-space sigma ---------------------- */

/* The Gaussian kernel UnsharpMaskFilter and CannyEdgeDetectionFilter build
 * (src/filters/edges.rs:414-423, 563-572): radius ceil(3 sigma) (sigma in
 * pixels, written to *radius), weights exp(-(d*d) / (2 sigma sigma)) over
 * d = -radius..radius divided by their sequential sum, f64, to `out`
 * (capacity `cap`). */
int itkg_index_gaussian_kernel(double sigma, double* out, uint64_t cap, uint64_t* radius);

/* UnsharpMaskFilter::generate_region, src/filters/edges.rs:406-452, for
 * scalar f32 or f64 pixels (the reference's to_f64 of a vector pixel keeps
 * only its first component): the blur of the request with the kernel above
 * on every axis, then out = orig + amount * (orig - blur) where
 * |orig - blur| > threshold, else orig, in f64, narrowed to the pixel type.
 * input covers output->region padded by the radius, clipped to `largest`.
 * arith selects the blur's arithmetic (EXACT: bit-identical).  Output and
 * input must not overlap. */
int itkg_unsharp_mask_generate_region(int32_t ndim, itkg_dtype dtype, double sigma, double amount,
                                      double threshold, const itkg_region* largest,
                                      const itkg_view* input, const itkg_view* output,
                                      itkg_arith arith, void* stream);

/* The same on dense host buffers (input: the padded, clipped request). */
int itkg_unsharp_mask_generate_region_host(int32_t ndim, itkg_dtype dtype, double sigma,
                                           double amount, double threshold,
                                           const itkg_region* largest, const void* input,
                                           const itkg_region* requested, void* output,
                                           itkg_arith arith);

/* CannyEdgeDetectionFilter::generate_region, src/filters/edges.rs:548-664,
 * 2-D scalar f32 / f64 (other ndim: ITKG_ERR_UNSUPPORTED; the reference
 * asserts D = 2): the whole image smoothed with the kernel above (arith
 * selects its arithmetic; EXACT is the reference's), f64 gradient by clamped
 * central differences, magnitude and atan2 direction, non-maximum
 * suppression, hysteresis between lower and upper; the request as 1 / 0.
 * input covers `largest` (the filter's input_region_for_output).  The
 * hysteresis repeats until a pass adds nothing, waiting on the stream after
 * each pass, so the call synchronises the stream, returns when the output
 * is written, and cannot be captured into a graph.  The direction bins are
 * those of a correctly rounded atan2 (itkg_canny_direction_bins): they equal
 * the reference's except where the reference's libm atan2 is off by more
 * than half an ulp for an angle within an ulp of 22.5 + 45 k degrees. */
int itkg_canny_generate_region(int32_t ndim, itkg_dtype dtype, double sigma, double lower,
                               double upper, const itkg_region* largest, const itkg_view* input,
                               const itkg_view* output, itkg_arith arith, void* stream);

/* The same on dense host buffers (input: the whole largest region; output:
 * the request). */
int itkg_canny_generate_region_host(int32_t ndim, itkg_dtype dtype, double sigma, double lower,
                                    double upper, const itkg_region* largest, const void* input,
                                    const itkg_region* requested, void* output, itkg_arith arith);

/* ---- recursive (IIR) Gaussian ----------------------------------------- */

/* yvv_coeffs, src/filters/recursive_gaussian.rs:33-47: the Young - van Vliet
 * third-order coefficients (B, a1, a2, a3) for sigma in pixels (clamped below
 * at 0.5), written to coef[0..3]. */
int itkg_yvv_coefficients(double sigma_px, double* coef);

/* RecursiveGaussianFilter::generate_region (axis = -1, every axis in order;
 * src/filters/recursive_gaussian.rs:147-176) and
 * RecursiveGaussianAxisFilter::generate_region (axis = 0..ndim-1;
 * recursive_gaussian.rs:219-240), scalar f32 / f64 pixels: the whole largest
 * region widened to f64; per filtered axis (of more than one sample), over
 * every line along it, a causal then an anticausal pass (recursive_gaussian.rs:
 * 55-86) with sigma / spacing[axis]; the requested region narrowed back.
 * input covers `largest` (the filters' input_region_for_output); output's
 * region is the request.  Bit-identical to the reference (f64 throughout). */
int itkg_recursive_gaussian_generate_region(int32_t ndim, itkg_dtype dtype, double sigma,
                                            const double* spacing, int32_t axis,
                                            const itkg_region* largest, const itkg_view* input,
                                            const itkg_view* output, void* stream);

/* The same on dense host buffers (input: the whole largest region). */
int itkg_recursive_gaussian_generate_region_host(int32_t ndim, itkg_dtype dtype, double sigma,
                                                 const double* spacing, int32_t axis,
                                                 const itkg_region* largest, const void* input,
                                                 const itkg_region* requested, void* output);

/* ---- adjoint (no counterpart in the reference) ------------------------- */

/* Transpose of itkg_gaussian_generate_region for the same (largest, sigma,
 * spacing, requested = grad_output->region): the passes are transposed in
 * reverse axis order; each clamp folds the clamped taps' weights onto the
 * edge voxel.  grad_input->region must contain input_region_for_output(
 * requested); only that subregion is written.  With accumulate != 0 the
 * result is added to grad_input (streamed pieces whose halos overlap). */
int itkg_gaussian_backward(int32_t ndim, itkg_dtype dtype, int32_t ncomp, double sigma,
                           const double* spacing, const itkg_region* largest,
                           const itkg_view* grad_output, const itkg_view* grad_input,
                           int32_t accumulate, void* stream);

/* The direction bin CannyEdgeDetectionFilter's non-maximum suppression
 * uses for each of n gradients (src/filters/edges.rs:602, 612-620: atan2 in
 * degrees modulo 180; 0 below 22.5 or from 157.5, 1 below 67.5, 2 below
 * 112.5, else 3), from device arrays gx, gy (f64) into device bins, exactly
 * as the Canny kernels bin them: next to a bin edge the bin is the one a
 * correctly rounded atan2 gives, independent of the device's atan2. */
int itkg_canny_direction_bins(const double* gx, const double* gy, int64_t n, uint8_t* bins,
                              void* stream);

/* itkg_gaussian_backward on dense host buffers (axis 0 fastest, ncomp
 * interleaved): grad_output covers `requested`, grad_input covers
 * *grad_input_region, which must contain input_region_for_output(
 * requested).  Only that subregion changes; with accumulate != 0 the result
 * is added to grad_input's contents.  For a host-Vec pipeline node's
 * backward; no counterpart in the reference. */
int itkg_gaussian_backward_host(int32_t ndim, itkg_dtype dtype, int32_t ncomp, double sigma,
                                const double* spacing, const itkg_region* largest,
                                const void* grad_output, const itkg_region* requested,
                                void* grad_input, const itkg_region* grad_input_region,
                                int32_t accumulate);

/* ---- batched many-volume form (no counterpart in the reference) -------- */

/* One whole 3-D f32 scalar volume: output = GaussianFilter{source: input,
 * sigma}.generate_region(largest_region) with per-axis spacing.  Both
 * buffers are dense device memory of size[0]*size[1]*size[2] floats.  For
 * the adjoint direction `input` is the gradient w.r.t. the output and
 * `output` receives the gradient w.r.t. the input. */
typedef struct {
    const float* input;
    float* output;
    int64_t size[3];
    double sigma;
    double spacing[3];
} itkg_volume;

typedef struct itkg_batch itkg_batch;

typedef enum { ITKG_BATCH_FORWARD = 0, ITKG_BATCH_BACKWARD = 1 } itkg_batch_direction;

/* Plans the launches over all volumes: per-volume radii and weights, packed
 * tile work lists, tile shapes from the autotuner's cache (or defaults).
 * Allocates the plan's device tables.  The volumes' buffers are bound at
 * creation; create a new plan for new buffers. */
int itkg_batch_create(int32_t n, const itkg_volume* volumes, itkg_batch_direction direction,
                      itkg_arith arith, itkg_batch** plan);
/* Runs the plan on `stream`.  The first run on a stream captures the launch
 * sequence into a HIP graph; later runs on that stream replay it.  Volumes
 * are grouped by radius; each group is one persistent launch over a packed
 * (volume, tile) work list. */
int itkg_batch_run(itkg_batch* plan, void* stream);
/* Autotuner: times the z-run length of a tile for every radius group of the
 * plan on `stream` and keeps the fastest (the plan's outputs are written
 * with valid results while it runs).  With the environment variable
 * ITKG_TUNE_CACHE naming a file, itkg_batch_create reads earlier choices
 * from it and this call appends its own (key: direction, arithmetic,
 * radius).  Captured graphs are dropped and re-captured on the next run. */
int itkg_batch_autotune(itkg_batch* plan, void* stream);
/* Voxel count covered by the plan (sum of volume sizes). */
int64_t itkg_batch_voxels(const itkg_batch* plan);
int itkg_batch_destroy(itkg_batch* plan);

/* ---- control ----------------------------------------------------------- */

/* Message of the last failing call on this thread ("" if none). */
const char* itkg_last_error(void);
/* Process-wide kernel selection (ITKG_PATH_AUTO by default). */
int itkg_set_path(itkg_path path);
/* ---- pixel conversions (the UnaryFilter casts around a Gaussian) ------- */

/* UnaryFilter::generate_region, src/filters/mod.rs:76-80, for the cast
 * closures of examples/blur_png.rs:35-37 (`|p: u8| p as f32` is
 * ITKG_CONVERT_AS to f32; `|p: f32| p.round().clamp(0.0, 255.0) as u8` is
 * ITKG_CONVERT_ROUND to u8), on the device: every pixel of output->region
 * (components interleaved, ncomp 1..4) is the conversion of the input pixel
 * at the same index.  input->region must contain output->region; device
 * views; strides in pixels.  Keeps a cast -> Gaussian -> cast chain
 * resident on the GPU. */
int itkg_convert_pixels(int32_t ndim, itkg_ptype in_type, itkg_ptype out_type, int32_t ncomp,
                        const itkg_view* input, const itkg_view* output, itkg_convert_mode mode,
                        void* stream);

/* The chain UnaryFilter(cast to f32) -> GaussianFilter -> UnaryFilter(cast)
 * on dense HOST buffers in one GPU round trip, so PCIe moves the pixels in
 * their own type (1 byte each for u8) instead of f32: `input` (in_type)
 * holds input_region_for_output(requested) of the f32 Gaussian, converted
 * with ITKG_CONVERT_AS; `output` (out_type) receives `requested`, converted
 * with out_mode.  Scalar pixels; arith as itkg_gaussian_generate_region.
 * Returns when the output is written. */
int itkg_cast_gaussian_cast_host(int32_t ndim, itkg_ptype in_type, itkg_ptype out_type,
                                 double sigma, const double* spacing, const itkg_region* largest,
                                 const void* input, const itkg_region* requested, void* output,
                                 itkg_arith arith, itkg_convert_mode out_mode);

/* Scratch memory of the entry points (per-axis intermediates, staging,
 * work images) comes from a stream-ordered pool per device that keeps up to
 * 8 GiB of freed memory between calls (environment ITKG_POOL_KEEP_BYTES sets
 * another limit).  This hands the current device's idle pool memory beyond
 * keep_bytes back to the driver, e.g. before allocating through another
 * allocator (torch's cannot see the pool), and frees the pinned host staging
 * ring of the _host entry points (72 MiB) unless a call is using it; the
 * next _host call creates it again.  No counterpart in the reference, which
 * allocates host Vecs per call. */
int itkg_trim_scratch(uint64_t keep_bytes);

/* Host staging of the _host entry points: requests above 32 MiB are split
 * into pieces along the slowest axis, and the input and output move in
 * 12 MiB chunks, overlapped with each other and with the pieces' kernels,
 * through a process-wide ring of pinned host slots (3 per direction) that
 * 6 copy threads per direction (at most half the host's cores) fill and
 * empty (ITKG_STAGE_CHUNK_MIB,
 * ITKG_STAGE_SLOTS and ITKG_STAGE_THREADS set others).  One call at a time uses
 * the ring; a call that finds it in use, or any call when the environment
 * sets ITKG_HOST_STAGING=pageable, copies straight from and to the caller's
 * pageable memory instead.  Results are the same either way.  This reports
 * whether the last such call went through the ring (1) or not (0).  No
 * counterpart in the reference (host Vecs, no device). */
int itkg_host_staging_info(int32_t* last_pinned);

/* Library version string. */
const char* itkg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* ITKG_H */
