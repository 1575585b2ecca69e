// Internal declarations shared by the C-ABI host code and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "itkg.h"

namespace itkg {

constexpr int kMaxTaps = 2 * ITKG_MAX_RADIUS + 1;

// Thread-local error slot behind itkg_last_error().
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
void clear_error();
// The calling thread's current error text.
std::string last_error_text();

// Arguments of one clamp-to-edge pass along `axis` (the convolve_axis of
// src/filters/conv.rs:11-41) from a strided source view to a strided
// destination view.  Coordinates are absolute image indices.
template <typename T>
struct AxisPassArgs {
    const T* src;          // pixel at src_lo
    T* dst;                // pixel at dst_lo
    int64_t n[ITKG_MAX_DIM];         // pass output extents (1 beyond ndim)
    int64_t out_lo[ITKG_MAX_DIM];    // pass output region index
    int64_t src_lo[ITKG_MAX_DIM];
    int64_t src_stride[ITKG_MAX_DIM];
    int64_t dst_lo[ITKG_MAX_DIM];
    int64_t dst_stride[ITKG_MAX_DIM];
    int64_t clamp_lo, clamp_hi;      // pass input extent on `axis`
    int64_t g_lo, g_hi;              // adjoint only: incoming-gradient extent on `axis`
    int64_t total;                   // product of n
    int32_t axis, radius;
    int32_t accumulate;              // adjoint only: dst += result
    T w[kMaxTaps];                   // (T)w_k, k = 0..2r
    T tail[ITKG_MAX_RADIUS + 1];     // adjoint only: (T)(sum_{k<m} w_k), m = 0..r
};

template <typename T>
hipError_t launch_axis_forward(const AxisPassArgs<T>& a, int ncomp, itkg_arith arith,
                               hipStream_t s);
// Adjoint pass; [g_lo, g_hi] is the gradient-of-output extent on the axis.
template <typename T>
hipError_t launch_axis_adjoint_range(const AxisPassArgs<T>& a, int ncomp, int64_t g_lo,
                                     int64_t g_hi, hipStream_t s);

// UnsharpMaskFilter's per-pixel step over region size n (axis 0 fastest;
// strides in pixels): out = orig + amount * (orig - blur) where
// |orig - blur| > threshold, else orig, in f64.
template <typename T>
struct UnsharpArgs {
    const T* in;
    const T* blur;
    T* out;
    int64_t n[4];
    int64_t in_stride[4], blur_stride[4], out_stride[4];
    double amount, threshold;
};
template <typename T>
hipError_t launch_unsharp(const UnsharpArgs<T>& a, hipStream_t s);

// Recursive Gaussian: `in` (strides in pixels) covers the whole region of
// sizes n; work is its dense f64 copy; out (strides) receives the sub-region
// of sizes m at offset off (pixels from the whole region's first corner).
template <typename T>
struct RecursiveArgs {
    const T* in;
    double* work;
    T* out;
    double* ckpt;  // causal state at each chunk start (recursive_ckpt_doubles)
    int64_t n[4], m[4], off[4];
    int64_t in_stride[4], out_stride[4];
};
// Line chunks of the recursive passes: rows (axis 0) and other axes.
constexpr int kRecRowChunk = 32;
constexpr int kRecLineChunk = 32;
// Doubles of the checkpoint table of one filtered axis: 3 per line per chunk
// after the first.
inline int64_t recursive_ckpt_doubles(int64_t total, int64_t len, int axis) {
    const int64_t K = axis == 0 ? kRecRowChunk : kRecLineChunk;
    const int64_t nch = (len + K - 1) / K;
    return nch > 1 ? total / len * (nch - 1) * 3 : 0;
}
// Widen, run each listed axis's causal + anticausal lines with its
// coefficients (B, a1, a2, a3), narrow the sub-region.
template <typename T>
hipError_t launch_recursive(const RecursiveArgs<T>& a, const int* axes, int naxes,
                            const double (*coef)[4], hipStream_t s);

// Canny after smoothing: `smoothed` is the dense nx x ny smoothed image;
// weak and edges are nx x ny byte maps of scratch and `changed` two device
// ints (the flag of a hysteresis batch's earlier passes, which nobody
// reads; the last pass of a batch raises a pinned host flag instead); out
// (strides in pixels) receives the mx x my request at offset (ox, oy).  The
// call synchronises the stream after each batch of hysteresis passes.
template <typename T>
struct CannyArgs {
    const T* smoothed;
    uint8_t* weak;   // nms >= lower (suppressed pixels as 0.0)
    uint8_t* edges;  // nms >= upper, then grown by hysteresis
    int* changed;
    T* out;
    int64_t nx, ny, ox, oy, mx, my;
    int64_t out_stride[2];
    double lower, upper;
};
template <typename T>
hipError_t launch_canny(const CannyArgs<T>& a, hipStream_t s);
// The direction bins of n gradients (device arrays), as the Canny kernels bin them.
hipError_t launch_canny_bins(const double* gx, const double* gy, uint8_t* bins, int64_t n,
                             hipStream_t s);

// Pixel conversion over a region of sizes n (1 beyond ndim); strides in
// pixels, ncomp interleaved components.
struct ConvertArgs {
    const void* in;
    void* out;
    int64_t n[4];
    int64_t in_stride[4], out_stride[4];
    int32_t ncomp;
    itkg_ptype in_type, out_type;
    itkg_convert_mode mode;
};
hipError_t launch_convert(const ConvertArgs& a, hipStream_t s);
size_t ptype_size(itkg_ptype t);

// Fused one-round-trip 3-D kernel for f32 scalar pixels (forward or adjoint).
// Destination pixel j reads source pixels j + shift + k - r, k = 0..2r
// (forward: clamped into the source; adjoint: zero outside it).
// The _S modes are FAST and adjoint on 32-row tiles (the others use 64-row
// tiles), for volumes small enough that shorter tiles buy longer z runs.
// FUSED_FWD_FAST_W and FUSED_FWD_EXACT_W are FAST and EXACT on 128 x 32
// tiles (plain launches, small radii): longer contiguous source rows, the
// same arithmetic per output, so the same bits as the other shapes.
enum FusedMode {
    FUSED_FWD_EXACT = 0,
    FUSED_FWD_FAST = 1,
    FUSED_ADJ = 2,
    FUSED_FWD_FAST_S = 3,
    FUSED_ADJ_S = 4,
    FUSED_FWD_FAST_W = 5,
    FUSED_FWD_EXACT_W = 6,
    // UnsharpMaskFilter's per-pixel step on the forward result, in the same
    // launch (64-row tiles, EXACT or FAST blur): with o the source pixel at
    // the output pixel's index and b its blurred value, both widened to f64,
    // the output is o + amount * (o - b) where |o - b| > threshold, else o.
    FUSED_FWD_EXACT_U = 7,
    FUSED_FWD_FAST_U = 8,
    FUSED_NMODES = 9
};
constexpr bool fused_is_adjoint(int mode) { return mode == FUSED_ADJ || mode == FUSED_ADJ_S; }
// EXACT arithmetic: every product and every sum rounded separately.
constexpr bool fused_is_exact(int mode) {
    return mode == FUSED_FWD_EXACT || mode == FUSED_FWD_EXACT_W || mode == FUSED_FWD_EXACT_U;
}
constexpr bool fused_is_unsharp(int mode) { return mode == FUSED_FWD_EXACT_U || mode == FUSED_FWD_FAST_U; }
// The plain forward mode whose tiles, registers and arithmetic a mode shares.
constexpr int fused_base_mode(int mode) {
    return mode == FUSED_FWD_EXACT_U ? FUSED_FWD_EXACT : (mode == FUSED_FWD_FAST_U ? FUSED_FWD_FAST : mode);
}
// Destination tile width (x) of a mode.
constexpr int fused_tile_x(int mode) {
    return mode == FUSED_FWD_FAST_W || mode == FUSED_FWD_EXACT_W ? 128 : 64;
}
constexpr int kFusedMaxR = 16;

template <int MAXR>
struct Fused3dArgsT {
    const float* src;        // source view pixel at the source region index
    float* dst;              // destination view pixel at the destination region index
    int64_t src_stride[3];
    int64_t dst_stride[3];
    int32_t n_src[3];        // source extents (forward: the clamp bounds)
    int32_t n_dst[3];        // destination extents
    int32_t shift[3];        // forward: requested.index - input.index; adjoint: its negation
    int32_t tz;              // destination planes per workgroup
    int32_t tiles_x, tiles_y, tiles_z;
    int32_t accumulate;      // adjoint: dst += result
    float wx[2 * MAXR + 1], wy[2 * MAXR + 1], wz[2 * MAXR + 1];
    float wz2[2 * (2 * MAXR + 1)];  // wz repeated twice: rotated-weight lookups
    float px[MAXR + 1], py[MAXR + 1], pz[MAXR + 1];  // adjoint edge sums
};
// Arguments of a plain launch (they travel in the kernel-argument segment,
// four of them per launch, so their size is bounded) ...
using Fused3dArgs = Fused3dArgsT<kFusedMaxR>;
// ... and of the persistent batched kernel (a device table, radii up to 24).
constexpr int kFusedMaxRBatch = 24;
using Fused3dArgsB = Fused3dArgsT<kFusedMaxRBatch>;

// One compiled instantiation of the fused kernel: x/y radius rxy, z radius
// rz (0 for 2-D images), its y tile height, and its launchers.  `batch` and
// `occupancy` are null for instantiations without a persistent form.
// One plain fused launch over the components of an interleaved pixel: each
// component is a scalar image with strides scaled by the component count.
constexpr int kFusedMaxComp = 4;
struct FusedMulti {
    Fused3dArgs a[kFusedMaxComp];
    int32_t nc;
    double amount, threshold;  // the unsharp modes' step
};

struct FusedEntry {
    int rxy, rz;
    int ty[FUSED_NMODES];  // tile rows per mode (FusedMode index; 0: mode not built)
    hipError_t (*launch)(const FusedMulti& m, int mode, hipStream_t s);
    hipError_t (*batch)(const Fused3dArgs* vols, const int2* work, int nwork, int grid_blocks,
                        int mode, hipStream_t s);
    int (*occupancy)(int mode);  // resident workgroups per CU of the persistent form
};
constexpr int kFusedTileX = 64;  // tile width of every mode but the _W ones
// The instantiation for (rxy, rz), or null.
const FusedEntry* fused_entry(int rxy, int rz);

// The single persistent launch of a batched plan (fused3d_mega_kernel):
// `fast` serves forward FAST, `adj` the adjoint.  Work items are int4
// (volume, tile, class, unused), class = radius r on the family's own tile
// shape or kMegaClassSmall + r (r <= 7) on 32-row tiles; qstart[0..8] bounds
// the 8 per-XCD queues in `work`; heads[0..8] are the queue heads and the
// finish count, zero before the first launch and reset by each launch.
constexpr int kMegaClassSmall = 32;
constexpr int kMegaClasses = 40;  // class index bound
hipError_t mega_prepare_fast(int lds_bytes);
hipError_t mega_prepare_adj(int lds_bytes);
// trace: null, or 3 ticks-and-workgroup words per work item (breakdowns)
hipError_t mega_launch_fast(const Fused3dArgsB* vols, const int4* work, const int* qstart,
                            int* heads, int grid, int lds_bytes, hipStream_t s,
                            unsigned long long* trace = nullptr);
hipError_t mega_launch_adj(const Fused3dArgsB* vols, const int4* work, const int* qstart,
                           int* heads, int grid, int lds_bytes, hipStream_t s,
                           unsigned long long* trace = nullptr);
int mega_occupancy_fast(int lds_bytes);
int mega_occupancy_adj(int lds_bytes);
// LDS bytes and tile rows of every class (arrays of kMegaClasses; classes
// the family lacks are left untouched).
void mega_classes_fast(int* lds, int* ty);
void mega_classes_adj(int* lds, int* ty);

// Per-call plan of GaussianFilter::generate_region (src/filters/gaussian.rs:
// 59-90): radii, f64 weights and the region of every pass.
struct Plan {
    int D = 0;
    int64_t radii[ITKG_MAX_DIM] = {0, 0, 0, 0};
    itkg_region bounds{}, req{}, in_reg{};
    itkg_region reg[ITKG_MAX_DIM + 1]{};  // reg[a]: input of pass a; reg[D] = req
    double w[ITKG_MAX_DIM][kMaxTaps];
};

int make_plan(int D, double sigma, const double* spacing, const itkg_region* largest,
              const itkg_region& req, Plan* p);
int64_t region_len(int D, const itkg_region& r);
itkg_view dense_view(int D, void* data, const itkg_region& r);
// Device scratch the per-axis path needs (two ping-pong intermediates).
size_t scratch_bytes(const Plan& p, itkg_dtype dt, int ncomp);
// True when the fused kernel serves this plan (and the path setting allows it).
bool use_fused(const Plan& p, itkg_dtype dt, int ncomp);
const FusedEntry* fused_for(const Plan& p);
// z run of a fused tile for a grid of tiles_xy tiles per z run over nz planes.
int fused_tz(int tiles_xy, int nz, int r);
// True when the fused kernel can address the source view of this call (in-
// plane offsets are 32-bit).
bool fused_src_ok(const Plan& p, const itkg_view& src, int ncomp, bool adjoint);
// Every extent fits the fused kernel's 32-bit coordinates.
bool fits_i32(const Plan& p);
// The path setting (itkg_set_path) allows the fused kernel.
bool fused_path_allowed();
template <class A>
void fill_fused_args(const Plan& p, const itkg_view& src_view, const itkg_view& dst_view,
                     int mode, int accumulate, int tz, int comp, int ncomp, A* out);
// Enqueue one forward / adjoint; `scratch` (nullable) holds scratch_bytes().
int forward_impl(const Plan& p, itkg_dtype dt, int ncomp, const itkg_view& in,
                 const itkg_view& out, itkg_arith arith, hipStream_t s, void* scratch);
int backward_impl(const Plan& p, itkg_dtype dt, int ncomp, const itkg_view& gout,
                  const itkg_view& gin, int accumulate, hipStream_t s, void* scratch);
// Several forward requests over one input view (itkg_gaussian_generate_regions).
int multi_forward(const std::vector<Plan>& plans, itkg_dtype dt, int ncomp, const itkg_view& in,
                  const itkg_view* outs, itkg_arith arith, hipStream_t s);

// A piecewise computation between dense host buffers and dense device
// buffers of the same byte layout (host_stage.cpp).  Piece k reads the input
// bytes [0, in_need[k]) and completes the output bytes [0, out_end[k])
// (both non-decreasing in k); compute(k, s) enqueues it on s.
struct StageStats {
    bool pinned = false;  // staged through the pinned ring (else pageable)
};
struct StagedJob {
    const void* host_in = nullptr;
    size_t in_bytes = 0;
    void* host_out = nullptr;
    size_t out_bytes = 0;
    void* dev_in = nullptr;
    const void* dev_out = nullptr;
    int pieces = 0;
    std::vector<size_t> in_need, out_end;
    std::function<int(int, hipStream_t)> compute;
    StageStats* stats = nullptr;
};
// Runs the job with the compute on s_comp (a stream the caller owns) and
// the copies on streams of its own; returns when the output is in host_out.
// One piece of at most one chunk each way runs serially on s_comp.
int run_staged(const StagedJob& job, hipStream_t s_comp);
// Frees the pinned staging ring if no call is using it.
int stage_trim();

}  // namespace itkg
