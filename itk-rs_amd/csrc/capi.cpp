// C ABI of libitkg: region planning, argument checking and kernel dispatch.
//
// The planning restates src/filters/gaussian.rs:24-90 (radius rule, f64
// weights, per-pass regions) and src/image.rs:28-64 (padded_per_axis,
// clipped_to); the kernels do the per-pixel work.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "itkg_impl.hpp"

namespace itkg {

namespace {
thread_local std::string g_err;
std::atomic<int> g_path{ITKG_PATH_AUTO};
}  // namespace

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? ITKG_ERR_OOM : ITKG_ERR_HIP, "%s: %s", what,
                hipGetErrorString(e));
}

void clear_error() { g_err.clear(); }

std::string last_error_text() { return g_err; }

// Scratch memory of the entry points comes from a stream-ordered pool per
// device that the library owns and that keeps up to kPoolKeepBytes of freed
// memory across synchronisations (ITKG_POOL_KEEP_BYTES overrides it).  A
// device's default pool returns freed memory to the driver at every
// synchronisation, and a caller that waits between calls would then pay for
// mapping each call's scratch again.  itkg_trim_scratch hands the idle part
// back (memory other allocators, such as torch's, cannot see otherwise).
constexpr uint64_t kPoolKeepBytes = 8ull << 30;
constexpr int kPoolDevices = 64;

namespace {
std::mutex g_pool_mu;
hipMemPool_t g_pools[kPoolDevices] = {};

uint64_t pool_keep_bytes() {
    const char* v = std::getenv("ITKG_POOL_KEEP_BYTES");
    if (!v || !*v) return kPoolKeepBytes;
    char* end = nullptr;
    const unsigned long long k = std::strtoull(v, &end, 10);
    return (end && *end == '\0') ? (uint64_t)k : kPoolKeepBytes;
}
}  // namespace

hipError_t pool_alloc(void** ptr, size_t bytes, hipStream_t s) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kPoolDevices) return hipErrorInvalidDevice;
    hipMemPool_t pool;
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        if (!g_pools[dev]) {
            hipMemPoolProps props{};
            props.allocType = hipMemAllocationTypePinned;
            props.location.type = hipMemLocationTypeDevice;
            props.location.id = dev;
            hipMemPool_t created = nullptr;
            if ((e = hipMemPoolCreate(&created, &props)) != hipSuccess) return e;
            uint64_t keep = pool_keep_bytes();
            if ((e = hipMemPoolSetAttribute(created, hipMemPoolAttrReleaseThreshold, &keep)) !=
                hipSuccess) {
                (void)hipMemPoolDestroy(created);
                return e;
            }
            g_pools[dev] = created;
        }
        pool = g_pools[dev];
    }
    return hipMallocFromPoolAsync(ptr, bytes, pool, s);
}

// Streams of the _host entry points, reused across calls: creating and
// destroying a stream costs every call, which matters for small images.  A
// free list per device; a stream in use belongs to one call, which has
// synchronised it before handing it back; the streams live until the
// process ends.
namespace {
std::mutex g_hs_mu;
std::vector<hipStream_t> g_hs_free[kPoolDevices];
}  // namespace

hipError_t acquire_host_stream(hipStream_t* s) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kPoolDevices) return hipErrorInvalidDevice;
    {
        std::lock_guard<std::mutex> lk(g_hs_mu);
        if (!g_hs_free[dev].empty()) {
            *s = g_hs_free[dev].back();
            g_hs_free[dev].pop_back();
            return hipSuccess;
        }
    }
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

void release_host_stream(hipStream_t s) {
    int dev = 0;
    if (!s) return;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kPoolDevices) {
        (void)hipStreamDestroy(s);
        return;
    }
    std::lock_guard<std::mutex> lk(g_hs_mu);
    g_hs_free[dev].push_back(s);
}

// Releases the current device's idle pool memory beyond keep_bytes.
int trim_pool(uint64_t keep_bytes) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kPoolDevices) return fail(ITKG_ERR_UNSUPPORTED, "device %d", dev);
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (!g_pools[dev]) return ITKG_OK;
    e = hipMemPoolTrimTo(g_pools[dev], (size_t)keep_bytes);
    return e == hipSuccess ? ITKG_OK : hip_fail(e, "hipMemPoolTrimTo");
}

// ---- region algebra (src/image.rs:17-64) --------------------------------

int64_t region_len(int D, const itkg_region& r) {
    int64_t n = 1;
    for (int d = 0; d < D; ++d) n *= r.size[d];
    return n;
}

itkg_region padded(int D, itkg_region r, const int64_t* radii) {
    for (int d = 0; d < D; ++d) {
        r.index[d] -= radii[d];
        r.size[d] += 2 * radii[d];
    }
    return r;
}

itkg_region clipped(int D, itkg_region r, const itkg_region& b) {
    for (int d = 0; d < D; ++d) {
        const int64_t se = r.index[d] + r.size[d], be = b.index[d] + b.size[d];
        const int64_t s = std::max(r.index[d], b.index[d]);
        const int64_t e = std::min(se, be);
        r.index[d] = s;
        r.size[d] = std::max<int64_t>(e - s, 0);
    }
    return r;
}

bool contains(int D, const itkg_region& outer, const itkg_region& inner) {
    for (int d = 0; d < D; ++d) {
        if (inner.size[d] == 0) continue;
        if (inner.index[d] < outer.index[d]) return false;
        if (inner.index[d] + inner.size[d] > outer.index[d] + outer.size[d]) return false;
    }
    return true;
}

bool region_ok(int D, const itkg_region& r) {
    for (int d = 0; d < D; ++d)
        if (r.size[d] < 0) return false;
    return true;
}

bool same_region(int D, const itkg_region& a, const itkg_region& b) {
    for (int d = 0; d < D; ++d)
        if (a.index[d] != b.index[d] || a.size[d] != b.size[d]) return false;
    return true;
}

// ---- kernel (src/filters/gaussian.rs:24-40) ------------------------------

uint64_t radius_of(double sigma, double spacing) {
    const double sigma_px = sigma / spacing;
    const double v = std::ceil(3.0 * sigma_px);
    if (!(v > 0.0)) return 0;
    if (v >= 18446744073709551616.0) return UINT64_MAX;
    return (uint64_t)v;
}

void weights_of(double sigma_px, uint64_t radius, double* k) {
    const uint64_t len = 2 * radius + 1;
    for (uint64_t i = 0; i < len; ++i) {
        const double x = (double)i - (double)radius;
        const double t = x / sigma_px;
        k[i] = std::exp(-0.5 * (t * t));
    }
    double sum = 0.0;
    for (uint64_t i = 0; i < len; ++i) sum += k[i];
    for (uint64_t i = 0; i < len; ++i) k[i] /= sum;
}

// The Gaussian kernel of UnsharpMaskFilter and CannyEdgeDetectionFilter
// (src/filters/edges.rs:414-423, 563-572): radius ceil(3 sigma) with sigma in
// index space, w_i = exp(-(d * d) / (2 sigma sigma)) for d = i - radius,
// divided by their sequential sum.  (sigma = 0 gives NaN weights, as there.)
uint64_t index_weights(double sigma, double* k, uint64_t cap) {
    const uint64_t radius = radius_of(sigma, 1.0);
    if (radius > (uint64_t)ITKG_MAX_RADIUS || 2 * radius + 1 > cap) return UINT64_MAX;
    const uint64_t len = 2 * radius + 1;
    for (uint64_t i = 0; i < len; ++i) {
        const double d = (double)i - (double)radius;
        k[i] = std::exp(-(d * d) / (2.0 * sigma * sigma));
    }
    double sum = 0.0;
    for (uint64_t i = 0; i < len; ++i) sum += k[i];
    for (uint64_t i = 0; i < len; ++i) k[i] /= sum;
    return radius;
}

// ---- per-call plan (src/filters/gaussian.rs:59-90) ----------------------

// The regions of a plan whose radii are set: the input region (the request
// padded by the radii, clipped to the largest region: gaussian.rs:52-57,
// mean.rs:49-57) and each pass's output region (gaussian.rs:79-85,
// mean.rs:71-80).
int plan_regions(int D, const itkg_region* largest, const itkg_region& req, Plan* p) {
    if (region_len(D, req) > 0 && !contains(D, *largest, req))
        return fail(ITKG_ERR_INVALID_ARG, "requested region is not inside the largest region");
    p->req = req;
    p->in_reg = clipped(D, padded(D, req, p->radii), *largest);
    p->reg[0] = p->in_reg;
    for (int a = 0; a < D; ++a) {
        if (a + 1 < D) {
            int64_t rem[ITKG_MAX_DIM] = {0, 0, 0, 0};
            for (int d = a + 1; d < D; ++d) rem[d] = p->radii[d];
            p->reg[a + 1] = clipped(D, padded(D, req, rem), *largest);
        } else {
            p->reg[a + 1] = req;
        }
    }
    return ITKG_OK;
}

int plan_common(int D, const itkg_region* largest, const itkg_region& req, Plan* p) {
    if (D < 1 || D > ITKG_MAX_DIM) return fail(ITKG_ERR_INVALID_ARG, "ndim %d not in [1, %d]", D, ITKG_MAX_DIM);
    if (!largest) return fail(ITKG_ERR_INVALID_ARG, "null largest region");
    if (!region_ok(D, *largest) || !region_ok(D, req))
        return fail(ITKG_ERR_INVALID_ARG, "negative region size");
    p->D = D;
    p->bounds = *largest;
    p->req = req;
    return ITKG_OK;
}

int make_plan(int D, double sigma, const double* spacing, const itkg_region* largest,
              const itkg_region& req, Plan* p) {
    if (!spacing) return fail(ITKG_ERR_INVALID_ARG, "null spacing");
    if (int rc = plan_common(D, largest, req, p)) return rc;
    for (int d = 0; d < D; ++d) {
        const uint64_t r = radius_of(sigma, spacing[d]);
        if (r > ITKG_MAX_RADIUS)
            return fail(ITKG_ERR_UNSUPPORTED, "axis %d radius %llu exceeds %d", d,
                        (unsigned long long)r, ITKG_MAX_RADIUS);
        p->radii[d] = (int64_t)r;
        weights_of(sigma / spacing[d], r, p->w[d]);
    }
    return plan_regions(D, largest, req, p);
}

// A plan from caller-given radii and weights (axis 0's 2 radii[0] + 1
// weights, then axis 1's, ...), as MeanFilter, BoxMeanFilterN and
// BinomialBlurFilter hand convolve_axis their own kernels.  The weights must
// be symmetric (w_k = w_{2r-k}): the adjoint kernels rely on it.
int make_plan_weights(int D, const uint64_t* radii, const double* weights,
                      const itkg_region* largest, const itkg_region& req, Plan* p,
                      bool check_symmetric = true) {
    if (!radii || !weights) return fail(ITKG_ERR_INVALID_ARG, "null radii or weights");
    if (int rc = plan_common(D, largest, req, p)) return rc;
    const double* w = weights;
    for (int d = 0; d < D; ++d) {
        const uint64_t r = radii[d];
        if (r > ITKG_MAX_RADIUS)
            return fail(ITKG_ERR_UNSUPPORTED, "axis %d radius %llu exceeds %d", d,
                        (unsigned long long)r, ITKG_MAX_RADIUS);
        p->radii[d] = (int64_t)r;
        for (uint64_t k = 0; k <= 2 * r; ++k) {
            if (check_symmetric && !(w[k] == w[2 * r - k]))
                return fail(ITKG_ERR_INVALID_ARG, "axis %d weights are not symmetric", d);
            p->w[d][k] = w[k];
        }
        w += 2 * r + 1;
    }
    return plan_regions(D, largest, req, p);
}

size_t dtype_size(itkg_dtype t) { return t == ITKG_F64 ? 8 : 4; }

// Pixel offset of `at` inside view v.
int64_t view_offset(int D, const itkg_view& v, const int64_t* at) {
    int64_t off = 0;
    for (int d = 0; d < D; ++d) off += (at[d] - v.region.index[d]) * v.stride[d];
    return off;
}

itkg_view dense_view(int D, void* data, const itkg_region& r) {
    itkg_view v{};
    v.data = data;
    v.region = r;
    int64_t s = 1;
    for (int d = 0; d < ITKG_MAX_DIM; ++d) {
        v.stride[d] = s;
        if (d < D) s *= r.size[d];
    }
    return v;
}

// Scratch needed by the per-axis path: two ping-pong buffers holding the
// largest intermediate region.
size_t scratch_bytes(const Plan& p, itkg_dtype dt, int ncomp) {
    int64_t m = 0;
    for (int a = 1; a < p.D; ++a) m = std::max(m, region_len(p.D, p.reg[a]));
    return (size_t)(2 * m) * ncomp * dtype_size(dt);
}

template <typename T>
void fill_pass(AxisPassArgs<T>& A, const Plan& p, int axis, const itkg_view& src,
               const itkg_view& dst, const itkg_region& out_reg) {
    const int D = p.D;
    std::memset(&A, 0, sizeof A);
    for (int d = 0; d < ITKG_MAX_DIM; ++d) {
        A.n[d] = d < D ? out_reg.size[d] : 1;
        A.out_lo[d] = d < D ? out_reg.index[d] : 0;
        A.src_lo[d] = d < D ? src.region.index[d] : 0;
        A.src_stride[d] = d < D ? src.stride[d] : 0;
        A.dst_lo[d] = d < D ? dst.region.index[d] : 0;
        A.dst_stride[d] = d < D ? dst.stride[d] : 0;
    }
    A.src = static_cast<const T*>(src.data);
    A.dst = static_cast<T*>(dst.data);
    A.total = region_len(D, out_reg);
    A.axis = axis;
    A.radius = (int32_t)p.radii[axis];
    for (int k = 0; k <= 2 * p.radii[axis]; ++k) A.w[k] = (T)p.w[axis][k];
}

int check_view(int D, const itkg_view* v, const itkg_region& need, const char* what) {
    if (!v) return fail(ITKG_ERR_INVALID_ARG, "%s view is null", what);
    if (!region_ok(D, v->region)) return fail(ITKG_ERR_INVALID_ARG, "%s view: negative size", what);
    if (region_len(D, need) > 0) {
        if (!v->data) return fail(ITKG_ERR_INVALID_ARG, "%s view has no data", what);
        if (!contains(D, v->region, need))
            return fail(ITKG_ERR_INVALID_ARG, "%s view does not cover the region it must hold", what);
    }
    return ITKG_OK;
}

template <typename T>
int forward_per_axis(const Plan& p, int ncomp, const itkg_view& in, const itkg_view& out,
                     itkg_arith arith, hipStream_t s, void* scratch) {
    const int D = p.D;
    void* own = nullptr;
    const size_t sb = scratch_bytes(p, sizeof(T) == 8 ? ITKG_F64 : ITKG_F32, ncomp);
    if (D > 1 && !scratch && sb > 0) {
        hipError_t e = pool_alloc(&own, sb, s);
        if (e != hipSuccess) return hip_fail(e, "pool_alloc(per-axis scratch)");
        scratch = own;
    }
    char* buf[2] = {static_cast<char*>(scratch), static_cast<char*>(scratch) + sb / 2};
    itkg_view cur = in;
    int rc = ITKG_OK;
    for (int a = 0; a < D && rc == ITKG_OK; ++a) {
        const itkg_region& oreg = p.reg[a + 1];
        itkg_view nxt = (a + 1 == D) ? out : dense_view(D, buf[a & 1], oreg);
        AxisPassArgs<T> A;
        fill_pass(A, p, a, cur, nxt, oreg);
        A.clamp_lo = p.reg[a].index[a];
        A.clamp_hi = p.reg[a].index[a] + p.reg[a].size[a] - 1;
        hipError_t e = launch_axis_forward<T>(A, ncomp, arith, s);
        if (e != hipSuccess) rc = hip_fail(e, "per-axis forward launch");
        cur = nxt;
    }
    if (own) (void)hipFreeAsync(own, s);
    return rc;
}

// Compute units of the current device (cached per device).
int device_cus() {
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int v = cache[dev].load();
    if (v > 0) return v;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    cache[dev].store(cus);
    return cus;
}

// z run of a fused tile.  A workgroup walks tz + 2r source planes, plus a
// short prologue.  A CU works through ceil(tiles / CUs) tiles: workgroups
// resident on one CU together share it and finish no sooner than one after
// the other (measured: two 32-row workgroups per CU at 512^3, r = 3, take
// 24% longer than one 64-row workgroup covering the same outputs).  The z
// runs are the equal splits ceil(nz / k); the one with the fewest plane
// steps per CU wins (the smaller tz on ties: shorter tails).
// ITKG_FUSED_TZ overrides.
int fused_tz(int tiles_xy, int nz, int r) {
    if (const char* e = std::getenv("ITKG_FUSED_TZ")) return std::max(1, std::atoi(e));
    nz = std::max(nz, 1);
    const int64_t cus = device_cus();
    int best_tz = nz;
    int64_t best = INT64_MAX;
    for (int k = 1; k <= std::min(nz, 1024); ++k) {
        const int tz = (nz + k - 1) / k;
        if (k > 1 && tz == (nz + k - 2) / (k - 1)) continue;  // same split as k - 1
        const int64_t tiles = (int64_t)tiles_xy * ((nz + tz - 1) / tz);
        const int64_t per_cu = (tiles + cus - 1) / cus;
        const int64_t cost = per_cu * (tz + 2 * r + 3);
        if (cost <= best) {
            best = cost;
            best_tz = tz;
        }
    }
    if (std::getenv("ITKG_DEBUG"))
        std::fprintf(stderr, "itkg: fused r=%d tiles_xy=%d nz=%d -> tz=%d\n", r, tiles_xy, nz, best_tz);
    return best_tz;
}

// The fused instantiation serving this plan (2-D images run as one plane
// with z radius 0), or null.
const FusedEntry* fused_for(const Plan& p) {
    if (p.D != 2 && p.D != 3) return nullptr;
    if (p.radii[0] != p.radii[1]) return nullptr;
    return fused_entry((int)p.radii[0], p.D == 3 ? (int)p.radii[2] : 0);
}

// Arguments of one fused launch for component `comp` of `ncomp` interleaved
// components (a component is a scalar image with strides scaled by ncomp).
// Forward: source = input view at the input region, destination = output
// (requested).  Adjoint: source = grad_output (requested), destination =
// grad_input at the input region.  tz <= 0 picks the z run with fused_tz.
// The caller keeps the radii within the argument block's MAXR.
template <class ARGS>
void fill_fused_args(const Plan& p, const itkg_view& src_view, const itkg_view& dst_view,
                     int mode, int accumulate, int tz, int comp, int ncomp, ARGS* out) {
    const bool adj = fused_is_adjoint(mode);
    const itkg_region& sreg = adj ? p.req : p.in_reg;
    const itkg_region& dreg = adj ? p.in_reg : p.req;
    const int D = p.D;
    ARGS& A = *out;
    std::memset(&A, 0, sizeof A);
    A.src = static_cast<const float*>(src_view.data) + view_offset(D, src_view, sreg.index) * ncomp +
            comp;
    A.dst = static_cast<float*>(dst_view.data) + view_offset(D, dst_view, dreg.index) * ncomp + comp;
    for (int d = 0; d < 3; ++d) {
        if (d < D) {
            A.src_stride[d] = src_view.stride[d] * ncomp;
            A.dst_stride[d] = dst_view.stride[d] * ncomp;
            A.n_src[d] = (int32_t)sreg.size[d];
            A.n_dst[d] = (int32_t)dreg.size[d];
            const int32_t off = (int32_t)(p.req.index[d] - p.in_reg.index[d]);
            A.shift[d] = adj ? -off : off;
        } else {  // a 2-D image is one plane
            A.n_src[d] = A.n_dst[d] = 1;
        }
    }
    const int rxy = (int)p.radii[0];
    const int rz = D == 3 ? (int)p.radii[2] : 0;
    for (int k = 0; k <= 2 * rxy; ++k) {
        A.wx[k] = (float)p.w[0][k];
        A.wy[k] = (float)p.w[1][k];
    }
    for (int k = 0; k <= 2 * rz; ++k) A.wz[k] = D == 3 ? (float)p.w[2][k] : 1.0f;
    for (int k = 0; k < 2 * (2 * rz + 1); ++k) A.wz2[k] = A.wz[k % (2 * rz + 1)];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int m = 0; m <= rxy; ++m) {
        A.px[m] = (float)acc[0];
        A.py[m] = (float)acc[1];
        acc[0] += p.w[0][m];
        acc[1] += p.w[1][m];
    }
    for (int m = 0; m <= rz && D == 3; ++m) {
        A.pz[m] = (float)acc[2];
        acc[2] += p.w[2][m];
    }
    A.accumulate = accumulate;
    const FusedEntry* fe = fused_for(p);
    const int TY = fe ? fe->ty[mode] : 32;
    const int TX = fused_tile_x(mode);
    A.tiles_x = (A.n_dst[0] + TX - 1) / TX;
    A.tiles_y = (A.n_dst[1] + TY - 1) / TY;
    A.tz = tz > 0 ? tz : fused_tz(A.tiles_x * A.tiles_y * ncomp, A.n_dst[2], rz);
    A.tiles_z = (A.n_dst[2] + A.tz - 1) / A.tz;
}
template void fill_fused_args<Fused3dArgs>(const Plan&, const itkg_view&, const itkg_view&, int,
                                           int, int, int, int, Fused3dArgs*);
template void fill_fused_args<Fused3dArgsB>(const Plan&, const itkg_view&, const itkg_view&, int,
                                            int, int, int, int, Fused3dArgsB*);

// Tile shape of a FAST, EXACT or adjoint launch: 64-row tiles (fewer halo
// rows per output) or, FAST and adjoint, 32-row tiles (twice the tiles, so
// longer z runs fill the GPU on smaller volumes), and 128 x 32 tiles for
// FAST at radii 1..5 and EXACT at radii 1..4.
// Modelled cost: tiles per CU x plane steps per tile x work per step, which
// grows with the tile's rows plus their x halo plus a fixed per-step part
// (loads, barrier), per 64 columns; the z run of each shape as fused_tz
// picks it.  A 128-wide tile's step is weighted by kWideWalk: each source
// row it loads is one longer contiguous segment, and its memory walk is
// cheaper (measured at 512^3, FAST: 4.5%, 7%, 2%, 6% and 1% faster than
// 64 x 64 tiles at radii 1..5, where the kernel is bound by the walk; 1.7%
// slower at radius 6, where the arithmetic of its 44-row x pass dominates,
// so it is not built there; EXACT: 6%, 2.5%, 1% faster and even at radii
// 1..4, 2% slower at 5).  ITKG_FUSED_SHAPE=64|32|128 overrides.
constexpr double kWideWalk = 0.84;
int pick_shape(const Plan& p, const FusedEntry* fe, int mode, int ncomp) {
    if (mode != FUSED_FWD_FAST && mode != FUSED_ADJ && mode != FUSED_FWD_EXACT) return mode;
    // the other shapes of this arithmetic (-1: none)
    const int small = mode == FUSED_FWD_FAST ? FUSED_FWD_FAST_S : (mode == FUSED_ADJ ? FUSED_ADJ_S : -1);
    const int wide = mode == FUSED_FWD_FAST ? FUSED_FWD_FAST_W : (mode == FUSED_FWD_EXACT ? FUSED_FWD_EXACT_W : -1);
    const bool wide_ok = wide >= 0 && fe->ty[wide] > 0;
    if (const char* e = std::getenv("ITKG_FUSED_SHAPE")) {
        const int v = std::atoi(e);
        if (v == 32 && small >= 0) return small;
        if (v == 128 && wide_ok) return wide;
        return mode;
    }
    const itkg_region& d = fused_is_adjoint(mode) ? p.in_reg : p.req;
    const int64_t nx = d.size[0], ny = p.D >= 2 ? d.size[1] : 1, nz = p.D == 3 ? d.size[2] : 1;
    const int rxy = (int)p.radii[0], rz = p.D == 3 ? (int)p.radii[2] : 0;
    auto cost = [&](int m) {
        const int ty = fe->ty[m];
        const int tx = fused_tile_x(m);
        const int64_t txy = ((nx + tx - 1) / tx) * ((ny + ty - 1) / ty) * ncomp;
        const int tz = fused_tz((int)std::min<int64_t>(txy, INT32_MAX), (int)nz, rz);
        const int64_t tiles = txy * ((nz + tz - 1) / tz);
        const int64_t cus = device_cus();
        const double per_cu = (double)((tiles + cus - 1) / cus);
        const double step = (3.0 * ty + 2.0 * rxy + 30.0) * (tx / 64) * (tx > 64 ? kWideWalk : 1.0);
        return per_cu * (double)(std::min<int64_t>(tz, nz) + 2 * rz + 3) * step;
    };
    int pick = small >= 0 && cost(small) < cost(mode) ? small : mode;
    if (wide_ok && cost(wide) < cost(pick)) pick = wide;
    if (std::getenv("ITKG_DEBUG"))
        std::fprintf(stderr, "itkg: fused shape %d x %d tiles (mode %d)\n", fused_tile_x(pick),
                     fe->ty[pick], pick);
    return pick;
}

int run_fused(const Plan& p, int ncomp, const itkg_view& src_view, const itkg_view& dst_view,
              int mode, int accumulate, hipStream_t s, double amount = 0.0, double threshold = 0.0) {
    const FusedEntry* fe = fused_for(p);
    mode = pick_shape(p, fe, mode, ncomp);
    FusedMulti m;
    std::memset(&m, 0, sizeof m);
    m.nc = ncomp;  // one launch over (tile, component) items
    m.amount = amount;
    m.threshold = threshold;
    for (int c = 0; c < ncomp; ++c)
        fill_fused_args(p, src_view, dst_view, mode, accumulate, 0, c, ncomp, &m.a[c]);
    hipError_t e = fe->launch(m, mode, s);
    if (e != hipSuccess) return hip_fail(e, "fused launch");
    return ITKG_OK;
}

bool fits_i32(const Plan& p) {
    for (int d = 0; d < p.D; ++d)
        if (p.in_reg.size[d] > INT32_MAX || p.req.size[d] > INT32_MAX) return false;
    return true;
}

bool fused_src_ok(const Plan& p, const itkg_view& src, int ncomp, bool adjoint) {
    const itkg_region& sreg = adjoint ? p.req : p.in_reg;
    int64_t extent = ncomp - 1;  // largest in-plane element offset
    for (int d = 0; d < std::min(p.D, 2); ++d) {
        if (src.stride[d] < 0) return false;
        if (sreg.size[d] > 0) extent += (sreg.size[d] - 1) * src.stride[d] * ncomp;
    }
    return extent < (INT64_C(1) << 30);
}

bool fused_path_allowed() { return g_path.load() != ITKG_PATH_PER_AXIS; }

bool use_fused(const Plan& p, itkg_dtype dt, int ncomp) {
    const int path = g_path.load();
    if (path == ITKG_PATH_PER_AXIS) return false;
    return dt == ITKG_F32 && ncomp >= 1 && ncomp <= 4 && fits_i32(p) && fused_for(p) != nullptr;
}

int forward_impl(const Plan& p, itkg_dtype dt, int ncomp, const itkg_view& in,
                 const itkg_view& out, itkg_arith arith, hipStream_t s, void* scratch) {
    if (region_len(p.D, p.req) == 0) return ITKG_OK;
    if (use_fused(p, dt, ncomp) && fused_src_ok(p, in, ncomp, false))
        return run_fused(p, ncomp, in, out, arith == ITKG_FAST ? FUSED_FWD_FAST : FUSED_FWD_EXACT, 0,
                         s);
    if (g_path.load() == ITKG_PATH_FUSED)
        return fail(ITKG_ERR_UNSUPPORTED, "fused path requested but not available for this call");
    return dt == ITKG_F64 ? forward_per_axis<double>(p, ncomp, in, out, arith, s, scratch)
                          : forward_per_axis<float>(p, ncomp, in, out, arith, s, scratch);
}

template <typename T>
int backward_per_axis(const Plan& p, int ncomp, const itkg_view& gout, const itkg_view& gin,
                      int accumulate, hipStream_t s, void* scratch) {
    const int D = p.D;
    void* own = nullptr;
    const size_t sb = scratch_bytes(p, sizeof(T) == 8 ? ITKG_F64 : ITKG_F32, ncomp);
    if (D > 1 && !scratch && sb > 0) {
        hipError_t e = pool_alloc(&own, sb, s);
        if (e != hipSuccess) return hip_fail(e, "pool_alloc(adjoint scratch)");
        scratch = own;
    }
    char* buf[2] = {static_cast<char*>(scratch), static_cast<char*>(scratch) + sb / 2};
    itkg_view cur = gout;
    int rc = ITKG_OK;
    for (int a = D - 1; a >= 0 && rc == ITKG_OK; --a) {
        const itkg_region& dreg = p.reg[a];  // gradient w.r.t. this pass's input
        itkg_view nxt = (a == 0) ? gin : dense_view(D, buf[a & 1], dreg);
        AxisPassArgs<T> A;
        fill_pass(A, p, a, cur, nxt, dreg);
        A.clamp_lo = p.reg[a].index[a];
        A.clamp_hi = p.reg[a].index[a] + p.reg[a].size[a] - 1;
        A.accumulate = (a == 0) ? accumulate : 0;
        double acc = 0.0;
        for (int m = 0; m <= p.radii[a]; ++m) {
            A.tail[m] = (T)acc;
            acc += p.w[a][m];
        }
        const int64_t g_lo = p.reg[a + 1].index[a];
        const int64_t g_hi = g_lo + p.reg[a + 1].size[a] - 1;
        hipError_t e = launch_axis_adjoint_range<T>(A, ncomp, g_lo, g_hi, s);
        if (e != hipSuccess) rc = hip_fail(e, "adjoint launch");
        cur = nxt;
    }
    if (own) (void)hipFreeAsync(own, s);
    return rc;
}

int backward_impl(const Plan& p, itkg_dtype dt, int ncomp, const itkg_view& gout,
                  const itkg_view& gin, int accumulate, hipStream_t s, void* scratch) {
    if (region_len(p.D, p.req) == 0) return ITKG_OK;
    if (use_fused(p, dt, ncomp) && fused_src_ok(p, gout, ncomp, true))
        return run_fused(p, ncomp, gout, gin, FUSED_ADJ, accumulate, s);
    if (g_path.load() == ITKG_PATH_FUSED)
        return fail(ITKG_ERR_UNSUPPORTED, "fused path requested but not available for this call");
    return dt == ITKG_F64 ? backward_per_axis<double>(p, ncomp, gout, gin, accumulate, s, scratch)
                          : backward_per_axis<float>(p, ncomp, gout, gin, accumulate, s, scratch);
}

// Several requests over one input view (stream_full's pieces): one
// persistent launch of the fused batch kernel over every piece's tiles when
// the fused path serves them all (3-D, f32 scalar pixels, isotropic radii,
// non-empty requests), otherwise one forward_impl per piece.  Each piece is
// filled exactly as its own launch would fill it (same arguments, same
// arithmetic), only the z run and the tile shape come from the work of all
// pieces together, so the results are bit-identical either way.
int multi_forward(const std::vector<Plan>& plans, itkg_dtype dt, int ncomp, const itkg_view& in,
                  const itkg_view* outs, itkg_arith arith, hipStream_t s) {
    const int n = (int)plans.size();
    const FusedEntry* fe = n > 1 ? fused_for(plans[0]) : nullptr;
    bool one = fe && fe->batch && plans[0].D == 3 && fe->rz == fe->rxy && ncomp == 1 &&
               use_fused(plans[0], dt, ncomp);
    int64_t nz_max = 0;
    for (int k = 0; k < n && one; ++k) {
        const Plan& p = plans[k];
        one = fused_for(p) == fe && fits_i32(p) && fused_src_ok(p, in, ncomp, false) &&
              region_len(3, p.req) > 0;
        nz_max = std::max<int64_t>(nz_max, p.req.size[2]);
    }
    if (!one) {
        for (int k = 0; k < n; ++k)
            if (int rc = forward_impl(plans[k], dt, ncomp, in, outs[k], arith, s, nullptr)) return rc;
        return ITKG_OK;
    }
    const int rz = (int)plans[0].radii[2], rxy = (int)plans[0].radii[0];
    // tile shape and z run from all pieces' tiles (as pick_shape and
    // fused_tz model one launch)
    auto tiles_xy = [&](int m) {
        int64_t t = 0;
        for (const Plan& p : plans)
            t += ((p.req.size[0] + kFusedTileX - 1) / kFusedTileX) *
                 ((p.req.size[1] + fe->ty[m] - 1) / fe->ty[m]);
        return (int)std::min<int64_t>(t, INT32_MAX);
    };
    int mode = arith == ITKG_EXACT ? FUSED_FWD_EXACT : FUSED_FWD_FAST;
    if (mode == FUSED_FWD_FAST) {
        auto cost = [&](int m) {
            const int txy = tiles_xy(m);
            const int tz = fused_tz(txy, (int)nz_max, rz);
            const int64_t tiles = (int64_t)txy * ((nz_max + tz - 1) / tz);
            const int64_t cus = device_cus();
            return (double)((tiles + cus - 1) / cus) * (double)(std::min<int64_t>(tz, nz_max) + 2 * rz + 3) *
                   (3.0 * fe->ty[m] + 2.0 * rxy + 30.0);
        };
        if (const char* e = std::getenv("ITKG_FUSED_SHAPE"))
            mode = std::atoi(e) == 32 ? FUSED_FWD_FAST_S : FUSED_FWD_FAST;
        else if (cost(FUSED_FWD_FAST_S) < cost(FUSED_FWD_FAST))
            mode = FUSED_FWD_FAST_S;
    }
    const int tz = fused_tz(tiles_xy(mode), (int)nz_max, rz);
    std::vector<Fused3dArgs> args((size_t)n);
    std::vector<int2> items;
    for (int k = 0; k < n; ++k) {
        fill_fused_args(plans[k], in, outs[k], mode, 0, tz, 0, 1, &args[k]);
        const int nt = args[k].tiles_x * args[k].tiles_y * args[k].tiles_z;
        for (int t = 0; t < nt; ++t) items.push_back(make_int2(k, t));
    }
    // workgroup b runs items b, b + G, ... on XCD b mod 8 (G a multiple of
    // 8): 8 contiguous chunks keep neighbouring tiles on one XCD's L2
    const int ni = (int)items.size(), q = (ni + 7) / 8;
    std::vector<int2> work((size_t)q * 8, make_int2(-1, 0));
    for (int i = 0; i < ni; ++i) work[(size_t)(i % q) * 8 + i / q] = items[i];
    const size_t ab = args.size() * sizeof(Fused3dArgs), wb = work.size() * sizeof(int2);
    void* block = nullptr;
    hipError_t e = pool_alloc(&block, (ab + 255) / 256 * 256 + wb, s);
    if (e != hipSuccess) return hip_fail(e, "pool_alloc(piece tables)");
    char* base = static_cast<char*>(block);
    // copies from pageable memory: the host buffers may go once they return
    e = hipMemcpyAsync(base, args.data(), ab, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(base + (ab + 255) / 256 * 256, work.data(), wb, hipMemcpyHostToDevice, s);
    int rc = ITKG_OK;
    if (e != hipSuccess) rc = hip_fail(e, "piece tables");
    if (!rc) {
        const int per_cu = std::max(1, fe->occupancy(mode));
        const int nwork = (int)work.size();
        const int grid = std::min(nwork, ((device_cus() * per_cu + 7) / 8) * 8);
        e = fe->batch(reinterpret_cast<const Fused3dArgs*>(base),
                      reinterpret_cast<const int2*>(base + (ab + 255) / 256 * 256), nwork, grid,
                      mode, s);
        if (e != hipSuccess) rc = hip_fail(e, "piece launch");
    }
    (void)hipFreeAsync(block, s);
    return rc;
}

// Host buffers, one piece: stage the input region in, run the plan, stage
// the requested region out, on a private stream.
int host_forward_serial(const Plan& p, itkg_dtype dtype, int ncomp, const void* input,
                        void* output, itkg_arith arith) {
    const int ndim = p.D;
    const int64_t nin = region_len(ndim, p.in_reg), nout = region_len(ndim, p.req);
    int rc = ITKG_OK;
    const size_t px = dtype_size(dtype) * ncomp;
    hipStream_t s = nullptr;
    hipError_t e = acquire_host_stream(&s);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    void *din = nullptr, *dout = nullptr;
    e = pool_alloc(&din, (size_t)nin * px, s);
    if (e == hipSuccess) e = pool_alloc(&dout, (size_t)nout * px, s);
    if (e != hipSuccess) {
        rc = hip_fail(e, "pool_alloc(host staging)");
    } else {
        e = hipMemcpyAsync(din, input, (size_t)nin * px, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) rc = hip_fail(e, "copy in");
        if (!rc) {
            const itkg_view vin = dense_view(ndim, din, p.in_reg);
            const itkg_view vout = dense_view(ndim, dout, p.req);
            rc = forward_impl(p, dtype, ncomp, vin, vout, arith, s, nullptr);
        }
        if (!rc) {
            e = hipMemcpyAsync(output, dout, (size_t)nout * px, hipMemcpyDeviceToHost, s);
            if (e != hipSuccess) rc = hip_fail(e, "copy out");
        }
    }
    if (din) (void)hipFreeAsync(din, s);
    if (dout) (void)hipFreeAsync(dout, s);
    e = hipStreamSynchronize(s);
    if (e != hipSuccess && !rc) rc = hip_fail(e, "hipStreamSynchronize");
    release_host_stream(s);
    return rc;
}

// Host buffers, pipelined (src/source.rs:81-97 applied to the transfers):
// the request is split into pieces along its slowest axis, and the staging
// of host_stage.cpp moves the input in and the output out in chunks while
// the pieces run.  Piece k's kernels wait for the input planes the piece
// reads; each piece is computed exactly as a streamed piece, so the output
// is bit-identical to the serial form.  Pieces hold about kHostPieceBytes of
// the larger of the two transfers.
constexpr int64_t kHostPieceBytes = 16ll << 20;
constexpr int kHostMaxPieces = 32;

struct HostPieces {
    std::vector<Plan> piece;
    std::vector<int64_t> in_planes_end;   // input planes (from p.in_reg) piece k reads: [0, end)
    std::vector<int64_t> out_planes_end;  // output planes (from p.req) done after piece k
    int64_t in_plane = 0, out_plane = 0;  // pixels per plane
};

// Splits p along its slowest axis into pieces of about kHostPieceBytes of
// the larger transfer (in_px / out_px bytes per pixel); 1 piece when the
// request is small.
int host_pieces(const Plan& p, size_t in_px, size_t out_px, HostPieces* hp) {
    const int D = p.D, ax = D - 1;
    const int64_t nin = region_len(D, p.in_reg), nout = region_len(D, p.req);
    const int64_t in_planes = p.in_reg.size[ax], out_planes = p.req.size[ax];
    const int64_t big = std::max<int64_t>(nin * (int64_t)in_px, nout * (int64_t)out_px);
    const int K = (int)std::max<int64_t>(
        1, std::min<int64_t>({(int64_t)kHostMaxPieces, out_planes, in_planes,
                              big / kHostPieceBytes}));
    hp->in_plane = nin / in_planes;
    hp->out_plane = nout / out_planes;
    hp->piece.assign(K, p);
    hp->in_planes_end.resize(K);
    hp->out_planes_end.resize(K);
    for (int k = 0; k < K; ++k) {
        const int64_t o0 = out_planes * k / K, o1 = out_planes * (k + 1) / K;
        itkg_region r = p.req;
        r.index[ax] = p.req.index[ax] + o0;
        r.size[ax] = o1 - o0;
        if (int rc = plan_regions(D, &p.bounds, r, &hp->piece[k])) return rc;
        const Plan& q = hp->piece[k];
        hp->in_planes_end[k] = q.in_reg.index[ax] + q.in_reg.size[ax] - p.in_reg.index[ax];
        hp->out_planes_end[k] = o1;
    }
    return ITKG_OK;
}

StageStats g_last_stage;  // of the last staged call (read by itkg_host_staging_info)
std::mutex g_last_stage_mu;

void note_stage(const StageStats& st) {
    std::lock_guard<std::mutex> lk(g_last_stage_mu);
    g_last_stage = st;
}

int host_forward(const Plan& p, itkg_dtype dtype, int ncomp, const void* input, void* output,
                 itkg_arith arith) {
    const int D = p.D;
    const int64_t nin = region_len(D, p.in_reg), nout = region_len(D, p.req);
    if (nout == 0) return ITKG_OK;
    if (!input || !output) return fail(ITKG_ERR_INVALID_ARG, "null host buffer");
    const size_t px = dtype_size(dtype) * ncomp;
    HostPieces hp;
    if (int rc = host_pieces(p, px, px, &hp)) return rc;
    const int K = (int)hp.piece.size();
    if (K <= 1) return host_forward_serial(p, dtype, ncomp, input, output, arith);

    hipStream_t s = nullptr;
    hipError_t e = acquire_host_stream(&s);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    void *din = nullptr, *dout = nullptr;
    e = pool_alloc(&din, (size_t)nin * px, s);
    if (e == hipSuccess) e = pool_alloc(&dout, (size_t)nout * px, s);
    // the copy streams use the buffers: in place before they start
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    int rc = ITKG_OK;
    if (e != hipSuccess) {
        rc = hip_fail(e, "pool_alloc(host staging)");
    } else {
        StagedJob job;
        job.host_in = input;
        job.in_bytes = (size_t)nin * px;
        job.host_out = output;
        job.out_bytes = (size_t)nout * px;
        job.dev_in = din;
        job.dev_out = dout;
        job.pieces = K;
        for (int k = 0; k < K; ++k) {
            job.in_need.push_back((size_t)(hp.in_planes_end[k] * hp.in_plane) * px);
            job.out_end.push_back((size_t)(hp.out_planes_end[k] * hp.out_plane) * px);
        }
        const itkg_view vin = dense_view(D, din, p.in_reg);
        const itkg_view vout = dense_view(D, dout, p.req);
        job.compute = [&](int k, hipStream_t st) {
            return forward_impl(hp.piece[k], dtype, ncomp, vin, vout, arith, st, nullptr);
        };
        StageStats stats;
        job.stats = &stats;
        rc = run_staged(job, s);
        note_stage(stats);
    }
    if (din) (void)hipFreeAsync(din, s);
    if (dout) (void)hipFreeAsync(dout, s);
    e = hipStreamSynchronize(s);
    if (e != hipSuccess && !rc) rc = hip_fail(e, "hipStreamSynchronize");
    release_host_stream(s);
    return rc;
}

bool ptype_ok(itkg_ptype t) { return t >= ITKG_PT_U8 && t <= ITKG_PT_F64; }

size_t ptype_size(itkg_ptype t) {
    switch (t) {
        case ITKG_PT_U8: return 1;
        case ITKG_PT_U16:
        case ITKG_PT_I16: return 2;
        case ITKG_PT_I32:
        case ITKG_PT_F32: return 4;
        default: return 8;
    }
}

// Enqueues the conversion of region `reg` from view `in` to view `out`.
int convert_views(int ndim, itkg_ptype in_type, itkg_ptype out_type, int ncomp,
                  const itkg_view& in, const itkg_view& out, const itkg_region& reg,
                  itkg_convert_mode mode, hipStream_t s) {
    ConvertArgs a{};
    int64_t at[ITKG_MAX_DIM] = {0, 0, 0, 0};
    for (int d = 0; d < ndim; ++d) at[d] = reg.index[d];
    a.in = static_cast<const char*>(in.data) + view_offset(ndim, in, at) * ncomp * (int64_t)ptype_size(in_type);
    a.out = static_cast<char*>(out.data) + view_offset(ndim, out, at) * ncomp * (int64_t)ptype_size(out_type);
    for (int d = 0; d < ITKG_MAX_DIM; ++d) {
        a.n[d] = d < ndim ? reg.size[d] : 1;
        a.in_stride[d] = d < ndim ? in.stride[d] : 0;
        a.out_stride[d] = d < ndim ? out.stride[d] : 0;
    }
    a.ncomp = ncomp;
    a.in_type = in_type;
    a.out_type = out_type;
    a.mode = mode;
    const hipError_t e = launch_convert(a, s);
    return e == hipSuccess ? ITKG_OK : hip_fail(e, "conversion launch");
}

bool args_ok(int ndim, itkg_dtype dt, int ncomp) {
    return ndim >= 1 && ndim <= ITKG_MAX_DIM && (dt == ITKG_F32 || dt == ITKG_F64) && ncomp >= 1 &&
           ncomp <= 4;
}

}  // namespace itkg

// Scratch of itkg_gaussian_generate_region_noalloc: the per-axis path's two
// ping-pong intermediates, in a file-scope device array instead of a pool
// allocation.  Each GPU the library runs on holds its own copy until the
// process ends.  In host code `scratch` names a host-side stand-in, not
// device memory: device code only ever receives the address that
// hipGetSymbolAddress returns for the current device (noalloc_scratch).
// Every pass writes its whole output region before the next pass reads it,
// so nothing has to start at zero; calls that overlap in time on one GPU
// share the array and race.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace itkg {
namespace {

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
int noalloc_scratch(void** out) {
    static std::mutex mu;
    static void* addr[kPoolDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kPoolDevices) return fail(ITKG_ERR_UNSUPPORTED, "device %d", dev);
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(ITKG_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
        addr[dev] = p;
    }
    *out = addr[dev];
    return ITKG_OK;
}

}  // namespace
}  // namespace itkg

using namespace itkg;

extern "C" {

uint64_t itkg_kernel_radius(double sigma, double spacing) { return radius_of(sigma, spacing); }

int itkg_build_kernel(double sigma_px, uint64_t radius, double* out, uint64_t cap) {
    clear_error();
    if (!out) return fail(ITKG_ERR_INVALID_ARG, "null output");
    if (radius > (uint64_t)ITKG_MAX_RADIUS)
        return fail(ITKG_ERR_UNSUPPORTED, "radius exceeds %d", ITKG_MAX_RADIUS);
    if (cap < 2 * radius + 1) return fail(ITKG_ERR_INVALID_ARG, "capacity below 2r+1");
    weights_of(sigma_px, radius, out);
    return ITKG_OK;
}

int itkg_input_region_for_output(int32_t ndim, double sigma, const double* spacing,
                                 const itkg_region* largest, const itkg_region* output,
                                 itkg_region* input) {
    clear_error();
    if (ndim < 1 || ndim > ITKG_MAX_DIM || !spacing || !largest || !output || !input)
        return fail(ITKG_ERR_INVALID_ARG, "bad arguments");
    int64_t radii[ITKG_MAX_DIM] = {0, 0, 0, 0};
    for (int d = 0; d < ndim; ++d) {
        const uint64_t r = radius_of(sigma, spacing[d]);
        radii[d] = r > (uint64_t)INT64_MAX / 4 ? INT64_MAX / 4 : (int64_t)r;
    }
    itkg_region out{};
    out = clipped(ndim, padded(ndim, *output, radii), *largest);
    for (int d = ndim; d < ITKG_MAX_DIM; ++d) {
        out.index[d] = 0;
        out.size[d] = 0;
    }
    *input = out;
    return ITKG_OK;
}

int itkg_split_region(int32_t ndim, const itkg_region* region, uint64_t num_pieces,
                      itkg_region* pieces, uint64_t cap, uint64_t* count) {
    clear_error();
    if (ndim < 1 || ndim > ITKG_MAX_DIM || !region || !count)
        return fail(ITKG_ERR_INVALID_ARG, "bad arguments");
    const uint64_t n = num_pieces < 1 ? 1 : num_pieces;
    *count = 0;
    if (region_len(ndim, *region) == 0) return ITKG_OK;
    const int slow = ndim - 1;
    const uint64_t total = (uint64_t)region->size[slow];
    const uint64_t actual = std::max<uint64_t>(std::min(n, total), 1);
    if (!pieces || cap < actual) return fail(ITKG_ERR_INVALID_ARG, "piece capacity below %llu",
                                             (unsigned long long)actual);
    const uint64_t base = total / actual, remainder = total % actual;
    int64_t start = region->index[slow];
    for (uint64_t i = 0; i < actual; ++i) {
        const uint64_t this_size = base + (i < remainder ? 1 : 0);
        itkg_region p = *region;
        p.index[slow] = start;
        p.size[slow] = (int64_t)this_size;
        pieces[i] = p;
        start += (int64_t)this_size;
    }
    *count = actual;
    return ITKG_OK;
}

int itkg_gaussian_generate_region(int32_t ndim, itkg_dtype dtype, int32_t ncomp, double sigma,
                                  const double* spacing, const itkg_region* largest,
                                  const itkg_view* input, const itkg_view* output,
                                  itkg_arith arith, void* stream) {
    clear_error();
    if (!args_ok(ndim, dtype, ncomp)) return fail(ITKG_ERR_INVALID_ARG, "bad ndim/dtype/ncomp");
    if (!output) return fail(ITKG_ERR_INVALID_ARG, "null output view");
    Plan p;
    int rc = make_plan(ndim, sigma, spacing, largest, output->region, &p);
    if (rc) return rc;
    if ((rc = check_view(ndim, input, p.in_reg, "input"))) return rc;
    if ((rc = check_view(ndim, output, p.req, "output"))) return rc;
    return forward_impl(p, dtype, ncomp, *input, *output, arith,
                        static_cast<hipStream_t>(stream), nullptr);
}

int itkg_gaussian_generate_regions(int32_t ndim, itkg_dtype dtype, int32_t ncomp, double sigma,
                                   const double* spacing, const itkg_region* largest,
                                   const itkg_view* input, int32_t n, const itkg_view* outputs,
                                   itkg_arith arith, void* stream) {
    clear_error();
    if (!args_ok(ndim, dtype, ncomp)) return fail(ITKG_ERR_INVALID_ARG, "bad ndim/dtype/ncomp");
    if (n < 0) return fail(ITKG_ERR_INVALID_ARG, "negative count");
    if (n == 0) return ITKG_OK;
    if (!outputs) return fail(ITKG_ERR_INVALID_ARG, "null output views");
    std::vector<Plan> plans((size_t)n);
    for (int k = 0; k < n; ++k) {
        int rc = make_plan(ndim, sigma, spacing, largest, outputs[k].region, &plans[k]);
        if (!rc) rc = check_view(ndim, input, plans[k].in_reg, "input");
        if (!rc) rc = check_view(ndim, &outputs[k], plans[k].req, "output");
        if (rc) return rc;
    }
    return multi_forward(plans, dtype, ncomp, *input, outputs, arith,
                         static_cast<hipStream_t>(stream));
}

int itkg_gaussian_generate_region_noalloc(int32_t ndim, itkg_dtype dtype, int32_t ncomp,
                                          double sigma, const double* spacing,
                                          const itkg_region* largest, const itkg_view* input,
                                          const itkg_view* output, itkg_arith arith,
                                          void* stream) {
    clear_error();
    if (!args_ok(ndim, dtype, ncomp)) return fail(ITKG_ERR_INVALID_ARG, "bad ndim/dtype/ncomp");
    if (!output) return fail(ITKG_ERR_INVALID_ARG, "null output view");
    Plan p;
    int rc = make_plan(ndim, sigma, spacing, largest, output->region, &p);
    if (rc) return rc;
    if ((rc = check_view(ndim, input, p.in_reg, "input"))) return rc;
    if ((rc = check_view(ndim, output, p.req, "output"))) return rc;
    // the fused kernel and 1-D calls take no scratch; the per-axis path takes
    // its two intermediates from the static array
    const bool fused = use_fused(p, dtype, ncomp) && fused_src_ok(p, *input, ncomp, false);
    const size_t need = (fused || ndim == 1) ? 0 : scratch_bytes(p, dtype, ncomp);
    if (need > kNoAllocScratchBytes)
        return fail(ITKG_ERR_UNSUPPORTED,
                    "this call needs %zu bytes of per-axis scratch; the allocation-free entry "
                    "point holds %zu",
                    need, (size_t)kNoAllocScratchBytes);
    void* scr = nullptr;
    if ((rc = noalloc_scratch(&scr))) return rc;
    return forward_impl(p, dtype, ncomp, *input, *output, arith, static_cast<hipStream_t>(stream),
                        scr);
}

int itkg_gaussian_generate_region_host(int32_t ndim, itkg_dtype dtype, int32_t ncomp,
                                       double sigma, const double* spacing,
                                       const itkg_region* largest, const void* input,
                                       const itkg_region* requested, void* output,
                                       itkg_arith arith) {
    clear_error();
    if (!args_ok(ndim, dtype, ncomp)) return fail(ITKG_ERR_INVALID_ARG, "bad ndim/dtype/ncomp");
    if (!requested) return fail(ITKG_ERR_INVALID_ARG, "null requested region");
    Plan p;
    int rc = make_plan(ndim, sigma, spacing, largest, *requested, &p);
    if (rc) return rc;
    return host_forward(p, dtype, ncomp, input, output, arith);
}

int itkg_box_kernel(uint64_t radius, double* out, uint64_t cap) {
    clear_error();
    if (!out) return fail(ITKG_ERR_INVALID_ARG, "null output");
    if (radius > (uint64_t)ITKG_MAX_RADIUS)
        return fail(ITKG_ERR_UNSUPPORTED, "radius exceeds %d", ITKG_MAX_RADIUS);
    if (cap < 2 * radius + 1) return fail(ITKG_ERR_INVALID_ARG, "capacity below 2r+1");
    const uint64_t len = 2 * radius + 1;
    for (uint64_t k = 0; k < len; ++k) out[k] = 1.0 / (double)len;
    return ITKG_OK;
}

int itkg_separable_input_region_for_output(int32_t ndim, const uint64_t* radii,
                                           const itkg_region* largest, const itkg_region* output,
                                           itkg_region* input) {
    clear_error();
    if (ndim < 1 || ndim > ITKG_MAX_DIM || !radii || !largest || !output || !input)
        return fail(ITKG_ERR_INVALID_ARG, "bad arguments");
    int64_t r[ITKG_MAX_DIM] = {0, 0, 0, 0};
    for (int d = 0; d < ndim; ++d)
        r[d] = radii[d] > (uint64_t)INT64_MAX / 4 ? INT64_MAX / 4 : (int64_t)radii[d];
    itkg_region out = clipped(ndim, padded(ndim, *output, r), *largest);
    for (int d = ndim; d < ITKG_MAX_DIM; ++d) {
        out.index[d] = 0;
        out.size[d] = 0;
    }
    *input = out;
    return ITKG_OK;
}

int itkg_separable_generate_region(int32_t ndim, itkg_dtype dtype, int32_t ncomp,
                                   const uint64_t* radii, const double* weights,
                                   const itkg_region* largest, const itkg_view* input,
                                   const itkg_view* output, itkg_arith arith, void* stream) {
    clear_error();
    if (!args_ok(ndim, dtype, ncomp)) return fail(ITKG_ERR_INVALID_ARG, "bad ndim/dtype/ncomp");
    if (!output) return fail(ITKG_ERR_INVALID_ARG, "null output view");
    Plan p;
    int rc = make_plan_weights(ndim, radii, weights, largest, output->region, &p);
    if (rc) return rc;
    if ((rc = check_view(ndim, input, p.in_reg, "input"))) return rc;
    if ((rc = check_view(ndim, output, p.req, "output"))) return rc;
    return forward_impl(p, dtype, ncomp, *input, *output, arith,
                        static_cast<hipStream_t>(stream), nullptr);
}

int itkg_separable_generate_region_host(int32_t ndim, itkg_dtype dtype, int32_t ncomp,
                                        const uint64_t* radii, const double* weights,
                                        const itkg_region* largest, const void* input,
                                        const itkg_region* requested, void* output,
                                        itkg_arith arith) {
    clear_error();
    if (!args_ok(ndim, dtype, ncomp)) return fail(ITKG_ERR_INVALID_ARG, "bad ndim/dtype/ncomp");
    if (!requested) return fail(ITKG_ERR_INVALID_ARG, "null requested region");
    Plan p;
    int rc = make_plan_weights(ndim, radii, weights, largest, *requested, &p);
    if (rc) return rc;
    return host_forward(p, dtype, ncomp, input, output, arith);
}

int itkg_index_gaussian_kernel(double sigma, double* out, uint64_t cap, uint64_t* radius) {
    clear_error();
    if (!out || !radius) return fail(ITKG_ERR_INVALID_ARG, "null output");
    const uint64_t r = radius_of(sigma, 1.0);
    if (r > (uint64_t)ITKG_MAX_RADIUS)
        return fail(ITKG_ERR_UNSUPPORTED, "radius exceeds %d", ITKG_MAX_RADIUS);
    if (cap < 2 * r + 1) return fail(ITKG_ERR_INVALID_ARG, "capacity below 2r+1");
    *radius = index_weights(sigma, out, cap);
    return ITKG_OK;
}

namespace {

// UnsharpMaskFilter::generate_region (src/filters/edges.rs:406-452) on
// device views: the blur of the requested region (the separable path with
// the filter's own weights; each pass clamps to the same extents as the
// reference's whole-buffer passes for every requested pixel), then the
// per-pixel step.
int unsharp_impl(int ndim, itkg_dtype dtype, double sigma, double amount, double threshold,
                 const itkg_region* largest, const itkg_view& in, const itkg_view& out,
                 itkg_arith arith, hipStream_t s) {
    double w[kMaxTaps * ITKG_MAX_DIM];
    const uint64_t r = radius_of(sigma, 1.0);
    if (r > (uint64_t)ITKG_MAX_RADIUS)
        return fail(ITKG_ERR_UNSUPPORTED, "radius exceeds %d", ITKG_MAX_RADIUS);
    uint64_t radii[ITKG_MAX_DIM] = {0, 0, 0, 0};
    for (int d = 0; d < ndim; ++d) {
        index_weights(sigma, w + d * (2 * r + 1), kMaxTaps);
        radii[d] = r;
    }
    Plan p;
    int rc = make_plan_weights(ndim, radii, w, largest, out.region, &p, false);
    if (rc) return rc;
    if ((rc = check_view(ndim, &in, p.in_reg, "input"))) return rc;
    if ((rc = check_view(ndim, &out, p.req, "output"))) return rc;
    const int64_t n = region_len(ndim, p.req);
    if (n == 0) return ITKG_OK;
    // f32 on the fused kernel: the step in the blur's own launch, one round
    // trip (the blur never reaches memory)
    const int umode = arith == ITKG_FAST ? FUSED_FWD_FAST_U : FUSED_FWD_EXACT_U;
    if (ndim >= 2 && use_fused(p, dtype, 1) && fused_src_ok(p, in, 1, false) &&
        fused_for(p)->ty[umode] > 0)
        return run_fused(p, 1, in, out, umode, 0, s, amount, threshold);
    const size_t px = dtype_size(dtype);
    void* blur = nullptr;
    hipError_t e = pool_alloc(&blur, (size_t)n * px, s);
    if (e != hipSuccess) return hip_fail(e, "pool_alloc(unsharp blur)");
    const itkg_view bv = dense_view(ndim, blur, p.req);
    rc = forward_impl(p, dtype, 1, in, bv, arith, s, nullptr);
    if (!rc) {
        auto fill = [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            UnsharpArgs<T> a{};
            a.in = static_cast<const T*>(in.data) + view_offset(ndim, in, p.req.index);
            a.blur = static_cast<const T*>(blur);
            a.out = static_cast<T*>(out.data) + view_offset(ndim, out, p.req.index);
            for (int d = 0; d < ITKG_MAX_DIM; ++d) {
                a.n[d] = d < ndim ? p.req.size[d] : 1;
                a.in_stride[d] = d < ndim ? in.stride[d] : 0;
                a.blur_stride[d] = d < ndim ? bv.stride[d] : 0;
                a.out_stride[d] = d < ndim ? out.stride[d] : 0;
            }
            a.amount = amount;
            a.threshold = threshold;
            return launch_unsharp<T>(a, s);
        };
        e = dtype == ITKG_F64 ? fill((double*)nullptr) : fill((float*)nullptr);
        if (e != hipSuccess) rc = hip_fail(e, "unsharp launch");
    }
    (void)hipFreeAsync(blur, s);
    return rc;
}

}  // namespace

int itkg_unsharp_mask_generate_region(int32_t ndim, itkg_dtype dtype, double sigma, double amount,
                                      double threshold, const itkg_region* largest,
                                      const itkg_view* input, const itkg_view* output,
                                      itkg_arith arith, void* stream) {
    clear_error();
    if (!args_ok(ndim, dtype, 1)) return fail(ITKG_ERR_INVALID_ARG, "bad ndim/dtype");
    if (!input || !output || !largest) return fail(ITKG_ERR_INVALID_ARG, "null argument");
    return unsharp_impl(ndim, dtype, sigma, amount, threshold, largest, *input, *output, arith,
                        static_cast<hipStream_t>(stream));
}

int itkg_unsharp_mask_generate_region_host(int32_t ndim, itkg_dtype dtype, double sigma,
                                           double amount, double threshold,
                                           const itkg_region* largest, const void* input,
                                           const itkg_region* requested, void* output,
                                           itkg_arith arith) {
    clear_error();
    if (!args_ok(ndim, dtype, 1)) return fail(ITKG_ERR_INVALID_ARG, "bad ndim/dtype");
    if (!largest || !requested) return fail(ITKG_ERR_INVALID_ARG, "null region");
    if (!region_ok(ndim, *largest) || !region_ok(ndim, *requested))
        return fail(ITKG_ERR_INVALID_ARG, "negative region size");
    int64_t rr[ITKG_MAX_DIM] = {0, 0, 0, 0};
    const uint64_t r = radius_of(sigma, 1.0);
    if (r > (uint64_t)ITKG_MAX_RADIUS)
        return fail(ITKG_ERR_UNSUPPORTED, "radius exceeds %d", ITKG_MAX_RADIUS);
    for (int d = 0; d < ndim; ++d) rr[d] = (int64_t)r;
    const itkg_region in_reg = clipped(ndim, padded(ndim, *requested, rr), *largest);
    const int64_t nin = region_len(ndim, in_reg), nout = region_len(ndim, *requested);
    if (nout == 0) return ITKG_OK;
    if (!input || !output) return fail(ITKG_ERR_INVALID_ARG, "null host buffer");
    const size_t px = dtype_size(dtype);
    hipStream_t s = nullptr;
    hipError_t e = acquire_host_stream(&s);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    int rc = ITKG_OK;
    void *din = nullptr, *dout = nullptr;
    e = pool_alloc(&din, (size_t)nin * px, s);
    if (e == hipSuccess) e = pool_alloc(&dout, (size_t)nout * px, s);
    if (e == hipSuccess) e = hipMemcpyAsync(din, input, (size_t)nin * px, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) rc = hip_fail(e, "unsharp host staging");
    if (!rc) {
        const itkg_view vin = dense_view(ndim, din, in_reg);
        const itkg_view vout = dense_view(ndim, dout, *requested);
        rc = unsharp_impl(ndim, dtype, sigma, amount, threshold, largest, vin, vout, arith, s);
    }
    if (!rc) {
        e = hipMemcpyAsync(output, dout, (size_t)nout * px, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) rc = hip_fail(e, "copy out");
    }
    if (din) (void)hipFreeAsync(din, s);
    if (dout) (void)hipFreeAsync(dout, s);
    e = hipStreamSynchronize(s);
    if (e != hipSuccess && !rc) rc = hip_fail(e, "hipStreamSynchronize");
    release_host_stream(s);
    return rc;
}

}  // extern "C"

namespace itkg {
namespace {

// yvv_coeffs (src/filters/recursive_gaussian.rs:33-47), the same expressions
// in the same order (f64::max ignores a NaN, as fmax does).
void yvv_coefficients(double sigma, double c[4]) {
    sigma = std::fmax(sigma, 0.5);
    const double q = sigma >= 2.5 ? 0.98711 * sigma - 0.96330
                                  : 3.97156 - 4.14554 * std::sqrt(1.0 - 0.26891 * sigma);
    const double q2 = q * q;
    const double q3 = q2 * q;
    const double b0 = 1.57825 + 2.44413 * q + 1.4281 * q2 + 0.422205 * q3;
    const double b1 = 2.44413 * q + 2.85619 * q2 + 1.26661 * q3;
    const double b2 = -(1.4281 * q2 + 1.26661 * q3);
    const double b3 = 0.422205 * q3;
    c[0] = 1.0 - (b1 + b2 + b3) / b0;
    c[1] = b1 / b0;
    c[2] = b2 / b0;
    c[3] = b3 / b0;
}

// RecursiveGaussianFilter / RecursiveGaussianAxisFilter::generate_region
// (recursive_gaussian.rs:147-176, 219-240) on device views: the input covers
// the whole largest region (the filters' input_region_for_output).
int recursive_impl(int ndim, itkg_dtype dtype, double sigma, const double* spacing, int axis,
                   const itkg_region& largest, const itkg_view& in, const itkg_view& out,
                   hipStream_t s) {
    if ((axis < -1 || axis >= ndim)) return fail(ITKG_ERR_INVALID_ARG, "axis %d of a %d-D image", axis, ndim);
    if (!spacing) return fail(ITKG_ERR_INVALID_ARG, "null spacing");
    if (!region_ok(ndim, largest) || !region_ok(ndim, out.region))
        return fail(ITKG_ERR_INVALID_ARG, "negative region size");
    if (region_len(ndim, out.region) > 0 && !contains(ndim, largest, out.region))
        return fail(ITKG_ERR_INVALID_ARG, "requested region is not inside the largest region");
    int rc = check_view(ndim, &in, largest, "input");
    if (rc) return rc;
    if ((rc = check_view(ndim, &out, out.region, "output"))) return rc;
    const int64_t total = region_len(ndim, largest);
    if (region_len(ndim, out.region) == 0 || total == 0) return ITKG_OK;
    int axes[ITKG_MAX_DIM];
    double coef[ITKG_MAX_DIM][4];
    int naxes = 0;
    bool any_line = false;  // a filtered axis with lines of more than one sample
    for (int d = 0; d < ndim; ++d) {
        if (axis >= 0 && d != axis) continue;
        axes[naxes] = d;
        yvv_coefficients(sigma / spacing[d], coef[naxes]);
        ++naxes;
        any_line = any_line || largest.size[d] > 1;
    }
    if (!any_line) {
        // the result is the input (every filtered line has one sample); the
        // widen / narrow copy grids take axis 1 and the plane count of axes
        // 2-3 as grid dimensions
        const int64_t n1 = ndim > 1 ? largest.size[1] : 1;
        const int64_t planes = (ndim > 2 ? largest.size[2] : 1) * (ndim > 3 ? largest.size[3] : 1);
        if (n1 > 65535 || planes > 65535)
            return fail(ITKG_ERR_UNSUPPORTED,
                        "recursive Gaussian with no filtered line longer than one sample: axis 1 "
                        "or axes 2-3 exceed 65535 rows");
    }
    // the f64 work image and, after it, the largest checkpoint table of the
    // filtered axes (one block)
    int64_t ck = 0;
    for (int k = 0; k < naxes; ++k)
        if (largest.size[axes[k]] > 1)
            ck = std::max(ck, recursive_ckpt_doubles(total, largest.size[axes[k]], axes[k]));
    void* work = nullptr;
    hipError_t e = pool_alloc(&work, (size_t)(total + ck) * sizeof(double), s);
    if (e != hipSuccess) return hip_fail(e, "pool_alloc(recursive work)");
    auto run = [&](auto* tag) {
        using T = std::remove_pointer_t<decltype(tag)>;
        RecursiveArgs<T> a{};
        a.in = static_cast<const T*>(in.data) + view_offset(ndim, in, largest.index);
        a.work = static_cast<double*>(work);
        a.ckpt = static_cast<double*>(work) + total;
        a.out = static_cast<T*>(out.data);
        for (int d = 0; d < ITKG_MAX_DIM; ++d) {
            a.n[d] = d < ndim ? largest.size[d] : 1;
            a.m[d] = d < ndim ? out.region.size[d] : 1;
            a.off[d] = d < ndim ? out.region.index[d] - largest.index[d] : 0;
            a.in_stride[d] = d < ndim ? in.stride[d] : 0;
            a.out_stride[d] = d < ndim ? out.stride[d] : 0;
        }
        return launch_recursive<T>(a, axes, naxes, coef, s);
    };
    e = dtype == ITKG_F64 ? run((double*)nullptr) : run((float*)nullptr);
    if (e != hipSuccess) rc = hip_fail(e, "recursive launch");
    (void)hipFreeAsync(work, s);
    return rc;
}

}  // namespace
}  // namespace itkg

extern "C" {

int itkg_yvv_coefficients(double sigma_px, double* coef) {
    clear_error();
    if (!coef) return fail(ITKG_ERR_INVALID_ARG, "null output");
    yvv_coefficients(sigma_px, coef);
    return ITKG_OK;
}

int itkg_recursive_gaussian_generate_region(int32_t ndim, itkg_dtype dtype, double sigma,
                                            const double* spacing, int32_t axis,
                                            const itkg_region* largest, const itkg_view* input,
                                            const itkg_view* output, void* stream) {
    clear_error();
    if (!args_ok(ndim, dtype, 1)) return fail(ITKG_ERR_INVALID_ARG, "bad ndim/dtype");
    if (!largest || !input || !output) return fail(ITKG_ERR_INVALID_ARG, "null argument");
    return recursive_impl(ndim, dtype, sigma, spacing, axis, *largest, *input, *output,
                          static_cast<hipStream_t>(stream));
}

int itkg_recursive_gaussian_generate_region_host(int32_t ndim, itkg_dtype dtype, double sigma,
                                                 const double* spacing, int32_t axis,
                                                 const itkg_region* largest, const void* input,
                                                 const itkg_region* requested, void* output) {
    clear_error();
    if (!args_ok(ndim, dtype, 1)) return fail(ITKG_ERR_INVALID_ARG, "bad ndim/dtype");
    if (!largest || !requested) return fail(ITKG_ERR_INVALID_ARG, "null region");
    if (!region_ok(ndim, *largest) || !region_ok(ndim, *requested))
        return fail(ITKG_ERR_INVALID_ARG, "negative region size");
    const int64_t nin = region_len(ndim, *largest), nout = region_len(ndim, *requested);
    if (nout == 0) return ITKG_OK;
    if (!input || !output) return fail(ITKG_ERR_INVALID_ARG, "null host buffer");
    const size_t px = dtype_size(dtype);
    hipStream_t s = nullptr;
    hipError_t e = acquire_host_stream(&s);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    int rc = ITKG_OK;
    void *din = nullptr, *dout = nullptr;
    e = pool_alloc(&din, (size_t)nin * px, s);
    if (e == hipSuccess) e = pool_alloc(&dout, (size_t)nout * px, s);
    if (e == hipSuccess) e = hipMemcpyAsync(din, input, (size_t)nin * px, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) rc = hip_fail(e, "recursive host staging");
    if (!rc) {
        const itkg_view vin = dense_view(ndim, din, *largest);
        const itkg_view vout = dense_view(ndim, dout, *requested);
        rc = recursive_impl(ndim, dtype, sigma, spacing, axis, *largest, vin, vout, s);
    }
    if (!rc) {
        e = hipMemcpyAsync(output, dout, (size_t)nout * px, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) rc = hip_fail(e, "copy out");
    }
    if (din) (void)hipFreeAsync(din, s);
    if (dout) (void)hipFreeAsync(dout, s);
    e = hipStreamSynchronize(s);
    if (e != hipSuccess && !rc) rc = hip_fail(e, "hipStreamSynchronize");
    release_host_stream(s);
    return rc;
}

int itkg_canny_generate_region(int32_t ndim, itkg_dtype dtype, double sigma, double lower,
                               double upper, const itkg_region* largest, const itkg_view* input,
                               const itkg_view* output, itkg_arith arith, void* stream) {
    clear_error();
    if (ndim != 2) return fail(ITKG_ERR_UNSUPPORTED, "Canny takes 2-D images");
    if (!args_ok(ndim, dtype, 1)) return fail(ITKG_ERR_INVALID_ARG, "bad dtype");
    if (!largest || !input || !output) return fail(ITKG_ERR_INVALID_ARG, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const itkg_region& L = *largest;
    if (!region_ok(2, L) || !region_ok(2, output->region))
        return fail(ITKG_ERR_INVALID_ARG, "negative region size");
    if (region_len(2, output->region) > 0 && !contains(2, L, output->region))
        return fail(ITKG_ERR_INVALID_ARG, "requested region is not inside the largest region");
    int rc = check_view(2, input, L, "input");
    if (rc) return rc;
    if ((rc = check_view(2, output, output->region, "output"))) return rc;
    const int64_t n = region_len(2, L);
    if (n == 0 || region_len(2, output->region) == 0) return ITKG_OK;
    if (L.size[1] > 65535 || output->region.size[1] > 65535)
        return fail(ITKG_ERR_UNSUPPORTED, "Canny: more than 65535 rows");
    // 1. smoothing (edges.rs:563-577): the index-space kernel over the whole
    // image, which is the separable path's whole-image request
    const uint64_t r = radius_of(sigma, 1.0);
    if (r > (uint64_t)ITKG_MAX_RADIUS) return fail(ITKG_ERR_UNSUPPORTED, "radius exceeds %d", ITKG_MAX_RADIUS);
    double w[2 * kMaxTaps];
    index_weights(sigma, w, kMaxTaps);
    index_weights(sigma, w + 2 * r + 1, kMaxTaps);
    const uint64_t radii[2] = {r, r};
    Plan p;
    if ((rc = make_plan_weights(2, radii, w, largest, L, &p, false))) return rc;
    const size_t px = dtype_size(dtype);
    // the scratch buffers (smoothed image, weak and strong edge maps, the
    // hysteresis flags) carved from one allocation: each stream-ordered
    // free costs the host about 0.1 ms, more than this filter's kernels at
    // moderate sizes
    auto up = [](size_t bytes) { return (bytes + 255) & ~(size_t)255; };
    const size_t o_weak = up((size_t)n * px), o_edges = o_weak + up((size_t)n);
    const size_t o_flag = o_edges + up((size_t)n), total_bytes = o_flag + 256;
    void* block = nullptr;
    void *sm = nullptr, *weak = nullptr, *edges = nullptr, *flag = nullptr;
    hipError_t e = pool_alloc(&block, total_bytes, s);
    if (e != hipSuccess) {
        rc = hip_fail(e, "pool_alloc(Canny scratch)");
    } else {
        char* base = static_cast<char*>(block);
        sm = base;
        weak = base + o_weak;
        edges = base + o_edges;
        flag = base + o_flag;
    }
    const itkg_view smv = dense_view(2, sm, L);
    if (!rc) rc = forward_impl(p, dtype, 1, *input, smv, arith, s, nullptr);
    // 2-4. gradient, suppression, hysteresis, the request as 1 / 0
    if (!rc) {
        auto run = [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            CannyArgs<T> a{};
            a.smoothed = static_cast<const T*>(sm);
            a.weak = static_cast<uint8_t*>(weak);
            a.edges = static_cast<uint8_t*>(edges);
            a.changed = static_cast<int*>(flag);
            a.out = static_cast<T*>(output->data);
            a.nx = L.size[0];
            a.ny = L.size[1];
            a.ox = output->region.index[0] - L.index[0];
            a.oy = output->region.index[1] - L.index[1];
            a.mx = output->region.size[0];
            a.my = output->region.size[1];
            a.out_stride[0] = output->stride[0];
            a.out_stride[1] = output->stride[1];
            a.lower = lower;
            a.upper = upper;
            return launch_canny<T>(a, s);
        };
        e = dtype == ITKG_F64 ? run((double*)nullptr) : run((float*)nullptr);
        if (e != hipSuccess) rc = hip_fail(e, "Canny launch");
    }
    if (block) (void)hipFreeAsync(block, s);
    return rc;
}

int itkg_canny_generate_region_host(int32_t ndim, itkg_dtype dtype, double sigma, double lower,
                                    double upper, const itkg_region* largest, const void* input,
                                    const itkg_region* requested, void* output, itkg_arith arith) {
    clear_error();
    if (ndim != 2) return fail(ITKG_ERR_UNSUPPORTED, "Canny takes 2-D images");
    if (!args_ok(ndim, dtype, 1)) return fail(ITKG_ERR_INVALID_ARG, "bad dtype");
    if (!largest || !requested) return fail(ITKG_ERR_INVALID_ARG, "null region");
    if (!region_ok(2, *largest) || !region_ok(2, *requested))
        return fail(ITKG_ERR_INVALID_ARG, "negative region size");
    const int64_t nin = region_len(2, *largest), nout = region_len(2, *requested);
    if (nout == 0) return ITKG_OK;
    if (!input || !output) return fail(ITKG_ERR_INVALID_ARG, "null host buffer");
    const size_t px = dtype_size(dtype);
    hipStream_t s = nullptr;
    hipError_t e = acquire_host_stream(&s);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    int rc = ITKG_OK;
    void* block = nullptr;  // input and output staging in one allocation
    const size_t o_out = ((size_t)nin * px + 255) & ~(size_t)255;
    e = pool_alloc(&block, o_out + (size_t)nout * px, s);
    void* din = block;
    void* dout = block ? static_cast<char*>(block) + o_out : nullptr;
    if (e == hipSuccess) e = hipMemcpyAsync(din, input, (size_t)nin * px, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) rc = hip_fail(e, "Canny host staging");
    if (!rc) {
        const itkg_view vin = dense_view(2, din, *largest);
        const itkg_view vout = dense_view(2, dout, *requested);
        rc = itkg_canny_generate_region(2, dtype, sigma, lower, upper, largest, &vin, &vout, arith, s);
    }
    if (!rc) {
        e = hipMemcpyAsync(output, dout, (size_t)nout * px, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) rc = hip_fail(e, "copy out");
    }
    if (block) (void)hipFreeAsync(block, s);
    e = hipStreamSynchronize(s);
    if (e != hipSuccess && !rc) rc = hip_fail(e, "hipStreamSynchronize");
    release_host_stream(s);
    return rc;
}

int itkg_gaussian_backward(int32_t ndim, itkg_dtype dtype, int32_t ncomp, double sigma,
                           const double* spacing, const itkg_region* largest,
                           const itkg_view* grad_output, const itkg_view* grad_input,
                           int32_t accumulate, void* stream) {
    clear_error();
    if (!args_ok(ndim, dtype, ncomp)) return fail(ITKG_ERR_INVALID_ARG, "bad ndim/dtype/ncomp");
    if (!grad_output) return fail(ITKG_ERR_INVALID_ARG, "null grad_output view");
    Plan p;
    int rc = make_plan(ndim, sigma, spacing, largest, grad_output->region, &p);
    if (rc) return rc;
    if ((rc = check_view(ndim, grad_output, p.req, "grad_output"))) return rc;
    if ((rc = check_view(ndim, grad_input, p.in_reg, "grad_input"))) return rc;
    return backward_impl(p, dtype, ncomp, *grad_output, *grad_input, accumulate,
                         static_cast<hipStream_t>(stream), nullptr);
}

int itkg_canny_direction_bins(const double* gx, const double* gy, int64_t n, uint8_t* bins,
                              void* stream) {
    clear_error();
    if (n < 0) return fail(ITKG_ERR_INVALID_ARG, "negative count");
    if (n == 0) return ITKG_OK;
    if (!gx || !gy || !bins) return fail(ITKG_ERR_INVALID_ARG, "null buffer");
    const hipError_t e = launch_canny_bins(gx, gy, bins, n, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? ITKG_OK : hip_fail(e, "canny bins launch");
}

int itkg_gaussian_backward_host(int32_t ndim, itkg_dtype dtype, int32_t ncomp, double sigma,
                                const double* spacing, const itkg_region* largest,
                                const void* grad_output, const itkg_region* requested,
                                void* grad_input, const itkg_region* grad_input_region,
                                int32_t accumulate) {
    clear_error();
    if (!args_ok(ndim, dtype, ncomp)) return fail(ITKG_ERR_INVALID_ARG, "bad ndim/dtype/ncomp");
    if (!requested || !grad_input_region) return fail(ITKG_ERR_INVALID_ARG, "null region");
    Plan p;
    int rc = make_plan(ndim, sigma, spacing, largest, *requested, &p);
    if (rc) return rc;
    itkg_region greg = *grad_input_region;
    for (int d = ndim; d < ITKG_MAX_DIM; ++d) {
        greg.index[d] = 0;
        greg.size[d] = 1;
    }
    if (!region_ok(ndim, greg)) return fail(ITKG_ERR_INVALID_ARG, "grad_input region: negative size");
    const int64_t nout = region_len(ndim, p.req), ngin = region_len(ndim, greg);
    if (nout == 0) return ITKG_OK;
    if (!contains(ndim, greg, p.in_reg))
        return fail(ITKG_ERR_INVALID_ARG,
                    "grad_input region does not cover input_region_for_output(requested)");
    if (!grad_output || !grad_input) return fail(ITKG_ERR_INVALID_ARG, "null host buffer");
    const size_t px = dtype_size(dtype) * ncomp;
    hipStream_t s = nullptr;
    hipError_t e = acquire_host_stream(&s);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    void *dgo = nullptr, *dgi = nullptr;
    e = pool_alloc(&dgo, (size_t)nout * px, s);
    if (e == hipSuccess) e = pool_alloc(&dgi, (size_t)ngin * px, s);
    if (e != hipSuccess) rc = hip_fail(e, "pool_alloc(backward staging)");
    // the grad_input buffer goes in whole when the kernels add to it or when
    // it is larger than the part they write (the rest comes back unchanged)
    const bool whole = accumulate || !same_region(ndim, greg, p.in_reg);
    if (!rc) {
        e = hipMemcpyAsync(dgo, grad_output, (size_t)nout * px, hipMemcpyHostToDevice, s);
        if (e == hipSuccess && whole)
            e = hipMemcpyAsync(dgi, grad_input, (size_t)ngin * px, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) rc = hip_fail(e, "copy in");
    }
    if (!rc) {
        const itkg_view vgo = dense_view(ndim, dgo, p.req), vgi = dense_view(ndim, dgi, greg);
        rc = backward_impl(p, dtype, ncomp, vgo, vgi, accumulate ? 1 : 0, s, nullptr);
    }
    if (!rc) {
        e = hipMemcpyAsync(grad_input, dgi, (size_t)ngin * px, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) rc = hip_fail(e, "copy out");
    }
    if (dgo) (void)hipFreeAsync(dgo, s);
    if (dgi) (void)hipFreeAsync(dgi, s);
    e = hipStreamSynchronize(s);
    if (e != hipSuccess && !rc) rc = hip_fail(e, "hipStreamSynchronize");
    release_host_stream(s);
    return rc;
}

int itkg_convert_pixels(int32_t ndim, itkg_ptype in_type, itkg_ptype out_type, int32_t ncomp,
                        const itkg_view* input, const itkg_view* output, itkg_convert_mode mode,
                        void* stream) {
    clear_error();
    if (ndim < 1 || ndim > ITKG_MAX_DIM || ncomp < 1 || ncomp > 4 || !ptype_ok(in_type) ||
        !ptype_ok(out_type) || (mode != ITKG_CONVERT_AS && mode != ITKG_CONVERT_ROUND))
        return fail(ITKG_ERR_INVALID_ARG, "bad ndim/type/ncomp/mode");
    if (!output) return fail(ITKG_ERR_INVALID_ARG, "null output view");
    int rc = check_view(ndim, output, output->region, "output");
    if (!rc) rc = check_view(ndim, input, output->region, "input");
    if (rc) return rc;
    return convert_views(ndim, in_type, out_type, ncomp, *input, *output, output->region, mode,
                         static_cast<hipStream_t>(stream));
}

int itkg_cast_gaussian_cast_host(int32_t ndim, itkg_ptype in_type, itkg_ptype out_type,
                                 double sigma, const double* spacing, const itkg_region* largest,
                                 const void* input, const itkg_region* requested, void* output,
                                 itkg_arith arith, itkg_convert_mode out_mode) {
    clear_error();
    if (!args_ok(ndim, ITKG_F32, 1) || !ptype_ok(in_type) || !ptype_ok(out_type) ||
        (out_mode != ITKG_CONVERT_AS && out_mode != ITKG_CONVERT_ROUND))
        return fail(ITKG_ERR_INVALID_ARG, "bad ndim/type/mode");
    if (!requested) return fail(ITKG_ERR_INVALID_ARG, "null requested region");
    Plan p;
    int rc = make_plan(ndim, sigma, spacing, largest, *requested, &p);
    if (rc) return rc;
    const int64_t nin = region_len(ndim, p.in_reg), nout = region_len(ndim, p.req);
    if (nout == 0) return ITKG_OK;
    if (!input || !output) return fail(ITKG_ERR_INVALID_ARG, "null buffer");
    const size_t pin = ptype_size(in_type), pout = ptype_size(out_type);
    HostPieces hp;
    if ((rc = host_pieces(p, pin, pout, &hp))) return rc;
    const int K = (int)hp.piece.size();
    hipStream_t s = nullptr;
    hipError_t e = acquire_host_stream(&s);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    // one block: input pixels, f32 input, f32 output, output pixels
    const size_t bi = (size_t)nin * pin, bf = (size_t)nin * 4, bo = (size_t)nout * 4,
                 bq = (size_t)nout * pout;
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    void* block = nullptr;
    e = pool_alloc(&block, up(bi) + up(bf) + up(bo) + up(bq), s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        if (block) (void)hipFreeAsync(block, s);
        (void)hipStreamSynchronize(s);
        release_host_stream(s);
        return hip_fail(e, "pool_alloc(cast staging)");
    }
    char* base = static_cast<char*>(block);
    void* din = base;
    void* dfi = base + up(bi);
    void* dfo = base + up(bi) + up(bf);
    void* dq = base + up(bi) + up(bf) + up(bo);
    const itkg_view vin = dense_view(ndim, din, p.in_reg), vfi = dense_view(ndim, dfi, p.in_reg);
    const itkg_view vfo = dense_view(ndim, dfo, p.req), vq = dense_view(ndim, dq, p.req);
    StagedJob job;
    job.host_in = input;
    job.in_bytes = bi;
    job.host_out = output;
    job.out_bytes = bq;
    job.dev_in = din;
    job.dev_out = dq;
    job.pieces = K;
    for (int k = 0; k < K; ++k) {
        job.in_need.push_back((size_t)(hp.in_planes_end[k] * hp.in_plane) * pin);
        job.out_end.push_back((size_t)(hp.out_planes_end[k] * hp.out_plane) * pout);
    }
    const int ax = ndim - 1;
    int64_t converted = 0;  // input planes already converted to f32
    job.compute = [&](int k, hipStream_t st) {
        const Plan& q = hp.piece[k];
        // the f32 input planes this piece reads that no earlier piece converted
        itkg_region cr = p.in_reg;
        cr.index[ax] = p.in_reg.index[ax] + converted;
        cr.size[ax] = hp.in_planes_end[k] - converted;
        int r = ITKG_OK;
        if (cr.size[ax] > 0) {
            r = convert_views(ndim, in_type, ITKG_PT_F32, 1, vin, vfi, cr, ITKG_CONVERT_AS, st);
            converted = hp.in_planes_end[k];
        }
        if (!r) r = forward_impl(q, ITKG_F32, 1, vfi, vfo, arith, st, nullptr);
        if (!r) r = convert_views(ndim, ITKG_PT_F32, out_type, 1, vfo, vq, q.req, out_mode, st);
        return r;
    };
    StageStats stats;
    job.stats = &stats;
    rc = run_staged(job, s);
    note_stage(stats);
    (void)hipFreeAsync(block, s);
    e = hipStreamSynchronize(s);
    if (e != hipSuccess && !rc) rc = hip_fail(e, "hipStreamSynchronize");
    release_host_stream(s);
    return rc;
}

int itkg_trim_scratch(uint64_t keep_bytes) {
    clear_error();
    int rc = trim_pool(keep_bytes);
    if (!rc) rc = stage_trim();
    return rc;
}

int itkg_host_staging_info(int32_t* last_pinned) {
    clear_error();
    if (!last_pinned) return fail(ITKG_ERR_INVALID_ARG, "null output");
    std::lock_guard<std::mutex> lk(g_last_stage_mu);
    *last_pinned = g_last_stage.pinned ? 1 : 0;
    return ITKG_OK;
}

const char* itkg_last_error(void) { return g_err.c_str(); }

int itkg_set_path(itkg_path path) {
    clear_error();
    if (path != ITKG_PATH_AUTO && path != ITKG_PATH_PER_AXIS && path != ITKG_PATH_FUSED)
        return fail(ITKG_ERR_INVALID_ARG, "bad path");
    g_path.store(path);
    return ITKG_OK;
}

const char* itkg_version(void) { return "itkg 0.1 (gfx950)"; }

}  // extern "C"

