// The single persistent launch of a batched plan, forward FAST (radii 1..24; 32-row tiles also for radii 1..7).
#include "fused3d.hpp"

namespace itkg {

hipError_t mega_prepare_fast(int lds_bytes) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&fused3d_mega_kernel<FUSED_FWD_FAST>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
}

hipError_t mega_launch_fast(const Fused3dArgsB* vols, const int4* work, const int* qstart,
                           int* heads, int grid, int lds_bytes, hipStream_t s,
                           unsigned long long* trace) {
    hipLaunchKernelGGL((fused3d_mega_kernel<FUSED_FWD_FAST>), dim3(grid), dim3(1024), (size_t)lds_bytes, s, vols,
                       work, qstart, heads, trace);
    return hipGetLastError();
}

int mega_occupancy_fast(int lds_bytes) {
    int blocks = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &blocks, reinterpret_cast<const void*>(&fused3d_mega_kernel<FUSED_FWD_FAST>), 1024, (size_t)lds_bytes);
    return e == hipSuccess && blocks > 0 ? blocks : 1;
}

// Host-side table (the device pass compiles no host function bodies).
void mega_classes_fast(int* lds, int* ty) {
    mega_table<FUSED_FWD_FAST>(std::make_integer_sequence<int, kMegaMaxR + 7>{}, lds, ty);
}

}  // namespace itkg
