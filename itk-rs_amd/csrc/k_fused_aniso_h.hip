// Fused kernel instantiations for anisotropic spacing: x/y radius 6..8
// with z radius 1 to 4 above it (slices finer than the in-plane spacing,
// z radius at most 15), plain launches.
#include "fused3d.hpp"

namespace itkg {
// Host-side table (the device pass compiles no host function bodies).
const FusedEntry* fused_table_aniso_h(int* count) {
    static const FusedEntry table[] = {
    make_entry_single<6, 7>(), make_entry_single<6, 8>(), make_entry_single<6, 9>(),
    make_entry_single<6, 10>(), make_entry_single<7, 8>(), make_entry_single<7, 9>(),
    make_entry_single<7, 10>(), make_entry_single<7, 11>(), make_entry_single<8, 9>(),
    make_entry_single<8, 10>(), make_entry_single<8, 11>(), make_entry_single<8, 12>()};
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}
}  // namespace itkg
