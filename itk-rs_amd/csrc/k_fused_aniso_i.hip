// Fused kernel instantiations for anisotropic spacing: x/y radius 9..11
// with z radius 1 to 4 above it (slices finer than the in-plane spacing,
// z radius at most 15), plain launches.
#include "fused3d.hpp"

namespace itkg {
// Host-side table (the device pass compiles no host function bodies).
const FusedEntry* fused_table_aniso_i(int* count) {
    static const FusedEntry table[] = {
    make_entry_single<9, 10>(), make_entry_single<9, 11>(), make_entry_single<9, 12>(),
    make_entry_single<9, 13>(), make_entry_single<10, 11>(), make_entry_single<10, 12>(),
    make_entry_single<10, 13>(), make_entry_single<10, 14>(), make_entry_single<11, 12>(),
    make_entry_single<11, 13>(), make_entry_single<11, 14>(), make_entry_single<11, 15>()};
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}
}  // namespace itkg
