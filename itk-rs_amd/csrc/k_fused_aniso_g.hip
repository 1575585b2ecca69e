// Fused kernel instantiations for anisotropic spacing: x/y radius 1..5
// with z radius 1 to 4 above it (slices finer than the in-plane spacing,
// z radius at most 15), plain launches.
#include "fused3d.hpp"

namespace itkg {
// Host-side table (the device pass compiles no host function bodies).
const FusedEntry* fused_table_aniso_g(int* count) {
    static const FusedEntry table[] = {
    make_entry_single<1, 2>(), make_entry_single<1, 3>(), make_entry_single<1, 4>(),
    make_entry_single<1, 5>(), make_entry_single<2, 3>(), make_entry_single<2, 4>(),
    make_entry_single<2, 5>(), make_entry_single<2, 6>(), make_entry_single<3, 4>(),
    make_entry_single<3, 5>(), make_entry_single<3, 6>(), make_entry_single<3, 7>(),
    make_entry_single<4, 5>(), make_entry_single<4, 6>(), make_entry_single<4, 7>(),
    make_entry_single<4, 8>(), make_entry_single<5, 6>(), make_entry_single<5, 7>(),
    make_entry_single<5, 8>(), make_entry_single<5, 9>()};
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}
}  // namespace itkg
