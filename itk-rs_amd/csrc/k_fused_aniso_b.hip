// Fused kernel instantiations for anisotropic spacing: x/y radius 7..8, every
// z radius below it (slices thicker than the in-plane spacing), plain launches.
#include "fused3d.hpp"

namespace itkg {
// Host-side table (the device pass compiles no host function bodies).
const FusedEntry* fused_table_aniso_b(int* count) {
    static const FusedEntry table[] = {
    make_entry_single<7, 1>(), make_entry_single<7, 2>(), make_entry_single<7, 3>(),
    make_entry_single<7, 4>(), make_entry_single<7, 5>(), make_entry_single<7, 6>(),
    make_entry_single<8, 1>(), make_entry_single<8, 2>(), make_entry_single<8, 3>(),
    make_entry_single<8, 4>(), make_entry_single<8, 5>(), make_entry_single<8, 6>(),
    make_entry_single<8, 7>()};
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}
}  // namespace itkg
