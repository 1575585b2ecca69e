// Fused kernel instantiations for anisotropic spacing: x/y radius 9, every
// z radius below it (slices thicker than the in-plane spacing), plain launches.
#include "fused3d.hpp"

namespace itkg {
// Host-side table (the device pass compiles no host function bodies).
const FusedEntry* fused_table_aniso_c(int* count) {
    static const FusedEntry table[] = {
    make_entry_single<9, 1>(), make_entry_single<9, 2>(), make_entry_single<9, 3>(),
    make_entry_single<9, 4>(), make_entry_single<9, 5>(), make_entry_single<9, 6>(),
    make_entry_single<9, 7>(), make_entry_single<9, 8>()};
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}
}  // namespace itkg
