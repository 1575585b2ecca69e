// Fused kernel instantiations for anisotropic spacing: x/y radius 11, every
// z radius below it (slices thicker than the in-plane spacing), plain launches.
#include "fused3d.hpp"

namespace itkg {
// Host-side table (the device pass compiles no host function bodies).
const FusedEntry* fused_table_aniso_e(int* count) {
    static const FusedEntry table[] = {
    make_entry_single<11, 1>(), make_entry_single<11, 2>(), make_entry_single<11, 3>(),
    make_entry_single<11, 4>(), make_entry_single<11, 5>(), make_entry_single<11, 6>(),
    make_entry_single<11, 7>(), make_entry_single<11, 8>(), make_entry_single<11, 9>(),
    make_entry_single<11, 10>()};
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}
}  // namespace itkg
