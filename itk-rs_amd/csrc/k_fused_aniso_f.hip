// Fused kernel instantiations for anisotropic spacing: x/y radius 12, every
// z radius below it (slices thicker than the in-plane spacing), plain launches.
#include "fused3d.hpp"

namespace itkg {
// Host-side table (the device pass compiles no host function bodies).
const FusedEntry* fused_table_aniso_f(int* count) {
    static const FusedEntry table[] = {
    make_entry_single<12, 1>(), make_entry_single<12, 2>(), make_entry_single<12, 3>(),
    make_entry_single<12, 4>(), make_entry_single<12, 5>(), make_entry_single<12, 6>(),
    make_entry_single<12, 7>(), make_entry_single<12, 8>(), make_entry_single<12, 9>(),
    make_entry_single<12, 10>(), make_entry_single<12, 11>()};
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}
}  // namespace itkg
