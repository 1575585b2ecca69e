// Fused kernel instantiations for anisotropic spacing: x/y radius 10, every
// z radius below it (slices thicker than the in-plane spacing), plain launches.
#include "fused3d.hpp"

namespace itkg {
// Host-side table (the device pass compiles no host function bodies).
const FusedEntry* fused_table_aniso_d(int* count) {
    static const FusedEntry table[] = {
    make_entry_single<10, 1>(), make_entry_single<10, 2>(), make_entry_single<10, 3>(),
    make_entry_single<10, 4>(), make_entry_single<10, 5>(), make_entry_single<10, 6>(),
    make_entry_single<10, 7>(), make_entry_single<10, 8>(), make_entry_single<10, 9>()};
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}
}  // namespace itkg
