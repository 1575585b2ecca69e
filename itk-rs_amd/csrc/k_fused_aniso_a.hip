// Fused kernel instantiations for anisotropic spacing: x/y radius 2..6, every
// z radius below it (slices thicker than the in-plane spacing), plain launches.
#include "fused3d.hpp"

namespace itkg {
// Host-side table (the device pass compiles no host function bodies).
const FusedEntry* fused_table_aniso_a(int* count) {
    static const FusedEntry table[] = {
    make_entry_single<2, 1>(), make_entry_single<3, 1>(), make_entry_single<3, 2>(),
    make_entry_single<4, 1>(), make_entry_single<4, 2>(), make_entry_single<4, 3>(),
    make_entry_single<5, 1>(), make_entry_single<5, 2>(), make_entry_single<5, 3>(),
    make_entry_single<5, 4>(), make_entry_single<6, 1>(), make_entry_single<6, 2>(),
    make_entry_single<6, 3>(), make_entry_single<6, 4>(), make_entry_single<6, 5>()};
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}
}  // namespace itkg
