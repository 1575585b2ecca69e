// Fused kernel instantiations: isotropic radii 7..11.
#include "fused3d.hpp"

namespace itkg {
// Host-side table (the device pass compiles no host function bodies).
const FusedEntry* fused_table_iso_b(int* count) {
    static const FusedEntry table[] = {make_entry<7, 7>(), make_entry<8, 8>(), make_entry<9, 9>(),
                                        make_entry<10, 10>(), make_entry<11, 11>()};
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}
}  // namespace itkg
