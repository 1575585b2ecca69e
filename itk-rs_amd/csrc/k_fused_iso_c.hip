// Fused kernel instantiations: isotropic radii 12..16.
#include "fused3d.hpp"

namespace itkg {
// Host-side table (the device pass compiles no host function bodies).
const FusedEntry* fused_table_iso_c(int* count) {
    static const FusedEntry table[] = {make_entry<12, 12>(), make_entry<13, 13>(),
                                        make_entry<14, 14>(), make_entry<15, 15>()};
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}
}  // namespace itkg
