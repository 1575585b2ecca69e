// Fused kernel instantiations: isotropic radii 1..6 (forward, adjoint, persistent).
#include "fused3d.hpp"

namespace itkg {
// Host-side table (the device pass compiles no host function bodies).
const FusedEntry* fused_table_iso_a(int* count) {
    static const FusedEntry table[] = {make_entry<1, 1>(), make_entry<2, 2>(), make_entry<3, 3>(),
                                        make_entry<4, 4>(), make_entry<5, 5>(), make_entry<6, 6>()};
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}
}  // namespace itkg
