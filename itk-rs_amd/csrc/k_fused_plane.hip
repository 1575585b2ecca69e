// Fused kernel instantiations without a persistent form:
//   - 2-D images (z radius 0, one plane): x/y radii 1..16;
//   - the anisotropic radii (13, 13, 4) of sigma 3 at spacing 0.7/0.7/2.5.
#include "fused3d.hpp"

namespace itkg {
// Host-side table (the device pass compiles no host function bodies).
const FusedEntry* fused_table_plane(int* count) {
    static const FusedEntry table[] = {
    make_entry_single<1, 0>(),  make_entry_single<2, 0>(),  make_entry_single<3, 0>(),
    make_entry_single<4, 0>(),  make_entry_single<5, 0>(),  make_entry_single<6, 0>(),
    make_entry_single<7, 0>(),  make_entry_single<8, 0>(),  make_entry_single<9, 0>(),
    make_entry_single<10, 0>(), make_entry_single<11, 0>(), make_entry_single<12, 0>(),
    make_entry_single<13, 0>(), make_entry_single<14, 0>(), make_entry_single<15, 0>(),
    make_entry_single<16, 0>(), make_entry_single<13, 4>()};
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}
}  // namespace itkg
