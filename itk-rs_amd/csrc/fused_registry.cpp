// Lookup of the fused kernel instantiations compiled into the library.
#include "itkg_impl.hpp"

namespace itkg {

const FusedEntry* fused_table_iso_a(int* count);
const FusedEntry* fused_table_iso_b(int* count);
const FusedEntry* fused_table_iso_c(int* count);
const FusedEntry* fused_table_plane(int* count);

const FusedEntry* fused_entry(int rxy, int rz) {
    const FusedEntry* (*tables[])(int*) = {fused_table_iso_a, fused_table_iso_b, fused_table_iso_c,
                                          fused_table_plane};
    for (auto get : tables) {
        int n = 0;
        const FusedEntry* t = get(&n);
        for (int i = 0; i < n; ++i)
            if (t[i].rxy == rxy && t[i].rz == rz) return &t[i];
    }
    return nullptr;
}

}  // namespace itkg
