// Lookup of the fused kernel instantiations compiled into the library.
#include "itkg_impl.hpp"

namespace itkg {

const FusedEntry* fused_table_iso_a(int* count);
const FusedEntry* fused_table_iso_b(int* count);
const FusedEntry* fused_table_iso_c(int* count);
const FusedEntry* fused_table_plane(int* count);
const FusedEntry* fused_table_aniso_a(int* count);
const FusedEntry* fused_table_aniso_b(int* count);
const FusedEntry* fused_table_aniso_c(int* count);
const FusedEntry* fused_table_aniso_d(int* count);
const FusedEntry* fused_table_aniso_e(int* count);
const FusedEntry* fused_table_aniso_f(int* count);
const FusedEntry* fused_table_aniso_g(int* count);
const FusedEntry* fused_table_aniso_h(int* count);
const FusedEntry* fused_table_aniso_i(int* count);

const FusedEntry* fused_entry(int rxy, int rz) {
    const FusedEntry* (*tables[])(int*) = {
        fused_table_iso_a,   fused_table_iso_b,   fused_table_iso_c,   fused_table_plane,
        fused_table_aniso_a, fused_table_aniso_b, fused_table_aniso_c, fused_table_aniso_d,
        fused_table_aniso_e, fused_table_aniso_f, fused_table_aniso_g, fused_table_aniso_h,
        fused_table_aniso_i};
    for (auto get : tables) {
        int n = 0;
        const FusedEntry* t = get(&n);
        for (int i = 0; i < n; ++i)
            if (t[i].rxy == rxy && t[i].rz == rz) return &t[i];
    }
    return nullptr;
}

}  // namespace itkg
