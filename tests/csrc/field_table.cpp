// C wrapper around the library's L2T/M2T field gather table (elfmm_ops.cpp)
// so the tests can check it against finite differences without a GPU.
#include <stdexcept>
#include <vector>

#include "../../elfmmlib3d_amd/csrc/elfmm_internal.h"

// Fills idx[n][3], c[n][3] for the n = (ncf rounded up to even) * 2 * 9
// entries; returns n, or -1 if the table has more than nmax entries or
// could not be built.
extern "C" int field_table(int p, int irr, int *idx, double *c, int nmax)
{
    std::vector<elfmm::FieldTerm> tab;
    try {
        elfmm::build_field_table(p, irr != 0, tab);
    } catch (const std::exception &) {
        return -1;
    }
    if ((int)tab.size() > nmax) return -1;
    for (size_t e = 0; e < tab.size(); e++)
        for (int t = 0; t < 3; t++) {
            idx[3 * e + t] = tab[e].idx[t];
            c[3 * e + t] = tab[e].c[t];
        }
    return (int)tab.size();
}
