// geometry.cpp -- host construction of the analytic rigid objects; the
// distance, closest-point and inside queries live in geom.h (shared with the
// device kernels of moving objects).
#include <cmath>

#include "engine.h"

namespace fdtd {

Geom make_geom(const fdtd_object &o)
{
    Geom g;
    g.type = o.type;
    for (int i = 0; i < 3; ++i) { g.c[i] = o.center[i]; g.a[i] = o.semi_axes[i]; }
    for (int i = 0; i < 9; ++i) g.R[i] = o.rotation[i];
    if (g.type == FDTD_OBJ_SPHERE) g.a[1] = g.a[2] = g.a[0];
    return g;
}

void geom_aabb(const Geom &g, double *mn, double *mx)
{
    for (int i = 0; i < 3; ++i) {
        double ext = 0;
        if (g.type == FDTD_OBJ_SPHERE) ext = g.a[0];
        else {
            for (int j = 0; j < 3; ++j) ext += (g.R[3 * i + j] * g.a[j]) * (g.R[3 * i + j] * g.a[j]);
            ext = std::sqrt(ext);
        }
        mn[i] = g.c[i] - ext;
        mx[i] = g.c[i] + ext;
    }
}

}  // namespace fdtd
