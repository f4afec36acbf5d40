// C entry points over the reference's own InterpolationScheme, compiled from the
// reference sources where they lie (see oracle/ref.mk).  TEST INFRASTRUCTURE ONLY.
#include <cstdio>
#include "interpolation_scheme.h"

static InterpolationScheme* g_scheme = nullptr;

extern "C" int ref_efield_init(const char* folder, const char* fields, const char* triangles,
                               float r0, float E0) {
    delete g_scheme;
    g_scheme = new InterpolationScheme(BackgroundEfield::singleEmitter);
    g_scheme->r0 = r0;
    g_scheme->E0 = E0;
    g_scheme->initializeInterpolationScheme(folder, fields, triangles);
    return 0;
}

extern "C" void ref_efield(float x, float y, float z, float* out) {
    vec3<float> e = g_scheme->Efield(x, y, z);
    out[0] = e.x;
    out[1] = e.y;
    out[2] = e.z;
}
