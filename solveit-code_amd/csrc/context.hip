// Handle lifetime, error reporting, workspace and reference constants.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "shp_internal.h"

namespace shp {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// K_CONST = (e*e) / (4*PI*eps0*amu): e*e is a float product, the denominator
// is long double because PI is (constants.h:29,46).
double k_const_host() {
    const long double pi = 3.14159265358979323846264338328L;
    const float e = 1.6021766e-19f, amu = 1.66053904e-27f, eps0 = 8.85419e-12f;
    float e2 = e * e;
    long double den = 4 * pi * eps0 * amu;
    return (double)((long double)e2 / den);
}

}  // namespace shp

void* shp_context::get(int slot, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (ws_bytes[slot] >= bytes) return ws[slot];
    if (ws[slot]) {
        (void)hipStreamSynchronize(stream);
        (void)hipFree(ws[slot]);
        ws[slot] = nullptr;
        ws_bytes[slot] = 0;
    }
    // Headroom: the FMM's box and unit counts grow from call to call as a
    // plume expands, and each regrowth costs a stream synchronisation and a
    // fresh allocation in the middle of an evaluation.
    size_t want = bytes + bytes / 2;
    hipError_t e = hipMalloc(&ws[slot], want);
    if (e != hipSuccess) {
        ws[slot] = nullptr;
        shp::set_error("workspace allocation of %zu bytes failed: %s", want, hipGetErrorString(e));
        return nullptr;
    }
    ws_bytes[slot] = want;
    return ws[slot];
}

extern "C" {

int shp_version(void) { return 1; }

const char* shp_last_error(void) { return shp::g_err; }

int shp_create(int device, shp_handle* out) {
    SHP_REQUIRE(out != nullptr, "shp_create: null output pointer");
    int ndev = 0;
    SHP_HIP_CHECK(hipGetDeviceCount(&ndev));
    SHP_REQUIRE(device >= 0 && device < ndev, "shp_create: device %d not present (%d devices)",
                device, ndev);
    SHP_HIP_CHECK(hipSetDevice(device));
    shp_context* h = new shp_context();
    h->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) h->num_cus = prop.multiProcessorCount;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&h->ev0);
    if (e == hipSuccess) e = hipEventCreate(&h->ev1);
    if (e == hipSuccess) e = hipEventCreate(&h->ev2);
    if (e == hipSuccess) e = hipEventCreate(&h->ev3);
    for (int k = 0; k < 8 && e == hipSuccess; k++) e = hipEventCreate(&h->fev[k]);
    if (e == hipSuccess) e = hipMalloc((void**)&h->fmm_stats, 8 * sizeof(unsigned long long));
    if (e != hipSuccess) {
        shp::set_error("shp_create: %s", hipGetErrorString(e));
        shp_destroy(h);
        return SHP_ERR_HIP;
    }
    const int rc = shp::direct_scratch_setup(device);  // per-device scratch of shp_direct_dev_noalloc
    if (rc != SHP_OK) {
        shp_destroy(h);
        return rc;
    }
    *out = h;
    return SHP_OK;
}

int shp_destroy(shp_handle h) {
    if (!h) return SHP_OK;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    shp::writer_release(h);
    for (int i = 0; i < shp::WS_NSLOTS; i++)
        if (h->ws[i]) (void)hipFree(h->ws[i]);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->ev2) (void)hipEventDestroy(h->ev2);
    if (h->ev3) (void)hipEventDestroy(h->ev3);
    for (int k = 0; k < 8; k++)
        if (h->fev[k]) (void)hipEventDestroy(h->fev[k]);
    if (h->fmm_stats) (void)hipFree(h->fmm_stats);
    if (h->res_f64) (void)hipFree(h->res_f64);
    if (h->res_i32) (void)hipFree(h->res_i32);
    if (h->res_aos) (void)hipFree(h->res_aos);
    if (h->res_level) (void)hipFree(h->res_level);
    if (h->res_e2) (void)hipFree(h->res_e2);
    if (h->res_ev) (void)hipFree(h->res_ev);
    if (h->res_evkey) (void)hipFree(h->res_evkey);
    if (h->res_evn) (void)hipFree(h->res_evn);
    shp::free_mesh(h);
    shp::fmm_release(h);
    shp::sph_release(h);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SHP_OK;
}

int shp_sync(shp_handle h) {
    SHP_REQUIRE(h != nullptr, "shp_sync: null handle");
    SHP_HIP_CHECK(hipStreamSynchronize(h->stream));
    return SHP_OK;
}

void* shp_stream(shp_handle h) { return h ? (void*)h->stream : nullptr; }

double shp_k_const(void) { return shp::k_const_host(); }
double shp_softening(void) { return shp::kSoftening; }
double shp_species_mass(int species) { return shp::species_mass(species); }

double shp_last_update_ms(shp_handle h) {
    if (!h || !h->timed_update) return -1.0;
    float ms = -1.0f;
    if (hipEventSynchronize(h->ev3) != hipSuccess) return -1.0;
    if (hipEventElapsedTime(&ms, h->ev2, h->ev3) != hipSuccess) return -1.0;
    return (double)ms;
}

double shp_last_kernel_ms(shp_handle h) {
    if (!h || !h->timed) return -1.0;
    float ms = -1.0f;
    if (hipEventSynchronize(h->ev1) != hipSuccess) return -1.0;
    if (hipEventElapsedTime(&ms, h->ev0, h->ev1) != hipSuccess) return -1.0;
    return (double)ms;
}

}  // extern "C"
