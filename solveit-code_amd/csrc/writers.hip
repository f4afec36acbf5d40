// Output of the resident state in the reference's binary formats:
//  * state snapshots, Integrator::saveData and saveFullCurrentConditions
//    (integration_schemes.cpp:197-331), both SaveDataOptions layouts;
//  * boundary-condition events, ChargedParticles::writeToEventsFile
//    (particles.cpp:666-684);
//  * the Efield1 / Efield2 arrays that the snapshots carry
//    (integration_schemes.cpp:350-355, particles.cpp:508-512).
// A snapshot is packed on the GPU in the file's byte layout, copied to pinned
// host memory on a second stream and written by a host thread, so the copy
// and the write overlap the steps that follow.
#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "shp_internal.h"

namespace shp {

// PhysicalConstant::AMU / PhysicalConstant::ELEM_CHARGE: two float constants,
// so the quotient is a float (constants.h:33-34).
constexpr float kAmuOverE = 1.66053904e-27f / 1.6021766e-19f;

// The float nearest to hi + lo, where lo is the rounding error of hi (|lo| at
// most half an ulp of hi): (float)hi unless hi lies exactly halfway between
// two floats, where the sign of lo decides.
__device__ __forceinline__ float round_pair_to_float(double hi, double lo) {
    float f = (float)hi;
    const double d = hi - (double)f;  // exact
    if (lo != 0.0 && d != 0.0) {
        const float g = d > 0.0 ? nextafterf(f, INFINITY) : nextafterf(f, -INFINITY);
        if (d == ((double)g - (double)f) * 0.5 && ((lo > 0.0) == (d > 0.0))) f = g;
    }
    return f;
}

// Efield1 of one component (integration_schemes.cpp:352-354): factor * a *
// getMass / q, with a and q the reference's floats, factor * a a double and
// the rest in long double (getMass returns long double), rounded to float
// once.  A double times an integer mass below 2^11 is exact in long double;
// here that exact product is hi + lo.  Charges of +-1 divide exactly; other
// charges are divided in double.
__device__ __forceinline__ float efield1_of(double a, double q, double m) {
    const float af = (float)a, qf = (float)q;
    const double p = (double)kAmuOverE * (double)af;
    double hi = p * m;
    double lo = fma(p, m, -hi);
    if (qf == 1.0f) return round_pair_to_float(hi, lo);
    if (qf == -1.0f) return round_pair_to_float(-hi, -lo);
    return (float)((hi + lo) / (double)qf);
}

// Words of one snapshot record and of the snapshot header.
__host__ __device__ constexpr int save_rec_words(int format) { return format == SHP_SAVE_ORIGINAL ? 15 : 14; }
__host__ __device__ constexpr int save_hdr_words(int format) { return format == SHP_SAVE_ORIGINAL ? 0 : 3; }
// The packed buffer: word 0 the record count, words 1-3 unused, the file's
// bytes from word kSaveOff.
constexpr int kSaveOff = 4;

// 1 for ion first + k when it is written: status 1 (inbound) or 2
// (affectsAcceleration) for saveData, every ion for saveFullCurrentConditions.
// Entry m is 0, so the exclusive sums end with the record count.
__global__ void k_save_flags(int first, int m, const int* __restrict__ status, int all, int* __restrict__ flag) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > m) return;
    int f = 0;
    if (k < m) {
        const int s = status[first + k];
        f = all || s == 1 || s == 2;
    }
    flag[k] = f;
}

struct SaveArgs {
    int format, first, m, last;
    float time;
    int n;  // resident ions (row stride of e2)
    const double *x, *y, *z, *vx, *vy, *vz, *q, *ax, *ay, *az;
    const int* species;
    const float* e2;  // [3][n]
};

// Writes record pos[k] for every flagged ion first + k, in the layout of the
// format: original = x y z vx vy vz E1 E2 species index lastParticle;
// newOriginal = index species x y z vx vy vz E1 E2 after the header
// (0xffffffff, lastParticle, currentTime).
__global__ void k_save_pack(SaveArgs A, const int* __restrict__ flag, const int* __restrict__ pos,
                            unsigned int* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        out[0] = (unsigned int)pos[A.m];
        if (A.format == SHP_SAVE_NEW_ORIGINAL) {
            out[kSaveOff] = 0xffffffffu;
            out[kSaveOff + 1] = (unsigned int)A.last;
            out[kSaveOff + 2] = __float_as_uint(A.time);
        }
    }
    if (k >= A.m || !flag[k]) return;
    const int i = A.first + k;
    const int rw = save_rec_words(A.format);
    unsigned int* r = out + kSaveOff + save_hdr_words(A.format) + (size_t)pos[k] * rw;
    const double m = species_mass(A.species[i]);
    const float f[12] = {(float)A.x[i],
                         (float)A.y[i],
                         (float)A.z[i],
                         (float)A.vx[i],
                         (float)A.vy[i],
                         (float)A.vz[i],
                         efield1_of(A.ax[i], A.q[i], m),
                         efield1_of(A.ay[i], A.q[i], m),
                         efield1_of(A.az[i], A.q[i], m),
                         A.e2[i],
                         A.e2[A.n + i],
                         A.e2[2 * (size_t)A.n + i]};
    const unsigned int sp = (unsigned int)A.species[i];
    if (A.format == SHP_SAVE_ORIGINAL) {
        for (int c = 0; c < 12; c++) r[c] = __float_as_uint(f[c]);
        r[12] = sp;
        r[13] = (unsigned int)i;
        r[14] = (unsigned int)A.last;
    } else {
        r[0] = (unsigned int)i;
        r[1] = sp;
        for (int c = 0; c < 12; c++) r[2 + c] = __float_as_uint(f[c]);
    }
}

// Efield1 and Efield2 of every resident ion as the reference's vec3<float>
// arrays (x, y, z per ion).
__global__ void k_fields_aos(SaveArgs A, float* __restrict__ e1, float* __restrict__ e2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const double m = species_mass(A.species[i]);
    e1[3 * (size_t)i] = efield1_of(A.ax[i], A.q[i], m);
    e1[3 * (size_t)i + 1] = efield1_of(A.ay[i], A.q[i], m);
    e1[3 * (size_t)i + 2] = efield1_of(A.az[i], A.q[i], m);
    e2[3 * (size_t)i] = A.e2[i];
    e2[3 * (size_t)i + 1] = A.e2[A.n + i];
    e2[3 * (size_t)i + 2] = A.e2[2 * (size_t)A.n + i];
}

__global__ void k_wr_zero(int* __restrict__ p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// One snapshot in flight: its packed device buffer, its pinned copy, the
// event that ends the copy and where its bytes go.
struct SnapSlot {
    unsigned int* dev = nullptr;
    size_t dev_words = 0;
    unsigned int* host = nullptr;
    size_t host_words = 0;
    hipEvent_t packed = nullptr, copied = nullptr;
    FILE* out = nullptr;
    int hdr_words = 0, rec_words = 0;
    bool busy = false;  // queued or being written
};

struct StateWriter {
    int device = 0;
    hipStream_t copy = nullptr;
    SnapSlot slot[2];
    int next = 0;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> queue;
    bool stop = false;
    std::string err;  // first failure of the writer thread

    void run() {
        (void)hipSetDevice(device);
        for (;;) {
            int k;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !queue.empty(); });
                if (queue.empty()) return;
                k = queue.front();
            }
            SnapSlot& s = slot[k];
            std::string e;
            if (hipEventSynchronize(s.copied) != hipSuccess) {
                e = "snapshot copy failed";
            } else {
                const size_t nw = (size_t)s.hdr_words + (size_t)s.host[0] * (size_t)s.rec_words;
                if (nw && fwrite(s.host + kSaveOff, sizeof(unsigned int), nw, s.out) != nw)
                    e = "snapshot write failed (fwrite)";
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                if (!e.empty() && err.empty()) err = e;
                queue.pop_front();
                s.busy = false;
            }
            cv.notify_all();
        }
    }
};

static int writer_get(shp_handle h, StateWriter** out) {
    if (!h->writer) {
        StateWriter* w = new StateWriter();
        w->device = h->device;
        hipError_t e = hipStreamCreateWithFlags(&w->copy, hipStreamNonBlocking);
        for (int k = 0; k < 2 && e == hipSuccess; k++) {
            e = hipEventCreateWithFlags(&w->slot[k].packed, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&w->slot[k].copied, hipEventDisableTiming);
        }
        if (e != hipSuccess) {
            for (int k = 0; k < 2; k++) {
                if (w->slot[k].packed) (void)hipEventDestroy(w->slot[k].packed);
                if (w->slot[k].copied) (void)hipEventDestroy(w->slot[k].copied);
            }
            if (w->copy) (void)hipStreamDestroy(w->copy);
            delete w;
            set_error("state writer: %s", hipGetErrorString(e));
            return SHP_ERR_HIP;
        }
        w->th = std::thread([w] { w->run(); });
        h->writer = w;
    }
    *out = h->writer;
    return SHP_OK;
}

// Waits until every queued snapshot is written; reports the first failure.
static int writer_drain(StateWriter* w) {
    std::unique_lock<std::mutex> lk(w->mu);
    w->cv.wait(lk, [&] { return w->queue.empty(); });
    if (!w->err.empty()) {
        set_error("%s", w->err.c_str());
        w->err.clear();
        return SHP_ERR_HIP;
    }
    return SHP_OK;
}

void writer_release(shp_handle h) {
    StateWriter* w = h->writer;
    if (!w) return;
    {
        std::lock_guard<std::mutex> lk(w->mu);
        w->stop = true;
    }
    w->cv.notify_all();
    if (w->th.joinable()) w->th.join();
    for (int k = 0; k < 2; k++) {
        SnapSlot& s = w->slot[k];
        if (s.dev) (void)hipFree(s.dev);
        if (s.host) (void)hipHostFree(s.host);
        if (s.packed) (void)hipEventDestroy(s.packed);
        if (s.copied) (void)hipEventDestroy(s.copied);
    }
    if (w->copy) (void)hipStreamDestroy(w->copy);
    delete w;
    h->writer = nullptr;
}

static SaveArgs save_args(shp_handle h) {
    const int n = h->res_n;
    const double* f = h->res_f64;
    SaveArgs A{};
    A.n = n;
    A.x = f;
    A.y = f + n;
    A.z = f + 2 * (size_t)n;
    A.vx = f + 3 * (size_t)n;
    A.vy = f + 4 * (size_t)n;
    A.vz = f + 5 * (size_t)n;
    A.q = f + 6 * (size_t)n;
    A.ax = f + 8 * (size_t)n;
    A.ay = f + 9 * (size_t)n;
    A.az = f + 10 * (size_t)n;
    A.species = h->res_i32 + n;
    A.e2 = h->res_e2;
    return A;
}

static int state_save(shp_handle h, int format, int first, int last, int all, float time, FILE* out,
                      const char* who) {
    SHP_REQUIRE(h && out, "%s: null argument", who);
    SHP_REQUIRE(format == SHP_SAVE_ORIGINAL || format == SHP_SAVE_NEW_ORIGINAL, "%s: format %d", who, format);
    SHP_REQUIRE(h->res_f64 || h->res_n == 0, "%s: no resident state (shp_state_upload_host)", who);
    const int n = h->res_n;
    SHP_REQUIRE(first >= 0 && first <= last && last <= n, "%s: range [%d, %d) outside [0, %d)", who, first, last,
                n);
    SHP_HIP_CHECK(hipSetDevice(h->device));
    StateWriter* w;
    int rc = writer_get(h, &w);
    if (rc) return rc;
    const int k = w->next;
    w->next ^= 1;
    SnapSlot& s = w->slot[k];
    {
        // the slot's previous snapshot must be written before it is reused
        std::unique_lock<std::mutex> lk(w->mu);
        w->cv.wait(lk, [&] { return !s.busy; });
    }
    const int m = all ? n : last - first;
    const int f0 = all ? 0 : first;
    const int rw = save_rec_words(format), hw = save_hdr_words(format);
    const size_t words = kSaveOff + hw + (size_t)m * rw;
    if (words > s.dev_words) {
        if (s.dev) SHP_HIP_CHECK(hipFree(s.dev));
        s.dev = nullptr;
        s.dev_words = 0;
        SHP_HIP_CHECK(hipMalloc((void**)&s.dev, words * sizeof(unsigned int)));
        s.dev_words = words;
    }
    if (words > s.host_words) {
        if (s.host) SHP_HIP_CHECK(hipHostFree(s.host));
        s.host = nullptr;
        s.host_words = 0;
        SHP_HIP_CHECK(hipHostMalloc((void**)&s.host, words * sizeof(unsigned int), hipHostMallocDefault));
        s.host_words = words;
    }
    SHP_WS(flag, int, h, WS_FLAGS, 2 * ((size_t)m + 1));
    int* pos = flag + m + 1;
    hipStream_t st = h->stream;
    hipLaunchKernelGGL(k_save_flags, dim3((m + 256) / 256), dim3(256), 0, st, f0, m, h->res_i32, all, flag);
    size_t tsz = 0;
    SHP_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tsz, flag, pos, m + 1, st));
    SHP_WS(cub, char, h, WS_CUB, tsz);
    SHP_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(cub, tsz, flag, pos, m + 1, st));
    SaveArgs A = save_args(h);
    A.format = format;
    A.first = f0;
    A.m = m;
    A.last = last;
    A.time = time;
    hipLaunchKernelGGL(k_save_pack, dim3(std::max(1, (m + 255) / 256)), dim3(256), 0, st, A, flag, pos, s.dev);
    SHP_HIP_CHECK(hipGetLastError());
    // the copy runs on the writer's stream, after the pack, beside the steps
    // that the caller enqueues next
    SHP_HIP_CHECK(hipEventRecord(s.packed, st));
    SHP_HIP_CHECK(hipStreamWaitEvent(w->copy, s.packed, 0));
    SHP_HIP_CHECK(hipMemcpyAsync(s.host, s.dev, words * sizeof(unsigned int), hipMemcpyDeviceToHost, w->copy));
    SHP_HIP_CHECK(hipEventRecord(s.copied, w->copy));
    {
        std::lock_guard<std::mutex> lk(w->mu);
        s.out = out;
        s.hdr_words = hw;
        s.rec_words = rw;
        s.busy = true;
        w->queue.push_back(k);
    }
    w->cv.notify_all();
    return SHP_OK;
}

}  // namespace shp

using namespace shp;

extern "C" {

int shp_state_save(shp_handle h, int format, int first, int last, float current_time, FILE* out) {
    return state_save(h, format, first, last, 0, current_time, out, "shp_state_save");
}

int shp_state_save_full(shp_handle h, int format, int last_particle, float current_time, FILE* out) {
    SHP_REQUIRE(h, "shp_state_save_full: null handle");
    SHP_REQUIRE(last_particle >= 0 && last_particle <= h->res_n, "shp_state_save_full: lastParticle %d",
                last_particle);
    return state_save(h, format, 0, last_particle, 1, current_time, out, "shp_state_save_full");
}

int shp_state_flush(shp_handle h) {
    SHP_REQUIRE(h, "shp_state_flush: null handle");
    if (!h->writer) return SHP_OK;
    return writer_drain(h->writer);
}

int shp_state_events(shp_handle h, FILE* out, int* nevents) {
    SHP_REQUIRE(h && out, "shp_state_events: null argument");
    if (nevents) *nevents = 0;
    if (!h->res_evn) return SHP_OK;
    SHP_HIP_CHECK(hipSetDevice(h->device));
    hipStream_t st = h->stream;
    int cnt = 0;
    SHP_HIP_CHECK(hipMemcpyAsync(&cnt, h->res_evn, sizeof(int), hipMemcpyDeviceToHost, st));
    SHP_HIP_CHECK(hipStreamSynchronize(st));
    const int ne = std::min(cnt, h->res_cap);
    if (ne > 0) {
        std::vector<unsigned int> rec((size_t)ne * kEventWords);
        std::vector<unsigned long long> key(ne);
        SHP_HIP_CHECK(hipMemcpyAsync(rec.data(), h->res_ev, rec.size() * sizeof(unsigned int),
                                     hipMemcpyDeviceToHost, st));
        SHP_HIP_CHECK(hipMemcpyAsync(key.data(), h->res_evkey, key.size() * sizeof(unsigned long long),
                                     hipMemcpyDeviceToHost, st));
        hipLaunchKernelGGL(k_wr_zero, dim3(1), dim3(64), 0, st, h->res_evn);
        SHP_HIP_CHECK(hipStreamSynchronize(st));
        // the order of the reference's passes: by drift, spatial before
        // temporal, then ion index
        std::vector<int> ord(ne);
        std::iota(ord.begin(), ord.end(), 0);
        std::sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
        for (int e : ord)
            SHP_REQUIRE(fwrite(&rec[(size_t)e * kEventWords], sizeof(unsigned int), kEventWords, out) ==
                            (size_t)kEventWords,
                        "shp_state_events: write failed (fwrite)");
    }
    if (nevents) *nevents = ne;
    return SHP_OK;
}

int shp_state_fields_host(shp_handle h, float* efield1, float* efield2) {
    SHP_REQUIRE(h && efield1 && efield2, "shp_state_fields_host: null argument");
    const int n = h->res_n;
    if (n == 0) return SHP_OK;
    SHP_HIP_CHECK(hipSetDevice(h->device));
    SHP_WS(e, float, h, WS_HOST_OUT, 6 * (size_t)n);
    const SaveArgs A = save_args(h);
    hipLaunchKernelGGL(k_fields_aos, dim3((n + 255) / 256), dim3(256), 0, h->stream, A, e, e + 3 * (size_t)n);
    SHP_HIP_CHECK(hipGetLastError());
    SHP_HIP_CHECK(hipMemcpyAsync(efield1, e, 3 * sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, h->stream));
    SHP_HIP_CHECK(hipMemcpyAsync(efield2, e + 3 * (size_t)n, 3 * sizeof(float) * (size_t)n, hipMemcpyDeviceToHost,
                                 h->stream));
    SHP_HIP_CHECK(hipStreamSynchronize(h->stream));
    return SHP_OK;
}

}  // extern "C"
