// engine.cpp -- implementation of the C ABI declared in include/spinwalk_engine.h.
//
// Host-side orchestration of the HIP walk: device selection and memory checks
// (the reference's monte_carlo constructor and device_helper), phantom and
// sequence upload (monte_carlo.cu:247-262, simulation_parameters.cuh:74-90), the
// per-sequence precomputation of event trigonometry, and the launch of the walk
// (monte_carlo.cu:324-333).  Compiled by g++ against the HIP runtime API.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <math.h>
#include <memory>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/spinwalk_engine.h"
#include "walk_args.hpp"
#include "walk_launch.hpp"
#include "phantom_gen.hpp"

#define SW_VERSION_STRING "spinwalk-mi355x 0.1.0 (drop-in for SpinWalk 1.21.0 sim)"

namespace {

constexpr double kDeg2Rad = 0.0174532925199433;   // simulation_parameters.cuh:24
constexpr double kRad2Deg = 57.2957795130823;     // simulation_parameters.cuh:25
constexpr double kGamma   = 267515315.;           // definitions.h:20
constexpr uint64_t kPairsMinBytes = 256ull << 20;   // records beyond the Infinity Cache: two attempts per trip

thread_local std::string g_tls_error;

}  // namespace

struct sw_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    unsigned long long *d_counters = nullptr;   // 4 x u64
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // scratch reused across calls
    void *d_sort = nullptr; size_t sort_bytes = 0;
    uint32_t *d_order = nullptr; size_t order_cap = 0;
    double *d_partial = nullptr; size_t partial_cap = 0;
    // device copies of sw_walk_host's buffers: M0, XYZ0, M1, XYZ1, T, normals, uniforms
    void *d_host[7] = {};
    size_t host_cap[7] = {};
    // Calls that use the engine-owned scratch above (counters, visiting order,
    // sort and reduction buffers, excitation tables) may come on different
    // streams.  Each records scratch_ev on its stream when it has enqueued its
    // last use; the next call on another stream waits for it first.
    hipEvent_t scratch_ev = nullptr;
    hipStream_t scratch_stream = nullptr;     // stream of the last recorded use
    // sw_walk_device_noalloc: the device address of this device's fixed scratch
    // (noalloc.hip), looked up at engine creation; na_err keeps a failed lookup.
    void *na_scratch = nullptr;
    hipError_t na_err = hipSuccess;
};

struct sw_phantom {
    sw_engine *eng = nullptr;
    int device = 0;
    uint64_t dims[3] = {0, 0, 0};
    float fov[3] = {0, 0, 0};
    uint64_t n = 0;
    uint8_t *mask = nullptr;
    float *field = nullptr;          // tesla per 1 T
    float half_scale = 1.f;          // float16/bfloat16 records hold field_T / half_scale
    float q_scale = 1.f;             // int16 records hold rint(field_T / q_scale)
    uint32_t n_labels = 0;
    bool has_field = false;
    // bricked voxel records gathered by the walk (walk_args.hpp)
    void *rec[5] = {};               // by SW_FIELD_*: none (label only), f32, f16, bf16, q16
    void *rec_packed = nullptr;      // float32 packed 5-byte records (production batches)
    void *rec_soa = nullptr;         // float32 structure-of-arrays records, same bricks (single production walks)
};

struct sw_sequence {
    sw_engine *eng = nullptr;
    int device = 0;
    uint32_t n_sub = 0, n_rf = 0, n_te = 0, n_deph = 0, n_grad = 0;
    double *diffusivity = nullptr;
    float *pXY = nullptr, *T1s = nullptr, *T2s = nullptr;
    double *T1d = nullptr, *T2d = nullptr;
    sw::RfPulse *rf = nullptr;
    int32_t *TE_t = nullptr, *deph_t = nullptr, *grad_t = nullptr;
    float *deph_deg = nullptr, *gx = nullptr, *gy = nullptr, *gz = nullptr;
    // excitation table cache, keyed by the parameters it depends on
    sw::Excitation *exc = nullptr;
    uint32_t exc_cap = 0;
    bool exc_valid = false;
    float key_lin = 0, key_quad = 0; int32_t key_ndummy = -1;
    float RF_FA0 = 0, RF_PH0 = 0;
    // The tables that shape a walk's trajectory and relaxation, byte for byte:
    // diffusivity, P_XY, T1/T2 and every event time.  Tiles whose sequences have
    // equal signatures (and equal draws, geometry and start positions) walk the
    // same trajectories and can share one fused walk.
    std::vector<char> traj_sig;
};

namespace {

int fail(sw_engine *e, int code, const std::string &msg) {
    if (e) e->err = msg; else g_tls_error = msg;
    return code;
}

#define SW_HIP(eng, expr)                                                                      \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return fail((eng), SW_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));  \
    } while (0)

template <class T>
hipError_t dev_upload(T **dst, const T *src, size_t n, hipStream_t s) {
    *dst = nullptr;
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc((void **)dst, n * sizeof(T));
    if (e != hipSuccess) return e;
    if (src) return hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, s);
    return hipSuccess;
}

void dev_free(void *p) { if (p) (void)hipFree(p); }

// xrot_withphase special cases (src/sim/kernels.cuh:162-181)
int32_t rot_kind_f(float ph) {
    if (ph == 0.0f) return sw::ROT_X;
    if (ph == 180.0) return sw::ROT_XNEG;
    if (ph == 90.0) return sw::ROT_Y;
    if (ph == -90.0 || ph == 270.0) return sw::ROT_YNEG;
    return sw::ROT_GENERAL;
}
int32_t rot_kind_d(double ph) {
    if (ph == 0.0) return sw::ROT_X;
    if (ph == 180.0) return sw::ROT_XNEG;
    if (ph == 90.0) return sw::ROT_Y;
    if (ph == -90.0 || ph == 270.0) return sw::ROT_YNEG;
    return sw::ROT_GENERAL;
}

// Excitation phase of dummy scan d (src/sim/kernels.cu:110-114), float path.
float exc_phase_f(float rf_phase, uint32_t d, float lin, float quad) {
    float v = (float)((double)(rf_phase + d * lin) + d * (d + 1) / 2.0 * quad);
    while (v > 360.0) v = (float)(v - 360.0);
    while (v < 0) v = (float)(v + 360.0);
    return v;
}
double exc_phase_d(float rf_phase, uint32_t d, float lin, float quad) {
    double v = (double)rf_phase + (double)d * lin + d * (d + 1) / 2.0 * quad;
    while (v > 360.0) v -= 360.0;
    while (v < 0) v += 360.0;
    return v;
}

// The excitation of every dummy scan and of the measured scan (kernels.cu:110-120).
std::vector<sw::Excitation> exc_table(const sw_sequence *q, const sw_params *p) {
    const uint32_t n = (uint32_t)p->n_dummy_scan + 1;
    std::vector<sw::Excitation> h(n);
    for (uint32_t d = 0; d < n; d++) {
        const float pf = exc_phase_f(q->RF_PH0, d, p->linear_phase_cycling, p->quadratic_phase_cycling);
        const double pd = exc_phase_d(q->RF_PH0, d, p->linear_phase_cycling, p->quadratic_phase_cycling);
        h[d].kind = rot_kind_f(pf);
        h[d].kind_d = rot_kind_d(pd);
        h[d].sp = sinf((float)(pf * kDeg2Rad));
        h[d].cp = cosf((float)(pf * kDeg2Rad));
        h[d].spd = sin(pd * kDeg2Rad);
        h[d].cpd = cos(pd * kDeg2Rad);
    }
    return h;
}

int ensure_exc(sw_engine *e, sw_sequence *q, const sw_params *p, hipStream_t s) {
    if (q->exc_valid && q->key_lin == p->linear_phase_cycling && q->key_quad == p->quadratic_phase_cycling &&
        q->key_ndummy == p->n_dummy_scan)
        return SW_OK;
    if (p->n_dummy_scan < 0) return fail(e, SW_E_INVALID, "n_dummy_scan < 0: call sw_params_prepare first");
    const uint32_t n = (uint32_t)p->n_dummy_scan + 1;
    const std::vector<sw::Excitation> h = exc_table(q, p);
    if (n > q->exc_cap) {
        SW_HIP(e, hipStreamSynchronize(s));
        dev_free(q->exc);
        q->exc = nullptr;
        SW_HIP(e, hipMalloc((void **)&q->exc, n * sizeof(sw::Excitation)));
        q->exc_cap = n;
    }
    SW_HIP(e, hipMemcpyAsync(q->exc, h.data(), n * sizeof(sw::Excitation), hipMemcpyHostToDevice, s));
    SW_HIP(e, hipStreamSynchronize(s));   // h goes out of scope
    q->exc_valid = true;
    q->key_lin = p->linear_phase_cycling;
    q->key_quad = p->quadratic_phase_cycling;
    q->key_ndummy = p->n_dummy_scan;
    return SW_OK;
}

int check_free_memory(sw_engine *e, size_t bytes) {   // device_helper.cu:78-101
    size_t fr = 0, tot = 0;
    SW_HIP(e, hipMemGetInfo(&fr, &tot));
    if (bytes > fr) {
        char b[256];
        snprintf(b, sizeof b, "not enough device memory: need %zu MB, free %zu MB", bytes >> 20, fr >> 20);
        return fail(e, SW_E_NOMEM, b);
    }
    return SW_OK;
}

// Orders a call's use of the engine-owned scratch after the previous call's,
// when that call ran on another stream.
int scratch_acquire(sw_engine *e, hipStream_t s) {
    if (e->scratch_stream && e->scratch_stream != s) SW_HIP(e, hipStreamWaitEvent(s, e->scratch_ev, 0));
    return SW_OK;
}
// Marks the point on `s` after which the scratch is free for the next call.
int scratch_release(sw_engine *e, hipStream_t s) {
    SW_HIP(e, hipEventRecord(e->scratch_ev, s));
    e->scratch_stream = s;
    return SW_OK;
}

}  // namespace

extern "C" {

SW_API const char *sw_version(void) { return SW_VERSION_STRING; }

SW_API int sw_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

SW_API int sw_device_info(int device_id, char *buf, size_t buflen) {
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, device_id) != hipSuccess) return SW_E_NODEVICE;
    size_t fr = 0, tot = 0;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device_id);
    (void)hipMemGetInfo(&fr, &tot);
    (void)hipSetDevice(cur);
    if (buf && buflen)
        snprintf(buf, buflen,
                 "Device %d: %s (%s)\n  compute units: %d\n  clock: %d MHz\n  global memory: %zu MB "
                 "(free %zu MB)\n  L2 cache: %d KB\n  wavefront: %d\n",
                 device_id, pr.name, pr.gcnArchName, pr.multiProcessorCount, pr.clockRate / 1000,
                 (size_t)pr.totalGlobalMem >> 20, fr >> 20, pr.l2CacheSize >> 10, pr.warpSize);
    return SW_OK;
}

SW_API int sw_engine_create(int device_id, sw_engine **out) {
    if (!out) return SW_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(nullptr, SW_E_NODEVICE, "no HIP device found");
    if (device_id < 0 || device_id >= n)
        return fail(nullptr, SW_E_NODEVICE, "device id " + std::to_string(device_id) + " is not available");
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, device_id) != hipSuccess)
        return fail(nullptr, SW_E_NODEVICE, "hipGetDeviceProperties failed");
    if (std::strncmp(pr.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, SW_E_NODEVICE, std::string("device is ") + pr.gcnArchName + ", this build targets gfx950");
    std::unique_ptr<sw_engine, void (*)(sw_engine *)> e(new sw_engine, sw_engine_destroy);
    e->device = device_id;
    if (hipSetDevice(device_id) != hipSuccess) return fail(nullptr, SW_E_HIP, "hipSetDevice failed");
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(nullptr, SW_E_HIP, "hipStreamCreate failed");
    if (hipMalloc((void **)&e->d_counters, 4 * sizeof(unsigned long long)) != hipSuccess)
        return fail(nullptr, SW_E_HIP, "hipMalloc counters failed");
    if (hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&e->scratch_ev, hipEventDisableTiming) != hipSuccess)
        return fail(nullptr, SW_E_HIP, "hipEventCreate failed");
    e->na_err = sw::noalloc_scratch_init(device_id, &e->na_scratch);   // reported by sw_walk_device_noalloc
    *out = e.release();
    return SW_OK;
}

SW_API void sw_engine_destroy(sw_engine *e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (void *p : e->d_host) dev_free(p);
    dev_free(e->d_counters);
    dev_free(e->d_sort);
    dev_free(e->d_order);
    dev_free(e->d_partial);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->scratch_ev) {
        (void)hipEventSynchronize(e->scratch_ev);
        (void)hipEventDestroy(e->scratch_ev);
    }
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

SW_API const char *sw_last_error(const sw_engine *e) {
    return e ? e->err.c_str() : g_tls_error.c_str();
}

// ---------------------------------------------------------------------------
// host transforms
// ---------------------------------------------------------------------------
SW_API int sw_params_prepare(sw_params *p, double *diffusivity, float RF_FA0_deg, float T1_0_ms) {
    if (!p || (p->n_substrate && !diffusivity) || p->timestep_us <= 0) return SW_E_INVALID;
    p->c = cosf((float)(RF_FA0_deg * kDeg2Rad));                 // simulation_parameters.cuh:229
    p->s = sinf((float)(RF_FA0_deg * kDeg2Rad));                 // :230
    p->n_timepoints = (uint32_t)(p->TR_us / p->timestep_us);     // :231
    if (p->seed == 0) p->seed = std::random_device{}();          // :233-234
    for (uint32_t i = 0; i < p->n_substrate; i++)                // :236-237
        diffusivity[i] = 1e-3 * sqrt(2. * diffusivity[i] * p->timestep_us);
    if (p->n_dummy_scan < 0)                                     // :239-242
        p->n_dummy_scan = (int32_t)(5.0 * T1_0_ms / float(p->TR_us * 1e-3));
    return SW_OK;
}

SW_API void sw_timing_scale(int32_t *v, uint32_t n, int32_t dt) {   // config_reader.cpp:39-46
    if (!v || dt == 0) return;
    for (uint32_t i = 0; i < n; i++) v[i] = v[i] / dt;
}

SW_API float sw_tesla2deg(float B0, int32_t dt) {                 // monte_carlo.cu:241
    return (float)(B0 * dt * 1e-6 * kGamma * kRad2Deg);
}

SW_API void sw_init_positions(uint64_t seed, const float fov[3], uint32_t n_spins, float *XYZ0) {
    std::mt19937 gen((std::mt19937::result_type)seed);           // monte_carlo.cu:142-151
    std::uniform_real_distribution<float> dx(0.01 * fov[0], 0.99 * fov[0]);
    std::uniform_real_distribution<float> dy(0.01 * fov[1], 0.99 * fov[1]);
    std::uniform_real_distribution<float> dz(0.01 * fov[2], 0.99 * fov[2]);
    for (uint32_t i = 0; i < n_spins; i++) {
        XYZ0[3 * (size_t)i + 0] = dx(gen);
        XYZ0[3 * (size_t)i + 1] = dy(gen);
        XYZ0[3 * (size_t)i + 2] = dz(gen);
    }
}

SW_API int sw_replay_fill(uint64_t seed, uint32_t spin_begin, uint32_t n, float *normals, uint64_t ns,
                          float *uniforms, uint64_t us, int n_threads) {
    if ((ns && !normals) || (us && !uniforms)) return SW_E_INVALID;
    n_threads = std::max(1, std::min(n_threads, 256));
    auto work = [&](uint32_t i0, uint32_t i1) {
        for (uint32_t i = i0; i < i1; i++) {
            const uint64_t sp = seed + spin_begin + i;
            // kernels.cu:82-88: two generators seeded seed+spin, each discarded by seed+spin
            std::mt19937 gr((std::mt19937::result_type)sp), gu((std::mt19937::result_type)sp);
            gr.discard(sp);
            gu.discard(sp);
            std::normal_distribution<float> nd(0.f, 1.0f);
            std::uniform_real_distribution<float> ud(0.f, 1.f);
            float *nr = normals + (uint64_t)i * ns;
            for (uint64_t k = 0; k < ns; k++) nr[k] = nd(gr);
            float *ur = uniforms + (uint64_t)i * us;
            for (uint64_t k = 0; k < us; k++) ur[k] = ud(gu);
        }
    };
    if (n_threads == 1 || n < 2) { work(0, n); return SW_OK; }
    std::vector<std::thread> th;
    const uint32_t chunk = (n + n_threads - 1) / n_threads;
    for (int t = 0; t < n_threads; t++) {
        const uint32_t a = t * chunk, b = std::min(n, a + chunk);
        if (a < b) th.emplace_back(work, a, b);
    }
    for (auto &t : th) t.join();
    return SW_OK;
}

// ---------------------------------------------------------------------------
// phantom
// ---------------------------------------------------------------------------
// HBM a phantom takes: mask, float32 field and the bricked records (for
// float32 8 bytes a voxel and twice 5.33 for the production copies, 4 for the
// label-only records, 3.56 for the 16-bit ones, rounded up).
static size_t phantom_bytes(uint64_t n, bool has_field, uint32_t formats) {
    size_t need = n;
    if (!has_field) return need + n * 4;
    need += n * 4;
    for (int f = SW_FIELD_F32; f <= SW_FIELD_Q16; f++)
        if (formats & (1u << f)) need += n * (f == SW_FIELD_F32 ? 8 + 6 + 6 : 4);   // f32: 8-byte, packed and SoA records
    return need;
}

static int phantom_finish(sw_engine *e, sw_phantom *ph, uint32_t formats) {
    hipStream_t s = e->stream;
    uint32_t st[2] = {0, 0};
    {
        uint32_t *d_stats = nullptr;
        SW_HIP(e, hipMalloc((void **)&d_stats, 2 * sizeof(uint32_t)));
        std::unique_ptr<uint32_t, void (*)(uint32_t *)> guard(d_stats, [](uint32_t *p) { dev_free(p); });
        SW_HIP(e, hipMemsetAsync(d_stats, 0, 2 * sizeof(uint32_t), s));
        SW_HIP(e, sw::launch_phantom_stats(ph->mask, ph->field, ph->n, d_stats, s));
        SW_HIP(e, hipMemcpyAsync(st, d_stats, sizeof st, hipMemcpyDeviceToHost, s));
        SW_HIP(e, hipStreamSynchronize(s));
    }
    ph->n_labels = st[0] + 1;                                    // monte_carlo.cu:113
    float fmax;
    std::memcpy(&fmax, &st[1], sizeof fmax);
    if (ph->has_field && (formats & (1u << SW_FIELD_Q16)) && !std::isfinite(fmax))
        return fail(e, SW_E_INVALID, "the int16 field copy needs a finite fieldmap");
    // power-of-two scale mapping max|field| to about 2^14
    int ex = 0;
    if (fmax > 0 && std::isfinite(fmax)) { std::frexp(fmax, &ex); ph->half_scale = std::ldexp(1.0f, ex - 15); }
    else ph->half_scale = 1.f;
    // fixed-point step mapping max|field| to 32767
    ph->q_scale = (fmax > 0 && std::isfinite(fmax)) ? (float)((double)fmax / 32767.0) : 1.f;
    // bricked records for the walk
    const uint64_t nx = ph->dims[0], ny = ph->dims[1], nz = ph->dims[2];
    for (int f = 0; f <= SW_FIELD_Q16; f++) {
        const bool want = (f == SW_FIELD_NONE) ? !ph->has_field : (ph->has_field && (formats & (1u << f)));
        if (!want) continue;
        size_t bytes;
        if (f == SW_FIELD_F16 || f == SW_FIELD_BF16 || f == SW_FIELD_Q16) {   // packed 3-byte records
            bytes = (size_t)sw::pack_bytes(nx, ny, nz, sw::kPack16, sw::pack_super(nx, ny, nz, sw::kPack16));
        } else {
            const sw::BrickShape sh = f == SW_FIELD_F32 ? sw::kBrick8 : sw::kBrick4;
            bytes = ((size_t)(sw::brick_count(nx, sh.lbx) * sw::brick_count(ny, sh.lby) * sw::brick_count(nz, sh.lbz))
                     << (sh.lbx + sh.lby + sh.lbz)) * (f == SW_FIELD_F32 ? 8 : 4);
        }
        SW_HIP(e, hipMalloc(&ph->rec[f], bytes));
        SW_HIP(e, hipMemsetAsync(ph->rec[f], 0, bytes, s));
        const int kind = f == SW_FIELD_F32 ? 0 : f == SW_FIELD_F16 ? 1 : f == SW_FIELD_BF16 ? 2 : f == SW_FIELD_Q16 ? 4 : 3;
        SW_HIP(e, sw::launch_records(ph->mask, ph->field, nx, ny, nz, kind,
                                     1.0f / (f == SW_FIELD_Q16 ? ph->q_scale : ph->half_scale), ph->rec[f], s));
        if (f == SW_FIELD_F32) {         // the production walks' copies: packed for batches, SoA for single walks
            const size_t bytes = (size_t)sw::pack_bytes(nx, ny, nz, sw::kPackF32, sw::pack_super(nx, ny, nz, sw::kPackF32));
            for (int soa = 0; soa < 2; soa++) {
                void *&r = soa ? ph->rec_soa : ph->rec_packed;
                SW_HIP(e, hipMalloc(&r, bytes));
                SW_HIP(e, hipMemsetAsync(r, 0, bytes, s));
                SW_HIP(e, sw::launch_f32_records(ph->mask, ph->field, nx, ny, nz, soa != 0, r, s));
            }
        }
    }
    SW_HIP(e, hipStreamSynchronize(s));
    return SW_OK;
}

static int phantom_create_common(sw_engine *e, const uint8_t *mask, const float *field, const uint64_t dims[3],
                                 const float fov[3], uint32_t formats, sw_phantom **out, hipMemcpyKind kind) {
    if (!e || !mask || !dims || !fov || !out) return fail(e, SW_E_INVALID, "null argument");
    *out = nullptr;
    SW_HIP(e, hipSetDevice(e->device));
    std::unique_ptr<sw_phantom, void (*)(sw_phantom *)> ph(new sw_phantom, sw_phantom_destroy);
    ph->eng = e;
    ph->device = e->device;
    ph->n = dims[0] * dims[1] * dims[2];
    if (ph->n == 0) return fail(e, SW_E_INVALID, "empty phantom");
    for (int i = 0; i < 3; i++) { ph->dims[i] = dims[i]; ph->fov[i] = fov[i]; }
    ph->has_field = field != nullptr;
    if (formats == 0) formats = 1u << SW_FIELD_F32;
    int rc = check_free_memory(e, phantom_bytes(ph->n, ph->has_field, formats) + (64u << 20));
    if (rc) return rc;
    hipStream_t s = e->stream;
    SW_HIP(e, hipMalloc((void **)&ph->mask, ph->n));
    SW_HIP(e, hipMemcpyAsync(ph->mask, mask, ph->n, kind, s));
    if (ph->has_field) {
        SW_HIP(e, hipMalloc((void **)&ph->field, ph->n * sizeof(float)));
        SW_HIP(e, hipMemcpyAsync(ph->field, field, ph->n * sizeof(float), kind, s));
    }
    SW_HIP(e, hipStreamSynchronize(s));
    rc = phantom_finish(e, ph.get(), formats);
    if (rc) return rc;
    *out = ph.release();
    return SW_OK;
}

SW_API int sw_phantom_create(sw_engine *e, const uint8_t *mask, const float *field_T, const uint64_t dims[3],
                             const float fov[3], uint32_t formats, sw_phantom **out) {
    return phantom_create_common(e, mask, field_T, dims, fov, formats, out, hipMemcpyHostToDevice);
}

SW_API int sw_phantom_create_device(sw_engine *e, const uint8_t *d_mask, const float *d_field_T,
                                    const uint64_t dims[3], const float fov[3], uint32_t formats, sw_phantom **out) {
    return phantom_create_common(e, d_mask, d_field_T, dims, fov, formats, out, hipMemcpyDeviceToDevice);
}

SW_API void sw_phantom_destroy(sw_phantom *ph) {
    if (!ph) return;
    (void)hipSetDevice(ph->device);
    dev_free(ph->mask);
    dev_free(ph->field);
    for (void *r : ph->rec) dev_free(r);
    dev_free(ph->rec_packed);
    dev_free(ph->rec_soa);
    delete ph;
}

SW_API int sw_phantom_info(const sw_phantom *ph, uint32_t *n_labels, float *half_scale) {
    if (!ph) return SW_E_INVALID;
    if (n_labels) *n_labels = ph->n_labels;
    if (half_scale) *half_scale = ph->half_scale;
    return SW_OK;
}

SW_API int sw_phantom_download(const sw_phantom *ph, uint8_t *mask, float *field_T) {
    if (!ph) return SW_E_INVALID;
    sw_engine *e = ph->eng;
    SW_HIP(e, hipSetDevice(ph->device));
    if (mask) SW_HIP(e, hipMemcpy(mask, ph->mask, ph->n, hipMemcpyDeviceToHost));
    if (field_T) {
        if (!ph->field) return fail(e, SW_E_INVALID, "phantom has no float32 field copy");
        SW_HIP(e, hipMemcpy(field_T, ph->field, ph->n * sizeof(float), hipMemcpyDeviceToHost));
    }
    return SW_OK;
}

SW_API int sw_phantom_generate(sw_engine *e, const uint64_t dims[3], const float fov_um[3], const sw_shape_spec *specs,
                               uint32_t n_specs, uint32_t formats, sw_phantom **out, double *achieved_vf) {
    if (!e || !dims || !fov_um || (n_specs && !specs) || !out) return fail(e, SW_E_INVALID, "null argument");
    *out = nullptr;
    SW_HIP(e, hipSetDevice(e->device));
    std::vector<sw::GenCylinder> cyl, axn;
    std::vector<sw::GenSphere> sph;
    for (uint32_t i = 0; i < n_specs; i++) {
        sw::ShapeSpec ss;
        ss.kind = specs[i].kind;
        if (specs[i].label == 0 || specs[i].label > 255) return fail(e, SW_E_INVALID, "shape label must be 1..255");
        const bool neg_ok = ss.kind == SW_SHAPE_AXONS;         // negative radius: random radii
        if (!(specs[i].radius_um > 0 || (neg_ok && specs[i].radius_um < 0)))
            return fail(e, SW_E_INVALID, "shape radius must be positive");
        if (specs[i].radius_min_um < 0 || (specs[i].radius_min_um > 0 && !neg_ok))
            return fail(e, SW_E_INVALID, "radius_min_um applies to axons and must not be negative");
        ss.radius_min_um = specs[i].radius_min_um;
        ss.label = (uint8_t)specs[i].label;
        ss.radius_um = specs[i].radius_um;
        ss.volume_fraction = specs[i].volume_fraction;
        ss.random_orientation = specs[i].random_orientation;
        ss.orientation_deg = specs[i].orientation_deg;
        ss.dchi = specs[i].dchi;
        ss.Y = specs[i].Y;
        ss.seed = specs[i].seed;
        ss.trunc_radii = specs[i].trunc_radii;
        double vf = 0;
        if (ss.kind == SW_SHAPE_CYLINDER) vf = sw::sample_cylinders(ss, fov_um, cyl);
        else if (ss.kind == SW_SHAPE_SPHERE) vf = sw::sample_spheres(ss, fov_um, cyl, sph);
        else if (ss.kind == SW_SHAPE_AXONS) vf = sw::sample_axons(ss, fov_um, axn);
        else return fail(e, SW_E_INVALID, "unknown shape kind");
        if (achieved_vf) achieved_vf[i] = vf;
    }
    // cell list of spheres; a cell is at least as large as every truncation radius
    float cell = 1.f;
    for (auto &s : sph) cell = std::max(cell, std::sqrt(s.trunc2));
    uint32_t nc[3];
    for (int i = 0; i < 3; i++) nc[i] = std::max(1u, (uint32_t)std::ceil(fov_um[i] / cell));
    const size_t ncells = (size_t)nc[0] * nc[1] * nc[2];
    std::vector<uint32_t> start(ncells + 1, 0), cid(sph.size());
    for (size_t k = 0; k < sph.size(); k++) {
        uint32_t c3[3];
        for (int i = 0; i < 3; i++) c3[i] = std::min(nc[i] - 1, (uint32_t)std::max(0.f, sph[k].c[i] / cell));
        cid[k] = (c3[0] * nc[1] + c3[1]) * nc[2] + c3[2];
        start[cid[k] + 1]++;
    }
    for (size_t c = 0; c < ncells; c++) start[c + 1] += start[c];
    std::vector<sw::GenSphere> sorted(sph.size());
    {
        std::vector<uint32_t> fillp(start.begin(), start.end() - 1);
        for (size_t k = 0; k < sph.size(); k++) sorted[fillp[cid[k]]++] = sph[k];
    }
    // 2-D cell list of axons over the x-y plane, cells as large as every truncation radius
    float acell = 1.f;
    for (auto &a : axn) acell = std::max(acell, std::sqrt(a.trunc2));
    const uint32_t anc[2] = {std::max(1u, (uint32_t)std::ceil(fov_um[0] / acell)),
                             std::max(1u, (uint32_t)std::ceil(fov_um[1] / acell))};
    std::vector<uint32_t> astart((size_t)anc[0] * anc[1] + 1, 0), acid(axn.size());
    for (size_t k = 0; k < axn.size(); k++) {
        const uint32_t cx = std::min(anc[0] - 1, (uint32_t)std::max(0.f, axn[k].p[0] / acell));
        const uint32_t cy = std::min(anc[1] - 1, (uint32_t)std::max(0.f, axn[k].p[1] / acell));
        acid[k] = cx * anc[1] + cy;
        astart[acid[k] + 1]++;
    }
    for (size_t c = 0; c + 1 < astart.size(); c++) astart[c + 1] += astart[c];
    std::vector<sw::GenCylinder> asorted(axn.size());
    {
        std::vector<uint32_t> fillp(astart.begin(), astart.end() - 1);
        for (size_t k = 0; k < axn.size(); k++) asorted[fillp[acid[k]]++] = axn[k];
    }
    std::unique_ptr<sw_phantom, void (*)(sw_phantom *)> ph(new sw_phantom, sw_phantom_destroy);
    ph->eng = e;
    ph->device = e->device;
    ph->n = dims[0] * dims[1] * dims[2];
    if (ph->n == 0) return fail(e, SW_E_INVALID, "empty phantom");
    for (int i = 0; i < 3; i++) { ph->dims[i] = dims[i]; ph->fov[i] = fov_um[i] * 1e-6f; }
    ph->has_field = true;
    if (formats == 0) formats = 1u << SW_FIELD_F32;
    int rc = check_free_memory(e, phantom_bytes(ph->n, true, formats) + (64u << 20));
    if (rc) return rc;
    hipStream_t s = e->stream;
    sw::GenCylinder *dcyl = nullptr;
    sw::GenSphere *dsph = nullptr;
    uint32_t *dstart = nullptr, *dastart = nullptr;
    sw::GenCylinder *daxn = nullptr;
    struct Tables {                      // the shape tables, freed on every return
        void **p[5];
        ~Tables() { for (void **q : p) dev_free(*q); }
    } tables{{(void **)&dcyl, (void **)&dsph, (void **)&dstart, (void **)&daxn, (void **)&dastart}};
    SW_HIP(e, hipMalloc((void **)&ph->mask, ph->n));
    SW_HIP(e, hipMalloc((void **)&ph->field, ph->n * sizeof(float)));
    SW_HIP(e, dev_upload(&dcyl, cyl.data(), cyl.size(), s));
    SW_HIP(e, dev_upload(&dsph, sorted.data(), sorted.size(), s));
    SW_HIP(e, dev_upload(&dstart, start.data(), start.size(), s));
    SW_HIP(e, dev_upload(&daxn, asorted.data(), asorted.size(), s));
    SW_HIP(e, dev_upload(&dastart, astart.data(), astart.size(), s));
    sw::GenArgs g;
    std::memset(&g, 0, sizeof g);
    g.n = ph->n;
    g.nx = (uint32_t)dims[0]; g.ny = (uint32_t)dims[1]; g.nz = (uint32_t)dims[2];
    for (int i = 0; i < 3; i++) g.vs[i] = fov_um[i] / (float)dims[i];
    g.cyl = dcyl; g.n_cyl = (uint32_t)cyl.size();
    g.sph = dsph; g.n_sph = (uint32_t)sorted.size();
    g.cell_start = dstart;
    for (int i = 0; i < 3; i++) g.nc[i] = nc[i];
    g.cell = cell;
    g.axn = daxn; g.n_axn = (uint32_t)asorted.size();
    g.axn_start = dastart;
    g.anc[0] = anc[0]; g.anc[1] = anc[1];
    g.acell = acell;
    g.background = 0;
    g.mask = ph->mask;
    g.field = ph->field;
    SW_HIP(e, sw::launch_phantom_gen(g, s));
    SW_HIP(e, hipStreamSynchronize(s));
    rc = phantom_finish(e, ph.get(), formats);
    if (rc) return rc;
    *out = ph.release();
    return SW_OK;
}

// ---------------------------------------------------------------------------
// sequence
// ---------------------------------------------------------------------------
SW_API int sw_sequence_create(sw_engine *e, const sw_params *p, const sw_seq *q, sw_sequence **out) {
    if (!e || !p || !q || !out) return fail(e, SW_E_INVALID, "null argument");
    *out = nullptr;
    if (q->n_substrate == 0 || !q->diffusivity || !q->T1_ms || !q->T2_ms || !q->pXY)
        return fail(e, SW_E_INVALID, "substrate tables are required");
    if (q->n_rf == 0 || !q->RF_FA_deg || !q->RF_PH_deg || !q->RF_t)
        return fail(e, SW_E_INVALID, "at least one RF pulse (the excitation) is required");
    if ((q->n_te && !q->TE_t) || (q->n_dephasing && (!q->dephasing_t || !q->dephasing_deg)) ||
        (q->n_gradient && (!q->gradient_t || !q->gradientX_mTm || !q->gradientY_mTm || !q->gradientZ_mTm)))
        return fail(e, SW_E_INVALID, "table pointer missing for a non-empty table");
    SW_HIP(e, hipSetDevice(e->device));
    std::unique_ptr<sw_sequence, void (*)(sw_sequence *)> sq(new sw_sequence, sw_sequence_destroy);
    sq->eng = e;
    sq->device = e->device;
    const uint32_t ns = q->n_substrate;
    sq->n_sub = ns; sq->n_rf = q->n_rf; sq->n_te = q->n_te; sq->n_deph = q->n_dephasing; sq->n_grad = q->n_gradient;
    sq->RF_FA0 = q->RF_FA_deg[0];
    sq->RF_PH0 = q->RF_PH_deg[0];
    std::vector<float> T1s(ns), T2s(ns);
    std::vector<double> T1d(ns), T2d(ns);
    for (uint32_t i = 0; i < ns; i++) {
        T1s[i] = (float)(q->T1_ms[i] * 1e-3);                    // kernels.cu:167-168
        T2s[i] = (float)(q->T2_ms[i] * 1e-3);
        T1d[i] = (double)q->T1_ms[i] * 1e-3;
        T2d[i] = (double)q->T2_ms[i] * 1e-3;
    }
    std::vector<sw::RfPulse> rf(q->n_rf);
    for (uint32_t k = 0; k < q->n_rf; k++) {
        const float fa = q->RF_FA_deg[k], phs = q->RF_PH_deg[k];
        sw::RfPulse &r = rf[k];
        std::memset(&r, 0, sizeof r);
        r.t = q->RF_t[k];
        r.kind = rot_kind_f(phs);
        r.kind_d = rot_kind_d((double)phs);
        r.s = (float)sin(fa * kDeg2Rad);                         // kernels.cuh:205
        r.c = (float)cos(fa * kDeg2Rad);
        r.sp = sinf((float)(phs * kDeg2Rad));                    // kernels.cuh:184-185
        r.cp = cosf((float)(phs * kDeg2Rad));
        r.sd = sin((double)fa * kDeg2Rad);
        r.cd = cos((double)fa * kDeg2Rad);
        r.spd = sin((double)phs * kDeg2Rad);
        r.cpd = cos((double)phs * kDeg2Rad);
    }
    {
        auto add = [&](const void *p, size_t n) {
            const char *c = (const char *)p;
            sq->traj_sig.insert(sq->traj_sig.end(), c, c + n);
        };
        const uint32_t counts[5] = {ns, q->n_rf, q->n_te, q->n_dephasing, q->n_gradient};
        add(counts, sizeof counts);
        add(q->diffusivity, ns * sizeof(double));
        add(q->pXY, (size_t)ns * ns * sizeof(float));
        add(T1s.data(), ns * sizeof(float)); add(T2s.data(), ns * sizeof(float));
        add(T1d.data(), ns * sizeof(double)); add(T2d.data(), ns * sizeof(double));
        add(q->RF_t, q->n_rf * sizeof(int32_t));
        if (q->n_te) add(q->TE_t, q->n_te * sizeof(int32_t));
        if (q->n_dephasing) add(q->dephasing_t, q->n_dephasing * sizeof(int32_t));
        if (q->n_gradient) add(q->gradient_t, q->n_gradient * sizeof(int32_t));
    }
    hipStream_t s = e->stream;
    SW_HIP(e, dev_upload(&sq->diffusivity, q->diffusivity, ns, s));
    SW_HIP(e, dev_upload(&sq->pXY, q->pXY, (size_t)ns * ns, s));
    SW_HIP(e, dev_upload(&sq->T1s, T1s.data(), ns, s));
    SW_HIP(e, dev_upload(&sq->T2s, T2s.data(), ns, s));
    SW_HIP(e, dev_upload(&sq->T1d, T1d.data(), ns, s));
    SW_HIP(e, dev_upload(&sq->T2d, T2d.data(), ns, s));
    SW_HIP(e, dev_upload(&sq->rf, rf.data(), rf.size(), s));
    SW_HIP(e, dev_upload(&sq->TE_t, q->TE_t, q->n_te, s));
    SW_HIP(e, dev_upload(&sq->deph_t, q->dephasing_t, q->n_dephasing, s));
    SW_HIP(e, dev_upload(&sq->deph_deg, q->dephasing_deg, q->n_dephasing, s));
    SW_HIP(e, dev_upload(&sq->grad_t, q->gradient_t, q->n_gradient, s));
    SW_HIP(e, dev_upload(&sq->gx, q->gradientX_mTm, q->n_gradient, s));
    SW_HIP(e, dev_upload(&sq->gy, q->gradientY_mTm, q->n_gradient, s));
    SW_HIP(e, dev_upload(&sq->gz, q->gradientZ_mTm, q->n_gradient, s));
    SW_HIP(e, hipStreamSynchronize(s));
    int rc = ensure_exc(e, sq.get(), p, s);
    if (rc) return rc;
    *out = sq.release();
    return SW_OK;
}

SW_API void sw_sequence_destroy(sw_sequence *q) {
    if (!q) return;
    (void)hipSetDevice(q->device);
    for (void *p : {(void *)q->diffusivity, (void *)q->pXY, (void *)q->T1s, (void *)q->T2s, (void *)q->T1d,
                    (void *)q->T2d, (void *)q->rf, (void *)q->TE_t, (void *)q->deph_t, (void *)q->deph_deg,
                    (void *)q->grad_t, (void *)q->gx, (void *)q->gy, (void *)q->gz, (void *)q->exc})
        dev_free(p);
    delete q;
}

// ---------------------------------------------------------------------------
// the walk
// ---------------------------------------------------------------------------
namespace {

// Visiting order: local spin indices sorted by the Morton index of the 8^3-voxel
// brick holding their start position.  Neighbouring lanes then gather from
// neighbouring voxels, and the XCD-aware block mapping of the walk kernel gives
// each XCD (own L2) a compact region of the phantom.
// Writes the visiting order of a's spins into out[0, n_local).
int build_order_into(sw_engine *e, const sw::WalkArgs &a, uint32_t *out, hipStream_t s,
                     int lb = sw::kOrderBrickLog2) {
    const uint32_t n = a.n_local;
    const sw::OrderPlan plan = sw::order_plan(n, a.nx, a.ny, a.nz, lb);
    if (plan.total_bytes > e->sort_bytes) {
        SW_HIP(e, hipStreamSynchronize(s));
        dev_free(e->d_sort);
        e->d_sort = nullptr;
        SW_HIP(e, hipMalloc(&e->d_sort, plan.total_bytes));
        e->sort_bytes = plan.total_bytes;
    }
    SW_HIP(e, sw::launch_spatial_order(plan, e->d_sort, a.XYZ0, a.xyz_scale, n, a.s2g, a.nx, a.ny, a.nz, out, s));
    return SW_OK;
}

int build_order(sw_engine *e, const sw::WalkArgs &a, hipStream_t s) {
    const uint32_t n = a.n_local;
    if (n > e->order_cap) {
        SW_HIP(e, hipStreamSynchronize(s));
        dev_free(e->d_order);
        e->d_order = nullptr;
        SW_HIP(e, hipMalloc((void **)&e->d_order, n * sizeof(uint32_t)));
        e->order_cap = n;
    }
    return build_order_into(e, a, e->d_order, s);
}

// Checks one walk's arguments and fills its WalkArgs, except exc, order and
// counters.  *field_out receives the record format the walk gathers.
int fill_walk_args(sw_engine *e, const sw_phantom *ph, const sw_sequence *q, const sw_params *p,
                   const sw_walk_opts *o, const float *d_M0, const float *d_XYZ0, void *d_M1, float *d_XYZ1,
                   uint8_t *d_T, sw::WalkArgs &a, int *field_out) {
    if (!ph || !q || !p || !o) return fail(e, SW_E_INVALID, "null argument");
    // ---- consistency checks (the reference trusts its caller; an ABI may not)
    for (int i = 0; i < 3; i++)
        if (p->phantom_size[i] != ph->dims[i]) return fail(e, SW_E_INVALID, "phantom_size does not match phantom");
    if (p->matrix_length != (int64_t)ph->n) return fail(e, SW_E_INVALID, "matrix_length does not match phantom");
    if (p->n_substrate != q->n_sub) return fail(e, SW_E_INVALID, "n_substrate does not match sequence tables");
    if (ph->n_labels > q->n_sub)                                  // monte_carlo.cu:113-118
        return fail(e, SW_E_INVALID, "mask has " + std::to_string(ph->n_labels) + " labels but only " +
                                         std::to_string(q->n_sub) + " substrates are configured");
    if (p->timestep_us <= 0) return fail(e, SW_E_INVALID, "timestep_us must be positive");
    if (p->n_timepoints >= (1u << 31))            // the walk keeps a flag in the top bit of its step count
        return fail(e, SW_E_INVALID, "n_timepoints must be below 2^31");
    if (p->n_dummy_scan < 0) return fail(e, SW_E_INVALID, "n_dummy_scan < 0: call sw_params_prepare first");
    if (!(p->fov[0] > 0 && p->fov[1] > 0 && p->fov[2] > 0)) return fail(e, SW_E_INVALID, "fov must be positive");
    if ((uint64_t)o->spin_offset + o->n_local > p->n_spins)
        return fail(e, SW_E_INVALID, "spin_offset + n_local exceeds n_spins");
    if (o->n_local && (!d_M0 || !d_XYZ0 || !d_XYZ1 || (q->n_te && (!d_M1 || !d_T))))
        return fail(e, SW_E_INVALID, "spin buffers are required");
    if (o->precision < 0 || o->precision > 2) return fail(e, SW_E_INVALID, "unknown precision");
    if (o->rng == SW_RNG_REPLAY && (!o->replay_normals || !o->replay_uniforms || o->replay_normal_stride < 3))
        return fail(e, SW_E_INVALID, "replay mode needs normal and uniform streams");
    if (o->rng != SW_RNG_REPLAY && o->rng != SW_RNG_PHILOX) return fail(e, SW_E_INVALID, "unknown rng");
    int field = ph->has_field ? o->field_format : SW_FIELD_NONE;
    if (field < 0 || field > SW_FIELD_Q16) return fail(e, SW_E_INVALID, "unknown field format");
    if (!ph->rec[field]) return fail(e, SW_E_INVALID, "the phantom was not created with this field format");
    if (o->precision == SW_PREC_F64 && field > SW_FIELD_F32)
        return fail(e, SW_E_INVALID, "the FP64 path gathers the float32 field");
    *field_out = field;

    std::memset(&a, 0, sizeof a);
    a.rec = ph->rec[field];
    const bool packed = field == SW_FIELD_F16 || field == SW_FIELD_BF16 || field == SW_FIELD_Q16 ||
                        (field == SW_FIELD_F32 && o->rng == SW_RNG_PHILOX);
    if (packed) {
        if (field == SW_FIELD_F32) {
            // a single production walk gathers the float32 field from the
            // structure-of-arrays bricks (a batch swaps in the packed records)
            if (!ph->rec_soa || !ph->rec_packed) return fail(e, SW_E_INVALID, "the phantom has no production float32 records");
            a.rec = ph->rec_soa;
        }
        const sw::PackShape &S = field == SW_FIELD_F32 ? sw::kPackF32 : sw::kPack16;
        const bool sup = sw::pack_super(ph->dims[0], ph->dims[1], ph->dims[2], S);
        a.super_blocks = sup ? 1u : 0u;
        a.nb_y = (uint32_t)sw::pack_pitch_y(ph->dims[1], S, sup);
        a.nb_z = (uint32_t)sw::pack_pitch_z(ph->dims[2], S, sup);
        // the walk forms the records' offsets with 32-bit strides
        if ((uint64_t)a.nb_y * a.nb_z * (sup ? sw::kSuperLines : 1) * 128 >= (1ull << 32))
            return fail(e, SW_E_INVALID, "the phantom's y-z planes are too large for the packed records (4 GiB a slab)");
    } else {
        const sw::BrickShape sh = field == SW_FIELD_F32 ? sw::kBrick8 : sw::kBrick4;
        a.nb_y = (uint32_t)sw::brick_count(ph->dims[1], sh.lby);
        a.nb_z = (uint32_t)sw::brick_count(ph->dims[2], sh.lbz);
    }
    a.t2d = sw_tesla2deg(p->B0, p->timestep_us);
    a.h2d = (field == SW_FIELD_Q16 ? ph->q_scale : ph->half_scale) * a.t2d;
    a.t2d_f64 = (double)p->B0 * p->timestep_us * 1e-6 * kGamma * kRad2Deg;
    a.nx = (int64_t)ph->dims[0]; a.ny = (int64_t)ph->dims[1]; a.nz = (int64_t)ph->dims[2];
    for (int i = 0; i < 3; i++) a.nd[i] = (double)ph->dims[i];
    a.matrix_length = p->matrix_length;
    for (int i = 0; i < 3; i++) {
        a.fov[i] = p->fov[i];
        a.s2g[i] = (double)p->phantom_size[i] / p->fov[i];       // kernels.cu:98
        a.s2gf[i] = (float)a.s2g[i];
        a.vs[i] = p->fov[i] / (double)p->phantom_size[i];
        a.vsf[i] = (float)a.vs[i];
        a.nf[i] = (float)p->phantom_size[i];
    }
    a.g2d = (float)(1e-3 * p->timestep_us * 1e-6 * kGamma * kRad2Deg);
    a.diffusivity = q->diffusivity; a.pXY = q->pXY; a.T1s = q->T1s; a.T2s = q->T2s; a.T1d = q->T1d; a.T2d = q->T2d;
    a.n_sub = q->n_sub;
    // The tissue tables go to LDS beside the per-axis tables of the 32-bit mode
    // when both fit its 64 KiB (a walk in the 64-bit mode ignores the flag).
    a.tissue_lds = (o->rng != SW_RNG_REPLAY && q->n_sub <= sw::kTissueLdsMax &&
                    ((uint64_t)p->phantom_size[0] + p->phantom_size[1] + p->phantom_size[2] +
                     (uint64_t)q->n_sub * q->n_sub + 5ull * q->n_sub) * 4 <= 65536) ? 1u : 0u;
    a.rf = q->rf; a.n_rf = q->n_rf;
    a.TE_t = q->TE_t; a.n_te = q->n_te;
    a.deph_t = q->deph_t; a.deph_deg = q->deph_deg; a.n_deph = q->n_deph;
    a.grad_t = q->grad_t; a.gx = q->gx; a.gy = q->gy; a.gz = q->gz; a.n_grad = q->n_grad;
    a.exc_s = p->s; a.exc_c = p->c;
    a.exc_sd = sin((double)q->RF_FA0 * kDeg2Rad);
    a.exc_cd = cos((double)q->RF_FA0 * kDeg2Rad);
    a.n_timepoints = p->n_timepoints;
    a.n_dummy = (uint32_t)p->n_dummy_scan;
    a.timestep_us = p->timestep_us;
    a.n_spins_total = p->n_spins;
    a.spin_offset = o->spin_offset;
    a.n_local = o->n_local;
    a.seed = p->seed;
    a.max_iterations = p->max_iterations;
    a.cross_fov = p->enCrossFOV ? 1 : 0;
    a.record_traj = p->enRecordTrajectory ? 1 : 0;
    a.xyz_scale = o->xyz_scale;
    a.gradient_scale = o->gradient_scale;
    a.traj_len = p->enRecordTrajectory ? (uint64_t)(p->n_dummy_scan + 1) * p->n_timepoints : 1;
    a.M0 = d_M0; a.XYZ0 = d_XYZ0; a.M1 = d_M1; a.XYZ1 = d_XYZ1; a.T = d_T;
    a.rn = o->replay_normals; a.rn_stride = o->replay_normal_stride;
    a.ru = o->replay_uniforms; a.ru_stride = o->replay_uniform_stride;
    // The step loop of a single production walk: two attempts per trip when the
    // records it gathers exceed the 256 MiB Infinity Cache, where each gather
    // waits on HBM and a second one in flight pays (both loops give the same
    // results).  Replayed walks and batches take one attempt per trip.
    if (o->step_loop < SW_STEP_LOOP_AUTO || o->step_loop > SW_STEP_LOOP_TWO)
        return fail(e, SW_E_INVALID, "unknown step_loop");
    a.pairs = 0;
    if (o->rng == SW_RNG_PHILOX) {
        if (o->step_loop == SW_STEP_LOOP_TWO) a.pairs = 1;
        else if (o->step_loop == SW_STEP_LOOP_AUTO) a.pairs = sw::walk_mode(a, field, 0).bytes > kPairsMinBytes ? 1u : 0u;
    }
    return SW_OK;
}

}  // namespace

SW_API int sw_walk_device(sw_engine *e, const sw_phantom *ph, const sw_sequence *cq, const sw_params *p,
                          const sw_walk_opts *o, const float *d_M0, const float *d_XYZ0, void *d_M1,
                          float *d_XYZ1, uint8_t *d_T, void *stream, sw_walk_stats *stats) {
    if (!e || !ph || !cq || !p || !o) return fail(e, SW_E_INVALID, "null argument");
    sw_sequence *q = const_cast<sw_sequence *>(cq);
    hipStream_t s = stream ? (hipStream_t)stream : e->stream;
    SW_HIP(e, hipSetDevice(e->device));
    sw::WalkArgs a;
    int field = 0;
    int rc = fill_walk_args(e, ph, q, p, o, d_M0, d_XYZ0, d_M1, d_XYZ1, d_T, a, &field);
    if (rc) return rc;
    // the counters, the visiting order and the excitation table are engine-owned
    if ((rc = scratch_acquire(e, s))) return rc;
    rc = ensure_exc(e, q, p, s);
    if (rc) return rc;
    a.exc = q->exc;
    a.counters = e->d_counters;
    a.order = nullptr;
    if (o->spatial_order && o->n_local > 1) {
        rc = build_order(e, a, s);
        if (rc) return rc;
        a.order = e->d_order;
    }
    SW_HIP(e, sw::launch_zero_u64(e->d_counters, 4, s));
    if (stats) SW_HIP(e, hipEventRecord(e->ev0, s));
    SW_HIP(e, sw::launch_walk(a, o->rng == SW_RNG_REPLAY ? 1 : 0, field, o->precision, s));
    unsigned long long c[4];
    if (stats) {
        SW_HIP(e, hipEventRecord(e->ev1, s));
        SW_HIP(e, hipMemcpyAsync(c, e->d_counters, sizeof c, hipMemcpyDeviceToHost, s));
    }
    if ((rc = scratch_release(e, s))) return rc;
    if (stats) {
        SW_HIP(e, hipStreamSynchronize(s));
        float ms = 0;
        SW_HIP(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
        stats->n_lost = c[0]; stats->n_replay_exhausted = c[1]; stats->n_steps = c[2]; stats->n_attempts = c[3];
        stats->kernel_ms = ms;
        if (c[1]) return fail(e, SW_E_REPLAY, std::to_string(c[1]) + " replayed spins ran out of recorded draws");
    }
    return SW_OK;
}

// The walk of sw_walk_device with no allocation: the counters, visiting order
// and sort buffers sit in the fixed device scratch of noalloc.hip instead of
// engine-owned memory.  Layout: [0, 256) counters, then the visiting order
// (4 bytes a spin), then the sort's buffers.  Nothing here allocates device
// memory; the sequence's excitation table must already match p (it is built
// by sw_sequence_create for the parameters given there, and by any earlier
// sw_walk_device call with these phase-cycling parameters).
SW_API int sw_walk_device_noalloc(sw_engine *e, const sw_phantom *ph, const sw_sequence *q, const sw_params *p,
                                  const sw_walk_opts *o, const float *d_M0, const float *d_XYZ0, void *d_M1,
                                  float *d_XYZ1, uint8_t *d_T, void *stream, sw_walk_stats *stats) {
    if (!e || !ph || !q || !p || !o) return fail(e, SW_E_INVALID, "null argument");
    hipStream_t s = stream ? (hipStream_t)stream : e->stream;
    SW_HIP(e, hipSetDevice(e->device));
    if (e->na_err != hipSuccess || !e->na_scratch)
        return fail(e, SW_E_HIP, std::string("no-allocation scratch unavailable: ") + hipGetErrorString(e->na_err));
    sw::WalkArgs a;
    int field = 0;
    int rc = fill_walk_args(e, ph, q, p, o, d_M0, d_XYZ0, d_M1, d_XYZ1, d_T, a, &field);
    if (rc) return rc;
    if (!q->exc_valid || q->key_lin != p->linear_phase_cycling || q->key_quad != p->quadratic_phase_cycling ||
        q->key_ndummy != p->n_dummy_scan)
        return fail(e, SW_E_INVALID, "the sequence's excitation table was built for other phase-cycling or dummy-scan "
                                     "parameters; create the sequence with these parameters");
    const size_t cap = sw::noalloc_scratch_bytes();
    char *base = (char *)e->na_scratch;
    a.exc = q->exc;
    a.counters = (unsigned long long *)base;
    a.order = nullptr;
    if (o->spatial_order && o->n_local > 1) {
        const sw::OrderPlan plan = sw::order_plan(o->n_local, a.nx, a.ny, a.nz);
        const size_t order_bytes = ((size_t)o->n_local * 4 + 255) & ~(size_t)255;
        if (256 + order_bytes + plan.total_bytes > cap)
            return fail(e, SW_E_INVALID, "n_local = " + std::to_string(o->n_local) + " with spatial_order needs " +
                                             std::to_string(256 + order_bytes + plan.total_bytes) +
                                             " bytes of scratch; the no-allocation scratch holds " +
                                             std::to_string(cap) + " (use spatial_order 0 or sw_walk_device)");
        uint32_t *order = (uint32_t *)(base + 256);
        SW_HIP(e, sw::launch_spatial_order(plan, base + 256 + order_bytes, a.XYZ0, a.xyz_scale, o->n_local, a.s2g,
                                           a.nx, a.ny, a.nz, order, s));
        a.order = order;
    }
    SW_HIP(e, sw::launch_zero_u64(a.counters, 4, s));
    if (stats) SW_HIP(e, hipEventRecord(e->ev0, s));
    SW_HIP(e, sw::launch_walk(a, o->rng == SW_RNG_REPLAY ? 1 : 0, field, o->precision, s));
    if (stats) {
        unsigned long long c[4];
        SW_HIP(e, hipEventRecord(e->ev1, s));
        SW_HIP(e, hipMemcpyAsync(c, a.counters, sizeof c, hipMemcpyDeviceToHost, s));
        SW_HIP(e, hipStreamSynchronize(s));
        float ms = 0;
        SW_HIP(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
        stats->n_lost = c[0]; stats->n_replay_exhausted = c[1]; stats->n_steps = c[2]; stats->n_attempts = c[3];
        stats->kernel_ms = ms;
        if (c[1]) return fail(e, SW_E_REPLAY, std::to_string(c[1]) + " replayed spins ran out of recorded draws");
    }
    return SW_OK;
}

// ---------------------------------------------------------------------------
// batched walk
// ---------------------------------------------------------------------------
struct sw_batch {
    sw_engine *eng = nullptr;
    int device = 0;
    uint32_t n_items = 0;
    sw::WalkArgs *d_items = nullptr;             // single walks
    sw::FusedArgs *d_fused = nullptr;            // fused walks of up to kMaxViews tiles
    sw::BatchChunk *d_chunks = nullptr;
    unsigned long long *d_counters = nullptr;    // [n_items][4]
    sw::Excitation *d_exc = nullptr;             // every tile's excitation table
    uint32_t *d_order = nullptr;                 // every tile's visiting order
    double *d_partial = nullptr;                 // echo-reduction scratch
    struct Group { bool cross; int views; uint32_t offset, n_slots; };
    std::vector<Group> groups;
    struct Reduce { const void *M1; int f64; const uint8_t *T; uint32_t n, n_te, n_sub; double *out; };
    std::vector<Reduce> reduces;
    int rng = 0, field = 0, prec = 0;
    bool idx32 = false;
    bool mixed = false;                          // 64-bit mode over items in both packed layouts
    size_t lds = 0;                              // LDS of a workgroup: per-axis tables, then tissue tables (32-bit mode)
    // what the chunk table is built from: per walk unit (a single tile, or a
    // fused group of tiles), its CROSS_FOV group, spins, steps per spin, phantom
    // records, views (1: d_items[index]; kMaxViews: d_fused[index])
    struct Tile { bool cross; uint32_t n_local; uint64_t cost; const void *rec; int views; uint32_t index; };
    std::vector<Tile> tiles;
    uint32_t n_fused_tiles = 0;                  // tiles walked inside fused units
    sw_batch_schedule sched{};
    size_t chunks_cap = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // The chunk table, counters and reduction scratch belong to the batch: a run
    // on one stream waits for the previous run on another (run_ev), and a new
    // schedule waits for the last run before it rewrites the chunk table.
    hipEvent_t run_ev = nullptr;
    hipStream_t run_stream = nullptr;
    // Visiting orders: each ordered tile's arguments and its slice of d_order,
    // rebuilt in place when a schedule changes the order brick (order_lb).
    std::vector<std::pair<sw::WalkArgs, uint32_t *>> orders;
    int order_lb = sw::kOrderBrickLog2;
    size_t lds_cu = 0, lds_block = 0;            // LDS per CU and per workgroup, bytes
};

namespace {

// Counter reset, the walk of every group and the echo sums, in stream order.
// The LDS of a batch workgroup: the per-axis tables, raised so that at most
// waves_per_simd waves fit each SIMD when the schedule asks for a cap.  A CU
// has 4 SIMDs and a workgroup chunk_spins / 64 waves, so the cap is
// k = 256 * waves_per_simd / chunk_spins workgroups per CU; k fit when
// k * L <= LDS per CU < (k + 1) * L, so L = LDS per CU / (k + 1/2).
size_t batch_lds(const sw_batch &b, const sw_batch_schedule &sc) {
    size_t lds = b.idx32 ? b.lds : 0;
    if (sc.waves_per_simd && b.lds_cu && sc.chunk_spins) {
        const size_t k = (size_t)256 * sc.waves_per_simd / sc.chunk_spins;
        const size_t l = (size_t)(2 * b.lds_cu / (2 * k + 1)) & ~(size_t)1023;
        lds = std::max(lds, l);
    }
    return lds;
}

hipError_t batch_enqueue(const sw_batch *b, hipStream_t s) {
    hipError_t err = sw::launch_zero_u64(b->d_counters, (int)(4 * b->n_items), s);
    if (err != hipSuccess) return err;
    const size_t lds = batch_lds(*b, b->sched);
    for (const auto &g : b->groups) {
        const void *items = g.views == 1 ? (const void *)b->d_items : (const void *)b->d_fused;
        err = sw::launch_walk_batch(items, b->d_chunks + g.offset, g.n_slots, b->sched.chunk_spins,
                                    b->sched.spins_per_lane, g.views, b->rng, b->field, b->prec, g.cross,
                                    b->idx32 ? sw::ADDR_32 : b->mixed ? sw::ADDR_64_MIXED : sw::ADDR_64,
                                    lds, s);
        if (err != hipSuccess) return err;
    }
    for (const auto &r : b->reduces) {
        err = sw::launch_echo_reduce(r.M1, r.f64, r.T, r.n, r.n_te, r.n_sub, b->d_partial, r.out, s);
        if (err != hipSuccess) return err;
    }
    return hipSuccess;
}

// The default batched schedule: chunks of 256 spins; tiles of one phantom
// interleave chunk by chunk in relative position (2^20 bands: finer than any
// part), so tiles that walk the same trajectories (phase-cycling or gradient
// scales of one seed) run the same spins side by side and share every line;
// tiles of different phantoms take turns in 4 bands.  Measured: the 1000^3
// bSSFP batch (4 tiles, one phantom) 2.676 s tile after tile, 2.511 s with 64
// bands, 2.029 s chunk by chunk; the 64-phantom batch, walked tile by tile,
// 163.5 ms tile after tile, 151.2 ms with 16 bands, 155.2 ms with 64; with
// fused walks the tuner picks 4 bands for FP32 (70.4 ms against 72.2 ms at 16)
// and 1 for Kahan + int16 (66.5 against 68.3).  sw_batch_tune searches around it.
constexpr sw_batch_schedule kDefaultSchedule = {256u, 1u << 20, 4u, 1u, 0u, 8u};

// The chunk table of a schedule: per CROSS_FOV group and walk kind (single or
// fused), each walk unit's chunks of sc.chunk_spins spins split into 8
// contiguous parts (one per XCD; a part is one spatial region of the unit,
// since chunks follow the visiting order).  Within
// an XCD's list, tiles of one cost class (a factor of two in steps per spin)
// take turns band by band through their parts, longer classes first.  The
// chunks resident on an XCD then start in one band of each tile's part: a
// narrow window of its phantom, whose lines stay in L2, and the same window for
// tiles that share a phantom.  A class whose tiles all walk one phantom uses
// sc.bands_shared bands per part, any other class sc.bands_mixed.  *kinds
// receives bit 0 if some class is shared, bit 1 if some class is mixed.
void build_chunk_table(const sw_batch &b, const sw_batch_schedule &sc, std::vector<sw::BatchChunk> &table,
                       std::vector<sw_batch::Group> &groups, unsigned *kinds) {
    table.clear();
    groups.clear();
    *kinds = 0;
    const uint32_t cs = sc.chunk_spins * sc.spins_per_lane;
    for (int grp = 0; grp < 4; grp++) {
        const int cross = grp & 1, views = grp < 2 ? 1 : sw::kMaxViews;
        std::vector<uint32_t> idx;
        for (uint32_t i = 0; i < (uint32_t)b.tiles.size(); i++)
            if (b.tiles[i].cross == (cross != 0) && b.tiles[i].views == views && b.tiles[i].n_local) idx.push_back(i);
        if (idx.empty()) continue;
        auto cost = [&](uint32_t i) { return b.tiles[i].cost; };
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) { return cost(x) > cost(y); });
        auto cls_of = [&](uint32_t i) {
            uint32_t c = 0;
            for (uint64_t v = cost(i); v > 1; v >>= 1) c++;
            return c;
        };
        auto shared = [&](uint32_t cls) {
            const void *rec = nullptr;
            for (uint32_t i : idx) {
                if (cls_of(i) != cls) continue;
                if (rec && b.tiles[i].rec != rec) return false;
                rec = b.tiles[i].rec;
            }
            return true;
        };
        struct Key { uint32_t cls, band, rank, item, c; };
        std::vector<Key> keys[8];
        for (uint32_t r = 0; r < (uint32_t)idx.size(); r++) {
            const uint32_t i = idx[r];
            const uint32_t cls = cls_of(i);
            const bool sh = shared(cls);
            *kinds |= sh ? 1u : 2u;
            const uint32_t bands = sh ? sc.bands_shared : sc.bands_mixed;
            const uint32_t nc = (b.tiles[i].n_local + cs - 1) / cs;
            for (uint32_t k = 0; k < 8; k++) {
                const uint32_t c0 = (uint32_t)((uint64_t)nc * k / 8), c1 = (uint32_t)((uint64_t)nc * (k + 1) / 8);
                for (uint32_t c = c0; c < c1; c++) {
                    const uint32_t band = (uint32_t)((uint64_t)(c - c0) * bands / (c1 - c0));
                    keys[k].push_back({cls, band, r, b.tiles[i].index, c});
                }
            }
        }
        std::vector<sw::BatchChunk> lists[8];
        size_t per = 0;
        for (uint32_t k = 0; k < 8; k++) {
            std::stable_sort(keys[k].begin(), keys[k].end(), [](const Key &x, const Key &y) {
                if (x.cls != y.cls) return x.cls > y.cls;
                if (x.band != y.band) return x.band < y.band;
                return x.rank < y.rank;
            });
            for (const Key &q : keys[k]) lists[k].push_back({q.item, q.c * cs});
            per = std::max(per, lists[k].size());
        }
        sw_batch::Group g;
        g.cross = cross != 0;
        g.views = views;
        g.offset = (uint32_t)table.size();
        g.n_slots = (uint32_t)(8 * per);
        for (auto &l : lists) {
            l.resize(per, sw::BatchChunk{sw::kNoItem, 0});
            table.insert(table.end(), l.begin(), l.end());
        }
        groups.push_back(g);
    }
}

bool schedule_ok(const sw_batch_schedule &sc) {
    return (sc.chunk_spins == 64 || sc.chunk_spins == 128 || sc.chunk_spins == 256) && sc.bands_shared >= 1 &&
           sc.bands_mixed >= 1 && (sc.spins_per_lane == 1 || sc.spins_per_lane == 2) &&
           (sc.waves_per_simd == 0 || (sc.waves_per_simd >= 2 && sc.waves_per_simd <= 8)) &&
           (sc.order_brick == 4 || sc.order_brick == 8 || sc.order_brick == 16 || sc.order_brick == 32);
}

int log2_of(uint32_t v) {
    int l = 0;
    while ((1u << l) < v) l++;
    return l;
}

// Makes sc the batch's schedule: builds and uploads its chunk table and, when
// `capture`, captures the batch as one HIP graph (counter reset, walks, echo
// sums).  Without capture the batch runs by plain launches until the next call.
int apply_schedule(sw_engine *e, sw_batch *b, const sw_batch_schedule &sc, hipStream_t s, bool capture) {
    std::vector<sw::BatchChunk> table;
    std::vector<sw_batch::Group> groups;
    unsigned kinds = 0;
    build_chunk_table(*b, sc, table, groups, &kinds);
    if (batch_lds(*b, sc) > b->lds_block)
        return fail(e, SW_E_INVALID, "waves_per_simd " + std::to_string(sc.waves_per_simd) +
                                         " needs more LDS per workgroup than the device allows");
    // the previous table and orders may be in use by a run on any stream
    if (b->run_stream) SW_HIP(e, hipEventSynchronize(b->run_ev));
    SW_HIP(e, hipStreamSynchronize(s));
    const int lb = log2_of(sc.order_brick);
    if (lb != b->order_lb && !b->orders.empty()) {
        int rc = scratch_acquire(e, s);              // build_order_into uses the engine's sort scratch
        if (rc) return rc;
        for (const auto &o : b->orders)
            if ((rc = build_order_into(e, o.first, o.second, s, lb))) return rc;
        if ((rc = scratch_release(e, s))) return rc;
        SW_HIP(e, hipStreamSynchronize(s));
    }
    b->order_lb = lb;
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    if (table.size() > b->chunks_cap) {
        dev_free(b->d_chunks);
        b->d_chunks = nullptr;
        b->chunks_cap = 0;
        SW_HIP(e, hipMalloc((void **)&b->d_chunks, table.size() * sizeof(sw::BatchChunk)));
        b->chunks_cap = table.size();
    }
    if (!table.empty())
        SW_HIP(e, hipMemcpyAsync(b->d_chunks, table.data(), table.size() * sizeof(sw::BatchChunk),
                                 hipMemcpyHostToDevice, s));
    SW_HIP(e, hipStreamSynchronize(s));          // the host table goes out of scope
    b->groups = groups;
    b->sched = sc;
    if (!capture) return SW_OK;
    // A batch runs from its graph or not at all: a failed capture is an error,
    // not a quiet fall-back to stream launches.
    const hipError_t be = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (be != hipSuccess)
        return fail(e, SW_E_HIP, std::string("batch graph capture could not begin: ") + hipGetErrorString(be));
    const hipError_t le = batch_enqueue(b, s);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(s, &g);
    if (le != hipSuccess || ce != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return fail(e, SW_E_HIP, std::string("batch graph capture failed: ") +
                                     hipGetErrorString(le != hipSuccess ? le : ce));
    }
    b->graph = g;
    const hipError_t ie = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        b->exec = nullptr;
        return fail(e, SW_E_HIP, std::string("batch graph instantiation failed: ") + hipGetErrorString(ie));
    }
    return SW_OK;
}

unsigned schedule_kinds(const sw_batch &b) {
    std::vector<sw::BatchChunk> table;
    std::vector<sw_batch::Group> groups;
    unsigned kinds = 0;
    build_chunk_table(b, b.sched, table, groups, &kinds);
    return kinds;
}

// Whether two tiles of a batch walk the same trajectories: the same draws
// (seed, spins, draw source), geometry, start positions, tissue tables and event
// times, with outputs apart.  Only RF and excitation parameters, gradient
// amplitudes, dephasing and the output buffers may differ.
bool same_trajectory(const sw_batch_item &ix, const sw::WalkArgs &x, const sw_batch_item &iy,
                     const sw::WalkArgs &y) {
    auto eq = [](const void *p, const void *q, size_t n) { return std::memcmp(p, q, n) == 0; };
    return ix.sequence->traj_sig == iy.sequence->traj_sig && ix.opts.spatial_order == iy.opts.spatial_order &&
           x.rec == y.rec && x.nb_y == y.nb_y && x.nb_z == y.nb_z && eq(&x.t2d, &y.t2d, sizeof x.t2d) &&
           eq(&x.h2d, &y.h2d, sizeof x.h2d) && eq(&x.t2d_f64, &y.t2d_f64, sizeof x.t2d_f64) && x.nx == y.nx &&
           x.ny == y.ny && x.nz == y.nz && x.matrix_length == y.matrix_length && eq(x.fov, y.fov, sizeof x.fov) &&
           eq(x.s2g, y.s2g, sizeof x.s2g) && x.n_timepoints == y.n_timepoints && x.n_dummy == y.n_dummy &&
           x.timestep_us == y.timestep_us && x.n_spins_total == y.n_spins_total && x.spin_offset == y.spin_offset &&
           x.n_local == y.n_local && x.seed == y.seed && x.max_iterations == y.max_iterations &&
           x.cross_fov == y.cross_fov && x.record_traj == y.record_traj && eq(&x.xyz_scale, &y.xyz_scale, 4) &&
           x.traj_len == y.traj_len && x.M0 == y.M0 && x.XYZ0 == y.XYZ0 && x.rn == y.rn &&
           x.rn_stride == y.rn_stride && x.ru == y.ru && x.ru_stride == y.ru_stride;
}

void batch_free(sw_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->run_ev) {
        (void)hipEventSynchronize(b->run_ev);
        (void)hipEventDestroy(b->run_ev);
    }
    for (void *x : {(void *)b->d_items, (void *)b->d_fused, (void *)b->d_chunks, (void *)b->d_counters,
                    (void *)b->d_exc, (void *)b->d_order, (void *)b->d_partial})
        dev_free(x);
    delete b;
}

}  // namespace

SW_API int sw_batch_create(sw_engine *e, const sw_batch_item *items, uint32_t n_items, sw_batch **out) {
    if (!e || !out || (n_items && !items)) return fail(e, SW_E_INVALID, "null argument");
    *out = nullptr;
    if (n_items == 0) return fail(e, SW_E_INVALID, "empty batch");
    SW_HIP(e, hipSetDevice(e->device));
    hipStream_t s = e->stream;
    std::unique_ptr<sw_batch, void (*)(sw_batch *)> b(new sw_batch, batch_free);
    b->eng = e;
    b->device = e->device;
    b->n_items = n_items;
    const sw_walk_opts &o0 = items[0].opts;
    b->rng = o0.rng == SW_RNG_REPLAY ? 1 : 0;
    b->prec = o0.precision;

    // ---- per-tile arguments, checked as sw_walk_device checks them
    std::vector<sw::WalkArgs> args(n_items);
    std::vector<sw::Excitation> exc;
    std::vector<size_t> exc_off(n_items), order_off(n_items);
    size_t n_order = 0, partial = 0;
    b->idx32 = true;
    for (uint32_t i = 0; i < n_items; i++) {
        const sw_batch_item &it = items[i];
        if (it.opts.rng != o0.rng || it.opts.precision != o0.precision)
            return fail(e, SW_E_INVALID, "tile " + std::to_string(i) + ": every tile of a batch shares rng and precision");
        int field = 0;
        int rc = fill_walk_args(e, it.phantom, it.sequence, &it.params, &it.opts, it.d_M0, it.d_XYZ0, it.d_M1,
                                it.d_XYZ1, it.d_T, args[i], &field);
        if (rc) { e->err = "tile " + std::to_string(i) + ": " + e->err; return rc; }
        // batches gather the float32 field from the packed records (same bricks)
        if (field == SW_FIELD_F32 && it.opts.rng == SW_RNG_PHILOX) args[i].rec = it.phantom->rec_packed;
        if (i == 0) b->field = field;
        else if (field != b->field)
            return fail(e, SW_E_INVALID, "tile " + std::to_string(i) + ": every tile of a batch gathers the same field format");
        const sw::WalkMode m = sw::walk_mode(args[i], field, b->rng);
        b->idx32 = b->idx32 && m.idx32;
        b->lds = std::max(b->lds, m.lds);
        exc_off[i] = exc.size();
        const std::vector<sw::Excitation> h = exc_table(it.sequence, &it.params);
        exc.insert(exc.end(), h.begin(), h.end());
        order_off[i] = n_order;
        if (it.opts.spatial_order && it.opts.n_local > 1) n_order += it.opts.n_local;
        if (it.d_echo_sums) {
            sw_batch::Reduce r;
            r.M1 = it.d_M1; r.f64 = it.opts.precision == SW_PREC_F64; r.T = it.d_T;
            r.n = it.opts.n_local; r.n_te = it.sequence->n_te; r.n_sub = it.sequence->n_sub; r.out = it.d_echo_sums;
            if (r.n_te && r.n_sub) {
                b->reduces.push_back(r);
                partial = std::max(partial, (size_t)sw::reduce_partial_len(r.n, r.n_te, r.n_sub));
            }
        }
    }
    if (!b->idx32) b->lds = 0;
    for (uint32_t i = 0; i < n_items && !b->idx32; i++) b->mixed = b->mixed || args[i].super_blocks;

    // ---- device tables
    SW_HIP(e, hipMalloc((void **)&b->d_counters, 4 * n_items * sizeof(unsigned long long)));
    SW_HIP(e, dev_upload(&b->d_exc, exc.data(), exc.size(), s));
    if (n_order) SW_HIP(e, hipMalloc((void **)&b->d_order, n_order * sizeof(uint32_t)));
    if (partial) SW_HIP(e, hipMalloc((void **)&b->d_partial, partial * sizeof(double)));
    int rc0 = scratch_acquire(e, s);             // build_order_into uses the engine's sort scratch
    if (rc0) return rc0;
    for (uint32_t i = 0; i < n_items; i++) {
        sw::WalkArgs &a = args[i];
        a.exc = b->d_exc + exc_off[i];
        a.counters = b->d_counters + 4 * (size_t)i;
        a.order = nullptr;
        if (items[i].opts.spatial_order && items[i].opts.n_local > 1) {
            uint32_t *dst = b->d_order + order_off[i];
            int rc = build_order_into(e, a, dst, s);
            if (rc) return rc;
            b->orders.push_back({a, dst});
            a.order = dst;
        }
    }
    // ---- walk units: tiles that walk the same trajectories fuse, up to
    // kMaxViews at a time (FP32 and Kahan phases; an FP64 view would not fit
    // the registers)
    std::vector<sw::WalkArgs> singles;
    std::vector<sw::FusedArgs> fused;
    std::vector<bool> taken(n_items, false);
    for (uint32_t i = 0; i < n_items; i++) {
        if (taken[i]) continue;
        taken[i] = true;
        std::vector<uint32_t> group{i};
        if (b->prec != SW_PREC_F64)
            for (uint32_t j = i + 1; j < n_items && group.size() < (size_t)sw::kMaxViews; j++)
                if (!taken[j] && same_trajectory(items[i], args[i], items[j], args[j])) {
                    taken[j] = true;
                    group.push_back(j);
                }
        const sw::WalkArgs &a0 = args[i];
        const uint64_t cost = (uint64_t)(a0.n_dummy + 1) * a0.n_timepoints;
        if (group.size() == 1) {
            b->tiles.push_back({a0.cross_fov != 0, a0.n_local, cost, a0.rec, 1, (uint32_t)singles.size()});
            singles.push_back(a0);
            continue;
        }
        sw::FusedArgs f;
        std::memset(&f, 0, sizeof f);
        f.a = a0;
        f.n_views = (uint32_t)group.size();
        for (int v = 1; v < sw::kMaxViews; v++) {
            const bool real = (size_t)v < group.size();
            const sw::WalkArgs &x = args[real ? group[v] : i];
            sw::WalkView &w = f.xv[v - 1];
            w.rf = x.rf; w.exc = x.exc; w.deph_deg = x.deph_deg; w.gx = x.gx; w.gy = x.gy; w.gz = x.gz;
            w.exc_s = x.exc_s; w.exc_c = x.exc_c; w.exc_sd = x.exc_sd; w.exc_cd = x.exc_cd;
            w.gradient_scale = x.gradient_scale;
            w.M1 = x.M1; w.XYZ1 = x.XYZ1; w.T = x.T;
            w.counters = real ? x.counters : nullptr;      // a repeat of view 0 counts nothing
        }
        b->tiles.push_back({a0.cross_fov != 0, a0.n_local, cost, a0.rec, sw::kMaxViews, (uint32_t)fused.size()});
        fused.push_back(f);
        b->n_fused_tiles += (uint32_t)group.size();
    }
    SW_HIP(e, dev_upload(&b->d_items, singles.data(), singles.size(), s));
    SW_HIP(e, dev_upload(&b->d_fused, fused.data(), fused.size(), s));
    if ((rc0 = scratch_release(e, s))) return rc0;
    SW_HIP(e, hipStreamSynchronize(s));        // host tables go out of scope
    SW_HIP(e, hipEventCreate(&b->ev0));
    SW_HIP(e, hipEventCreate(&b->ev1));
    SW_HIP(e, hipEventCreateWithFlags(&b->run_ev, hipEventDisableTiming));
    {
        int v = 0;
        SW_HIP(e, hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, e->device));
        b->lds_cu = (size_t)v;
        SW_HIP(e, hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device));
        b->lds_block = (size_t)v;
    }
    int rc = apply_schedule(e, b.get(), kDefaultSchedule, s, true);
    if (rc) return rc;
    *out = b.release();
    return SW_OK;
}

SW_API int sw_batch_run(sw_batch *b, void *stream, sw_walk_stats *per_item, sw_walk_stats *total) {
    if (!b) return SW_E_INVALID;
    sw_engine *e = b->eng;
    SW_HIP(e, hipSetDevice(b->device));
    hipStream_t s = stream ? (hipStream_t)stream : e->stream;
    if (!b->exec) return fail(e, SW_E_INVALID, "the batch has no graph (its last schedule change failed)");
    const bool stats = per_item || total;
    if (b->run_stream && b->run_stream != s) SW_HIP(e, hipStreamWaitEvent(s, b->run_ev, 0));
    if (stats) SW_HIP(e, hipEventRecord(b->ev0, s));
    SW_HIP(e, hipGraphLaunch(b->exec, s));
    std::vector<unsigned long long> c;
    if (stats) {
        SW_HIP(e, hipEventRecord(b->ev1, s));
        c.resize(4 * (size_t)b->n_items);
        SW_HIP(e, hipMemcpyAsync(c.data(), b->d_counters, c.size() * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost, s));
    }
    SW_HIP(e, hipEventRecord(b->run_ev, s));
    b->run_stream = s;
    if (!stats) return SW_OK;
    SW_HIP(e, hipStreamSynchronize(s));
    float ms = 0;
    SW_HIP(e, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    sw_walk_stats t;
    std::memset(&t, 0, sizeof t);
    for (uint32_t i = 0; i < b->n_items; i++) {
        sw_walk_stats x;
        x.n_lost = c[4 * i]; x.n_replay_exhausted = c[4 * i + 1]; x.n_steps = c[4 * i + 2]; x.n_attempts = c[4 * i + 3];
        x.kernel_ms = ms;
        if (per_item) per_item[i] = x;
        t.n_lost += x.n_lost; t.n_replay_exhausted += x.n_replay_exhausted;
        t.n_steps += x.n_steps; t.n_attempts += x.n_attempts;
    }
    t.kernel_ms = ms;
    if (total) *total = t;
    if (t.n_replay_exhausted)
        return fail(e, SW_E_REPLAY, std::to_string(t.n_replay_exhausted) + " replayed spins ran out of recorded draws");
    return SW_OK;
}

SW_API void sw_batch_destroy(sw_batch *b) { batch_free(b); }

SW_API int sw_batch_info(const sw_batch *b, uint32_t *n_walks, uint32_t *n_fused_tiles) {
    if (!b) return SW_E_INVALID;
    if (n_walks) *n_walks = (uint32_t)b->tiles.size();
    if (n_fused_tiles) *n_fused_tiles = b->n_fused_tiles;
    return SW_OK;
}

SW_API int sw_batch_graph_info(const sw_batch *b, int32_t *uses_graph, uint32_t *n_nodes) {
    if (!b) return SW_E_INVALID;
    if (uses_graph) *uses_graph = b->exec != nullptr;
    if (n_nodes) {
        size_t n = 0;
        if (b->graph && hipGraphGetNodes(b->graph, nullptr, &n) != hipSuccess) n = 0;
        *n_nodes = (uint32_t)n;
    }
    return SW_OK;
}

SW_API int sw_batch_get_schedule(const sw_batch *b, sw_batch_schedule *out) {
    if (!b || !out) return SW_E_INVALID;
    *out = b->sched;
    return SW_OK;
}

SW_API int sw_batch_set_schedule(sw_batch *b, const sw_batch_schedule *sc) {
    if (!b) return SW_E_INVALID;
    sw_engine *e = b->eng;
    if (!sc || !schedule_ok(*sc))
        return fail(e, SW_E_INVALID, "schedule needs chunk_spins 64, 128 or 256, at least one band, spins_per_lane 1 "
                                     "or 2, waves_per_simd 0 or 2-8 and order_brick 4, 8, 16 or 32");
    SW_HIP(e, hipSetDevice(b->device));
    return apply_schedule(e, b, *sc, e->stream, true);
}

// Coordinate search over the schedule: starting from the current one, each
// parameter in turn (chunk size, the band counts of the class kinds the batch
// has, spins per lane, resident waves per SIMD, visiting-order brick, then every
// pair of chunk size and waves) takes the candidate value that times fastest, a
// candidate having to beat the best by 1 %.  Each trial is the faster
// of two full runs of the batch, timed with HIP events;
// schedules change only which spins run where and when, so every trial leaves
// the batch's outputs as any run does.
SW_API int sw_batch_tune(sw_batch *b, void *stream, uint32_t max_trials, sw_batch_schedule *best_out,
                         float *best_ms, float *first_ms) {
    if (!b) return SW_E_INVALID;
    sw_engine *e = b->eng;
    SW_HIP(e, hipSetDevice(b->device));
    hipStream_t s = stream ? (hipStream_t)stream : e->stream;
    uint32_t trials = 0;
    // A trial is the faster of two full runs under the candidate: the first run
    // after a schedule change meets caches that the last schedule left, and
    // the runs of one schedule differ by about 1 %, the margin a candidate must beat.
    auto timed = [&](const sw_batch_schedule &sc, float *ms) -> int {
        int rc = apply_schedule(e, b, sc, s, false);
        if (rc) return rc;
        if (b->run_stream && b->run_stream != s) SW_HIP(e, hipStreamWaitEvent(s, b->run_ev, 0));
        float t_min = 0;
        for (int k = 0; k < 2; k++) {
            SW_HIP(e, hipEventRecord(b->ev0, s));
            SW_HIP(e, batch_enqueue(b, s));
            SW_HIP(e, hipEventRecord(b->ev1, s));
            SW_HIP(e, hipEventSynchronize(b->ev1));
            float t_k = 0;
            SW_HIP(e, hipEventElapsedTime(&t_k, b->ev0, b->ev1));
            t_min = k == 0 ? t_k : std::min(t_min, t_k);
        }
        SW_HIP(e, hipEventRecord(b->run_ev, s));
        b->run_stream = s;
        *ms = t_min;
        trials++;
        return SW_OK;
    };
    sw_batch_schedule best = b->sched;
    float t_best = 0, t = 0;
    int rc = timed(best, &t_best);                 // warm-up: caches, code objects
    if (rc) return rc;
    rc = timed(best, &t_best);
    if (rc) return rc;
    const float t_first = t_best;
    const unsigned kinds = schedule_kinds(*b);
    auto consider = [&](sw_batch_schedule cand) -> int {
        if (max_trials && trials >= max_trials) return SW_OK;
        if (std::memcmp(&cand, &best, sizeof cand) == 0) return SW_OK;
        if (batch_lds(*b, cand) > b->lds_block) return SW_OK;     // a cap this device cannot express
        int r = timed(cand, &t);
        if (r) return r;
        if (t < 0.99f * t_best) { best = cand; t_best = t; }
        return SW_OK;
    };
    auto chunks = [&]() -> int {
        for (uint32_t c : {256u, 128u, 64u}) {
            sw_batch_schedule x = best; x.chunk_spins = c;
            if (int r = consider(x)) return r;
        }
        return SW_OK;
    };
    auto caps = [&]() -> int {
        for (uint32_t v : {5u, 4u, 3u, 0u}) {
            sw_batch_schedule x = best; x.waves_per_simd = v;
            if (int r = consider(x)) return r;
        }
        return SW_OK;
    };
    if ((rc = chunks())) return rc;
    if (kinds & 1u)
        for (uint32_t v : {1u << 20, 64u, 16u, 4u}) {
            sw_batch_schedule x = best; x.bands_shared = v;
            if ((rc = consider(x))) return rc;
        }
    if (kinds & 2u)
        for (uint32_t v : {16u, 64u, 4u, 1u}) {
            sw_batch_schedule x = best; x.bands_mixed = v;
            if ((rc = consider(x))) return rc;
        }
    for (uint32_t v : {2u, 1u}) {
        sw_batch_schedule x = best; x.spins_per_lane = v;
        if ((rc = consider(x))) return rc;
    }
    if ((rc = caps())) return rc;
    for (uint32_t v : {16u, 4u, 32u, 8u}) {
        sw_batch_schedule x = best; x.order_brick = v;
        if ((rc = consider(x))) return rc;
    }
    // The chunk size and the cap interact (a cap counts workgroups of the
    // chunk's size), so a pass over one and then the other can stop short of
    // the best pair: every pair is tried from the point reached.
    {
        const sw_batch_schedule from = best;
        for (uint32_t c : {256u, 128u, 64u})
            for (uint32_t v : {5u, 4u, 3u, 0u}) {
                sw_batch_schedule x = from; x.chunk_spins = c; x.waves_per_simd = v;
                if ((rc = consider(x))) return rc;
            }
    }
    rc = apply_schedule(e, b, best, s, true);
    if (rc) return rc;
    if (best_out) *best_out = best;
    if (best_ms) *best_ms = t_best;
    if (first_ms) *first_ms = t_first;
    return SW_OK;
}

SW_API int sw_walk_host(sw_engine *e, const sw_phantom *ph, const sw_sequence *q, const sw_params *p,
                        const sw_walk_opts *o, const float *M0, const float *XYZ0, void *M1, float *XYZ1,
                        uint8_t *T, const float *rn, const float *ru, sw_walk_stats *stats) {
    if (!e || !o || !p || !ph || !q) return fail(e, SW_E_INVALID, "null argument");
    // checked here as well as by sw_walk_device: the buffer sizes below depend on them
    if (p->n_dummy_scan < 0) return fail(e, SW_E_INVALID, "n_dummy_scan < 0: call sw_params_prepare first");
    if (o->n_local && (!M0 || !XYZ0 || !XYZ1 || (q->n_te && (!M1 || !T))))
        return fail(e, SW_E_INVALID, "spin buffers are required");
    if (o->rng == SW_RNG_REPLAY && o->n_local && (!rn || !ru))
        return fail(e, SW_E_INVALID, "replay mode needs normal and uniform streams");
    SW_HIP(e, hipSetDevice(e->device));
    const size_t n = o->n_local, nte = q->n_te;
    const size_t traj = p->enRecordTrajectory ? (size_t)(p->n_dummy_scan + 1) * p->n_timepoints : 1;
    const bool replay = o->rng == SW_RNG_REPLAY;
    // bytes of each device copy, in the order of sw_engine::d_host
    const size_t bytes[7] = {12 * n, 12 * n, (o->precision == SW_PREC_F64 ? 8 : 4) * 3 * n * nte, 12 * n * traj,
                             n * nte, replay ? 4 * n * o->replay_normal_stride : 0,
                             replay ? 4 * n * o->replay_uniform_stride : 0};
    hipStream_t s = e->stream;
    // The device copies persist between calls and only grow: a call reuses the
    // previous call's buffers once the stream has finished with them.
    size_t grow = 0;
    for (int i = 0; i < 7; i++) if (bytes[i] > e->host_cap[i]) grow += bytes[i];
    if (grow) {
        SW_HIP(e, hipStreamSynchronize(s));
        for (int i = 0; i < 7; i++)
            if (bytes[i] > e->host_cap[i]) {
                dev_free(e->d_host[i]);
                e->d_host[i] = nullptr;
                e->host_cap[i] = 0;
            }
        int rc = check_free_memory(e, grow);
        if (rc) return rc;
        for (int i = 0; i < 7; i++)
            if (bytes[i] > e->host_cap[i]) {
                SW_HIP(e, hipMalloc(&e->d_host[i], bytes[i]));
                e->host_cap[i] = bytes[i];
            }
    }
    float *dM0 = (float *)e->d_host[0], *dX0 = (float *)e->d_host[1], *dX1 = (float *)e->d_host[3];
    void *dM1 = e->d_host[2];
    uint8_t *dT = (uint8_t *)e->d_host[4];
    const void *src[7] = {M0, XYZ0, M1, XYZ1, T, rn, ru};
    for (int i = 0; i < 7; i++)
        if (bytes[i]) SW_HIP(e, hipMemcpyAsync(e->d_host[i], src[i], bytes[i], hipMemcpyHostToDevice, s));
    sw_walk_opts oo = *o;
    if (replay) {
        oo.replay_normals = (const float *)e->d_host[5];
        oo.replay_uniforms = (const float *)e->d_host[6];
    }
    sw_walk_stats st;
    std::memset(&st, 0, sizeof st);
    const int rc = sw_walk_device(e, ph, q, p, &oo, dM0, dX0, dM1, dX1, dT, s, &st);
    if (rc == SW_OK || rc == SW_E_REPLAY) {
        if (bytes[2]) SW_HIP(e, hipMemcpyAsync(M1, dM1, bytes[2], hipMemcpyDeviceToHost, s));
        if (bytes[4]) SW_HIP(e, hipMemcpyAsync(T, dT, bytes[4], hipMemcpyDeviceToHost, s));
        if (bytes[3]) SW_HIP(e, hipMemcpyAsync(XYZ1, dX1, bytes[3], hipMemcpyDeviceToHost, s));
    }
    SW_HIP(e, hipStreamSynchronize(s));
    if (stats) *stats = st;
    return rc;
}

SW_API int sw_echo_reduce_device(sw_engine *e, const void *d_M1, int m1_dtype, const uint8_t *d_T, uint32_t n_spins,
                                 uint32_t n_te, uint32_t n_sub, double *d_out, void *stream) {
    if (!e || !d_out || (n_spins && (!d_M1 || !d_T))) return fail(e, SW_E_INVALID, "null argument");
    hipStream_t s = stream ? (hipStream_t)stream : e->stream;
    SW_HIP(e, hipSetDevice(e->device));
    int rc = scratch_acquire(e, s);               // d_partial is engine-owned
    if (rc) return rc;
    const size_t need = sw::reduce_partial_len(n_spins, n_te, n_sub);
    if (need > e->partial_cap) {
        SW_HIP(e, hipStreamSynchronize(s));
        dev_free(e->d_partial);
        e->d_partial = nullptr;
        SW_HIP(e, hipMalloc((void **)&e->d_partial, need * sizeof(double)));
        e->partial_cap = need;
    }
    SW_HIP(e, sw::launch_echo_reduce(d_M1, m1_dtype == SW_DTYPE_F64, d_T, n_spins, n_te, n_sub, e->d_partial, d_out,
                                     s));
    return scratch_release(e, s);
}

// diagnostic: the production draws (not part of the reference interface)
SW_API int sw_philox_draws_device(sw_engine *e, uint64_t seed, uint32_t spin0, uint32_t n_spins, uint32_t n_att,
                                  float *d_out, void *stream) {
    if (!e || !d_out) return fail(e, SW_E_INVALID, "null argument");
    hipStream_t s = stream ? (hipStream_t)stream : e->stream;
    SW_HIP(e, hipSetDevice(e->device));
    SW_HIP(e, sw::launch_philox_draws(seed, spin0, n_spins, n_att, d_out, s));
    return SW_OK;
}

SW_API int sw_probe_gather_device(sw_engine *e, const void *d_buf, uint64_t bytes, uint32_t n_lanes,
                                  uint32_t n_reads, uint64_t seed, uint64_t *d_sink, void *stream, float *ms) {
    if (!e || !d_buf || !d_sink || !ms) return fail(e, SW_E_INVALID, "null argument");
    if (bytes < 128 || n_lanes == 0 || n_reads == 0 || n_reads % 8)
        return fail(e, SW_E_INVALID, "probe needs bytes >= 128, lanes > 0 and reads a positive multiple of 8");
    hipStream_t s = stream ? (hipStream_t)stream : e->stream;
    SW_HIP(e, hipSetDevice(e->device));
    SW_HIP(e, hipEventRecord(e->ev0, s));
    SW_HIP(e, sw::launch_gather_probe(d_buf, bytes, n_lanes, n_reads, seed, d_sink, s));
    SW_HIP(e, hipEventRecord(e->ev1, s));
    SW_HIP(e, hipEventSynchronize(e->ev1));
    SW_HIP(e, hipEventElapsedTime(ms, e->ev0, e->ev1));
    return SW_OK;
}

SW_API int sw_stream_sync(sw_engine *e, void *stream) {
    if (!e) return SW_E_INVALID;
    SW_HIP(e, hipStreamSynchronize(stream ? (hipStream_t)stream : e->stream));
    return SW_OK;
}

}  // extern "C"
