// driver.cpp -- monte_carlo::run for the MI355X engine, and the host C ABI of
// include/spinwalk_host.h.
//
// Mirrors src/sim/monte_carlo.cu:199-355: per config, read + prepare the
// parameters; per phantom, read the HDF5 phantom (fieldmap optional, mask, fov),
// initialise positions (file or seeded uniform within 98 % of the FoV) and
// magnetisation ((0,0,1), G1), upload the phantom and the output buffers (the
// output buffers keep the previous phantom's values, so spins abandoned before an
// echo keep stale entries exactly as in the reference), run every scale as one
// walk launch into its output slice, then write M, XYZ, T, scales and TE.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/spinwalk_host.h"
#include "config.hpp"
#include "h5io.hpp"
#include "seqgen.hpp"

namespace fs = std::filesystem;

namespace {

thread_local std::string g_err;

void info(const std::string &m) { std::fprintf(stderr, "[spinwalk] %s\n", m.c_str()); }

int fail(int code, const std::string &m) {
    g_err = m;
    std::fprintf(stderr, "[spinwalk] error: %s\n", m.c_str());
    return code;
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t n) { return n ? hipMalloc(&p, n) : hipSuccess; }
};

// Host copy of a trajectory chunk (hundreds of MB) split over host threads: one
// thread copies a few GB/s, well below what the PCIe transfers it sits between
// move.
void par_memcpy(void *dst, const void *src, size_t bytes) {
    constexpr size_t kPiece = 32u << 20;
    const size_t n_pieces = (bytes + kPiece - 1) / kPiece;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t n_threads = std::min<size_t>({n_pieces, (size_t)hw, 16});
    if (n_threads <= 1) { std::memcpy(dst, src, bytes); return; }
    std::vector<std::thread> th;
    th.reserve(n_threads);
    for (size_t t = 0; t < n_threads; t++)
        th.emplace_back([=] {
            for (size_t i = t; i < n_pieces; i += n_threads) {
                const size_t off = i * kPiece, len = std::min(kPiece, bytes - off);
                std::memcpy((char *)dst + off, (const char *)src + off, len);
            }
        });
    for (auto &x : th) x.join();
}

struct RunOpts {
    int32_t rng = SW_RNG_PHILOX, precision = SW_PREC_F32, field_format = SW_FIELD_F32, spatial_order = 1;
    uint32_t traj_chunk_spins = 0;
};

// Pinned host memory (hipHostMalloc), freed on scope exit.
struct PinnedBuf {
    void *p = nullptr;
    ~PinnedBuf() { if (p) (void)hipHostFree(p); }
    hipError_t alloc(size_t n) { return n ? hipHostMalloc(&p, n, hipHostMallocDefault) : hipSuccess; }
};

struct StreamPair {
    hipStream_t walk = nullptr, copy = nullptr;
    hipEvent_t walked[2] = {nullptr, nullptr}, copied[2] = {nullptr, nullptr};
    ~StreamPair() {
        for (hipEvent_t e : {walked[0], walked[1], copied[0], copied[1]}) if (e) (void)hipEventDestroy(e);
        if (walk) (void)hipStreamDestroy(walk);
        if (copy) (void)hipStreamDestroy(copy);
    }
    hipError_t create() {
        hipError_t e = hipStreamCreateWithFlags(&walk, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&copy, hipStreamNonBlocking);
        for (int i = 0; i < 2 && e == hipSuccess; i++) {
            e = hipEventCreateWithFlags(&walked[i], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&copied[i], hipEventDisableTiming);
        }
        return e;
    }
};

int run_phantom(sw_engine *eng, swhost::Config &c, size_t iph, std::vector<float> &M1, std::vector<float> &XYZ1,
                std::vector<uint8_t> &T, const RunOpts &ro) {
    const std::string file = c.files.at("PHANTOM")[iph];
    info("Simulating phantom: " + file);
    std::string err;
    std::vector<unsigned char> mbytes, fbytes, vbytes;
    std::vector<size_t> mdims, fdims, vdims;
    const bool has_field = swhost::h5_exists(file, "fieldmap");                // monte_carlo.cu:101-102
    if (has_field && !swhost::h5_read(file, "fieldmap", swhost::H5T::F32, true, 0, fbytes, fdims, err))
        return fail(SW_E_IO, err);
    if (!swhost::h5_read(file, "mask", swhost::H5T::U8, true, 0, mbytes, mdims, err)) return fail(SW_E_IO, err);
    if (!swhost::h5_read(file, "fov", swhost::H5T::F32, false, 3, vbytes, vdims, err)) return fail(SW_E_IO, err);
    if (mdims.size() != 3) return fail(SW_E_IO, "mask must be a 3-D dataset");
    if (has_field && fbytes.size() != mbytes.size() * 4) return fail(SW_E_IO, "fieldmap and mask differ in size");
    float fov[3];
    std::memcpy(fov, vbytes.data(), sizeof fov);
    const uint32_t n_sub_mask = *std::max_element(mbytes.begin(), mbytes.end()) + 1u;   // :113-118
    if (n_sub_mask > c.p.n_substrate)
        return fail(SW_E_CONFIG, "The number of substrate types in the mask does not match the number of substrate "
                                 "types in the config file: " + std::to_string(n_sub_mask) + " vs " +
                                 std::to_string(c.p.n_substrate));
    const uint32_t n = c.p.n_spins;
    // initial positions (monte_carlo.cu:125-153)
    std::vector<float> XYZ0(3 * (size_t)n);
    const std::string xf = c.files.at("XYZ0")[iph];
    if (!xf.empty()) {
        if (!swhost::h5_read(xf, "XYZ", swhost::H5T::F32, false, 3 * (size_t)n, vbytes, vdims, err))
            return fail(SW_E_IO, err);
        std::memcpy(XYZ0.data(), vbytes.data(), XYZ0.size() * 4);
        for (size_t i = 0; i < XYZ0.size(); i++)
            if (XYZ0[i] < 0 || XYZ0[i] > fov[i % 3]) return fail(SW_E_INVALID, "Initial positions are outside the FoV.");
    } else {
        sw_init_positions(c.p.seed, fov, n, XYZ0.data());
    }
    // magnetisation: a given M0 file is read, then (0,0,1) is used (monte_carlo.cu:155-166)
    const std::string mf = c.files.at("M0")[iph];
    if (!mf.empty() && !swhost::h5_read(mf, "M", swhost::H5T::F32, false, 3 * (size_t)n, vbytes, vdims, err))
        return fail(SW_E_IO, err);
    std::vector<float> M0(3 * (size_t)n, 0.f);
    for (size_t i = 0; i < n; i++) M0[3 * i + 2] = 1.f;

    const uint64_t dims[3] = {mdims[0], mdims[1], mdims[2]};
    uint32_t formats = 1u << SW_FIELD_F32;
    if (ro.field_format > SW_FIELD_F32 && ro.field_format <= SW_FIELD_Q16) formats |= 1u << ro.field_format;
    sw_phantom *ph = nullptr;
    int rc = sw_phantom_create(eng, (const uint8_t *)mbytes.data(), has_field ? (const float *)fbytes.data() : nullptr,
                               dims, fov, formats, &ph);
    if (rc) return fail(rc, std::string("phantom upload: ") + sw_last_error(eng));
    std::unique_ptr<sw_phantom, void (*)(sw_phantom *)> ph_guard(ph, sw_phantom_destroy);
    c.p.fieldmap_exist = has_field;
    c.p.matrix_length = (int64_t)(dims[0] * dims[1] * dims[2]);
    for (int i = 0; i < 3; i++) c.p.phantom_size[i] = dims[i];

    const size_t n_te = c.TE.size(), n_sc = c.scales.size();
    const size_t traj = c.p.enRecordTrajectory ? (size_t)(c.p.n_dummy_scan + 1) * c.p.n_timepoints : 1;
    const size_t m1_slice = 3 * n_te * (size_t)n, t_slice = n_te * (size_t)n, x_slice = 3 * (size_t)n * traj;
    // Recorded trajectories stream: spins are walked in chunks into two device
    // buffers in turn, and each chunk's trajectory goes back through pinned memory
    // on a second stream while the next chunk walks.  Final positions alone
    // (traj == 1) stay on the device for every scale.
    const bool streamed = c.p.enRecordTrajectory != 0;
    uint32_t chunk = n;
    if (streamed) {
        const uint64_t per_spin = 12ull * traj;
        const uint64_t cap = ro.traj_chunk_spins ? ro.traj_chunk_spins : std::max<uint64_t>(1, (512ull << 20) / per_spin);
        chunk = (uint32_t)std::min<uint64_t>(n, cap);
    }
    const size_t xc = 3 * (size_t)chunk * traj;                       // floats per chunk buffer
    DevBuf dM0, dX0, dM1, dX1, dT, dRN, dRU, dXa, dXb;
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    const size_t need = 24 * (size_t)n + 4 * m1_slice * n_sc + t_slice * n_sc +
                        (streamed ? 2 * 4 * xc : 4 * x_slice * n_sc);
    if (need > fr) return fail(SW_E_NOMEM, "not enough device memory for the spin buffers");
    const bool x_failed = streamed ? (dXa.alloc(4 * xc) != hipSuccess || dXb.alloc(4 * xc) != hipSuccess)
                                   : dX1.alloc(XYZ1.size() * 4) != hipSuccess;
    if (dM0.alloc(M0.size() * 4) || dX0.alloc(XYZ0.size() * 4) || dM1.alloc(M1.size() * 4) || dT.alloc(T.size()) ||
        x_failed)
        return fail(SW_E_HIP, "hipMalloc failed");
    // the device outputs start from the host buffers (the reference's d_M1 = M1)
    if (hipMemcpy(dM0.p, M0.data(), M0.size() * 4, hipMemcpyHostToDevice) ||
        hipMemcpy(dX0.p, XYZ0.data(), XYZ0.size() * 4, hipMemcpyHostToDevice) ||
        (M1.size() && hipMemcpy(dM1.p, M1.data(), M1.size() * 4, hipMemcpyHostToDevice)) ||
        (!streamed && hipMemcpy(dX1.p, XYZ1.data(), XYZ1.size() * 4, hipMemcpyHostToDevice)) ||
        (T.size() && hipMemcpy(dT.p, T.data(), T.size(), hipMemcpyHostToDevice)))
        return fail(SW_E_HIP, "hipMemcpy failed");
    PinnedBuf pin[2];
    StreamPair sp;
    if (streamed && (pin[0].alloc(4 * xc) || pin[1].alloc(4 * xc) || sp.create()))
        return fail(SW_E_HIP, "trajectory staging buffers");

    sw_walk_opts o{};
    o.rng = ro.rng;
    o.precision = ro.precision == SW_PREC_F64 ? SW_PREC_F32 : ro.precision;   // M is written as float32
    o.field_format = has_field ? ro.field_format : SW_FIELD_NONE;
    o.spatial_order = ro.spatial_order;
    uint64_t rn_stride = 0, ru_stride = 0;
    if (ro.rng == SW_RNG_REPLAY) {
        // the reference CPU path's own draws, enough for every attempt but a
        // pathological number of rejections
        const uint64_t steps = (uint64_t)(c.p.n_dummy_scan + 1) * c.p.n_timepoints;
        rn_stride = 9 * steps + 96;
        ru_stride = 3 * steps + 64;
        std::vector<float> rn(rn_stride * n), ru(ru_stride * n);
        sw_replay_fill(c.p.seed, 0, n, rn.data(), rn_stride, ru.data(), ru_stride, 16);
        if (dRN.alloc(rn.size() * 4) || dRU.alloc(ru.size() * 4) ||
            hipMemcpy(dRN.p, rn.data(), rn.size() * 4, hipMemcpyHostToDevice) ||
            hipMemcpy(dRU.p, ru.data(), ru.size() * 4, hipMemcpyHostToDevice))
            return fail(SW_E_HIP, "replay upload failed");
        o.replay_normal_stride = rn_stride;
        o.replay_uniform_stride = ru_stride;
    }
    sw_sequence *seq = nullptr;
    std::unique_ptr<sw_sequence *, void (*)(sw_sequence **)> seq_guard(&seq, [](sw_sequence **q) {
        if (*q) sw_sequence_destroy(*q);
    });
    const sw_seq q = c.seq_view();
    auto t0 = std::chrono::steady_clock::now();
    uint64_t lost = 0, steps = 0;
    // one trajectory chunk in flight per buffer: (scale, first spin, spins)
    struct Pending { bool live = false; size_t is = 0; uint32_t s0 = 0, cnt = 0; } pend[2];
    auto land = [&](int b) -> int {                  // chunk in buffer b -> host XYZ1
        if (!pend[b].live) return SW_OK;
        if (hipEventSynchronize(sp.copied[b]) != hipSuccess) return fail(SW_E_HIP, "trajectory download");
        par_memcpy(XYZ1.data() + x_slice * pend[b].is + 3 * (size_t)pend[b].s0 * traj, pin[b].p,
                   12 * (size_t)pend[b].cnt * traj);
        pend[b].live = false;
        return SW_OK;
    };
    // Without recorded trajectories every scale of the phantom is one tile of a
    // single batched launch (the same results as one walk per scale).
    const bool batched = !streamed && n_sc > 1 && n > 0;
    std::vector<sw_batch_item> tiles;
    size_t k = 0;                                                             // chunk counter
    for (size_t is = 0; is < n_sc; is++) {                                     // :273-337
        const float scale = c.scales[is];
        sw_params p = c.p;
        o.xyz_scale = 1.f;
        o.gradient_scale = 1.f;
        for (int i = 0; i < 3; i++) p.fov[i] = fov[i];
        if (c.scale_type == 0) {                                               // FoV (:277-285)
            for (int i = 0; i < 3; i++) p.fov[i] = scale * fov[i];
            o.xyz_scale = scale;
        } else if (c.scale_type == 1) {                                        // gradients (:287-301)
            o.gradient_scale = scale;
        } else {                                                               // phase cycling (:302-305)
            p.linear_phase_cycling = c.linear_phase_cycling_orig * scale;
            p.quadratic_phase_cycling = c.quadratic_phase_cycling_orig;
        }
        if (!seq) {
            rc = sw_sequence_create(eng, &p, &q, &seq);
            if (rc) return fail(rc, std::string("sequence upload: ") + sw_last_error(eng));
        }
        for (uint32_t s0 = 0; s0 < n || (n == 0 && s0 == 0); s0 += chunk, k++) {
            const uint32_t cnt = n ? std::min<uint32_t>(chunk, n - s0) : 0;
            o.spin_offset = s0;
            o.n_local = cnt;
            if (ro.rng == SW_RNG_REPLAY) {
                o.replay_normals = (const float *)dRN.p + rn_stride * s0;
                o.replay_uniforms = (const float *)dRU.p + ru_stride * s0;
            }
            const float *m0 = (const float *)dM0.p + 3 * (size_t)s0, *x0 = (const float *)dX0.p + 3 * (size_t)s0;
            float *m1 = (float *)dM1.p + m1_slice * is + 3 * n_te * (size_t)s0;
            uint8_t *tt = (uint8_t *)dT.p + t_slice * is + n_te * (size_t)s0;
            sw_walk_stats st{};
            if (batched) {
                sw_batch_item it;
                std::memset(&it, 0, sizeof it);
                it.phantom = ph; it.sequence = seq; it.params = p; it.opts = o;
                it.d_M0 = m0; it.d_XYZ0 = x0; it.d_M1 = m1; it.d_T = tt;
                it.d_XYZ1 = (float *)dX1.p + x_slice * is + 3 * (size_t)s0;
                tiles.push_back(it);
                break;
            }
            if (!streamed) {
                rc = sw_walk_device(eng, ph, seq, &p, &o, m0, x0, m1, (float *)dX1.p + x_slice * is + 3 * (size_t)s0,
                                    tt, nullptr, &st);
            } else {
                const int b = (int)(k & 1);
                rc = land(b);                                   // the chunk that used this buffer before
                if (rc) return rc;
                // stale host contents first (abandoned spins keep them), then walk
                const size_t bytes = 12 * (size_t)cnt * traj;
                par_memcpy(pin[b].p, XYZ1.data() + x_slice * is + 3 * (size_t)s0 * traj, bytes);
                float *dx = (float *)(b ? dXb.p : dXa.p);
                if (hipMemcpyAsync(dx, pin[b].p, bytes, hipMemcpyHostToDevice, sp.walk) != hipSuccess)
                    return fail(SW_E_HIP, "trajectory upload");
                rc = sw_walk_device(eng, ph, seq, &p, &o, m0, x0, m1, dx, tt, sp.walk, &st);
                if (rc == SW_OK) {
                    if (hipEventRecord(sp.walked[b], sp.walk) != hipSuccess ||
                        hipStreamWaitEvent(sp.copy, sp.walked[b], 0) != hipSuccess ||
                        hipMemcpyAsync(pin[b].p, dx, bytes, hipMemcpyDeviceToHost, sp.copy) != hipSuccess ||
                        hipEventRecord(sp.copied[b], sp.copy) != hipSuccess)
                        return fail(SW_E_HIP, "trajectory download");
                    pend[b].live = true; pend[b].is = is; pend[b].s0 = s0; pend[b].cnt = cnt;
                }
            }
            if (rc) return fail(rc, std::string("walk: ") + sw_last_error(eng));
            lost += st.n_lost;
            steps += st.n_steps;
            if (n == 0) break;
        }
        if (!batched)
            info("scale " + std::to_string(scale) + ": " + std::to_string(steps) + " spin-steps so far, " +
                 std::to_string(lost) + " spins lost so far");
    }
    if (batched) {
        sw_batch *bt = nullptr;
        rc = sw_batch_create(eng, tiles.data(), (uint32_t)tiles.size(), &bt);
        if (rc) return fail(rc, std::string("batch: ") + sw_last_error(eng));
        sw_walk_stats tot{};
        rc = sw_batch_run(bt, nullptr, nullptr, &tot);
        sw_batch_destroy(bt);
        if (rc) return fail(rc, std::string("walk: ") + sw_last_error(eng));
        lost += tot.n_lost;
        steps += tot.n_steps;
        info(std::to_string(n_sc) + " scales in one launch: " + std::to_string(steps) + " spin-steps in " +
             std::to_string(tot.kernel_ms) + " ms, " + std::to_string(lost) + " spins lost");
    }
    for (int b = 0; b < 2; b++) {
        rc = land(b);
        if (rc) return rc;
    }
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    info("Simulation took " + std::to_string(el) + " seconds (" + std::to_string(steps / std::max(el, 1e-9)) +
         " spin-steps/s).");
    if ((M1.size() && hipMemcpy(M1.data(), dM1.p, M1.size() * 4, hipMemcpyDeviceToHost)) ||
        (!streamed && hipMemcpy(XYZ1.data(), dX1.p, XYZ1.size() * 4, hipMemcpyDeviceToHost)) ||
        (T.size() && hipMemcpy(T.data(), dT.p, T.size(), hipMemcpyDeviceToHost)))
        return fail(SW_E_HIP, "result download failed");
    (void)lost;
    // save (monte_carlo.cu:168-197)
    const std::string out = c.outputs[iph];
    info("Saving the results to " + out);
    std::error_code ec;
    fs::remove(out, ec);
    const size_t ns = n_sc;
    if (!swhost::h5_write(out, "M", swhost::H5T::F32, {ns, (size_t)n, n_te, 3}, M1.data(), err) ||
        !swhost::h5_write(out, "XYZ", swhost::H5T::F32, {ns, (size_t)n, traj, 3}, XYZ1.data(), err) ||
        !swhost::h5_write(out, "T", swhost::H5T::U8, {ns, (size_t)n, n_te, 1}, T.data(), err) ||
        !swhost::h5_write(out, "scales", swhost::H5T::F32, {ns, 1, 1, 1}, c.scales.data(), err))
        return fail(SW_E_IO, err);
    std::vector<float> TE_s;
    for (int32_t t : c.TE) TE_s.push_back((float)(t * c.p.timestep_us * 1e-6));
    if (!swhost::h5_write(out, "TE", swhost::H5T::F32, {TE_s.size(), 1, 1, 1}, TE_s.data(), err))
        return fail(SW_E_IO, err);
    return SW_OK;
}

int run_config(sw_engine *eng, const char *path, const RunOpts &ro) {
    swhost::Config c;
    std::string err;
    if (!swhost::load_config(path, c, err)) return fail(SW_E_CONFIG, err);
    const size_t n = c.p.n_spins, n_te = c.TE.size(), n_sc = c.scales.size();
    const size_t traj = c.p.enRecordTrajectory ? (size_t)(c.p.n_dummy_scan + 1) * c.p.n_timepoints : 1;
    // allocate_memory (monte_carlo.cu:61-70); these host buffers carry over between phantoms
    std::vector<float> M1(3 * n_te * n * n_sc, 0.f), XYZ1(3 * n * traj * n_sc, 0.f);
    std::vector<uint8_t> T(n_te * n * n_sc, 0);
    for (size_t i = 0; i < c.files.at("PHANTOM").size(); i++) {
        int rc = run_phantom(eng, c, i, M1, XYZ1, T, ro);
        if (rc) return rc;
    }
    return SW_OK;
}

}  // namespace

struct sw_config {
    swhost::Config c;
};

extern "C" {

SW_API const char *sw_host_last_error(void) { return g_err.c_str(); }

SW_API int sw_config_load(const char *path, sw_config **out, char **err) {
    if (!path || !out) return SW_E_INVALID;
    *out = nullptr;
    std::unique_ptr<sw_config> c(new sw_config);
    std::string e;
    if (!swhost::load_config(path, c->c, e)) {
        g_err = e;
        if (err) *err = strdup(e.c_str());
        return SW_E_CONFIG;
    }
    *out = c.release();
    return SW_OK;
}

SW_API void sw_config_destroy(sw_config *c) { delete c; }
SW_API void sw_config_free_string(char *s) { free(s); }

SW_API int sw_config_params(const sw_config *c, sw_params *p, sw_seq *q) {
    if (!c) return SW_E_INVALID;
    if (p) *p = c->c.p;
    if (q) *q = c->c.seq_view();
    return SW_OK;
}

SW_API int sw_config_scales(const sw_config *c, const float **scales, uint32_t *n, int32_t *what) {
    if (!c) return SW_E_INVALID;
    if (scales) *scales = c->c.scales.data();
    if (n) *n = (uint32_t)c->c.scales.size();
    if (what) *what = c->c.scale_type;
    return SW_OK;
}

SW_API uint32_t sw_config_n_files(const sw_config *c, const char *key) {
    if (!c || !key || !c->c.files.count(key)) return 0;
    return (uint32_t)c->c.files.at(key).size();
}

SW_API const char *sw_config_file(const sw_config *c, const char *key, uint32_t i) {
    if (!c || !key || !c->c.files.count(key) || i >= c->c.files.at(key).size()) return nullptr;
    return c->c.files.at(key)[i].c_str();
}

SW_API const char *sw_config_output(const sw_config *c, uint32_t i) {
    if (!c || i >= c->c.outputs.size()) return nullptr;
    return c->c.outputs[i].c_str();
}

SW_API const char *sw_config_seq_name(const sw_config *c) { return c ? c->c.seq_name.c_str() : nullptr; }

SW_API int sw_run_config(sw_engine *eng, const char *path) {
    if (!eng || !path) return SW_E_INVALID;
    return run_config(eng, path, RunOpts{});
}

SW_API int sw_run_config_ex(sw_engine *eng, const char *path, int32_t rng, int32_t precision, int32_t field_format,
                            int32_t spatial_order) {
    if (!eng || !path) return SW_E_INVALID;
    RunOpts ro;
    ro.rng = rng;
    ro.precision = precision;
    ro.field_format = field_format;
    ro.spatial_order = spatial_order;
    return run_config(eng, path, ro);
}

SW_API int sw_run_config_opts(sw_engine *eng, const char *path, const sw_run_options *opt) {
    if (!eng || !path || !opt) return SW_E_INVALID;
    RunOpts ro;
    ro.rng = opt->rng;
    ro.precision = opt->precision;
    ro.field_format = opt->field_format;
    ro.spatial_order = opt->spatial_order;
    ro.traj_chunk_spins = opt->trajectory_chunk_spins;
    return run_config(eng, path, ro);
}

SW_API int sw_sim_execute(int use_cpu, uint32_t device_id, const char *const *files, uint32_t n_files) {
    if (use_cpu) return fail(SW_E_INVALID, "this engine has no CPU path; run without -p");
    sw_engine *eng = nullptr;
    int rc = sw_engine_create((int)device_id, &eng);
    if (rc) return fail(rc, std::string("device ") + std::to_string(device_id) + ": " + sw_last_error(nullptr));
    for (uint32_t i = 0; i < n_files && rc == SW_OK; i++) {
        std::fprintf(stderr, "<%s>\n", fs::path(files[i]).filename().string().c_str());
        rc = sw_run_config(eng, files[i]);
    }
    sw_engine_destroy(eng);
    if (rc == SW_OK) std::fprintf(stderr, "Simulation completed successfully.\n");
    return rc;
}

// HDF5 helpers for callers that have no HDF5 binding of their own.
SW_API int sw_h5_read(const char *file, const char *name, int32_t dtype, void *buf, uint64_t cap_elems,
                      uint64_t *dims, uint32_t *ndims) {
    std::vector<unsigned char> bytes;
    std::vector<size_t> d;
    std::string err;
    const swhost::H5T t = dtype == 0 ? swhost::H5T::F32 : (dtype == 1 ? swhost::H5T::U8 : swhost::H5T::F64);
    if (!swhost::h5_read(file, name, t, true, 0, bytes, d, err)) return fail(SW_E_IO, err);
    const size_t es = dtype == 0 ? 4 : (dtype == 1 ? 1 : 8);
    if (ndims) *ndims = (uint32_t)d.size();
    if (dims) for (size_t i = 0; i < d.size() && i < 8; i++) dims[i] = d[i];
    if (buf) {
        if (bytes.size() / es > cap_elems) return fail(SW_E_INVALID, "buffer too small");
        std::memcpy(buf, bytes.data(), bytes.size());
    }
    return SW_OK;
}

SW_API int sw_h5_write(const char *file, const char *name, int32_t dtype, const void *buf, const uint64_t *dims,
                       uint32_t ndims) {
    std::string err;
    const swhost::H5T t = dtype == 0 ? swhost::H5T::F32 : (dtype == 1 ? swhost::H5T::U8 : swhost::H5T::F64);
    std::vector<size_t> d(dims, dims + ndims);
    if (!swhost::h5_write(file, name, t, d, buf, err)) return fail(SW_E_IO, err);
    return SW_OK;
}

// Sequence generators of the reference's `config` and `dwi` subcommands.
SW_API int sw_generate_config(const char *seq_name, uint32_t TE_us, uint32_t timestep_us,
                              const char *const *phantoms, uint32_t n_phantoms, const char *output) {
    if (!seq_name || !output || (n_phantoms && !phantoms)) return fail(SW_E_INVALID, "null argument");
    if (TE_us == 0 || timestep_us == 0) return fail(SW_E_INVALID, "TE and the time step must be positive");
    std::vector<std::string> ph;
    for (uint32_t i = 0; i < n_phantoms; i++) ph.emplace_back(phantoms[i] ? phantoms[i] : "");
    std::string err;
    if (!swhost::generate_config(seq_name, TE_us, timestep_us, ph, output, err)) return fail(SW_E_CONFIG, err);
    return SW_OK;
}

SW_API int sw_generate_pgse(const char *config_file, const double *b_values, uint32_t n_b, const float dir[3],
                            uint32_t start_ms, uint32_t delta_ms, uint32_t DELTA_ms) {
    if (!config_file || !b_values || n_b == 0 || !dir) return fail(SW_E_INVALID, "null or empty argument");
    std::string err;
    if (!swhost::generate_pgse(config_file, std::vector<double>(b_values, b_values + n_b),
                               std::vector<float>(dir, dir + 3), start_ms, delta_ms, DELTA_ms, err))
        return fail(SW_E_CONFIG, err);
    return SW_OK;
}

}  // extern "C"
