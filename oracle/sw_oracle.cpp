/*
 * sw_oracle.cpp -- CPU oracle for the SpinWalk `sim` hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see sw_oracle.h): the checker for the HIP engine and
 * the timed CPU baseline, never part of the product.
 *
 * A restatement of the reference CPU path.  The float walk reproduces the
 * reference's mix of float and double arithmetic expression by expression
 * (C++ usual arithmetic conversions), so positions, tissue labels and phases are
 * the reference's bit for bit on the same toolchain.  Build flags matter: the
 * Makefile compiles with -O2 -ffp-contract=off and no -march, so that no fused
 * multiply-add changes a rounding.
 *
 * Choice recorded here: the reference calls unqualified `exp` on a float
 * argument inside namespace sim (kernels.cu:51).  With libstdc++ <cmath> only the
 * global C `double exp(double)` is found by that lookup, so the argument is
 * promoted and the result narrowed; this oracle does the same.  The difference
 * from expf is at most one float ulp and touches only M, not trajectories.
 */
#include "sw_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <math.h>
#include <random>
#include <sstream>
#include <iterator>
#include <thread>
#include <vector>

namespace {

constexpr double kDeg2Rad = 0.0174532925199433;  // simulation_parameters.cuh:24
constexpr double kRad2Deg = 57.2957795130823;    // simulation_parameters.cuh:25
constexpr double kGamma   = 267515315.;          // definitions.h:20, rad/s/T

// ---------------------------------------------------------------------------
// Rotations and relaxation, float flavour (kernels.cuh:77-230).
// ---------------------------------------------------------------------------
inline void xrot_f(float s, float c, const float *m0, float *m1) {   // kernels.cuh:82-87
    m1[0] = m0[0];
    m1[1] = c * m0[1] - s * m0[2];
    m1[2] = s * m0[1] + c * m0[2];
}
inline void yrot_f(float s, float c, const float *m0, float *m1) {   // kernels.cuh:108-113
    m1[0] = c * m0[0] + s * m0[2];
    m1[1] = m0[1];
    m1[2] = -s * m0[0] + c * m0[2];
}
inline void zrot_f(float s, float c, const float *m0, float *m1) {   // kernels.cuh:134-139
    m1[0] = c * m0[0] - s * m0[1];
    m1[1] = s * m0[0] + c * m0[1];
    m1[2] = m0[2];
}
// theta overloads: sinf/cosf of a float argument formed as float*double (kernels.cuh:95-100)
inline float sin_deg_f(float theta) { return sinf((float)(theta * kDeg2Rad)); }
inline float cos_deg_f(float theta) { return cosf((float)(theta * kDeg2Rad)); }

inline void xrot_withphase_f(float s, float c, float ph, const float *m0, float *m1) {  // kernels.cuh:160-195
    if (ph == 0.0f)                { xrot_f(s, c, m0, m1);  return; }
    if (ph == 180.0)               { xrot_f(-s, c, m0, m1); return; }
    if (ph == 90.0)                { yrot_f(s, c, m0, m1);  return; }
    if (ph == -90.0 || ph == 270.0){ yrot_f(-s, c, m0, m1); return; }
    float t[3];
    const float sp = sin_deg_f(ph), cp = cos_deg_f(ph);
    zrot_f(-sp, cp, m0, m1);
    xrot_f(s, c, m1, t);
    zrot_f(sp, cp, t, m1);
}
// (theta, phase) overload uses double sin/cos narrowed to float (kernels.cuh:203-206)
inline void xrot_withphase_deg_f(float theta, float ph, const float *m0, float *m1) {
    xrot_withphase_f((float)sin(theta * kDeg2Rad), (float)cos(theta * kDeg2Rad), ph, m0, m1);
}
inline void relax_f(float e1, float e2, const float *m0, float *m1) {  // kernels.cuh:214-219
    m1[0] = m0[0] * e2;
    m1[1] = m0[1] * e2;
    m1[2] = (float)(1. + e1 * (m0[2] - 1.));
}
inline void dephase_relax_f(const float *m0, float *m1, float phase, float T1, float T2, float t) {  // kernels.cu:45-52
    zrot_f(sin_deg_f(phase), cos_deg_f(phase), m0, m1);
    if (T1 >= 0 && T2 >= 0)
        relax_f((float)::exp((double)(-t / T1)), (float)::exp((double)(-t / T2)), m1, m1);
}

// ---------------------------------------------------------------------------
// Double flavour for the FP64 twin: same operations, no narrowing.
// ---------------------------------------------------------------------------
inline void xrot_d(double s, double c, const double *m0, double *m1) {
    m1[0] = m0[0]; m1[1] = c * m0[1] - s * m0[2]; m1[2] = s * m0[1] + c * m0[2];
}
inline void yrot_d(double s, double c, const double *m0, double *m1) {
    m1[0] = c * m0[0] + s * m0[2]; m1[1] = m0[1]; m1[2] = -s * m0[0] + c * m0[2];
}
inline void zrot_d(double s, double c, const double *m0, double *m1) {
    m1[0] = c * m0[0] - s * m0[1]; m1[1] = s * m0[0] + c * m0[1]; m1[2] = m0[2];
}
inline void xrot_withphase_d(double s, double c, double ph, const double *m0, double *m1) {
    if (ph == 0.0)                 { xrot_d(s, c, m0, m1);  return; }
    if (ph == 180.0)               { xrot_d(-s, c, m0, m1); return; }
    if (ph == 90.0)                { yrot_d(s, c, m0, m1);  return; }
    if (ph == -90.0 || ph == 270.0){ yrot_d(-s, c, m0, m1); return; }
    double t[3];
    const double sp = sin(ph * kDeg2Rad), cp = cos(ph * kDeg2Rad);
    zrot_d(-sp, cp, m0, m1);
    xrot_d(s, c, m1, t);
    zrot_d(sp, cp, t, m1);
}
inline void dephase_relax_d(const double *m0, double *m1, double phase, double T1, double T2, double t) {
    zrot_d(sin(phase * kDeg2Rad), cos(phase * kDeg2Rad), m0, m1);
    if (T1 >= 0 && T2 >= 0) {
        const double e1 = exp(-t / T1), e2 = exp(-t / T2);
        m1[0] *= e2; m1[1] *= e2; m1[2] = 1. + e1 * (m1[2] - 1.);
    }
}

inline int64_t sub2ind3(int64_t x, int64_t y, int64_t z, int64_t nx, int64_t ny, int64_t nz) {
    (void)nx;
    return x * nz * ny + y * nz + z;   // kernels.cuh:53-60 (ROW_MAJOR)
}

// ---------------------------------------------------------------------------
// One spin of the walk (kernels.cu:72-233).  MODE_F32 is the reference; the
// FP64 twin shares every trajectory decision and differs only in M/phase/field.
// ---------------------------------------------------------------------------
struct Draws {
    std::mt19937 gen_r, gen_u;
    std::normal_distribution<float> nd{0.f, 1.0f};
    std::uniform_real_distribution<float> ud{0.f, 1.f};
    Draws(uint64_t seed_plus_spin) : gen_r((std::mt19937::result_type)seed_plus_spin),
                                     gen_u((std::mt19937::result_type)seed_plus_spin) {}
};

template <bool F64>
struct Real { using type = float; };
template <>
struct Real<true> { using type = double; };

template <bool F64>
bool walk_spin(const orc_params &p, const orc_tables &q, const float *field_f32,
               const float *field_T, double f64_factor, const uint8_t *mask,
               const float *M0, const float *XYZ0,
               typename Real<F64>::type *M1, float *XYZ1, uint8_t *T,
               uint32_t spin_no, double *t_discard)
{
    using R = typename Real<F64>::type;
    const size_t traj = p.enRecordTrajectory ? (size_t)(p.n_dummy_scan + 1) * p.n_timepoints : 1;
    float *xyz1 = XYZ1 + 3 * (size_t)spin_no * traj;                 // kernels.cu:75

    auto t0 = std::chrono::steady_clock::now();
    Draws d(p.seed + spin_no);                                        // kernels.cu:82-85
    d.gen_r.discard(p.seed + spin_no);                                // kernels.cu:87
    d.gen_u.discard(p.seed + spin_no);                                // kernels.cu:88
    if (t_discard)
        *t_discard += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    uint32_t itr = 0;
    R field = 0., T1 = 0., T2 = 0., time_elapsed = 0.;               // kernels.cu:91
    const float rf_phase = q.RF_PH_deg[0];
    R m0[3], m1[3];
    double xyz_old[3], xyz_new[3], scale2grid[3];
    for (uint32_t i = 0, shift = 3 * spin_no; i < 3; i++) {          // kernels.cu:95-99
        xyz1[i] = XYZ0[shift + i];
        xyz_old[i] = xyz_new[i] = xyz1[i];
        m0[i] = M0[shift + i];
        scale2grid[i] = (double)p.phantom_size[i] / p.fov[i];
    }
    const int64_t nx = (int64_t)p.phantom_size[0], ny = (int64_t)p.phantom_size[1],
                  nz = (int64_t)p.phantom_size[2];
    int64_t indx = sub2ind3((int64_t)(xyz1[0] * scale2grid[0]), (int64_t)(xyz1[1] * scale2grid[1]),
                            (int64_t)(xyz1[2] * scale2grid[2]), nx, ny, nz);   // kernels.cu:102
    if (indx < 0 || indx >= p.matrix_length)
        return false;   // the reference reads out of bounds here; an oracle refuses instead
    uint8_t ts, ts_old;
    ts = ts_old = mask[indx];                                         // kernels.cu:103
    double diffusivity_scale = q.diffusivity[ts_old];                 // kernels.cu:104
    (void)ts;

    for (uint32_t dummy_scan = 0; dummy_scan < (uint32_t)p.n_dummy_scan + 1; dummy_scan++) {   // :107
        const bool is_lastscan = (dummy_scan == (uint32_t)p.n_dummy_scan);
        // :110-114  rf_phase + d*lin (float) + d(d+1)/2.0*quad (double), narrowed
        float new_rf_phase = (float)((double)(rf_phase + dummy_scan * p.linear_phase_cycling) +
                                     dummy_scan * (dummy_scan + 1) / 2.0 * p.quadratic_phase_cycling);
        while (new_rf_phase > 360.0) new_rf_phase = (float)(new_rf_phase - 360.0);
        while (new_rf_phase < 0)     new_rf_phase = (float)(new_rf_phase + 360.0);

        if (F64) {                                                     // :117
            // the twin forms the excitation's phase and sin/cos in double
            double ph = (double)rf_phase + (double)dummy_scan * p.linear_phase_cycling +
                        dummy_scan * (dummy_scan + 1) / 2.0 * p.quadratic_phase_cycling;
            while (ph > 360.0) ph -= 360.0;
            while (ph < 0) ph += 360.0;
            const double fa0 = q.RF_FA_deg[0];
            double a[3] = {(double)m0[0], (double)m0[1], (double)m0[2]}, b[3];
            xrot_withphase_d(sin(fa0 * kDeg2Rad), cos(fa0 * kDeg2Rad), ph, a, b);
            for (int i = 0; i < 3; i++) m1[i] = (R)b[i];
        } else {
            float a[3] = {(float)m0[0], (float)m0[1], (float)m0[2]}, b[3];
            xrot_withphase_f(p.s, p.c, new_rf_phase, a, b);
            for (int i = 0; i < 3; i++) m1[i] = (R)b[i];
        }
        for (int i = 0; i < 3; i++) m0[i] = m1[i];                     // :119-120

        int64_t ind = 0, ind_old = p.matrix_length + 1;                // :123-126
        uint32_t current_timepoint = 0, old_timepoint = 0;
        uint16_t current_rf = 1, current_te = 0, counter_dephasing = 0, counter_gradient = 0;
        R accumulated_phase = 0.f;

        while (current_timepoint < p.n_timepoints) {                   // :128
            for (uint8_t i = 0; i < 3 && diffusivity_scale != 0.; i++) {   // :130-137
                double rnd_wlk = d.nd(d.gen_r) * diffusivity_scale;
                xyz_new[i] = xyz_old[i] + rnd_wlk;
                if (xyz_new[i] < 0)
                    xyz_new[i] += (p.enCrossFOV ? p.fov[i] : 2 * std::abs(rnd_wlk));
                else if (xyz_new[i] >= p.fov[i])
                    xyz_new[i] -= (p.enCrossFOV ? p.fov[i] : 2 * std::abs(rnd_wlk));
            }
            ind = sub2ind3((int64_t)(xyz_new[0] * scale2grid[0]), (int64_t)(xyz_new[1] * scale2grid[1]),
                           (int64_t)(xyz_new[2] * scale2grid[2]), nx, ny, nz);   // :140
            if (ind >= p.matrix_length || ind < 0) {                   // :141-147
                for (int i = 0; i < 3; i++) xyz1[i] = (float)xyz_old[i];
                return false;
            }
            if (ind != ind_old) {                                      // :150
                ts = mask[ind];                                        // :152
                if (ts != ts_old) {                                    // :153-163
                    if (d.ud(d.gen_u) >= q.pXY[ts_old * p.n_substrate + ts]) {
                        if (itr++ > p.max_iterations)
                            return false;
                        continue;
                    }
                    ts_old = ts;
                }
                ind_old = ind;                                         // :165-169
                if (F64)
                    field = field_T ? (R)((double)field_T[ind] * f64_factor) : (R)0;
                else
                    field = field_f32 != nullptr ? (R)field_f32[ind] : (R)0.f;
                T1 = (R)(q.T1_ms[ts_old] * 1e-3);
                T2 = (R)(q.T2_ms[ts_old] * 1e-3);
                diffusivity_scale = q.diffusivity[ts_old];
            }
            accumulated_phase += field;                                // :171
            itr = 0;

            if (counter_dephasing < q.n_dephasing &&
                (uint32_t)q.dephasing_t[counter_dephasing] == current_timepoint) {   // :175-178
                if (F64)
                    accumulated_phase += (double)spin_no * q.dephasing_deg[counter_dephasing] / (double)p.n_spins;
                else
                    accumulated_phase += (float)spin_no * q.dephasing_deg[counter_dephasing] / (float)p.n_spins;
                counter_dephasing++;
            }
            if (counter_gradient < q.n_gradient &&
                (uint32_t)q.gradient_t[counter_gradient] == current_timepoint) {     // :181-187
                const float Gx = q.gradientX_mTm[counter_gradient];
                const float Gy = q.gradientY_mTm[counter_gradient];
                const float Gz = q.gradientZ_mTm[counter_gradient];
                accumulated_phase = (R)((double)accumulated_phase +
                    (Gx * xyz_new[0] + Gy * xyz_new[1] + Gz * xyz_new[2]) * 1e-3 * p.timestep_us * 1e-6 * kGamma * kRad2Deg);
                counter_gradient++;
            }
            if (current_rf < q.n_rf && (uint32_t)q.RF_t[current_rf] == current_timepoint) {   // :190-199
                time_elapsed = (R)((current_timepoint - old_timepoint) * (uint32_t)p.timestep_us * 1e-6);
                if (F64) {
                    double a[3] = {(double)m0[0], (double)m0[1], (double)m0[2]}, b[3], c2[3];
                    dephase_relax_d(a, b, accumulated_phase, T1, T2, time_elapsed);
                    const double th = q.RF_FA_deg[current_rf];
                    xrot_withphase_d(sin(th * kDeg2Rad), cos(th * kDeg2Rad), q.RF_PH_deg[current_rf], b, c2);
                    for (int i = 0; i < 3; i++) { m1[i] = (R)b[i]; m0[i] = (R)c2[i]; }
                } else {
                    float a[3] = {(float)m0[0], (float)m0[1], (float)m0[2]}, b[3], c2[3];
                    dephase_relax_f(a, b, (float)accumulated_phase, (float)T1, (float)T2, (float)time_elapsed);
                    xrot_withphase_deg_f(q.RF_FA_deg[current_rf], q.RF_PH_deg[current_rf], b, c2);
                    for (int i = 0; i < 3; i++) { m1[i] = (R)b[i]; m0[i] = (R)c2[i]; }
                }
                accumulated_phase = 0;
                old_timepoint = current_timepoint;
                current_rf++;
            }
            if (is_lastscan && current_te < q.n_te &&
                (uint32_t)q.TE_t[current_te] == current_timepoint) {   // :202-215
                time_elapsed = (R)((current_timepoint - old_timepoint) * (uint32_t)p.timestep_us * 1e-6);
                if (F64) {
                    double a[3] = {(double)m0[0], (double)m0[1], (double)m0[2]}, b[3];
                    dephase_relax_d(a, b, accumulated_phase, T1, T2, time_elapsed);
                    for (int i = 0; i < 3; i++) m1[i] = (R)b[i];
                } else {
                    float a[3] = {(float)m0[0], (float)m0[1], (float)m0[2]}, b[3];
                    dephase_relax_f(a, b, (float)accumulated_phase, (float)T1, (float)T2, (float)time_elapsed);
                    for (int i = 0; i < 3; i++) m1[i] = (R)b[i];
                }
                const size_t shift = 3 * (size_t)q.n_te * spin_no + 3 * (size_t)current_te;
                for (int i = 0; i < 3; i++) M1[shift + i] = m0[i] = m1[i];
                T[(size_t)spin_no * q.n_te + current_te] = ts_old;
                accumulated_phase = 0;
                old_timepoint = current_timepoint;
                current_te++;
            }
            if (p.enRecordTrajectory && (current_timepoint != 0 || dummy_scan != 0))   // :218-221
                xyz1 += 3;
            for (int i = 0; i < 3; i++) {
                xyz_old[i] = xyz_new[i];
                xyz1[i] = (float)xyz_old[i];
            }
            current_timepoint++;
        }
        time_elapsed = (R)((current_timepoint - old_timepoint) * (uint32_t)p.timestep_us * 1e-6);   // :226-231
        if (F64) {
            double a[3] = {(double)m0[0], (double)m0[1], (double)m0[2]}, b[3];
            dephase_relax_d(a, b, accumulated_phase, T1, T2, time_elapsed);
            for (int i = 0; i < 3; i++) m1[i] = (R)b[i];
        } else {
            float a[3] = {(float)m0[0], (float)m0[1], (float)m0[2]}, b[3];
            dephase_relax_f(a, b, (float)accumulated_phase, (float)T1, (float)T2, (float)time_elapsed);
            for (int i = 0; i < 3; i++) m1[i] = (R)b[i];
        }
        for (int i = 0; i < 3; i++) m0[i] = m1[i];
    }
    return true;
}

template <bool F64>
uint32_t run_spins(const orc_params *p, const orc_tables *q, const float *field_f32,
                   const float *field_T, double f64_factor, const uint8_t *mask,
                   const float *M0, const float *XYZ0, typename Real<F64>::type *M1,
                   float *XYZ1, uint8_t *T, uint32_t b, uint32_t e, int n_threads,
                   uint8_t *lost_flags, double *timing)
{
    if (e <= b) return 0;
    n_threads = std::max(1, n_threads);
    std::atomic<uint32_t> next{b}, lost{0};
    std::vector<double> disc(n_threads, 0.0);
    auto t0 = std::chrono::steady_clock::now();
    auto worker = [&](int tid) {
        for (;;) {
            const uint32_t s0 = next.fetch_add(16);
            if (s0 >= e) break;
            const uint32_t s1 = std::min(e, s0 + 16);
            for (uint32_t s = s0; s < s1; s++) {
                const bool ok = walk_spin<F64>(*p, *q, field_f32, field_T, f64_factor, mask, M0, XYZ0,
                                               M1, XYZ1, T, s, timing ? &disc[tid] : nullptr);
                if (!ok) { lost++; if (lost_flags) lost_flags[s] = 1; }
            }
        }
    };
    if (n_threads == 1) worker(0);
    else {
        std::vector<std::thread> th;
        for (int t = 0; t < n_threads; t++) th.emplace_back(worker, t);
        for (auto &t : th) t.join();
    }
    if (timing) {
        double s = 0; for (double v : disc) s += v;
        timing[0] = s / n_threads;   // wall-clock share of the discard across threads
        timing[1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return lost.load();
}

}  // namespace

extern "C" {

int64_t orc_sub2ind3(int64_t x, int64_t y, int64_t z, int64_t nx, int64_t ny, int64_t nz) {
    return sub2ind3(x, y, z, nx, ny, nz);
}
int64_t orc_sub2ind4(int64_t x, int64_t y, int64_t z, int64_t o, int64_t nx, int64_t ny, int64_t nz, int64_t no) {
    (void)nx;
    return x * no * nz * ny + y * nz * no + z * no + o;   // kernels.cuh:39-46 (ROW_MAJOR)
}
void orc_xrot_sc(float s, float c, const float *m0, float *m1) { xrot_f(s, c, m0, m1); }
void orc_yrot_sc(float s, float c, const float *m0, float *m1) { yrot_f(s, c, m0, m1); }
void orc_zrot_sc(float s, float c, const float *m0, float *m1) { zrot_f(s, c, m0, m1); }
void orc_xrot_deg(float t, const float *m0, float *m1) { xrot_f(sin_deg_f(t), cos_deg_f(t), m0, m1); }
void orc_yrot_deg(float t, const float *m0, float *m1) { yrot_f(sin_deg_f(t), cos_deg_f(t), m0, m1); }
void orc_zrot_deg(float t, const float *m0, float *m1) { zrot_f(sin_deg_f(t), cos_deg_f(t), m0, m1); }
void orc_xrot_withphase_sc(float s, float c, float ph, const float *m0, float *m1) { xrot_withphase_f(s, c, ph, m0, m1); }
void orc_xrot_withphase_deg(float t, float ph, const float *m0, float *m1) { xrot_withphase_deg_f(t, ph, m0, m1); }
void orc_relax(float e1, float e2, const float *m0, float *m1) { relax_f(e1, e2, m0, m1); }
void orc_dephase_relax(const float *m0, float *m1, float ph, float T1, float T2, float t) {
    dephase_relax_f(m0, m1, ph, T1, T2, t);
}

int orc_prepare(float RF_FA0_deg, int32_t TR_us, int32_t timestep_us, float T1_0_ms,
                double *diffusivity, uint32_t n_substrate, int32_t *n_dummy_scan,
                float *c, float *s, uint32_t *n_timepoints)
{
    if (timestep_us <= 0) return 1;
    *c = cosf((float)(RF_FA0_deg * kDeg2Rad));                        // simulation_parameters.cuh:229
    *s = sinf((float)(RF_FA0_deg * kDeg2Rad));                        // :230
    *n_timepoints = (uint32_t)(TR_us / timestep_us);                  // :231
    for (uint32_t i = 0; i < n_substrate; i++)                        // :236-237
        diffusivity[i] = 1e-3 * sqrt(2. * diffusivity[i] * timestep_us);
    if (*n_dummy_scan < 0)                                            // :239-242
        *n_dummy_scan = (int32_t)(5.0 * T1_0_ms / float(TR_us * 1e-3));
    return 0;
}

void orc_timing_scale(int32_t *v, uint32_t n, int32_t dt) {          // config_reader.cpp:39-46
    for (uint32_t i = 0; i < n; i++) v[i] = v[i] / dt;
}

float orc_tesla2deg(float B0, int32_t dt) {                          // monte_carlo.cu:241
    return (float)(B0 * dt * 1e-6 * kGamma * kRad2Deg);
}

void orc_scale_field(float *f, uint64_t n, float k) {                // monte_carlo.cu:243-244
    for (uint64_t i = 0; i < n; i++) f[i] = f[i] * k;
}

void orc_init_positions(uint64_t seed, const float fov[3], uint32_t n_spins, float *XYZ0) {
    std::mt19937 gen((std::mt19937::result_type)seed);                // monte_carlo.cu:142-151
    std::uniform_real_distribution<float> dx(0.01 * fov[0], 0.99 * fov[0]);
    std::uniform_real_distribution<float> dy(0.01 * fov[1], 0.99 * fov[1]);
    std::uniform_real_distribution<float> dz(0.01 * fov[2], 0.99 * fov[2]);
    for (uint32_t i = 0; i < n_spins; i++) {
        XYZ0[3 * i + 0] = dx(gen);
        XYZ0[3 * i + 1] = dy(gen);
        XYZ0[3 * i + 2] = dz(gen);
    }
}

int orc_str2vec_i32(const char *s, int32_t *out, int cap) {          // config_reader.cpp:16-23
    std::istringstream ss(s);
    int n = 0;
    for (std::istream_iterator<float> it(ss), end; it != end; ++it) {
        if (n < cap) out[n] = static_cast<int32_t>(*it);
        n++;
    }
    return n;
}

uint32_t orc_sim(const orc_params *p, const orc_tables *q, const float *field, const uint8_t *mask,
                 const float *M0, const float *XYZ0, float *M1, float *XYZ1, uint8_t *T,
                 uint32_t b, uint32_t e, int n_threads, uint8_t *lost_flags, double *timing)
{
    return run_spins<false>(p, q, field, nullptr, 0.0, mask, M0, XYZ0, M1, XYZ1, T, b, e,
                            n_threads, lost_flags, timing);
}

uint32_t orc_sim_f64(const orc_params *p, const orc_tables *q, float B0, const float *field_T,
                     const uint8_t *mask, const float *M0, const float *XYZ0, double *M1,
                     float *XYZ1, uint8_t *T, uint32_t b, uint32_t e, int n_threads,
                     uint8_t *lost_flags)
{
    const double k = (double)B0 * p->timestep_us * 1e-6 * kGamma * kRad2Deg;
    return run_spins<true>(p, q, nullptr, field_T, k, mask, M0, XYZ0, M1, XYZ1, T, b, e,
                           n_threads, lost_flags, nullptr);
}

void orc_record_draws(uint64_t seed, uint32_t spin, float *normals, uint64_t nn,
                      float *uniforms, uint64_t nu) {
    Draws d(seed + spin);
    d.gen_r.discard(seed + spin);
    d.gen_u.discard(seed + spin);
    for (uint64_t i = 0; i < nn; i++) normals[i] = d.nd(d.gen_r);
    for (uint64_t i = 0; i < nu; i++) uniforms[i] = d.ud(d.gen_u);
}

void orc_mt19937_raw(uint32_t seed, uint32_t *out, uint64_t n) {
    std::mt19937 g(seed);
    for (uint64_t i = 0; i < n; i++) out[i] = g();
}

void orc_echo_reduce(const float *M1, const uint8_t *T, uint32_t n_spins, uint32_t n_te,
                     uint32_t n_sub, double *out) {
    std::memset(out, 0, sizeof(double) * 3 * n_te * n_sub);
    for (uint32_t s = 0; s < n_spins; s++)
        for (uint32_t k = 0; k < n_te; k++) {
            const uint8_t lab = T[(size_t)s * n_te + k];
            if (lab >= n_sub) continue;
            double *o = out + 3 * ((size_t)k * n_sub + lab);
            o[0] += M1[3 * ((size_t)s * n_te + k) + 0];
            o[1] += M1[3 * ((size_t)s * n_te + k) + 1];
            o[2] += 1.0;
        }
}

}  // extern "C"
