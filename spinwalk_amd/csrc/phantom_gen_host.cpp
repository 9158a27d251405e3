// phantom_gen_host.cpp -- seeded sampling of vessel shapes for the device
// phantom generator.
//
// Like the reference generator (src/phantom/phantom_cylinder.cpp:85-130,
// src/phantom/phantom_sphere.cpp:79-113) shapes are drawn with std::mt19937 and
// rejected when they overlap an accepted shape, until the target volume
// fraction is reached; unlike it, cylinder axes may point in any direction.
// Volume is accounted analytically: pi a^2 times the chord of the axis through
// the box for a cylinder, 4/3 pi a^3 for a sphere.
#include <algorithm>
#include <cmath>
#include <random>
#include <unordered_map>

#include "phantom_gen.hpp"

namespace sw {
namespace {

constexpr double kPi = 3.14159265358979323846;

// chord length of the line p + t u through the box [0,L]^3 (slab method)
double chord(const double p[3], const double u[3], const float L[3]) {
    double t0 = -1e300, t1 = 1e300;
    for (int i = 0; i < 3; i++) {
        if (std::fabs(u[i]) < 1e-12) {
            if (p[i] < 0 || p[i] > L[i]) return 0.0;
            continue;
        }
        double a = (0 - p[i]) / u[i], b = (L[i] - p[i]) / u[i];
        if (a > b) std::swap(a, b);
        t0 = std::max(t0, a);
        t1 = std::min(t1, b);
    }
    return t1 > t0 ? t1 - t0 : 0.0;
}

// distance between two infinite lines
double line_distance(const float p1[3], const float u1[3], const float p2[3], const float u2[3]) {
    const double n[3] = {(double)u1[1] * u2[2] - (double)u1[2] * u2[1], (double)u1[2] * u2[0] - (double)u1[0] * u2[2],
                         (double)u1[0] * u2[1] - (double)u1[1] * u2[0]};
    const double d[3] = {(double)p2[0] - p1[0], (double)p2[1] - p1[1], (double)p2[2] - p1[2]};
    const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (nn < 1e-9) {   // parallel: distance from p2 to line 1
        const double t = d[0] * u1[0] + d[1] * u1[1] + d[2] * u1[2];
        const double q[3] = {d[0] - t * u1[0], d[1] - t * u1[1], d[2] - t * u1[2]};
        return std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    }
    return std::fabs(d[0] * n[0] + d[1] * n[1] + d[2] * n[2]) / nn;
}

double point_line_distance(const float c[3], const GenCylinder &cy) {
    const double d[3] = {(double)c[0] - cy.p[0], (double)c[1] - cy.p[1], (double)c[2] - cy.p[2]};
    const double t = d[0] * cy.u[0] + d[1] * cy.u[1] + d[2] * cy.u[2];
    const double q[3] = {d[0] - t * cy.u[0], d[1] - t * cy.u[1], d[2] - t * cy.u[2]};
    return std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
}

}  // namespace

double sample_cylinders(const ShapeSpec &s, const float L[3], std::vector<GenCylinder> &out) {
    std::mt19937 gen(s.seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    std::normal_distribution<double> nrm(0.0, 1.0);
    const double a = s.radius_um, box = (double)L[0] * L[1] * L[2];
    const double target = s.volume_fraction / 100.0 * box;
    const double k = 2 * kPi * (1.0 - s.Y) * s.dchi;
    const double tr = (s.trunc_radii > 0 ? s.trunc_radii : 1e30) * a;
    double vol = 0;
    size_t start = out.size();
    for (long attempt = 0; vol < target && attempt < 2000000; attempt++) {
        double u[3];
        if (s.random_orientation) {
            do { u[0] = nrm(gen); u[1] = nrm(gen); u[2] = nrm(gen); } while (u[0] * u[0] + u[1] * u[1] + u[2] * u[2] < 1e-12);
        } else {
            const double th = s.orientation_deg * kPi / 180.0;   // axis tilted from z about y
            u[0] = std::sin(th); u[1] = 0; u[2] = std::cos(th);
        }
        const double nu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (double &x : u) x /= nu;
        double p[3] = {uni(gen) * L[0], uni(gen) * L[1], uni(gen) * L[2]};
        const double ch = chord(p, u, L);
        if (ch <= 0) continue;
        GenCylinder c{};
        for (int i = 0; i < 3; i++) { c.p[i] = (float)p[i]; c.u[i] = (float)u[i]; }
        bool ok = true;
        for (size_t j = start; j < out.size() && ok; j++)
            if (line_distance(c.p, c.u, out[j].p, out[j].u) < a + std::sqrt((double)out[j].a2)) ok = false;
        if (!ok) continue;
        const double ct = u[2];                                   // B0 along z
        const double st2 = 1.0 - ct * ct;
        double bp[3] = {0.0 - ct * u[0], 0.0 - ct * u[1], 1.0 - ct * u[2]};
        const double nb = std::sqrt(bp[0] * bp[0] + bp[1] * bp[1] + bp[2] * bp[2]);
        for (int i = 0; i < 3; i++) c.bp[i] = nb > 1e-9 ? (float)(bp[i] / nb) : 0.f;
        c.a2 = (float)(a * a);
        c.k_in = (float)(k * (ct * ct - 1.0 / 3.0));
        c.k_out = nb > 1e-9 ? (float)(k * a * a * st2) : 0.f;
        c.trunc2 = (float)std::min(tr * tr, 1e30);
        c.label = s.label;
        out.push_back(c);
        vol += kPi * a * a * ch;
    }
    return 100.0 * vol / box;
}

double sample_spheres(const ShapeSpec &s, const float L[3], const std::vector<GenCylinder> &cyl,
                      std::vector<GenSphere> &out) {
    std::mt19937 gen(s.seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    const double a = s.radius_um, box = (double)L[0] * L[1] * L[2];
    const double target = s.volume_fraction / 100.0 * box;
    const double k = 4 * kPi * (1.0 - s.Y) * s.dchi;
    const double tr = (s.trunc_radii > 0 ? s.trunc_radii : 10.0) * a;
    const double cell = std::max(2.0 * a, 1.0);
    const int nc[3] = {(int)std::ceil(L[0] / cell), (int)std::ceil(L[1] / cell), (int)std::ceil(L[2] / cell)};
    std::unordered_multimap<long, uint32_t> grid;
    auto key = [&](int x, int y, int z) { return ((long)x * nc[1] + y) * nc[2] + z; };
    double vol = 0;
    const size_t start = out.size();
    for (long attempt = 0; vol < target && attempt < 50000000; attempt++) {
        GenSphere g{};
        for (int i = 0; i < 3; i++) g.c[i] = (float)(uni(gen) * L[i]);
        const int cx = std::min((int)(g.c[0] / cell), nc[0] - 1), cy = std::min((int)(g.c[1] / cell), nc[1] - 1),
                  cz = std::min((int)(g.c[2] / cell), nc[2] - 1);
        bool ok = true;
        for (int ox = -1; ox <= 1 && ok; ox++)
            for (int oy = -1; oy <= 1 && ok; oy++)
                for (int oz = -1; oz <= 1 && ok; oz++) {
                    auto r = grid.equal_range(key(cx + ox, cy + oy, cz + oz));
                    for (auto it = r.first; it != r.second && ok; ++it) {
                        const GenSphere &o = out[it->second];
                        const double d2 = std::pow(g.c[0] - o.c[0], 2) + std::pow(g.c[1] - o.c[1], 2) +
                                          std::pow(g.c[2] - o.c[2], 2);
                        if (d2 < std::pow(a + std::sqrt((double)o.a2), 2)) ok = false;
                    }
                }
        for (size_t j = 0; j < cyl.size() && ok; j++)
            if (point_line_distance(g.c, cyl[j]) < a + std::sqrt((double)cyl[j].a2)) ok = false;
        if (!ok) continue;
        g.a2 = (float)(a * a);
        g.k_a3 = (float)(k * a * a * a);
        g.trunc2 = (float)(tr * tr);
        g.label = s.label;
        grid.emplace(key(cx, cy, cz), (uint32_t)out.size());
        out.push_back(g);
        vol += 4.0 / 3.0 * kPi * a * a * a;
    }
    (void)start;
    return 100.0 * vol / box;
}

// Axons: the reference cylinder generator's geometry (src/phantom/phantom_cylinder.cpp:
// 85-130, 183-275): cylinders parallel to z, B0 tilted from z by orientation_deg
// in the x-z plane, centres uniform in the x-y square, rejected when they fall
// inside an accepted axon.  With random radii (radius_um < 0, or radius_min_um > 0)
// a radius is drawn in [radius_min, |radius|] and shrunk to the gap when it
// overlaps a neighbour (check_cylinder_overlap, :24-60); a radius shrunk below
// radius_min is rejected.  The field is truncated at trunc_radii radii (default
// 20, the reference's bounding box of 20 radii).  A 2-D cell list of edge
// 2 |radius| keeps the overlap test local.
double sample_axons(const ShapeSpec &s, const float L[3], std::vector<GenCylinder> &out) {
    std::mt19937 gen(s.seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    const double rmax = std::fabs(s.radius_um);
    const bool rnd = s.radius_um < 0 || s.radius_min_um > 0;
    const double rmin = s.radius_min_um > 0 ? std::min((double)s.radius_min_um, rmax) : 0.0;
    const double area = (double)L[0] * L[1];
    const double target = s.volume_fraction / 100.0 * area;
    const double k = 2 * kPi * (1.0 - s.Y) * s.dchi;
    const double th = s.orientation_deg * kPi / 180.0;
    const double ct = std::cos(th), st2 = 1.0 - ct * ct;
    const double trr = s.trunc_radii > 0 ? s.trunc_radii : 20.0;
    const double cell = 2.0 * rmax;
    const int ncx = std::max(1, (int)std::ceil(L[0] / cell)), ncy = std::max(1, (int)std::ceil(L[1] / cell));
    std::vector<std::vector<uint32_t>> grid((size_t)ncx * ncy);
    struct Disk { double x, y, a; };
    std::vector<Disk> disks;
    double acc = 0;
    long misses = 0;
    // stop at the target, or after 2e6 draws in a row that place nothing (jammed)
    while (acc < target && misses < 2000000) {
        const double x = uni(gen) * L[0], y = uni(gen) * L[1];
        double a = rnd ? rmin + uni(gen) * (rmax - rmin) : rmax;
        const int cx = std::min((int)(x / cell), ncx - 1), cy = std::min((int)(y / cell), ncy - 1);
        bool ok = a > 0;
        for (int ox = -1; ox <= 1 && ok; ox++)
            for (int oy = -1; oy <= 1 && ok; oy++) {
                const int gx = cx + ox, gy = cy + oy;
                if (gx < 0 || gy < 0 || gx >= ncx || gy >= ncy) continue;
                for (uint32_t j : grid[(size_t)gx * ncy + gy]) {
                    const double d = std::hypot(x - disks[j].x, y - disks[j].y);
                    if (d <= disks[j].a) { ok = false; break; }
                    if (d < disks[j].a + a) {
                        if (!rnd) { ok = false; break; }
                        a = d - disks[j].a;
                        if (a < rmin || a <= 0) { ok = false; break; }
                    }
                }
            }
        if (!ok) { misses++; continue; }
        misses = 0;
        grid[(size_t)cx * ncy + cy].push_back((uint32_t)disks.size());
        disks.push_back({x, y, a});
        acc += kPi * a * a;
    }
    for (const Disk &dk : disks) {
        GenCylinder c{};
        c.p[0] = (float)dk.x; c.p[1] = (float)dk.y; c.p[2] = 0.f;
        c.u[0] = 0.f; c.u[1] = 0.f; c.u[2] = 1.f;
        c.bp[0] = 1.f; c.bp[1] = 0.f; c.bp[2] = 0.f;              // B0 projected on the x-y plane
        c.a2 = (float)(dk.a * dk.a);
        c.k_in = (float)(k * (ct * ct - 1.0 / 3.0));
        c.k_out = (float)(k * dk.a * dk.a * st2);
        // parallel to B0 there is no outside field: the lookup then only needs the radius
        c.trunc2 = st2 > 0 ? (float)(trr * trr * dk.a * dk.a) : c.a2;
        c.label = s.label;
        out.push_back(c);
    }
    return 100.0 * acc / area;
}

}  // namespace sw
