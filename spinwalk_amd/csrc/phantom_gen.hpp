// phantom_gen.hpp -- shapes and kernel arguments of the device phantom generator.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <vector>

namespace sw {

struct GenCylinder {
    float p[3];      // a point on the axis (um)
    float u[3];      // unit axis
    float bp[3];     // unit projection of B0 onto the plane normal to u (0 if u || B0)
    float a2;        // radius^2
    float k_in;      // 2 pi (1-Y) dChi (cos^2 theta - 1/3)
    float k_out;     // 2 pi (1-Y) dChi a^2 sin^2 theta
    float trunc2;    // field truncation radius^2 (um^2)
    uint8_t label, pad[3];
};

struct GenSphere {
    float c[3];
    float a2;
    float k_a3;      // 4 pi (1-Y) dChi a^3
    float trunc2;
    uint8_t label, pad[3];
};

struct GenArgs {
    uint64_t n;
    uint32_t nx, ny, nz;
    float vs[3];                 // voxel size (um)
    const GenCylinder *cyl; uint32_t n_cyl;
    const GenSphere *sph; uint32_t n_sph;
    const uint32_t *cell_start;  // [ncells+1], spheres sorted by cell
    uint32_t nc[3];
    float cell;                  // cell edge (um) >= sphere truncation radius
    // axons: cylinders parallel to z, sorted by their cell in the x-y plane
    const GenCylinder *axn; uint32_t n_axn;
    const uint32_t *axn_start;   // [anc[0]*anc[1]+1]
    uint32_t anc[2];
    float acell;                 // cell edge (um) >= axon truncation radius
    uint8_t background;
    uint8_t *mask;
    float *field;
};

hipError_t launch_phantom_gen(const GenArgs &g, hipStream_t s);

// Host-side shape sampling (phantom_gen_host.cpp).  Returns the analytic volume
// fraction placed (percent).
struct ShapeSpec {
    int32_t kind;               // 0 cylinder, 1 sphere
    uint8_t label;
    float radius_um;
    float volume_fraction;      // percent of the box
    int32_t random_orientation; // cylinders: isotropic axes when 1
    float orientation_deg;      // cylinders: axis tilt from B0 when not random
    float dchi, Y;
    uint32_t seed;
    float trunc_radii;          // field truncation in radii (0: default)
    float radius_min_um;        // > 0 (or radius_um < 0): random radii, shrunk to fit
};

double sample_cylinders(const ShapeSpec &s, const float fov_um[3], std::vector<GenCylinder> &out);
double sample_spheres(const ShapeSpec &s, const float fov_um[3], const std::vector<GenCylinder> &cyl,
                      std::vector<GenSphere> &out);
double sample_axons(const ShapeSpec &s, const float fov_um[3], std::vector<GenCylinder> &out);

}  // namespace sw
