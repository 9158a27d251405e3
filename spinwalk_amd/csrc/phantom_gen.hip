// phantom_gen.hip -- voxelised vessel phantoms generated on the device.
//
// The reference generates phantoms on the host with OpenMP: random
// non-overlapping cylinders parallel to z with B0 tilted by an orientation
// angle (src/phantom/phantom_cylinder.cpp:85-130, 183-275) and random spheres
// (src/phantom/phantom_sphere.cpp:79-198), each adding the closed-form field of
// a magnetised cylinder or sphere to every voxel of a bounding box.  Here every
// voxel is one lane that sums the closed forms of all cylinders (any axis) and
// of the spheres and of the axons (cylinders parallel to z) in the neighbouring
// cells of their cell lists, and writes the label of the shape that contains it.  Fields are in tesla per 1 T of B0 (B0 along z):
//   cylinder, outside: 2 pi (1-Y) dChi (a^2/r^2) cos(2 phi) sin^2(theta)
//   cylinder, inside:  2 pi (1-Y) dChi (cos^2(theta) - 1/3)
//   sphere, outside:   4 pi (1-Y) dChi (a^3/r^3) (cos^2(phi) - 1/3), 0 inside
// (phantom_cylinder.cpp:254-265, phantom_sphere.cpp:185-187).
#include <hip/hip_runtime.h>

#include "phantom_gen.hpp"

namespace sw {
namespace {

__global__ void gen_kernel(GenArgs g) {
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.n) return;
    const uint32_t iz = (uint32_t)(v % g.nz);
    const uint32_t iy = (uint32_t)((v / g.nz) % g.ny);
    const uint32_t ix = (uint32_t)(v / ((uint64_t)g.nz * g.ny));
    const float x = ((float)ix + 0.5f) * g.vs[0];
    const float y = ((float)iy + 0.5f) * g.vs[1];
    const float z = ((float)iz + 0.5f) * g.vs[2];
    uint8_t label = g.background;
    float field = 0.f;
    for (uint32_t c = 0; c < g.n_cyl; c++) {
        const GenCylinder cy = g.cyl[c];
        const float dx = x - cy.p[0], dy = y - cy.p[1], dz = z - cy.p[2];
        const float t = dx * cy.u[0] + dy * cy.u[1] + dz * cy.u[2];
        const float qx = dx - t * cy.u[0], qy = dy - t * cy.u[1], qz = dz - t * cy.u[2];
        const float r2 = qx * qx + qy * qy + qz * qz;
        if (r2 <= cy.a2) {
            label = cy.label;
            field += cy.k_in;
        } else if (r2 < cy.trunc2) {
            const float pb = qx * cy.bp[0] + qy * cy.bp[1] + qz * cy.bp[2];
            field += cy.k_out * (2.f * pb * pb / r2 - 1.f) / r2;      // a^2/r^2 cos(2 phi)
        }
    }
    if (g.n_sph) {
        const int cx = min((int)(x / g.cell), (int)g.nc[0] - 1);
        const int cyi = min((int)(y / g.cell), (int)g.nc[1] - 1);
        const int cz = min((int)(z / g.cell), (int)g.nc[2] - 1);
        for (int ox = -1; ox <= 1; ox++) {
            const int ax = cx + ox;
            if (ax < 0 || ax >= (int)g.nc[0]) continue;
            for (int oy = -1; oy <= 1; oy++) {
                const int ay = cyi + oy;
                if (ay < 0 || ay >= (int)g.nc[1]) continue;
                for (int oz = -1; oz <= 1; oz++) {
                    const int az = cz + oz;
                    if (az < 0 || az >= (int)g.nc[2]) continue;
                    const uint32_t cell = ((uint32_t)ax * g.nc[1] + (uint32_t)ay) * g.nc[2] + (uint32_t)az;
                    for (uint32_t k = g.cell_start[cell]; k < g.cell_start[cell + 1]; k++) {
                        const GenSphere s = g.sph[k];
                        const float dx = x - s.c[0], dy = y - s.c[1], dz = z - s.c[2];
                        const float r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 <= s.a2) {
                            label = s.label;
                        } else if (r2 < s.trunc2) {
                            const float inv = 1.f / r2;
                            field += s.k_a3 * inv * sqrtf(inv) * (dz * dz * inv - (1.f / 3.f));
                        }
                    }
                }
            }
        }
    }
    if (g.n_axn) {
        const int cx = min((int)(x / g.acell), (int)g.anc[0] - 1);
        const int cyi = min((int)(y / g.acell), (int)g.anc[1] - 1);
        for (int ox = -1; ox <= 1; ox++) {
            const int ax = cx + ox;
            if (ax < 0 || ax >= (int)g.anc[0]) continue;
            for (int oy = -1; oy <= 1; oy++) {
                const int ay = cyi + oy;
                if (ay < 0 || ay >= (int)g.anc[1]) continue;
                const uint32_t cell = (uint32_t)ax * g.anc[1] + (uint32_t)ay;
                for (uint32_t k = g.axn_start[cell]; k < g.axn_start[cell + 1]; k++) {
                    const GenCylinder cy = g.axn[k];
                    const float qx = x - cy.p[0], qy = y - cy.p[1];
                    const float r2 = qx * qx + qy * qy;
                    if (r2 <= cy.a2) {
                        label = cy.label;
                        field += cy.k_in;
                    } else if (r2 < cy.trunc2) {
                        field += cy.k_out * (2.f * qx * qx / r2 - 1.f) / r2;   // B0 projection along x
                    }
                }
            }
        }
    }
    g.mask[v] = label;
    g.field[v] = field;
}

}  // namespace

hipError_t launch_phantom_gen(const GenArgs &g, hipStream_t s) {
    if (g.n == 0) return hipSuccess;
    hipLaunchKernelGGL(gen_kernel, dim3((unsigned)((g.n + 255) / 256)), dim3(256), 0, s, g);
    return hipGetLastError();
}

}  // namespace sw
