// Slope and aspect of the terrain at each cell (no counterpart in the
// reference, whose sky view factor is for a horizontal plane,
// mo_lradtopo_horayzon.cpp:320-332).
//
// The surface normal at a cell is the area-weighted sum of the normals of
// the mesh triangles that have the cell's circumcentre as a corner (FP64
// from the FP32 scene, summed in a fixed order through a CSR list, so the
// result is deterministic).  slope = angle between that normal and the
// cell's sphere normal; aspect = azimuth, clockwise from north, of the
// normal's horizontal part (the downhill direction), in [0, 360); 0 for a
// level cell.

#include "hz_internal.h"

namespace hz {
namespace {

__global__ __launch_bounds__(256) void k_slope_aspect(const float* __restrict__ vert,
                                                      const int* __restrict__ tri,
                                                      const int* __restrict__ off,
                                                      const int* __restrict__ adj,
                                                      const CellFrame* __restrict__ cells, int ncell,
                                                      double* __restrict__ slope,
                                                      double* __restrict__ aspect) {
    int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncell) return;
    const CellFrame f = cells[c];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int q = off[c]; q < off[c + 1]; q++) {
        const int t = adj[q];
        const int a = tri[3 * t], b = tri[3 * t + 1], d = tri[3 * t + 2];
        const double ax = vert[3 * a], ay = vert[3 * a + 1], az = vert[3 * a + 2];
        const double ux = vert[3 * b] - ax, uy = vert[3 * b + 1] - ay, uz = vert[3 * b + 2] - az;
        const double vx = vert[3 * d] - ax, vy = vert[3 * d + 1] - ay, vz = vert[3 * d + 2] - az;
        double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
        // orient upwards (the triangulation's winding is not part of the mesh contract)
        if (nx * f.n[0] + ny * f.n[1] + nz * f.n[2] < 0.0) {
            nx = -nx;
            ny = -ny;
            nz = -nz;
        }
        sx += nx;
        sy += ny;
        sz += nz;
    }
    const double m = sqrt(sx * sx + sy * sy + sz * sz);
    if (!(m > 0.0)) {  // cell is not a corner of any triangle (mesh border)
        slope[c] = 0.0;
        aspect[c] = 0.0;
        return;
    }
    sx /= m;
    sy /= m;
    sz /= m;
    const double up = sx * f.n[0] + sy * f.n[1] + sz * f.n[2];
    const double nn = sx * f.north[0] + sy * f.north[1] + sz * f.north[2];
    // east = north x n
    const double ex = f.north[1] * f.n[2] - f.north[2] * f.n[1];
    const double ey = f.north[2] * f.n[0] - f.north[0] * f.n[2];
    const double ez = f.north[0] * f.n[1] - f.north[1] * f.n[0];
    const double ne = sx * ex + sy * ey + sz * ez;
    const double kDeg = 57.29577951308232;
    slope[c] = atan2(sqrt(nn * nn + ne * ne), up) * kDeg;
    double asp = atan2(ne, nn) * kDeg;
    if (asp < 0.0) asp += 360.0;
    aspect[c] = (nn == 0.0 && ne == 0.0) ? 0.0 : asp;
}

}  // namespace

void slope_aspect(const float* vert, const int* tri, const int* off, const int* adj,
                  const CellFrame* cells, int ncell, double* slope, double* aspect,
                  hipStream_t st) {
    hipLaunchKernelGGL(k_slope_aspect, dim3((ncell + 255) / 256), dim3(256), 0, st, vert, tri, off,
                       adj, cells, ncell, slope, aspect);
}

}  // namespace hz
