// Host FP64 prologue: ICON cells/vertices -> triangle mesh -> spherical
// ECEF -> local ENU -> FP32 scene and per-cell ray frames.
//
// This is the only FP64 geometry of the library.  It follows
//   build_triangle_mesh      mo_lradtopo_horayzon.cpp:228-317
//   lonlat2ecef              :158-169
//   ENU origin + ecef2enu    :178-201, :610-644
//   per-cell frame / origin  :693-722
// with the same arithmetic (operation order included), so that the FP32
// vertices and ray origins handed to the device are the values the reference
// hands to Embree (:406-410, :715-720).  Work is split over the library's
// host worker pool; the only order-dependent reduction (the ENU origin sum,
// :627-631) stays sequential in cell order and runs on the calling thread
// while the pool builds the triangles.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "hz_internal.h"

namespace hz {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kRadEarth = 6371229.0;  // reference :584

inline double deg2rad(double a) { return (a / 180.0) * kPi; }

// sin and cos of one argument through glibc's sincos: the reference is built
// with gcc (setup.py:18-19), which turns its paired sin(x) / cos(x) calls
// into one sincos(x) call, and glibc's sincos does not always return the
// bits of separate sin and cos calls.
inline void sin_cos(double x, double* s, double* c) { ::sincos(x, s, c); }

// f(b, e) over [0, n) on the library's worker pool (small loops run here)
template <class F>
void parallel_for(long long n, F&& f) {
    if (n < 20000) {
        f(0LL, n);
        return;
    }
    pool_for(n, f);
}

// Valid adjacent cells of ICON vertex v, ordered clockwise from the
// latitude axis (reference :243-261; for <= 6 keys std::sort keeps equal
// angles in input order, which this insertion sort reproduces).
//
// The reference sorts by atan2(dlon, dlat) mapped to [0, 2 pi).  Only the
// order matters, so the directions are compared exactly instead: by the
// half-turn they lie in ([0, pi): dlon > 0, or dlon == 0 and dlat >= 0),
// then by the sign of their cross product.  Where directions lie closer than
// 1e-9 rad to each other or to the half-turn boundary, the rounding of
// atan2 could order them otherwise (or tie them), and the atan2 keys decide.
inline int sorted_cells(const double* clon, const double* clat, const double* vlon,
                        const double* vlat, const int* cov, int nvert, int v, int* cell,
                        double* hsum, const double* hsurf) {
    double dy[6], dx[6], r2[6];
    int half[6];
    int c6[6];
    int n = 0;
    double hs = 0.0;
    bool exact = true;
    constexpr double kSep2 = 1e-18;  // (1e-9 rad)^2, compared with squared quantities
    for (int j = 0; j < 6; j++) {
        int c = cov[(size_t)nvert * j + v];
        if (c == -2) continue;
        const double y = clon[c] - vlon[v], x = clat[c] - vlat[v];
        dy[n] = y;
        dx[n] = x;
        r2[n] = x * x + y * y;
        half[n] = (y > 0.0 || (y == 0.0 && x >= 0.0)) ? 0 : 1;
        if (!(y * y > kSep2 * r2[n])) exact = false;
        c6[n] = c;
        hs += hsurf[c];
        n++;
    }
    for (int i = 0; exact && i < n; i++)
        for (int k = i + 1; k < n; k++)
            if (half[i] == half[k] &&
                !((dx[i] * dy[k] - dy[i] * dx[k]) * (dx[i] * dy[k] - dy[i] * dx[k]) >
                  kSep2 * (r2[i] * r2[k]))) {
                exact = false;
                break;
            }
    double ang[6];
    if (!exact)
        for (int i = 0; i < n; i++) {
            double a = std::atan2(dy[i], dx[i]);
            if (a < 0.0) a += 2.0 * kPi;
            ang[i] = a;
        }
    auto before = [&](int a, int b) {
        if (!exact) return ang[a] < ang[b];
        if (half[a] != half[b]) return half[a] < half[b];
        return dx[a] * dy[b] - dy[a] * dx[b] > 0.0;
    };
    int idx[6];
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 1; i < n; i++) {
        int t = idx[i];
        int j = i - 1;
        while (j >= 0 && before(t, idx[j])) {
            idx[j + 1] = idx[j];
            j--;
        }
        idx[j + 1] = t;
    }
    for (int i = 0; i < n; i++) cell[i] = c6[idx[i]];
    *hsum = hs;
    return n;
}

}  // namespace

void lattice_table(double hori_acc, double2* out, double* sin_h) {
    for (int k = 0; k <= kMaxLattice; k++) {
        double sn, cs;
        sin_cos((double)k * (2.0 * hori_acc), &sn, &cs);
        out[k].x = cs;
        out[k].y = sn;
        sin_h[k] = std::sin((double)(2 * k + 1) * hori_acc);
    }
}

WalkParams make_walk_params(int num_hori, int refine_factor, int svf_type, double dist_search,
                            int ncell) {
    WalkParams p;
    int L = num_hori * refine_factor;
    p.hori_acc = deg2rad(0.25);
    p.elev_thresh = deg2rad(-85.0);
    sin_cos(deg2rad(360.0) / (double)L, &p.azim_sin, &p.azim_cos);
    sin_cos(2.0 * p.hori_acc, &p.elev_sin_2ha, &p.elev_cos_2ha);
    double shift = 0.0;
    if (refine_factor != 1)
        shift = -(deg2rad(360.0) / (2.0 * num_hori)) + (deg2rad(360.0) / (2.0 * (double)L));
    sin_cos(-shift, &p.shift_sin, &p.shift_cos);
    p.tfar = (float)dist_search;
    p.tfar_box = p.tfar * 1.0001f + 1.0f;
    p.num_hori = num_hori;
    p.refine_factor = refine_factor;
    p.svf_type = svf_type;
    p.ncell = ncell;
    p.elev_cs = nullptr;
    p.pred = nullptr;
    p.pred_stride = 0;
    p.sin_h = nullptr;
    p.reserve_blocks = 0;
    p.range = nullptr;
    p.start = nullptr;
    return p;
}

void enu_frame_from_sum(const double* P, int ncell, EnuFrame& E) {
    // centroid of the cell points, summed in cell order (reference :624-634)
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int i = 0; i < ncell; i++) {
        sx += P[3 * (size_t)i];
        sy += P[3 * (size_t)i + 1];
        sz += P[3 * (size_t)i + 2];
    }
    double rad = std::sqrt(sx * sx + sy * sy + sz * sz);
    E.lon0 = std::atan2(sy, sx);
    E.lat0 = std::asin(sz / rad);
    // ECEF -> ENU about the surface point below the origin (reference :178-201)
    sin_cos(E.lon0, &E.so, &E.co);
    sin_cos(E.lat0, &E.sa, &E.ca);
    const double R = kRadEarth;
    E.x0 = R * E.ca * E.co;
    E.y0 = R * E.ca * E.so;
    E.z0 = R * E.sa;
    auto enu = [&](double x, double y, double z, double* r) {
        double dx = x - E.x0, dy = y - E.y0, dz = z - E.z0;
        r[0] = -E.so * dx + E.co * dy;
        r[1] = -E.sa * E.co * dx - E.sa * E.so * dy + E.ca * dz;
        r[2] = E.ca * E.co * dx + E.ca * E.so * dy + E.sa * dz;
    };
    // earth centre and north pole in ENU (reference :643-644)
    enu(0.0, 0.0, 0.0, E.centre);
    enu(0.0, 0.0, R, E.pole);
}

void enu_frame_host(const double* clon, const double* clat, const double* hsurf, int ncell,
                    EnuFrame& E) {
    // spherical ECEF of the cell points (reference :158-169), then the origin
    HostVec<double> P((size_t)ncell * 3);
    const double R = kRadEarth;
    parallel_for(ncell, [&](long long b, long long e) {
        for (long long i = b; i < e; i++) {
            double sl, cl, sp, cp;
            sin_cos(clon[i], &sl, &cl);
            sin_cos(clat[i], &sp, &cp);
            double h = hsurf[i];
            P[3 * i] = (R + h) * cp * cl;
            P[3 * i + 1] = (R + h) * cp * sl;
            P[3 * i + 2] = (R + h) * sp;
        }
    });
    enu_frame_from_sum(P.data(), ncell, E);
}

int prologue(const double* clon, const double* clat, const double* hsurf, const double* vlon,
             const double* vlat, const int* cov, int ncell, int nvert, int grid_type,
             double ray_org_elev, HostScene& out, std::string& err) {
    const int rc = prologue_scene(clon, clat, hsurf, vlon, vlat, cov, ncell, nvert, grid_type,
                                  out, err);
    if (rc) return rc;
    prologue_frames(cov, ncell, nvert, ray_org_elev, out);
    return 0;
}

int prologue_scene(const double* clon, const double* clat, const double* hsurf,
                   const double* vlon, const double* vlat, const int* cov, int ncell, int nvert,
                   int grid_type, HostScene& out, std::string& err) {
    if (ncell < 1 || nvert < 0) {
        err = "num_cell must be >= 1 and num_vertex >= 0";
        return -1;
    }
    // --- pass 1: triangles (and added vertices) per ICON vertex -------------
    HostVec<int>& ntri_v = out.vcount;  // triangles of vertex v, then added vertices
    ntri_v.resize((size_t)2 * nvert);
    int* nadd_v = ntri_v.data() + nvert;
    parallel_for(nvert, [&](long long b, long long e) {
        for (long long v = b; v < e; v++) {
            int n = 0;
            for (int j = 0; j < 6; j++)
                if (cov[(size_t)nvert * j + v] != -2) n++;
            if (grid_type == 0) {
                ntri_v[v] = n >= 3 ? n - 2 : 0;
                nadd_v[v] = 0;
            } else {
                ntri_v[v] = n == 6 ? 6 : 0;
                nadd_v[v] = n == 6 ? 1 : 0;
            }
        }
    });
    HostVec<long long>& toff = out.voff;  // prefix sums of both counts
    toff.resize((size_t)2 * (nvert + 1));
    long long* aoff = toff.data() + (nvert + 1);
    toff[0] = aoff[0] = 0;
    for (int v = 0; v < nvert; v++) {
        toff[v + 1] = toff[v] + ntri_v[v];
        aoff[v + 1] = aoff[v] + nadd_v[v];
    }
    long long ntri = toff[nvert];
    long long nmv = (long long)ncell + aoff[nvert];
    if (ntri > (long long)kLeafFirstMask) {
        err = "mesh has more triangles than the library supports (2^28)";
        return -1;
    }
    // mesh vertices: the cell points (reference :234-238) straight to
    // spherical ECEF (:158-169); the added vertices (:300-301) as lon, lat,
    // elevation first
    HostVec<double>& P = out.work;  // every entry is written below
    P.resize((size_t)nmv * 3);
    out.tri.resize((size_t)ntri * 3);
    const double R = kRadEarth;
    auto ecef = [R](double lon, double lat, double h, double* p) {
        double sl, cl, sp, cp;
        sin_cos(lon, &sl, &cl);
        sin_cos(lat, &sp, &cp);
        p[0] = (R + h) * cp * cl;
        p[1] = (R + h) * cp * sl;
        p[2] = (R + h) * sp;
    };
    parallel_for(ncell, [&](long long b, long long e) {
        for (long long i = b; i < e; i++) ecef(clon[i], clat[i], hsurf[i], &P[3 * i]);
    });
    // --- pass 2: fill triangles (reference :243-312), on the pool, while
    // this thread sums the cell points in cell order for the ENU origin
    // (reference :624-634; pass 2 writes only the added vertices' entries)
    auto pass2 = [&](long long b, long long e) {
        int cell[6];
        double hs;
        for (long long v = b; v < e; v++) {
            if (ntri_v[v] == 0) continue;
            int n = sorted_cells(clon, clat, vlon, vlat, cov, nvert, (int)v, cell, &hs, hsurf);
            int* t = &out.tri[(size_t)toff[v] * 3];
            if (grid_type == 0) {
                for (int j = 1; j + 1 < n; j++) {  // fan about the first cell
                    *t++ = cell[0];
                    *t++ = cell[j];
                    *t++ = cell[j + 1];
                }
            } else {
                long long add = (long long)ncell + aoff[v];
                P[3 * add] = vlon[v];
                P[3 * add + 1] = vlat[v];
                P[3 * add + 2] = hs / 6.0;
                for (int j = 0; j < 6; j++) {
                    *t++ = cell[j];
                    *t++ = cell[(j + 1) % 6];
                    *t++ = (int)add;
                }
            }
        }
    };
    TaskGroup grp;
    const int nt2 = nvert < 20000 ? 0 : pool_threads();
    if (nt2 == 0) {
        pass2(0, nvert);
    } else {
        const long long chunk = ((long long)nvert + nt2 - 1) / nt2;
        for (int t = 0; t < nt2; t++) {
            const long long b = t * chunk, e = std::min<long long>(nvert, b + chunk);
            if (b < e) pool_submit(grp, [&pass2, b, e]() { pass2(b, e); });
        }
    }
    EnuFrame E;
    enu_frame_from_sum(P.data(), ncell, E);
    pool_wait(grp);
    out.lon0 = E.lon0;
    out.lat0 = E.lat0;
    const double so = E.so, co = E.co, sa = E.sa, ca = E.ca;
    const double x0 = E.x0, y0 = E.y0, z0 = E.z0;
    // --- ENU (reference :178-201, :643-644) ----------------------------------
    auto enu = [&](double x, double y, double z, double* r) {
        double dx = x - x0, dy = y - y0, dz = z - z0;
        r[0] = -so * dx + co * dy;
        r[1] = -sa * co * dx - sa * so * dy + ca * dz;
        r[2] = ca * co * dx + ca * so * dy + sa * dz;
    };
    std::memcpy(out.centre, E.centre, sizeof(out.centre));
    std::memcpy(out.pole, E.pole, sizeof(out.pole));
    out.vert32.resize((size_t)nmv * 3);
    std::mutex bmu;
    float bx[4] = {3e38f, 3e38f, -3e38f, -3e38f};
    parallel_for(nmv, [&](long long b, long long e) {
        float lb[4] = {3e38f, 3e38f, -3e38f, -3e38f};
        for (long long i = b; i < e; i++) {
            double* p = &P[3 * i];
            if (i >= ncell) ecef(p[0], p[1], p[2], p);  // an added vertex: ECEF first
            double r[3];
            enu(p[0], p[1], p[2], r);
            p[0] = r[0];
            p[1] = r[1];
            p[2] = r[2];
            // FP32 scene vertex (reference :406-410)
            const float x = (float)r[0], y = (float)r[1];
            out.vert32[3 * i] = x;
            out.vert32[3 * i + 1] = y;
            out.vert32[3 * i + 2] = (float)r[2];
            lb[0] = std::min(lb[0], x);
            lb[1] = std::min(lb[1], y);
            lb[2] = std::max(lb[2], x);
            lb[3] = std::max(lb[3], y);
        }
        std::lock_guard<std::mutex> g(bmu);
        bx[0] = std::min(bx[0], lb[0]);
        bx[1] = std::min(bx[1], lb[1]);
        bx[2] = std::max(bx[2], lb[2]);
        bx[3] = std::max(bx[3], lb[3]);
    });
    if (nmv == 0) bx[0] = bx[1] = bx[2] = bx[3] = 0.0f;
    std::memcpy(out.bounds_xy, bx, sizeof(bx));
    out.num_mesh_vertices = nmv;
    return 0;
}

void prologue_frames(const int* cov, int ncell, int nvert, double ray_org_elev,
                     HostScene& out) {
    const double* P = out.work.data();
    const double* centre = out.centre;
    const double* pole = out.pole;
    // --- per-cell frame (reference :693-722) ---------------------------------
    out.frame.resize(ncell);
    parallel_for(ncell, [&](long long b, long long e) {
        for (long long i = b; i < e; i++) {
            const double* p = &P[3 * i];
            CellFrame& f = out.frame[i];
            double n[3] = {p[0] - centre[0], p[1] - centre[1], p[2] - centre[2]};
            double m = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            n[0] /= m;
            n[1] /= m;
            n[2] /= m;
            double vn[3] = {pole[0] - p[0], pole[1] - p[1], pole[2] - p[2]};
            double dp = vn[0] * n[0] + vn[1] * n[1] + vn[2] * n[2];
            double N[3] = {vn[0] - dp * n[0], vn[1] - dp * n[1], vn[2] - dp * n[2]};
            double mN = std::sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
            N[0] /= mN;
            N[1] /= mN;
            N[2] /= mN;
            for (int a = 0; a < 3; a++) {
                f.n[a] = n[a];
                f.north[a] = N[a];
                f.org[a] = (float)(p[a] + n[a] * ray_org_elev);
            }
            f.pad = 0;
        }
    });
    // --- cells at mesh holes ---------------------------------------------------
    // A cell that is a corner of an ICON vertex with fewer than six cells (the
    // domain border, pentagons) sits next to a hole of the mesh: its rays can
    // leave through it and its walk then descends towards -85 deg, several
    // times the queries of an interior cell.  Flag it (CellFrame::pad bit 0)
    // so that the cell order can start these long searches first.
    for (int v = 0; v < nvert; v++) {
        int n = 0;
        for (int j = 0; j < 6; j++)
            if (cov[(size_t)nvert * j + v] != -2) n++;
        if (n == 6) continue;
        for (int j = 0; j < 6; j++) {
            const int c = cov[(size_t)nvert * j + v];
            if (c >= 0 && c < ncell) out.frame[c].pad |= 1;
        }
    }
}

void hole_cells(const int* cov, int ncell, int nvert, HostVec<int>& out) {
    out.clear();
    for (int v = 0; v < nvert; v++) {
        int n = 0;
        for (int j = 0; j < 6; j++)
            if (cov[(size_t)nvert * j + v] != -2) n++;
        if (n == 6) continue;
        for (int j = 0; j < 6; j++) {
            const int c = cov[(size_t)nvert * j + v];
            if (c >= 0 && c < ncell) out.push_back(c);
        }
    }
}

void cell_adjacency(const int* tri, long long ntri, int ncell, std::vector<int>& off,
                    std::vector<int>& adj) {
    off.assign((size_t)ncell + 1, 0);
    for (long long t = 0; t < ntri; t++)
        for (int k = 0; k < 3; k++) {
            const int v = tri[3 * t + k];
            if (v < ncell) off[v + 1]++;
        }
    for (int c = 0; c < ncell; c++) off[c + 1] += off[c];
    adj.assign((size_t)off[ncell], 0);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (long long t = 0; t < ntri; t++)
        for (int k = 0; k < 3; k++) {
            const int v = tri[3 * t + k];
            if (v < ncell) adj[fill[v]++] = (int)t;
        }
}

}  // namespace hz
