// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference horizon / sky-view-factor path
// (ChristianSteger/HORAYZON_extpar, mo_lradtopo_horayzon.cpp).  It is the
// checker for the HIP library: only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it.  The product path never links,
// loads or calls anything in this file.
//
// Contents and the reference lines each part follows:
//   * mesh builder (grid_type 0 / 1)      mo_lradtopo_horayzon.cpp:228-317
//   * spherical ECEF, ENU origin, ENU     :158-201, :610-644
//   * per-cell frame and FP32 origin      :693-722
//   * reference elevation walk            :490-555 (Rodrigues rotations :119-130)
//   * clip / sector mean / SVF            :333-358, :736-754
//   * occlusion query                     :446-460 (Embree rtcOccluded1, which
//     is third-party and absent; restated as the watertight ray/triangle test
//     of Woop, Benthin & Wald, JCGT 2(1) 2013, two-sided, t in (0, tfar])
//   * FP64 brute-force occlusion (Moller & Trumbore 1997) and an FP64
//     half-plane "true horizon" sweep, used to pin the FP32 decisions.
//
// Parity status of each function (what pins it) is listed in the header
// comment of tests/test_oracle.py.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off; no FMA contraction so
// the FP32 / FP64 operation order written below is the one executed).

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

const double kPi = 3.14159265358979323846;

inline double deg2rad(double a) { return (a / 180.0) * kPi; }
inline double rad2deg(double a) { return (a / kPi) * 180.0; }
// Paired sin / cos through glibc's sincos, which is what gcc emits for the
// reference's paired calls (its results differ from separate sin and cos in
// the last bit for a small fraction of arguments).
inline void sin_cos(double x, double* s, double* c) { ::sincos(x, s, c); }

struct V3 {
    double x, y, z;
};

inline double dot3(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross3(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline void normalise(V3& a) {
    double m = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    a.x /= m;
    a.y /= m;
    a.z /= m;
}
// Rodrigues rotation of v about unit axis k by the angle whose sine / cosine
// are given (reference :119-130).
inline V3 rodrigues(const V3& v, const V3& k, double s, double c) {
    double term = dot3(k, v) * (1.0 - c);
    V3 r;
    r.x = v.x * c + (k.y * v.z - k.z * v.y) * s + k.x * term;
    r.y = v.y * c + (k.z * v.x - k.x * v.z) * s + k.y * term;
    r.z = v.z * c + (k.x * v.y - k.y * v.x) * s + k.z * term;
    return r;
}

// ---------------------------------------------------------------------------
// Mesh (reference :228-317)
// ---------------------------------------------------------------------------

// Stable insertion sort of up to 6 angles; returns the permutation.  The
// reference uses std::sort (:137-146), which for so few elements performs an
// insertion sort that keeps equal keys in input order, as this does.
inline void angle_order(const double* ang, int n, int* idx) {
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 1; i < n; i++) {
        int t = idx[i];
        int j = i - 1;
        while (j >= 0 && ang[t] < ang[idx[j]]) {
            idx[j + 1] = idx[j];
            j--;
        }
        idx[j + 1] = t;
    }
}

struct Mesh {
    std::vector<V3> vert;       // lon, lat, elev -> later ENU [m]
    std::vector<int> tri;       // 3 per triangle
    V3 centre_enu, pole_enu;    // earth centre and north pole in ENU
    double lon0 = 0, lat0 = 0;  // ENU origin
};

void build_mesh(const double* clon, const double* clat, const double* hsurf,
                const double* vlon, const double* vlat, const int* cov, int ncell,
                int nvert, int grid_type, Mesh& m) {
    m.vert.resize(ncell);
    for (int i = 0; i < ncell; i++) m.vert[i] = {clon[i], clat[i], hsurf[i]};
    m.tri.clear();
    double ang[6];
    int cell[6], ord[6];
    for (int v = 0; v < nvert; v++) {
        int n = 0;
        double hsum = 0.0;
        for (int j = 0; j < 6; j++) {
            int c = cov[(size_t)nvert * j + v];
            if (c == -2) continue;
            // clockwise angle from the latitude axis (reference :249-254)
            double a = std::atan2(clon[c] - vlon[v], clat[c] - vlat[v]);
            if (a < 0.0) a += 2.0 * kPi;
            ang[n] = a;
            cell[n] = c;
            hsum += hsurf[c];
            n++;
        }
        if (grid_type == 0) {
            if (n < 3) continue;
            angle_order(ang, n, ord);
            for (int j = 1; j + 1 < n; j++) {  // fan about the first cell
                m.tri.push_back(cell[ord[0]]);
                m.tri.push_back(cell[ord[j]]);
                m.tri.push_back(cell[ord[j + 1]]);
            }
        } else {
            if (n != 6) continue;
            int extra = (int)m.vert.size();
            m.vert.push_back({vlon[v], vlat[v], hsum / 6.0});
            angle_order(ang, n, ord);
            for (int j = 0; j < 6; j++) {
                m.tri.push_back(cell[ord[j]]);
                m.tri.push_back(cell[ord[(j + 1) % 6]]);
                m.tri.push_back(extra);
            }
        }
    }
}

// Spherical ECEF then ENU about the centroid of the cell points
// (reference :158-201, :610-644).
void to_enu(Mesh& m, int ncell, double R) {
    for (auto& p : m.vert) {
        double sl, cl, sp, cp;
        sin_cos(p.x, &sl, &cl);
        sin_cos(p.y, &sp, &cp);
        double h = p.z;
        p = {(R + h) * cp * cl, (R + h) * cp * sl, (R + h) * sp};
    }
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int i = 0; i < ncell; i++) {
        sx += m.vert[i].x;
        sy += m.vert[i].y;
        sz += m.vert[i].z;
    }
    double rad = std::sqrt(sx * sx + sy * sy + sz * sz);
    m.lon0 = std::atan2(sy, sx);
    m.lat0 = std::asin(sz / rad);
    double so, co, sa, ca;
    sin_cos(m.lon0, &so, &co);
    sin_cos(m.lat0, &sa, &ca);
    double x0 = R * ca * co;
    double y0 = R * ca * so;
    double z0 = R * sa;
    auto enu = [&](V3 p) {
        double dx = p.x - x0, dy = p.y - y0, dz = p.z - z0;
        V3 r;
        r.x = -so * dx + co * dy;
        r.y = -sa * co * dx - sa * so * dy + ca * dz;
        r.z = ca * co * dx + ca * so * dy + sa * dz;
        return r;
    };
    for (auto& p : m.vert) p = enu(p);
    m.centre_enu = enu({0.0, 0.0, 0.0});
    m.pole_enu = enu({0.0, 0.0, R});
}

// ---------------------------------------------------------------------------
// FP32 watertight occlusion test (Woop, Benthin & Wald 2013)
// ---------------------------------------------------------------------------

struct Ray32 {
    float o[3], d[3];
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float tfar;
    float inv[3];
};

inline void ray_setup(Ray32& r) {
    float ad[3] = {std::fabs(r.d[0]), std::fabs(r.d[1]), std::fabs(r.d[2])};
    int kz = 0;
    if (ad[1] > ad[kz]) kz = 1;
    if (ad[2] > ad[kz]) kz = 2;
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    if (r.d[kz] < 0.0f) std::swap(kx, ky);
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.Sx = r.d[kx] / r.d[kz];
    r.Sy = r.d[ky] / r.d[kz];
    r.Sz = 1.0f / r.d[kz];
    for (int a = 0; a < 3; a++) {
        float dd = r.d[a];
        if (std::fabs(dd) < 1e-20f) dd = std::copysign(1e-20f, dd);
        r.inv[a] = 1.0f / dd;
    }
}

inline bool tri_hit32(const Ray32& r, const float* v0, const float* v1, const float* v2) {
    float A[3] = {v0[0] - r.o[0], v0[1] - r.o[1], v0[2] - r.o[2]};
    float B[3] = {v1[0] - r.o[0], v1[1] - r.o[1], v1[2] - r.o[2]};
    float C[3] = {v2[0] - r.o[0], v2[1] - r.o[1], v2[2] - r.o[2]};
    float Ax = A[r.kx] - r.Sx * A[r.kz];
    float Ay = A[r.ky] - r.Sy * A[r.kz];
    float Bx = B[r.kx] - r.Sx * B[r.kz];
    float By = B[r.ky] - r.Sy * B[r.kz];
    float Cx = C[r.kx] - r.Sx * C[r.kz];
    float Cy = C[r.ky] - r.Sy * C[r.kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * A[r.kz];
    float Bz = r.Sz * B[r.kz];
    float Cz = r.Sz * C[r.kz];
    float T = U * Az + V * Bz + W * Cz;
    float Ts = det < 0.0f ? -T : T;
    float ad = std::fabs(det);
    return (Ts > 0.0f) && (Ts <= r.tfar * ad);
}

// ---------------------------------------------------------------------------
// CPU BVH (binned SAH, BVH2, <= 4 triangles per leaf, conservative boxes)
// ---------------------------------------------------------------------------

struct BNode {
    float lo[3], hi[3];
    int left;   // internal: index of first child (second = left + 1); leaf: first prim
    int count;  // 0: internal, > 0: leaf primitive count
};

struct Scene {
    std::vector<float> vx;   // 3 floats per mesh vertex
    std::vector<int> tri;    // 3 per triangle
    std::vector<float> tv;   // 9 floats per triangle, in BVH leaf order
    std::vector<BNode> nodes;
    int ntri = 0;
};

// Every box is grown by this margin so that no ray accepted by the
// triangle test can be rejected by the (differently rounded) slab test.
inline float box_margin(float a) { return 1.0f + 1e-5f * std::fabs(a); }

void build_bvh(Scene& s) {
    int nt = s.ntri;
    s.nodes.clear();
    s.tv.assign((size_t)nt * 9, 0.0f);
    if (nt == 0) return;
    std::vector<float> cen((size_t)nt * 3), blo((size_t)nt * 3), bhi((size_t)nt * 3);
    std::vector<int> prim(nt);
    for (int t = 0; t < nt; t++) {
        prim[t] = t;
        for (int a = 0; a < 3; a++) {
            float p0 = s.vx[3 * s.tri[3 * t] + a], p1 = s.vx[3 * s.tri[3 * t + 1] + a],
                  p2 = s.vx[3 * s.tri[3 * t + 2] + a];
            blo[3 * t + a] = std::min(p0, std::min(p1, p2));
            bhi[3 * t + a] = std::max(p0, std::max(p1, p2));
            cen[3 * t + a] = 0.5f * (blo[3 * t + a] + bhi[3 * t + a]);
        }
    }
    struct Item {
        int node, first, count;
    };
    s.nodes.push_back({});
    std::vector<Item> work{{0, 0, nt}};
    const int NB = 16;
    while (!work.empty()) {
        Item it = work.back();
        work.pop_back();
        BNode nd;
        float clo[3] = {1e30f, 1e30f, 1e30f}, chi[3] = {-1e30f, -1e30f, -1e30f};
        for (int a = 0; a < 3; a++) {
            nd.lo[a] = 1e30f;
            nd.hi[a] = -1e30f;
        }
        for (int i = it.first; i < it.first + it.count; i++) {
            int t = prim[i];
            for (int a = 0; a < 3; a++) {
                nd.lo[a] = std::min(nd.lo[a], blo[3 * t + a]);
                nd.hi[a] = std::max(nd.hi[a], bhi[3 * t + a]);
                clo[a] = std::min(clo[a], cen[3 * t + a]);
                chi[a] = std::max(chi[a], cen[3 * t + a]);
            }
        }
        for (int a = 0; a < 3; a++) {
            nd.lo[a] -= box_margin(nd.lo[a]);
            nd.hi[a] += box_margin(nd.hi[a]);
        }
        int best_axis = -1, best_split = 0;
        if (it.count > 4) {
            double best = 1e300;
            for (int a = 0; a < 3; a++) {
                float ext = chi[a] - clo[a];
                if (!(ext > 0.0f)) continue;
                int cnt[NB] = {0};
                float bl[NB][3], bh[NB][3];
                for (int b = 0; b < NB; b++)
                    for (int q = 0; q < 3; q++) {
                        bl[b][q] = 1e30f;
                        bh[b][q] = -1e30f;
                    }
                for (int i = it.first; i < it.first + it.count; i++) {
                    int t = prim[i];
                    int b = (int)((cen[3 * t + a] - clo[a]) / ext * NB);
                    b = std::min(std::max(b, 0), NB - 1);
                    cnt[b]++;
                    for (int q = 0; q < 3; q++) {
                        bl[b][q] = std::min(bl[b][q], blo[3 * t + q]);
                        bh[b][q] = std::max(bh[b][q], bhi[3 * t + q]);
                    }
                }
                auto area = [](const float* l, const float* h) {
                    double dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
                    if (dx < 0) return 0.0;
                    return dx * dy + dy * dz + dz * dx;
                };
                // sweep
                double ra[NB];
                int rc[NB];
                float rl[3] = {1e30f, 1e30f, 1e30f}, rh[3] = {-1e30f, -1e30f, -1e30f};
                int c = 0;
                for (int b = NB - 1; b > 0; b--) {
                    c += cnt[b];
                    for (int q = 0; q < 3; q++) {
                        rl[q] = std::min(rl[q], bl[b][q]);
                        rh[q] = std::max(rh[q], bh[b][q]);
                    }
                    ra[b] = area(rl, rh);
                    rc[b] = c;
                }
                float ll[3] = {1e30f, 1e30f, 1e30f}, lh[3] = {-1e30f, -1e30f, -1e30f};
                c = 0;
                for (int b = 0; b < NB - 1; b++) {
                    c += cnt[b];
                    for (int q = 0; q < 3; q++) {
                        ll[q] = std::min(ll[q], bl[b][q]);
                        lh[q] = std::max(lh[q], bh[b][q]);
                    }
                    if (c == 0 || rc[b + 1] == 0) continue;
                    double cost = area(ll, lh) * c + ra[b + 1] * rc[b + 1];
                    if (cost < best) {
                        best = cost;
                        best_axis = a;
                        best_split = b + 1;
                    }
                }
            }
        }
        if (best_axis < 0) {
            if (it.count > 4) {
                // degenerate centroids: split in the middle of the list
                best_axis = 3;
                best_split = it.count / 2;
            }
        }
        if (best_axis < 0) {
            nd.left = it.first;
            nd.count = it.count;
            s.nodes[it.node] = nd;
            continue;
        }
        int mid;
        if (best_axis == 3) {
            mid = it.first + best_split;
        } else {
            int a = best_axis;
            float ext = chi[a] - clo[a];
            auto pm = std::partition(prim.begin() + it.first, prim.begin() + it.first + it.count,
                                     [&](int t) {
                                         int b = (int)((cen[3 * t + a] - clo[a]) / ext * NB);
                                         b = std::min(std::max(b, 0), NB - 1);
                                         return b < best_split;
                                     });
            mid = (int)(pm - prim.begin());
            if (mid == it.first || mid == it.first + it.count) mid = it.first + it.count / 2;
        }
        int l = (int)s.nodes.size();
        s.nodes.push_back({});
        s.nodes.push_back({});
        nd.left = l;
        nd.count = 0;
        s.nodes[it.node] = nd;
        work.push_back({l, it.first, mid - it.first});
        work.push_back({l + 1, mid, it.first + it.count - mid});
    }
    for (int i = 0; i < nt; i++) {
        int t = prim[i];
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++) s.tv[(size_t)9 * i + 3 * k + a] = s.vx[3 * s.tri[3 * t + k] + a];
    }
}

inline bool box_hit(const Ray32& r, const float* lo, const float* hi, float tfar_box) {
    float t0 = -1.0f, t1 = tfar_box;
    for (int a = 0; a < 3; a++) {
        float ta = (lo[a] - r.o[a]) * r.inv[a];
        float tb = (hi[a] - r.o[a]) * r.inv[a];
        t0 = std::max(t0, std::min(ta, tb));
        t1 = std::min(t1, std::max(ta, tb));
    }
    return t0 <= t1;
}

bool occluded_bvh(const Scene& s, const Ray32& r, long long* ntests) {
    if (s.ntri == 0) return false;
    float tfar_box = r.tfar * 1.0001f + 1.0f;
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const BNode& n = s.nodes[stack[--sp]];
        if (!box_hit(r, n.lo, n.hi, tfar_box)) continue;
        if (n.count) {
            for (int i = n.left; i < n.left + n.count; i++) {
                const float* t = &s.tv[(size_t)9 * i];
                if (ntests) (*ntests)++;
                if (tri_hit32(r, t, t + 3, t + 6)) return true;
            }
        } else {
            if (sp + 2 > 128) return true;  // cannot happen for depth < 126
            stack[sp++] = n.left + 1;
            stack[sp++] = n.left;
        }
    }
    return false;
}

bool occluded_brute(const Scene& s, const Ray32& r) {
    for (int t = 0; t < s.ntri; t++) {
        const float* v0 = &s.vx[3 * s.tri[3 * t]];
        const float* v1 = &s.vx[3 * s.tri[3 * t + 1]];
        const float* v2 = &s.vx[3 * s.tri[3 * t + 2]];
        if (tri_hit32(r, v0, v1, v2)) return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// Driver pieces
// ---------------------------------------------------------------------------

struct Frame {
    V3 n, north;
    float org[3];
};

struct Setup {
    Mesh mesh;
    Scene scene;
    std::vector<Frame> frame;
    int ncell = 0, L = 0, rf = 1, nhori = 0, svf_type = 2;
    double hori_acc, thresh, azim_sin, azim_cos, es2, ec2, sh_s, sh_c;
    float tfar;
    // cos, sin of k * 0.5 deg, k = 0 .. 180 (lattice directions of modes 1, 2)
    double lat_c[181], lat_s[181];
    const float* pred = nullptr;  // learned start [cell][24] deg (mode 2)
};

const double kRadEarth = 6371229.0;

void setup_domain(Setup& S, const double* clon, const double* clat, const double* hsurf,
                  const double* vlon, const double* vlat, const int* cov, int ncell, int nvert,
                  int nhori, int grid_type, double dist_search, double ray_org_elev, int rf,
                  int svf_type, bool bvh) {
    S.ncell = ncell;
    S.nhori = nhori;
    S.rf = rf;
    S.L = nhori * rf;
    S.svf_type = svf_type;
    build_mesh(clon, clat, hsurf, vlon, vlat, cov, ncell, nvert, grid_type, S.mesh);
    to_enu(S.mesh, ncell, kRadEarth);
    Scene& sc = S.scene;
    sc.vx.resize(S.mesh.vert.size() * 3);
    for (size_t i = 0; i < S.mesh.vert.size(); i++) {
        sc.vx[3 * i] = (float)S.mesh.vert[i].x;
        sc.vx[3 * i + 1] = (float)S.mesh.vert[i].y;
        sc.vx[3 * i + 2] = (float)S.mesh.vert[i].z;
    }
    sc.tri = S.mesh.tri;
    sc.ntri = (int)(S.mesh.tri.size() / 3);
    if (bvh) build_bvh(sc);
    S.frame.resize(ncell);
    const V3& c = S.mesh.centre_enu;
    const V3& q = S.mesh.pole_enu;
    for (int i = 0; i < ncell; i++) {
        const V3& p = S.mesh.vert[i];
        Frame& f = S.frame[i];
        f.n = {p.x - c.x, p.y - c.y, p.z - c.z};
        normalise(f.n);
        V3 vn = {q.x - p.x, q.y - p.y, q.z - p.z};
        double dp = dot3(vn, f.n);
        f.north = {vn.x - dp * f.n.x, vn.y - dp * f.n.y, vn.z - dp * f.n.z};
        normalise(f.north);
        f.org[0] = (float)(p.x + f.n.x * ray_org_elev);
        f.org[1] = (float)(p.y + f.n.y * ray_org_elev);
        f.org[2] = (float)(p.z + f.n.z * ray_org_elev);
    }
    S.hori_acc = deg2rad(0.25);
    S.thresh = deg2rad(-85.0);
    sin_cos(deg2rad(360.0) / (double)S.L, &S.azim_sin, &S.azim_cos);
    sin_cos(2.0 * S.hori_acc, &S.es2, &S.ec2);
    double shift = 0.0;
    if (rf != 1) shift = -(deg2rad(360.0) / (2.0 * nhori)) + (deg2rad(360.0) / (2.0 * (double)S.L));
    sin_cos(-shift, &S.sh_s, &S.sh_c);
    S.tfar = (float)dist_search;
    for (int k = 0; k <= 180; k++) sin_cos((double)k * (2.0 * S.hori_acc), &S.lat_s[k], &S.lat_c[k]);
}

inline bool cast(const Setup& S, const Frame& f, const V3& d, bool brute, long long* ntests) {
    Ray32 r;
    r.o[0] = f.org[0];
    r.o[1] = f.org[1];
    r.o[2] = f.org[2];
    r.d[0] = (float)d.x;
    r.d[1] = (float)d.y;
    r.d[2] = (float)d.z;
    r.tfar = S.tfar;
    ray_setup(r);
    return brute ? occluded_brute(S.scene, r) : occluded_bvh(S.scene, r, ntests);
}

// Reference walk for one cell (reference :490-555 and :736-754).
// mode 0: the reference walk (carried elevation, -85 deg floor).
// mode 1: clip-aware per-azimuth search h = 0.25 + 0.5 * max{k >= 0 : hit(k)},
//         walking up from k = 0 (equal to mode 0 when hits are monotone).
long long walk_cell(const Setup& S, int i, int mode, bool brute, double* horizon,
                    double* skyview, double* hcell, long long* ntests) {
    const Frame& f = S.frame[i];
    long long rays = 0;
    V3 dir = rodrigues(f.north, f.n, S.sh_s, S.sh_c);
    double e = 0.0;
    const double a = S.hori_acc;
    for (int j = 0; j < S.L; j++) {
        V3 ax = cross3(dir, f.n);
        normalise(ax);
        double h;
        if (mode == 0) {
            bool hit = cast(S, f, dir, brute, ntests);
            rays++;
            if (hit) {
                while (hit) {
                    if (e > kPi) break;  // guard for non-heightfield input (library flags it)
                    e += 2.0 * a;
                    dir = rodrigues(dir, ax, S.es2, S.ec2);
                    hit = cast(S, f, dir, brute, ntests);
                    rays++;
                }
                h = e - a;
            } else {
                while (!hit && e > S.thresh) {
                    e -= 2.0 * a;
                    dir = rodrigues(dir, ax, -S.es2, S.ec2);
                    hit = cast(S, f, dir, brute, ntests);
                    rays++;
                }
                h = e + a;
            }
        } else {
            // lattice search over the horizontal direction dir of this azimuth:
            // direction of index k = cos(k * 0.5 deg) dir + sin(k * 0.5 deg) n.
            // mode 1 starts at k = 0, mode 2 at the index nearest to the
            // predicted horizon; climb while hits, descend while misses
            auto ldir = [&](int k) {
                return V3{S.lat_c[k] * dir.x + S.lat_s[k] * f.n.x,
                          S.lat_c[k] * dir.y + S.lat_s[k] * f.n.y,
                          S.lat_c[k] * dir.z + S.lat_s[k] * f.n.z};
            };
            int k = 0;
            if (mode == 2) {
                float kf = std::floor(S.pred[(size_t)i * 24 + j] * 2.0f + 0.5f);
                k = !(kf >= 0.0f) ? 0 : (kf > 180.0f ? 180 : (int)kf);
            }
            int top = -1;
            bool hit = cast(S, f, ldir(k), brute, ntests);
            rays++;
            if (hit) {
                while (true) {
                    if (k >= 180) {  // guard for non-heightfield input
                        top = k;
                        break;
                    }
                    k++;
                    hit = cast(S, f, ldir(k), brute, ntests);
                    rays++;
                    if (!hit) {
                        top = k - 1;
                        break;
                    }
                }
            } else {
                while (k > 0) {
                    k--;
                    hit = cast(S, f, ldir(k), brute, ntests);
                    rays++;
                    if (hit) {
                        top = k;
                        break;
                    }
                }
            }
            h = top >= 0 ? (double)(2 * top + 1) * a : 0.0;
        }
        if (h < 0.0) h = 0.0;
        hcell[j] = h;
        dir = rodrigues(dir, f.n, -S.azim_sin, S.azim_cos);
    }
    for (int J = 0; J < S.nhori; J++) {
        double m = 0.0;
        for (int k = 0; k < S.rf; k++) m += hcell[J * S.rf + k];
        horizon[(size_t)J * S.ncell + i] = rad2deg(m) / (double)S.rf;
    }
    double svf = 0.0;
    for (int j = 0; j < S.L; j++) {
        double sn = std::sin(hcell[j]);
        if (S.svf_type == 0)
            svf += (1.0 - sn);
        else if (S.svf_type == 1)
            svf += (1.0 - (sn * sn));
        else
            svf += (1.0 - (sn * sn * sn));
    }
    skyview[i] = svf / (double)S.L;
    return rays;
}

// ---------------------------------------------------------------------------
// Certified lattice search (restates the library's certify.hip and the
// certified run of capi.cpp; no reference counterpart: it reproduces the
// reference walk, :521-546, with the lattice search where that is exact).
// Cells next to a mesh hole (a corner of an ICON vertex with fewer than six
// cells) and cells with such a cell within reach run the reference walk;
// every other cell sees hits monotone in elevation and runs the lattice
// search, whose results then equal the walk's.
// ---------------------------------------------------------------------------

struct CertSetup {
    std::vector<float> px, py, pz;  // boundary points (FP32 scene vertices)
    std::vector<char> flag;         // per cell: next to a mesh hole
    double lw = 0.0, reach = 0.0, r2max = 0.0;
};

// Boundary points; lw: longest triangle edge between two of them (FP32)
// with the library's margins; r2max: squared horizontal reach.
void cert_setup(const Setup& S, const int* cov, int nvert, CertSetup& C) {
    const int ncell = S.ncell;
    C.flag.assign(ncell, 0);
    for (int v = 0; v < nvert; v++) {
        int n = 0;
        for (int j = 0; j < 6; j++)
            if (cov[(size_t)nvert * j + v] != -2) n++;
        if (n == 6) continue;
        for (int j = 0; j < 6; j++) {
            int c = cov[(size_t)nvert * j + v];
            if (c >= 0 && c < ncell) C.flag[c] = 1;
        }
    }
    const std::vector<float>& vx = S.scene.vx;
    for (int i = 0; i < ncell; i++)
        if (C.flag[i]) {
            C.px.push_back(vx[3 * i]);
            C.py.push_back(vx[3 * i + 1]);
            C.pz.push_back(vx[3 * i + 2]);
        }
    float best = 0.0f;
    for (int t = 0; t < S.scene.ntri; t++)
        for (int e = 0; e < 3; e++) {
            int a = S.scene.tri[3 * t + e], b = S.scene.tri[3 * t + (e + 1) % 3];
            if (a >= ncell || b >= ncell || !C.flag[a] || !C.flag[b]) continue;
            float dx = vx[3 * b] - vx[3 * a], dy = vx[3 * b + 1] - vx[3 * a + 1],
                  dz = vx[3 * b + 2] - vx[3 * a + 2];
            best = std::max(best, std::sqrt(dx * dx + dy * dy + dz * dz));
        }
    C.lw = (double)best * 1.001 + 1.0;
    C.reach = (double)S.tfar;
    const double ulim = C.reach + C.lw;
    C.r2max = ulim * ulim + C.lw * C.lw;
}

// Does cell i have a boundary point within horizontal distance sqrt(r2max)
// in its own frame (east = north x normal)?  Brute force over the points.
bool cert_near(const Setup& S, const CertSetup& C, int i) {
    const Frame& f = S.frame[i];
    const double ox = f.org[0], oy = f.org[1], oz = f.org[2];
    const double n0 = f.n.x, n1 = f.n.y, n2 = f.n.z;
    const double N0 = f.north.x, N1 = f.north.y, N2 = f.north.z;
    const double E0 = N1 * n2 - N2 * n1, E1 = N2 * n0 - N0 * n2, E2 = N0 * n1 - N1 * n0;
    for (size_t q = 0; q < C.px.size(); q++) {
        const double rx = (double)C.px[q] - ox, ry = (double)C.py[q] - oy,
                     rz = (double)C.pz[q] - oz;
        const double uN = rx * N0 + ry * N1 + rz * N2;
        const double uE = rx * E0 + ry * E1 + rz * E2;
        if (uN * uN + uE * uE <= C.r2max) return true;
    }
    return false;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

struct orc_stats {
    long long num_rays;
    long long num_tri_tests;
    long long num_triangles;
    long long num_mesh_vertices;
    double lon0, lat0;
    double t_setup, t_trace;
    int nthreads;
    int cells_done;
    int cells_walked;  // certified modes: cells whose walk was cast in part
    int pad;
};

// Mesh only (reference :228-317).  Call with tri == NULL to get the counts.
// vert receives lon, lat, elev per mesh vertex (FP64) when non-NULL.
int orc_mesh(const double* clon, const double* clat, const double* hsurf, const double* vlon,
             const double* vlat, const int* cov, int ncell, int nvert, int grid_type, int* ntri,
             int* nmeshvert, double* vert, int* tri) {
    Mesh m;
    build_mesh(clon, clat, hsurf, vlon, vlat, cov, ncell, nvert, grid_type, m);
    *ntri = (int)(m.tri.size() / 3);
    *nmeshvert = (int)m.vert.size();
    if (tri) std::memcpy(tri, m.tri.data(), m.tri.size() * sizeof(int));
    if (vert)
        for (size_t i = 0; i < m.vert.size(); i++) {
            vert[3 * i] = m.vert[i].x;
            vert[3 * i + 1] = m.vert[i].y;
            vert[3 * i + 2] = m.vert[i].z;
        }
    return 0;
}

// ENU prologue (reference :610-644): mesh vertices in ENU (FP64), the FP32
// scene vertices, ENU earth centre / north pole, ENU origin lon/lat, and the
// per-cell frame (n, north: FP64; origin: FP32).
int orc_enu(const double* clon, const double* clat, const double* hsurf, const double* vlon,
            const double* vlat, const int* cov, int ncell, int nvert, int grid_type,
            double ray_org_elev, double* vert_enu, float* vert32, double* centre_pole_origin,
            double* frame_n, double* frame_north, float* org) {
    Setup S;
    setup_domain(S, clon, clat, hsurf, vlon, vlat, cov, ncell, nvert, 24, grid_type, 4e4,
                 ray_org_elev, 1, 2, false);
    for (size_t i = 0; i < S.mesh.vert.size(); i++) {
        if (vert_enu) {
            vert_enu[3 * i] = S.mesh.vert[i].x;
            vert_enu[3 * i + 1] = S.mesh.vert[i].y;
            vert_enu[3 * i + 2] = S.mesh.vert[i].z;
        }
        if (vert32)
            for (int a = 0; a < 3; a++) vert32[3 * i + a] = S.scene.vx[3 * i + a];
    }
    if (centre_pole_origin) {
        double* c = centre_pole_origin;
        c[0] = S.mesh.centre_enu.x;
        c[1] = S.mesh.centre_enu.y;
        c[2] = S.mesh.centre_enu.z;
        c[3] = S.mesh.pole_enu.x;
        c[4] = S.mesh.pole_enu.y;
        c[5] = S.mesh.pole_enu.z;
        c[6] = S.mesh.lon0;
        c[7] = S.mesh.lat0;
    }
    for (int i = 0; i < ncell; i++) {
        const Frame& f = S.frame[i];
        if (frame_n) {
            frame_n[3 * i] = f.n.x;
            frame_n[3 * i + 1] = f.n.y;
            frame_n[3 * i + 2] = f.n.z;
        }
        if (frame_north) {
            frame_north[3 * i] = f.north.x;
            frame_north[3 * i + 1] = f.north.y;
            frame_north[3 * i + 2] = f.north.z;
        }
        if (org)
            for (int a = 0; a < 3; a++) org[3 * i + a] = f.org[a];
    }
    return 0;
}

// Full horizon / SVF computation with the reference signature plus:
//   mode      0 = reference walk, 1 = clip-aware lattice climb from 0 deg,
//             2 = lattice search started at pred[cell * 24 + azimuth] [deg],
//             3 / 4 = modes 1 / 2 certified: the reference walk on the cells
//             with a mesh-boundary point in reach, the lattice search on the
//             others (brute is ignored)
//   brute     1 = test every triangle for every ray (no BVH)
//   nthreads  worker threads (<= 0: hardware concurrency)
//   cell_begin/cell_end  restrict the work to a cell range (outputs for other
//             cells are left untouched); cell_end <= 0 means ncell
//   time_budget_s  > 0: stop handing out cells after this many seconds;
//             stats->cells_done reports how many cells were finished (from
//             cell_begin on, in order of the chunks handed out)
int orc_horizon_svf(const double* clon, const double* clat, const double* hsurf,
                    const double* vlon, const double* vlat, const int* cov, double* horizon,
                    double* skyview, int ncell, int nvert, int nhori, int grid_type,
                    double dist_search, double ray_org_elev, int rf, int svf_type, int mode,
                    int brute, int nthreads, int cell_begin, int cell_end, double time_budget_s,
                    const float* pred, orc_stats* st) {
    double t0 = now_s();
    Setup S;
    setup_domain(S, clon, clat, hsurf, vlon, vlat, cov, ncell, nvert, nhori, grid_type,
                 dist_search, ray_org_elev, rf, svf_type, !brute);
    S.pred = pred;
    if ((mode == 2 || mode == 4) && (!pred || nhori * rf != 24)) return -1;
    if (mode < 0 || mode > 4) return -1;
    CertSetup C;
    if (mode >= 3) cert_setup(S, cov, nvert, C);
    std::atomic<int> walked_cells(0);
    double t1 = now_s();
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    if (cell_end <= 0 || cell_end > ncell) cell_end = ncell;
    if (cell_begin < 0) cell_begin = 0;
    std::atomic<int> next(cell_begin);
    std::atomic<long long> rays(0), tests(0);
    std::atomic<int> done(0);
    const int chunk = 16;
    auto worker = [&]() {
        std::vector<double> hc(S.L);
        long long r = 0, nt = 0;
        int d = 0;
        while (true) {
            if (time_budget_s > 0 && now_s() - t1 > time_budget_s) break;
            int b = next.fetch_add(chunk);
            if (b >= cell_end) break;
            int e = std::min(b + chunk, cell_end);
            for (int i = b; i < e; i++) {
                if (mode >= 3) {
                    // certified: the walk near mesh boundaries, the lattice
                    // search (mode 1 / 2) elsewhere
                    const bool w = C.flag[i] || cert_near(S, C, i);
                    r += walk_cell(S, i, w ? 0 : mode - 2, false, horizon, skyview, hc.data(),
                                   &nt);
                    walked_cells += w ? 1 : 0;
                } else {
                    r += walk_cell(S, i, mode, brute != 0, horizon, skyview, hc.data(), &nt);
                }
                d++;
            }
        }
        rays += r;
        tests += nt;
        done += d;
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(worker);
    for (auto& t : th) t.join();
    double t2 = now_s();
    if (st) {
        st->num_rays = rays.load();
        st->num_tri_tests = tests.load();
        st->num_triangles = S.scene.ntri;
        st->num_mesh_vertices = (long long)S.mesh.vert.size();
        st->lon0 = S.mesh.lon0;
        st->lat0 = S.mesh.lat0;
        st->t_setup = t1 - t0;
        st->t_trace = t2 - t1;
        st->nthreads = nthreads;
        st->cells_done = done.load();
        st->cells_walked = walked_cells.load();
    }
    return 0;
}



// Batch of independent rays against a triangle soup (FP32 scene vertices and
// triangle indices): hit[i] = OR over triangles of the watertight test.
// use_bvh selects the CPU BVH instead of testing every triangle.
int orc_occluded32(const float* vert, int nv, const int* tri, int ntri, const float* org,
                   const float* dir, int nray, float tfar, int use_bvh, unsigned char* hit) {
    Scene s;
    s.vx.assign(vert, vert + 3 * (size_t)nv);
    s.tri.assign(tri, tri + 3 * (size_t)ntri);
    s.ntri = ntri;
    if (use_bvh) build_bvh(s);
    for (int i = 0; i < nray; i++) {
        Ray32 r;
        for (int a = 0; a < 3; a++) {
            r.o[a] = org[3 * i + a];
            r.d[a] = dir[3 * i + a];
        }
        r.tfar = tfar;
        ray_setup(r);
        hit[i] = (use_bvh ? occluded_bvh(s, r, nullptr) : occluded_brute(s, r)) ? 1 : 0;
    }
    return 0;
}

// FP64 Moller-Trumbore two-sided occlusion over every triangle (no BVH).
// margin[i] receives, for the decisive triangle (first hit, or the nearest
// miss), the smallest barycentric slack in units of the triangle's size, so
// callers can tell a grazing decision from a clear one.
int orc_occluded64(const double* vert, int nv, const int* tri, int ntri, const double* org,
                   const double* dir, int nray, double tfar, unsigned char* hit, double* margin) {
    (void)nv;
    for (int i = 0; i < nray; i++) {
        const double* o = org + 3 * i;
        const double* d = dir + 3 * i;
        bool h = false;
        double best = 1e300;
        for (int t = 0; t < ntri && !h; t++) {
            const double* p0 = vert + 3 * tri[3 * t];
            const double* p1 = vert + 3 * tri[3 * t + 1];
            const double* p2 = vert + 3 * tri[3 * t + 2];
            double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
            double pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2],
                            d[0] * e2[1] - d[1] * e2[0]};
            double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
            if (det == 0.0) continue;
            double inv = 1.0 / det;
            double tv[3] = {o[0] - p0[0], o[1] - p0[1], o[2] - p0[2]};
            double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
            double qv[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2],
                            tv[0] * e1[1] - tv[1] * e1[0]};
            double v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) * inv;
            double tt = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
            double slack = std::min(std::min(u, v), 1.0 - u - v);
            bool inside = slack >= 0.0 && tt > 0.0 && tt <= tfar;
            if (inside) {
                h = true;
                best = slack;
            } else if (tt > 0.0 && tt <= tfar) {
                best = std::min(best, -slack);
            }
        }
        hit[i] = h ? 1 : 0;
        if (margin) margin[i] = best;
    }
    return 0;
}

// FP64 "true horizon": for each requested cell and each of the L = nhori*rf
// azimuths (same azimuth convention as the walk), the largest elevation
// angle [rad] of the mesh (FP32 scene vertices, evaluated in FP64) within
// the vertical half-plane of that azimuth and within dist_search of the
// FP32 ray origin.  Void azimuths return -pi/2.
//
// brute != 0 evaluates every triangle at every azimuth.  brute == 0 (the
// default) visits, per azimuth, only the triangles that can meet its
// half-plane inside the search ball: a triangle whose bounding box lies
// farther than dist_search from the origin cannot, and one whose horizontal
// projection (in the cell's frame) does not contain the origin meets the
// half-plane of azimuth phi only if phi lies in the arc its corners span as
// seen from the origin (widened by 1e-9 rad).  Every visited triangle runs
// the same FP64 arithmetic as in the brute-force form and the maximum does
// not depend on the order, so both forms return the same bits.

namespace {

// Largest elevation angle of the part of one triangle that lies in the
// half-plane w = 0, u >= 0 and within |(u, z)| <= sqrt(R2); P[k] = (u, w, z)
// of corner k.  -pi/2 where the triangle does not reach it.
double seg_max_elev(const double P[3][3], double R2) {
    double pts[3][2];
    int np = 0;
    for (int k = 0; k < 3 && np < 3; k++) {
        const double* A = P[k];
        const double* B = P[(k + 1) % 3];
        if (A[1] == 0.0) {
            pts[np][0] = A[0];
            pts[np][1] = A[2];
            np++;
        }
        if ((A[1] < 0.0 && B[1] > 0.0) || (A[1] > 0.0 && B[1] < 0.0)) {
            double s = A[1] / (A[1] - B[1]);
            pts[np][0] = A[0] + s * (B[0] - A[0]);
            pts[np][1] = A[2] + s * (B[2] - A[2]);
            np++;
        }
    }
    const double none = -kPi / 2;
    if (np < 2) return none;
    double a0 = pts[0][0], z0 = pts[0][1], a1 = pts[1][0], z1 = pts[1][1];
    // keep u >= 0
    if (a0 < 0 && a1 < 0) return none;
    if (a0 < 0 || a1 < 0) {
        double s = a0 / (a0 - a1);
        double zc = z0 + s * (z1 - z0);
        if (a0 < 0) {
            a0 = 0.0;
            z0 = zc;
        } else {
            a1 = 0.0;
            z1 = zc;
        }
    }
    // clip to the search radius: |(u, z)| <= tfar
    auto inside = [&](double u, double z) { return u * u + z * z <= R2; };
    bool i0 = inside(a0, z0), i1 = inside(a1, z1);
    if (!i0 && !i1) {
        // the segment may still cross the disc; test closest point
        double du = a1 - a0, dz = z1 - z0;
        double L2 = du * du + dz * dz;
        if (L2 == 0) return none;
        double s = -(a0 * du + z0 * dz) / L2;
        if (s <= 0 || s >= 1) return none;
        double cu = a0 + s * du, cz = z0 + s * dz;
        if (!inside(cu, cz)) return none;
        // both ends outside: clip both
        double disc = std::sqrt(std::max(0.0, R2 - (cu * cu + cz * cz)) / L2);
        double s0 = s - disc, s1 = s + disc;
        double na0 = a0 + s0 * du, nz0 = z0 + s0 * dz;
        double na1 = a0 + s1 * du, nz1 = z0 + s1 * dz;
        a0 = na0;
        z0 = nz0;
        a1 = na1;
        z1 = nz1;
    } else if (!i0 || !i1) {
        if (!i0) {
            std::swap(a0, a1);
            std::swap(z0, z1);
        }
        double du = a1 - a0, dz = z1 - z0;
        double A2 = du * du + dz * dz;
        double Bq = 2 * (a0 * du + z0 * dz);
        double Cq = a0 * a0 + z0 * z0 - R2;
        double s = (-Bq + std::sqrt(std::max(0.0, Bq * Bq - 4 * A2 * Cq))) / (2 * A2);
        a1 = a0 + s * du;
        z1 = z0 + s * dz;
    }
    double e0 = std::atan2(z0, a0), e1 = std::atan2(z1, a1);
    return std::max(e0, e1);
}

}  // namespace

int orc_true_horizon(const double* clon, const double* clat, const double* hsurf,
                     const double* vlon, const double* vlat, const int* cov, int ncell, int nvert,
                     int nhori, int grid_type, double dist_search, double ray_org_elev, int rf,
                     const int* cells, int ncells_req, double* out, int nthreads, int brute) {
    Setup S;
    setup_domain(S, clon, clat, hsurf, vlon, vlat, cov, ncell, nvert, nhori, grid_type,
                 dist_search, ray_org_elev, rf, 2, false);
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int> next(0);
    const double R2 = (double)S.tfar * (double)S.tfar;
    const int ntri = S.scene.ntri;
    const int nmv = (int)(S.scene.vx.size() / 3);
    // triangle bounding boxes (FP64 of the FP32 corners)
    std::vector<double> box((size_t)ntri * 6);
    for (int t = 0; t < ntri; t++) {
        double* bx = &box[(size_t)t * 6];
        for (int a = 0; a < 3; a++) {
            bx[a] = 1e300;
            bx[3 + a] = -1e300;
        }
        for (int k = 0; k < 3; k++) {
            const float* v = &S.scene.vx[3 * S.scene.tri[3 * t + k]];
            for (int a = 0; a < 3; a++) {
                bx[a] = std::min(bx[a], (double)v[a]);
                bx[3 + a] = std::max(bx[3 + a], (double)v[a]);
            }
        }
    }
    auto worker = [&]() {
        std::vector<double> ang;            // horizontal angle of each vertex from the origin
        std::vector<unsigned char> axis;    // vertex within 1 m of the origin's vertical axis
        std::vector<std::vector<int>> bin(S.L);
        while (true) {
            int q = next.fetch_add(1);
            if (q >= ncells_req) break;
            int i = cells[q];
            const Frame& f = S.frame[i];
            double o[3] = {f.org[0], f.org[1], f.org[2]};
            V3 east = cross3(f.north, f.n);
            double shift = 0.0;
            if (S.rf != 1)
                shift = -(deg2rad(360.0) / (2.0 * S.nhori)) + (deg2rad(360.0) / (2.0 * (double)S.L));
            const double dphi = 2.0 * kPi / S.L;
            auto tri_elev = [&](int t, const V3& hdir, const V3& side) {
                double P[3][3];  // (u, w, z) per corner
                for (int k = 0; k < 3; k++) {
                    const float* v = &S.scene.vx[3 * S.scene.tri[3 * t + k]];
                    V3 r = {v[0] - o[0], v[1] - o[1], v[2] - o[2]};
                    P[k][0] = dot3(r, hdir);
                    P[k][1] = dot3(r, side);
                    P[k][2] = dot3(r, f.n);
                }
                return seg_max_elev(P, R2);
            };
            if (!brute) {
                ang.resize(nmv);
                axis.resize(nmv);
                for (int v = 0; v < nmv; v++) {
                    const float* p = &S.scene.vx[3 * v];
                    V3 r = {p[0] - o[0], p[1] - o[1], p[2] - o[2]};
                    const double U = dot3(r, f.north), W = dot3(r, east);
                    ang[v] = std::atan2(W, U);
                    axis[v] = (U * U + W * W) < 1.0;  // within 1 m: every azimuth
                }
                for (auto& b : bin) b.clear();
                const double eps = 1e-9;
                for (int t = 0; t < ntri; t++) {
                    const double* bx = &box[(size_t)t * 6];
                    double d2 = 0.0;
                    for (int a = 0; a < 3; a++) {
                        const double d = std::max(std::max(bx[a] - o[a], o[a] - bx[3 + a]), 0.0);
                        d2 += d * d;
                    }
                    if (d2 > R2) continue;
                    const int* tv = &S.scene.tri[3 * t];
                    bool all = axis[tv[0]] || axis[tv[1]] || axis[tv[2]];
                    double lo = 0.0, span = 2.0 * kPi;
                    if (!all) {
                        // the arc the three corner angles span: the complement of the
                        // largest gap between them; a gap of at most pi means the
                        // projection may hold the origin, which every azimuth meets
                        double a3[3] = {ang[tv[0]], ang[tv[1]], ang[tv[2]]};
                        std::sort(a3, a3 + 3);
                        const double g[3] = {a3[1] - a3[0], a3[2] - a3[1], a3[0] + 2.0 * kPi - a3[2]};
                        int m = 0;
                        for (int k = 1; k < 3; k++)
                            if (g[k] > g[m]) m = k;
                        if (g[m] <= kPi + 1e-6) {
                            all = true;
                        } else {
                            lo = a3[(m + 1) % 3];
                            span = 2.0 * kPi - g[m];
                        }
                    }
                    if (all) {
                        for (int j = 0; j < S.L; j++) bin[j].push_back(t);
                        continue;
                    }
                    const long k0 = (long)std::ceil((lo - eps - shift) / dphi);
                    const long k1 = (long)std::floor((lo + span + eps - shift) / dphi);
                    if (k1 - k0 + 1 >= S.L) {
                        for (int j = 0; j < S.L; j++) bin[j].push_back(t);
                        continue;
                    }
                    for (long k = k0; k <= k1; k++) bin[(int)(((k % S.L) + S.L) % S.L)].push_back(t);
                }
            }
            for (int j = 0; j < S.L; j++) {
                double phi = shift + j * (2.0 * kPi / S.L);
                V3 hdir = {std::cos(phi) * f.north.x + std::sin(phi) * east.x,
                           std::cos(phi) * f.north.y + std::sin(phi) * east.y,
                           std::cos(phi) * f.north.z + std::sin(phi) * east.z};
                V3 side = cross3(f.n, hdir);  // normal of the vertical plane
                double best = -kPi / 2;
                if (brute) {
                    for (int t = 0; t < ntri; t++) best = std::max(best, tri_elev(t, hdir, side));
                } else {
                    for (int t : bin[j]) best = std::max(best, tri_elev(t, hdir, side));
                }
                out[(size_t)q * S.L + j] = best;
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) th.emplace_back(worker);
    for (auto& t : th) t.join();
    return 0;
}

}  // extern "C"
