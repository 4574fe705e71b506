// Horizon search kernels: the elevation search of every cell (all nhori *
// refine_factor azimuths) with the clip, sector mean and sky view factor
// fused behind it.  Replaces the TBB loop and the per-ray Embree queries of
// the reference (mo_lradtopo_horayzon.cpp:686-761, :490-555, :446-460).
//
// Numerics.  The walk state (ray direction, elevation angle, rotation axes)
// is FP64 and is updated with the reference's Rodrigues rotations in the
// reference's operation order (:119-130, :507-551); the file is compiled
// with -ffp-contract=off, so the FP32 ray handed to the occlusion query is
// the one the reference builds.  The occlusion query is the watertight
// ray/triangle test of Woop, Benthin & Wald (JCGT 2013), two-sided, accepting
// 0 < t <= (float)dist_search, over a BVH whose boxes are conservative
// (see lbvh.hip): its answer equals the OR of the triangle test over every
// triangle, independent of the tree.
//
// k_trace_pt (default): persistent lanes.  Each lane runs one cell's search
// as a state machine and keeps one occlusion query in flight; the wavefront
// alternates between traversal steps (one node per live lane) and a refill
// phase once fewer than `refill` lanes hold a live query, in which finished
// lanes consume their answer, advance the search and issue the next query
// (first against the cell's last occluding triangle, which answers about a
// third of the hit queries without a traversal) or take the next cell.
// Cells come from one work queue per XCD over its share of the Morton
// order (workgroups are dealt round-robin over the 8 XCDs), so each L2 sees
// a spatially compact stretch of cells.  The FP64 search state lives in a
// per-lane global area while the lane traverses, which keeps the traversal
// loop within 128 VGPRs (4 waves per SIMD).  The lattice searches (climb,
// learned start) begin every azimuth afresh, so for them a work item is one
// (cell, azimuth) pair: a chunk of 64 items is one azimuth of 64
// neighbouring cells (nearly parallel rays from nearby origins), the launch
// has 24x more items to share out, and k_items_finish forms the sector means
// and SVF from the per-item results.  The walk carries its elevation from one
// azimuth to the next and keeps whole cells as items.
// k_trace: one lane per cell with the azimuths and queries in lock step
// (kept as the simple reference schedule of the same arithmetic).
//
// The traversal stack keeps its first kLdsStack entries in LDS
// (entry-major, so lanes hit distinct banks) and deeper entries in a private
// array.  The lattice searches and the near-cell walk beside them step with
// the next node's boxes requested early, into an LDS node buffer
// (trav_step_pf, with 16 LDS stack entries).

#include "hz_internal.h"

namespace hz {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kLdsStack = 32;
// The Karras hierarchy is at most 63 code bits + 32 index bits = 95 levels
// deep: a binary step pushes at most one entry per level, a four-wide step
// (half as many levels) at most three, so 3 * 48 entries bound both.
constexpr int kStackMax = 160;
constexpr int kHintRounds = 8;  // queries answered by the last occluder per refill

struct D3 {
    double x, y, z;
};

__device__ __forceinline__ D3 rodrigues(const D3& v, const D3& k, double s, double c) {
    double term = (k.x * v.x + k.y * v.y + k.z * v.z) * (1.0 - c);
    D3 r;
    r.x = v.x * c + (k.y * v.z - k.z * v.y) * s + k.x * term;
    r.y = v.y * c + (k.z * v.x - k.x * v.z) * s + k.y * term;
    r.z = v.z * c + (k.x * v.y - k.y * v.x) * s + k.z * term;
    return r;
}

__device__ __forceinline__ D3 unit_cross(const D3& a, const D3& b) {
    D3 c = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    double m = sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
    c.x /= m;
    c.y /= m;
    c.z /= m;
    return c;
}

// Direction at lattice elevation k of the horizontal azimuth direction dir:
// cos(k * 0.5 deg) dir + sin(k * 0.5 deg) n (cos / sin from the host table).
__device__ __forceinline__ D3 lattice_dir(const D3& dir, const D3& n, const double2* cs, int k) {
    const double2 t = cs[k];
    return {t.x * dir.x + t.y * n.x, t.x * dir.y + t.y * n.y, t.x * dir.z + t.y * n.z};
}

// Lattice index nearest to a predicted horizon p [deg]: the search that
// starts there needs two queries when p is within 0.25 deg of the result.
__device__ __forceinline__ int start_index(float p) {
    float kf = floorf(p * 2.0f + 0.5f);
    if (!(kf >= 0.0f)) return 0;
    if (kf > (float)kMaxLattice) return kMaxLattice;
    return (int)kf;
}

// sin(h) of a horizon value: h is 0 or (2k + 1) * hori_acc up to the walk's
// accumulated rounding, so the host table of glibc values is used (a 1-ulp
// difference in h changes the SVF by ~1e-16).
__device__ __forceinline__ double sin_horizon(const WalkParams& P, double h) {
    if (!(h > 0.0)) return 0.0;
    int k = (int)rint(h / (2.0 * P.hori_acc) - 0.5);
    k = k < 0 ? 0 : (k > kMaxLattice ? kMaxLattice : k);
    return P.sin_h[k];
}

// Lane index (0..63), recomputed where it is used: the volatile asm keeps the
// compiler from holding it in a register across the traversal loop.
__device__ __forceinline__ int lane_now() {
    int l;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
    return l;
}

__device__ __forceinline__ float sel3(float a, float b, float c, int k) {
    return k == 0 ? a : (k == 1 ? b : c);
}

// Per-ray constants of the watertight test and of the slab test.
struct Ray {
    float o0, o1, o2;
    float Sx, Sy, Sz;
    int kk;  // the watertight test's axis permutation: kx | ky << 2 | kz << 4
    float ix, iy, iz;     // 1 / direction (zero components replaced by +-1e-20)
    float oix, oiy, oiz;  // origin * (1 / direction)
    // z' = z + (x^2 + y^2) kZCurv along the ray (the binary trees' z planes):
    // qa t^2 + qb t + qc, lowest at t = tv with value zv
    float qa, qb, qc, tv, zv;
    float tfar, tfar_box;
};

__device__ __forceinline__ Ray make_ray(float o0, float o1, float o2, float d0, float d1,
                                        float d2, float tfar, float tfar_box) {
    Ray r;
    float a0 = fabsf(d0), a1 = fabsf(d1), a2 = fabsf(d2);
    int kz = a1 > a0 ? 1 : 0;
    float akz = kz ? a1 : a0;
    if (a2 > akz) kz = 2;
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dkz = sel3(d0, d1, d2, kz);
    if (dkz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    float dkx = sel3(d0, d1, d2, kx), dky = sel3(d0, d1, d2, ky);
    r.Sx = dkx / dkz;
    r.Sy = dky / dkz;
    r.Sz = 1.0f / dkz;
    r.kk = kx | ky << 2 | kz << 4;
    r.o0 = o0;
    r.o1 = o1;
    r.o2 = o2;
    float e0 = fabsf(d0) < 1e-20f ? copysignf(1e-20f, d0) : d0;
    float e1 = fabsf(d1) < 1e-20f ? copysignf(1e-20f, d1) : d1;
    float e2 = fabsf(d2) < 1e-20f ? copysignf(1e-20f, d2) : d2;
    // slab-test reciprocals: the hardware approximation (1 ulp) is enough, the
    // boxes carry a margin of at least 1 m; the watertight test's Sx, Sy, Sz
    // above are IEEE divisions
    r.ix = __builtin_amdgcn_rcpf(e0);
    r.iy = __builtin_amdgcn_rcpf(e1);
    r.iz = __builtin_amdgcn_rcpf(e2);
    r.oix = o0 * r.ix;
    r.oiy = o1 * r.iy;
    r.oiz = o2 * r.iz;
    // FP32 is enough: the z' planes carry the boxes' margin of at least 1 m,
    // far above the rounding of these terms (mm over +-400 km)
    const float kc = (float)kZCurv;
    r.qa = (d0 * d0 + d1 * d1) * kc;
    r.qb = d2 + 2.0f * (o0 * d0 + o1 * d1) * kc;
    r.qc = o2 + (o0 * o0 + o1 * o1) * kc;
    // the vertex: the hardware reciprocal's 1-ulp error moves zv by far less
    // than a millimetre wherever tv lies inside a box's interval (|qb| is
    // then below 2 qa tfar)
    const bool vtx = r.qa > 0.0f;  // a vertical ray: z' linear, no vertex
    const float h = __builtin_amdgcn_rcpf(2.0f * r.qa);
    r.tv = vtx ? -r.qb * h : -3.0e38f;
    r.zv = vtx ? __builtin_fmaf(0.5f * r.qb, r.tv, r.qc) : 0.0f;
    r.tfar = tfar;  // uniform kernel arguments: scalar registers
    r.tfar_box = tfar_box;
    return r;
}

// Watertight two-sided test; operation order fixed (no contraction).
__device__ __forceinline__ bool tri_hit(const Ray& r, float v0x, float v0y, float v0z, float v1x,
                                        float v1y, float v1z, float v2x, float v2y, float v2z) {
    float A0 = v0x - r.o0, A1 = v0y - r.o1, A2 = v0z - r.o2;
    float B0 = v1x - r.o0, B1 = v1y - r.o1, B2 = v1z - r.o2;
    float C0 = v2x - r.o0, C1 = v2y - r.o1, C2 = v2z - r.o2;
    const int kx = r.kk & 3, ky = (r.kk >> 2) & 3, kz = r.kk >> 4;
    float Akx = sel3(A0, A1, A2, kx), Aky = sel3(A0, A1, A2, ky), Akz = sel3(A0, A1, A2, kz);
    float Bkx = sel3(B0, B1, B2, kx), Bky = sel3(B0, B1, B2, ky), Bkz = sel3(B0, B1, B2, kz);
    float Ckx = sel3(C0, C1, C2, kx), Cky = sel3(C0, C1, C2, ky), Ckz = sel3(C0, C1, C2, kz);
    float Ax = Akx - r.Sx * Akz;
    float Ay = Aky - r.Sy * Akz;
    float Bx = Bkx - r.Sx * Bkz;
    float By = Bky - r.Sy * Bkz;
    float Cx = Ckx - r.Sx * Ckz;
    float Cy = Cky - r.Sy * Ckz;
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
    float Az = r.Sz * Akz;
    float Bz = r.Sz * Bkz;
    float Cz = r.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float Ts = det < 0.0f ? -T : T;
    float ad = fabsf(det);
    return (Ts > 0.0f) && (Ts <= r.tfar * ad);
}

template <bool COUNT>
__device__ __forceinline__ bool leaf_hit(const Ray& r, const float4* __restrict__ tris, int ref,
                                         unsigned& ntest, int& hit_leaf) {
    int code = ~ref;
    int first = code & kLeafFirstMask;
    int cnt = (code >> kLeafFirstBits) + 1;
    for (int k = 0; k < cnt; k++) {
        const float4* t = tris + 3 * (first + k);
        float4 t0 = t[0], t1 = t[1], t2 = t[2];
        if (COUNT) ntest++;
        if (tri_hit(r, t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w, t2.x)) {
            hit_leaf = first + k;
            return true;
        }
    }
    return false;
}

// The two child boxes of a binary node (a = x planes, b = y planes, c = z'
// planes: lo0 lo1 hi0 hi1 each) against the ray: x / y slabs give the
// ray's interval [ta, tb] in the box's x/y footprint (within [-1, tfar_box]);
// over it the ray's z' is convex in t, so its range is spanned by the two
// ends and the vertex tv when inside; the box is hit when that range meets
// the box's z' range.  ta is the entry (for the nearer-first order) and g
// the height of the box top above the ray's lowest point in the box.
__device__ __forceinline__ void two_boxes(const Ray& r, const float4& a, const float4& b,
                                          const float4& c, bool& h0, bool& h1, float& ta0,
                                          float& ta1, float& g0, float& g1) {
    const float x0l = __builtin_fmaf(a.x, r.ix, -r.oix), x0h = __builtin_fmaf(a.z, r.ix, -r.oix);
    const float x1l = __builtin_fmaf(a.y, r.ix, -r.oix), x1h = __builtin_fmaf(a.w, r.ix, -r.oix);
    const float y0l = __builtin_fmaf(b.x, r.iy, -r.oiy), y0h = __builtin_fmaf(b.z, r.iy, -r.oiy);
    const float y1l = __builtin_fmaf(b.y, r.iy, -r.oiy), y1h = __builtin_fmaf(b.w, r.iy, -r.oiy);
    ta0 = fmaxf(fmaxf(fminf(x0l, x0h), fminf(y0l, y0h)), -1.0f);
    ta1 = fmaxf(fmaxf(fminf(x1l, x1h), fminf(y1l, y1h)), -1.0f);
    const float tb0 = fminf(fminf(fmaxf(x0l, x0h), fmaxf(y0l, y0h)), r.tfar_box);
    const float tb1 = fminf(fminf(fmaxf(x1l, x1h), fmaxf(y1l, y1h)), r.tfar_box);
    const float za0 = __builtin_fmaf(__builtin_fmaf(r.qa, ta0, r.qb), ta0, r.qc);
    const float zb0 = __builtin_fmaf(__builtin_fmaf(r.qa, tb0, r.qb), tb0, r.qc);
    const float za1 = __builtin_fmaf(__builtin_fmaf(r.qa, ta1, r.qb), ta1, r.qc);
    const float zb1 = __builtin_fmaf(__builtin_fmaf(r.qa, tb1, r.qb), tb1, r.qc);
    float lo0 = fminf(za0, zb0), lo1 = fminf(za1, zb1);
    const float hi0 = fmaxf(za0, zb0), hi1 = fmaxf(za1, zb1);
    if (r.tv > ta0 && r.tv < tb0) lo0 = fminf(lo0, r.zv);
    if (r.tv > ta1 && r.tv < tb1) lo1 = fminf(lo1, r.zv);
    h0 = ta0 <= tb0 && lo0 <= c.z && hi0 >= c.x;
    h1 = ta1 <= tb1 && lo1 <= c.w && hi1 >= c.y;
    g0 = c.z - lo0;
    g1 = c.w - lo1;
}

// The searches that take the taller-box order below: the walk and the learned
// start, whose queries sit near the horizon, where distant high ground
// occludes (C2: walk trace 31.6 -> 30.4 ms, uncertified learned items
// 16.7 -> 15.9 ms).  The climb's queries start at 0 deg and mostly hit the
// ground nearby, where the nearer child first is better (certified climb
// 27.4 ms against 29.6-30.2 ms with the taller-box order).
constexpr bool zorder(int mode) { return mode != 1; }

// The child to descend into first when both children of a node are
// interior and hit.  With a last occluder `hint` (leaf-order index, -1:
// none) and the node's leaf split (child 0's leaves end at `split`, -1:
// unknown): the child whose leaf range holds the hint, or lies on its side
// in leaf order.  A search's next query goes a step up or down from the last
// one, and its occluder is often that triangle or one near it (C2: walk
// trace 30.4 -> 29.7 ms, certified climb 27.3 -> 25.2-26.3 ms).  Otherwise,
// with ZORD, the box whose top rises higher above the ray where the ray is
// lowest inside it (g): for near-horizontal rays over terrain that box holds
// an occluder more often than the nearer one (C2: walk trace 31.6 -> 30.4
// ms); else the nearer box.  Only the order changes: an any-hit answer is
// the OR over the triangles either way.
__device__ __forceinline__ bool first_child(int split, int hint, float ta0, float ta1, float g0,
                                            float g1, bool zord) {
    if (hint >= 0 && split >= 0) return hint <= split;
    return zord ? g0 >= g1 : ta0 <= ta1;
}

// One step of the any-hit traversal: fetch node `node`,
// test both child boxes, test the triangles of hit leaf children at once,
// then descend (with both children hit: into first_child's choice) or pop.  Returns 0 while the query is open, 1 on a hit, 2
// when the stack empties without one.  The stack keeps its first kLdsStack
// entries in LDS at stk[k * BLOCK + lane] and the rest in the private array ovf.
template <bool COUNT, int BLOCK, bool MULTI, bool ZORD = true>
__device__ __forceinline__ int trav_step(const BvhNode* __restrict__ nodes,
                                         const float4* __restrict__ tris, int* stk, int* ovf,
                                         const Ray& r, int& node, int& sp, unsigned& nvisit,
                                         unsigned& ntest, int& err, int& hit_leaf) {
    const float4* np = reinterpret_cast<const float4*>(nodes + (node & kLeafFirstMask));
    const float4 a = np[0], b = np[1], c = np[2], d = np[3];
    // the second half is read only for nodes with inline triangles, and only
    // their leaf children test it: elsewhere its value is never used, so it
    // is left unset rather than cleared on every step
    float4 e4, f4, g4, h4;
    if (node & kWideBit) {  // inline triangle(s): second half of the node
        e4 = np[4];
        f4 = np[5];
        g4 = np[6];
        h4 = np[7];
    }
    if (COUNT) nvisit++;
    // child boxes: a = (lx0, lx1, hx0, hx1), b = (ly0, ly1, hy0, hy1), c = (lz0, lz1, hz0, hz1)
    float tmin0, tmin1, g0, g1;
    bool h0, h1;
    two_boxes(r, a, b, c, h0, h1, tmin0, tmin1, g0, g1);
    const int r0 = __float_as_int(d.x), r1 = __float_as_int(d.y);
    // leaf children: single triangles are inline (slot 0: d.zw e4 f4.xyz,
    // slot 1: f4.w g4 h4); larger leaves (MULTI only) live in the tris array.
    // Each slot has its own test site on the loaded registers: a shared site
    // would need nine selects and a loop per step, and some lane of the wave
    // tests a slot in nearly every step.
    const bool l0 = h0 && r0 < 0, l1 = h1 && r1 < 0;
    if (l0 || l1) {
        if (l0) {
            if (MULTI && r0 != ~(~r0 & kLeafFirstMask)) {
                if (leaf_hit<COUNT>(r, tris, r0, ntest, hit_leaf)) return 1;
            } else {
                if (COUNT) ntest++;
                if (tri_hit(r, d.z, d.w, e4.x, e4.y, e4.z, e4.w, f4.x, f4.y, f4.z)) {
                    hit_leaf = ~r0;  // a single-triangle leaf: its leaf-order index
                    return 1;
                }
            }
        }
        if (l1) {
            if (MULTI && r1 != ~(~r1 & kLeafFirstMask)) {
                if (leaf_hit<COUNT>(r, tris, r1, ntest, hit_leaf)) return 1;
            } else {
                if (COUNT) ntest++;
                if (tri_hit(r, f4.w, g4.x, g4.y, g4.z, g4.w, h4.x, h4.y, h4.z, h4.w)) {
                    hit_leaf = ~r1;
                    return 1;
                }
            }
        }
        h0 = h0 && r0 >= 0;
        h1 = h1 && r1 >= 0;
    }
    if (h0 && h1) {
        const bool first0 = first_child(__float_as_int(d.w), hit_leaf, tmin0, tmin1, g0, g1, ZORD);
        int nearc = first0 ? r0 : r1;
        int farc = first0 ? r1 : r0;
        if (sp >= kStackMax) {
            err |= 1;
            return 1;
        }
        if (sp < kLdsStack)
            stk[sp * BLOCK + lane_now()] = farc;
        else
            ovf[sp - kLdsStack] = farc;
        sp++;
        node = nearc;
    } else if (h0) {
        node = r0;
    } else if (h1) {
        node = r1;
    } else {
        if (sp == 0) return 2;
        sp--;
        node = sp < kLdsStack ? stk[sp * BLOCK + lane_now()] : ovf[sp - kLdsStack];
    }
    return 0;
}

// Binary tree with single-triangle leaves, lattice searches and the walk of
// the near cells beside them: the next node's
// first half (child boxes and references, 64 bytes) is loaded straight into
// LDS by wavefront DMA loads (16 bytes per lane per instruction, lane l's
// bytes at base + 16 l), so a step can request it before its triangle tests
// without holding it in registers.  Each wavefront has a 4 KB node buffer:
// quarter k of lane l's first half at nb[k * 64 + l].
__device__ __forceinline__ void dma_boxes(const BvhNode* __restrict__ nodes, int node,
                                          float4* nb) {
    const float4* np = reinterpret_cast<const float4*>(nodes + (node & kLeafFirstMask));
#pragma unroll
    for (int k = 0; k < 4; k++)
        __builtin_amdgcn_global_load_lds(static_cast<const void*>(np + k),
                                         (__attribute__((address_space(3))) void*)(nb + k * 64),
                                         16, 0, 0);
}

// trav_step for single-triangle leaves with the next fetch issued early.  In
// an any-hit query the triangle tests only decide whether the query ends, so
// the next node (first hit interior child, or the stack top) is known once
// the child boxes are tested: its first half is requested into the node
// buffer before the two inline-triangle test sites run, and the tests overlap
// the fetch.  On entry the node buffer holds (or is loading) node `node`'s
// first half; its second half (inline triangles, same 128-byte line) is read
// into registers at the start of the step and is needed only after the box
// tests and the stack.  The LDS stack holds the first DEP entries.  Same
// return values, stack, hits and counts as trav_step; a fetch issued by a
// step whose query then ends is never read.
template <bool COUNT, int BLOCK, int DEP, bool ZORD = true>
__device__ __forceinline__ int trav_step_pf(const BvhNode* __restrict__ nodes, int* stk,
                                            int* ovf, float4* nb, const Ray& r, int& node,
                                            int& sp, unsigned& nvisit, unsigned& ntest, int& err,
                                            int& hit_leaf) {
    if (COUNT) nvisit++;
    const int lane = lane_now();
    // the node buffer is written by the DMA loads issued before this step: wait
    // for them (the compiler does not order LDS reads after those loads across
    // the loop, so the wait is explicit, and the clobber keeps the reads below it)
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    const float4 a = nb[lane], b = nb[64 + lane], c = nb[128 + lane], d = nb[192 + lane];
    float4 e4, f4, g4, h4;  // left unset (never read) for nodes without inline triangles
    if (node & kWideBit) {
        const float4* np = reinterpret_cast<const float4*>(nodes + (node & kLeafFirstMask));
        e4 = np[4];
        f4 = np[5];
        g4 = np[6];
        h4 = np[7];
    }
    float tmin0, tmin1, g0, g1;
    bool h0, h1;
    two_boxes(r, a, b, c, h0, h1, tmin0, tmin1, g0, g1);
    const int r0 = __float_as_int(d.x), r1 = __float_as_int(d.y);
    const float t0 = d.z, t1 = d.w;                     // slot 0's first two coordinates
    const bool l0 = h0 && r0 < 0, l1 = h1 && r1 < 0;    // hit leaf children
    const bool i0 = h0 && r0 >= 0, i1 = h1 && r1 >= 0;  // hit interior children
    // next node first
    bool empty = false, over = false;
    int next = 0;
    if (i0 && i1) {
        const bool first0 = first_child(__float_as_int(d.w), hit_leaf, tmin0, tmin1, g0, g1, ZORD);
        const int farc = first0 ? r1 : r0;
        next = first0 ? r0 : r1;
        if (sp >= kStackMax) {
            over = true;
        } else {
            if (sp < DEP)
                stk[sp * BLOCK + lane] = farc;
            else
                ovf[sp - DEP] = farc;
            sp++;
        }
    } else if (i0) {
        next = r0;
    } else if (i1) {
        next = r1;
    } else if (sp == 0) {
        empty = true;
    } else {
        sp--;
        if (sp < DEP) {
            next = stk[sp * BLOCK + lane];
        } else {
            next = ovf[sp - DEP];
            // a wait the compiler sees (vmcnt 0): on this rare path the private
            // load is the newest, and without it the merged wait before the
            // request below would cover the inline-triangle loads on every path
            __builtin_amdgcn_s_waitcnt(0xF70);
        }
    }
    // the inline triangles have arrived before the request (a wait the
    // compiler sees: the tests below then wait for nothing, where its own
    // wait for them would also wait for the request); the request is made on
    // every path (the root, a resident line, when the query ends here)
    __builtin_amdgcn_s_waitcnt(0xF70);
    // the buffer's reads of this step are done before the request overwrites it
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    dma_boxes(nodes, (empty || over) ? 0 : next, nb);
    // then this node's inline triangles (slot 0: d.zw e4 f4.xyz, slot 1:
    // f4.w g4 h4), each at its own test site
    if (l0) {
        if (COUNT) ntest++;
        if (tri_hit(r, t0, t1, e4.x, e4.y, e4.z, e4.w, f4.x, f4.y, f4.z)) {
            hit_leaf = ~r0;  // a single-triangle leaf: its leaf-order index
            return 1;
        }
    }
    if (l1) {
        if (COUNT) ntest++;
        if (tri_hit(r, f4.w, g4.x, g4.y, g4.z, g4.w, h4.x, h4.y, h4.z, h4.w)) {
            hit_leaf = ~r1;
            return 1;
        }
    }
    if (over) {
        err |= 1;
        return 1;
    }
    if (empty) return 2;
    node = next;
    return 0;
}

// One step of the any-hit traversal of the four-wide tree: fetch node `node`
// (64 bytes, or 128 with its inline triangles), test the quantized boxes of
// its children, test the triangles of hit leaf children at once, continue
// with the nearest hit interior child and push the others.  Same return
// values and stack as trav_step.
__device__ __forceinline__ float qbyte(unsigned q, int k) {
    return (float)((q >> (8 * k)) & 0xffu);
}

template <bool COUNT, int BLOCK>
__device__ __forceinline__ int trav_step4(const Bvh4Node* __restrict__ nodes,
                                          const float4* __restrict__ tris, int* stk, int* ovf,
                                          const Ray& r, int& node, int& sp, unsigned& nvisit,
                                          unsigned& ntest, int& err, int& hit_leaf) {
    const float4* np = reinterpret_cast<const float4*>(nodes + (node & kLeafFirstMask));
    const float4 a = np[0], b = np[1], c = np[2], d = np[3];
    // the second half is read only for nodes with inline triangles, and only
    // their leaf children test it: elsewhere its value is never used, so it
    // is left unset rather than cleared on every step
    float4 e4, f4, g4, h4;
    if (node & kWideBit) {  // inline triangle(s): second half of the node
        e4 = np[4];
        f4 = np[5];
        g4 = np[6];
        h4 = np[7];
    }
    if (COUNT) nvisit++;
    const unsigned meta = __float_as_uint(a.w);
    const int nch = (int)(meta >> 24);
    // the ray parameter of plane org + q * 2^e is q * (2^e / dir) + (org / dir - o / dir)
    const float ax = __uint_as_float((meta & 0xffu) << 23) * r.ix;
    const float ay = __uint_as_float(((meta >> 8) & 0xffu) << 23) * r.iy;
    const float az = __uint_as_float(((meta >> 16) & 0xffu) << 23) * r.iz;
    const float bx = __builtin_fmaf(a.x, r.ix, -r.oix);
    const float by = __builtin_fmaf(a.y, r.iy, -r.oiy);
    const float bz = __builtin_fmaf(a.z, r.iz, -r.oiz);
    // entry / exit planes by the sign of the direction
    const unsigned qlx = __float_as_uint(b.x), qhx = __float_as_uint(b.y);
    const unsigned qly = __float_as_uint(b.z), qhy = __float_as_uint(b.w);
    const unsigned qlz = __float_as_uint(c.x), qhz = __float_as_uint(c.y);
    const unsigned nx = r.ix >= 0.0f ? qlx : qhx, fx = r.ix >= 0.0f ? qhx : qlx;
    const unsigned ny = r.iy >= 0.0f ? qly : qhy, fy = r.iy >= 0.0f ? qhy : qly;
    const unsigned nz = r.iz >= 0.0f ? qlz : qhz, fz = r.iz >= 0.0f ? qhz : qlz;
    const int r0 = __float_as_int(c.z), r1 = __float_as_int(c.w);
    const int r2 = __float_as_int(d.x), r3 = __float_as_int(d.y);
    const unsigned leafm = (r0 < 0 ? 1u : 0u) | (r1 < 0 ? 2u : 0u) | (r2 < 0 ? 4u : 0u) |
                           (r3 < 0 ? 8u : 0u);
    // hit children, and the nearest hit interior child (entry distance)
    unsigned hit = 0u;
    int best = -1;
    float tb = 3.0e38f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float t0 = fmaxf(fmaxf(__builtin_fmaf(qbyte(nx, k), ax, bx),
                                     __builtin_fmaf(qbyte(ny, k), ay, by)),
                               fmaxf(__builtin_fmaf(qbyte(nz, k), az, bz), -1.0f));
        const float t1 = fminf(fminf(__builtin_fmaf(qbyte(fx, k), ax, bx),
                                     __builtin_fmaf(qbyte(fy, k), ay, by)),
                               fminf(__builtin_fmaf(qbyte(fz, k), az, bz), r.tfar_box));
        if (k < nch && t0 <= t1) {
            hit |= 1u << k;
            if (!((leafm >> k) & 1u) && t0 < tb) {
                tb = t0;
                best = k;
            }
        }
    }
    unsigned lh = hit & leafm;
    if (lh) {
        // hit leaf children: the two inline triangles (slot 0: d.zw e4 f4.xyz,
        // slot 1: f4.w g4 h4), each with a test site of its own, then leaves
        // in the leaf-order array
        int c0 = -1, c1 = -1;
        unsigned rest = 0u;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!((lh >> k) & 1u)) continue;
            const int code = ~(k == 0 ? r0 : (k == 1 ? r1 : (k == 2 ? r2 : r3)));
            if (code & kB4InlineBit) {
                if (code & 1) c1 = code; else c0 = code;
            } else {
                rest |= 1u << k;
            }
        }
        if (c0 >= 0) {
            if (COUNT) ntest++;
            if (tri_hit(r, d.z, d.w, e4.x, e4.y, e4.z, e4.w, f4.x, f4.y, f4.z)) {
                hit_leaf = (c0 >> 1) & kB4FirstMask;
                return 1;
            }
        }
        if (c1 >= 0) {
            if (COUNT) ntest++;
            if (tri_hit(r, f4.w, g4.x, g4.y, g4.z, g4.w, h4.x, h4.y, h4.z, h4.w)) {
                hit_leaf = (c1 >> 1) & kB4FirstMask;
                return 1;
            }
        }
#pragma unroll 1
        while (rest) {
            const int k = __builtin_ctz(rest);
            rest &= rest - 1u;
            const int code = ~(k == 0 ? r0 : (k == 1 ? r1 : (k == 2 ? r2 : r3)));
            const int first = code & kB4FirstMask;
            const int cnt = (code >> kB4CountShift) + 1;
#pragma unroll 1
            for (int j = 0; j < cnt; j++) {
                const float4* t = tris + 3 * (first + j);
                const float4 t0 = t[0], t1 = t[1], t2 = t[2];
                if (COUNT) ntest++;
                if (tri_hit(r, t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w, t2.x)) {
                    hit_leaf = first + j;
                    return 1;
                }
            }
        }
    }
    const unsigned ih = hit & ~leafm;
    if (!ih) {
        if (sp == 0) return 2;
        sp--;
        node = sp < kLdsStack ? stk[sp * BLOCK + lane_now()] : ovf[sp - kLdsStack];
        return 0;
    }
    // nearest hit interior child next, the others onto the stack
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (!((ih >> k) & 1u) || k == best) continue;
        if (sp >= kStackMax) {
            err |= 1;
            return 1;
        }
        const int rk = k == 0 ? r0 : (k == 1 ? r1 : (k == 2 ? r2 : r3));
        if (sp < kLdsStack)
            stk[sp * BLOCK + lane_now()] = rk;
        else
            ovf[sp - kLdsStack] = rk;
        sp++;
    }
    node = best == 0 ? r0 : (best == 1 ? r1 : (best == 2 ? r2 : r3));
    return 0;
}

// Whole any-hit query, first against the triangle (leaf-order index hint,
// -1: none) that occluded this lane's last hit; hint is set to the
// occluding triangle of a hit.
template <bool COUNT, int BLOCK>
__device__ __forceinline__ bool occluded_hint(const BvhNode* __restrict__ nodes,
                                              const float4* __restrict__ tris, int* stk,
                                              const Ray& r, int& hint, unsigned& nvisit,
                                              unsigned& ntest, int& err) {
    if (hint >= 0) {
        const float4* t = tris + 3 * (size_t)hint;
        const float4 t0 = t[0], t1 = t[1], t2 = t[2];
        if (COUNT) ntest++;
        if (tri_hit(r, t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w, t2.x)) return true;
    }
    int ovf[kStackMax - kLdsStack];
    int node = kWideBit;  // the root is read in full (128 bytes)
    int sp = 0;
    for (int guard = 0;; guard++) {
        if (guard > (1 << 28)) {
            err |= 4;
            return true;
        }
        int hit_leaf = 0;
        int s = trav_step<COUNT, BLOCK, true>(nodes, tris, stk, ovf, r, node, sp, nvisit, ntest,
                                              err, hit_leaf);
        if (s) {
            if (s == 1) hint = hit_leaf;
            return s == 1;
        }
    }
}

// Whole any-hit query.
template <bool COUNT, int BLOCK, bool ZORD = true>
__device__ __forceinline__ bool occluded(const BvhNode* __restrict__ nodes,
                                         const float4* __restrict__ tris, int* stk, const Ray& r,
                                         unsigned& nvisit, unsigned& ntest, int& err) {
    int ovf[kStackMax - kLdsStack];
    int node = kWideBit;  // the root is read in full (128 bytes)
    int sp = 0;
    for (int guard = 0;; guard++) {
        // a BVH of n triangles has n - 1 interior nodes; more fetches than
        // that in one query can only come from a malformed tree
        if (guard > (1 << 28)) {
            err |= 4;
            return true;
        }
        int hit_leaf = 0;
        int s = trav_step<COUNT, BLOCK, true, ZORD>(nodes, tris, stk, ovf, r, node, sp, nvisit, ntest,
                                              err, hit_leaf);
        if (s) return s == 1;
    }
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int MODE, bool COUNT, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace(const BvhNode* __restrict__ nodes,
                                                 const float4* __restrict__ tris, int num_tri,
                                                 const CellFrame* __restrict__ cells,
                                                 const int* __restrict__ order, WalkParams P,
                                                 int nblocks, double* __restrict__ horizon,
                                                 double* __restrict__ skyview,
                                                 TraceCounters* __restrict__ cnt) {
    __shared__ int stk_all[kLdsStack * BLOCK];
    // XCD-aware remap: gridDim.x is a multiple of 8; XCD x gets the
    // contiguous block range [x * per, (x + 1) * per).
    const int per = gridDim.x >> 3;
    const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int idx = lb * BLOCK + threadIdx.x;
    unsigned rays = 0, nvisit = 0, ntest = 0;
    int err = 0;
    if (lb < nblocks && idx < P.ncell) {
        int* stk = stk_all + (__builtin_amdgcn_readfirstlane(threadIdx.x) & ~63);
        const int cell = order ? order[idx] : idx;
        const CellFrame f = cells[cell];
        const D3 n = {f.n[0], f.n[1], f.n[2]};
        const D3 north = {f.north[0], f.north[1], f.north[2]};
        const float o0 = f.org[0], o1 = f.org[1], o2 = f.org[2];
        const double a = P.hori_acc;
        const int L = P.num_hori * P.refine_factor;
        const bool have = num_tri > 0;

        auto cast = [&](const D3& d) -> bool {
            rays++;
            if (!have) return false;
            Ray r = make_ray(o0, o1, o2, (float)d.x, (float)d.y, (float)d.z, P.tfar,
                             P.tfar_box);
            return occluded<COUNT, BLOCK, zorder(MODE)>(nodes, tris, stk, r, nvisit, ntest, err);
        };

        D3 dir = rodrigues(north, n, P.shift_sin, P.shift_cos);
        double e = 0.0;
        double sector = 0.0, svf = 0.0;
        int ksec = 0, jsec = 0;
        for (int j = 0; j < L; j++) {
            const D3 ax = unit_cross(dir, n);
            double h;
            if (MODE == 0) {
                // reference walk (:521-546)
                bool hit = cast(dir);
                if (hit) {
                    while (hit) {
                        if (e > kPi) {  // a ray past the zenith still hits: not a heightfield
                            err |= 2;
                            break;
                        }
                        e += 2.0 * a;
                        dir = rodrigues(dir, ax, P.elev_sin_2ha, P.elev_cos_2ha);
                        hit = cast(dir);
                    }
                    h = e - a;
                } else {
                    while (!hit && e > P.elev_thresh) {
                        e -= 2.0 * a;
                        dir = rodrigues(dir, ax, -P.elev_sin_2ha, P.elev_cos_2ha);
                        hit = cast(dir);
                    }
                    h = e + a;
                }
            } else {
                // clip-aware climb over the lattice from the horizontal direction
                int k = 0;
                bool hit = cast(lattice_dir(dir, n, P.elev_cs, 0));
                if (!hit) {
                    h = 0.0;
                } else {
                    int top = 0;
                    while (true) {
                        if (k >= kMaxLattice) {  // straight up still hits: not a heightfield
                            err |= 2;
                            top = k;
                            break;
                        }
                        k++;
                        if (!cast(lattice_dir(dir, n, P.elev_cs, k))) {
                            top = k - 1;
                            break;
                        }
                    }
                    h = (double)(2 * top + 1) * a;
                }
            }
            if (h < 0.0) h = 0.0;  // (:737-741)
            sector += h;
            if (++ksec == P.refine_factor) {  // sector mean in degrees (:744-751)
                horizon[(size_t)jsec * P.ncell + cell] = ((sector / kPi) * 180.0) / (double)P.refine_factor;
                sector = 0.0;
                ksec = 0;
                jsec++;
            }
            double s = sin_horizon(P, h);  // SVF (:333-358)
            if (P.svf_type == 0)
                svf += (1.0 - s);
            else if (P.svf_type == 1)
                svf += (1.0 - (s * s));
            else
                svf += (1.0 - (s * s * s));
            dir = rodrigues(dir, n, -P.azim_sin, P.azim_cos);  // next azimuth (:550-551)
        }
        skyview[cell] = svf / (double)L;
    }
    unsigned long long wr = wave_sum(rays);
    unsigned long long wn = COUNT ? wave_sum(nvisit) : 0ull;
    unsigned long long wt = COUNT ? wave_sum(ntest) : 0ull;
    int werr = __any(err != 0);
    if ((threadIdx.x & 63) == 0) {
        if (wr) atomicAdd(&cnt->rays, wr);
        if (COUNT) {
            atomicAdd(&cnt->nodes, wn);
            atomicAdd(&cnt->tris, wt);
        }
        if (werr) atomicOr(&cnt->error, 1ull);
    }
}

// ---------------------------------------------------------------------------
// Persistent kernel
//
// Every lane runs the elevation search of its current cell as a small state
// machine and keeps one occlusion query in flight.  The wave alternates
// between a traversal phase (one BVH step per iteration for every lane with
// a live query) and a refill phase entered once fewer than `refill` lanes
// still have a live query: lanes whose query ended consume its answer,
// advance their search (next elevation step, next azimuth, or the end of the
// cell with its sector means and SVF written out) and issue their next query;
// lanes whose cell is finished take the next cell from the wave's chunk of
// the Morton-ordered cell list (with ITEMS: the next (cell, azimuth) item).
// A wave takes as many cells as its lanes ask for (at most 64) from one
// atomic counter per XCD, each over its own contiguous eighth of that list;
// a wave whose eighth is used up moves on to the next one.  This keeps SIMD lanes
// busy although searches need 1 to ~200 queries per azimuth, and keeps each
// XCD's L2 working set spatially compact.
// ---------------------------------------------------------------------------

constexpr int kChunk = 64;
constexpr int kItemCells = 64;  // cells per block of work items (ITEMS)
enum { P_AZSTART = 0, P_FIRST = 1, P_UP = 2, P_DOWN = 3 };

__global__ void k_clear_state(TraceState* ts) {
    int* w = reinterpret_cast<int*>(ts);
    for (int i = threadIdx.x; i < (int)(sizeof(TraceState) / 4); i += blockDim.x) w[i] = 0;
}

// Work queues only (the counters of the launch before are kept).
__global__ void k_clear_queue(TraceState* ts) {
    for (int i = threadIdx.x; i < 8 * kQueueStride; i += blockDim.x) ts->queue[i] = 0;
}

// Per-lane FP64 search state kept in global memory between queries (it is
// needed only when a query ends), so that the traversal loop runs with the
// registers of the traversal alone.  Field f of lane g is spill[f * lanes + g].
// Fields S_I0.. hold two ints each: {cell, phase}, {azimuth j, lattice index},
// {sector step, sector index}, {batch domain, unused}.
enum {
    S_DIR = 0, S_AX = 3, S_E = 6, S_SECTOR = 7, S_SVF = 8,
    S_I0 = 9, S_I1 = 10, S_I2 = 11, S_I3 = 12, S_FIELDS = 13
};
static_assert(S_FIELDS == S_FIELDS_HOST, "spill layout");

// TREE: 0 binary tree with single-triangle leaves, 1 binary tree with larger
// leaves, 2 four-wide tree (Bvh4Node records behind `nodes`).
// BATCH: the cells of several domains (same search parameters) form one
// index space; doms[d] covers global indices [doms[d].start, doms[d + 1].start)
// and each lane follows its own cell's domain (scene, outputs).
// ITEMS (search modes 1 and 2, one domain): the work items are (cell,
// azimuth) pairs instead of cells, since these searches start every azimuth
// afresh; item results go to hitem and k_items_finish forms the sector means
// and the SVF.
// P.range (device {first, count}; NULL: all cells): the positions of the
// trace order this launch works on (its cells, or the items of its cells).
template <int MODE, bool COUNT, int BLOCK, int TREE, bool BATCH, bool ITEMS = false,
          bool PFW = false>
__global__ __launch_bounds__(BLOCK, 4) void k_trace_pt(const BvhNode* __restrict__ nodes,
                                                    const float4* __restrict__ tris, int num_tri,
                                                    const CellFrame* __restrict__ cells,
                                                    const int* __restrict__ order, WalkParams P,
                                                    int refill, int chunk,
                                                    double* __restrict__ horizon,
                                                    double* __restrict__ skyview,
                                                    TraceState* __restrict__ ts,
                                                    double* __restrict__ spill,
                                                    const BatchDomain* __restrict__ doms, int ndom,
                                                    int total, double* __restrict__ hitem) {
    static_assert(!ITEMS || (MODE != 0 && !BATCH), "items: lattice searches of one domain");
    // Lattice searches over the binary tree with single-triangle leaves step
    // with the next node's boxes requested early (trav_step_pf), and so does
    // the walk with PFW (a part of the cells beside another launch); a walk
    // over all cells keeps the plain step, which is faster there (it fills
    // the GPU's memory path alone, where the early request only adds work).
    constexpr bool PF = TREE == 0 && (MODE != 0 || PFW);
    // LDS: the stack's first kLdsDep entries per lane, and with PF a 64-byte
    // node buffer per lane (a 16-entry stack then keeps the LDS at 128 bytes
    // per lane; the deepest stack of a C2 query is 13 entries, deeper ones
    // continue in ovf)
    constexpr int kLdsDep = PF ? 16 : kLdsStack;
    __shared__ int stk_all[kLdsDep * BLOCK];
    __shared__ float4 nbuf_all[PF ? 4 * BLOCK : 1];
    // the wave's part of the stack: stk[k * BLOCK + lane] (the base is uniform)
    const int wbase = __builtin_amdgcn_readfirstlane(threadIdx.x) & ~63;
    int* stk = stk_all + wbase;
    float4* nbuf = nbuf_all + (PF ? 4 * wbase : 0);
    int ovf[kStackMax - kLdsDep];
    const int lane = threadIdx.x & 63;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    const int L = P.num_hori * P.refine_factor;
    // size of the work index space (the launcher keeps ncell * L below 2^31)
    // positions [ibase, ibase + npos) of the trace order (P.range; all cells
    // without it): items of those cells, or those cells
    const int ibase = (!BATCH && P.range) ? P.range[0] : 0;
    const int npos = (!BATCH && P.range) ? P.range[1] : P.ncell;
    const int N = BATCH ? total : (ITEMS ? npos * L : npos);
    const double a = P.hori_acc;
    int* queue = ts->queue;
    const size_t lanes = (size_t)gridDim.x * BLOCK;
    double* sv = spill + (size_t)blockIdx.x * BLOCK + threadIdx.x;
#define SV(f) sv[(size_t)(f) * lanes]
#define SI(f) (*reinterpret_cast<int2*>(&sv[(size_t)(f) * lanes]))

    // wave-uniform queue state
    int cur = blockIdx.x & 7;  // workgroups are dealt round-robin over the 8 XCDs
    int tries = 0;
    int qb = 0, qe = 0;
    bool exhausted = false;

    // lane state kept in registers across the traversal (the rest of the
    // search state lives in the spill area)
    bool has_cell = false, live = false, last_hit = false;
    // a query issued since the lane's last traversal step (binary tree with
    // single-triangle leaves: its first node is not in the node buffer yet)
    bool fresh = false;
    int hint = -1;  // leaf-order index of the triangle that occluded the cell's last hit
    Ray ray = make_ray(0.f, 0.f, 0.f, 1.f, 0.f, 0.f, P.tfar, P.tfar_box);
    int node = kWideBit, sp = 0, steps = 0;
    unsigned rays = 0, nvisit = 0, ntest = 0;
    int err = 0;
    unsigned long long cyc_refill = 0, cyc_trav = 0, iters = 0;  // COUNT only (wave-uniform)
    // the scene of the lane's domain (batch) or of the single domain
    const BvhNode* lnodes = nodes;
    const float4* ltris = tris;
    bool have = num_tri > 0;

    while (true) {
        // ---------------- refill ----------------
        unsigned long long c0 = 0;
        if (COUNT) c0 = __builtin_amdgcn_s_memtime();
        for (int pass = 0; pass < 4; pass++) {
            const bool want = !live && !has_cell;
            const unsigned long long need = __ballot(want);
            if (need && !exhausted) {
                if (qb >= qe) {
                    while (tries < 8) {
                        const int rlo = (int)(((long long)cur * N) >> 3);
                        const int rhi = (int)(((long long)(cur + 1) * N) >> 3);
                        // only as many as the wave's lanes ask for now (at most
                        // `chunk`): a wave holds no spare cells when the queues
                        // run dry, which keeps the end of the launch short
                        const int take = min(chunk, __popcll(need));
                        int old = 0;
                        if (lane == 0) old = atomicAdd(&queue[cur * kQueueStride], take);
                        old = __shfl(old, 0, 64);
                        if (rlo + old < rhi) {
                            qb = rlo + old;
                            qe = min(qb + take, rhi);
                            break;
                        }
                        cur = (cur + 1) & 7;
                        tries++;
                    }
                    if (tries >= 8) exhausted = true;
                }
                if (!exhausted) {
                    const int avail = qe - qb;
                    const int rank = __popcll(need & lt_mask);
                    if (want && rank < avail) {
                        const int g = qb + rank;
                        int cell = 0, dom = 0, item_j = 0, item_p = 0;
                        const CellFrame* lcells = cells;
                        if (BATCH) {
                            int lo = 0, hi = ndom - 1;  // last domain with start <= g
                            while (lo < hi) {
                                const int mid = (lo + hi + 1) >> 1;
                                if (doms[mid].start <= g) lo = mid; else hi = mid - 1;
                            }
                            dom = lo;
                            const BatchDomain& bd = doms[dom];
                            const int l = g - bd.start;
                            cell = bd.order ? bd.order[l] : l;
                            lnodes = bd.nodes;
                            ltris = bd.tris;
                            have = bd.num_tri > 0;
                            lcells = bd.cells;
                        } else if (ITEMS) {
                            // blocks of kItemCells cells x L azimuths, azimuth-major
                            // inside a block: a chunk of 64 items is one azimuth of
                            // 64 neighbouring cells
                            const int blk = g / (kItemCells * L);
                            const int rem = g - blk * (kItemCells * L);
                            const int nb = min(kItemCells, npos - blk * kItemCells);
                            item_j = rem / nb;
                            item_p = ibase + blk * kItemCells + (rem - item_j * nb);
                            cell = order ? order[item_p] : item_p;
                        } else {
                            cell = order ? order[ibase + g] : ibase + g;
                        }
                        has_cell = true;
                        hint = -1;
                        const CellFrame* f = lcells + cell;
                        const D3 n = {f->n[0], f->n[1], f->n[2]};
                        const D3 north = {f->north[0], f->north[1], f->north[2]};
                        D3 dir = rodrigues(north, n, P.shift_sin, P.shift_cos);
                        // the azimuth's direction by the same sequence of rotations
                        // as the cell-by-cell search (:550-551)
                        if (ITEMS)
                            for (int t = 0; t < item_j; t++)
                                dir = rodrigues(dir, n, -P.azim_sin, P.azim_cos);
                        SV(S_DIR) = dir.x;
                        SV(S_DIR + 1) = dir.y;
                        SV(S_DIR + 2) = dir.z;
                        SV(S_E) = 0.0;
                        SV(S_SECTOR) = 0.0;
                        SV(S_SVF) = 0.0;
                        SI(S_I0) = make_int2(cell, P_AZSTART);
                        SI(S_I1) = make_int2(ITEMS ? item_j : 0, 0);
                        SI(S_I2) = make_int2(0, ITEMS ? item_p : 0);
                        if (BATCH) SI(S_I3) = make_int2(dom, 0);
                    }
                    qb += min(__popcll(need), avail);
                }
            }
            if (has_cell && !live) {
                // consume the last answer and issue the next query of this cell
                const int2 i0 = SI(S_I0), i1 = SI(S_I1), i2 = SI(S_I2);
                const int cell = i0.x;
                int phase = i0.y, j = i1.x, kk = i1.y, ksec = i2.x, jsec = i2.y;
                const CellFrame* lcells = cells;
                double* lhorizon = horizon;
                double* lskyview = skyview;
                int lncell = P.ncell;
                if (BATCH) {
                    const BatchDomain& bd = doms[SI(S_I3).x];
                    lcells = bd.cells;
                    lhorizon = bd.horizon;
                    lskyview = bd.skyview;
                    lncell = bd.ncell;
                }
                const CellFrame* fr = lcells + cell;
                const D3 n = {fr->n[0], fr->n[1], fr->n[2]};
                D3 dir = {SV(S_DIR), SV(S_DIR + 1), SV(S_DIR + 2)};
                D3 ax = {0.0, 0.0, 0.0};
                double e = 0.0;
                if (MODE == 0) {
                    if (phase != P_AZSTART) ax = {SV(S_AX), SV(S_AX + 1), SV(S_AX + 2)};
                    e = SV(S_E);
                }
                // the per-azimuth fields (sector sums, azimuth counters, axis) are
                // read when an azimuth ends and written back only when one ended
                // or started here; a step of a run rewrites dir, e and the phase
                double sector = 0.0, svf = 0.0;
                bool az_read = false, az_dirty = false;
                D3 q = dir;  // direction of the query to issue
                bool issue = false;
                // the last occluder, fetched now so that the load overlaps the search
                // state update below
                float4 ht0 = make_float4(0.f, 0.f, 0.f, 0.f), ht1 = ht0, ht2 = ht0;
                if (hint >= 0) {
                    const float4* t = ltris + 3 * (size_t)hint;
                    ht0 = t[0];
                    ht1 = t[1];
                    ht2 = t[2];
                }
                // a query that the cell's last occluding triangle already blocks is
                // answered here, without a traversal (the answer is the OR over all
                // triangles either way); a few in a row, then the tree decides
                for (int hint_round = 0;; hint_round++) {
                issue = false;
                while (true) {
                    double h = 0.0;
                    if (phase == P_AZSTART) {
                        {
                            az_dirty = true;
                            phase = P_FIRST;
                            if (MODE == 0) {
                                ax = unit_cross(dir, n);
                                q = dir;
                            } else {
                                kk = MODE == 1
                                         ? 0
                                         : start_index(P.pred[(size_t)cell * P.pred_stride + j]);
                                q = lattice_dir(dir, n, P.elev_cs, kk);
                            }
                            issue = true;
                            break;
                        }
                    } else if (MODE == 0) {  // reference walk (:521-546)
                        if (phase == P_FIRST) {
                            if (last_hit) {
                                phase = P_UP;
                                e += 2.0 * a;
                                dir = rodrigues(dir, ax, P.elev_sin_2ha, P.elev_cos_2ha);
                                q = dir;
                                issue = true;
                                break;
                            }
                            if (e > P.elev_thresh) {
                                phase = P_DOWN;
                                e -= 2.0 * a;
                                dir = rodrigues(dir, ax, -P.elev_sin_2ha, P.elev_cos_2ha);
                                q = dir;
                                issue = true;
                                break;
                            }
                            h = e + a;
                        } else if (phase == P_UP) {
                            if (last_hit) {
                                if (!(e > kPi)) {
                                    e += 2.0 * a;
                                    dir = rodrigues(dir, ax, P.elev_sin_2ha, P.elev_cos_2ha);
                                        q = dir;
                                    issue = true;
                                    break;
                                }
                                err |= 2;  // a ray past the zenith still hits: not a heightfield
                            }
                            h = e - a;
                        } else {
                            if (!last_hit && e > P.elev_thresh) {
                                e -= 2.0 * a;
                                dir = rodrigues(dir, ax, -P.elev_sin_2ha, P.elev_cos_2ha);
                                q = dir;
                                issue = true;
                                break;
                            }
                            h = e + a;
                        }
                    } else {
                        // lattice search: climb while the query at index kk hits, descend
                        // (learned start only) while it misses; the result is the top hit
                        // index, or 0 deg when index 0 misses
                        int top = -1;
                        if (last_hit) {
                            if (phase == P_DOWN) {
                                top = kk;
                            } else if (kk >= kMaxLattice) {
                                err |= 2;  // straight up still hits: not a heightfield
                                top = kk;
                            } else {
                                phase = P_UP;
                                kk++;
                                q = lattice_dir(dir, n, P.elev_cs, kk);
                                issue = true;
                                break;
                            }
                        } else {
                            if (phase == P_UP) {
                                top = kk - 1;
                            } else if (kk > 0) {
                                phase = P_DOWN;
                                kk--;
                                q = lattice_dir(dir, n, P.elev_cs, kk);
                                issue = true;
                                break;
                            }
                        }
                        h = top >= 0 ? (double)(2 * top + 1) * a : 0.0;
                    }
                    // azimuth finished: clip, sector mean, SVF (:736-754)
                    if (h < 0.0) h = 0.0;
                    if (ITEMS) {  // item p = jsec of the spill: its result, then the next item
                        hitem[(size_t)j * P.ncell + jsec] = h;
                        has_cell = false;
                        break;
                    }
                    if (!az_read) {
                        sector = SV(S_SECTOR);
                        svf = SV(S_SVF);
                        az_read = true;
                    }
                    az_dirty = true;
                    sector += h;
                    if (++ksec == P.refine_factor) {
                        lhorizon[(size_t)jsec * lncell + cell] =
                            ((sector / kPi) * 180.0) / (double)P.refine_factor;
                        sector = 0.0;
                        ksec = 0;
                        jsec++;
                    }
                    const double sn = sin_horizon(P, h);
                    if (P.svf_type == 0)
                        svf += (1.0 - sn);
                    else if (P.svf_type == 1)
                        svf += (1.0 - (sn * sn));
                    else
                        svf += (1.0 - (sn * sn * sn));
                    dir = rodrigues(dir, n, -P.azim_sin, P.azim_cos);  // (:550-551)
                    if (++j == L) {
                        lskyview[cell] = svf / (double)L;
                        has_cell = false;
                        break;
                    }
                    phase = P_AZSTART;
                }
                if (!issue) break;
                rays++;
                ray = make_ray(fr->org[0], fr->org[1], fr->org[2], (float)q.x, (float)q.y,
                               (float)q.z, P.tfar, P.tfar_box);
                if (hint >= 0 && hint_round < kHintRounds) {
                    if (COUNT) ntest++;
                    if (tri_hit(ray, ht0.x, ht0.y, ht0.z, ht0.w, ht1.x, ht1.y, ht1.z, ht1.w,
                                ht2.x)) {
                        last_hit = true;
                        continue;
                    }
                }
                break;
                }
                if (issue) {
                    SI(S_I0) = make_int2(cell, phase);
                    node = kWideBit;  // the root is read in full
                    if constexpr (TREE != 2 && !BATCH) {
                        if (P.start) node = P.start[cell];  // below the levels out of reach
                    }
                    sp = 0;
                    steps = 0;
                    live = true;
                    fresh = true;
                    if (MODE == 0) {
                        SV(S_DIR) = dir.x;
                        SV(S_DIR + 1) = dir.y;
                        SV(S_DIR + 2) = dir.z;
                        SV(S_E) = e;
                    } else {
                        SI(S_I1) = make_int2(j, kk);  // the lattice index moves every query
                    }
                    if (az_dirty) {
                        if (MODE == 0) {
                            SI(S_I1) = make_int2(j, kk);
                            SV(S_AX) = ax.x;
                            SV(S_AX + 1) = ax.y;
                            SV(S_AX + 2) = ax.z;
                        } else {
                            SV(S_DIR) = dir.x;
                            SV(S_DIR + 1) = dir.y;
                            SV(S_DIR + 2) = dir.z;
                        }
                        SI(S_I2) = make_int2(ksec, jsec);
                        if (az_read) {
                            SV(S_SECTOR) = sector;
                            SV(S_SVF) = svf;
                        }
                    }
                }
            }
            if (!__any(!live && !has_cell && !exhausted)) break;
        }
        // keep the search state in memory, not in registers, across the traversal
        __asm__ __volatile__("" ::: "memory");
        if (!__any(live)) break;
        // ---------------- traversal ----------------
        unsigned long long c1 = 0;
        if (COUNT) {
            c1 = __builtin_amdgcn_s_memtime();
            cyc_refill += c1 - c0;
        }
        // once the queues are empty, refill as soon as fewer lanes are live
        // than hold a cell (scaled like the refill threshold), so that a lane
        // that still has queries to make does not wait for the wave's slowest
        int thresh = refill;
        if (exhausted) thresh = max(1, (__popcll(__ballot(has_cell)) * refill + 63) >> 6);
        // PF: each step requests the next node's boxes before its triangle
        // tests, so a lane whose query goes on from the last phase has them in
        // its node buffer; a new query's root is requested here
        if constexpr (PF) {
            if (live && have && fresh) dma_boxes(lnodes, node, nbuf);
            fresh = false;
        }
        while (true) {
            if (COUNT) iters++;
            if (live) {
                int s = 2;
                if (have) {
                    if constexpr (PF)
                        s = trav_step_pf<COUNT, BLOCK, kLdsDep, zorder(MODE)>(
                            lnodes, stk, ovf, nbuf, ray, node, sp, nvisit, ntest, err, hint);
                    else if constexpr (TREE == 2)
                        s = trav_step4<COUNT, BLOCK>(reinterpret_cast<const Bvh4Node*>(lnodes),
                                                     ltris, stk, ovf, ray, node, sp, nvisit,
                                                     ntest, err, hint);
                    else
                        s = trav_step<COUNT, BLOCK, TREE == 1, zorder(MODE)>(
                            lnodes, ltris, stk, ovf, ray, node, sp, nvisit, ntest, err, hint);
                    // a BVH of n triangles has n - 1 interior nodes; more steps than
                    // that in one query can only come from a malformed tree
                    if (s == 0 && ++steps > (1 << 28)) {
                        err |= 4;
                        s = 1;
                    }
                }
                if (s) {
                    live = false;
                    last_hit = (s == 1);
                }
            }
            if (__popcll(__ballot(live)) < thresh) break;
        }
        if (COUNT) cyc_trav += __builtin_amdgcn_s_memtime() - c1;
    }
#undef SI
#undef SV
    const unsigned long long wr = wave_sum(rays);
    const unsigned long long wn = COUNT ? wave_sum(nvisit) : 0ull;
    const unsigned long long wt = COUNT ? wave_sum(ntest) : 0ull;
    const int werr = __any(err != 0);
    if (lane == 0) {
        if (wr) atomicAdd(&ts->cnt.rays, wr);
        if (COUNT) {
            atomicAdd(&ts->cnt.nodes, wn);
            atomicAdd(&ts->cnt.tris, wt);
            atomicAdd(&ts->cnt.cyc_refill, cyc_refill);
            atomicAdd(&ts->cnt.cyc_trav, cyc_trav);
            atomicAdd(&ts->cnt.trav_iters, iters);
        }
        if (werr) atomicOr(&ts->cnt.error, 1ull);
    }
}

// ---------------------------------------------------------------------------
// Cooperative walk: the reference walk of listed cells, one cell per team of
// TEAM lanes.  Within an azimuth the walk's queries lie on a run of fixed
// rotations about the azimuth's axis, so their directions do not depend on
// the answers: a lane reaches step m of the run by the walk's own m
// rotations and elevation updates from the run's current state (the same
// FP64 operations in the same order, so the same bits).  Round 0 casts the
// azimuth's first query (lane 0) with up-steps (lanes 1 .. U) and down-steps
// (lanes U + 1 .. TEAM - 1) beside it; later rounds cast the next TEAM steps
// of the run the first answer chose.  The walk then takes the first answer
// that ends its run (:521-546), or the zenith / -85 deg limit, exactly as
// one lane would, so a run of r steps costs about r / TEAM rounds instead of
// r dependent queries.
// ---------------------------------------------------------------------------
template <int TEAM, bool COUNT, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_walk_team(
    const BvhNode* __restrict__ nodes, const float4* __restrict__ tris, int num_tri,
    const CellFrame* __restrict__ cells, const int* __restrict__ order,
    WalkParams P, double* __restrict__ horizon, double* __restrict__ skyview,
    TraceState* __restrict__ ts) {
    static_assert(TEAM >= 2 && TEAM <= 64 && (TEAM & (TEAM - 1)) == 0, "team size");
    __shared__ int stk_all[kLdsStack * BLOCK];
    int* stk = stk_all + (__builtin_amdgcn_readfirstlane(threadIdx.x) & ~63);
    const int lane = threadIdx.x & 63;
    const int t = lane & (TEAM - 1);
    const int tb = lane - t;
    const unsigned long long full = TEAM == 64 ? ~0ull : ((1ull << TEAM) - 1ull);
    unsigned long long rays = 0, spec = 0;
    unsigned nvisit = 0, ntest = 0;
    int err = 0;
    // cells from one counter (queue[0]) over the positions of P.range;
    // every lane of a team takes the same cell
    const int count = P.range[1];
    while (true) {
        int idx = 0;
        if (t == 0) idx = atomicAdd(&ts->queue[0], 1);
        idx = __shfl(idx, tb, 64);
        if (idx >= count) break;
        const int p = P.range[0] + idx;
        int hint = -1;  // this lane's last occluding triangle in this cell
        const int cell = order ? order[p] : p;
        const CellFrame& f = cells[cell];
        const D3 n = {f.n[0], f.n[1], f.n[2]};
        const D3 north = {f.north[0], f.north[1], f.north[2]};
        const float o0 = f.org[0], o1 = f.org[1], o2 = f.org[2];
        const double a = P.hori_acc;
        const int L = P.num_hori * P.refine_factor;
        const bool have = num_tri > 0;
        D3 dir = rodrigues(north, n, P.shift_sin, P.shift_cos);
        double e = 0.0;
        double sector = 0.0, svf = 0.0;
        int ksec = 0, jsec = 0;
        for (int j = 0; j < L; j++) {
            const D3 ax = unit_cross(dir, n);
            double h;
            bool done = false;
            // round 0 casts the azimuth's first query with U up-steps and D
            // down-steps beside it
            constexpr int U = TEAM / 2 - 1, D = TEAM - 1 - U;
            int sgn = 0;  // the run's direction once known (+1 up, -1 down)
            for (int round = 0; !done; round++) {
                // this lane's step: round 0 = start / up / down lanes
                int m, ls;
                if (round == 0) {
                    if (t == 0) {
                        m = 0;
                        ls = 0;
                    } else if (t <= U) {
                        m = t;
                        ls = 1;
                    } else {
                        m = t - U;
                        ls = -1;
                    }
                } else {
                    m = t + 1;
                    ls = sgn;
                }
                D3 d = dir;
                double ee = e;
                bool valid = true;
                for (int q = 0; q < m; q++) {
                    if (ls > 0) {
                        if (ee > kPi) {  // the walk stops past the zenith
                            valid = false;
                            break;
                        }
                        ee += 2.0 * a;
                        d = rodrigues(d, ax, P.elev_sin_2ha, P.elev_cos_2ha);
                    } else {
                        if (!(ee > P.elev_thresh)) {  // the -85 deg floor
                            valid = false;
                            break;
                        }
                        ee -= 2.0 * a;
                        d = rodrigues(d, ax, -P.elev_sin_2ha, P.elev_cos_2ha);
                    }
                }
                bool hit = false;
                if (valid && have) {
                    const Ray r = make_ray(o0, o1, o2, (float)d.x, (float)d.y, (float)d.z, P.tfar,
                                           P.tfar_box);
                    hit = occluded_hint<COUNT, BLOCK>(nodes, tris, stk, r, hint, nvisit, ntest,
                                                      err);
                }
                const unsigned long long hb = (__ballot(hit) >> tb) & full;
                const unsigned long long vb = (__ballot(valid) >> tb) & full;
                const int cast = __popcll(vb);
                // steps of the run this round, bit r-1 = step r
                unsigned long long vm, am;
                int R;
                if (round == 0) {
                    sgn = (hb & 1ull) ? 1 : -1;
                    if (sgn > 0) {
                        R = U;
                        vm = (vb >> 1) & ((1ull << U) - 1ull);  // U <= 63
                        am = ~(hb >> 1) & vm;  // misses end an up run
                    } else {
                        R = D;
                        vm = D ? ((vb >> (U + 1)) & ((1ull << D) - 1ull)) : 0ull;
                        am = D ? ((hb >> (U + 1)) & vm) : 0ull;  // hits end a down run
                    }
                } else {
                    R = TEAM;
                    vm = vb;
                    am = sgn > 0 ? (~hb & vb) : (hb & vb);
                }
                const unsigned long long rmask = R == 64 ? ~0ull : ((1ull << R) - 1ull);
                const unsigned long long inval = ~vm & rmask;
                const int fi = inval ? __builtin_ctzll(inval) : R;  // first invalid step - 1
                const int fa = am ? __builtin_ctzll(am) : R;        // first ending answer - 1
                int steps, used;
                if (fa < fi) {
                    steps = fa + 1;
                    used = fa + 1;
                    done = true;
                } else if (fi < R) {
                    steps = fi;
                    used = fi;
                    done = true;
                    if (sgn > 0) err |= 2;  // a ray past the zenith still hits
                } else {
                    steps = R;
                    used = R;
                }
                if (round == 0) used += 1;  // the azimuth's first query
                if (t == 0) {
                    rays += (unsigned long long)used;
                    spec += (unsigned long long)(cast - used);
                }
                for (int q = 0; q < steps; q++) {  // the run's state after `steps` steps
                    if (sgn > 0) {
                        e += 2.0 * a;
                        dir = rodrigues(dir, ax, P.elev_sin_2ha, P.elev_cos_2ha);
                    } else {
                        e -= 2.0 * a;
                        dir = rodrigues(dir, ax, -P.elev_sin_2ha, P.elev_cos_2ha);
                    }
                }
                h = sgn > 0 ? e - a : e + a;
            }
            if (h < 0.0) h = 0.0;  // (:737-741)
            sector += h;
            if (++ksec == P.refine_factor) {  // (:744-751)
                if (t == 0)
                    horizon[(size_t)jsec * P.ncell + cell] =
                        ((sector / kPi) * 180.0) / (double)P.refine_factor;
                sector = 0.0;
                ksec = 0;
                jsec++;
            }
            const double sn = sin_horizon(P, h);  // (:333-358)
            if (P.svf_type == 0)
                svf += (1.0 - sn);
            else if (P.svf_type == 1)
                svf += (1.0 - (sn * sn));
            else
                svf += (1.0 - (sn * sn * sn));
            dir = rodrigues(dir, n, -P.azim_sin, P.azim_cos);  // (:550-551)
        }
        if (t == 0) skyview[cell] = svf / (double)L;
    }
    const unsigned long long wr = wave_sum(rays), wsp = wave_sum(spec);
    const unsigned long long wn = COUNT ? wave_sum(nvisit) : 0ull;
    const unsigned long long wt = COUNT ? wave_sum(ntest) : 0ull;
    const int werr = __any(err != 0);
    if (lane == 0) {
        if (wr) atomicAdd(&ts->cnt.rays, wr);
        if (wsp) atomicAdd(&ts->cnt.spec_rays, wsp);
        if (COUNT) {
            atomicAdd(&ts->cnt.nodes, wn);
            atomicAdd(&ts->cnt.tris, wt);
        }
        if (werr) atomicOr(&ts->cnt.error, 1ull);
    }
}

// Sector means and SVF of the item results of one domain, cell by cell in
// the order of the cell-by-cell search (:736-754).
__global__ void k_items_finish(const double* __restrict__ hitem, const int* __restrict__ order,
                               WalkParams P, double* __restrict__ horizon,
                               double* __restrict__ skyview) {
    const int pr = blockIdx.x * 256 + threadIdx.x;
    if (pr >= (P.range ? P.range[1] : P.ncell)) return;
    const int p = (P.range ? P.range[0] : 0) + pr;
    const int cell = order ? order[p] : p;
    const int L = P.num_hori * P.refine_factor;
    double sector = 0.0, svf = 0.0;
    int ksec = 0, jsec = 0;
    for (int j = 0; j < L; j++) {
        const double h = hitem[(size_t)j * P.ncell + p];
        sector += h;
        if (++ksec == P.refine_factor) {
            horizon[(size_t)jsec * P.ncell + cell] =
                ((sector / kPi) * 180.0) / (double)P.refine_factor;
            sector = 0.0;
            ksec = 0;
            jsec++;
        }
        const double sn = sin_horizon(P, h);
        if (P.svf_type == 0)
            svf += (1.0 - sn);
        else if (P.svf_type == 1)
            svf += (1.0 - (sn * sn));
        else
            svf += (1.0 - (sn * sn * sn));
    }
    skyview[cell] = svf / (double)L;
}

template <int MODE, bool COUNT, int BLOCK, int TREE, bool BATCH, bool ITEMS = false,
          bool PFW = false>
int pt_grid(int ncell, int reserve = 0) {
    static const int occ_grid = []() {
        int dev = 0, cus = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per,
                reinterpret_cast<const void*>(
                    &k_trace_pt<MODE, COUNT, BLOCK, TREE, BATCH, ITEMS, PFW>),
                BLOCK, 0) != hipSuccess)
            return 0;
        return std::max(1, per) * std::max(1, cus);
    }();
    int want = (ncell + BLOCK - 1) / BLOCK;
    int g = occ_grid > 0 ? std::min(std::max(8, occ_grid - reserve), want) : want;
    return std::max(8, (g + 7) / 8 * 8);
}

template <int MODE, bool COUNT, int BLOCK, int TREE>
void launch_pt(const DeviceBvh& bvh, const CellFrame* cells, const int* order,
               const WalkParams& p, int refill, double* horizon, double* skyview,
               TraceState* ts, double* spill, double* hitem, bool finish, hipStream_t st) {
    if constexpr (MODE != 0) {
        if (hitem) {
            const int items = p.ncell * p.num_hori * p.refine_factor;
            hipLaunchKernelGGL((k_trace_pt<MODE, COUNT, BLOCK, TREE, false, true>),
                               dim3(pt_grid<MODE, COUNT, BLOCK, TREE, false, true>(
                                   items, p.reserve_blocks)),
                               dim3(BLOCK), 0, st, bvh.nodes, bvh.tris, bvh.num_tri, cells, order,
                               p, refill, kChunk, horizon, skyview, ts, spill, nullptr, 0, 0,
                               hitem);
            if (finish)
                hipLaunchKernelGGL(k_items_finish, dim3((p.ncell + 255) / 256), dim3(256), 0, st,
                                   hitem, order, p, horizon, skyview);
            return;
        }
    }
    if constexpr (MODE == 0 && TREE == 0) {
        if (p.range) {  // the near cells' walk of the certified searches, beside the lattice
            hipLaunchKernelGGL((k_trace_pt<MODE, COUNT, BLOCK, TREE, false, false, true>),
                               dim3(pt_grid<MODE, COUNT, BLOCK, TREE, false, false, true>(
                                   p.ncell)),
                               dim3(BLOCK), 0, st, bvh.nodes, bvh.tris, bvh.num_tri, cells, order,
                               p, refill, kChunk, horizon, skyview, ts, spill, nullptr, 0, 0,
                               nullptr);
            return;
        }
    }
    hipLaunchKernelGGL((k_trace_pt<MODE, COUNT, BLOCK, TREE, false>),
                       dim3(pt_grid<MODE, COUNT, BLOCK, TREE, false>(p.ncell)), dim3(BLOCK), 0,
                       st, bvh.nodes, bvh.tris, bvh.num_tri, cells, order, p, refill, kChunk,
                       horizon, skyview, ts, spill, nullptr, 0, 0, nullptr);
}

template <int MODE, int BLOCK, int TREE>
void launch_pt_batch(const BatchDomain* doms, int ndom, int total, const WalkParams& p,
                     int refill, int chunk, TraceState* ts, double* spill, hipStream_t st) {
    hipLaunchKernelGGL((k_trace_pt<MODE, false, BLOCK, TREE, true>),
                       dim3(pt_grid<MODE, false, BLOCK, TREE, true>(total)), dim3(BLOCK), 0, st,
                       nullptr, nullptr, 0, nullptr, nullptr, p, refill, chunk, nullptr, nullptr,
                       ts, spill, doms, ndom, total, nullptr);
}

inline int tree_kind(bool multi, int width) { return width == 4 ? 2 : (multi ? 1 : 0); }

template <int MODE, bool COUNT, int BLOCK>
void launch_one(const DeviceBvh& bvh, const CellFrame* cells, const int* order,
                const WalkParams& p, int kernel, bool multi, int refill, double* horizon,
                double* skyview, TraceState* ts, double* spill, double* hitem, bool finish,
                hipStream_t st) {
    if (kernel == 0) {
        const int tk = tree_kind(multi, bvh.width);
        if (tk == 2)
            launch_pt<MODE, COUNT, BLOCK, 2>(bvh, cells, order, p, refill, horizon, skyview, ts,
                                             spill, hitem, finish, st);
        else if (tk == 1)
            launch_pt<MODE, COUNT, BLOCK, 1>(bvh, cells, order, p, refill, horizon, skyview, ts,
                                             spill, hitem, finish, st);
        else
            launch_pt<MODE, COUNT, BLOCK, 0>(bvh, cells, order, p, refill, horizon, skyview, ts,
                                             spill, hitem, finish, st);
        return;
    }
    if constexpr (MODE == 2) {
        return;  // the learned start exists only in the persistent kernel
    }
    TraceCounters* cnt = &ts->cnt;
    int nb = (p.ncell + BLOCK - 1) / BLOCK;
    int g = (nb + 7) / 8 * 8;
    hipLaunchKernelGGL((k_trace<MODE, COUNT, BLOCK>), dim3(g), dim3(BLOCK), 0, st, bvh.nodes,
                       bvh.tris, bvh.num_tri, cells, order, p, nb, horizon, skyview, cnt);
}

template <int MODE, bool COUNT>
void launch_block(int block, const DeviceBvh& bvh, const CellFrame* cells, const int* order,
                  const WalkParams& p, int kernel, bool multi, int refill, double* horizon,
                  double* skyview, TraceState* ts, double* spill, double* hitem, bool finish,
                  hipStream_t st) {
    if (block == 64)
        launch_one<MODE, COUNT, 64>(bvh, cells, order, p, kernel, multi, refill, horizon, skyview,
                                    ts, spill, hitem, finish, st);
    else if (block == 256)
        launch_one<MODE, COUNT, 256>(bvh, cells, order, p, kernel, multi, refill, horizon,
                                     skyview, ts, spill, hitem, finish, st);
    else
        launch_one<MODE, COUNT, 128>(bvh, cells, order, p, kernel, multi, refill, horizon,
                                     skyview, ts, spill, hitem, finish, st);
}

template <int MODE, int BLOCK>
void batch_tree(int tk, const BatchDomain* doms, int ndom, int total, const WalkParams& p,
                int refill, int chunk, TraceState* ts, double* spill, hipStream_t st) {
    if (tk == 2)
        launch_pt_batch<MODE, BLOCK, 2>(doms, ndom, total, p, refill, chunk, ts, spill, st);
    else if (tk == 1)
        launch_pt_batch<MODE, BLOCK, 1>(doms, ndom, total, p, refill, chunk, ts, spill, st);
    else
        launch_pt_batch<MODE, BLOCK, 0>(doms, ndom, total, p, refill, chunk, ts, spill, st);
}

template <int MODE>
void batch_block(int block, int tk, const BatchDomain* doms, int ndom, int total,
                 const WalkParams& p, int refill, int chunk, TraceState* ts, double* spill,
                 hipStream_t st) {
    if (block == 64)
        batch_tree<MODE, 64>(tk, doms, ndom, total, p, refill, chunk, ts, spill, st);
    else if (block == 256)
        batch_tree<MODE, 256>(tk, doms, ndom, total, p, refill, chunk, ts, spill, st);
    else
        batch_tree<MODE, 128>(tk, doms, ndom, total, p, refill, chunk, ts, spill, st);
}

template <int MODE, int BLOCK>
int batch_grid_tree(int tk, int total) {
    if (tk == 2) return pt_grid<MODE, false, BLOCK, 2, true>(total) * BLOCK;
    if (tk == 1) return pt_grid<MODE, false, BLOCK, 1, true>(total) * BLOCK;
    return pt_grid<MODE, false, BLOCK, 0, true>(total) * BLOCK;
}

template <int MODE>
int batch_grid_block(int block, int tk, int total) {
    if (block == 64) return batch_grid_tree<MODE, 64>(tk, total);
    if (block == 256) return batch_grid_tree<MODE, 256>(tk, total);
    return batch_grid_tree<MODE, 128>(tk, total);
}

}  // namespace

size_t trace_spill_doubles(int ncell) {
    // the persistent grid never holds more lanes than the device can keep
    // resident (32 waves of 64 lanes per CU) nor more than ncell + 8 blocks
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    size_t lanes = std::min((size_t)cus * 2048, (size_t)ncell + 8 * 256);
    return lanes * S_FIELDS;
}

int batch_grid(int mode, int block, bool multi, int width, int total) {
    const int tk = tree_kind(multi, width);
    return mode == 0 ? batch_grid_block<0>(block, tk, total) : batch_grid_block<1>(block, tk, total);
}

int launch_trace_batch(const BatchDomain* doms, int ndom, int total, bool multi, int width,
                       const WalkParams& p, int mode, int block, int refill, int chunk,
                       TraceState* ts, double* spill, hipStream_t st, std::string& err) {
    if (refill < 1 || refill > 64) refill = kDefaultRefill;
    if (chunk < 1) chunk = kChunk;
    const int tk = tree_kind(multi, width);
    hipLaunchKernelGGL(k_clear_state, dim3(1), dim3(256), 0, st, ts);
    if (mode == 0)
        batch_block<0>(block, tk, doms, ndom, total, p, refill, chunk, ts, spill, st);
    else if (mode == 1)
        batch_block<1>(block, tk, doms, ndom, total, p, refill, chunk, ts, spill, st);
    else {
        err = "the learned search is not available in batches";
        return -1;
    }
    HZ_HIP(hipGetLastError(), err);
    return 0;
}

template <int TEAM, bool COUNT>
void launch_team(const DeviceBvh& bvh, const CellFrame* cells, const int* order,
                 const WalkParams& p, double* horizon, double* skyview, TraceState* ts,
                 int max_cells, int grid_blocks, hipStream_t st) {
    constexpr int B = 128;
    const long long lanes = (long long)max_cells * TEAM;
    const int grid = (int)std::max<long long>(
        1, std::min<long long>(grid_blocks, (lanes + B - 1) / B));
    hipLaunchKernelGGL(k_clear_queue, dim3(1), dim3(256), 0, st, ts);
    hipLaunchKernelGGL((k_walk_team<TEAM, COUNT, B>), dim3(grid), dim3(B), 0, st, bvh.nodes,
                       bvh.tris, bvh.num_tri, cells, order, p, horizon, skyview, ts);
}

template <bool COUNT>
void team_size(int team, const DeviceBvh& bvh, const CellFrame* cells, const int* order,
               const WalkParams& p, double* horizon, double* skyview, TraceState* ts,
               int max_cells, int grid_blocks, hipStream_t st) {
    switch (team) {
        case 2: launch_team<2, COUNT>(bvh, cells, order, p, horizon, skyview, ts, max_cells, grid_blocks, st); break;
        case 4: launch_team<4, COUNT>(bvh, cells, order, p, horizon, skyview, ts, max_cells, grid_blocks, st); break;
        case 8: launch_team<8, COUNT>(bvh, cells, order, p, horizon, skyview, ts, max_cells, grid_blocks, st); break;
        case 32: launch_team<32, COUNT>(bvh, cells, order, p, horizon, skyview, ts, max_cells, grid_blocks, st); break;
        case 64: launch_team<64, COUNT>(bvh, cells, order, p, horizon, skyview, ts, max_cells, grid_blocks, st); break;
        default: launch_team<16, COUNT>(bvh, cells, order, p, horizon, skyview, ts, max_cells, grid_blocks, st); break;
    }
}

void launch_clear_state(TraceState* ts, hipStream_t st) {
    hipLaunchKernelGGL(k_clear_state, dim3(1), dim3(256), 0, st, ts);
}

int launch_walk_team(const DeviceBvh& bvh, const CellFrame* cells, const int* order,
                     const WalkParams& p, int team, bool counters, double* horizon,
                     double* skyview, TraceState* ts, int max_cells, int grid_blocks,
                     hipStream_t st, std::string& err) {
    if (bvh.width != 2) {
        err = "the cooperative walk traverses binary trees only";
        return -1;
    }
    max_cells = std::max(1, std::min(max_cells, p.ncell));
    if (counters)
        team_size<true>(team, bvh, cells, order, p, horizon, skyview, ts, max_cells,
                        grid_blocks, st);
    else
        team_size<false>(team, bvh, cells, order, p, horizon, skyview, ts, max_cells,
                         grid_blocks, st);
    HZ_HIP(hipGetLastError(), err);
    return 0;
}

namespace {

__global__ __launch_bounds__(256) void k_start_nodes(const BvhNode* __restrict__ nodes,
                                                    const CellFrame* __restrict__ cells,
                                                    int ncell, float reach,
                                                    int* __restrict__ start) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncell) return;
    const float ox = cells[c].org[0], oy = cells[c].org[1];
    const float r2 = reach * reach;
    int ref = kWideBit;  // the root
    for (int depth = 0; depth < 128; depth++) {  // a chain of interior nodes: < 128 levels
        const float4* np = reinterpret_cast<const float4*>(nodes + (ref & kLeafFirstMask));
        const float4 a = np[0], b = np[1], d = np[3];  // x planes, y planes, references
        const bool in0 = ox >= a.x && ox <= a.z && oy >= b.x && oy <= b.z;
        const bool in1 = ox >= a.y && ox <= a.w && oy >= b.y && oy <= b.w;
        if (in0 == in1) break;  // both children hold the origin, or neither
        // the other child's x/y box: its distance from the origin
        const float lx = in0 ? a.y : a.x, hx = in0 ? a.w : a.z;
        const float ly = in0 ? b.y : b.x, hy = in0 ? b.w : b.z;
        const float dx = fmaxf(fmaxf(lx - ox, ox - hx), 0.0f);
        const float dy = fmaxf(fmaxf(ly - oy, oy - hy), 0.0f);
        if (dx * dx + dy * dy <= r2) break;  // within reach: its subtree stays in
        const int child = __float_as_int(in0 ? d.x : d.y);
        if (child < 0) break;  // a leaf
        ref = child;
    }
    start[c] = ref;
}

}  // namespace

void launch_start_nodes(const DeviceBvh& bvh, const CellFrame* cells, int ncell, float reach,
                        int* start, hipStream_t stream) {
    if (ncell <= 0 || bvh.width != 2 || bvh.num_tri <= 0) return;
    hipLaunchKernelGGL(k_start_nodes, dim3((ncell + 255) / 256), dim3(256), 0, stream, bvh.nodes,
                       cells, ncell, reach, start);
}

int launch_trace(const DeviceBvh& bvh, const CellFrame* cells, const int* order,
                 const WalkParams& p, int mode, int kernel, int block, int refill, bool counters,
                 double* horizon, double* skyview, TraceState* ts, double* spill, double* hitem,
                 bool items_finish, hipStream_t st, std::string& err) {
    const bool multi = bvh.leaf_size > 1;
    if (refill < 1 || refill > 64) refill = kDefaultRefill;
    hipLaunchKernelGGL(k_clear_state, dim3(1), dim3(256), 0, st, ts);
    if (mode == 2 && (kernel == 1 || !p.pred)) {
        err = "the learned search needs the persistent kernel and predictions";
        return -1;
    }
    if (kernel == 1 && bvh.width != 2) {
        err = "the lane-per-cell kernel traverses binary trees only";
        return -1;
    }
#define HZ_LAUNCH(M)                                                                          \
    do {                                                                                      \
        if (counters)                                                                         \
            launch_block<M, true>(block, bvh, cells, order, p, kernel, multi, refill, horizon,   \
                                  skyview, ts, spill, hitem, items_finish, st);               \
        else                                                                                  \
            launch_block<M, false>(block, bvh, cells, order, p, kernel, multi, refill, horizon,  \
                                   skyview, ts, spill, hitem, items_finish, st);              \
    } while (0)
    if (mode == 0)
        HZ_LAUNCH(0);
    else if (mode == 1)
        HZ_LAUNCH(1);
    else
        HZ_LAUNCH(2);
#undef HZ_LAUNCH
    HZ_HIP(hipGetLastError(), err);
    return 0;
}

}  // namespace hz
