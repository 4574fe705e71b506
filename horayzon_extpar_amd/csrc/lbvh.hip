// On-device LBVH builder (replaces the reference's Embree rtcCommitScene,
// mo_lradtopo_horayzon.cpp:395-432).
//
//   1. k_tri_bounds   triangle boxes and per-workgroup centroid bounds,
//      k_reduce_bounds  the scene's centroid bounds (no atomics)
//   2. k_morton       63-bit Morton codes of the centroids in a cube of the
//                     largest extent (a heightfield then splits mostly in x/y)
//   3. rocprim radix sort of (code, triangle) pairs
//   4. k_karras       Karras 2012 hierarchy ("Maximizing parallelism in the
//                     construction of BVHs, octrees and k-d trees", HPG 2012),
//                     duplicate codes broken by leaf index
//   5. interior boxes without a bottom-up pass: a Karras node covers the
//      contiguous leaf range [first, last], so its box is the union of that
//      range.  k_block_scan forms prefix and suffix unions inside 256-leaf
//      blocks, k_sparse_level a sparse table over block unions, and
//      k_node_boxes answers every node's range in O(1) (short in-block
//      ranges are unioned directly).  No atomics and no fences.
//   6. k_pack         128-byte traversal nodes holding both children's
//                     conservative boxes and inline single-triangle leaves;
//                     subtrees of <= leaf_size triangles collapse into one
//                     leaf; larger leaves' triangles written in leaf order
//
// Every box is grown by 1 m + 1e-5 |coordinate| so that the FP32 slab test
// can never reject a ray that the watertight triangle test accepts: the
// occlusion answer is then the OR of the triangle test over all triangles,
// whatever the tree looks like.

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "hz_internal.h"

namespace hz {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ unsigned f2ord(float f) {
    unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(unsigned u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_init_bounds(unsigned* b) {
    int t = threadIdx.x;
    if (t < 3) b[t] = 0xffffffffu;        // min x, y, z
    else if (t < 6) b[t] = 0u;            // max x, y, z
    else if (t < 8) b[t] = 0u;            // depth, spare
}

// Triangle boxes (original order) and, per workgroup, the min / max of the
// triangle centroids: part[block * 8 + 0..2] = min, [3..5] = max.
__global__ __launch_bounds__(kBlock) void k_tri_bounds(const float* __restrict__ vert,
                                                      const int* __restrict__ tri, int n,
                                                      float4* __restrict__ lo,
                                                      float4* __restrict__ hi,
                                                      float* __restrict__ part) {
    __shared__ float red[6][kBlock / 64];
    int i = blockIdx.x * kBlock + threadIdx.x;
    float cmin[3] = {3.0e38f, 3.0e38f, 3.0e38f}, cmax[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    if (i < n) {
        int a = tri[3 * i], b = tri[3 * i + 1], c = tri[3 * i + 2];
        float l[3], h[3];
        for (int k = 0; k < 3; k++) {
            float p0 = vert[3 * a + k], p1 = vert[3 * b + k], p2 = vert[3 * c + k];
            l[k] = fminf(p0, fminf(p1, p2));
            h[k] = fmaxf(p0, fmaxf(p1, p2));
            float cc = 0.5f * (l[k] + h[k]);
            cmin[k] = cc;
            cmax[k] = cc;
        }
        lo[i] = make_float4(l[0], l[1], l[2], 0.0f);
        hi[i] = make_float4(h[0], h[1], h[2], 0.0f);
    }
    for (int k = 0; k < 3; k++) {
        for (int off = 32; off > 0; off >>= 1) {
            cmin[k] = fminf(cmin[k], __shfl_xor(cmin[k], off, 64));
            cmax[k] = fmaxf(cmax[k], __shfl_xor(cmax[k], off, 64));
        }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; k++) {
            red[k][w] = cmin[k];
            red[3 + k][w] = cmax[k];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float v = red[k][0];
        for (int q = 1; q < kBlock / 64; q++) v = k < 3 ? fminf(v, red[k][q]) : fmaxf(v, red[k][q]);
        part[blockIdx.x * 8 + k] = v;
    }
}

// Binary trees: replace the z bounds of triangle boxes by bounds of
// z' = z + (x^2 + y^2) kZCurv.  z' is convex over a triangle, so its maximum
// is at a vertex.  Below its linear interpolation between the vertices it
// dips by kZCurv times the barycentric variance of the vertices' x/y
// positions, at most diam^2 / 3 (an equilateral triangle's circumradius
// squared), so the lowest vertex z' minus kZCurv diam^2 / 3 bounds it from
// below, diam^2 taken as the x/y box's diagonal squared.  FP64, rounded
// outward; k_pack grows them like the other planes.  Runs after the Morton
// codes (which keep the ENU centroids).
__global__ __launch_bounds__(kBlock) void k_tri_zcurv(const float* __restrict__ vert,
                                                     const int* __restrict__ tri, int n,
                                                     float4* __restrict__ lo,
                                                     float4* __restrict__ hi) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double zl = 1e300, zh = -1e300, xl = 1e300, xh = -1e300, yl = 1e300, yh = -1e300;
    for (int k = 0; k < 3; k++) {
        const int v = tri[3 * i + k];
        const double x = vert[3 * v], y = vert[3 * v + 1], z = vert[3 * v + 2];
        const double zc = z + (x * x + y * y) * kZCurv;
        zl = fmin(zl, zc);
        zh = fmax(zh, zc);
        xl = fmin(xl, x);
        xh = fmax(xh, x);
        yl = fmin(yl, y);
        yh = fmax(yh, y);
    }
    zl -= ((xh - xl) * (xh - xl) + (yh - yl) * (yh - yl)) * kZCurv / 3.0;
    lo[i].z = __double2float_rd(zl);
    hi[i].z = __double2float_ru(zh);
}

// One workgroup: reduce nb partials into the encoded bounds b[0..5].
__global__ __launch_bounds__(kBlock) void k_reduce_bounds(const float* __restrict__ part, int nb,
                                                         unsigned* __restrict__ b) {
    __shared__ float red[6][kBlock];
    float v[6] = {3.0e38f, 3.0e38f, 3.0e38f, -3.0e38f, -3.0e38f, -3.0e38f};
    for (int q = threadIdx.x; q < nb; q += kBlock)
        for (int k = 0; k < 6; k++)
            v[k] = k < 3 ? fminf(v[k], part[q * 8 + k]) : fmaxf(v[k], part[q * 8 + k]);
    for (int k = 0; k < 6; k++) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 6; k++)
                red[k][threadIdx.x] = k < 3 ? fminf(red[k][threadIdx.x], red[k][threadIdx.x + s])
                                            : fmaxf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 6) b[threadIdx.x] = f2ord(red[threadIdx.x][0]);
}

__device__ __forceinline__ uint64_t spread21(uint64_t x) {
    x &= 0x1fffffull;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__device__ __forceinline__ uint64_t spread32(uint64_t x);

// mode 0: 3D codes (21 bits per axis) in a cube of the largest extent;
// mode 1: 2D codes (32 bits per axis) of x / y only
__global__ __launch_bounds__(kBlock) void k_morton(const float4* __restrict__ lo,
                                                  const float4* __restrict__ hi, int n,
                                                  const unsigned* __restrict__ bounds, int mode,
                                                  uint64_t* __restrict__ keys,
                                                  int* __restrict__ vals) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float bmin[3], ext = 0.0f;
    for (int k = 0; k < 3; k++) {
        bmin[k] = ord2f(bounds[k]);
        ext = fmaxf(ext, ord2f(bounds[3 + k]) - bmin[k]);
    }
    float inv = ext > 0.0f ? 1.0f / ext : 0.0f;
    float4 l = lo[i], h = hi[i];
    float c[3] = {0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
    uint64_t code = 0;
    if (mode == 1) {
        float e2 = fmaxf(ord2f(bounds[3]) - bmin[0], ord2f(bounds[4]) - bmin[1]);
        float i2 = e2 > 0.0f ? 1.0f / e2 : 0.0f;
        float qx = fminf(fmaxf((c[0] - bmin[0]) * i2, 0.0f), 1.0f);
        float qy = fminf(fmaxf((c[1] - bmin[1]) * i2, 0.0f), 1.0f);
        uint64_t ux = (uint64_t)fminf(qx * 4294967040.0f, 4294967040.0f);
        uint64_t uy = (uint64_t)fminf(qy * 4294967040.0f, 4294967040.0f);
        keys[i] = ((spread32(uy) << 1) | spread32(ux)) >> 1;  // 63 bits
        vals[i] = i;
        return;
    }
    for (int k = 0; k < 3; k++) {
        float q = (c[k] - bmin[k]) * inv;
        q = fminf(fmaxf(q, 0.0f), 1.0f);
        uint64_t v = (uint64_t)fminf(q * 2097152.0f, 2097151.0f);
        code |= spread21(v) << (2 - k);
    }
    keys[i] = code;
    vals[i] = i;
}

__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz((unsigned)(i ^ j));
    return __clzll(a ^ b);
}

// Interior node i of n-1; children encoded: >= 0 interior, < 0 leaf ~index.
__global__ __launch_bounds__(kBlock) void k_karras(const uint64_t* __restrict__ keys, int n,
                                                  int* __restrict__ child,
                                                  int* __restrict__ range,
                                                  int* __restrict__ parent) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0;
    int div = 2;
    int t = (l + div - 1) / div;
    while (true) {
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
        div <<= 1;
        t = (l + div - 1) / div;
    }
    int g = i + s * d + (d < 0 ? -1 : 0);
    int first = min(i, j), last = max(i, j);
    int left = (first == g) ? ~g : g;
    int right = (last == g + 1) ? ~(g + 1) : g + 1;
    child[2 * i] = left;
    child[2 * i + 1] = right;
    range[2 * i] = first;
    range[2 * i + 1] = last;
    // parent slots: leaves [0, n), interior [n, 2n - 1)
    parent[left < 0 ? ~left : n + left] = i;
    parent[right < 0 ? ~right : n + right] = i;
    if (i == 0) parent[n] = -1;
}

struct Box {
    float4 lo, hi;
};
__device__ __forceinline__ Box box_empty() {
    return {make_float4(3.0e38f, 3.0e38f, 3.0e38f, 0.f), make_float4(-3.0e38f, -3.0e38f, -3.0e38f, 0.f)};
}
__device__ __forceinline__ Box box_union(const Box& a, const Box& b) {
    return {make_float4(fminf(a.lo.x, b.lo.x), fminf(a.lo.y, b.lo.y), fminf(a.lo.z, b.lo.z), 0.f),
            make_float4(fmaxf(a.hi.x, b.hi.x), fmaxf(a.hi.y, b.hi.y), fmaxf(a.hi.z, b.hi.z), 0.f)};
}

// Per 256-leaf block of the sorted leaf order: the leaf boxes (sorted), their
// inclusive prefix and suffix unions within the block, and the block union
// (level 0 of the sparse table).
__global__ __launch_bounds__(kBlock) void k_block_scan(const int* __restrict__ sorted, int n,
                                                      const float4* __restrict__ lo,
                                                      const float4* __restrict__ hi,
                                                      Box* __restrict__ leaf, Box* __restrict__ pre,
                                                      Box* __restrict__ suf, Box* __restrict__ st0) {
    __shared__ Box sp[kBlock], ss[kBlock];
    const int k = blockIdx.x * kBlock + threadIdx.x;
    Box bx = box_empty();
    if (k < n) {
        int t = sorted[k];
        bx = {lo[t], hi[t]};
        leaf[k] = bx;
    }
    sp[threadIdx.x] = bx;
    ss[threadIdx.x] = bx;
    __syncthreads();
    for (int d = 1; d < kBlock; d <<= 1) {
        Box a = sp[threadIdx.x], b = ss[threadIdx.x];
        if (threadIdx.x >= d) a = box_union(a, sp[threadIdx.x - d]);
        if (threadIdx.x + d < kBlock) b = box_union(b, ss[threadIdx.x + d]);
        __syncthreads();
        sp[threadIdx.x] = a;
        ss[threadIdx.x] = b;
        __syncthreads();
    }
    if (k < n) {
        pre[k] = sp[threadIdx.x];
        suf[k] = ss[threadIdx.x];
    }
    if (threadIdx.x == 0) st0[blockIdx.x] = ss[0];
}

// dst[b] = src[b] U src[b + half] for b + half < nb (else src[b]).
__global__ __launch_bounds__(kBlock) void k_sparse_level(const Box* __restrict__ src, int nb,
                                                        int half, Box* __restrict__ dst) {
    int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= nb) return;
    dst[b] = b + half < nb ? box_union(src[b], src[b + half]) : src[b];
}

// Box of every interior node from its leaf range; written to lo/hi[n + i].
__global__ __launch_bounds__(kBlock) void k_node_boxes(const int* __restrict__ range, int n,
                                                      const Box* __restrict__ leaf,
                                                      const Box* __restrict__ pre,
                                                      const Box* __restrict__ suf,
                                                      const Box* __restrict__ st, int nb,
                                                      float4* __restrict__ lo,
                                                      float4* __restrict__ hi) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    const int a = range[2 * i], b = range[2 * i + 1];
    const int ba = a / kBlock, bb = b / kBlock;
    Box r;
    if (ba == bb) {
        r = leaf[a];
        for (int k = a + 1; k <= b; k++) r = box_union(r, leaf[k]);
    } else {
        r = box_union(suf[a], pre[b]);
        if (bb - ba > 1) {
            const int l = ba + 1, h = bb - 1;
            const int j = 31 - __clz(h - l + 1);  // floor(log2(length))
            r = box_union(r, box_union(st[(size_t)j * nb + l], st[(size_t)j * nb + h - (1 << j) + 1]));
        }
    }
    lo[n + i] = r.lo;
    hi[n + i] = r.hi;
}

__device__ __forceinline__ float grow_lo(float v) { return v - (1.0f + 1e-5f * fabsf(v)); }
__device__ __forceinline__ float grow_hi(float v) { return v + (1.0f + 1e-5f * fabsf(v)); }

__device__ __forceinline__ void put_tri(float* dst, const float* __restrict__ vert,
                                        const int* __restrict__ tri, int t) {
    for (int k = 0; k < 3; k++) {
        int v = tri[3 * t + k];
        dst[3 * k] = vert[3 * v];
        dst[3 * k + 1] = vert[3 * v + 1];
        dst[3 * k + 2] = vert[3 * v + 2];
    }
}

__global__ __launch_bounds__(kBlock) void k_pack(const int* __restrict__ sorted, int n,
                                                const float4* __restrict__ lo,
                                                const float4* __restrict__ hi,
                                                const int* __restrict__ child,
                                                const int* __restrict__ range, int leaf_size,
                                                const float* __restrict__ vert,
                                                const int* __restrict__ tri,
                                                BvhNode* __restrict__ nodes) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    float4 l[2], h[2];
    int ref[2];
    BvhNode nd;
    for (int k = 0; k < 18; k++) nd.tri[k] = 0.0f;
    for (int s = 0; s < 2; s++) {
        int c = child[2 * i + s];
        if (c < 0) {
            int leaf = ~c;
            int slot = sorted[leaf];
            l[s] = lo[slot];
            h[s] = hi[slot];
            ref[s] = ~leaf;  // count 1: inline triangle
            put_tri(&nd.tri[9 * s], vert, tri, slot);
        } else {
            l[s] = lo[n + c];
            h[s] = hi[n + c];
            int first = range[2 * c], last = range[2 * c + 1];
            int cnt = last - first + 1;
            if (cnt <= leaf_size) {
                ref[s] = ~(first | ((cnt - 1) << kLeafFirstBits));
            } else {
                bool wide = child[2 * c] < 0 || child[2 * c + 1] < 0;
                ref[s] = c | (wide ? kWideBit : 0);
            }
        }
    }
    nd.lx0 = grow_lo(l[0].x);
    nd.lx1 = grow_lo(l[1].x);
    nd.hx0 = grow_hi(h[0].x);
    nd.hx1 = grow_hi(h[1].x);
    nd.ly0 = grow_lo(l[0].y);
    nd.ly1 = grow_lo(l[1].y);
    nd.hy0 = grow_hi(h[0].y);
    nd.hy1 = grow_hi(h[1].y);
    nd.lz0 = grow_lo(l[0].z);
    nd.lz1 = grow_lo(l[1].z);
    nd.hz0 = grow_hi(h[0].z);
    nd.hz1 = grow_hi(h[1].z);
    nd.r0 = ref[0];
    nd.r1 = ref[1];
    // both children interior (the node has no inline triangle): tri[1] holds
    // the leaf-order index of child 0's last leaf, child 1's leaves follow it
    // (the traversal descends first towards the triangle of the last hit)
    if (ref[0] >= 0 && ref[1] >= 0) nd.tri[1] = __int_as_float(range[2 * child[2 * i] + 1]);
    nodes[i] = nd;
}

// One triangle: both children of the root are that triangle, inline.
__global__ void k_pack_single(const float4* __restrict__ lo, const float4* __restrict__ hi,
                              const float* __restrict__ vert, const int* __restrict__ tri,
                              BvhNode* __restrict__ nodes) {
    if (threadIdx.x != 0) return;
    float4 l = lo[0], h = hi[0];
    BvhNode nd;
    nd.lx0 = nd.lx1 = grow_lo(l.x);
    nd.ly0 = nd.ly1 = grow_lo(l.y);
    nd.lz0 = nd.lz1 = grow_lo(l.z);
    nd.hx0 = nd.hx1 = grow_hi(h.x);
    nd.hy0 = nd.hy1 = grow_hi(h.y);
    nd.hz0 = nd.hz1 = grow_hi(h.z);
    nd.r0 = nd.r1 = ~0;
    put_tri(&nd.tri[0], vert, tri, 0);
    put_tri(&nd.tri[9], vert, tri, 0);
    nodes[0] = nd;
}

__global__ __launch_bounds__(kBlock) void k_tris(const int* __restrict__ sorted, int n,
                                                const float* __restrict__ vert,
                                                const int* __restrict__ tri,
                                                float4* __restrict__ out) {
    int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    int t = sorted[k];
    int a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
    out[3 * k] = make_float4(vert[3 * a], vert[3 * a + 1], vert[3 * a + 2], vert[3 * b]);
    out[3 * k + 1] = make_float4(vert[3 * b + 1], vert[3 * b + 2], vert[3 * c], vert[3 * c + 1]);
    out[3 * k + 2] = make_float4(vert[3 * c + 2], __int_as_float(t), 0.0f, 0.0f);
}

// --- four-wide collapse -----------------------------------------------------
//
// Binary interior node i becomes a four-wide node when its depth is even and
// it holds more than leaf_size triangles (the root always does).  Its
// children are its binary children, each opened into its own two children
// unless it is a leaf or holds at most leaf_size triangles (then it is one
// leaf of the four-wide node).  An opened child's interior children sit at
// even depth again, so every four-wide child is itself a four-wide node.

__device__ __forceinline__ int leaf_count(const int* __restrict__ range, int c) {
    return range[2 * c + 1] - range[2 * c] + 1;
}

__global__ __launch_bounds__(kBlock) void k_b4_flag(const int* __restrict__ parent,
                                                   const int* __restrict__ range, int n,
                                                   int leaf_size, int* __restrict__ flag) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    int d = 0;
    int p = parent[n + i];
    while (p >= 0 && d < 4096) {  // the Karras hierarchy has depth < 64 + log2(n)
        d++;
        p = parent[n + p];
    }
    flag[i] = (i == 0 || ((d & 1) == 0 && leaf_count(range, i) > leaf_size)) ? 1 : 0;
}

__global__ void k_b4_count(const int* __restrict__ flag, const int* __restrict__ idx, int m,
                           unsigned* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = (unsigned)(idx[m - 1] + flag[m - 1]);
}

// Does the four-wide node made from binary node g hold a single-triangle leaf
// (and therefore inline triangles)?
__device__ __forceinline__ bool b4_has_single(const int* __restrict__ child,
                                              const int* __restrict__ range, int g,
                                              int leaf_size) {
    for (int s = 0; s < 2; s++) {
        const int c = child[2 * g + s];
        if (c < 0) return true;
        if (leaf_count(range, c) <= leaf_size) continue;
        if (child[2 * c] < 0 || child[2 * c + 1] < 0) return true;
    }
    return false;
}

// Quantize one axis: origin lo, scale 2^e with 255 * 2^e >= (hi - lo); child
// planes rounded outward (exact arithmetic in FP64: the products q * 2^e are
// exact and the sums carry at most one FP64 rounding, far below the 1 m box
// margin).
__device__ __forceinline__ unsigned b4_axis(const float* clo, const float* chi, int nch, float lo,
                                            float hi, unsigned& qlo, unsigned& qhi) {
    const double ext = (double)hi - (double)lo;
    const double need = ext * (1.0 + 0x1p-20) / 255.0;
    int e = -100;
    if (need > 0x1p-100) {
        e = ilogb(need);
        if (ldexp(1.0, e) < need) e++;
    }
    const double sc = ldexp(1.0, e);
    qlo = 0u;
    qhi = 0u;
    for (int k = 0; k < 4; k++) {
        unsigned ql = 255u, qh = 0u;  // empty slot
        if (k < nch) {
            double a = floor(((double)clo[k] - (double)lo) / sc);
            a = fmin(fmax(a, 0.0), 255.0);
            while (a > 0.0 && (double)lo + a * sc > (double)clo[k]) a -= 1.0;
            double b = ceil(((double)chi[k] - (double)lo) / sc);
            b = fmin(fmax(b, 0.0), 255.0);
            while (b < 255.0 && (double)lo + b * sc < (double)chi[k]) b += 1.0;
            ql = (unsigned)a;
            qh = (unsigned)b;
        }
        qlo |= ql << (8 * k);
        qhi |= qh << (8 * k);
    }
    return (unsigned)(e + 127);
}

__global__ __launch_bounds__(kBlock) void k_pack4(const int* __restrict__ sorted, int n,
                                                 const float4* __restrict__ lo,
                                                 const float4* __restrict__ hi,
                                                 const int* __restrict__ child,
                                                 const int* __restrict__ range, int leaf_size,
                                                 const int* __restrict__ flag,
                                                 const int* __restrict__ idx,
                                                 const float* __restrict__ vert,
                                                 const int* __restrict__ tri,
                                                 Bvh4Node* __restrict__ nodes) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1 || !flag[i]) return;
    // children: kind 0 = interior (binary index), 1 = single leaf (sorted
    // leaf index), 2 = leaf range (binary index)
    int kid[4], kind[4];
    int nch = 0;
    for (int s = 0; s < 2; s++) {
        const int c = child[2 * i + s];
        if (c < 0) {
            kid[nch] = ~c;
            kind[nch++] = 1;
        } else if (leaf_count(range, c) <= leaf_size) {
            kid[nch] = c;
            kind[nch++] = 2;
        } else {
            for (int t = 0; t < 2; t++) {
                const int g = child[2 * c + t];
                if (g < 0) {
                    kid[nch] = ~g;
                    kind[nch++] = 1;
                } else {
                    kid[nch] = g;
                    kind[nch++] = leaf_count(range, g) <= leaf_size ? 2 : 0;
                }
            }
        }
    }
    float clo[3][4], chi[3][4];
    float blo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, bhi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (int k = 0; k < nch; k++) {
        const int slot = kind[k] == 1 ? sorted[kid[k]] : n + kid[k];
        const float4 l = lo[slot], h = hi[slot];
        const float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z};
        for (int a = 0; a < 3; a++) {
            clo[a][k] = grow_lo(lv[a]);
            chi[a][k] = grow_hi(hv[a]);
            blo[a] = fminf(blo[a], clo[a][k]);
            bhi[a] = fmaxf(bhi[a], chi[a][k]);
        }
    }
    Bvh4Node nd;
    for (int k = 0; k < 18; k++) nd.tri[k] = 0.0f;
    unsigned meta = (unsigned)nch << 24;
    for (int a = 0; a < 3; a++) {
        nd.org[a] = blo[a];
        meta |= b4_axis(clo[a], chi[a], nch, blo[a], bhi[a], nd.q[2 * a], nd.q[2 * a + 1]) << (8 * a);
    }
    nd.meta = meta;
    int inl = 0;
    for (int k = 0; k < 4; k++) {
        int r = 0;
        if (k < nch) {
            if (kind[k] == 0) {
                r = idx[kid[k]] | (b4_has_single(child, range, kid[k], leaf_size) ? kWideBit : 0);
            } else if (kind[k] == 1) {
                if (inl < 2) {
                    put_tri(&nd.tri[9 * inl], vert, tri, sorted[kid[k]]);
                    r = ~(kB4InlineBit | (kid[k] << 1) | inl);
                    inl++;
                } else {
                    r = ~kid[k];  // one triangle in the leaf-order array
                }
            } else {
                const int first = range[2 * kid[k]];
                r = ~(first | ((leaf_count(range, kid[k]) - 1) << kB4CountShift));
            }
        }
        nd.ref[k] = r;
    }
    nodes[idx[i]] = nd;
}

// One triangle: the root holds it as its only child, inline.
__global__ void k_pack4_single(const float4* __restrict__ lo, const float4* __restrict__ hi,
                               const float* __restrict__ vert, const int* __restrict__ tri,
                               Bvh4Node* __restrict__ nodes, unsigned* __restrict__ count) {
    if (threadIdx.x != 0) return;
    const float4 l = lo[0], h = hi[0];
    const float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z};
    Bvh4Node nd;
    for (int k = 0; k < 18; k++) nd.tri[k] = 0.0f;
    unsigned meta = 1u << 24;
    for (int a = 0; a < 3; a++) {
        float cl[4] = {grow_lo(lv[a]), 0.f, 0.f, 0.f}, ch[4] = {grow_hi(hv[a]), 0.f, 0.f, 0.f};
        nd.org[a] = cl[0];
        meta |= b4_axis(cl, ch, 1, cl[0], ch[0], nd.q[2 * a], nd.q[2 * a + 1]) << (8 * a);
    }
    nd.meta = meta;
    nd.ref[0] = ~(kB4InlineBit | 0);
    nd.ref[1] = nd.ref[2] = nd.ref[3] = 0;
    put_tri(&nd.tri[0], vert, tri, 0);
    nodes[0] = nd;
    *count = 1u;
}

// Depth of every interior node's deepest leaf path: each leaf walks to the
// root counting levels (diagnostics and stack sizing).
__global__ __launch_bounds__(kBlock) void k_depth(const int* __restrict__ parent, int n,
                                                 unsigned* __restrict__ out) {
    int leaf = blockIdx.x * kBlock + threadIdx.x;
    if (leaf >= n) return;
    int p = parent[leaf];
    unsigned d = 0;
    while (p >= 0 && d < 256) {
        d++;
        p = parent[n + p];
    }
    // d = number of interior ancestors; deepest interior level = d - 1
    for (int off = 32; off > 0; off >>= 1) d = max(d, (unsigned)__shfl_xor((int)d, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, d);
}

// --- cell ordering ---------------------------------------------------------

// Per workgroup, the min / max of the cells' ray origins in x and y, laid
// out like k_tri_bounds' partials (z slots 0) for k_reduce_bounds.
__global__ __launch_bounds__(kBlock) void k_cell_bounds(const CellFrame* __restrict__ cells,
                                                       int n, float* __restrict__ part) {
    __shared__ float red[4][kBlock / 64];
    int i = blockIdx.x * kBlock + threadIdx.x;
    float mn[2] = {3.0e38f, 3.0e38f}, mx[2] = {-3.0e38f, -3.0e38f};
    if (i < n) {
        for (int k = 0; k < 2; k++) mn[k] = mx[k] = cells[i].org[k];
    }
    for (int k = 0; k < 2; k++)
        for (int off = 32; off > 0; off >>= 1) {
            mn[k] = fminf(mn[k], __shfl_xor(mn[k], off, 64));
            mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], off, 64));
        }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 2; k++) {
            red[k][w] = mn[k];
            red[2 + k][w] = mx[k];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;  // 0, 1: min x, y; 3, 4: max x, y; 2, 5: z
        float v = 0.0f;
        if (k != 2 && k != 5) {
            const int r = k < 3 ? k : k - 1;
            v = red[r][0];
            for (int q = 1; q < kBlock / 64; q++) v = k < 3 ? fminf(v, red[r][q]) : fmaxf(v, red[r][q]);
        }
        part[blockIdx.x * 8 + k] = v;
    }
}

__device__ __forceinline__ uint64_t spread32(uint64_t x) {
    x &= 0xffffffffull;
    x = (x | (x << 16)) & 0x0000ffff0000ffffull;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

__global__ __launch_bounds__(kBlock) void k_cell_morton(const CellFrame* __restrict__ cells,
                                                       int n, const unsigned* __restrict__ bounds,
                                                       uint64_t* __restrict__ keys,
                                                       int* __restrict__ vals) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float bx = ord2f(bounds[0]), by = ord2f(bounds[1]);
    float ext = fmaxf(ord2f(bounds[3]) - bx, ord2f(bounds[4]) - by);
    float inv = ext > 0.0f ? 1.0f / ext : 0.0f;
    float qx = fminf(fmaxf((cells[i].org[0] - bx) * inv, 0.0f), 1.0f);
    float qy = fminf(fmaxf((cells[i].org[1] - by) * inv, 0.0f), 1.0f);
    uint64_t ux = (uint64_t)fminf(qx * 16777216.0f, 16777215.0f);
    uint64_t uy = (uint64_t)fminf(qy * 16777216.0f, 16777215.0f);
    // cells next to a mesh hole (flag bit 0) sort first: their searches are
    // the longest, and starting them first shortens the end of the launch;
    // then cells with a mesh-boundary point in reach (flag bit 1, set by the
    // certificate of the lattice searches), then the rest
    const int fl = cells[i].pad;
    const uint64_t cls = (fl & 1) ? 0ull : ((fl & 2) ? 1ull : 2ull);
    keys[i] = (cls << 48) | (spread32(uy) << 1) | spread32(ux);
    vals[i] = i;
}

inline int grid(int n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

// ---------------------------------------------------------------------------
// SAH rebuild of the top of the tree (option): the interior nodes whose leaf
// range holds more than T triangles form a connected top part of the Morton
// tree; the subtrees hanging below it (and leaves hanging from it) are cut
// off as K items, the top is rebuilt over the items by a binned surface-area
// heuristic on the host (K ≈ 8k, microseconds of work), and the K - 1 new
// nodes are written into the slots of the K - 1 old top nodes, the root into
// slot 0.  Item boxes are the grown boxes their old parents stored; a new
// node's box is the union of grown boxes, a superset as the traversal needs.
// The answer of a query stays the OR over all triangles, so results do not
// change, only which nodes a ray visits.
// ---------------------------------------------------------------------------

struct SahItem {
    float lo[3], hi[3];
    int ref;
    int pad;
    float tri[9];  // the inline triangle of a single-triangle leaf
    float pad2[3];
};
static_assert(sizeof(SahItem) == 80, "SahItem layout");

namespace {

__global__ __launch_bounds__(kBlock) void k_sah_cut(const int* __restrict__ range,
                                                   const int* __restrict__ child, int n, int T,
                                                   const BvhNode* __restrict__ nodes,
                                                   SahItem* __restrict__ items,
                                                   int* __restrict__ top,
                                                   unsigned* __restrict__ counts, int cap) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    if (range[2 * i + 1] - range[2 * i] + 1 <= T) return;
    const unsigned ti = atomicAdd(&counts[0], 1u);
    if ((int)ti < cap) top[ti] = i;
    const BvhNode& nd = nodes[i];
    for (int s = 0; s < 2; s++) {
        const int c = child[2 * i + s];
        if (c >= 0 && range[2 * c + 1] - range[2 * c] + 1 > T) continue;  // a top node
        const unsigned k = atomicAdd(&counts[1], 1u);
        if ((int)k >= cap) continue;
        SahItem it;
        it.lo[0] = s ? nd.lx1 : nd.lx0;
        it.hi[0] = s ? nd.hx1 : nd.hx0;
        it.lo[1] = s ? nd.ly1 : nd.ly0;
        it.hi[1] = s ? nd.hy1 : nd.hy0;
        it.lo[2] = s ? nd.lz1 : nd.lz0;
        it.hi[2] = s ? nd.hz1 : nd.hz0;
        it.ref = s ? nd.r1 : nd.r0;
        it.pad = 0;
        for (int q = 0; q < 9; q++) it.tri[q] = nd.tri[9 * s + q];
        it.pad2[0] = it.pad2[1] = it.pad2[2] = 0.f;
        items[k] = it;
    }
}

__global__ __launch_bounds__(kBlock) void k_sah_scatter(const BvhNode* __restrict__ src,
                                                       const int* __restrict__ slot, int m,
                                                       BvhNode* __restrict__ nodes) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < m) nodes[slot[i]] = src[i];
}

// Binned SAH build over the items on the host.  Node k of `out` has children
// a / b: >= 0 another new node, < 0 item ~x.
struct SahBuild {
    const std::vector<SahItem>& it;
    std::vector<int2> kids;
    std::vector<float> box;  // 6 floats per new node: lo xyz, hi xyz
    std::vector<char> wide;  // the node holds an inline triangle

    explicit SahBuild(const std::vector<SahItem>& items) : it(items) {}

    static float area(const float* b) {
        const float dx = std::max(0.f, b[3] - b[0]), dy = std::max(0.f, b[4] - b[1]),
                    dz = std::max(0.f, b[5] - b[2]);
        return dx * dy + dy * dz + dz * dx;
    }
    static void grow(float* b, const float* lo, const float* hi) {
        for (int a = 0; a < 3; a++) {
            b[a] = std::min(b[a], lo[a]);
            b[3 + a] = std::max(b[3 + a], hi[a]);
        }
    }
    static void empty(float* b) {
        b[0] = b[1] = b[2] = 3e38f;
        b[3] = b[4] = b[5] = -3e38f;
    }
    // build over idx[lo, hi); returns the child reference (node or ~item)
    int build(std::vector<int>& idx, int lo, int hi) {
        if (hi - lo == 1) return ~idx[lo];
        float cb[6];
        empty(cb);
        for (int k = lo; k < hi; k++) {
            const SahItem& x = it[idx[k]];
            float c[3];
            for (int a = 0; a < 3; a++) c[a] = 0.5f * (x.lo[a] + x.hi[a]);
            grow(cb, c, c);
        }
        int axis = 0;
        for (int a = 1; a < 3; a++)
            if (cb[3 + a] - cb[a] > cb[3 + axis] - cb[axis]) axis = a;
        const float ext = cb[3 + axis] - cb[axis];
        int mid = lo + (hi - lo) / 2;
        if (ext > 0.f) {
            constexpr int kBins = 16;
            float bb[kBins][6];
            int bc[kBins] = {};
            for (auto& b : bb) empty(b);
            auto bin = [&](const SahItem& x) {
                const float c = 0.5f * (x.lo[axis] + x.hi[axis]);
                int k = (int)((c - cb[axis]) / ext * kBins);
                return std::min(kBins - 1, std::max(0, k));
            };
            for (int k = lo; k < hi; k++) {
                const SahItem& x = it[idx[k]];
                const int b = bin(x);
                grow(bb[b], x.lo, x.hi);
                bc[b]++;
            }
            float best = 3e38f;
            int split = -1;
            float lb[6];
            empty(lb);
            int ln = 0;
            for (int s = 0; s < kBins - 1; s++) {
                grow(lb, bb[s], bb[s] + 3);
                ln += bc[s];
                float rb[6];
                empty(rb);
                int rn = 0;
                for (int t = s + 1; t < kBins; t++) {
                    grow(rb, bb[t], bb[t] + 3);
                    rn += bc[t];
                }
                if (ln == 0 || rn == 0) continue;
                const float cost = area(lb) * ln + area(rb) * rn;
                if (cost < best) {
                    best = cost;
                    split = s;
                }
            }
            if (split >= 0) {
                const auto it_mid = std::partition(
                    idx.begin() + lo, idx.begin() + hi,
                    [&](int q) { return bin(it[q]) <= split; });
                mid = (int)(it_mid - idx.begin());
            } else {
                std::sort(idx.begin() + lo, idx.begin() + hi, [&](int p, int q) {
                    return it[p].lo[axis] + it[p].hi[axis] < it[q].lo[axis] + it[q].hi[axis];
                });
            }
        }
        const int me = (int)kids.size();
        kids.push_back(make_int2(0, 0));
        box.resize(box.size() + 6);
        wide.push_back(0);
        const int a = build(idx, lo, mid);
        const int b = build(idx, mid, hi);
        kids[me] = make_int2(a, b);
        float* mb = &box[(size_t)me * 6];
        empty(mb);
        for (int r : {a, b}) {
            if (r >= 0) {
                const float* cbx = &box[(size_t)r * 6];
                grow(mb, cbx, cbx + 3);
            } else {
                const SahItem& x = it[~r];
                grow(mb, x.lo, x.hi);
                const int code = ~x.ref;
                if (x.ref < 0 && (code >> kLeafFirstBits) == 0) wide[me] = 1;
            }
        }
        return me;
    }
};

}  // namespace

// Rebuilds the top of the binary tree in out.nodes (after k_pack) with the
// SAH; w holds the Karras arrays of the build.  Synchronous on st.
static int lbvh_sah_top(LbvhWork& w, int n, int leaf_size, hipStream_t st, DeviceBvh& out,
                        std::string& err) {
    // trees of fewer than 4096 triangles keep the Morton top
    if (n < 4096) return 0;
    const int T = std::max(std::max(leaf_size, 8), n / 8192);
    const int cap = 2 * (n / T) + 16;  // items, and top nodes (one fewer)
    // scratch in the build's range-union area (96 n bytes or more), unused once
    // the nodes are packed: items + top indices take about 21 n bytes, the new
    // nodes + slots about 33 n bytes
    SahItem* d_items = reinterpret_cast<SahItem*>(w.boxes);
    int* d_top = reinterpret_cast<int*>(d_items + cap);
    unsigned* d_cnt = w.bounds + 12;  // two free words of the bounds block
    HZ_HIP(hipMemsetAsync(d_cnt, 0, 2 * sizeof(unsigned), st), err);
    hipLaunchKernelGGL(k_sah_cut, dim3(grid(n - 1)), dim3(kBlock), 0, st, w.range, w.child, n, T,
                       out.nodes, d_items, d_top, d_cnt, cap);
    HZ_HIP(hipGetLastError(), err);
    unsigned cnt[2] = {0u, 0u};
    HZ_HIP(hipMemcpyAsync(cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, st), err);
    HZ_HIP(hipStreamSynchronize(st), err);
    const int ntop = (int)cnt[0], nitem = (int)cnt[1];
    if (nitem > cap || ntop + 1 != nitem || ntop < 1) {
        err = "SAH top rebuild: inconsistent cut (" + std::to_string(ntop) + " top nodes, " +
              std::to_string(nitem) + " items)";
        return -2;
    }
    std::vector<SahItem> items(nitem);
    std::vector<int> top(ntop);
    HZ_HIP(hipMemcpyAsync(items.data(), d_items, nitem * sizeof(SahItem), hipMemcpyDeviceToHost,
                          st), err);
    HZ_HIP(hipMemcpyAsync(top.data(), d_top, ntop * sizeof(int), hipMemcpyDeviceToHost, st), err);
    HZ_HIP(hipStreamSynchronize(st), err);
    std::vector<int> idx(nitem);
    for (int k = 0; k < nitem; k++) idx[k] = k;
    SahBuild b(items);
    b.kids.reserve(ntop);
    b.build(idx, 0, nitem);
    if ((int)b.kids.size() != ntop) {
        err = "SAH top rebuild: node count mismatch";
        return -2;
    }
    // slots: new node 0 (the root) -> slot 0, the others -> the other old top slots
    std::sort(top.begin(), top.end());
    if (top[0] != 0) {
        err = "SAH top rebuild: the root is not a top node";
        return -2;
    }
    std::vector<BvhNode> nn(ntop);
    for (int k = 0; k < ntop; k++) {
        BvhNode& nd = nn[k];
        std::memset(&nd, 0, sizeof(nd));
        const int kid[2] = {b.kids[k].x, b.kids[k].y};
        float bx[2][6];
        int ref[2];
        for (int s = 0; s < 2; s++) {
            const int r = kid[s];
            if (r >= 0) {
                std::memcpy(bx[s], &b.box[(size_t)r * 6], sizeof(bx[s]));
                ref[s] = top[r] | (b.wide[r] ? kWideBit : 0);
            } else {
                const SahItem& x = items[~r];
                std::memcpy(bx[s], x.lo, 3 * sizeof(float));
                std::memcpy(bx[s] + 3, x.hi, 3 * sizeof(float));
                ref[s] = x.ref;
                const int code = ~x.ref;
                if (x.ref < 0 && (code >> kLeafFirstBits) == 0)
                    std::memcpy(&nd.tri[9 * s], x.tri, 9 * sizeof(float));
            }
        }
        nd.lx0 = bx[0][0];
        nd.lx1 = bx[1][0];
        nd.hx0 = bx[0][3];
        nd.hx1 = bx[1][3];
        nd.ly0 = bx[0][1];
        nd.ly1 = bx[1][1];
        nd.hy0 = bx[0][4];
        nd.hy1 = bx[1][4];
        nd.lz0 = bx[0][2];
        nd.lz1 = bx[1][2];
        nd.hz0 = bx[0][5];
        nd.hz1 = bx[1][5];
        nd.r0 = ref[0];
        nd.r1 = ref[1];
        // the children's leaves are not two leaf-order ranges here: no split
        if (ref[0] >= 0 && ref[1] >= 0) {
            const int none = -1;
            std::memcpy(&nd.tri[1], &none, sizeof(none));
        }
    }
    // upload through the same scratch (items are no longer needed there)
    BvhNode* d_nn = reinterpret_cast<BvhNode*>(w.boxes);
    int* d_slot = reinterpret_cast<int*>(d_nn + ntop);
    HZ_HIP(hipMemcpyAsync(d_nn, nn.data(), ntop * sizeof(BvhNode), hipMemcpyHostToDevice, st),
           err);
    HZ_HIP(hipMemcpyAsync(d_slot, top.data(), ntop * sizeof(int), hipMemcpyHostToDevice, st), err);
    hipLaunchKernelGGL(k_sah_scatter, dim3(grid(ntop)), dim3(kBlock), 0, st, d_nn, d_slot, ntop,
                       out.nodes);
    HZ_HIP(hipGetLastError(), err);
    HZ_HIP(hipStreamSynchronize(st), err);
    return 0;
}

int lbvh_reserve(LbvhWork& w, int n, std::string& err) {
    if (n <= w.capacity) return 0;
    if (w.borrowed) {
        err = "the build work area holds " + std::to_string(w.capacity) + " triangles, " +
              std::to_string(n) + " asked";
        return -1;
    }
    lbvh_free(w);
    int cap = n;
    size_t nn = (size_t)cap;
    HZ_HIP(hipMalloc(&w.keys[0], nn * 8), err);
    HZ_HIP(hipMalloc(&w.keys[1], nn * 8), err);
    HZ_HIP(hipMalloc(&w.vals[0], nn * 4), err);
    HZ_HIP(hipMalloc(&w.vals[1], nn * 4), err);
    HZ_HIP(hipMalloc(&w.aabb_lo, nn * 2 * sizeof(float4)), err);
    HZ_HIP(hipMalloc(&w.aabb_hi, nn * 2 * sizeof(float4)), err);
    HZ_HIP(hipMalloc(&w.child, nn * 2 * 4), err);
    HZ_HIP(hipMalloc(&w.range, nn * 2 * 4), err);
    HZ_HIP(hipMalloc(&w.parent, nn * 2 * 4), err);
    const size_t nb = (nn + kBlock - 1) / kBlock;
    int levels = 1;
    while ((size_t)1 << levels < nb) levels++;
    HZ_HIP(hipMalloc(&w.boxes, (3 * nn + (size_t)(levels + 1) * nb) * 2 * sizeof(float4)), err);
    HZ_HIP(hipMalloc(&w.part, (nb + 1) * 8 * sizeof(float)), err);
    HZ_HIP(hipMalloc(&w.bounds, 16 * 4), err);
    size_t tmp = 0;
    HZ_HIP(rocprim::radix_sort_pairs(nullptr, tmp, w.keys[0], w.keys[1], w.vals[0], w.vals[1],
                                     (unsigned)cap, 0, 64),
           err);
    w.sort_tmp_bytes = tmp;
    HZ_HIP(hipMalloc(&w.sort_tmp, tmp), err);
    size_t stmp = 0;
    HZ_HIP(rocprim::exclusive_scan(nullptr, stmp, w.vals[0], reinterpret_cast<int*>(w.keys[0]), 0,
                                   (size_t)cap, rocprim::plus<int>()),
           err);
    w.scan_tmp_bytes = stmp;
    HZ_HIP(hipMalloc(&w.scan_tmp, stmp), err);
    w.capacity = cap;
    return 0;
}

size_t lbvh_bind(LbvhWork& w, int n, char* base, std::string& err) {
    const size_t nn = (size_t)std::max(n, 1);
    const size_t nb = (nn + kBlock - 1) / kBlock;
    int levels = 1;
    while ((size_t)1 << levels < nb) levels++;
    // the temporary storage the sort and the scan ask for at this size
    size_t tmp = 0, stmp = 0;
    uint64_t* kq = nullptr;
    int* vq = nullptr;
    if (rocprim::radix_sort_pairs(nullptr, tmp, kq, kq, vq, vq, (unsigned)nn, 0, 64) !=
            hipSuccess ||
        rocprim::exclusive_scan(nullptr, stmp, vq, vq, 0, nn, rocprim::plus<int>()) !=
            hipSuccess) {
        err = "rocprim temporary storage query failed";
        return 0;
    }
    const size_t sizes[] = {nn * 8, nn * 8, nn * 4, nn * 4, nn * 2 * sizeof(float4),
                            nn * 2 * sizeof(float4), nn * 2 * 4, nn * 2 * 4, nn * 2 * 4,
                            (3 * nn + (size_t)(levels + 1) * nb) * 2 * sizeof(float4),
                            (nb + 1) * 8 * sizeof(float), 16 * 4, tmp, stmp};
    void** slots[] = {(void**)&w.keys[0], (void**)&w.keys[1], (void**)&w.vals[0],
                      (void**)&w.vals[1], (void**)&w.aabb_lo, (void**)&w.aabb_hi,
                      (void**)&w.child, (void**)&w.range, (void**)&w.parent, &w.boxes,
                      (void**)&w.part, (void**)&w.bounds, &w.sort_tmp, &w.scan_tmp};
    size_t off = 0;
    for (int i = 0; i < 14; i++) {
        if (base) *slots[i] = base + off;
        off += (sizes[i] + 255) & ~(size_t)255;
    }
    if (base) {
        w.sort_tmp_bytes = tmp;
        w.scan_tmp_bytes = stmp;
        w.capacity = (int)nn;
        w.borrowed = true;
    }
    return off;
}

void lbvh_free(LbvhWork& w) {
    if (w.borrowed) {
        w = LbvhWork();
        return;
    }
    void* p[] = {w.keys[0], w.keys[1], w.vals[0], w.vals[1], w.aabb_lo, w.aabb_hi, w.child,
                 w.range, w.parent, w.boxes, w.part, w.bounds, w.sort_tmp, w.scan_tmp};
    for (void* q : p)
        if (q) (void)hipFree(q);
    w = LbvhWork();
}

int lbvh_build(LbvhWork& w, const float* vert, const int* tri, int n, int leaf_size,
               int morton_mode, int width, hipStream_t st, DeviceBvh& out, bool want_depth,
               std::string& err, bool sah_top) {
    out.num_tri = n;
    out.num_nodes = 0;
    out.depth = 0;
    out.leaf_size = std::max(1, std::min(leaf_size, kMaxLeafSize));
    out.width = width == 4 ? 4 : 2;
    out.d_num_nodes = nullptr;
    if (n == 0) return 0;
    if (out.width == 4 && n >= kMaxTriangles4) {
        err = "four-wide trees hold fewer than 2^27 triangles";
        return -1;
    }
    if (lbvh_reserve(w, n, err)) return -3;
    if (leaf_size < 1) leaf_size = 1;
    if (leaf_size > kMaxLeafSize) leaf_size = kMaxLeafSize;
    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, st, w.bounds);
    hipLaunchKernelGGL(k_tri_bounds, dim3(grid(n)), dim3(kBlock), 0, st, vert, tri, n, w.aabb_lo,
                       w.aabb_hi, w.part);
    hipLaunchKernelGGL(k_reduce_bounds, dim3(1), dim3(kBlock), 0, st, w.part, grid(n), w.bounds);
    if (out.width == 2) {
        // binary trees bound z' in z (after the Morton codes below, which read
        // the ENU centroids: the codes of n == 1 are not needed for its node)
        if (n > 1) {
            hipLaunchKernelGGL(k_morton, dim3(grid(n)), dim3(kBlock), 0, st, w.aabb_lo, w.aabb_hi,
                               n, w.bounds, morton_mode, w.keys[0], w.vals[0]);
        }
        hipLaunchKernelGGL(k_tri_zcurv, dim3(grid(n)), dim3(kBlock), 0, st, vert, tri, n,
                           w.aabb_lo, w.aabb_hi);
    }
    if (n == 1) {
        if (out.width == 4) {
            hipLaunchKernelGGL(k_pack4_single, dim3(1), dim3(64), 0, st, w.aabb_lo, w.aabb_hi,
                               vert, tri, reinterpret_cast<Bvh4Node*>(out.nodes), w.bounds + 8);
            out.num_nodes = -1;
            out.d_num_nodes = w.bounds + 8;
        } else {
            hipLaunchKernelGGL(k_pack_single, dim3(1), dim3(64), 0, st, w.aabb_lo, w.aabb_hi,
                               vert, tri, out.nodes);
            out.num_nodes = 1;
        }
        hipLaunchKernelGGL(k_morton, dim3(1), dim3(kBlock), 0, st, w.aabb_lo, w.aabb_hi, n,
                           w.bounds, morton_mode, w.keys[0], w.vals[0]);
        hipLaunchKernelGGL(k_tris, dim3(1), dim3(kBlock), 0, st, w.vals[0], n, vert, tri,
                           out.tris);
        HZ_HIP(hipGetLastError(), err);
        return 0;
    }
    if (out.width == 4)
        hipLaunchKernelGGL(k_morton, dim3(grid(n)), dim3(kBlock), 0, st, w.aabb_lo, w.aabb_hi, n,
                           w.bounds, morton_mode, w.keys[0], w.vals[0]);
    size_t tmp = w.sort_tmp_bytes;
    HZ_HIP(rocprim::radix_sort_pairs(w.sort_tmp, tmp, w.keys[0], w.keys[1], w.vals[0], w.vals[1],
                                     (unsigned)n, 0, 63, st),
           err);
    const uint64_t* keys = w.keys[1];
    const int* sorted = w.vals[1];
    hipLaunchKernelGGL(k_karras, dim3(grid(n - 1)), dim3(kBlock), 0, st, keys, n, w.child,
                       w.range, w.parent);
    {
        const int nb = grid(n);
        Box* boxes = reinterpret_cast<Box*>(w.boxes);
        Box *leaf = boxes, *pre = boxes + n, *suf = boxes + 2 * (size_t)n, *stab = boxes + 3 * (size_t)n;
        hipLaunchKernelGGL(k_block_scan, dim3(nb), dim3(kBlock), 0, st, sorted, n, w.aabb_lo,
                           w.aabb_hi, leaf, pre, suf, stab);
        int j = 1;
        for (; (1 << j) <= nb; j++)
            hipLaunchKernelGGL(k_sparse_level, dim3(grid(nb)), dim3(kBlock), 0, st,
                               stab + (size_t)(j - 1) * nb, nb, 1 << (j - 1), stab + (size_t)j * nb);
        hipLaunchKernelGGL(k_node_boxes, dim3(grid(n - 1)), dim3(kBlock), 0, st, w.range, n, leaf,
                           pre, suf, stab, nb, w.aabb_lo, w.aabb_hi);
    }
    if (out.width == 4) {
        // the sort's input buffers are free now: flags in vals[0], numbering in keys[0]
        int* flag = w.vals[0];
        int* idx = reinterpret_cast<int*>(w.keys[0]);
        hipLaunchKernelGGL(k_b4_flag, dim3(grid(n - 1)), dim3(kBlock), 0, st, w.parent, w.range,
                           n, leaf_size, flag);
        size_t stmp = w.scan_tmp_bytes;
        HZ_HIP(rocprim::exclusive_scan(w.scan_tmp, stmp, flag, idx, 0, (size_t)(n - 1),
                                       rocprim::plus<int>(), st),
               err);
        hipLaunchKernelGGL(k_b4_count, dim3(1), dim3(64), 0, st, flag, idx, n - 1, w.bounds + 8);
        hipLaunchKernelGGL(k_pack4, dim3(grid(n - 1)), dim3(kBlock), 0, st, sorted, n, w.aabb_lo,
                           w.aabb_hi, w.child, w.range, leaf_size, flag, idx, vert, tri,
                           reinterpret_cast<Bvh4Node*>(out.nodes));
        out.num_nodes = -1;
        out.d_num_nodes = w.bounds + 8;
    } else {
        hipLaunchKernelGGL(k_pack, dim3(grid(n - 1)), dim3(kBlock), 0, st, sorted, n, w.aabb_lo,
                           w.aabb_hi, w.child, w.range, leaf_size, vert, tri, out.nodes);
        out.num_nodes = n - 1;
    }
    hipLaunchKernelGGL(k_tris, dim3(grid(n)), dim3(kBlock), 0, st, sorted, n, vert, tri,
                       out.tris);
    if (sah_top && out.width == 2) {
        const int rc = lbvh_sah_top(w, n, leaf_size, st, out, err);
        if (rc) return rc;
    }
    if (want_depth) {
        hipLaunchKernelGGL(k_depth, dim3(grid(n)), dim3(kBlock), 0, st, w.parent, n,
                           w.bounds + 6);
        unsigned d = 0;
        HZ_HIP(hipMemcpyAsync(&d, w.bounds + 6, 4, hipMemcpyDeviceToHost, st), err);
        HZ_HIP(hipStreamSynchronize(st), err);
        out.depth = d > 0 ? (int)d - 1 : 0;
    }
    HZ_HIP(hipGetLastError(), err);
    return 0;
}

int order_reserve(LbvhWork& w, int ncell, std::string& err) { return lbvh_reserve(w, ncell, err); }

int order_cells(LbvhWork& w, const CellFrame* cells, int n, int* order, hipStream_t st,
                std::string& err) {
    if (lbvh_reserve(w, n, err)) return -3;
    hipLaunchKernelGGL(k_cell_bounds, dim3(grid(n)), dim3(kBlock), 0, st, cells, n, w.part);
    hipLaunchKernelGGL(k_reduce_bounds, dim3(1), dim3(kBlock), 0, st, w.part, grid(n), w.bounds);
    hipLaunchKernelGGL(k_cell_morton, dim3(grid(n)), dim3(kBlock), 0, st, cells, n, w.bounds,
                       w.keys[0], w.vals[0]);
    size_t tmp = w.sort_tmp_bytes;
    HZ_HIP(rocprim::radix_sort_pairs(w.sort_tmp, tmp, w.keys[0], w.keys[1], w.vals[0], order,
                                     (unsigned)n, 0, 50, st),
           err);
    HZ_HIP(hipGetLastError(), err);
    return 0;
}

}  // namespace hz
