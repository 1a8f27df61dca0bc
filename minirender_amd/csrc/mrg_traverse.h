// BVH traversal for one ray per lane: four-wide 64 B nodes from HBM, the current node in VGPRs, a
// stack of (child reference, entry distance) pairs in LDS (a kernel's `depth` entries per lane) that spills to a
// per-lane global region for deeper entries (the builder bounds the worst-case stack by kStackTotal).
// Children of a node are visited nearest first (a five-exchange sorting network on entry distance).
// Two drivers: trace_ray runs one ray to its end per call (per-ray loop); trace_step_dyn advances
// a ray by one step (one node, or a wave-wide triangle step over parked leaves) for drivers that
// refill lanes as their rays finish.
//
// Box test: conservative slab test on the quantized child boxes of a node (GpuNodeQ: plane =
// origin + q * scale, exact floats). Per node and axis, A = (origin - o) * (1/d) (two roundings,
// at most 2^-23 |A| off) and B = scale * (1/d) (exact: scale is a power of two); a child's slab
// distance is q * B + A in one FMA, with A moved by 2^-21 |A| toward the ray for near planes and away
// from it for far planes, and the entry/exit distances then widened by 2^-20 relative (the FMA's and
// the reciprocal's rounding). So a box is never culled when the watertight triangle test would accept
// a triangle inside it; closest-hit results therefore do not depend on the tree (nearest t, ties to
// the lower scene-order index; bvh.cpp:273-276).
// Triangle test: the FP32 form of RayQuery::intersect (intersection.cpp:108-152), operation for
// operation, in the render kernels; the ray-query kernels decide hits with its FP64 form (tri_test64).
#pragma once

#include "mrg_layout.h"
#include "mrg_math.h"

namespace mrg {

struct RayPre {
    F3       o, d, inv;
    uint32_t axes;  // the watertight test's axis permutation kx | ky << 2 | kz << 4 (one register)
    float    sx, sy, sz;
    __device__ __forceinline__ int kx() const { return static_cast<int>(axes & 3u); }
    __device__ __forceinline__ int ky() const { return static_cast<int>((axes >> 2) & 3u); }
    __device__ __forceinline__ int kz() const { return static_cast<int>(axes >> 4); }
};

__device__ __forceinline__ RayPre ray_setup(F3 o, F3 d) {
    RayPre r;
    r.o = o;
    r.d = d;
    r.inv.x = d.x != 0.0f ? 1.0f / d.x : __builtin_copysignf(1e30f, d.x);
    r.inv.y = d.y != 0.0f ? 1.0f / d.y : __builtin_copysignf(1e30f, d.y);
    r.inv.z = d.z != 0.0f ? 1.0f / d.z : __builtin_copysignf(1e30f, d.z);
    const F3 ad = vabs(d);
    int kz = ad.x > ad.y ? 0 : 1;
    if (ad.z > (kz == 0 ? ad.x : ad.y)) kz = 2;
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    if (comp(d, kz) < 0.0f) {
        const int t = kx;
        kx          = ky;
        ky          = t;
    }
    r.axes = static_cast<uint32_t>(kx) | static_cast<uint32_t>(ky) << 2 | static_cast<uint32_t>(kz) << 4;
    const float dz = comp(d, kz);
    r.sx = comp(d, kx) / dz;
    r.sy = comp(d, ky) / dz;
    r.sz = 1.0f / dz;
    return r;
}

// One child of a quantized node: near/far plane bytes per axis (byte k of each word), the per-node
// slab terms; returns the widened entry distance or +inf for a miss.
template <int K>
__device__ __forceinline__ float child_enter(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t fx, uint32_t fy, uint32_t fz, float bx,
                                             float by, float bz, float anx, float any_, float anz, float afx, float afy, float afz,
                                             float tmin, float tmax) {
    constexpr uint32_t sh = 8u * K;
    const float n_x = __builtin_fmaf(static_cast<float>((nx >> sh) & 255u), bx, anx);
    const float n_y = __builtin_fmaf(static_cast<float>((ny >> sh) & 255u), by, any_);
    const float n_z = __builtin_fmaf(static_cast<float>((nz >> sh) & 255u), bz, anz);
    const float f_x = __builtin_fmaf(static_cast<float>((fx >> sh) & 255u), bx, afx);
    const float f_y = __builtin_fmaf(static_cast<float>((fy >> sh) & 255u), by, afy);
    const float f_z = __builtin_fmaf(static_cast<float>((fz >> sh) & 255u), bz, afz);
    float tn = __builtin_fmaxf(__builtin_fmaxf(n_x, n_y), __builtin_fmaxf(n_z, tmin));
    float tf = __builtin_fminf(__builtin_fminf(f_x, f_y), __builtin_fminf(f_z, tmax));
    // widen by 2^-20 of the magnitude: one FMA with an |x| source modifier per side
    tn = __builtin_fmaf(-__builtin_fabsf(tn), 0x1p-20f, tn);
    tf = __builtin_fmaf(__builtin_fabsf(tf), 0x1p-20f, tf);
    return tn <= tf ? tn : __builtin_inff();
}

// The four child boxes of node `np` (GpuNodeQ): entry distances (+inf for a miss or an unused slot)
// and child references.
__device__ __forceinline__ void node_boxes(const RayPre& r, const float4* np, float tmin, float tmax, float& t0, float& t1, float& t2,
                                           float& t3, uint32_t& r0, uint32_t& r1, uint32_t& r2, uint32_t& r3) {
    const float4 a  = gld(np + 0);                                     // origin.xyz, scale.x
    const float4 b  = gld(np + 1);                                     // scale.y, scale.z, qlo.x, qlo.y
    const uint4  c  = gld(reinterpret_cast<const uint4*>(np) + 2);     // qlo.z, qhi.x, qhi.y, qhi.z
    const uint4  rf = gld(reinterpret_cast<const uint4*>(np) + 3);
    const float  ax = (a.x - r.o.x) * r.inv.x, ay = (a.y - r.o.y) * r.inv.y, az = (a.z - r.o.z) * r.inv.z;
    const float  bx = a.w * r.inv.x, by = b.x * r.inv.y, bz = b.y * r.inv.z;
    const float  anx = __builtin_fmaf(-__builtin_fabsf(ax), 0x1p-21f, ax), afx = __builtin_fmaf(__builtin_fabsf(ax), 0x1p-21f, ax);
    const float  any_ = __builtin_fmaf(-__builtin_fabsf(ay), 0x1p-21f, ay), afy = __builtin_fmaf(__builtin_fabsf(ay), 0x1p-21f, ay);
    const float  anz = __builtin_fmaf(-__builtin_fabsf(az), 0x1p-21f, az), afz = __builtin_fmaf(__builtin_fabsf(az), 0x1p-21f, az);
    // near planes are the low ones along a positive direction component, the high ones along a negative one
    const uint32_t lox = __float_as_uint(b.z), loy = __float_as_uint(b.w), loz = c.x;
    const bool     sx = r.inv.x < 0.0f, sy = r.inv.y < 0.0f, sz = r.inv.z < 0.0f;
    const uint32_t nx = sx ? c.y : lox, fx = sx ? lox : c.y;
    const uint32_t ny = sy ? c.z : loy, fy = sy ? loy : c.z;
    const uint32_t nz = sz ? c.w : loz, fz = sz ? loz : c.w;
    t0 = child_enter<0>(nx, ny, nz, fx, fy, fz, bx, by, bz, anx, any_, anz, afx, afy, afz, tmin, tmax);
    t1 = child_enter<1>(nx, ny, nz, fx, fy, fz, bx, by, bz, anx, any_, anz, afx, afy, afz, tmin, tmax);
    t2 = child_enter<2>(nx, ny, nz, fx, fy, fz, bx, by, bz, anx, any_, anz, afx, afy, afz, tmin, tmax);
    t3 = child_enter<3>(nx, ny, nz, fx, fy, fz, bx, by, bz, anx, any_, anz, afx, afy, afz, tmin, tmax);
    r0 = rf.x;
    r1 = rf.y;
    r2 = rf.z;
    r3 = rf.w;
    if (r1 == kEmptyRef) t1 = __builtin_inff();
    if (r2 == kEmptyRef) t2 = __builtin_inff();
    if (r3 == kEmptyRef) t3 = __builtin_inff();
}

__device__ __forceinline__ float dop(float a, float b, float c, float d) {
    const float cd   = c * d;
    const float corr = __builtin_fmaf(-c, d, cd);
    return __builtin_fmaf(a, b, -cd) + corr;
}

// A candidate hit: distance and the three edge functions with their sum (the barycentrics are
// a / det, b / det, c / det; the divisions wait until the ray's closest hit is known).
struct TriHit {
    float t, a, b, c, det;
};

__device__ __forceinline__ void permute(const RayPre& r, F3 v, float& px, float& py, float& pz) {
    const F3    t  = v - r.o;
    const float tx = comp(t, r.kx()), ty = comp(t, r.ky()), tz = comp(t, r.kz());
    px = tx - r.sx * tz;
    py = ty - r.sy * tz;
    pz = tz * r.sz;
}

// FP32 watertight test with the reference's error-bound rejection. Returns false on a miss.
__device__ __forceinline__ bool tri_test(const RayPre& r, F3 v0, F3 v1, F3 v2, float tmin, float tmax, TriHit& out) {
    float p0x, p0y, p0z, p1x, p1y, p1z, p2x, p2y, p2z;
    permute(r, v0, p0x, p0y, p0z);
    permute(r, v1, p1x, p1y, p1z);
    permute(r, v2, p2x, p2y, p2z);
    const float a = dop(p1x, p2y, p1y, p2x);
    const float b = dop(p2x, p0y, p2y, p0x);
    const float c = dop(p0x, p1y, p0y, p1x);
    if ((a < 0.0f || b < 0.0f || c < 0.0f) && (a > 0.0f || b > 0.0f || c > 0.0f)) return false;
    const float det = (a + b) + c;
    if (det == 0.0f) return false;
    const float t = ((a * p0z + b * p1z) + c * p2z) / det;
    if (t < tmin || t > tmax) return false;
    constexpr float g2 = (2.0f * 0x1p-24f) / (1.0f - 2.0f * 0x1p-24f);
    constexpr float g3 = (3.0f * 0x1p-24f) / (1.0f - 3.0f * 0x1p-24f);
    constexpr float g5 = (5.0f * 0x1p-24f) / (1.0f - 5.0f * 0x1p-24f);
    const float x    = max3(__builtin_fabsf(p0x), __builtin_fabsf(p1x), __builtin_fabsf(p2x));
    const float y    = max3(__builtin_fabsf(p0y), __builtin_fabsf(p1y), __builtin_fabsf(p2y));
    const float z    = max3(__builtin_fabsf(p0z), __builtin_fabsf(p1z), __builtin_fabsf(p2z));
    const float ex   = g5 * (x + z);
    const float ey   = g5 * (y + z);
    const float ee   = 2.0f * ((g2 * x * y + ey * x) + ex * y);
    const float edge = max3(__builtin_fabsf(a), __builtin_fabsf(b), __builtin_fabsf(c));
    const float et   = 3.0f * ((g3 * edge * z + ee * z) + g3 * z * edge) / __builtin_fabsf(det);
    if (t <= et) return false;
    out.t   = t;
    out.a   = a;
    out.b   = b;
    out.c   = c;
    out.det = det;
    return true;
}

// ---- FP64 hit decisions for ray queries
// The reference's RayQuery::intersect (intersection.cpp:108-152) in double on the float ray and
// vertices widened to double, operation for operation as the FP64 oracle computes it, so a query's
// hit ids, distances and barycentrics are the FP64 intersector's bit for bit. Traversal stays in
// float: the slab test is conservative (it never culls a box holding a hit the FP64 test accepts at a
// distance within the float bound, which is rounded up from the FP64 best hit). Every triangle the
// traversal reaches goes to the FP64 test: a float pre-test that rejected provable FP64 misses first
// measured slower (in a triangle step some lane almost always needs the FP64 test, so the wave runs it
// anyway and the pre-test only adds instructions).
struct TriHit64 {
    double t, a, b, c, det;
};

__device__ __forceinline__ double dop64(double a, double b, double c, double d) {
    const double cd   = c * d;
    const double corr = __builtin_fma(-c, d, cd);
    return __builtin_fma(a, b, -cd) + corr;
}

__device__ __forceinline__ double comp64(double x, double y, double z, int i) { return i == 0 ? x : (i == 1 ? y : z); }

__device__ __forceinline__ bool tri_test64_body(const RayPre& r, F3 v0, F3 v1, F3 v2, double tmin, double tmax, TriHit64& out) {
    const int    kx = r.kx(), ky = r.ky(), kz = r.kz();
    const double dkz = comp(r.d, kz);
    const double sx = double(comp(r.d, kx)) / dkz, sy = double(comp(r.d, ky)) / dkz, sz = 1.0 / dkz;
    double p[3][3];
    const F3 v[3] = {v0, v1, v2};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double tx = double(v[i].x) - double(r.o.x), ty = double(v[i].y) - double(r.o.y), tz = double(v[i].z) - double(r.o.z);
        const double c  = comp64(tx, ty, tz, kz);
        p[i][0] = comp64(tx, ty, tz, kx) - sx * c;
        p[i][1] = comp64(tx, ty, tz, ky) - sy * c;
        p[i][2] = c * sz;
    }
    const double a = dop64(p[1][0], p[2][1], p[1][1], p[2][0]);
    const double b = dop64(p[2][0], p[0][1], p[2][1], p[0][0]);
    const double c = dop64(p[0][0], p[1][1], p[0][1], p[1][0]);
    if ((a < 0.0 || b < 0.0 || c < 0.0) && (a > 0.0 || b > 0.0 || c > 0.0)) return false;
    const double det = (a + b) + c;
    if (det == 0.0) return false;
    const double t = ((a * p[0][2] + b * p[1][2]) + c * p[2][2]) / det;
    if (t < tmin || t > tmax) return false;
    // gamma_bound(n) = n (eps / 2) / (1 - n (eps / 2)) with the double epsilon
    constexpr double e2 = 2.0 * 0x1p-53, e3 = 3.0 * 0x1p-53, e5 = 5.0 * 0x1p-53;
    constexpr double g2 = e2 / (1.0 - e2), g3 = e3 / (1.0 - e3), g5 = e5 / (1.0 - e5);
    const double x    = __builtin_fmax(__builtin_fmax(__builtin_fabs(p[0][0]), __builtin_fabs(p[1][0])), __builtin_fabs(p[2][0]));
    const double y    = __builtin_fmax(__builtin_fmax(__builtin_fabs(p[0][1]), __builtin_fabs(p[1][1])), __builtin_fabs(p[2][1]));
    const double z    = __builtin_fmax(__builtin_fmax(__builtin_fabs(p[0][2]), __builtin_fabs(p[1][2])), __builtin_fabs(p[2][2]));
    const double ex   = g5 * (x + z);
    const double ey   = g5 * (y + z);
    const double ee   = 2.0 * ((g2 * x * y + ey * x) + ex * y);
    const double edge = __builtin_fmax(__builtin_fmax(__builtin_fabs(a), __builtin_fabs(b)), __builtin_fabs(c));
    const double et   = 3.0 * ((g3 * edge * z + ee * z) + g3 * z * edge) / __builtin_fabs(det);
    if (t <= et) return false;
    out.t   = t;
    out.a   = a;
    out.b   = b;
    out.c   = c;
    out.det = det;
    return true;
}

// Out of line, with its inputs by value: the call keeps the FP64 temporaries out of the register budget
// of the traversal loop (inlined, they raise it by about 30 VGPRs); it runs for every triangle a
// query ray reaches (4.8 per ray on the icosphere microbench).
__device__ __attribute__((noinline)) TriHit64 tri_test64_call(float ox, float oy, float oz, float dx, float dy, float dz, uint32_t axes,
                                                              float v0x, float v0y, float v0z, float v1x, float v1y, float v1z,
                                                              float v2x, float v2y, float v2z, double tmin, double tmax) {
    RayPre r;
    r.o    = mk(ox, oy, oz);
    r.d    = mk(dx, dy, dz);
    r.axes = axes;
    TriHit64 out;
    if (!tri_test64_body(r, mk(v0x, v0y, v0z), mk(v1x, v1y, v1z), mk(v2x, v2y, v2z), tmin, tmax, out)) out.det = 0.0;  // det == 0: a miss
    return out;
}

__device__ __forceinline__ bool tri_test64(const RayPre& r, F3 v0, F3 v1, F3 v2, double tmin, double tmax, TriHit64& out) {
    out = tri_test64_call(r.o.x, r.o.y, r.o.z, r.d.x, r.d.y, r.d.z, r.axes, v0.x, v0.y, v0.z, v1.x, v1.y, v1.z, v2.x, v2.y, v2.z, tmin,
                          tmax);
    return out.det != 0.0;
}

// Single-sided culling of the FP64 traversal (bvh.cpp:259-271): true when the ray reaches the front
// of the triangle (dot(d, outward) < 0), outward = (v1 - v0) x (v2 - v0), negated for a mirrored instance.
__device__ __forceinline__ bool front_side64(const RayPre& r, F3 v0, F3 v1, F3 v2, bool mirrored) {
    const double e1x = double(v1.x) - double(v0.x), e1y = double(v1.y) - double(v0.y), e1z = double(v1.z) - double(v0.z);
    const double e2x = double(v2.x) - double(v0.x), e2y = double(v2.y) - double(v0.y), e2z = double(v2.z) - double(v0.z);
    const double s   = mirrored ? -1.0 : 1.0;
    const double ox = (e1y * e2z - e2y * e1z) * s, oy = (e1z * e2x - e2z * e1x) * s, oz = (e1x * e2y - e2x * e1y) * s;
    return ((double(r.d.x) * ox + double(r.d.y) * oy) + double(r.d.z) * oz) < 0.0;
}

// Smallest float >= x (x >= 0): the float cull bound of an FP64 distance.
__device__ __forceinline__ float float_at_least(double x) {
    const float f = static_cast<float>(x);
    return static_cast<double>(f) < x ? __uint_as_float(__float_as_uint(f) + 1u) : f;
}

// Triangle record flags (GpuTri::flags): bit 0 single-sided, bit 1 mirrored instance.
constexpr uint32_t kTriSingleSided = 1u;
constexpr uint32_t kTriMirrored    = 2u;

// Every traversal kernel runs 256-thread blocks: the LDS stack row pitch is a compile-time
// constant, so an entry's address is a shift and an add (no per-push integer multiply).
constexpr uint32_t kTraversalBlock = 256;

struct StackView {
    uint32_t* lds_ref;   // [depth][kTraversalBlock]
    float*    lds_t;
    uint32_t* spill_ref; // [kStackTotal - depth][lanes]
    float*    spill_t;
    uint32_t  lane;      // threadIdx.x
    uint32_t  glane;     // global lane id (spill column)
    uint32_t  lanes;     // total lanes of the grid (spill row pitch)
    uint32_t  depth;     // entries per lane in LDS (kernel-uniform)
};

__device__ __forceinline__ void stack_push(const StackView& s, uint32_t& sp, uint32_t ref, float t) {
    if (sp >= kStackTotal) return;  // unreachable: the builder bounds the worst-case stack by kStackTotal
    if (sp < s.depth) {
        s.lds_ref[sp * kTraversalBlock + s.lane] = ref;
        s.lds_t[sp * kTraversalBlock + s.lane]   = t;
    } else {
        const uint32_t k = (sp - s.depth) * s.lanes + s.glane;
        gst(s.spill_ref + k, ref);
        gst(s.spill_t + k, t);
    }
    ++sp;
}

__device__ __forceinline__ void stack_pop(const StackView& s, uint32_t& sp, uint32_t& ref, float& t) {
    --sp;
    if (sp < s.depth) {
        ref = s.lds_ref[sp * kTraversalBlock + s.lane];
        t   = s.lds_t[sp * kTraversalBlock + s.lane];
    } else {
        const uint32_t k = (sp - s.depth) * s.lanes + s.glane;
        ref              = gld(s.spill_ref + k);
        t                = gld(s.spill_t + k);
    }
}

__device__ __forceinline__ void sort2(float& ta, uint32_t& ra, float& tb, uint32_t& rb) {
    const bool     swap = tb < ta;
    const float    t    = swap ? tb : ta;
    const uint32_t q    = swap ? rb : ra;
    tb                  = swap ? ta : tb;
    rb                  = swap ? ra : rb;
    ta                  = t;
    ra                  = q;
}

struct TraceResult {
    float    t, b0, b1, b2;  // during traversal b0..b2 hold the edge functions of the best hit
    float    det;            // their sum; finish_barycentrics divides once at the end
    uint32_t slot;           // leaf-order triangle slot, 0xffffffff on a miss
    uint32_t prim;           // scene-order primitive index
};

__device__ __forceinline__ void finish_barycentrics(TraceResult& res) {
    if (res.slot == 0xffffffffu) return;
    res.b0 = res.b0 / res.det;
    res.b1 = res.b1 / res.det;
    res.b2 = res.b2 / res.det;
}

// The pointers traversal needs from a scene.
struct TraceScene {
    const GpuNodeQ* nodes;
    const GpuTri*   tris;
    uint32_t        node_count;
};

__device__ __forceinline__ TraceScene trace_scene(const GpuScene& sc) { return TraceScene{sc.nodes, sc.tris, sc.node_count}; }

// Per-lane traversal state of the step driver: the ray, the best hit so far, the next node to
// visit, a parked leaf and the stack depth (the stack itself lives in LDS / the spill area of this
// lane). `ref` is an inner node, kEmptyRef (nothing to visit), or a leaf waiting for the park slot.
struct TraceState {
    RayPre      r;
    float       tmin;
    TraceResult res;
    uint32_t    ref;
    uint32_t    leaf;  // parked leaf, kEmptyRef when none (kEmptyRef is never a leaf of a scene below 2^27 triangles)
    uint32_t    sp;
    double      best64;  // FP64 hit decisions (EXACT): the best hit's FP64 distance; res.t is the float at least it
};

// `skip`: for an any-hit ray, a scene-order primitive to ignore (the emitter a shadow segment ends
// on); it waits in res.prim, which an any-hit ray only overwrites with the hit that finishes it.
__device__ __forceinline__ void trace_init(TraceState& st, F3 o, F3 d, float tmin, float tmax, uint32_t skip = 0xffffffffu) {
    st.r        = ray_setup(o, d);
    st.tmin     = tmin;
    st.res.t    = tmax;
    st.res.slot = 0xffffffffu;
    st.res.prim = skip;
    st.res.b0 = st.res.b1 = st.res.b2 = 0.0f;
    st.res.det = 1.0f;
    st.ref  = 0;  // the root is always an inner node
    st.leaf = kEmptyRef;
    st.sp   = 0;
    st.best64 = tmax;
}

// Leaf batching of the step driver: a lane that reaches a leaf parks it (one parked leaf per lane)
// and goes on with the inner nodes of its stack; the wave tests triangles when at least LEAF_BATCH
// lanes hold a parked leaf, or when the lanes that can only wait for a triangle step (a parked leaf
// and no inner node to visit) are at least as many as those with an inner node. Each step is then
// either a box step or a triangle step for the whole wave, and triangle tests run with many lanes
// at once instead of the few lanes that happen to reach a leaf in the same step. Neither result
// depends on the order of the leaves: the closest hit is the nearest t with ties to the lower
// scene-order index, and the any-hit result is only whether a hit exists; a box tested before a
// parked leaf has lowered the closest distance is tested against a larger bound, which costs work,
// never a hit.

// Where a closest-hit ray keeps the edge functions of its best hit so far: in its TraceResult
// (KeepInResult), or wherever a caller's functor puts them (the render traversal writes them to the
// path's hit record at once, so they hold no registers while the ray goes on).
struct KeepInResult {
    __device__ __forceinline__ void operator()(TraceResult& res, const TriHit& h) const {
        res.b0  = h.a;
        res.b1  = h.b;
        res.b2  = h.c;
        res.det = h.det;
    }
};

// One traversal step (all lanes of the wave that trace call it together): the parked leaves' triangles
// or one inner node per lane, then the lane parks a leaf it reached and takes stack entries until it
// has an inner node to visit. Returns true once the ray is finished: closest hit (any = false) or any
// hit (any = true) in [tmin, tmax], result in st.res; a closest hit's edge functions go to `keep`
// undivided (with KeepInResult the caller applies finish_barycentrics where it needs them). `any` may differ between the lanes of a
// wave. EXACT: hit decisions in FP64 (tri_test64; st.best64 holds the best distance, `keep` is not
// called: the caller recomputes the winner's FP64 hit where it needs its barycentrics).
template <int LEAF_BATCH, bool STATS, bool EXACT, class Keep>
__device__ __forceinline__ bool trace_step_dyn(const TraceScene& sc, TraceState& st, const StackView& stk, uint64_t& bounds_tests,
                                               uint64_t& tri_tests, const bool any, const Keep& keep) {
    if (sc.node_count == 0) return true;
    const RayPre& r          = st.r;
    const bool    parked     = st.leaf != kEmptyRef;
    const bool    node       = st.ref != kEmptyRef && !is_leaf(st.ref);
    const int     n_parked   = __builtin_popcountll(__ballot(parked));
    const int     n_node     = __builtin_popcountll(__ballot(node));
    const int     n_stuck    = __builtin_popcountll(__ballot(parked && !node));
    const bool    leaf_phase = n_parked >= LEAF_BATCH || n_stuck >= n_node;
    if (leaf_phase) {
        if (parked) {
            const uint32_t first = leaf_first(st.leaf), count = leaf_count(st.leaf);
            st.leaf = kEmptyRef;
            for (uint32_t k = 0; k < count; ++k) {
                const uint32_t slot = first + k;
                const float4*  tp   = reinterpret_cast<const float4*>(sc.tris + slot);
                const float4   a    = gld(tp + 0);
                const float4   b    = gld(tp + 1);
                const float4   c    = gld(tp + 2);
                if (STATS) tri_tests += 1;
                const F3 v0 = mk(a.x, a.y, a.z), v1 = mk(b.x, b.y, b.z), v2 = mk(c.x, c.y, c.z);
                if (EXACT) {
                    TriHit64 h;
                    if (!tri_test64(r, v0, v1, v2, st.tmin, st.best64, h)) continue;
                    const uint32_t flags = __float_as_uint(b.w);
                    if ((flags & kTriSingleSided) && !front_side64(r, v0, v1, v2, (flags & kTriMirrored) != 0u)) continue;
                    const uint32_t prim = __float_as_uint(a.w);
                    if (any && prim == st.res.prim) continue;
                    if (st.res.slot == 0xffffffffu || h.t < st.best64 || (h.t == st.best64 && prim < st.res.prim)) {
                        st.best64   = h.t;
                        st.res.t    = float_at_least(h.t);
                        st.res.slot = slot;
                        st.res.prim = prim;
                        if (any) return true;
                    }
                    continue;
                }
                TriHit h;
                if (!tri_test(r, v0, v1, v2, st.tmin, st.res.t, h)) continue;
                const uint32_t flags = __float_as_uint(b.w);
                if (flags & kTriSingleSided) {
                    const F3 e1      = v1 - v0;
                    const F3 e2      = v2 - v0;
                    F3       outward = cross(e1, e2) * ((flags & kTriMirrored) ? -1.0f : 1.0f);
                    if (dot(r.d, outward) >= 0.0f) continue;
                }
                const uint32_t prim = __float_as_uint(a.w);
                if (any && prim == st.res.prim) continue;
                if (st.res.slot == 0xffffffffu || h.t < st.res.t || (h.t == st.res.t && prim < st.res.prim)) {
                    st.res.t    = h.t;
                    st.res.slot = slot;
                    st.res.prim = prim;
                    if (any) return true;
                    keep(st.res, h);
                }
            }
        }
    } else if (node) {
        float    t0, t1, t2, t3;
        uint32_t r0, r1, r2, r3;
        node_boxes(r, reinterpret_cast<const float4*>(sc.nodes + st.ref), st.tmin, st.res.t, t0, t1, t2, t3, r0, r1, r2, r3);
        if (STATS) bounds_tests += 4;
        // sorting network on (t, ref): nearest first, misses (+inf) last
        sort2(t0, r0, t1, r1);
        sort2(t2, r2, t3, r3);
        sort2(t0, r0, t2, r2);
        sort2(t1, r1, t3, r3);
        sort2(t1, r1, t2, r2);
        st.ref = kEmptyRef;
        if (t0 != __builtin_inff()) {
            if (t3 != __builtin_inff()) stack_push(stk, st.sp, r3, t3);
            if (t2 != __builtin_inff()) stack_push(stk, st.sp, r2, t2);
            if (t1 != __builtin_inff()) stack_push(stk, st.sp, r1, t1);
            st.ref = r0;
        }
    }
    // Take stack entries (skipping those that start beyond the current closest hit) while ref holds
    // nothing, or a leaf that the free park slot can take. kEmptyRef carries the leaf bit, so both
    // cases read "ref is a leaf, and ref is empty or the slot is free"; parking is a select, since
    // moving an empty ref into a free slot leaves it free.
    for (;;) {
        // non-short-circuit operators: one mask computation instead of nested branches
        if (!((st.sp != 0) & is_leaf(st.ref) & ((st.ref == kEmptyRef) | (st.leaf == kEmptyRef)))) break;
        st.leaf = st.leaf == kEmptyRef ? st.ref : st.leaf;
        float    t;
        uint32_t ref;
        stack_pop(stk, st.sp, ref, t);
        st.ref = t > st.res.t ? kEmptyRef : ref;
    }
    if (is_leaf(st.ref) && st.leaf == kEmptyRef) {  // the stack ran out with a leaf in ref
        st.leaf = st.ref;
        st.ref  = kEmptyRef;
    }
    return st.ref == kEmptyRef && st.leaf == kEmptyRef;
}

// Closest hit (ANY = false) or any hit (ANY = true) in [tmin, tmax]; an any-hit query ignores the
// scene-order primitive `skip`.
template <bool ANY, bool STATS>
__device__ __forceinline__ TraceResult trace_ray(const GpuScene& sc, const RayPre& r, float tmin, float tmax,
                                                 const StackView& stk, uint64_t* bounds_tests, uint64_t* tri_tests,
                                                 uint32_t skip = 0xffffffffu) {
    TraceResult res;
    res.t    = tmax;
    res.slot = 0xffffffffu;
    res.prim = ANY ? skip : 0xffffffffu;
    res.b0 = res.b1 = res.b2 = 0.0f;
    res.det = 1.0f;
    if (sc.node_count == 0) return res;
    uint32_t sp  = 0;
    uint32_t ref = 0;
    for (;;) {
        if (!is_leaf(ref)) {
            float    t0, t1, t2, t3;
            uint32_t r0, r1, r2, r3;
            node_boxes(r, reinterpret_cast<const float4*>(sc.nodes + ref), tmin, res.t, t0, t1, t2, t3, r0, r1, r2, r3);
            if (STATS) *bounds_tests += 4;
            // sorting network on (t, ref): nearest first, misses (+inf) last
            sort2(t0, r0, t1, r1);
            sort2(t2, r2, t3, r3);
            sort2(t0, r0, t2, r2);
            sort2(t1, r1, t3, r3);
            sort2(t1, r1, t2, r2);
            if (t0 != __builtin_inff()) {
                if (t3 != __builtin_inff()) stack_push(stk, sp, r3, t3);
                if (t2 != __builtin_inff()) stack_push(stk, sp, r2, t2);
                if (t1 != __builtin_inff()) stack_push(stk, sp, r1, t1);
                ref = r0;
                continue;
            }
        } else {
            const uint32_t first = leaf_first(ref), count = leaf_count(ref);
            for (uint32_t k = 0; k < count; ++k) {
                const uint32_t slot = first + k;
                const float4*  tp   = reinterpret_cast<const float4*>(sc.tris + slot);
                const float4   a    = gld(tp + 0);
                const float4   b    = gld(tp + 1);
                const float4   c    = gld(tp + 2);
                if (STATS) *tri_tests += 1;
                TriHit h;
                const F3 v0 = mk(a.x, a.y, a.z), v1 = mk(b.x, b.y, b.z), v2 = mk(c.x, c.y, c.z);
                if (!tri_test(r, v0, v1, v2, tmin, res.t, h)) continue;
                const uint32_t flags = __float_as_uint(b.w);
                if (flags & kTriSingleSided) {
                    const F3 e1      = v1 - v0;
                    const F3 e2      = v2 - v0;
                    F3       outward = cross(e1, e2) * ((flags & kTriMirrored) ? -1.0f : 1.0f);
                    if (dot(r.d, outward) >= 0.0f) continue;
                }
                const uint32_t prim = __float_as_uint(a.w);
                if (ANY && prim == res.prim) continue;
                if (res.slot == 0xffffffffu || h.t < res.t || (h.t == res.t && prim < res.prim)) {
                    res.t    = h.t;
                    res.b0   = h.a;
                    res.b1   = h.b;
                    res.b2   = h.c;
                    res.det  = h.det;
                    res.slot = slot;
                    res.prim = prim;
                    if (ANY) {
                        finish_barycentrics(res);
                        return res;
                    }
                }
            }
        }
        // pop, skipping entries that start beyond the current closest hit
        for (;;) {
            if (sp == 0) {
                finish_barycentrics(res);
                return res;
            }
            float t;
            stack_pop(stk, sp, ref, t);
            if (!(t > res.t)) break;
        }
    }
}

}  // namespace mrg
