// pale_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A sequential CPU restatement of the reference renderer's hot path
// (M-Gjerde/DifferentiablePointRendering, "Pale"): scene build with the
// median-split BVH, TLAS, light CDF, the photon-mapping forward pass, the
// photon grid, the camera gather, post-processing and the opacity adjoint.
// It is the checker of the HIP implementation and the CPU baseline of the
// benchmark.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg may load it; the product never does.
//
// Execution order is the reference's sequential item order: every atomic
// append slot of the reference equals the item's rank (PrimalKernels.cpp:57,
// 152, 181, 249, 311; KernelHelpers.h:1702, 1741), which is what a
// one-work-item-at-a-time SYCL device would produce.  This is the only
// well-defined "reference RNG sequence" (the reference seeds per-item RNGs
// from queue slots).
//
// Floating point: IEEE-754 binary32 with no contraction (built with
// -ffp-contract=off).  normalize() is v * (1 / sqrt(|v|^2)) with correctly
// rounded sqrt and division where the reference calls sycl::rsqrt
// (GPUDataTypes.h:350-356); sin/cos/exp/pow are evaluated in binary64 and
// rounded to binary32 where the reference calls the SYCL float versions,
// whose ulp behaviour is implementation-defined (parity at ulp level is not
// pinned by any reference test).
//
// Build-defined semantics where the reference is undefined:
//  * A NullTransmittance gradient event whose endpoint is a surfel reads the
//    material of a point-cloud instance, which is out of bounds in the
//    reference (AdjointKernels.cpp:446-447 with SceneBuild.cpp:246).  Here
//    that endpoint is shaded as a Lambertian with reflectance
//    alpha_r * albedo of the endpoint surfel and no emission.
//  * An empty photon map (backward before any forward) gathers zero.
//  * The shape gradient, never written by the reference
//    (Sensors.ixx:150-152), is returned as zeros.

#include "../include/pale_abi.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include <omp.h>

namespace orc {

// ------------------------------------------------------------------ math --
struct F3 {
    float x, y, z;
};

static inline F3 mk(float x, float y, float z) { return F3{x, y, z}; }
static inline F3 operator+(F3 a, F3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline F3 operator-(F3 a, F3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline F3 operator*(F3 a, F3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline F3 operator*(F3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline F3 operator*(float s, F3 a) { return a * s; }
static inline F3 operator-(F3 a) { return mk(-a.x, -a.y, -a.z); }
// GPUDataTypes.h:303-306: division by a scalar multiplies by its reciprocal.
static inline F3 divs(F3 a, float s) {
    const float inv = 1.0f / s;
    return mk(a.x * inv, a.y * inv, a.z * inv);
}
// GPUDataTypes.h:287-293: component-wise division.
static inline F3 divv(F3 a, F3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline F3 cross(F3 a, F3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// GPUDataTypes.h:350-356 (rsqrt restated as 1/sqrt, see header).
static inline F3 normalize(F3 v) {
    const float sum = v.x * v.x + v.y * v.y + v.z * v.z;
    const float inv = 1.0f / std::sqrt(sum);
    return v * inv;
}
static inline float length(F3 v) { return std::sqrt(dot(v, v)); }
static inline F3 vmin(F3 a, F3 b) { return mk(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)); }
static inline F3 vmax(F3 a, F3 b) { return mk(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)); }
static inline float comp(F3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
static inline float fsum(F3 v) { return v.x + v.y + v.z; }  // GPUDataTypes.h:358-360

static inline float sin_f(float x) { return (float)std::sin((double)x); }
static inline float cos_f(float x) { return (float)std::cos((double)x); }
static inline float exp_f(float x) { return (float)std::exp((double)x); }
static inline float pow_f(float b, float e) { return (float)std::pow((double)b, (double)e); }

static const float kPi = 3.14159265358979323846f;       // M_PIf
static const float kInvPi = 0.318309886183790671538f;   // M_1_PIf
static const float kInf = std::numeric_limits<float>::infinity();
static const float kFltMax = std::numeric_limits<float>::max();
static const uint32_t kInvalid = 0xFFFFFFFFu;

struct M4 {
    float m[16];  // row-major
};
// Matrix x vec4 as row dot products (GPUDataTypes.h:48-59).
static inline void mulv(const M4& M, float x, float y, float z, float w, float out[4]) {
    for (int i = 0; i < 4; ++i) {
        const float* r = &M.m[4 * i];
        out[i] = r[0] * x + r[1] * y + r[2] * z + r[3] * w;
    }
}
// KernelHelpers.h:240-246
static inline F3 toWorldPoint(F3 p, const M4& o2w) {
    float h[4];
    mulv(o2w, p.x, p.y, p.z, 1.0f, h);
    return divs(mk(h[0], h[1], h[2]), h[3]);
}
// KernelHelpers.h:229-234
static inline F3 transformPoint(const M4& M, F3 p) {
    float h[4];
    mulv(M, p.x, p.y, p.z, 1.0f, h);
    const float invW = 1.0f / h[3];
    return mk(h[0] * invW, h[1] * invW, h[2] * invW);
}
// KernelHelpers.h:235-239
static inline F3 transformDirection(const M4& M, F3 d) {
    float h[4];
    mulv(M, d.x, d.y, d.z, 0.0f, h);
    return normalize(mk(h[0], h[1], h[2]));
}

// ------------------------------------------------------------------- rng --
// KernelHelpers.h:12-24
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t v) : s(v) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
};
// KernelHelpers.h:73-106
struct Xorshift128 {
    uint32_t x, y, z, w;
    explicit Xorshift128(uint64_t seed) {
        SplitMix64 sm(seed);
        x = (uint32_t)sm.next();
        y = (uint32_t)sm.next();
        z = (uint32_t)sm.next();
        w = (uint32_t)sm.next();
        if (!(x | y | z | w)) w = 0xA511E9B3u;
    }
    uint32_t nextUint() {
        uint32_t t = x ^ (x << 11);
        x = y;
        y = z;
        z = w;
        w = (w ^ (w >> 19)) ^ (t ^ (t >> 8));
        return w;
    }
    float nextFloat() { return (float)(nextUint() >> 8) * 0x1.0p-24f; }
};
// KernelHelpers.h:109-115
static inline uint64_t perItemSeed(uint64_t base, uint64_t id) {
    SplitMix64 sm(base ^ (0x9E3779B97F4A7C15ull + id));
    return sm.next();
}

// --------------------------------------------------------------- scene --
struct Ray {
    F3 o, d;
};
struct Tri {
    uint32_t v0, v1, v2;
    F3 centroid;
};
struct Node {
    F3 mn, mx;
    uint32_t leftFirst, count;
};
struct TNode {
    F3 mn, mx;
    uint32_t left, count, right;
};
struct Range {
    uint32_t first, count;
};
enum GeoType { kMesh = 0, kCloud = 1 };
struct Instance {
    int type;
    uint32_t geom, material, xf, blas;
};
struct Material {
    F3 base;
    float power;
};
struct Xf {
    M4 o2w, w2o;
};
struct Light {
    uint32_t geom, xf, mat, triOffset, triCount;
    F3 color;
    float power, area;
};
struct EmTri {
    uint32_t tri;
    float area, cdf;
};
struct Camera {
    std::string name;
    uint32_t W, H;
    float fovy;
    M4 view, invView;
    F3 pos, fwd;
    bool adjoint;
};
// Point (GPUDataStructures.h:36-51) after collectPointCloudGeometry.
struct Point {
    F3 p, tu, tv;
    float su, sv;
    F3 albedo;
    float alpha_t, alpha_r, opacity, beta, shape;
};
// PointGeometry of the asset (raw values as given by the caller).
struct AssetPoint {
    F3 p, tu, tv;
    float su, sv;
    F3 albedo;
    float opacity, beta, shape;
};
struct MeshData {
    std::vector<F3> verts;
    std::vector<uint32_t> idx;
    M4 o2w, w2o;
    Material mat;
};

struct Photon {
    F3 pos, power, dir;
};

struct Settings {
    uint32_t photons, bounces, forwardPasses, gatherPasses, adjointBounces, adjointPasses;
    uint64_t seed;
    float radius;
    uint32_t rrStart;
    float qNull, qReflect, qTransmit;
    bool decorrelate;
    uint32_t mapCapacity;
    bool fullGradients;
    bool adjointDecorrelate;
    uint32_t adjointSeed;
    bool debugImages;
};

struct Grid {
    float radius = 0.0f;
    F3 cell{0, 0, 0}, origin{0, 0, 0};
    int res[3] = {0, 0, 0};
    uint32_t total = 0;
    std::vector<uint32_t> cellStart, cellEnd, sorted;
};

struct WorldHit {
    bool hit = false;
    float t = kFltMax;
    float alpha = 0.0f;
    uint32_t prim = kInvalid, inst = kInvalid;
    F3 pos{0, 0, 0}, n{0, 0, 0};
};

struct Renderer {
    Settings S;
    // Scene sources (the reference's Scene + assets).
    std::vector<MeshData> meshes;
    std::vector<Camera> cams;
    std::vector<AssetPoint> asset;
    // Build products (SceneBuild::BuildProducts).
    std::vector<F3> verts;
    std::vector<Tri> tris;
    std::vector<Range> meshRanges;  // first tri / tri count (per mesh)
    std::vector<uint32_t> meshVertBase;
    std::vector<Material> materials;
    std::vector<Xf> xfs;
    std::vector<Instance> inst;
    uint32_t cloudInstance = kInvalid;
    // Test mode: render_forward keeps the previous photon map.
    bool freezeMap = false;
    uint64_t photonDraws = 0;
    std::vector<Point> points;
    std::vector<uint32_t> perm;  // perm[slot] = asset index before the build
    std::vector<Node> blas;
    std::vector<Range> blasRanges;
    std::vector<TNode> tlas;
    std::vector<Light> lights;
    std::vector<EmTri> emTris;
    uint32_t rayCapacity = 0;
    // Photon map (persistent between forward and backward, as in the reference).
    std::vector<Photon> photons;
    uint32_t photonCapacity = 0;
    uint64_t photonCounter = 0;  // attempted deposits (uint32 in the reference)
    uint32_t photonCount = 0;
    bool mapValid = false;
    Grid grid;
    bool gridConfigured = false;
    // Gradients.
    std::vector<float> gradOpacity;
    std::vector<float> debugImg;  // debug opacity images of the last backward (RGBA per pixel)
    std::vector<size_t> debugOff;
    std::vector<std::string> debugNames;
    // Full gradients (this build's extension; the reference returns zeros):
    // per surfel position 3, tangent_u 3, tangent_v 3, scale 2, albedo 3,
    // beta 1.
    std::vector<float> gradFull;
    pale_stats stats{};
    // Bounded-sample mode of the CPU baseline: trace photons/photonStride
    // photons and only pixels with index % pixelStride == pixelOffset
    // (several handles with different offsets split one sample).
    uint32_t photonStride = 1, pixelStride = 1, pixelOffset = 0;
};

static thread_local std::string g_err;

// Work units of the reference algorithm (per-query counts of the
// algorithmic byte model): popped nodes, slab tests, primitive tests,
// gathers, gather cells, non-empty cells, photons scanned.
struct Work {
    uint64_t pops = 0, aabb = 0, prims = 0, gathers = 0, cells = 0, cellsNE = 0, photons = 0;
};
static thread_local Work g_workClass[3];
static thread_local Work* g_workPtr = &g_workClass[0];
#define g_work (*g_workPtr)
static void workClass(int c) { g_workPtr = &g_workClass[c]; }
static void addWork(Work& a, const Work& b) {
    a.pops += b.pops;
    a.aabb += b.aabb;
    a.prims += b.prims;
    a.gathers += b.gathers;
    a.cells += b.cells;
    a.cellsNE += b.cellsNE;
    a.photons += b.photons;
}

// Threads of the oracle's parallel loops (0 = OpenMP's default).  Only work
// that the reference runs as independent work-items without shared state is
// spread over threads: closest-hit queries, camera pixels and irradiance
// gathers.  Everything whose result depends on the sequential item order
// (queue slots, deposit slots, the reservoir, gradient sums) stays in one
// sequential loop, so the results do not depend on the thread count.
static int g_threads = 0;
template <class F>
static void parallelItems(size_t n, F&& body) {
    Work* caller = g_workPtr;
    Work total{};
    const int nt = g_threads > 0 ? g_threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
    {
        Work local{};
        Work* saved = g_workPtr;
        g_workPtr = &local;
#pragma omp for schedule(dynamic, 16)
        for (long long i = 0; i < (long long)n; ++i) body((size_t)i);
        g_workPtr = saved;
#pragma omp critical
        addWork(total, local);
    }
    addWork(*caller, total);
}

// ------------------------------------------------------- BVH (BVH.ixx) --
struct AABB {
    F3 mn, mx;
};

// BVH.ixx:67-80
static void updateBoundsFromBoxes(const std::vector<AABB>& boxes, const std::vector<uint32_t>& order,
                                  Node& node) {
    F3 bmin = mk(kInf, kInf, kInf), bmax = mk(-kInf, -kInf, -kInf);
    for (uint32_t i = 0; i < node.count; ++i) {
        const AABB& b = boxes[order[node.leftFirst + i]];
        bmin = vmin(bmin, b.mn);
        bmax = vmax(bmax, b.mx);
    }
    node.mn = bmin;
    node.mx = bmax;
}

// BVH.ixx:82-127 (recursion order and child-pair allocation are kept: the
// node numbering depends on them).
static void subdivideBoxes(const std::vector<AABB>& boxes, const std::vector<F3>& cents, std::vector<Node>& nodes,
                           std::vector<uint32_t>& order, uint32_t nodeIndex, uint32_t maxLeaf) {
    Node node = nodes[nodeIndex];
    if (node.count <= maxLeaf) return;
    const F3 e = node.mx - node.mn;
    const int axis = (e.y > e.x && e.y > e.z) ? 1 : (e.z > e.x && e.z > e.y) ? 2 : 0;
    auto b = order.begin() + node.leftFirst;
    auto en = b + node.count;
    auto mid = b + (node.count >> 1);
    std::nth_element(b, mid, en, [&](uint32_t a, uint32_t c) { return comp(cents[a], axis) < comp(cents[c], axis); });
    const uint32_t leftCount = (uint32_t)(mid - b);
    if (leftCount == 0 || leftCount == node.count) return;
    const uint32_t L = (uint32_t)nodes.size(), R = L + 1;
    nodes.push_back(Node{});
    nodes.push_back(Node{});
    nodes[L].leftFirst = node.leftFirst;
    nodes[L].count = leftCount;
    nodes[R].leftFirst = node.leftFirst + leftCount;
    nodes[R].count = node.count - leftCount;
    nodes[nodeIndex].leftFirst = L;
    nodes[nodeIndex].count = 0;
    updateBoundsFromBoxes(boxes, order, nodes[L]);
    updateBoundsFromBoxes(boxes, order, nodes[R]);
    subdivideBoxes(boxes, cents, nodes, order, L, maxLeaf);
    subdivideBoxes(boxes, cents, nodes, order, R, maxLeaf);
}

// BVH.ixx:49-63
static void buildFromBoxes(const std::vector<AABB>& boxes, const std::vector<F3>& cents, std::vector<Node>& nodes,
                           std::vector<uint32_t>& order, uint32_t maxLeaf) {
    const uint32_t n = (uint32_t)boxes.size();
    nodes.clear();
    nodes.reserve(std::max<uint32_t>(1u, n * 2));
    order.resize(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    nodes.push_back(Node{});
    nodes[0].leftFirst = 0;
    nodes[0].count = n;
    updateBoundsFromBoxes(boxes, order, nodes[0]);
    subdivideBoxes(boxes, cents, nodes, order, 0, maxLeaf);
}

// Triangle BVH, BVH.ixx:26-45 and :141-220 (bounds from vertices, split on
// triangle centroids).
static void updateBoundsTri(const std::vector<Tri>& t, const std::vector<F3>& v, const std::vector<uint32_t>& idx,
                            Node& node) {
    F3 bmin = mk(kInf, kInf, kInf), bmax = mk(-kInf, -kInf, -kInf);
    for (uint32_t i = 0; i < node.count; ++i) {
        const Tri& T = t[idx[node.leftFirst + i]];
        bmin = vmin(bmin, v[T.v0]);
        bmin = vmin(bmin, v[T.v1]);
        bmin = vmin(bmin, v[T.v2]);
        bmax = vmax(bmax, v[T.v0]);
        bmax = vmax(bmax, v[T.v1]);
        bmax = vmax(bmax, v[T.v2]);
    }
    node.mn = bmin;
    node.mx = bmax;
}
static void subdivideTri(const std::vector<Tri>& t, const std::vector<F3>& v, std::vector<Node>& nodes,
                         std::vector<uint32_t>& idx, uint32_t ni, uint32_t maxLeaf) {
    Node node = nodes[ni];
    if (node.count <= maxLeaf) return;
    const F3 e = node.mx - node.mn;
    const int axis = (e.y > e.x && e.y > e.z) ? 1 : (e.z > e.x && e.z > e.y) ? 2 : 0;
    auto s = idx.begin() + node.leftFirst;
    auto en = s + node.count;
    auto mid = s + (node.count >> 1);
    std::nth_element(s, mid, en, [&](uint32_t a, uint32_t b) {
        return comp(t[a].centroid, axis) < comp(t[b].centroid, axis);
    });
    const uint32_t leftCount = (uint32_t)(mid - s);
    if (leftCount == 0 || leftCount == node.count) return;
    const uint32_t L = (uint32_t)nodes.size(), R = L + 1;
    nodes.push_back(Node{});
    nodes.push_back(Node{});
    nodes[L].leftFirst = node.leftFirst;
    nodes[L].count = leftCount;
    nodes[R].leftFirst = node.leftFirst + leftCount;
    nodes[R].count = node.count - leftCount;
    nodes[ni].leftFirst = L;
    nodes[ni].count = 0;
    updateBoundsTri(t, v, idx, nodes[L]);
    updateBoundsTri(t, v, idx, nodes[R]);
    subdivideTri(t, v, nodes, idx, L, maxLeaf);
    subdivideTri(t, v, nodes, idx, R, maxLeaf);
}

// surfelObjectAabbBeta (SceneBuild.cpp:395-444).
static AABB surfelAabb(const Point& s) {
    const F3 tu = normalize(s.tu);
    const F3 tv = normalize(s.tv);
    const F3 n = normalize(cross(tu, tv));
    const float rU = 1.0001f * s.su;
    const float rV = 1.0001f * s.sv;
    auto ext = [&](int a) {
        const float cu = comp(tu, a), cv = comp(tv, a), cn = std::fabs(comp(n, a));
        const float proj = std::sqrt((rU * cu) * (rU * cu) + (rV * cv) * (rV * cv));
        return proj + 0.001f * cn;
    };
    const F3 h = mk(ext(0), ext(1), ext(2));
    return AABB{s.p - h, s.p + h};
}

static float triangleArea(F3 p0, F3 p1, F3 p2) { return 0.5f * length(cross(p1 - p0, p2 - p0)); }

static inline float clampf(float v, float lo, float hi) { return std::min(std::max(v, lo), hi); }

// SceneBuild::build (SceneBuild.cpp:26-46) with the pieces the hot path
// reads: geometry, instances, point collection, lights, BLAS, TLAS.
static void buildScene(Renderer& R) {
    // collectGeometry (SceneBuild.cpp:72-127)
    R.verts.clear();
    R.tris.clear();
    R.meshRanges.clear();
    R.meshVertBase.clear();
    const float oneThird = 1.0f / 3.0f;
    for (const MeshData& m : R.meshes) {
        const uint32_t vb = (uint32_t)R.verts.size(), tb = (uint32_t)R.tris.size();
        for (const F3& v : m.verts) R.verts.push_back(v);
        for (size_t i = 0; i + 2 < m.idx.size(); i += 3) {
            Tri t{m.idx[i] + vb, m.idx[i + 1] + vb, m.idx[i + 2] + vb, {0, 0, 0}};
            t.centroid = (R.verts[t.v0] + R.verts[t.v1] + R.verts[t.v2]) * oneThird;
            R.tris.push_back(t);
        }
        R.meshRanges.push_back(Range{tb, (uint32_t)(m.idx.size() / 3)});
        R.meshVertBase.push_back(vb);
    }
    // collectInstances (SceneBuild.cpp:129-177)
    R.materials.clear();
    R.xfs.clear();
    R.inst.clear();
    for (uint32_t i = 0; i < R.meshes.size(); ++i) {
        R.materials.push_back(R.meshes[i].mat);
        R.xfs.push_back(Xf{R.meshes[i].o2w, R.meshes[i].w2o});
        R.inst.push_back(Instance{kMesh, i, i, i, 0});
    }
    // collectPointCloudGeometry (SceneBuild.cpp:181-227)
    std::vector<Point> pts(R.asset.size());
    for (size_t i = 0; i < R.asset.size(); ++i) {
        const AssetPoint& a = R.asset[i];
        Point p{};
        p.p = a.p;
        p.tu = normalize(a.tu);
        p.tv = normalize(a.tv);
        p.su = a.su;
        p.sv = a.sv;
        p.albedo = mk(clampf(a.albedo.x, 0, 1), clampf(a.albedo.y, 0, 1), clampf(a.albedo.z, 0, 1));
        p.opacity = clampf(a.opacity, 0, 1);
        p.alpha_r = 1.0f;
        p.alpha_t = 0.0f;
        p.beta = a.beta;
        p.shape = clampf(a.shape, -5.0f, 5.0f);
        pts[i] = p;
    }
    // collectPointCloudInstances (SceneBuild.cpp:229-251): identity transform.
    {
        M4 I{};
        for (int k = 0; k < 16; ++k) I.m[k] = (k % 5 == 0) ? 1.0f : 0.0f;
        R.xfs.push_back(Xf{I, I});
        R.inst.push_back(Instance{kCloud, 0, kInvalid, (uint32_t)R.xfs.size() - 1, 0});
        R.cloudInstance = (uint32_t)R.inst.size() - 1;
    }
    // collectLights (SceneBuild.cpp:263-343): built before the BLAS build
    // reorders triangles; the emissive-triangle indices refer to that order.
    R.lights.clear();
    R.emTris.clear();
    for (const Instance& in : R.inst) {
        if (in.type == kCloud) continue;
        const Material& mat = R.materials[in.material];
        if (!(mat.power > 0.0f)) continue;
        const Range& mr = R.meshRanges[in.geom];
        if (mr.count == 0) continue;
        const Xf& xf = R.xfs[in.xf];
        const uint32_t off = (uint32_t)R.emTris.size();
        float total = 0.0f;
        for (uint32_t t = 0; t < mr.count; ++t) {
            const Tri& T = R.tris[mr.first + t];
            const F3 p0 = toWorldPoint(R.verts[T.v0], xf.o2w);
            const F3 p1 = toWorldPoint(R.verts[T.v1], xf.o2w);
            const F3 p2 = toWorldPoint(R.verts[T.v2], xf.o2w);
            const float a = triangleArea(p0, p1, p2);
            total += a;
            R.emTris.push_back(EmTri{mr.first + t, a, 0.0f});
        }
        if (total <= 0.0f) {
            R.emTris.resize(off);
            continue;
        }
        float run = 0.0f;
        for (uint32_t t = 0; t < mr.count; ++t) {
            run += R.emTris[off + t].area;
            R.emTris[off + t].cdf = run / total;
        }
        R.emTris[off + mr.count - 1].cdf = 1.0f;
        Light L{};
        L.geom = in.geom;
        L.xf = in.xf;
        L.mat = in.material;
        L.triOffset = off;
        L.triCount = mr.count;
        L.power = mat.power;
        L.color = mat.base;
        L.area = total;
        R.lights.push_back(L);
    }
    // buildBottomLevelBVHs (SceneBuild.cpp:642-772)
    R.blas.clear();
    R.blasRanges.clear();
    std::vector<uint32_t> meshToBlas(R.meshes.size());
    for (uint32_t mi = 0; mi < R.meshes.size(); ++mi) {
        const Range mr = R.meshRanges[mi];
        const uint32_t vb = R.meshVertBase[mi];
        std::vector<F3> lv(R.verts.begin() + vb, R.verts.begin() + vb + R.meshes[mi].verts.size());
        std::vector<Tri> lt;
        for (uint32_t t = 0; t < mr.count; ++t) {
            Tri T = R.tris[mr.first + t];
            T.v0 -= vb;
            T.v1 -= vb;
            T.v2 -= vb;
            lt.push_back(T);
        }
        std::vector<Node> nodes;
        std::vector<uint32_t> order(lt.size());
        for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
        nodes.push_back(Node{});
        nodes[0].leftFirst = 0;
        nodes[0].count = (uint32_t)lt.size();
        updateBoundsTri(lt, lv, order, nodes[0]);
        subdivideTri(lt, lv, nodes, order, 0, 4);
        for (uint32_t k = 0; k < order.size(); ++k) {
            Tri T = lt[order[k]];
            T.v0 += vb;
            T.v1 += vb;
            T.v2 += vb;
            R.tris[mr.first + k] = T;
        }
        const uint32_t first = (uint32_t)R.blas.size();
        for (Node& nd : nodes)
            if (nd.count > 0) nd.leftFirst += mr.first;
        R.blas.insert(R.blas.end(), nodes.begin(), nodes.end());
        meshToBlas[mi] = (uint32_t)R.blasRanges.size();
        R.blasRanges.push_back(Range{first, (uint32_t)nodes.size()});
    }
    {
        std::vector<AABB> boxes(pts.size());
        std::vector<F3> cents(pts.size());
        for (size_t i = 0; i < pts.size(); ++i) {
            boxes[i] = surfelAabb(pts[i]);
            cents[i] = pts[i].p;
        }
        std::vector<Node> nodes;
        std::vector<uint32_t> order;
        buildFromBoxes(boxes, cents, nodes, order, 4);
        R.points.resize(pts.size());
        for (size_t k = 0; k < order.size(); ++k) R.points[k] = pts[order[k]];
        R.perm = order;
        const uint32_t first = (uint32_t)R.blas.size();
        R.blas.insert(R.blas.end(), nodes.begin(), nodes.end());
        const uint32_t bi = (uint32_t)R.blasRanges.size();
        R.blasRanges.push_back(Range{first, (uint32_t)nodes.size()});
        for (Instance& in : R.inst) in.blas = (in.type == kMesh) ? meshToBlas[in.geom] : bi;
    }
    // buildTLAS (SceneBuild.cpp:555-640)
    struct Box {
        F3 mn, mx;
        uint32_t inst;
    };
    std::vector<Box> boxes;
    for (uint32_t i = 0; i < R.inst.size(); ++i) {
        const Instance& in = R.inst[i];
        const Xf& xf = R.xfs[in.xf];
        const Node& root = R.blas[R.blasRanges[in.blas].first];
        F3 wmin = mk(kFltMax, kFltMax, kFltMax), wmax = mk(-kFltMax, -kFltMax, -kFltMax);
        for (int c = 0; c < 8; ++c) {
            const bool bx = c & 4, by = c & 2, bz = c & 1;
            const F3 po = mk(bx ? root.mx.x : root.mn.x, by ? root.mx.y : root.mn.y, bz ? root.mx.z : root.mn.z);
            const F3 pw = toWorldPoint(po, xf.o2w);
            wmin = vmin(wmin, pw);
            wmax = vmax(wmax, pw);
        }
        boxes.push_back(Box{wmin, wmax, i});
    }
    R.tlas.clear();
    std::function<uint32_t(uint32_t, uint32_t)> build = [&](uint32_t start, uint32_t end) -> uint32_t {
        const uint32_t n = (uint32_t)R.tlas.size();
        R.tlas.push_back(TNode{});
        F3 bmin = mk(kFltMax, kFltMax, kFltMax), bmax = mk(-kFltMax, -kFltMax, -kFltMax);
        for (uint32_t i = start; i < end; ++i) {
            bmin = vmin(bmin, boxes[i].mn);
            bmax = vmax(bmax, boxes[i].mx);
        }
        R.tlas[n].mn = bmin;
        R.tlas[n].mx = bmax;
        const uint32_t count = end - start;
        if (count == 1) {
            R.tlas[n].count = 1;
            R.tlas[n].left = boxes[start].inst;
            R.tlas[n].right = 0;
        } else {
            R.tlas[n].count = 0;
            F3 cmin = mk(kFltMax, kFltMax, kFltMax), cmax = mk(-kFltMax, -kFltMax, -kFltMax);
            for (uint32_t i = start; i < end; ++i) {
                const F3 c = (boxes[i].mn + boxes[i].mx) * 0.5f;
                cmin = vmin(cmin, c);
                cmax = vmax(cmax, c);
            }
            const F3 ext = cmax - cmin;
            const int axis = (ext.x > ext.y && ext.x > ext.z) ? 0 : (ext.y > ext.z) ? 1 : 2;
            const float pivot = (comp(cmin, axis) + comp(cmax, axis)) * 0.5f;
            auto it = std::partition(boxes.begin() + start, boxes.begin() + end, [&](const Box& b) {
                const F3 c = (b.mn + b.mx) * 0.5f;
                return comp(c, axis) < pivot;
            });
            uint32_t mid = (uint32_t)(it - boxes.begin());
            if (mid == start || mid == end) mid = start + count / 2;
            const uint32_t l = build(start, mid);
            const uint32_t r = build(mid, end);
            R.tlas[n].left = l;
            R.tlas[n].right = r;
        }
        return n;
    };
    if (!boxes.empty()) build(0, (uint32_t)boxes.size());
}

// configurePhotonGrid (PathTracer.cpp:184-221), once at the first setScene.
static void configureGrid(Renderer& R) {
    Grid& g = R.grid;
    g.radius = R.S.radius;
    const float cell = 0.5f * g.radius;
    g.cell = mk(cell, cell, cell);
    const F3 pad = mk(g.radius, g.radius, g.radius);
    g.origin = R.tlas[0].mn - pad;
    const F3 gmax = R.tlas[0].mx + pad;
    const F3 ext = gmax - g.origin;
    auto axisCount = [](float e, float c) { return (int)std::ceil(e / std::fmax(c, 1e-6f)); };
    g.res[0] = axisCount(ext.x, cell);
    g.res[1] = axisCount(ext.y, cell);
    g.res[2] = axisCount(ext.z, cell);
    const uint64_t total = (uint64_t)g.res[0] * (uint64_t)g.res[1] * (uint64_t)g.res[2];
    if (total == 0 || total > 0xFFFFFFFFull)
        throw std::runtime_error("Photon grid resolution too high; increase r or tighten AABB.");
    g.total = (uint32_t)total;
    R.gridConfigured = true;
}

// PathTracer::setScene / ensureRayCapacity / allocatePhotonMap
// (PathTracer.cpp:22-159).
static void setScene(Renderer& R) {
    uint32_t cap = 1;
    while (cap < R.S.photons) cap <<= 1u;
    R.rayCapacity = std::max(R.rayCapacity, cap);
    if (!R.gridConfigured) {
        const uint64_t requested = (uint64_t)R.S.photons * (uint64_t)(R.S.forwardPasses * R.S.bounces);
        const uint64_t maxPhotons = (10ull * 1024ull * 1024ull * 1024ull) / 64ull;
        R.photonCapacity = R.S.mapCapacity ? R.S.mapCapacity : (uint32_t)std::min(requested, maxPhotons);
        configureGrid(R);
    }
}

// -------------------------------------------------- intersection kernels --
// KernelHelpers.h:140-147
static inline F3 safeInvDir(F3 d) {
    return mk(std::fabs(d.x) < 1e-6f ? 1e30f : 1.0f / d.x, std::fabs(d.y) < 1e-6f ? 1e30f : 1.0f / d.y,
              std::fabs(d.z) < 1e-6f ? 1e30f : 1.0f / d.z);
}
// KernelHelpers.h:149-214: the TLAS variant clamps the entry to 0, the BVH
// variant to 1e-6.
static inline bool slab(const Ray& r, F3 mn, F3 mx, F3 inv, float limit, float clampLo, float& tEntry) {
    const F3 t0 = (mn - r.o) * inv;
    const F3 t1 = (mx - r.o) * inv;
    const float txmin = std::fmin(t0.x, t1.x), txmax = std::fmax(t0.x, t1.x);
    const float tymin = std::fmin(t0.y, t1.y), tymax = std::fmax(t0.y, t1.y);
    const float tzmin = std::fmin(t0.z, t1.z), tzmax = std::fmax(t0.z, t1.z);
    const float tmin = std::fmax(std::fmax(txmin, tymin), tzmin);
    const float tmax = std::fmin(std::fmin(txmax, tymax), tzmax);
    if (tmin > tmax) return false;
    if (tmax < 0.0f) return false;
    if (tmin > limit) return false;
    tEntry = std::fmax(tmin, clampLo);
    return true;
}

// SmallStack (GPUDataTypes.h:393-412): drops pushes on overflow, pop of an
// empty stack returns 0.
template <int N>
struct SmallStack {
    uint32_t data[N];
    int sp = 0;
    void push(uint32_t v) {
        if (sp >= N) return;
        data[sp++] = v;
    }
    uint32_t pop() { return sp <= 0 ? 0u : data[--sp]; }
    bool empty() const { return sp == 0; }
};
template <typename S>
static inline void pushNearFar(S& st, uint32_t l, float tl, uint32_t r, float tr) {
    if (tl <= tr) {
        st.push(r);
        st.push(l);
    } else {
        st.push(l);
        st.push(r);
    }
}

struct LocalHit {
    F3 pos{0, 0, 0};
    float t = kFltMax, alpha = 0.0f;
    uint32_t prim = kInvalid;
};

// KernelHelpers.h:248-277
static bool intersectTriangle(const Ray& r, F3 v0, F3 v1, F3 v2, float& t, float tMin) {
    const F3 e1 = v1 - v0, e2 = v2 - v0;
    const F3 h = cross(r.d, e2);
    const float a = dot(e1, h);
    if (std::fabs(a) < 1.0e-4f) return false;
    const float f = 1.0f / a;
    const F3 s = r.o - v0;
    const float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    const F3 q = cross(s, e1);
    const float v = f * dot(r.d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    const float tt = f * dot(e2, q);
    if (tt <= tMin) return false;
    t = tt;
    return true;
}

// KernelHelpers.h:637-669 (with opacityBeta :623-634 and phiInverse :514-520).
static bool intersectSurfel(const Ray& r, const Point& s, float tMin, float tMax, float& tOut, F3& hitOut,
                            float& alphaOut) {
    const F3 u = normalize(s.tu);
    const F3 v = normalize(s.tv - u * dot(u, s.tv));
    const F3 n = normalize(cross(u, v));
    const float nd = dot(n, r.d);
    if (std::fabs(nd) < 1e-6f) return false;
    const float t = dot(n, s.p - r.o) / nd;
    if (t <= tMin || t >= tMax) return false;
    const F3 x = r.o + r.d * t;
    const F3 rr = x - s.p;
    const float uu = dot(s.tu, rr) / s.su;
    const float vv = dot(s.tv, rr) / s.sv;
    const float r2 = uu * uu + vv * vv;
    if (r2 >= 1.0f) return false;
    const float base = 1.0f - r2;
    const float e = 4.0f * exp_f(s.beta);
    alphaOut = pow_f(base, e);
    hitOut = x;
    tOut = t;
    return true;
}

struct Counters {
    uint64_t queries = 0;
};

// intersectBLASMesh (IntersectionKernels.h:37-105)
static bool blasMesh(const Renderer& R, const Ray& ro, uint32_t bi, LocalHit& out, const Xf& xf) {
    const Range br = R.blasRanges[bi];
    const Node* nodes = R.blas.data() + br.first;
    if (br.count <= 1 && nodes[0].count == 0) return false;  // empty mesh: build-defined (no hit)
    float best = kInf;
    bool any = false;
    const F3 inv = safeInvDir(ro.d);
    SmallStack<512> st;
    st.push(0);
    while (!st.empty()) {
        const uint32_t ni = st.pop();
        const Node& nd = nodes[ni];
        float te = 0.0f;
        g_work.pops++;
        g_work.aabb++;
        if (!slab(ro, nd.mn, nd.mx, inv, best, 1e-6f, te)) continue;
        if (nd.count == 0) {
            g_work.aabb += 2;
            const uint32_t l = nd.leftFirst, r = nd.leftFirst + 1;
            float tl = kInf, tr = kInf;
            const bool hl = slab(ro, nodes[l].mn, nodes[l].mx, inv, best, 1e-6f, tl);
            const bool hr = slab(ro, nodes[r].mn, nodes[r].mx, inv, best, 1e-6f, tr);
            if (hl && hr) pushNearFar(st, l, tl, r, tr);
            else if (hl) st.push(l);
            else if (hr) st.push(r);
            continue;
        }
        g_work.prims += nd.count;
        for (uint32_t i = 0; i < nd.count; ++i) {
            const uint32_t ti = nd.leftFirst + i;
            const Tri& T = R.tris[ti];
            float t = kFltMax;
            if (intersectTriangle(ro, R.verts[T.v0], R.verts[T.v1], R.verts[T.v2], t, 1e-4f) && t < best &&
                t > 0.0f) {
                best = t;
                any = true;
                out.t = t;
                out.prim = ti;
                out.pos = toWorldPoint(ro.o + ro.d * t, xf.o2w);
            }
        }
    }
    return any;
}

// intersectBLASPointCloudFirstHit (IntersectionKernels.h:585-675)
static bool blasCloud(const Renderer& R, const Ray& ro, uint32_t bi, LocalHit& out) {
    if (R.points.empty()) return false;  // empty cloud: build-defined (no hit)
    const Range br = R.blasRanges[bi];
    const Node* nodes = R.blas.data() + br.first;
    bool any = false;
    float best = kInf, bestAlpha = 0.0f;
    uint32_t bestIdx = 0;
    F3 bestHit = mk(0, 0, 0);
    const F3 inv = safeInvDir(ro.d);
    SmallStack<256> st;
    st.push(0);
    while (!st.empty()) {
        const uint32_t ni = st.pop();
        const Node& nd = nodes[ni];
        float te = 0.0f;
        g_work.pops++;
        g_work.aabb++;
        if (!slab(ro, nd.mn, nd.mx, inv, best, 1e-6f, te)) continue;
        if (nd.count == 0) {
            g_work.aabb += 2;
            const uint32_t l = nd.leftFirst, r = nd.leftFirst + 1;
            float tl = kInf, tr = kInf;
            const bool hl = slab(ro, nodes[l].mn, nodes[l].mx, inv, best, 1e-6f, tl);
            const bool hr = slab(ro, nodes[r].mn, nodes[r].mx, inv, best, 1e-6f, tr);
            if (hl && hr) pushNearFar(st, l, tl, r, tr);
            else if (hl) st.push(l);
            else if (hr) st.push(r);
            continue;
        }
        g_work.prims += nd.count;
        for (uint32_t k = 0; k < nd.count; ++k) {
            const uint32_t si = nd.leftFirst + k;
            float t = 0.0f, a = 0.0f;
            F3 x = mk(0, 0, 0);
            if (!intersectSurfel(ro, R.points[si], 1e-5f, best, t, x, a)) continue;
            any = true;
            best = t;
            bestIdx = si;
            bestHit = x;
            bestAlpha = a;
        }
    }
    if (!any) return false;
    out.t = best;
    out.prim = bestIdx;
    out.alpha = bestAlpha;
    out.pos = bestHit;
    return true;
}

// KernelHelpers.h:217-227
static inline Ray toObjectSpace(const Ray& w, const Xf& xf) {
    float h[4];
    mulv(xf.w2o, w.o.x, w.o.y, w.o.z, 1.0f, h);
    Ray r;
    r.o = divs(mk(h[0], h[1], h[2]), h[3]);
    mulv(xf.w2o, w.d.x, w.d.y, w.d.z, 0.0f, h);
    r.d = normalize(mk(h[0], h[1], h[2]));
    return r;
}

// intersectScene in FirstHit mode (IntersectionKernels.h:954-1097) followed
// by buildIntersectionNormal (KernelHelpers.h:480-503).
static bool intersectScene(const Renderer& R, const Ray& rw, WorldHit& out) {
    out = WorldHit{};
    if (R.tlas.empty()) return false;
    const F3 inv = safeInvDir(rw.d);
    SmallStack<256> st;
    st.push(0);
    float bestW = kInf;
    bool found = false;
    while (!st.empty()) {
        const uint32_t ni = st.pop();
        const TNode& nd = R.tlas[ni];
        float te = 0.0f;
        g_work.pops++;
        g_work.aabb++;
        if (!slab(rw, nd.mn, nd.mx, inv, bestW, 0.0f, te)) continue;
        if (nd.count == 0) {
            g_work.aabb += 2;
            const uint32_t l = nd.left, r = nd.right;
            float tl = kInf, tr = kInf;
            const bool hl = slab(rw, R.tlas[l].mn, R.tlas[l].mx, inv, bestW, 0.0f, tl);
            const bool hr = slab(rw, R.tlas[r].mn, R.tlas[r].mx, inv, bestW, 0.0f, tr);
            if (hl && hr) pushNearFar(st, l, tl, r, tr);
            else if (hl) st.push(l);
            else if (hr) st.push(r);
            continue;
        }
        const uint32_t ii = nd.left;
        const Instance& in = R.inst[ii];
        const Xf& xf = R.xfs[in.xf];
        const Ray ro = toObjectSpace(rw, xf);
        LocalHit lh;
        const bool acc = (in.type == kMesh) ? blasMesh(R, ro, in.blas, lh, xf) : blasCloud(R, ro, in.blas, lh);
        if (acc) {
            const float tw = dot(lh.pos - rw.o, rw.d);
            if (tw > 0.0f && tw < bestW) {
                bestW = tw;
                found = true;
                out.hit = true;
                out.t = tw;
                out.pos = lh.pos;
                out.inst = ii;
                out.prim = lh.prim;
                out.alpha = lh.alpha;
            }
        }
    }
    if (!found) return false;
    const Instance& in = R.inst[out.inst];
    if (in.type == kMesh) {
        const Tri& T = R.tris[out.prim];
        const Xf& xf = R.xfs[in.xf];
        const F3 p0 = toWorldPoint(R.verts[T.v0], xf.o2w);
        const F3 p1 = toWorldPoint(R.verts[T.v1], xf.o2w);
        const F3 p2 = toWorldPoint(R.verts[T.v2], xf.o2w);
        out.n = normalize(cross(p1 - p0, p2 - p0));
    } else {
        const Point& s = R.points[out.prim];
        out.n = normalize(cross(s.tu, s.tv));
    }
    return true;
}

// ------------------------------------------------------------- sampling --
// sampleCosineHemisphere(rng, n, dir, pdf) (KernelHelpers.h:317-337)
static void sampleCosine(Xorshift128& rng, F3 n, F3& dir, float& pdf) {
    const float u1 = rng.nextFloat();
    const float u2 = rng.nextFloat();
    const float z = std::sqrt(1.0f - u1);
    const float r = std::sqrt(1.0f - z * z);
    const float phi = 2.0f * kPi * u2;
    const float x = r * cos_f(phi);
    const float y = r * sin_f(phi);
    const F3 up = std::fabs(n.z) < 0.999f ? mk(0, 0, 1) : mk(1, 0, 0);
    const F3 tang = normalize(cross(up, n));
    const F3 bit = cross(n, tang);
    dir = normalize(tang * x + bit * y + n * z);
    pdf = std::fmax(0.0f, dot(dir, n)) / kPi;
}

// sampleUniformHemisphereAroundNormal (KernelHelpers.h:551-596) with the
// Frisvad basis of buildOrthonormalBasisFromNormal (:526-549).
static void sampleUniformHemi(Xorshift128& rng, F3 normal, F3& dir, float& pdf) {
    const F3 n = normalize(normal);
    const float u1 = rng.nextFloat();
    const float u2 = rng.nextFloat();
    const float z = u1;
    const float phi = 2.0f * kPi * u2;
    const float r = std::sqrt(std::fmax(0.0f, 1.0f - z * z));
    const float x = r * cos_f(phi);
    const float y = r * sin_f(phi);
    const float sign = std::copysign(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    const F3 t = mk(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    const F3 bt = mk(b, sign + n.y * n.y * a, -n.y);
    dir = normalize(t * x + bt * y + n * z);
    pdf = 1.0f / (2.0f * kPi);
}

// applyRussianRoulette (KernelHelpers.h:1612-1657)
static bool russianRoulette(Xorshift128& rng, uint32_t bounce, F3& T, uint32_t start) {
    if (bounce < start) return true;
    if (!std::isfinite(T.x) || !std::isfinite(T.y) || !std::isfinite(T.z)) return false;
    const float lum = 0.2126f * T.x + 0.7152f * T.y + 0.0722f * T.z;
    const float p = std::min(std::max(lum, 0.0f), 0.99f);
    if (rng.nextFloat() > p) return false;
    float c = 1.0f / p;
    c = std::fmin(c, 10.0f);
    T = T * c;
    return true;
}

struct RayState {
    Ray ray;
    F3 T;
    uint32_t bounce;
    uint32_t pixel;
};

// depositPhotonSurface (KernelHelpers.h:1688-1727)
static void deposit(Renderer& R, F3 pos, F3 dir, F3 flux, Xorshift128& rng) {
    const uint32_t attempt = (uint32_t)R.photonCounter;
    R.photonCounter += 1;
    uint32_t slot;
    if (attempt < R.photonCapacity) {
        slot = attempt;
    } else {
        const float u = rng.nextFloat();
        const float accept = (float)R.photonCapacity / (float)(attempt + 1u);
        if (u >= accept) return;
        slot = rng.nextUint() % R.photonCapacity;
    }
    R.photons[slot] = Photon{pos, flux, dir};
}

// ------------------------------------------------------- forward pass --
// sampleMeshAreaLight (KernelHelpers.h:340-449)
static bool sampleLight(const Renderer& R, Xorshift128& rng, F3& pos, F3& nrm, F3& dir, F3& power) {
    const uint32_t lc = (uint32_t)R.lights.size();
    if (lc == 0) return false;
    const float ul = rng.nextFloat();
    const uint32_t li = std::min((uint32_t)(ul * (float)lc), lc - 1u);
    const Light& L = R.lights[li];
    if (L.triCount == 0u || L.area <= 0.0f) return false;
    const float ut = rng.nextFloat();
    uint32_t lo = 0u, hi = L.triCount - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1u;
        if (ut <= R.emTris[L.triOffset + mid].cdf) hi = mid;
        else lo = mid + 1u;
    }
    const EmTri& et = R.emTris[L.triOffset + lo];
    const Tri& T = R.tris[et.tri];
    const F3 v0 = R.verts[T.v0], v1 = R.verts[T.v1], v2 = R.verts[T.v2];
    const float u1 = rng.nextFloat();
    const float u2 = rng.nextFloat();
    const float s1 = std::sqrt(u1);
    const float b0 = 1.0f - s1, b1 = s1 * (1.0f - u2), b2 = s1 * u2;
    const F3 xo = v0 * b0 + v1 * b1 + v2 * b2;
    const Xf& xf = R.xfs[L.xf];
    const F3 p0 = toWorldPoint(v0, xf.o2w), p1 = toWorldPoint(v1, xf.o2w), p2 = toWorldPoint(v2, xf.o2w);
    const F3 e0 = p1 - p0, e1 = p2 - p0;
    F3 n = mk(e0.y * e1.z - e0.z * e1.y, e0.z * e1.x - e0.x * e1.z, e0.x * e1.y - e0.y * e1.x);
    const float nl = std::sqrt(dot(n, n));
    if (nl <= 0.0f) return false;
    n = divs(n, nl);
    float pdf = 0.0f;
    sampleCosine(rng, n, dir, pdf);
    pos = toWorldPoint(xo, xf.o2w);
    nrm = n;
    power = L.color * L.power;
    return true;
}

// One photon bounce (launchIntersectKernel, PrimalKernels.cpp:65-321).
// The closest hits of a bounce are computed for all rays first (in
// parallel: each depends only on its ray), then the rays are shaded in
// queue order, where deposits, the reservoir and the next queue depend on
// the order.
static void photonBounce(Renderer& R, const std::vector<RayState>& in, std::vector<RayState>& out, uint32_t number) {
    const Settings& S = R.S;
    std::vector<WorldHit> hits(in.size());
    std::vector<char> hit(in.size());
    parallelItems(in.size(), [&](size_t i) { hit[i] = intersectScene(R, in[i].ray, hits[i]) ? 1 : 0; });
    R.stats.photon_queries += in.size();
    for (uint32_t i = 0; i < in.size(); ++i) {
        Xorshift128 rng(perItemSeed(number, i));
        const RayState& rs = in[i];
        if (!hit[i]) continue;
        const WorldHit& h = hits[i];
        const Instance& inst = R.inst[h.inst];
        if (inst.type == kMesh) {
            deposit(R, h.pos, rs.ray.d, rs.T, rng);
            R.stats.photons_deposited++;
            const Material& m = R.materials[inst.material];
            F3 dir = rs.ray.d;
            float pdf = 0.0f;
            sampleCosine(rng, h.n, dir, pdf);
            RayState nx{};
            nx.ray.o = h.pos + h.n * 1e-6f;
            nx.ray.d = dir;
            nx.bounce = rs.bounce + 1;
            nx.pixel = rs.pixel;
            nx.T = rs.T * m.base;
            if (!russianRoulette(rng, nx.bounce, nx.T, S.rrStart)) continue;
            out.push_back(nx);
        } else {
            const float u = rng.nextFloat();
            const Point& s = R.points[h.prim];
            if (u < S.qNull) {
                const float att = 1.0f - h.alpha * s.opacity;
                const float w = att / S.qNull;
                RayState nx{};
                nx.ray.o = h.pos + rs.ray.d * 1e-5f;
                nx.ray.d = rs.ray.d;
                nx.bounce = rs.bounce + 1;
                nx.pixel = rs.pixel;
                nx.T = rs.T * w;
                if (!russianRoulette(rng, nx.bounce, nx.T, S.rrStart)) continue;
                out.push_back(nx);
            } else if (u < S.qNull + S.qReflect) {
                const F3 nc = normalize(cross(s.tu, s.tv));
                const float sc = dot(nc, -rs.ray.d);
                const int sign = sc >= 0.0f ? 1 : -1;
                const F3 no = (float)sign * nc;
                deposit(R, h.pos, rs.ray.d, divs(rs.T, S.qReflect), rng);
                R.stats.photons_deposited++;
                F3 dir = rs.ray.d;
                float pdf = 0.0f;
                sampleUniformHemi(rng, no, dir, pdf);
                const float cosT = std::fmax(0.0f, dot(dir, no));
                const F3 fs = (s.alpha_r * s.albedo) * kInvPi;
                const float alpha = h.alpha * s.opacity;
                const F3 w = divs(((alpha / S.qReflect) * fs) * cosT, pdf);
                RayState nx{};
                nx.ray.o = h.pos + no * 1e-5f;
                nx.ray.d = dir;
                nx.bounce = rs.bounce + 1;
                nx.pixel = rs.pixel;
                nx.T = rs.T * w;
                if (!russianRoulette(rng, nx.bounce, nx.T, S.rrStart)) continue;
                out.push_back(nx);
            } else if (u < S.qNull + S.qReflect + S.qTransmit) {
                const F3 nc = normalize(cross(s.tu, s.tv));
                const float sc = dot(nc, -rs.ray.d);
                const int sign = sc >= 0.0f ? 1 : -1;
                const F3 no = (float)sign * nc;
                F3 dir = rs.ray.d;
                float pdf = 0.0f;
                sampleUniformHemi(rng, no, dir, pdf);
                const float cosT = std::fmax(0.0f, dot(dir, no));
                const F3 fs = (s.alpha_t * s.albedo) * kInvPi;
                const float alpha = h.alpha * s.opacity;
                const F3 w = divs(((alpha / S.qReflect) * fs) * cosT, pdf);
                RayState nx{};
                nx.ray.o = h.pos + (-no) * 1e-5f;
                nx.ray.d = dir;
                nx.bounce = rs.bounce + 1;
                nx.pixel = rs.pixel;
                nx.T = rs.T * w;
                if (!russianRoulette(rng, nx.bounce, nx.T, S.rrStart)) continue;
                out.push_back(nx);
            }
        }
    }
}

// KernelHelpers.h:897-918 (worldToCell == worldToCellClamped)
static inline void worldToCell(const Grid& g, F3 p, int c[3]) {
    const F3 f = divv(p - g.origin, g.cell);
    c[0] = std::min(std::max((int)std::floor(f.x), 0), g.res[0] - 1);
    c[1] = std::min(std::max((int)std::floor(f.y), 0), g.res[1] - 1);
    c[2] = std::min(std::max((int)std::floor(f.z), 0), g.res[2] - 1);
}
static inline uint32_t linearCell(const Grid& g, const int c[3]) {
    return (uint32_t)((c[2] * g.res[1] + c[1]) * g.res[0] + c[0]);
}

// buildPhotonCellRangesAndOrdering (PrimalKernels.cpp:796-1018) in sequential
// order: the scatter's per-cell offsets grow with the photon index.
static void buildGrid(Renderer& R) {
    Grid& g = R.grid;
    const uint32_t n = R.photonCount;
    std::vector<uint32_t> cellId(n), cnt(g.total, 0u);
    for (uint32_t i = 0; i < n; ++i) {
        int c[3];
        worldToCell(g, R.photons[i].pos, c);
        cellId[i] = linearCell(g, c);
        cnt[cellId[i]]++;
    }
    g.cellStart.assign(g.total, kInvalid);
    g.cellEnd.assign(g.total, kInvalid);
    uint32_t run = 0;
    for (uint32_t c = 0; c < g.total; ++c) {
        if (cnt[c]) {
            g.cellStart[c] = run;
            g.cellEnd[c] = run + cnt[c];
        }
        run += cnt[c];
    }
    g.sorted.assign(n, 0u);
    std::vector<uint32_t> off(g.total, 0u);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t c = cellId[i];
        g.sorted[g.cellStart[c] + off[c]++] = i;
    }
}

// gatherDiffuseIrradianceAtPoint (KernelHelpers.h:990-1037)
static F3 gather(const Renderer& R, F3 q, F3 n) {
    g_work.gathers++;
    if (!R.mapValid) return mk(0, 0, 0);
    const Grid& g = R.grid;
    const float r = g.radius;
    const float r2 = r * r;
    const float invArea = 1.0f / (kPi * r2);
    int lo[3], hi[3];
    worldToCell(g, q - mk(r, r, r), lo);
    worldToCell(g, q + mk(r, r, r), hi);
    F3 E = mk(0, 0, 0);
    for (int cz = lo[2]; cz <= hi[2]; ++cz)
        for (int cy = lo[1]; cy <= hi[1]; ++cy)
            for (int cx = lo[0]; cx <= hi[0]; ++cx) {
                const int c[3] = {cx, cy, cz};
                const uint32_t id = linearCell(g, c);
                const uint32_t st = g.cellStart[id];
                g_work.cells++;
                if (st == kInvalid) continue;
                const uint32_t en = g.cellEnd[id];
                g_work.cellsNE++;
                g_work.photons += en - st;
                for (uint32_t j = st; j < en; ++j) {
                    const Photon& ph = R.photons[g.sorted[j]];
                    const F3 d = ph.pos - q;
                    const float cosine = dot(n, -ph.dir);
                    if (cosine < 0.0f) continue;
                    const float d2 = dot(d, d);
                    if (d2 > r2) continue;
                    E = E + ph.power * invArea;
                }
            }
    const uint32_t T = (uint32_t)R.photonCounter;
    const uint32_t D = std::min(T, R.photonCapacity);
    if (D == 0u) return mk(0, 0, 0);
    return E * ((float)T / (float)D);
}

// makePrimaryRayFromPixelJitteredFov (KernelHelpers.h:764-796); the focal
// length is a host-side constant of the camera.
struct CamConst {
    float fx, fy;
};
static CamConst camConst(const Camera& c) {
    const float W = (float)c.W, H = (float)c.H;
    const float rad = c.fovy * 0.01745329251994329576923690768489f;  // glm::radians
    const float fy = 0.5f * H / (float)std::tan((double)(0.5f * rad));
    return CamConst{fy * (W / H), fy};
}
static Ray primaryRay(const Camera& c, const CamConst& k, float px, float py, float jx, float jy) {
    const float W = (float)c.W, H = (float)c.H;
    const float u = px + 0.5f + jx;
    const float v = py + 0.5f + jy;
    const float vf = H - v;
    const float ndcX = 2.0f * u / W - 1.0f;
    const float ndcY = 2.0f * vf / H - 1.0f;
    const F3 dc = normalize(mk(ndcX * (0.5f * W) / k.fx, ndcY * (0.5f * H) / k.fy, -1.0f));
    Ray r;
    r.d = transformDirection(c.invView, dc);
    r.o = transformPoint(c.invView, mk(0, 0, 0));
    return r;
}

// launchCameraGatherKernel (PrimalKernels.cpp:500-653) for every pass.
// A served view passes its own pass count and always decorrelates.
static void cameraGather(Renderer& R, const Camera& c, uint32_t baseSeed, std::vector<float>& fb, uint32_t passes,
                         bool decorrelate) {
    const uint32_t W = c.W, H = c.H, npx = W * H;
    fb.assign((size_t)npx * 4, 0.0f);
    const CamConst k = camConst(c);
    const float spp = (float)passes;
    const float invSpp = 1.0f / spp;
    std::vector<uint32_t> pix;
    for (uint32_t pi = 0; pi < npx; ++pi)
        if (pi % R.pixelStride == R.pixelOffset) pix.push_back(pi);
    std::vector<uint64_t> queries(pix.size());
    for (uint32_t pass = 0; pass < passes; ++pass) {
        const uint64_t seed = decorrelate ? (uint64_t)(uint32_t)(baseSeed + pass) : (uint64_t)baseSeed;
        // Pixels are independent: each adds to its own framebuffer entry,
        // once per pass, in pass order.
        parallelItems(pix.size(), [&](size_t ix) {
            const uint32_t pi = pix[ix];
            const uint32_t px = pi % W, py = pi / W;
            Xorshift128 rng(perItemSeed(seed, pi));
            const float jx = rng.nextFloat() - 0.5f;
            const float jy = rng.nextFloat() - 0.5f;
            Ray ray = primaryRay(c, k, (float)px, (float)py, jx, jy);
            F3 C = mk(0, 0, 0);
            float tr = 1.0f;
            uint64_t nq = 0;
            while (true) {
                WorldHit h;
                nq++;
                if (!intersectScene(R, ray, h)) break;
                const Instance& in = R.inst[h.inst];
                if (in.type == kCloud) {
                    const Point& s = R.points[h.prim];
                    F3 n = normalize(cross(s.tu, s.tv));
                    if (dot(n, -ray.d) < 0.0f) n = -n;
                    const F3 E = gather(R, h.pos, n);
                    const F3 Lo = E * ((s.alpha_r * s.albedo) * kInvPi);
                    const float ae = s.opacity * h.alpha;
                    C = C + (tr * ae) * Lo;
                    tr *= (1.0f - ae);
                    ray.o = h.pos + ray.d * 1e-4f;
                    continue;
                }
                const Material& m = R.materials[in.material];
                if (m.power > 0.0f) {
                    const F3 Le = m.base * m.power;
                    C = C + tr * vmin(Le, mk(1.0f, 1.0f, 1.0f));
                } else {
                    const F3 E = gather(R, h.pos, h.n);
                    const F3 Lo = (m.base * kInvPi) * E;
                    C = C + tr * Lo;
                }
                tr = 0.0f;
                break;
            }
            float* f = &fb[(size_t)pi * 4];
            f[0] = C.x * invSpp + f[0];
            f[1] = C.y * invSpp + f[1];
            f[2] = C.z * invSpp + f[2];
            f[3] = 1.0f * invSpp + f[3];
            queries[ix] += nq;
        });
    }
    for (uint64_t q : queries) R.stats.gather_queries += q;
}

// launchPostProcessKernel (PostprocessKernel.cpp:11-106), LDR float output.
static void postprocess(const std::vector<float>& hdr, float* ldr, size_t npx) {
    const float invGamma = 1.0f / 2.2f;
    for (size_t i = 0; i < npx; ++i) {
        float c[3];
        for (int k = 0; k < 3; ++k) {
            float v = hdr[i * 4 + k] * 1.0f;
            v = std::fmax(v, 0.0f);
            v = pow_f(v, invGamma);
            v = std::min(std::max(v, 0.0f), 1.0f);
            c[k] = v;
        }
        ldr[i * 4 + 0] = c[0];
        ldr[i * 4 + 1] = c[1];
        ldr[i * 4 + 2] = c[2];
        ldr[i * 4 + 3] = 1.0f;
    }
}

// submitPhotonMappingKernel (SyclBridge.cpp:82-166).
static void renderForward(Renderer& R, const std::vector<uint32_t>& camIdx, float* const* ldr, float* const* raw) {
    const Settings& S = R.S;
    std::mt19937_64 seedGen(S.seed);
    if (R.freezeMap && R.mapValid) {
        // Test mode: keep the photon map of the previous render and only
        // re-render the cameras (same camera seeds).
        for (uint64_t k = 0; k < R.photonDraws; ++k) (void)seedGen();
        workClass(1);
        std::vector<float> fb;
        for (size_t k = 0; k < camIdx.size(); ++k) {
            const uint32_t number = (uint32_t)seedGen();
            const Camera& c = R.cams[camIdx[k]];
            cameraGather(R, c, number, fb, R.S.gatherPasses, R.S.decorrelate);
            const size_t npx = (size_t)c.W * c.H;
            if (raw && raw[k]) std::memcpy(raw[k], fb.data(), npx * 4 * sizeof(float));
            if (ldr && ldr[k]) postprocess(fb, ldr[k], npx);
        }
        return;
    }
    uint64_t draws = 0;
    R.photonCounter = 0;
    R.photons.assign(R.photonCapacity, Photon{});
    const uint32_t P = std::max(1u, S.photons / R.photonStride);
    const uint32_t emitted = P * S.forwardPasses;
    const float invEmitted = 1.0f / (float)emitted;
    workClass(0);
    for (uint32_t pass = 0; pass < S.forwardPasses; ++pass) {
        uint32_t number = (uint32_t)seedGen();
        ++draws;
        std::vector<RayState> q, nq;
        q.reserve(P);
        // launchRayGenEmitterKernel (PrimalKernels.cpp:13-62)
        for (uint32_t i = 0; i < P; ++i) {
            Xorshift128 rng(perItemSeed(number, i));
            F3 pos, n, dir, power;
            if (!sampleLight(R, rng, pos, n, dir, power)) continue;
            RayState rs{};
            rs.ray.o = pos;
            rs.ray.d = dir;
            rs.T = (power * (float)R.lights.size()) * invEmitted;
            rs.bounce = 0;
            rs.pixel = kInvalid;
            q.push_back(rs);
        }
        for (uint32_t b = 0; b < S.bounces && !q.empty(); ++b) {
            number = (uint32_t)seedGen();
            ++draws;
            nq.clear();
            photonBounce(R, q, nq, number);
            q.swap(nq);
        }
    }
    R.photonCount = (uint32_t)std::min<uint64_t>(R.photonCounter, R.photonCapacity);
    R.photonDraws = draws;
    buildGrid(R);
    R.mapValid = true;
    R.stats.grid_cells = R.grid.total;
    std::vector<float> fb;
    workClass(1);
    for (size_t k = 0; k < camIdx.size(); ++k) {
        const uint32_t number = (uint32_t)seedGen();
        const Camera& c = R.cams[camIdx[k]];
        cameraGather(R, c, number, fb, R.S.gatherPasses, R.S.decorrelate);
        const size_t npx = (size_t)c.W * c.H;
        if (raw && raw[k]) std::memcpy(raw[k], fb.data(), npx * 4 * sizeof(float));
        if (ldr && ldr[k]) postprocess(fb, ldr[k], npx);
    }
}

// ---------------------------------------------------------- adjoint pass --
enum PendingKind { kNone = 0, kNullT = 1, kReflectS = 2, kTransmitS = 3, kProjection = 4 };
struct Pending {
    int kind = kNone;
    uint32_t prim = kInvalid, inst = kInvalid;
    float alpha = 0.0f;
    F3 pos{0, 0, 0}, T{0, 0, 0};
    F3 ro{0, 0, 0}, rd{0, 0, 0};  // the ray that hit `prim`
};
struct Completed {
    int kind = kNone;
    uint32_t prim = 0, inst = 0;
    float alpha = 0.0f;
    F3 pos{0, 0, 0}, normal{0, 0, 0}, T{0, 0, 0};
    F3 ro{0, 0, 0}, rd{0, 0, 0};  // the ray that hit the surfel whose alpha the event weighs
    uint32_t endInst = 0, endPrim = 0;
    int endType = -1;
    float endAlpha = 0.0f, endPdf = 0.0f, endCos = 0.0f;
    F3 endPos{0, 0, 0}, endNormal{0, 0, 0};
    uint32_t pixel = 0;  // image pixel of the path (debug images)
};
struct ARay {
    Ray ray;
    F3 T;
    uint32_t bounce, pixel, pathId;
};

// ---------------------------------------------------- full gradients --
// Forward-mode derivatives of alpha_geom (intersectSurfel, KernelHelpers.h:
// 637-669) with respect to the 14 geometric parameters of one surfel:
// center (0-2), tangent_u (3-5), tangent_v (6-8), s_u (9), s_v (10),
// beta (11).  Dual numbers carry value and gradient through the same
// expressions the test evaluates; the build's GPU code uses a hand-written
// reverse pass, so the two agree only if both are right.
struct Dual {
    double v;
    double g[12];
};
static Dual dconst(double v) {
    Dual r{};
    r.v = v;
    return r;
}
static Dual dvar(double v, int i) {
    Dual r = dconst(v);
    r.g[i] = 1.0;
    return r;
}
static Dual operator+(const Dual& a, const Dual& b) {
    Dual r{};
    r.v = a.v + b.v;
    for (int i = 0; i < 12; ++i) r.g[i] = a.g[i] + b.g[i];
    return r;
}
static Dual operator-(const Dual& a, const Dual& b) {
    Dual r{};
    r.v = a.v - b.v;
    for (int i = 0; i < 12; ++i) r.g[i] = a.g[i] - b.g[i];
    return r;
}
static Dual operator*(const Dual& a, const Dual& b) {
    Dual r{};
    r.v = a.v * b.v;
    for (int i = 0; i < 12; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
    return r;
}
static Dual operator/(const Dual& a, const Dual& b) {
    Dual r{};
    r.v = a.v / b.v;
    for (int i = 0; i < 12; ++i) r.g[i] = (a.g[i] * b.v - a.v * b.g[i]) / (b.v * b.v);
    return r;
}
static Dual dsqrt(const Dual& a) {
    Dual r{};
    r.v = std::sqrt(a.v);
    for (int i = 0; i < 12; ++i) r.g[i] = a.g[i] / (2.0 * r.v);
    return r;
}
struct D3 {
    Dual x, y, z;
};
static D3 operator+(const D3& a, const D3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static D3 operator-(const D3& a, const D3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static D3 operator*(const D3& a, const Dual& s) { return {a.x * s, a.y * s, a.z * s}; }
static Dual ddot(const D3& a, const D3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static D3 dcross(const D3& a, const D3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static D3 dnormalize(const D3& a) { return a * (dconst(1.0) / dsqrt(ddot(a, a))); }
static D3 dconst3(F3 v) { return {dconst(v.x), dconst(v.y), dconst(v.z)}; }
static D3 dvar3(F3 v, int i) { return {dvar(v.x, i), dvar(v.y, i + 1), dvar(v.z, i + 2)}; }

// alpha_geom of surfel s for the object-space ray (o, d) and its gradient
// (12 values, order above).  False if the ray does not hit the support.
static bool alphaGradient(const Point& s, F3 o, F3 d, float grad[12]) {
    const D3 p = dvar3(s.p, 0), tu = dvar3(s.tu, 3), tv = dvar3(s.tv, 6);
    const Dual su = dvar(s.su, 9), sv = dvar(s.sv, 10);
    const D3 uh = dnormalize(tu);
    const D3 vh = dnormalize(tv - uh * ddot(uh, tv));
    const D3 n = dnormalize(dcross(uh, vh));
    const D3 O = dconst3(o), Dd = dconst3(d);
    const Dual nd = ddot(n, Dd);
    const Dual t = ddot(n, p - O) / nd;
    const D3 r = (O + Dd * t) - p;
    const Dual u = ddot(tu, r) / su, v = ddot(tv, r) / sv;
    const Dual R2 = u * u + v * v;
    if (!(R2.v < 1.0)) return false;
    const double k = 4.0 * std::exp((double)s.beta);
    const double a = std::pow(1.0 - R2.v, k);
    const double dadR2 = -k * a / (1.0 - R2.v);
    for (int i = 0; i < 11; ++i) grad[i] = (float)(dadR2 * R2.g[i]);
    grad[11] = (float)(a * k * std::log(1.0 - R2.v));
    return true;
}

// Adds the full gradient of one event: `w` is the event's scalar weight
// sum_rgb(T * Lo) * invSpp (the factor of alpha_geom * eta in its
// contribution), `sgn` its sign.
// What an event's gradient needs besides its own record: the irradiance at
// its shading point and, with full gradients, d(alpha_geom)/d(geometry) of
// the surfel it weighs.  Computed for all events of a bounce in parallel.
struct EventPre {
    F3 E{0, 0, 0};
    float g[12] = {};
    bool gv = false;
};
static void geometryGradient(const Renderer& R, uint32_t prim, F3 ro, F3 rd, EventPre& pre) {
    const Point& s = R.points[prim];
    Ray wr;
    wr.o = ro;
    wr.d = rd;
    const Ray obj = toObjectSpace(wr, R.xfs[R.inst[R.cloudInstance].xf]);
    pre.gv = alphaGradient(s, obj.o, obj.d, pre.g);
}
static void addGeometryGradient(Renderer& R, uint32_t prim, const EventPre& pre, float w) {
    const Point& s = R.points[prim];
    if (!pre.gv) return;
    const float* g = pre.g;
    float* out = &R.gradFull[(size_t)prim * 16];
    const float f = s.opacity * w;
    for (int i = 0; i < 11; ++i) out[i] += g[i] * f;
    out[15] += g[11] * f;
}

// launchAdjointTransportKernel (AdjointKernels.cpp:389-509), one event.
// E is the irradiance gathered at the event's shading point (the surfel hit
// for projection events, the endpoint otherwise).
// Returns the event's opacity-gradient contribution (what the reference also
// adds to the debug opacity image, AdjointKernels.cpp:434-505).
static float transport(Renderer& R, const Completed& c, float invSpp, const EventPre& pre) {
    const F3 E = pre.E;
    float val = 0.0f;
    if (c.kind == kProjection) {
        const Point& s = R.points[c.prim];
        const F3 fr = (s.alpha_r * s.albedo) * kInvPi;
        const F3 Lr = fr * E;
        const F3 g = (c.alpha * c.T) * Lr;
        val = fsum(g) * invSpp;
        R.gradOpacity[c.prim] += val;
        if (R.S.fullGradients) {
            addGeometryGradient(R, c.prim, pre, fsum(c.T * Lr) * invSpp);
            // d/d(albedo_c) of alpha * eta * (alpha_r * albedo_c / pi) * E_c
            const F3 ga = ((c.alpha * s.opacity * s.alpha_r * kInvPi) * (c.T * E)) * invSpp;
            float* out = &R.gradFull[(size_t)c.prim * 16];
            out[11] += ga.x;
            out[12] += ga.y;
            out[13] += ga.z;
        }
    }
    if (c.kind == kNullT) {
        F3 base;
        float power;
        if (c.endType == kMesh) {
            const Material& m = R.materials[R.inst[c.endInst].material];
            base = m.base;
            power = m.power;
        } else {
            const Point& e = R.points[c.endPrim];
            base = e.alpha_r * e.albedo;
            power = 0.0f;
        }
        const F3 Lr = (E * base) * kInvPi;
        F3 Le = mk(0, 0, 0);
        if (power > 0.0f) {
            const Light& L0 = R.lights[0];
            Le = (base * (power / (kPi * L0.area))) * c.endCos;
        }
        const F3 Lo = Le + Lr;
        const F3 g = ((-c.alpha) * c.T) * Lo;
        val = fsum(g) * invSpp;
        R.gradOpacity[c.prim] += val;
        if (R.S.fullGradients) addGeometryGradient(R, c.prim, pre, -(fsum(c.T * Lo) * invSpp));
    }
    if (c.kind == kReflectS) {
        const Material& m = R.materials[R.inst[c.inst].material];
        const F3 Lo = (E * m.base) * kInvPi;
        const F3 p = divs(c.T, c.endPdf);
        const F3 g = (p * Lo) * c.endAlpha;
        val = fsum(g) * invSpp;
        R.gradOpacity[c.endPrim] += val;
        if (R.S.fullGradients) addGeometryGradient(R, c.endPrim, pre, fsum(p * Lo) * invSpp);
    }
    return val;
}

// submitAdjointKernel (SyclBridge.cpp:169-294) for the given cameras.  The
// adjoint kernels seed from PathTracerSettings::random.number, default 42
// (AdjointKernels.cpp:27,94; GPUDataStructures.h:395-398; setting
// adjoint_seed).  With adjoint_decorrelate (no reference counterpart)
// sample k seeds from adjoint_seed + k instead, so that the samples of a
// pixel are independent.
static void renderBackward(Renderer& R, const std::vector<const Camera*>& camList,
                           const std::vector<const float*>& dloss, uint32_t passes) {
    const Settings& S = R.S;
    workClass(2);
    R.gradOpacity.assign(R.points.size(), 0.0f);
    R.gradFull.assign(S.fullGradients ? R.points.size() * 16 : 0, 0.0f);
    // Debug opacity images of these cameras, cleared per call
    // (SyclBridge.cpp:183-205).
    R.debugOff.assign(1, 0);
    R.debugNames.clear();
    for (const Camera* c : camList) {
        R.debugOff.push_back(R.debugOff.back() + (size_t)c->W * c->H);
        R.debugNames.push_back(c->name);
    }
    R.debugImg.assign(S.debugImages ? R.debugOff.back() * 4 : 0, 0.0f);
    const float invSpp = 1.0f / (float)passes;
    for (size_t k = 0; k < camList.size(); ++k) {
        const Camera& c = *camList[k];
        const CamConst kc = camConst(c);
        const uint32_t W = c.W, H = c.H, npx = W * H;
        for (uint32_t spp = 0; spp < passes; ++spp) {
            const uint64_t number = S.adjointDecorrelate ? (uint64_t)(uint32_t)(S.adjointSeed + spp)
                                                         : (uint64_t)S.adjointSeed;
            // launchRayGenAdjointKernel (AdjointKernels.cpp:15-84)
            std::vector<ARay> q, nq;
            q.reserve(npx / R.pixelStride + 1);
            for (uint32_t pi = 0; pi < npx; ++pi) {
                if (pi % R.pixelStride != R.pixelOffset) continue;
                Xorshift128 rng(perItemSeed(number, pi));
                const float jx = rng.nextFloat() - 0.5f;
                const float jy = rng.nextFloat() - 0.5f;
                ARay a{};
                a.ray = primaryRay(c, kc, (float)(pi % W), (float)(pi / W), jx, jy);
                a.T = mk(dloss[k][(size_t)pi * 3 + 0], dloss[k][(size_t)pi * 3 + 1], dloss[k][(size_t)pi * 3 + 2]);
                a.bounce = 0;
                a.pixel = pi;
                a.pathId = pi;
                q.push_back(a);
            }
            std::vector<Pending> pend(npx);
            for (uint32_t b = 0; b < S.adjointBounces && !q.empty(); ++b) {
                nq.clear();
                std::vector<Completed> done;
                // launchAdjointIntersectKernel (AdjointKernels.cpp:87-386):
                // closest hits of all rays (in parallel), then each ray in
                // queue order.
                std::vector<WorldHit> hits(q.size());
                std::vector<char> hit(q.size());
                parallelItems(q.size(), [&](size_t i) { hit[i] = intersectScene(R, q[i].ray, hits[i]) ? 1 : 0; });
                R.stats.adjoint_queries += q.size();
                for (uint32_t ri = 0; ri < q.size(); ++ri) {
                    Xorshift128 rng(perItemSeed(number, ri));
                    const ARay& rs = q[ri];
                    if (!hit[ri]) continue;
                    const WorldHit& h = hits[ri];
                    const Instance& ein = R.inst[h.inst];
                    F3 no = mk(0, 0, 0), dir = mk(0, 0, 0);
                    float pdf = 0.0f;
                    if (ein.type == kMesh) {
                        sampleCosine(rng, h.n, dir, pdf);
                        no = h.n;
                    } else {
                        const Point& s = R.points[h.prim];
                        const F3 nc = normalize(cross(s.tu, s.tv));
                        const float sc = dot(nc, -rs.ray.d);
                        const int sign = sc >= 0.0f ? 1 : -1;
                        no = (float)sign * nc;
                        sampleUniformHemi(rng, no, dir, pdf);
                    }
                    if (rs.pathId < R.rayCapacity) {
                        Pending p = pend[rs.pathId];
                        const bool reflectDone = p.kind == kReflectS && ein.type == kCloud;
                        const bool nullDone = p.kind == kNullT;
                        if (reflectDone || nullDone) {
                            Completed e{};
                            e.kind = p.kind;
                            e.prim = p.prim;
                            e.inst = p.inst;
                            e.alpha = p.alpha;
                            e.pos = p.pos;
                            e.T = p.T * (reflectDone ? S.qReflect : S.qNull);
                            e.endAlpha = h.alpha;
                            e.endInst = h.inst;
                            e.endPrim = h.prim;
                            e.endPos = h.pos;
                            e.endNormal = no;
                            e.endPdf = pdf;
                            e.endType = ein.type;
                            e.endCos = std::fmax(0.0f, dot(dir, no));
                            // Null: the ray that crossed `prim`; ReflectScatter:
                            // this ray, which hit the endpoint surfel.
                            e.ro = nullDone ? p.ro : rs.ray.o;
                            e.rd = nullDone ? p.rd : rs.ray.d;
                            e.pixel = rs.pixel;
                            done.push_back(e);
                            pend[rs.pathId] = Pending{};
                        }
                    }
                    if (ein.type == kMesh) {
                        const Material& m = R.materials[ein.material];
                        ARay nx{};
                        nx.ray.o = h.pos + h.n * 1e-6f;
                        nx.ray.d = dir;
                        nx.bounce = rs.bounce + 1;
                        nx.pixel = rs.pixel;
                        nx.pathId = rs.pathId;
                        nx.T = rs.T * m.base;
                        if (!russianRoulette(rng, nx.bounce, nx.T, S.rrStart)) continue;
                        if (rs.pathId < R.rayCapacity) {
                            Pending p{};
                            p.kind = kReflectS;
                            p.inst = h.inst;
                            p.pos = h.pos;
                            p.T = rs.T * m.base;
                            pend[rs.pathId] = p;
                        }
                        nq.push_back(nx);
                    } else {
                        const Point& s = R.points[h.prim];
                        const float u = rng.nextFloat();
                        if (u < S.qNull) {
                            const float att = 1.0f - h.alpha * s.opacity;
                            const float w = att / S.qNull;
                            ARay nx{};
                            nx.ray.o = h.pos + rs.ray.d * 1e-5f;
                            nx.ray.d = rs.ray.d;
                            nx.bounce = rs.bounce + 1;
                            nx.pixel = rs.pixel;
                            nx.pathId = rs.pathId;
                            nx.T = rs.T * w;
                            if (!russianRoulette(rng, nx.bounce, nx.T, S.rrStart)) continue;
                            if (rs.pathId < R.rayCapacity) {
                                Pending p{};
                                p.kind = kNullT;
                                p.prim = h.prim;
                                p.alpha = h.alpha;
                                p.pos = h.pos;
                                p.T = divs(rs.T, S.qNull);
                                p.ro = rs.ray.o;
                                p.rd = rs.ray.d;
                                pend[rs.pathId] = p;
                            }
                            nq.push_back(nx);
                        } else if (u < S.qNull + S.qReflect) {
                            const F3 fs = (s.alpha_r * s.albedo) * kInvPi;
                            const float cosT = std::fmax(0.0f, dot(dir, no));
                            const float alpha = h.alpha * s.opacity;
                            const F3 mult = divs((alpha / S.qReflect) * (fs * cosT), pdf);
                            ARay nx{};
                            nx.ray.o = h.pos + no * 1e-5f;
                            nx.ray.d = dir;
                            nx.bounce = rs.bounce + 1;
                            nx.pixel = rs.pixel;
                            nx.pathId = rs.pathId;
                            nx.T = rs.T * mult;
                            if (!russianRoulette(rng, nx.bounce, nx.T, S.rrStart)) continue;
                            if (b == 0 && rs.pathId < R.rayCapacity) {
                                Completed e{};
                                e.kind = kProjection;
                                e.prim = h.prim;
                                e.alpha = h.alpha;
                                e.pos = h.pos;
                                e.normal = no;
                                e.T = divs(rs.T, S.qReflect);
                                e.ro = rs.ray.o;
                                e.rd = rs.ray.d;
                                e.pixel = rs.pixel;
                                done.push_back(e);
                            }
                            nq.push_back(nx);
                        } else if (u < S.qNull + S.qReflect + S.qTransmit) {
                            const float alpha = h.alpha * s.opacity;
                            const float w = (alpha * s.alpha_t) / S.qTransmit;
                            const F3 nc = normalize(cross(s.tu, s.tv));
                            const float sc = dot(nc, -rs.ray.d);
                            const int sign = sc >= 0.0f ? 1 : -1;
                            const F3 ot = (float)sign * nc;
                            F3 d2 = rs.ray.d;
                            float p2 = 0.0f;
                            sampleUniformHemi(rng, ot, d2, p2);
                            ARay nx{};
                            nx.ray.o = h.pos + (-ot) * 1e-5f;
                            nx.ray.d = d2;
                            nx.bounce = rs.bounce + 1;
                            nx.pixel = rs.pixel;
                            nx.pathId = rs.pathId;
                            nx.T = (rs.T * w) * s.albedo;
                            if (!russianRoulette(rng, nx.bounce, nx.T, S.rrStart)) continue;
                            nq.push_back(nx);
                        }
                    }
                }
                R.stats.gradient_events += done.size();
                // Irradiance at every event (in parallel), then the gradient
                // sums in event order.
                std::vector<EventPre> pre(done.size());
                parallelItems(done.size(), [&](size_t i) {
                    const Completed& e = done[i];
                    pre[i].E = e.kind == kProjection ? gather(R, e.pos, e.normal) : gather(R, e.endPos, e.endNormal);
                    if (S.fullGradients)
                        geometryGradient(R, e.kind == kReflectS ? e.endPrim : e.prim, e.ro, e.rd, pre[i]);
                });
                for (size_t i = 0; i < done.size(); ++i) {
                    const float v = transport(R, done[i], invSpp, pre[i]);
                    if (S.debugImages) {
                        // atomicAddFloat4ToImage (KernelHelpers.h:1489-1502)
                        float* px = &R.debugImg[4 * (R.debugOff[k] + done[i].pixel)];
                        px[0] += v;
                        px[1] += v;
                        px[2] += v;
                        px[3] = 1.0f;
                    }
                }
                q.swap(nq);
            }
        }
    }
}

// ------------------------------------------------------------ host logic --
static Settings toSettings(const pale_settings& s) {
    Settings S{};
    S.photons = s.photons;
    S.bounces = s.bounces;
    S.forwardPasses = s.forward_passes;
    S.gatherPasses = s.gather_passes;
    S.adjointBounces = s.adjoint_bounces;
    S.adjointPasses = s.adjoint_passes;
    S.seed = s.seed;
    S.radius = s.gather_radius;
    S.rrStart = s.rr_start;
    S.qNull = s.q_null;
    S.qReflect = s.q_reflect;
    S.qTransmit = s.q_transmit;
    S.decorrelate = s.spp_decorrelate != 0;
    S.mapCapacity = s.photon_map_capacity;
    S.fullGradients = s.full_gradients != 0;
    S.adjointDecorrelate = s.adjoint_decorrelate != 0;
    S.adjointSeed = s.adjoint_seed;
    S.debugImages = s.debug_images != 0;
    return S;
}

static M4 toM4(const float* m) {
    M4 r;
    std::memcpy(r.m, m, sizeof(r.m));
    return r;
}

static std::vector<AssetPoint> toAsset(const pale_points_desc& p) {
    std::vector<AssetPoint> out(p.count);
    for (uint32_t i = 0; i < p.count; ++i) {
        AssetPoint a{};
        a.p = mk(p.position[3 * i], p.position[3 * i + 1], p.position[3 * i + 2]);
        a.tu = mk(p.tangent_u[3 * i], p.tangent_u[3 * i + 1], p.tangent_u[3 * i + 2]);
        a.tv = mk(p.tangent_v[3 * i], p.tangent_v[3 * i + 1], p.tangent_v[3 * i + 2]);
        a.su = p.scale[2 * i];
        a.sv = p.scale[2 * i + 1];
        a.albedo = mk(p.albedo[3 * i], p.albedo[3 * i + 1], p.albedo[3 * i + 2]);
        a.opacity = p.opacity[i];
        a.beta = p.beta[i];
        a.shape = p.shape ? p.shape[i] : 0.0f;
        out[i] = a;
    }
    return out;
}

// rebuild_bvh (pale_bindings.cpp:1058-1079)
static void rebuild(Renderer& R) {
    buildScene(R);
    setScene(R);
}

}  // namespace orc

using namespace orc;

static Camera cameraFrom(const pale_camera_desc& cd) {
    Camera c;
    c.name = std::string(cd.name, strnlen(cd.name, PALE_NAME_LEN));
    c.W = cd.width;
    c.H = cd.height;
    c.fovy = cd.fovy_deg;
    c.view = toM4(cd.view);
    c.invView = toM4(cd.inv_view);
    c.pos = mk(cd.position[0], cd.position[1], cd.position[2]);
    c.fwd = mk(cd.forward[0], cd.forward[1], cd.forward[2]);
    c.adjoint = cd.use_for_adjoint != 0;
    return c;
}

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void* orc_create(const pale_scene_desc* scene, const pale_settings* settings) {
    try {
        if (!scene || !settings) throw std::runtime_error("null scene or settings");
        Renderer* R = new Renderer();
        R->S = toSettings(*settings);
        for (uint32_t i = 0; i < scene->mesh_count; ++i) {
            const pale_mesh_desc& md = scene->meshes[i];
            MeshData m;
            for (uint32_t v = 0; v < md.vertex_count; ++v)
                m.verts.push_back(mk(md.positions[3 * v], md.positions[3 * v + 1], md.positions[3 * v + 2]));
            m.idx.assign(md.indices, md.indices + 3 * md.triangle_count);
            m.o2w = toM4(md.object_to_world);
            m.w2o = toM4(md.world_to_object);
            m.mat = Material{mk(md.base_color[0], md.base_color[1], md.base_color[2]), md.power};
            R->meshes.push_back(m);
        }
        for (uint32_t i = 0; i < scene->camera_count; ++i) R->cams.push_back(cameraFrom(scene->cameras[i]));
        R->asset = toAsset(scene->points);
        rebuild(*R);
        return R;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_destroy(void* h) { delete static_cast<Renderer*>(h); }

uint32_t orc_point_count(void* h) { return (uint32_t)static_cast<Renderer*>(h)->points.size(); }

int orc_get_points(void* h, float* pos, float* tu, float* tv, float* sc, float* alb, float* op, float* be,
                   float* sh) {
    Renderer& R = *static_cast<Renderer*>(h);
    for (size_t i = 0; i < R.points.size(); ++i) {
        const Point& p = R.points[i];
        if (pos) { pos[3 * i] = p.p.x; pos[3 * i + 1] = p.p.y; pos[3 * i + 2] = p.p.z; }
        if (tu) { tu[3 * i] = p.tu.x; tu[3 * i + 1] = p.tu.y; tu[3 * i + 2] = p.tu.z; }
        if (tv) { tv[3 * i] = p.tv.x; tv[3 * i + 1] = p.tv.y; tv[3 * i + 2] = p.tv.z; }
        if (sc) { sc[2 * i] = p.su; sc[2 * i + 1] = p.sv; }
        if (alb) { alb[3 * i] = p.albedo.x; alb[3 * i + 1] = p.albedo.y; alb[3 * i + 2] = p.albedo.z; }
        if (op) op[i] = p.opacity;
        if (be) be[i] = p.beta;
        if (sh) sh[i] = p.shape;
    }
    return 0;
}

int orc_get_surfel_bvh(void* h, uint32_t* nodeCount, float* bounds, uint32_t* leftFirst, uint32_t* count,
                       uint32_t* perm) {
    Renderer& R = *static_cast<Renderer*>(h);
    const Range br = R.blasRanges.back();
    *nodeCount = br.count;
    for (uint32_t i = 0; bounds && i < br.count; ++i) {
        const Node& n = R.blas[br.first + i];
        float* b = bounds + 6 * i;
        b[0] = n.mn.x; b[1] = n.mn.y; b[2] = n.mn.z;
        b[3] = n.mx.x; b[4] = n.mx.y; b[5] = n.mx.z;
        leftFirst[i] = n.leftFirst;
        count[i] = n.count;
    }
    if (perm) std::memcpy(perm, R.perm.data(), R.perm.size() * sizeof(uint32_t));
    return 0;
}

int orc_trace_rays(void* h, uint32_t n, const float* o, const float* d, int32_t* inst, int32_t* prim, float* t,
                   float* pos, float* alpha) {
    Renderer& R = *static_cast<Renderer*>(h);
    for (uint32_t i = 0; i < n; ++i) {
        Ray r;
        r.o = mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
        r.d = mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        WorldHit hh;
        const bool hit = intersectScene(R, r, hh);
        inst[i] = hit ? (int32_t)hh.inst : -1;
        prim[i] = hit ? (int32_t)hh.prim : -1;
        t[i] = hit ? hh.t : 0.0f;
        pos[3 * i] = hit ? hh.pos.x : 0.0f;
        pos[3 * i + 1] = hit ? hh.pos.y : 0.0f;
        pos[3 * i + 2] = hit ? hh.pos.z : 0.0f;
        alpha[i] = hit ? hh.alpha : 0.0f;
    }
    return 0;
}

static void copyWork(Renderer& R) {
    for (int c = 0; c < 3; ++c) {
        const Work& w = g_workClass[c];
        const uint64_t v[7] = {w.pops, w.aabb, w.prims, w.gathers, w.cells, w.cellsNE, w.photons};
        for (int u = 0; u < 7; ++u) R.stats.work[c][u] = v[u];
    }
}

static std::vector<uint32_t> selectCams(Renderer& R, const char* name) {
    std::vector<uint32_t> idx;
    const std::string nm = name ? name : "";
    for (uint32_t i = 0; i < R.cams.size(); ++i)
        if (nm.empty() || R.cams[i].name == nm) idx.push_back(i);
    return idx;
}

int orc_render_forward(void* h, const char* cam, float* const* ldr, float* const* raw, uint32_t* rendered) {
    try {
        Renderer& R = *static_cast<Renderer*>(h);
        R.stats = pale_stats{};
        for (Work& w : g_workClass) w = Work{};
        const std::vector<uint32_t> idx = selectCams(R, cam);
        const auto t0 = std::chrono::steady_clock::now();
        renderForward(R, idx, ldr, raw);
        R.stats.forward_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        copyWork(R);
        if (rendered) *rendered = (uint32_t)idx.size();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PALE_ERR_INTERNAL;
    }
}

int orc_render_backward(void* h, uint32_t n, const char* const* names, const float* const* dloss, float* opacityGrad) {
    try {
        Renderer& R = *static_cast<Renderer*>(h);
        std::vector<uint32_t> idx;
        std::vector<const float*> src;
        for (uint32_t c = 0; c < R.cams.size(); ++c) {
            if (!R.cams[c].adjoint) continue;
            for (uint32_t k = 0; k < n; ++k)
                if (R.cams[c].name == names[k]) {
                    idx.push_back(c);
                    src.push_back(dloss[k]);
                    break;
                }
        }
        for (uint32_t c : idx) {
            const uint32_t need = R.cams[c].W * R.cams[c].H;
            uint32_t cap = 1;
            while (cap < need) cap <<= 1u;
            R.rayCapacity = std::max(R.rayCapacity, cap);
        }
        R.stats.adjoint_queries = 0;
        R.stats.gradient_events = 0;
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<const Camera*> camList;
        for (uint32_t c : idx) camList.push_back(&R.cams[c]);
        renderBackward(R, camList, src, R.S.adjointPasses);
        R.stats.backward_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        copyWork(R);
        if (opacityGrad) std::memcpy(opacityGrad, R.gradOpacity.data(), R.gradOpacity.size() * sizeof(float));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PALE_ERR_INTERNAL;
    }
}

// Served views (no reference counterpart): the photon stage alone, then any
// camera over that map with spp decorrelated passes seeded seed + pass, and
// its adjoint with spp samples.
int orc_build_photon_map(void* h) {
    try {
        Renderer& R = *static_cast<Renderer*>(h);
        R.stats = pale_stats{};
        for (Work& w : g_workClass) w = Work{};
        const bool frozen = R.freezeMap;
        R.freezeMap = false;
        renderForward(R, {}, nullptr, nullptr);
        R.freezeMap = frozen;
        copyWork(R);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PALE_ERR_INTERNAL;
    }
}

int orc_render_view(void* h, const pale_camera_desc* cd, uint32_t spp, uint32_t seed, float* ldr, float* raw) {
    try {
        Renderer& R = *static_cast<Renderer*>(h);
        if (!R.mapValid || spp == 0) return PALE_ERR_INVALID_ARG;
        const Camera c = cameraFrom(*cd);
        std::vector<float> fb;
        workClass(1);
        cameraGather(R, c, seed, fb, spp, true);
        const size_t npx = (size_t)c.W * c.H;
        if (raw) std::memcpy(raw, fb.data(), npx * 4 * sizeof(float));
        if (ldr) postprocess(fb, ldr, npx);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PALE_ERR_INTERNAL;
    }
}

int orc_render_view_backward(void* h, const pale_camera_desc* cd, uint32_t spp, const float* dloss,
                             float* opacityGrad) {
    try {
        Renderer& R = *static_cast<Renderer*>(h);
        if (!R.mapValid || spp == 0) return PALE_ERR_INVALID_ARG;
        const Camera c = cameraFrom(*cd);
        uint32_t cap = 1;
        while (cap < c.W * c.H) cap <<= 1u;
        R.rayCapacity = std::max(R.rayCapacity, cap);
        renderBackward(R, {&c}, {dloss}, spp);
        if (opacityGrad) std::memcpy(opacityGrad, R.gradOpacity.data(), R.gradOpacity.size() * sizeof(float));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PALE_ERR_INTERNAL;
    }
}

// Full gradients of the last backward pass, 16 floats per surfel: position
// (3), tangent_u (3), tangent_v (3), scale (2), albedo (3), unused, beta.
int orc_get_debug_image(void* h, const char* name, float* out) {
    const Renderer& R = *static_cast<Renderer*>(h);
    if (!R.S.debugImages) return -1;
    for (size_t k = 0; k < R.debugNames.size(); ++k) {
        if (R.debugNames[k] != name) continue;
        std::memcpy(out, R.debugImg.data() + 4 * R.debugOff[k], (R.debugOff[k + 1] - R.debugOff[k]) * 4 * sizeof(float));
        return 0;
    }
    return -1;
}

int orc_full_gradients(void* h, float* out) {
    try {
        Renderer& R = *static_cast<Renderer*>(h);
        if (R.gradFull.empty()) return PALE_ERR_INVALID_ARG;
        std::memcpy(out, R.gradFull.data(), R.gradFull.size() * sizeof(float));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PALE_ERR_INTERNAL;
    }
}

int orc_get_photons(void* h, uint32_t* count, uint64_t* attempted, float* pos, float* power, float* dir) {
    Renderer& R = *static_cast<Renderer*>(h);
    *count = R.photonCount;
    *attempted = R.photonCounter;
    for (uint32_t i = 0; i < R.photonCount; ++i) {
        const Photon& p = R.photons[i];
        if (pos) { pos[3 * i] = p.pos.x; pos[3 * i + 1] = p.pos.y; pos[3 * i + 2] = p.pos.z; }
        if (power) { power[3 * i] = p.power.x; power[3 * i + 1] = p.power.y; power[3 * i + 2] = p.power.z; }
        if (dir) { dir[3 * i] = p.dir.x; dir[3 * i + 1] = p.dir.y; dir[3 * i + 2] = p.dir.z; }
    }
    return 0;
}

// Grid geometry and the irradiance estimate at arbitrary query points
// (used to pin the gather against brute force).
int orc_grid_info(void* h, float* origin, float* cell, int32_t* res) {
    Renderer& R = *static_cast<Renderer*>(h);
    origin[0] = R.grid.origin.x; origin[1] = R.grid.origin.y; origin[2] = R.grid.origin.z;
    cell[0] = R.grid.cell.x; cell[1] = R.grid.cell.y; cell[2] = R.grid.cell.z;
    res[0] = R.grid.res[0]; res[1] = R.grid.res[1]; res[2] = R.grid.res[2];
    return 0;
}

int orc_gather(void* h, uint32_t n, const float* q, const float* nrm, float* out) {
    Renderer& R = *static_cast<Renderer*>(h);
    for (uint32_t i = 0; i < n; ++i) {
        const F3 E = gather(R, mk(q[3 * i], q[3 * i + 1], q[3 * i + 2]), mk(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]));
        out[3 * i] = E.x;
        out[3 * i + 1] = E.y;
        out[3 * i + 2] = E.z;
    }
    return 0;
}

int orc_get_stats(void* h, pale_stats* out) {
    *out = static_cast<Renderer*>(h)->stats;
    return 0;
}

// Host-side boundary logic mirrored for the parity tests of the shim.
int orc_apply_points(void* h, const pale_points_desc* p) {
    Renderer& R = *static_cast<Renderer*>(h);
    if (p->count != R.points.size()) {
        g_err = "apply_point_optimization expects consistent point count";
        return PALE_ERR_SIZE;
    }
    for (uint32_t i = 0; i < p->count; ++i) {
        Point& q = R.points[i];
        q.p = mk(p->position[3 * i], p->position[3 * i + 1], p->position[3 * i + 2]);
        q.tu = mk(p->tangent_u[3 * i], p->tangent_u[3 * i + 1], p->tangent_u[3 * i + 2]);
        q.tv = mk(p->tangent_v[3 * i], p->tangent_v[3 * i + 1], p->tangent_v[3 * i + 2]);
        q.su = p->scale[2 * i];
        q.sv = p->scale[2 * i + 1];
        q.albedo = mk(p->albedo[3 * i], p->albedo[3 * i + 1], p->albedo[3 * i + 2]);
        q.opacity = p->opacity[i];
        q.beta = p->beta[i];
        AssetPoint& a = R.asset[i];
        a.p = q.p;
        a.tu = q.tu;
        a.tv = q.tv;
        a.su = q.su;
        a.sv = q.sv;
        a.albedo = q.albedo;
        a.opacity = q.opacity;
        a.beta = q.beta;
    }
    setScene(R);
    return 0;
}

int orc_freeze_map(void* h, int32_t on) {
    static_cast<Renderer*>(h)->freezeMap = on != 0;
    return 0;
}

// Threads of the oracle's parallel loops in this process (0 = OpenMP's
// default); results do not depend on it.
int orc_set_threads(int32_t n) {
    g_threads = n > 0 ? n : 0;
    return 0;
}

// Diagnostic for the tests: applyRussianRoulette (KernelHelpers.h:1612-1657)
// on n throughputs T[3*i..], each with the per-item stream (seed, i).  Out:
// alive[i], the throughput after the call, and next[i] = the stream's next
// float after the call (which tells how many draws the call consumed).
int orc_russian_roulette(uint32_t n, const float* T, uint32_t bounce, uint32_t rrStart, uint64_t seed,
                         int32_t* alive, float* Tout, float* next) {
    for (uint32_t i = 0; i < n; ++i) {
        Xorshift128 rng(perItemSeed(seed, i));
        F3 t = mk(T[3 * i], T[3 * i + 1], T[3 * i + 2]);
        alive[i] = russianRoulette(rng, bounce, t, rrStart) ? 1 : 0;
        Tout[3 * i] = t.x;
        Tout[3 * i + 1] = t.y;
        Tout[3 * i + 2] = t.z;
        next[i] = rng.nextFloat();
    }
    return 0;
}

int orc_set_sample(void* h, uint32_t photonStride, uint32_t pixelStride, uint32_t pixelOffset) {
    Renderer& R = *static_cast<Renderer*>(h);
    R.photonStride = std::max(1u, photonStride);
    R.pixelStride = std::max(1u, pixelStride);
    R.pixelOffset = pixelOffset % R.pixelStride;
    return 0;
}

int orc_rebuild_bvh(void* h) {
    rebuild(*static_cast<Renderer*>(h));
    return 0;
}

// The asset replaced by the given points (any count), then rebuilt, with the
// photon grid and capacities kept from the first build: what the binding's
// add_points / remove_points / set_point_* leave behind (they edit the
// PointAsset and rebuild, pale_bindings.cpp:1081-1587; PathTracer::setScene
// configures the grid once, PathTracer.cpp:22-53).  Shape is taken as given.
int orc_replace_points(void* h, const pale_points_desc* p) {
    try {
        Renderer& R = *static_cast<Renderer*>(h);
        R.asset = toAsset(*p);
        rebuild(R);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PALE_ERR_INTERNAL;
    }
}


// RNG primitives, exposed for known-answer tests.
void orc_splitmix64(uint64_t seed, uint32_t n, uint64_t* out) {
    SplitMix64 sm(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = sm.next();
}
void orc_xorshift128_state(uint32_t x, uint32_t y, uint32_t z, uint32_t w, uint32_t n, uint32_t* out) {
    Xorshift128 r(0);
    r.x = x; r.y = y; r.z = z; r.w = w;
    for (uint32_t i = 0; i < n; ++i) out[i] = r.nextUint();
}
void orc_item_rng(uint64_t base, uint64_t id, uint32_t n, float* out) {
    Xorshift128 r(perItemSeed(base, id));
    for (uint32_t i = 0; i < n; ++i) out[i] = r.nextFloat();
}
void orc_mt19937_64(uint64_t seed, uint32_t n, uint64_t* out) {
    std::mt19937_64 g(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = g();
}

}  // extern "C"
