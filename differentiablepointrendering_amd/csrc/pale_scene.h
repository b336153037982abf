// pale_scene.h -- host-side scene model and build products.
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <exception>
#include <functional>
#include <mutex>
#include <unistd.h>
#include <string>
#include <thread>
#include <vector>

#include "pale_common.h"

namespace pale {

// A surfel as the caller/asset holds it (PointGeometry of the reference's
// PLY asset): tangents are not yet normalised, values not yet clamped.
struct AssetPoint {
    F3 p, tu, tv;
    float su, sv;
    F3 albedo;
    float opacity, beta, shape;
};

// A surfel after collectPointCloudGeometry (SceneBuild.cpp:181-227).
struct HostPoint {
    F3 p, tu, tv;
    float su, sv;
    F3 albedo;
    float alpha_t, alpha_r, opacity, beta, shape;
};

// The per-surfel records of the host scene build and of the device build,
// one definition for both so that they agree bit for bit.

// surfelObjectAabbBeta (SceneBuild.cpp:395-444): the surfel's box.
PALE_HD void surfelBoxOf(const HostPoint& s, F3& mn, F3& mx) {
    const F3 tu = normalize(s.tu);
    const F3 tv = normalize(s.tv);
    const F3 n = normalize(cross(tu, tv));
    const float rU = 1.0001f * s.su;
    const float rV = 1.0001f * s.sv;
    const float cu[3] = {tu.x, tu.y, tu.z}, cv[3] = {tv.x, tv.y, tv.z}, cn[3] = {n.x, n.y, n.z};
    float h[3];
    for (int a = 0; a < 3; ++a)
        h[a] = sqrtf((rU * cu[a]) * (rU * cu[a]) + (rV * cv[a]) * (rV * cv[a])) + 0.001f * fabsf(cn[a]);
    const F3 hv = mk3(h[0], h[1], h[2]);
    mn = s.p - hv;
    mx = s.p + hv;
}

// Geometry record: intersectSurfel's frame (KernelHelpers.h:646-648) and
// the opacity exponent 4 e^beta.
PALE_HD SurfelGeo surfelGeoOf(const HostPoint& s) {
    const F3 u = normalize(s.tu);
    const F3 v = normalize(s.tv - u * dot(u, s.tv));
    const F3 n = normalize(cross(u, v));
    SurfelGeo g{};
    g.px = s.p.x; g.py = s.p.y; g.pz = s.p.z; g.su = s.su;
    g.nx = n.x; g.ny = n.y; g.nz = n.z; g.sv = s.sv;
    g.tux = s.tu.x; g.tuy = s.tu.y; g.tuz = s.tu.z; g.e4b = 4.0f * exp_r(s.beta);
    g.tvx = s.tv.x; g.tvy = s.tv.y; g.tvz = s.tv.z; g.pad = 0.0f;
    return g;
}

// Shading record: the canonical normal normalize(cross(tanU, tanV))
// (KernelHelpers.h:497-501), opacity, albedo, alpha_r.
PALE_HD SurfelShade surfelShadeOf(const HostPoint& s) {
    const F3 ns = normalize(cross(s.tu, s.tv));
    SurfelShade h{};
    h.nsx = ns.x; h.nsy = ns.y; h.nsz = ns.z; h.opacity = s.opacity;
    h.ar = s.albedo.x; h.ag = s.albedo.y; h.ab = s.albedo.z; h.alpha_r = s.alpha_r;
    return h;
}

struct HostMesh {
    std::vector<F3> verts;
    std::vector<uint32_t> idx;
    float o2w[16];
    float w2o[16];
    F3 base;
    float power;
    std::string name;
};

struct HostCamera {
    std::string name;
    uint32_t W, H;
    float fovy;
    float view[16], invView[16];
    F3 pos, fwd;
    bool adjoint;
    float fx, fy;  // focal lengths of makePrimaryRayFromPixelJitteredFov
};

struct RefNode {  // BVHNode (GPUDataStructures.h:53-63)
    F3 mn, mx;
    uint32_t leftFirst, count;
};

struct Range {
    uint32_t first, count;
};

struct HostTri {
    uint32_t v0, v1, v2;
    F3 centroid;
};

struct HostInstance {
    uint32_t type, geom, material, blas;
    float o2w[16], w2o[16];
};

struct HostLight {
    uint32_t instance, triOffset, triCount;
    F3 color;
    float power, area;
};

struct HostEmTri {
    uint32_t tri;
    float area, cdf;
};

// SceneBuild::BuildProducts restricted to what the hot path reads.
struct BuildProducts {
    std::vector<F3> verts;
    std::vector<HostTri> tris;  // BVH order per mesh after the BLAS build
    std::vector<Range> meshRanges;
    std::vector<HostInstance> instances;
    std::vector<HostPoint> points;  // BVH order
    std::vector<uint32_t> perm;     // perm[slot] = asset index
    std::vector<RefNode> blas;
    std::vector<Range> blasRanges;
    std::vector<TlasNode> tlas;
    std::vector<HostLight> lights;
    std::vector<HostEmTri> emTris;
    uint32_t blasMaxDepth = 0;  // deepest BLAS node (root = 0)
    uint32_t tlasMaxDepth = 0;
    // The surfel BLAS was built on the device, which already holds the
    // surfel records and the cloud's wide nodes (after the meshes' ones).
    bool cloudOnDevice = false;
    std::vector<HostPoint> collected;  // scratch of a host build: the points in asset order
};

// A device build of the surfel BLAS.  staging(n) is where the scene build
// collects the n points (pinned memory), or null when the device build does
// not take a cloud of n points.  build() appends the reference nodes to
// `blas` and fills the permutation (perm[slot] = index in pts), both equal to
// the host build's, and the tree depth, and leaves the surfel records and
// the wide nodes (the cloud's first at wide index wideBase) on the device;
// false (nothing appended) = build on the host.
struct CloudBuild {
    virtual ~CloudBuild() = default;
    virtual HostPoint* staging(uint32_t n) = 0;
    virtual bool build(const HostPoint* pts, uint32_t n, uint32_t wideBase, std::vector<RefNode>& blas,
                       std::vector<uint32_t>& perm, uint32_t& depth) = 0;
};

// The shape of the surfel BLAS over n surfels, which depends on n only (a
// node of c > 4 surfels splits into c/2 and the rest), in the tables a
// device build works from (the kernels of pale_bvh.hip).
struct CloudShape {
    std::vector<RefNode> topo;         // ranges and child pairs (BVH.ixx numbering), no boxes
    std::vector<uint4> big;            // {first, count, ref, slot}: nodes of more than `subtree`
                                       // surfels, level by level (slot = index here)
    std::vector<uint32_t> levelBig;    // big[levelBig[d] .. levelBig[d + 1]) = level d
    std::vector<uint4> chunks;         // {slot, first, end, -}: box-fit pieces of the big nodes
    std::vector<uint32_t> levelChunk;  // chunks of level d, as levelBig
    std::vector<uint4> roots;          // {first, count, ref, preorder rank}: nodes of at most
                                       // `subtree` surfels under a big node (or the root)
    std::vector<uint32_t> internalOf;  // internal nodes of a subtree over c <= subtree surfels
    std::vector<uint4> wide;           // {children pair, left ref, right ref, -} per internal node
                                       // in reference order; a ref without the leaf bit is a
                                       // wide index relative to the first internal node
    uint32_t depth = 0;                // deepest node (root = 0)
};
CloudShape cloudShape(uint32_t n, uint32_t subtree, uint32_t fitChunk);

// Builds everything from the meshes and the asset points into bp (the
// surfel BLAS with `cloud` when given and it accepts the scene).
void buildScene(const std::vector<HostMesh>& meshes, const std::vector<AssetPoint>& asset, BuildProducts& bp,
                CloudBuild* cloud = nullptr);

// Device-layout images of the build products (host memory).
struct DeviceImage {
    std::vector<SurfelGeo> surfGeo;
    std::vector<SurfelShade> surfShade;
    std::vector<WideNode> nodes;
    std::vector<TriGeo> tris;
    std::vector<DevInstance> instances;
    std::vector<TlasNode> tlas;
    std::vector<DevMaterial> materials;
    std::vector<DevLight> lights;
    std::vector<DevEmTri> emTris;
    uint32_t stackDepth = 0;  // LDS stack entries one lane needs
    // With BuildProducts::cloudOnDevice: surfGeo / surfShade are sized but
    // not filled, and only nodes[0, meshWide) (the meshes' wide nodes) are;
    // the device holds the rest.
    uint32_t meshWide = 0;
};

// Rebuilt in place: the surfel records keep their memory (after a device
// build they are left as they are, not current, until refreshSurfels).
void makeDeviceImage(const BuildProducts& bp, const std::vector<HostMesh>& meshes, DeviceImage& img);

// Refresh only the surfel records of an image (apply_point_optimization).
void refreshSurfels(const std::vector<HostPoint>& pts, DeviceImage& img);

void cameraFocal(HostCamera& c);

// Host worker threads kept for the life of the process (up to 15; the
// calling thread is the 16th), so that a parallel loop does not pay for
// creating and joining its threads.  One loop runs on them at a time.
class HostPool {
   public:
    static HostPool& instance() {
        static HostPool pool;
        return pool;
    }
    // f(k) for k < n on the workers and the calling thread; false (nothing
    // run) when another thread's loop holds the workers.  An exception from
    // f stops the loop (items not yet started are skipped) and is rethrown
    // here once every worker has finished.
    bool tryRun(size_t n, const std::function<void(size_t)>& f) {
        if (getpid() != pid) return false;  // a forked child has none of the workers
        std::unique_lock<std::mutex> owner(runMu, std::try_to_lock);
        if (!owner.owns_lock()) return false;
        {
            std::lock_guard<std::mutex> g(mu);
            job = &f;
            count = n;
            next.store(0);
            pending = workers.size();
            error = nullptr;
            ++gen;
        }
        cv.notify_all();
        drain();
        std::unique_lock<std::mutex> g(mu);
        doneCv.wait(g, [&] { return pending == 0; });
        job = nullptr;
        const std::exception_ptr e = error;
        error = nullptr;
        g.unlock();
        if (e) std::rethrow_exception(e);
        return true;
    }
    size_t threads() const { return workers.size() + 1; }
    static bool& inLoop() {
        thread_local bool flag = false;
        return flag;
    }

   private:
    HostPool() : pid(getpid()) {
        const size_t hw = std::max<unsigned>(1u, std::thread::hardware_concurrency());
        const size_t nt = std::min<size_t>(hw, 16);
        for (size_t t = 1; t < nt; ++t) workers.emplace_back([this] { loop(); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv.notify_all();
        for (auto& t : workers) t.join();
    }
    void drain() {
        inLoop() = true;
        for (size_t k; (k = next.fetch_add(1)) < count;) {
            try {
                (*job)(k);
            } catch (...) {
                std::lock_guard<std::mutex> g(mu);
                if (!error) error = std::current_exception();
                next.store(count);
            }
        }
        inLoop() = false;
    }
    void loop() {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> g(mu);
        while (true) {
            cv.wait(g, [&] { return stop || gen != seen; });
            if (stop) return;
            seen = gen;
            g.unlock();
            drain();
            g.lock();
            if (--pending == 0) doneCv.notify_one();
        }
    }
    std::mutex runMu, mu;
    std::condition_variable cv, doneCv;
    std::vector<std::thread> workers;
    const std::function<void(size_t)>* job = nullptr;
    std::atomic<size_t> next{0};
    size_t count = 0, pending = 0;
    std::exception_ptr error;
    uint64_t gen = 0;
    bool stop = false;
    pid_t pid;
};

// f(k) for k < n on up to 16 threads (in any order): the pool's, or, when a
// loop of another thread holds the pool, threads of its own; a loop started
// inside a loop's task runs on the calling thread.
template <typename F>
inline void parallelFor(size_t n, F&& f) {
    if (n <= 1 || HostPool::inLoop()) {
        for (size_t k = 0; k < n; ++k) f(k);
        return;
    }
    const std::function<void(size_t)> fn = [&](size_t k) { f(k); };
    if (HostPool::instance().tryRun(n, fn)) return;
    const size_t hw = std::max<unsigned>(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(std::min<size_t>(hw, 16), n);
    std::atomic<size_t> next{0};
    std::exception_ptr error;
    std::mutex errorMu;
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (size_t k; (k = next.fetch_add(1)) < n;) {
                try {
                    f(k);
                } catch (...) {
                    std::lock_guard<std::mutex> g(errorMu);
                    if (!error) error = std::current_exception();
                    next.store(n);
                }
            }
        });
    for (auto& x : th) x.join();
    if (error) std::rethrow_exception(error);
}


}  // namespace pale
