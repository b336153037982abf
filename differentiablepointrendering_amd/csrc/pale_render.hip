// pale_render.hip -- renderer handle, device memory, pass orchestration and
// the extern "C" boundary declared in include/pale_abi.h.
//
// Orchestration restates the reference's PathTracer / SyclBridge
// (PathTracer.cpp:22-356, SyclBridge.cpp:82-294) as one HIP stream per
// handle: every launch of a forward or backward pass is queued back to back;
// queue counts, seeds and deposit counts stay in device memory (DevState), so
// the host synchronises once per call, when results are copied out.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <mutex>
#include <vector>

#include "pale_abi.h"
#include "pale_kernels.h"
#include "pale_scene.h"

using namespace pale;

namespace {

thread_local std::string g_err;

struct StatusError : std::runtime_error {
    int code;
    StatusError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(x)                                                                                   \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) throw StatusError(PALE_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device bytes held by every renderer of the process (pale_device_bytes).
std::atomic<uint64_t> g_deviceBytes{0};

struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    bool pinned = false;  // host memory from hipHostMalloc
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    DevMem(DevMem&& o) noexcept : p(o.p), bytes(o.bytes), pinned(o.pinned) {
        o.p = nullptr;
        o.bytes = 0;
    }
    ~DevMem() { release(); }
    void release() {
        if (p) {
            if (pinned) {
                (void)hipHostFree(p);
            } else {
                (void)hipFree(p);
                g_deviceBytes -= bytes;
            }
        }
        p = nullptr;
        bytes = 0;
    }
    void reserve(size_t b) {
        if (b <= bytes && p) return;
        release();
        HIPCHECK(hipMalloc(&p, std::max<size_t>(b, 16)));
        bytes = std::max<size_t>(b, 16);
        g_deviceBytes += bytes;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

template <typename T>
void upload(hipStream_t s, DevMem& m, const std::vector<T>& v) {
    m.reserve(v.size() * sizeof(T));
    if (!v.empty()) HIPCHECK(hipMemcpyAsync(m.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

uint32_t nextPow2(uint32_t v) {
    uint32_t c = 1;
    while (c < v) c <<= 1u;
    return c;
}

F3 ld3h(const float* p) { return mk3(p[0], p[1], p[2]); }

}  // namespace

// ---------------------------------------------- device BVH build --
// The surfel BLAS built on the device (kernels in pale_bvh.hip), equal node
// for node and slot for slot to the host build.  The tree's shape depends
// only on the surfel count, so it is computed on the host once per count:
// every node's range and reference index, the nodes of each level that
// split with a whole workgroup, the roots of the subtrees one thread builds,
// and the wide-node table.  Per build the device computes the boxes, the
// order of the surfels and, from them, the surfel records and wide nodes.
class DeviceCloudBuild final : public CloudBuild {
   public:
    static constexpr uint32_t kSubtree = 32;      // subtrees this small: one thread each
    static constexpr uint32_t kFitChunk = 16384;  // surfels per fit workgroup
    static constexpr uint32_t kWideMin = 65536;   // larger nodes take their first steps on many workgroups
    static constexpr uint32_t kWideTile = 8192;   // positions per workgroup in those steps

    hipStream_t stream = nullptr;
    uint32_t minCount = 65536;  // smaller clouds are built on the host
    DevMem* geo = nullptr;      // the renderer's surfel records and wide nodes
    DevMem* shade = nullptr;
    DevMem* nodes = nullptr;

    HostPoint* staging(uint32_t n) override {
        if (n < std::max<uint32_t>(minCount, kMaxLeafPrims + 1)) return nullptr;
        pinned(stage, (size_t)n * sizeof(HostPoint));
        return stage.as<HostPoint>();
    }

    bool build(const HostPoint* pts, uint32_t n, uint32_t wideBase, std::vector<RefNode>& blas,
               std::vector<uint32_t>& perm, uint32_t& depthOut) override {
        if (pts != stage.as<HostPoint>() || (size_t)n * sizeof(HostPoint) > stage.bytes)
            throw StatusError(PALE_ERR_INVALID_ARG, "device BVH build: points not in its staging memory");
        if (n != shapeN) shape(n);
        constexpr size_t kChunk = 1u << 15;
        dPts.reserve((size_t)n * sizeof(HostPoint));
        HIPCHECK(hipMemcpyAsync(dPts.p, pts, (size_t)n * sizeof(HostPoint), hipMemcpyHostToDevice, stream));
        for (DevMem* m : {&kx, &ky, &kz, &kv}) m->reserve((size_t)n * sizeof(float));
        for (DevMem* m : {&ord, &lpos, &rpos}) m->reserve((size_t)n * sizeof(uint32_t));
        boxes.reserve((size_t)n * 6 * sizeof(float));
        refBox.reserve(sh.topo.size() * 6 * sizeof(float));
        acc.reserve(std::max<size_t>(1, sh.big.size()) * 6 * sizeof(uint32_t));
        flags.reserve(sizeof(uint32_t));
        geo->reserve((size_t)n * sizeof(SurfelGeo));
        shade->reserve((size_t)n * sizeof(SurfelShade));
        nodes->reserve((size_t)(wideBase + sh.wide.size()) * sizeof(WideNode));
        launchFill(stream, flags.as<uint32_t>(), 1, 0u);
        launchBvhPrims(stream, dPts.as<HostPoint>(), n, kx.as<float>(), ky.as<float>(), kz.as<float>(),
                       boxes.as<float>(), ord.as<uint32_t>(), flags.as<uint32_t>());
        launchBvhInitFit(stream, acc.as<uint32_t>(), (uint32_t)sh.big.size());
        for (size_t d = 0; d + 1 < sh.levelBig.size(); ++d) {
            launchBvhFit(stream, dChunks.as<uint4>() + sh.levelChunk[d], sh.levelChunk[d + 1] - sh.levelChunk[d],
                         ord.as<uint32_t>(), boxes.as<float>(), acc.as<uint32_t>());
            const uint32_t nNodes = sh.levelBig[d + 1] - sh.levelBig[d];
            const uint4* lvl = dBig.as<uint4>() + sh.levelBig[d];
            const BvhWideState* resume = nullptr;
            if (levelMax[d] > kWideMin) {
                uint32_t steps = 3u;  // log2(count / kWideMin) + 3: halving steps, and some to spare
                for (uint32_t r = levelMax[d] / kWideMin; r > 1u; r >>= 1) ++steps;
                launchBvhWideSteps(stream, lvl, nNodes, dTiles.as<uint4>() + levelTile[d], levelTile[d + 1] - levelTile[d],
                                   kWideTile, steps, kWideMin, acc.as<uint32_t>(), kx.as<float>(), ky.as<float>(),
                                   kz.as<float>(), kv.as<float>(), ord.as<uint32_t>(), lpos.as<uint32_t>(),
                                   rpos.as<uint32_t>(), refBox.as<float>(), wstate.as<BvhWideState>(),
                                   tileCnt.as<uint2>());
                resume = wstate.as<BvhWideState>();
            }
            launchBvhSelect(stream, lvl, nNodes, levelMax[d], acc.as<uint32_t>(), kx.as<float>(), ky.as<float>(),
                            kz.as<float>(), kv.as<float>(), ord.as<uint32_t>(), lpos.as<uint32_t>(), rpos.as<uint32_t>(),
                            refBox.as<float>(), resume);
        }
        launchBvhSubtrees(stream, dRoots.as<uint4>(), (uint32_t)sh.roots.size(), dInternal.as<uint32_t>(),
                          kx.as<float>(), ky.as<float>(), kz.as<float>(), boxes.as<float>(), kv.as<float>(),
                          ord.as<uint32_t>(), refBox.as<float>());
        launchBvhWide(stream, dWide.as<uint4>(), (uint32_t)sh.wide.size(), refBox.as<float>(), wideBase,
                      nodes->as<WideNode>());
        launchBvhRecords(stream, dPts.as<HostPoint>(), ord.as<uint32_t>(), n, geo->as<SurfelGeo>(),
                         shade->as<SurfelShade>());
        HIPCHECK(hipGetLastError());
        const size_t permB = (size_t)n * sizeof(uint32_t), boxB = sh.topo.size() * 6 * sizeof(float);
        pinned(back, permB + boxB + 16);
        char* hb = back.as<char>();
        HIPCHECK(hipMemcpyAsync(hb, ord.p, permB, hipMemcpyDeviceToHost, stream));
        HIPCHECK(hipMemcpyAsync(hb + permB, refBox.p, boxB, hipMemcpyDeviceToHost, stream));
        HIPCHECK(hipMemcpyAsync(hb + permB + boxB, flags.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        HIPCHECK(hipStreamSynchronize(stream));
        uint32_t fl;
        std::memcpy(&fl, hb + permB + boxB, sizeof(fl));
        if (fl) return false;  // non-finite or -0 box coordinates: the host builds this scene
        perm.resize(n);
        const uint32_t* hperm = reinterpret_cast<const uint32_t*>(hb);
        const float* hbox = reinterpret_cast<const float*>(hb + permB);
        const size_t base = blas.size(), nn = sh.topo.size();
        blas.resize(base + nn);
        RefNode* out = blas.data() + base;
        parallelFor((std::max<size_t>(n, nn) + kChunk - 1) / kChunk, [&](size_t c) {
            const size_t a = c * kChunk;
            if (a < n) std::memcpy(perm.data() + a, hperm + a, (std::min<size_t>(n, a + kChunk) - a) * sizeof(uint32_t));
            for (size_t i = a; i < std::min(nn, a + kChunk); ++i) {
                const float* b = hbox + 6 * i;
                out[i] = RefNode{mk3(b[0], b[1], b[2]), mk3(b[3], b[4], b[5]), sh.topo[i].leftFirst, sh.topo[i].count};
            }
        });
        depthOut = sh.depth;
        return true;
    }

   private:
    // shape of the tree for shapeN surfels
    uint32_t shapeN = kInvalid;
    CloudShape sh;
    std::vector<uint32_t> levelMax;   // largest node of each level
    std::vector<uint32_t> levelTile;  // tiles of the multi-workgroup steps, level by level
    DevMem dBig, dChunks, dRoots, dInternal, dWide, dTiles;
    // per build
    DevMem dPts, kx, ky, kz, kv, ord, lpos, rpos, boxes, refBox, acc, flags, wstate, tileCnt;
    DevMem stage, back;  // pinned host staging

    static void pinned(DevMem& m, size_t b) {
        if (m.p && m.bytes >= b) return;
        m.release();
        HIPCHECK(hipHostMalloc(&m.p, std::max<size_t>(b, 16), hipHostMallocDefault));
        m.bytes = std::max<size_t>(b, 16);
        m.pinned = true;
    }

    void shape(uint32_t n) {
        sh = cloudShape(n, kSubtree, kFitChunk);
        levelMax.assign(sh.levelBig.size(), 0);
        for (size_t d = 0; d + 1 < sh.levelBig.size(); ++d)
            for (uint32_t b = sh.levelBig[d]; b < sh.levelBig[d + 1]; ++b) levelMax[d] = std::max(levelMax[d], sh.big[b].y);
        // Tiles of the levels whose nodes take multi-workgroup steps:
        // {node index in the level, first, end, tile index in the node}, a
        // node's tiles one after the other.
        std::vector<uint4> tiles;
        levelTile.assign(sh.levelBig.size(), 0);
        size_t maxNodes = 1, maxTiles = 1;
        for (size_t d = 0; d + 1 < sh.levelBig.size(); ++d) {
            levelTile[d] = (uint32_t)tiles.size();
            if (levelMax[d] <= kWideMin) continue;
            maxNodes = std::max<size_t>(maxNodes, sh.levelBig[d + 1] - sh.levelBig[d]);
            const uint32_t t0 = (uint32_t)tiles.size();
            for (uint32_t b = sh.levelBig[d]; b < sh.levelBig[d + 1]; ++b) {
                const uint4 nd = sh.big[b];
                uint32_t k = 0;
                for (uint32_t a = nd.x; a < nd.x + nd.y; a += kWideTile, ++k)
                    tiles.push_back(make_uint4(b - sh.levelBig[d], a, std::min(a + kWideTile, nd.x + nd.y), k));
            }
            maxTiles = std::max<size_t>(maxTiles, tiles.size() - t0);
        }
        levelTile.back() = (uint32_t)tiles.size();
        wstate.reserve(maxNodes * sizeof(BvhWideState));
        tileCnt.reserve(maxTiles * sizeof(uint2));
        auto put = [&](DevMem& m, const auto& v) {
            m.reserve(std::max<size_t>(1, v.size()) * sizeof(v[0]));
            if (!v.empty()) HIPCHECK(hipMemcpyAsync(m.p, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice, stream));
        };
        put(dBig, sh.big);
        put(dChunks, sh.chunks);
        put(dRoots, sh.roots);
        put(dInternal, sh.internalOf);
        put(dWide, sh.wide);
        put(dTiles, tiles);
        HIPCHECK(hipStreamSynchronize(stream));
        shapeN = n;
    }
};

// ------------------------------------------- no-allocation adjoint --
// pale_render_view_backward_noalloc keeps every intermediate of the adjoint
// of one view (ray queues, pending states, event staging and pool, flags,
// compaction tiles) in this file-scope device array instead of buffers the
// library allocates: views of at most kNoAllocMaxPixels pixels.  Each GPU
// holds its own copy for the life of the process.  In host code `scratch` is
// only a symbol: its device address comes from hipGetSymbolAddress, resolved
// once per device (noAllocScratch) and kept in the handle.
constexpr uint32_t kNoAllocMaxPixels = 65536;  // 256 x 256
// Per pixel: 9 queue arrays (16 B), 2 flag bytes, pending 16 + 16 + 8 B,
// 4 staging + 4 pool event arrays (16 B), 4 B of transport order; plus the
// compaction tiles.
constexpr size_t kNoAllocScratchBytes =
    (size_t)kNoAllocMaxPixels * (9 * 16 + 1 + 1 + 16 + 16 + 8 + 4 * 16 + 4 * 16 + 4) + 512;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratchMu;
constexpr int kMaxDevices = 64;
void* g_scratchAddr[kMaxDevices] = {};

// Device address of `scratch` on the current device (resolved once per
// device).
char* noAllocScratch() {
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw StatusError(PALE_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (!g_scratchAddr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw StatusError(PALE_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratchAddr[dev] = p;
    }
    return static_cast<char*>(g_scratchAddr[dev]);
}
}  // namespace

struct pale_handle {
    pale_settings S{};
    std::vector<HostMesh> meshes;
    std::vector<HostCamera> cams;
    std::vector<AssetPoint> asset;
    BuildProducts bp;
    DeviceImage img;
    DeviceCloudBuild cloudBuild;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    DevMem geo, shade, nodes, tris, inst, tlas, mats, lights, emTris;
    DevScene scene{};

    // Photon map (persistent between forward and backward passes, as the
    // reference's RenderIntermediatesGPU::map is).
    bool gridConfigured = false;
    GridParams grid{};
    float radius = 0.0f;
    uint32_t photonCap = 0;
    DevMem rawPosPow, rawDirPow, rawPowZ, srtPosPow, srtDirPow, srtPowZ, cellId, sortedIdx, cellCount, cellStart,
        gridTiles, gridTotal;
    bool mapValid = false;
    // Reservoir replacement once the map is full (only when the capacity is
    // below photons * passes * bounces, the most deposits a render makes).
    bool reservoirOn = false;
    DevMem resvAttempt, resvSlot, resvOwner;

    // Ray queues and staging.
    uint32_t queueCap = 0;
    DevMem qA[3], qB[3], stg[3], extFlags, depStg[3], depFlags, tiles, counts;
    DevMem walkSpill;  // layer-walk stack entries beyond the LDS part (deep BVHs)
    uint2* walkSpillPtr() const { return walkSpill.p ? walkSpill.as<uint2>() : nullptr; }
    uint32_t walkLds() const { return tune.walk_lds_entries ? tune.walk_lds_entries : kWalkLdsEntries; }
    // The photon and adjoint traces keep kTraceLdsEntries stack entries per
    // lane in LDS and the rest in traceSpill: 16 KB of LDS per workgroup
    // instead of the whole stack's (23 KB at 1M surfels), so that LDS allows
    // the 4 waves per SIMD their registers allow.
    static constexpr uint32_t kTraceLdsEntries = 16;
    DevMem traceSpill;
    size_t traceSpillBytes(uint32_t rays) const {
        if (img.stackDepth <= kTraceLdsEntries || !rays) return 0;
        const size_t lanes = (size_t)((rays + kTraceBlock - 1) / kTraceBlock) * kTraceBlock;
        return (size_t)(img.stackDepth - kTraceLdsEntries) * lanes * sizeof(uint2);
    }
    // t with the split stack when the spill area (sized with the queues)
    // covers a launch of `rays` lanes; else the whole stack in LDS.
    TraceParams withTraceSpill(TraceParams t, uint32_t rays) const {
        const size_t need = traceSpillBytes(rays);
        if (need && traceSpill.p && traceSpill.bytes >= need) {
            t.spill = traceSpill.as<uint2>();
            t.inLds = kTraceLdsEntries;
        }
        return t;
    }
    // The spill area of the layer walk for the current scene and tuning
    // (before any launch: a captured graph holds its address).
    void ensureWalkSpill() {
        if (!scene.nTlas) return;
        if (const size_t sb = walkSpillBytes(scene, walkLds())) walkSpill.reserve(sb);
    }
    // Adjoint.
    uint32_t adjCap = 0, poolCap = 0;
    DevMem pend0, pend1, pend2, pend3, pend4, evStg[6], evCmp[6], evCell, evSorted, evFlags, dloss, gradOpacity,
        gradFull;
    // Host gradient outputs: the split arrays on the device, then pinned.
    DevMem gradStage, gradHost;
    bool adjFull = false;  // event rays allocated (full gradients or debug images)
    // Camera gather: pixel state, entries of one round, their cell sort.
    uint32_t pxCap = 0, entryCap = 0;
    DevMem viewFb, viewLdr;  // images of pale_render_view (pass by pass)
    float4* outFb = nullptr;  // where the last pale_render_view left its images
    float4* outLdr = nullptr;
    // Batched views (pale_render_views): descriptors of the current batch,
    // the framebuffers of all views, the captured graph of a batch's rounds.
    DevMem bTiles, bSegs, bRuns, bHdr, bCams, bSegFirst, bSegEnd, bGroups, bGroupSegs, viewsFb, viewsLdr;
    pale_batch_tuning tune{};
    hipGraphExec_t bGraph = nullptr;
    std::vector<uint8_t> bGraphKey;
    DevMem pxOt, pxDl, pxCs, pxCnt, pxOff, pxTiles, entQa, entQb, entComp, entContrib, entPos, entCell, entSorted,
        entStart, flagHost;
    // Device state and seed table.
    DevMem state, seedTable, cursors;
    // Parameters applied from device memory (pale_apply_points_device): their
    // raw copy, 16 floats per surfel; the host mirror (bp.points, asset,
    // img) is refreshed from it when host code next needs it.
    DevMem rawPts;
    bool hostStale = false;
    // Debug gradient images of the last render_backward (debug_images): RGBA
    // per pixel over its cameras in call order, and each camera's first pixel.
    DevMem debugImg;
    std::vector<size_t> debugOff;
    std::vector<uint32_t> debugCams;
    // Device address of this GPU's no-allocation scratch array.
    char* scratchBase = nullptr;
    void syncHostPoints() {
        if (!hostStale) return;
        const size_t n = bp.points.size();
        std::vector<float> r(n * 16);
        if (n) HIPCHECK(hipMemcpyAsync(r.data(), rawPts.p, n * 16 * sizeof(float), hipMemcpyDeviceToHost, stream));
        HIPCHECK(hipStreamSynchronize(stream));
        for (size_t i = 0; i < n; ++i) {
            const float* v = &r[16 * i];
            HostPoint& q = bp.points[i];
            q.p = mk3(v[0], v[1], v[2]);
            q.tu = mk3(v[3], v[4], v[5]);
            q.tv = mk3(v[6], v[7], v[8]);
            q.su = v[9];
            q.sv = v[10];
            q.albedo = mk3(v[11], v[12], v[13]);
            q.opacity = v[14];
            q.beta = v[15];
            AssetPoint& a = asset[i];
            a.p = q.p;
            a.tu = q.tu;
            a.tv = q.tv;
            a.su = q.su;
            a.sv = q.sv;
            a.albedo = q.albedo;
            a.opacity = q.opacity;
            a.beta = q.beta;
        }
        refreshSurfels(bp.points, img);
        hostStale = false;
    }
    // Per-camera framebuffers.
    std::vector<DevMem> fb, ldr;
    pale_stats stats{};
    bool countWork = false;
    uint32_t camRounds = 0;

    // Per-kernel-class device timing: event pairs recorded around launches on
    // the handle's stream, read after the call's final synchronisation.
    enum TimedClass { kTPhoton = 0, kTGather, kTAdjTrace, kTTransport, kTGrid, kTCamTrace, kTCamSort, kTClasses };
    std::vector<hipEvent_t> evPool;
    std::vector<int> evClass;
    size_t evUsed = 0;
    double classMs[kTClasses] = {};
    uint32_t classLaunches[kTClasses] = {};

    size_t tBegin(int cls) {
        if (evUsed + 2 > evPool.size()) {
            for (int k = 0; k < 64; ++k) {
                hipEvent_t e;
                HIPCHECK(hipEventCreate(&e));
                evPool.push_back(e);
            }
        }
        const size_t i = evUsed;
        evUsed += 2;
        if (evClass.size() < evUsed / 2) evClass.resize(evUsed / 2);
        evClass[i / 2] = cls;
        HIPCHECK(hipEventRecord(evPool[i], stream));
        return i;
    }
    void tEnd(size_t i) { HIPCHECK(hipEventRecord(evPool[i + 1], stream)); }
    void tReset() {
        evUsed = 0;
        for (int k = 0; k < kTClasses; ++k) {
            classMs[k] = 0.0;
            classLaunches[k] = 0;
        }
    }
    void tCollect() {
        for (size_t i = 0; i < evUsed; i += 2) {
            float ms = 0.0f;
            HIPCHECK(hipEventElapsedTime(&ms, evPool[i], evPool[i + 1]));
            classMs[evClass[i / 2]] += ms;
            classLaunches[evClass[i / 2]] += 1;
        }
        evUsed = 0;
    }

    ~pale_handle() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (bGraph) (void)hipGraphExecDestroy(bGraph);
        for (hipEvent_t e : evPool) (void)hipEventDestroy(e);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    DevState* st() const { return state.as<DevState>(); }
    uint32_t* stField(size_t off) const { return reinterpret_cast<uint32_t*>(state.as<char>() + off); }

    static RayQueue rq(DevMem* m) { return RayQueue{m[0].as<float4>(), m[1].as<float4>(), m[2].as<float4>()}; }
    static PhotonArrays pa(DevMem& a, DevMem& b, DevMem& c) {
        return PhotonArrays{a.as<float4>(), b.as<float4>(), c.as<float>()};
    }
    // Events carry e4 / e5 (the ray of the surfel they weigh, and the image
    // pixel in e5.w) with full gradients or debug images.
    bool eventRays() const { return S.full_gradients != 0 || S.debug_images != 0; }
    EventArrays ea(DevMem* m) const {
        const bool rays = eventRays();
        return EventArrays{m[0].as<float4>(), m[1].as<float4>(), m[2].as<float4>(), m[3].as<float4>(),
                           rays ? m[4].as<float4>() : nullptr, rays ? m[5].as<float4>() : nullptr};
    }

    TraceParams tp(float invSpp = 1.0f, const BatchSeg* segs = nullptr, const uint32_t* segFirst = nullptr,
                   uint32_t seed = 42u) const {
        return TraceParams{S.rr_start, S.q_null, S.q_reflect, S.q_transmit, S.full_gradients ? 1u : 0u,
                           invSpp, segs, segFirst, nullptr, nullptr, nullptr, 0u, seed,
                           S.debug_images ? 1u : 0u, 0u};
    }
    // Base seed of adjoint sample k: Random::number, the reference's
    // constant 42 (AdjointKernels.cpp:27,94; setting adjoint_seed), or
    // adjoint_seed + k with adjoint_decorrelate.
    uint32_t adjointSeed(uint32_t k) const { return S.adjoint_decorrelate ? S.adjoint_seed + k : S.adjoint_seed; }

    DevGrid devGrid() const {
        DevGrid g{};
        g.posPow = srtPosPow.as<float4>();
        g.dirPow = srtDirPow.as<float4>();
        g.powZ = srtPowZ.as<float>();
        g.cellStart = cellStart.as<uint32_t>();
        g.ox = grid.ox;
        g.oy = grid.oy;
        g.oz = grid.oz;
        g.cx = g.cy = g.cz = grid.cell;
        g.rx = grid.rx;
        g.ry = grid.ry;
        g.rz = grid.rz;
        g.radius = radius;
        g.attempted = reinterpret_cast<const uint32_t*>(state.as<char>() + offsetof(DevState, photonCount));
        g.capacity = photonCap;
        g.valid = mapValid ? 1u : 0u;
        return g;
    }

    DevCamera devCam(const HostCamera& c) const {
        DevCamera d{};
        std::copy(c.invView, c.invView + 16, d.invView);
        d.W = c.W;
        d.H = c.H;
        d.fx = c.fx;
        d.fy = c.fy;
        return d;
    }

    // Upload of the build products (SceneUpload::allocateAndUpload,
    // SceneUpload.ixx:321-476).
    void uploadScene() {
        if (!bp.cloudOnDevice) {
            upload(stream, geo, img.surfGeo);
            upload(stream, shade, img.surfShade);
        }
        upload(stream, nodes, img.nodes);  // after a device build: the meshes' part
        upload(stream, tris, img.tris);
        upload(stream, inst, img.instances);
        upload(stream, tlas, img.tlas);
        upload(stream, mats, img.materials);
        upload(stream, lights, img.lights);
        upload(stream, emTris, img.emTris);
        scene.geo = geo.as<SurfelGeo>();
        scene.shade = shade.as<SurfelShade>();
        scene.nodes = nodes.as<WideNode>();
        scene.tris = tris.as<TriGeo>();
        scene.inst = inst.as<DevInstance>();
        scene.tlas = tlas.as<TlasNode>();
        scene.mats = mats.as<DevMaterial>();
        scene.lights = lights.as<DevLight>();
        scene.emTris = emTris.as<DevEmTri>();
        scene.nInst = (uint32_t)img.instances.size();
        scene.cloudInst = kInvalid;
        for (uint32_t i = 0; i < img.instances.size(); ++i)
            if (img.instances[i].type == kGeoCloud) scene.cloudInst = i;
        scene.nTlas = (uint32_t)img.tlas.size();
        scene.nLights = (uint32_t)img.lights.size();
        scene.stackDepth = img.stackDepth;
        if ((size_t)img.stackDepth * sizeof(uint2) * kTraceBlock > 65536)
            throw StatusError(PALE_ERR_INVALID_ARG, "BVH too deep for the LDS traversal stack");
        ensureWalkSpill();
        gradOpacity.reserve(std::max<size_t>(1, bp.points.size()) * sizeof(float));
        HIPCHECK(hipStreamSynchronize(stream));
    }

    void rebuild() {
        const auto t0 = std::chrono::steady_clock::now();
        cloudBuild.stream = stream;
        cloudBuild.geo = &geo;
        cloudBuild.shade = &shade;
        cloudBuild.nodes = &nodes;
        cloudBuild.minCount = S.bvh_build == 1 ? 0u : 65536u;
        buildScene(meshes, asset, bp, S.bvh_build == 2 ? nullptr : &cloudBuild);
        makeDeviceImage(bp, meshes, img);
        uploadScene();
        setScene();
        stats.bvh_on_device = bp.cloudOnDevice ? 1u : 0u;
        stats.bvh_build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    // PathTracer::setScene (PathTracer.cpp:22-53): queue capacity, photon map
    // and grid (configured once, on the first call).
    void setScene() {
        ensureQueues(nextPow2(std::max(S.photons, 1u)));
        if (!gridConfigured) {
            // allocatePhotonMap (PathTracer.cpp:123-157): photons * passes *
            // bounces, clamped to 10 GiB of 64-byte reference records.
            const uint64_t requested = (uint64_t)S.photons * (uint64_t)S.forward_passes * (uint64_t)S.bounces;
            const uint64_t maxPhotons = (10ull * 1024ull * 1024ull * 1024ull) / 64ull;
            photonCap = S.photon_map_capacity ? S.photon_map_capacity : (uint32_t)std::min(requested, maxPhotons);
            if ((uint64_t)S.photons * (uint64_t)S.forward_passes > 0xFFFFFFFFull)
                throw StatusError(PALE_ERR_INVALID_ARG, "photons * forward_passes exceeds 2^32");
            reservoirOn = (uint64_t)photonCap < requested;
            if (reservoirOn) resvOwner.reserve((size_t)photonCap * sizeof(uint32_t));
            // configurePhotonGrid (PathTracer.cpp:184-221)
            radius = S.gather_radius;
            const float cell = 0.5f * radius;
            const TlasNode& root = bp.tlas[0];
            const F3 pad = mk3(radius, radius, radius);
            const F3 origin = mk3(root.mnx, root.mny, root.mnz) - pad;
            const F3 gmax = mk3(root.mxx, root.mxy, root.mxz) + pad;
            const F3 ext = gmax - origin;
            auto axisCount = [](float e, float c) { return (int)std::ceil(e / std::fmax(c, 1e-6f)); };
            grid.ox = origin.x;
            grid.oy = origin.y;
            grid.oz = origin.z;
            grid.cell = cell;
            grid.rx = axisCount(ext.x, cell);
            grid.ry = axisCount(ext.y, cell);
            grid.rz = axisCount(ext.z, cell);
            const uint64_t total = (uint64_t)grid.rx * (uint64_t)grid.ry * (uint64_t)grid.rz;
            if (total == 0 || total >= 0xFFFFFFFFull)
                throw StatusError(PALE_ERR_INVALID_ARG, "Photon grid resolution too high; increase r or tighten AABB.");
            grid.total = (uint32_t)total;
            const size_t pc = std::max<uint32_t>(photonCap, 1u);
            rawPosPow.reserve(pc * sizeof(float4));
            rawDirPow.reserve(pc * sizeof(float4));
            rawPowZ.reserve(pc * sizeof(float));
            srtPosPow.reserve(pc * sizeof(float4));
            srtDirPow.reserve(pc * sizeof(float4));
            srtPowZ.reserve(pc * sizeof(float));
            cellId.reserve(pc * sizeof(uint32_t));
            sortedIdx.reserve(pc * sizeof(uint32_t));
            cellCount.reserve((size_t)grid.total * sizeof(uint32_t));
            cellStart.reserve(((size_t)grid.total + 1) * sizeof(uint32_t));
            gridTiles.reserve(((size_t)grid.total / kTile + 2) * sizeof(uint32_t));
            gridTotal.reserve(sizeof(uint32_t) * 4);
            launchFill(stream, cellCount.as<uint32_t>(), grid.total, 0u);
            HIPCHECK(hipStreamSynchronize(stream));
            gridConfigured = true;
        }
    }

    // Buffer sizing runs either for real (Allocate) or as a dry run that
    // adds up the growth a request would cause (Probe: HBM admission).
    struct Allocate {
        static constexpr bool kDry = false;
        void operator()(DevMem& m, size_t b) const { m.reserve(b); }
    };
    struct Probe {
        static constexpr bool kDry = true;
        uint64_t add = 0;
        std::vector<std::pair<const DevMem*, size_t>> planned;  // sizes this dry run has grown
        void operator()(const DevMem& m, size_t b) {
            // DevMem::reserve: no change if allocated and large enough, else
            // a new allocation of max(b, 16) bytes in place of the old one.
            size_t have = m.p ? m.bytes : 0;
            bool allocated = m.p != nullptr;
            auto it = std::find_if(planned.begin(), planned.end(), [&](const auto& x) { return x.first == &m; });
            if (it != planned.end()) {
                have = it->second;
                allocated = true;
            }
            if (allocated && b <= have) return;
            const size_t now = std::max<size_t>(b, 16);
            add += now - have;
            if (it != planned.end())
                it->second = now;
            else
                planned.emplace_back(&m, now);
        }
    };

    template <typename A>
    void sizeQueues(uint32_t cap, A& a) {
        if (cap <= queueCap) return;
        for (int k = 0; k < 3; ++k) {
            a(qA[k], (size_t)cap * sizeof(float4));
            a(qB[k], (size_t)cap * sizeof(float4));
            a(stg[k], (size_t)cap * sizeof(float4));
        }
        a(extFlags, cap);
        a(depStg[0], (size_t)cap * sizeof(float4));
        a(depStg[1], (size_t)cap * sizeof(float4));
        a(depStg[2], (size_t)cap * sizeof(float));
        a(depFlags, cap);
        a(tiles, ((size_t)cap / kTile + 4) * sizeof(uint32_t));
        a(counts, 16 * sizeof(uint32_t));
        a(resvAttempt, (size_t)cap * sizeof(uint32_t));
        a(resvSlot, (size_t)cap * sizeof(uint32_t));
        if (const size_t sb = traceSpillBytes(cap)) a(traceSpill, sb);
        if (!A::kDry) queueCap = cap;
    }
    void ensureQueues(uint32_t cap) {
        Allocate a;
        sizeQueues(cap, a);
    }

    template <typename A>
    void sizeAdjoint(uint32_t npx, A& a) {
        sizeQueues(nextPow2(npx), a);
        a(entStart, ((size_t)grid.total + 1) * 4);
        uint32_t have = adjCap;
        if (eventRays() && !adjFull) have = 0;  // also allocate the event rays
        if (npx <= have) return;
        a(pend0, (size_t)npx * sizeof(float4));
        a(pend1, (size_t)npx * sizeof(float4));
        a(pend2, (size_t)npx * sizeof(uint2));
        // Event pool: the events of up to 24 bounces of npx paths wait for
        // one cell-sorted transport launch.
        const uint32_t pool = (uint32_t)std::min<uint64_t>((uint64_t)npx * 24u, 1u << 28);
        const int arrays = eventRays() ? 6 : 4;
        for (int k = 0; k < arrays; ++k) {
            a(evStg[k], (size_t)npx * sizeof(float4));
            a(evCmp[k], (size_t)pool * sizeof(float4));
        }
        if (S.full_gradients) {
            a(pend3, (size_t)npx * sizeof(float4));
            a(pend4, (size_t)npx * sizeof(float4));
        }
        a(evCell, (size_t)pool * 4);
        a(evSorted, (size_t)pool * 4);
        a(evFlags, npx);
        if (!A::kDry) {
            poolCap = pool;
            if (eventRays()) adjFull = true;
            adjCap = npx;
        }
    }
    void ensureAdjoint(uint32_t npx) {
        Allocate a;
        sizeAdjoint(npx, a);
    }

    // Device bytes a W x H view of spp passes (with its adjoint) would add:
    // the buffer sizing of renderView / renderViewBackward, run dry.
    uint64_t viewGrowth(uint32_t W, uint32_t H, uint32_t spp, bool adjoint) {
        Probe p;
        const uint32_t npx = W * H;
        if (!S.batch_passes) {
            sizeCamera(npx, p);
            p(cursors, kTraceRegions * sizeof(uint32_t));
            p(viewFb, (size_t)npx * sizeof(float4));
            p(viewLdr, (size_t)npx * sizeof(float4));
            if (adjoint) {
                sizeAdjoint(npx, p);
                p(dloss, (size_t)npx * 3 * sizeof(float));
                if (S.full_gradients) p(gradFull, std::max<size_t>(1, bp.points.size() * 16) * sizeof(float));
            }
            return p.add;
        }
        HostCamera c{};
        c.W = W;
        c.H = H;
        return viewsGrowth({c}, {std::max(spp, 1u)}, {adjoint});
    }

    // Device bytes a render_views of these views (then, for those marked,
    // each one's adjoint) would add: the batch cutting and buffer sizing of
    // the real calls, run dry.
    uint64_t viewsGrowth(const std::vector<HostCamera>& vc, const std::vector<uint32_t>& spp,
                         const std::vector<bool>& adjoint) {
        Probe p;
        const ForwardPlan fp = cutForward(vc, spp, [](uint32_t, uint32_t) { return 0u; });
        sizeForward(fp, vc.size(), p);
        for (size_t v = 0; v < vc.size(); ++v) {
            if (!adjoint[v]) continue;
            const BackwardPlan bpl = cutBackward({vc[v]}, {spp[v]});
            sizeBackward(bpl, 1, p);
        }
        return p.add;
    }

    // Frees the buffers that grow with the image size (camera gather, view
    // images, adjoint); the scene, BVH, photon map and photon-stage queues
    // stay.  They are allocated again by the next view.
    std::vector<DevMem*> viewBuffers() {
        std::vector<DevMem*> v = {&pxOt,    &pxDl,    &pxCs,    &pxCnt,  &pxOff, &pxTiles, &entQa,  &entQb,
                                  &entComp, &entContrib, &entPos, &entCell, &entSorted, &viewFb, &viewLdr, &pend0,
                                  &pend1,   &pend2,   &pend3,   &pend4,  &evCell, &evSorted, &evFlags, &dloss,
                                  &bTiles,  &bSegs,   &bRuns,   &bHdr,   &bCams, &bSegFirst, &viewsFb, &viewsLdr};
        for (int k = 0; k < 6; ++k) {
            v.push_back(&evStg[k]);
            v.push_back(&evCmp[k]);
        }
        return v;
    }
    uint64_t viewBufferBytes() {
        uint64_t b = 0;
        for (DevMem* m : viewBuffers()) b += m->p ? m->bytes : 0;
        return b;
    }
    void releaseViewBuffers() {
        for (DevMem* m : viewBuffers()) m->release();
        pxCap = entryCap = 0;
        adjCap = poolCap = 0;
        adjFull = false;
    }

    void resetState() {
        launchFill(stream, state.as<uint32_t>(), sizeof(DevState) / 4, 0u);
    }

    void readStats(bool forward) {
        DevState h{};
        HIPCHECK(hipMemcpy(&h, state.p, sizeof(DevState), hipMemcpyDeviceToHost));
        tCollect();
        if (forward) {
            stats.photon_queries = h.photonQ;
            stats.gather_queries = h.gatherQ;
            stats.photons_deposited = h.photonCount;
            stats.photon_trace_ms = classMs[kTPhoton];
            stats.gather_ms = classMs[kTGather];
            stats.grid_ms = classMs[kTGrid];
            stats.photon_trace_launches = classLaunches[kTPhoton];
            stats.gather_launches = classLaunches[kTGather];
            stats.camera_trace_ms = classMs[kTCamTrace];
            stats.camera_sort_ms = classMs[kTCamSort];
            stats.camera_trace_launches = classLaunches[kTCamTrace];
            stats.camera_rounds = camRounds;
        } else {
            stats.adjoint_queries = h.adjointQ;
            stats.gradient_events = h.events;
            stats.adjoint_trace_ms = classMs[kTAdjTrace];
            stats.adjoint_transport_ms = classMs[kTTransport];
            stats.adjoint_trace_launches = classLaunches[kTAdjTrace];
            stats.adjoint_transport_launches = classLaunches[kTTransport];
        }
        for (int c = 0; c < 3; ++c)
            for (int u = 0; u < 7; ++u) stats.work[c][u] = h.work[c][u];
        stats.grid_cells = grid.total;
    }

    // Camera-gather buffers for images of up to npx pixels.  The entry
    // buffer holds gather_entry_budget (default 24) layers per pixel on
    // average: one round in practice; deeper images take more rounds.
    template <typename A>
    void sizeCamera(uint32_t npx, A& a, uint32_t perPixelOverride = 0) {
        if (npx > pxCap) {
            a(pxOt, (size_t)npx * 16);
            a(pxDl, (size_t)npx * 16);
            a(pxCs, (size_t)npx * 16);
            a(pxCnt, (size_t)npx * 4);
            a(pxOff, ((size_t)npx + 1) * 4);
            a(pxTiles, ((size_t)npx / kTile + 2) * 4);
            if (!A::kDry) pxCap = npx;
        }
        const uint32_t budget = perPixelOverride ? perPixelOverride : S.gather_entry_budget;
        const uint64_t perPixel = budget ? budget : 24u;
        const uint64_t floor = budget ? 64u : (1u << 16);
        const uint32_t want = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((uint64_t)npx * perPixel, floor), 1u << 28);
        if (want != entryCap) {
            a(entQa, (size_t)want * 16);
            a(entQb, (size_t)want * 16);
            a(entComp, (size_t)want * 16);
            a(entContrib, (size_t)want * 16);
            a(entPos, (size_t)want * 4);
            a(entCell, (size_t)want * 4);
            a(entSorted, (size_t)want * 4);
            if (!A::kDry) entryCap = want;
        }
        a(entStart, ((size_t)grid.total + 1) * 4);
    }

    void ensureCamera(uint32_t npx, uint32_t perPixelOverride = 0) {
        Allocate a;
        sizeCamera(npx, a, perPixelOverride);
        if (!flagHost.p) {
            HIPCHECK(hipHostMalloc(&flagHost.p, 16, hipHostMallocDefault));
            flagHost.bytes = 16;
            flagHost.pinned = true;
        }
    }
    PixelArrays pxa() const {
        return PixelArrays{pxOt.as<float4>(), pxDl.as<float4>(), pxCs.as<float4>(), pxCnt.as<uint32_t>(),
                           pxOff.as<uint32_t>()};
    }
    EntryArrays enta() const {
        return EntryArrays{entQa.as<float4>(),    entQb.as<float4>(),    entComp.as<float4>(),
                           entContrib.as<float4>(), entPos.as<uint32_t>(), entCell.as<uint32_t>(),
                           entSorted.as<uint32_t>()};
    }

    // One gather pass of one camera (launchCameraGatherKernel,
    // PrimalKernels.cpp:500-653), in rounds until every pixel is finished.
    void cameraPass(const DevCamera& dc, uint32_t pass, float invSpp, float4* fbp, int decorrelate) {
        const uint32_t npx = dc.W * dc.H;
        DevState* st = this->st();
        const PixelArrays px = pxa();
        const EntryArrays ent = enta();
        const DevGrid G = devGrid();
        uint32_t* entries = stField(offsetof(DevState, entries));
        launchCamInit(stream, dc, st, pass, decorrelate, px);
        for (uint32_t round = 0;; ++round) {
            launchFill(stream, entries, 2, 0u);  // entries, overflow
            launchFill(stream, cursors.as<uint32_t>(), kTraceRegions, 0u);
            const size_t tt = tBegin(kTCamTrace);
            launchCamTrace(stream, scene, st, npx, dc.W, px, ent, entryCap, cursors.as<uint32_t>(), walkSpillPtr(),
                           walkLds(), countWork);
            tEnd(tt);
            const size_t ts = tBegin(kTCamSort);
            launchExclusiveScan(stream, px.cnt, px.off, npx, pxTiles.as<uint32_t>(), gridTotal.as<uint32_t>());
            launchEntrySort(stream, grid, st, entryCap, px, ent, cellCount.as<uint32_t>(), entStart.as<uint32_t>(),
                            gridTiles.as<uint32_t>(), gridTotal.as<uint32_t>());
            tEnd(ts);
            const size_t tk = tBegin(kTGather);
            launchEntryGather(stream, G, st, entryCap, entStart.as<uint32_t>() + grid.total, ent, countWork);
            tEnd(tk);
            const size_t tc = tBegin(kTCamSort);
            launchCamComposite(stream, npx, px, ent, invSpp, fbp);
            tEnd(tc);
            ++camRounds;
            HIPCHECK(hipMemcpyAsync(flagHost.p, stField(offsetof(DevState, overflow)), 4, hipMemcpyDeviceToHost,
                                    stream));
            HIPCHECK(hipStreamSynchronize(stream));
            if (*flagHost.as<uint32_t>() == 0u) break;
            if (round > 4096) throw StatusError(PALE_ERR_INTERNAL, "camera gather: no progress");
        }
    }

    // submitPhotonMappingKernel (SyclBridge.cpp:82-166) for the given cameras.
    void renderForward(const std::vector<uint32_t>& camIdx) {
        // Seed schedule: mt19937_64(seed), truncated to 32 bits
        // (SyclBridge.cpp:83-143, Random::number is uint32_t).
        const uint32_t tlen = uploadSeedTable((uint32_t)camIdx.size());
        resetState();
        tReset();
        HIPCHECK(hipEventRecord(ev0, stream));
        tracePhotons(tlen);
        camRounds = 0;
        if (S.batch_passes && S.gather_passes > 0 && !camIdx.empty()) {
            // Every camera draws one seed from the table where the photon
            // stage left the cursor (k_seed_next); its passes use it
            // (+ pass when decorrelated).  All passes of all cameras run as
            // batches; each image equals the pass-by-pass one.
            DevState hs{};
            HIPCHECK(hipMemcpyAsync(&hs, state.p, sizeof(DevState), hipMemcpyDeviceToHost, stream));
            HIPCHECK(hipStreamSynchronize(stream));
            std::vector<HostCamera> vc;
            std::vector<uint32_t> base;
            for (size_t k = 0; k < camIdx.size(); ++k) {
                vc.push_back(cams[camIdx[k]]);
                const size_t c = (size_t)hs.cursor + k;
                base.push_back(c < tlen ? seedHost[c] : 0u);
            }
            const std::vector<uint32_t> spp(vc.size(), S.gather_passes);
            const bool dec = S.spp_decorrelate != 0;
            ForwardPlan plan =
                planForward(vc, spp, [&](uint32_t v, uint32_t p) { return dec ? base[v] + p : base[v]; });
            runForward(plan);
            for (size_t k = 0; k < camIdx.size(); ++k) {
                const size_t off = plan.fbOff[k], bytes = (plan.fbOff[k + 1] - off) * sizeof(float4);
                HIPCHECK(hipMemcpyAsync(fb[camIdx[k]].p, viewsFb.as<float4>() + off, bytes, hipMemcpyDeviceToDevice,
                                        stream));
                HIPCHECK(hipMemcpyAsync(ldr[camIdx[k]].p, viewsLdr.as<float4>() + off, bytes,
                                        hipMemcpyDeviceToDevice, stream));
            }
            finishForward();
            return;
        }
        DevState* st = this->st();
        const float invSpp = 1.0f / (float)S.gather_passes;
        cursors.reserve(kTraceRegions * sizeof(uint32_t));
        for (size_t k = 0; k < camIdx.size(); ++k) {
            const HostCamera& c = cams[camIdx[k]];
            const uint32_t npx = c.W * c.H;
            ensureCamera(npx);
            launchSeedNext(stream, st, seedTable.as<uint32_t>(), tlen, -1);
            launchFillF(stream, fb[camIdx[k]].as<float>(), npx * 4, 0.0f);
            const DevCamera dc = devCam(c);
            for (uint32_t pass = 0; pass < S.gather_passes; ++pass)
                cameraPass(dc, pass, invSpp, fb[camIdx[k]].as<float4>(), S.spp_decorrelate);
            launchPostprocess(stream, fb[camIdx[k]].as<float4>(), ldr[camIdx[k]].as<float4>(), npx);
        }
        finishForward();
    }

    uint32_t uploadSeedTable(uint32_t cameras) {
        std::mt19937_64 gen(S.seed);
        seedHost.assign((size_t)S.forward_passes * (1 + S.bounces) + cameras + 1, 0u);
        for (uint32_t& v : seedHost) v = (uint32_t)gen();
        upload(stream, seedTable, seedHost);
        HIPCHECK(hipStreamSynchronize(stream));
        return (uint32_t)seedHost.size();
    }
    std::vector<uint32_t> seedHost;  // the seed table on the host

    void finishForward() {
        HIPCHECK(hipEventRecord(ev1, stream));
        HIPCHECK(hipStreamSynchronize(stream));
        HIPCHECK(hipGetLastError());
        float ms = 0.0f;
        HIPCHECK(hipEventElapsedTime(&ms, ev0, ev1));
        stats.forward_ms = ms;
        readStats(true);
    }

    // The photon stage of submitPhotonMappingKernel (SyclBridge.cpp:82-141):
    // emission, bounces, grid.
    void tracePhotons(uint32_t tlen) {
        const uint32_t P = S.photons;
        if (P > queueCap) ensureQueues(nextPow2(P));
        DevState* st = this->st();
        const uint32_t emitted = S.photons * S.forward_passes;
        const float invEmitted = 1.0f / (float)emitted;
        uint32_t* emitCount = counts.as<uint32_t>();
        launchFill(stream, emitCount, 1, P);
        uint32_t* qCount = stField(offsetof(DevState, qCount));
        uint32_t* extTotal = stField(offsetof(DevState, extTotal));
        uint32_t* depTotal = stField(offsetof(DevState, depTotal));
        uint32_t* photonCount = stField(offsetof(DevState, photonCount));
        const Reservoir rs{reservoirOn ? resvAttempt.as<uint32_t>() : nullptr, resvSlot.as<uint32_t>(),
                           resvOwner.as<uint32_t>(), photonCap};
        if (reservoirOn) launchFill(stream, resvOwner.as<uint32_t>(), photonCap, 0u);
        DevMem* in = qA;
        DevMem* out = qB;
        for (uint32_t pass = 0; pass < S.forward_passes; ++pass) {
            launchSeedNext(stream, st, seedTable.as<uint32_t>(), tlen, -1);
            launchEmit(stream, scene, st, P, invEmitted, rq(stg), extFlags.as<uint8_t>());
            launchCompactRays(stream, extFlags.as<uint8_t>(), emitCount, P, tiles.as<uint32_t>(), extTotal, rq(stg),
                              rq(in), queueCap);
            launchSetQueue(stream, st, extTotal);
            for (uint32_t b = 0; b < S.bounces; ++b) {
                launchSeedNext(stream, st, seedTable.as<uint32_t>(), tlen, (int)b);
                const size_t te = tBegin(kTPhoton);
                launchPhotonTrace(stream, scene, st, P, rq(in), rq(stg), extFlags.as<uint8_t>(),
                                  pa(depStg[0], depStg[1], depStg[2]), depFlags.as<uint8_t>(),
                                  withTraceSpill(tp(), P), rs, tiles.as<uint32_t>(), countWork);
                tEnd(te);
                launchCompactRays(stream, extFlags.as<uint8_t>(), qCount, P, tiles.as<uint32_t>(), extTotal, rq(stg),
                                  rq(out), queueCap);
                launchCompactPhotons(stream, depFlags.as<uint8_t>(), qCount, P, tiles.as<uint32_t>(), depTotal,
                                     photonCount, pa(depStg[0], depStg[1], depStg[2]),
                                     pa(rawPosPow, rawDirPow, rawPowZ), photonCap);
                launchReservoirWrite(stream, st, P, rs, pa(depStg[0], depStg[1], depStg[2]),
                                     pa(rawPosPow, rawDirPow, rawPowZ));
                launchFinishPhotonBounce(stream, st, photonCap);
                std::swap(in, out);
            }
        }
        // buildPhotonCellRangesAndOrdering (PrimalKernels.cpp:1007-1018)
        const size_t tg = tBegin(kTGrid);
        launchBuildGrid(stream, grid, st, photonCap, pa(rawPosPow, rawDirPow, rawPowZ),
                        pa(srtPosPow, srtDirPow, srtPowZ), cellId.as<uint32_t>(), sortedIdx.as<uint32_t>(),
                        cellCount.as<uint32_t>(), cellStart.as<uint32_t>(), gridTiles.as<uint32_t>(),
                        gridTotal.as<uint32_t>());
        tEnd(tg);
        mapValid = true;
    }

    // Photon map only (serving: views reuse it).
    void buildPhotonMap() {
        const uint32_t tlen = uploadSeedTable(0);
        resetState();
        tReset();
        camRounds = 0;
        HIPCHECK(hipEventRecord(ev0, stream));
        tracePhotons(tlen);
        finishForward();
    }

    // ---------------------------------------------------------- batches --
    // Views x passes walked together: segments in pass-major order (every
    // view's pass 0, then pass 1, ...) are cut into batches of at most
    // slots_per_batch pixel slots; a batch is one persistent layer walk per
    // round over all its tiles, and a view's framebuffer is updated pass by
    // pass in order (k_batch_resolve), so each image equals renderView's.
    struct HostBatch {
        std::vector<uint4> tiles;
        std::vector<BatchSeg> segs;
        std::vector<BatchRun> runs;
        BatchHeader hdr{};
    };

    BatchDesc batchDesc() const {
        return BatchDesc{bTiles.as<uint4>(), bSegs.as<BatchSeg>(), bRuns.as<BatchRun>(), bHdr.as<BatchHeader>(),
                         bCams.as<DevCamera>()};
    }

    // One round of a batch.  timed: per-stage HIP events and work counting
    // (eager launches only; a captured round records neither).
    void batchRound(uint32_t maxSlots, float fraction, bool timed) {
        const PixelArrays px = pxa();
        const EntryArrays ent = enta();
        const bool count = timed && countWork;
        launchFill(stream, stField(offsetof(DevState, entries)), 2, 0u);  // entries, overflow
        launchFill(stream, cursors.as<uint32_t>(), kTraceRegions, 0u);
        const size_t tt = timed ? tBegin(kTCamTrace) : 0;
        launchCamTraceBatch(stream, scene, st(), batchDesc(), maxSlots, px, ent, entryCap, cursors.as<uint32_t>(),
                            walkSpillPtr(), walkLds(), fraction, count);
        if (timed) tEnd(tt);
        const size_t ts = timed ? tBegin(kTCamSort) : 0;
        launchExclusiveScan(stream, px.cnt, px.off, maxSlots, pxTiles.as<uint32_t>(), gridTotal.as<uint32_t>());
        launchEntrySort(stream, grid, st(), entryCap, px, ent, cellCount.as<uint32_t>(), entStart.as<uint32_t>(),
                        gridTiles.as<uint32_t>(), gridTotal.as<uint32_t>());
        if (timed) tEnd(ts);
        const size_t tk = timed ? tBegin(kTGather) : 0;
        launchEntryGather(stream, devGrid(), st(), entryCap, entStart.as<uint32_t>() + grid.total, ent, count);
        if (timed) tEnd(tk);
        const size_t tc = timed ? tBegin(kTCamSort) : 0;
        launchCamCompositeBatch(stream, maxSlots, bHdr.as<BatchHeader>(), px, ent);
        if (timed) tEnd(tc);
    }

    template <typename T>
    static void keyAdd(std::vector<uint8_t>& k, const T& v) {
        const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
        k.insert(k.end(), p, p + sizeof(T));
    }

    // The graph of a batch's first R rounds (init included); kernels read the
    // batch's sizes from the header, so one graph serves every batch of the
    // same capacities.  Recaptured when any argument it holds changes.
    void ensureBatchGraph(uint32_t maxTiles, uint32_t maxSlots, uint32_t rounds, float fraction) {
        std::vector<uint8_t> key;
        const PixelArrays px = pxa();
        const EntryArrays ent = enta();
        const BatchDesc bd = batchDesc();
        const DevGrid G = devGrid();
        keyAdd(key, px);
        keyAdd(key, ent);
        keyAdd(key, bd);
        keyAdd(key, scene);
        keyAdd(key, G);
        keyAdd(key, grid);
        keyAdd(key, state.p);
        keyAdd(key, cursors.p);
        keyAdd(key, walkSpill.p);
        keyAdd(key, walkLds());
        keyAdd(key, entStart.p);
        keyAdd(key, pxTiles.p);
        keyAdd(key, gridTiles.p);
        keyAdd(key, gridTotal.p);
        keyAdd(key, cellCount.p);
        keyAdd(key, maxTiles);
        keyAdd(key, maxSlots);
        keyAdd(key, entryCap);
        keyAdd(key, rounds);
        keyAdd(key, fraction);
        if (bGraph && key == bGraphKey) return;
        if (bGraph) {
            HIPCHECK(hipGraphExecDestroy(bGraph));
            bGraph = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        launchBatchInit(stream, bd, maxTiles, px);
        for (uint32_t r = 0; r < rounds; ++r) batchRound(maxSlots, fraction, false);
        HIPCHECK(hipStreamEndCapture(stream, &g));
        const hipError_t e = hipGraphInstantiate(&bGraph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHECK(e);
        bGraphKey = key;
    }

    // Segments (view, pass) in pass-major order cut into batches, their
    // descriptors, buffers and (graph_rounds > 0) the batch graph.  seedOf
    // gives the seed of a pass.
    struct ForwardPlan {
        std::vector<HostBatch> batches;
        std::vector<size_t> fbOff;
        uint32_t maxTiles = 0, maxSlots = 0, rounds = 0;
        size_t maxSegs = 1, maxRuns = 1;
        float fraction = 1.0f;
    };

    // Buffers of a forward plan; A = Allocate, or Probe for the dry run of
    // admission (the same calls, so the two cannot disagree).
    template <typename A>
    void sizeForward(const ForwardPlan& plan, size_t nViews, A& a) {
        // Entries per slot: the tuning's, else the settings' per-pixel budget.
        sizeCamera(std::max(plan.maxSlots, 1u), a, tune.entries_per_slot);
        a(cursors, kTraceRegions * sizeof(uint32_t));
        a(bTiles, (size_t)std::max(plan.maxTiles, 1u) * sizeof(uint4));
        a(bSegs, plan.maxSegs * sizeof(BatchSeg));
        a(bRuns, plan.maxRuns * sizeof(BatchRun));
        a(bHdr, sizeof(BatchHeader));
        a(bCams, std::max<size_t>(nViews, 1) * sizeof(DevCamera));
        a(viewsFb, std::max<size_t>(plan.fbOff.back(), 1) * sizeof(float4));
        a(viewsLdr, std::max<size_t>(plan.fbOff.back(), 1) * sizeof(float4));
    }

    template <typename SeedFn>
    ForwardPlan planForward(const std::vector<HostCamera>& vc, const std::vector<uint32_t>& spp, SeedFn seedOf) {
        ForwardPlan plan = cutForward(vc, spp, seedOf);
        Allocate a;
        sizeForward(plan, vc.size(), a);
        ensureCamera(std::max(plan.maxSlots, 1u), tune.entries_per_slot);  // and the pinned round flag
        const uint32_t n = (uint32_t)vc.size();
        std::vector<DevCamera> dcs(n);
        for (uint32_t v = 0; v < n; ++v) dcs[v] = devCam(vc[v]);
        upload(stream, bCams, dcs);
        HIPCHECK(hipStreamSynchronize(stream));
        if (plan.rounds) ensureBatchGraph(plan.maxTiles, plan.maxSlots, plan.rounds, plan.fraction);
        return plan;
    }

    // Default forward batch (slots_per_batch = 0): room for 16 passes of the
    // largest view, between 4 Mi and 32 Mi slots.  The passes of a pixel that
    // share a batch are walked side by side by one wave, so large views gain
    // from more of them per batch (a 1920x1080 view of 3M surfels at 64 spp:
    // 2 passes per batch 9.6 s, 16 passes 6.1 s); the entry buffer stays
    // bounded by its 2^28 cap and further passes resume in extra rounds.
    static uint64_t defaultForwardBudget(const std::vector<HostCamera>& vc) {
        uint64_t npx = 0;
        for (const HostCamera& c : vc) npx = std::max<uint64_t>(npx, (uint64_t)c.W * c.H);
        return std::min<uint64_t>(std::max<uint64_t>(npx * 16u, 4u << 20), 32u << 20);
    }

    // Segments (view, pass) in pass-major order cut into batches of at most
    // slots_per_batch pixel slots (a pass larger than that is a batch of its
    // own), with their descriptors.
    template <typename SeedFn>
    ForwardPlan cutForward(const std::vector<HostCamera>& vc, const std::vector<uint32_t>& spp, SeedFn seedOf) const {
        ForwardPlan plan;
        const uint32_t n = (uint32_t)vc.size();
        plan.fbOff.assign(n + 1, 0);
        uint32_t maxSpp = 0;
        for (uint32_t v = 0; v < n; ++v) {
            if (spp[v] == 0) throw StatusError(PALE_ERR_INVALID_ARG, "render_views: spp must be >= 1");
            plan.fbOff[v + 1] = plan.fbOff[v] + (size_t)vc[v].W * vc[v].H;
            maxSpp = std::max(maxSpp, spp[v]);
        }
        if (plan.fbOff[n] > 0xFFFFFFFFull) throw StatusError(PALE_ERR_INVALID_ARG, "render_views: too many pixels");
        const uint64_t budget = tune.slots_per_batch ? tune.slots_per_batch : defaultForwardBudget(vc);
        plan.fraction = tune.grid_fraction > 0.0f ? std::min(tune.grid_fraction, 1.0f) : 1.0f;
        plan.rounds = countWork ? 0u : tune.graph_rounds;
        std::vector<HostBatch>& batches = plan.batches;
        uint64_t cur = budget + 1;
        for (uint32_t p = 0; p < maxSpp; ++p) {
            for (uint32_t v = 0; v < n; ++v) {
                if (spp[v] <= p) continue;
                const HostCamera& c = vc[v];
                const uint32_t npx = c.W * c.H;
                if (cur > 0 && cur + npx > budget) {
                    batches.emplace_back();
                    cur = 0;
                }
                HostBatch& b = batches.back();
                const uint32_t tx = (c.W + 7u) / 8u, ty = (c.H + 7u) / 8u;
                const uint32_t si = (uint32_t)b.segs.size();
                BatchSeg sg{v, p, c.W, c.H, tx, (uint32_t)cur, seedOf(v, p), (uint32_t)plan.fbOff[v],
                            1.0f / (float)spp[v], {0, 0, 0}};
                b.segs.push_back(sg);
                for (uint32_t t = 0; t < tx * ty; ++t)
                    b.tiles.push_back(make_uint4(si, t, (uint32_t)b.tiles.size(), 1u));
                cur += npx;
            }
        }
        for (HostBatch& b : batches) {
            // One run per view present, its segments in pass order (in
            // pass-major order a view's segments of a batch are not adjacent).
            std::vector<std::vector<uint32_t>> byView(n);
            for (uint32_t s2 = 0; s2 < b.segs.size(); ++s2) byView[b.segs[s2].view].push_back(s2);
            std::vector<BatchSeg> ordered;
            uint32_t resolve = 0;
            for (uint32_t v = 0; v < n; ++v) {
                if (byView[v].empty()) continue;
                BatchRun r{};
                r.view = v;
                r.firstSeg = (uint32_t)ordered.size();
                r.numSegs = (uint32_t)byView[v].size();
                r.npx = vc[v].W * vc[v].H;
                r.fbOffset = (uint32_t)plan.fbOff[v];
                r.resolveBase = resolve;
                r.invSpp = 1.0f / (float)spp[v];
                for (uint32_t s2 : byView[v]) ordered.push_back(b.segs[s2]);
                resolve += r.npx;
                b.runs.push_back(r);
            }
            b.segs.swap(ordered);
            // Tiles view by view, then tile by tile (row-major), each tile as
            // one group of K units whose slots take its pixels K at a time,
            // pass fastest: a contiguous range of slots (a layer-walk region)
            // covers a band of one image across its passes, and the lanes of
            // a wave walk a few pixels in all their passes side by side.
            std::vector<uint4> tiles;
            tiles.reserve(b.tiles.size());
            for (const BatchRun& r : b.runs) {
                const BatchSeg& s0 = b.segs[r.firstSeg];
                const uint32_t nt = s0.tilesX * ((s0.H + 7u) / 8u);
                for (uint32_t t = 0; t < nt; ++t) {
                    const uint32_t first = (uint32_t)tiles.size();
                    for (uint32_t k = 0; k < r.numSegs; ++k)
                        tiles.push_back(make_uint4(r.firstSeg, t, first, r.numSegs));
                }
            }
            b.tiles.swap(tiles);
            uint32_t slots = 0;
            for (const BatchSeg& sg : b.segs) slots += sg.W * sg.H;
            b.hdr = BatchHeader{(uint32_t)b.tiles.size(), slots, (uint32_t)b.segs.size(), (uint32_t)b.runs.size(),
                                resolve, {0, 0, 0}};
            plan.maxTiles = std::max(plan.maxTiles, b.hdr.nTiles);
            plan.maxSlots = std::max(plan.maxSlots, slots);
            plan.maxSegs = std::max(plan.maxSegs, b.segs.size());
            plan.maxRuns = std::max(plan.maxRuns, b.runs.size());
        }
        return plan;
    }

    // The batches of a plan into viewsFb / viewsLdr (all views, concatenated).
    void runForward(ForwardPlan& plan) {
        const uint32_t total = (uint32_t)plan.fbOff.back();
        launchFill(stream, pxCnt.as<uint32_t>(), std::max(plan.maxSlots, 1u), 0u);
        launchFillF(stream, viewsFb.as<float>(), total * 4u, 0.0f);
        const BatchDesc bd = batchDesc();
        for (HostBatch& b : plan.batches) {
            upload(stream, bTiles, b.tiles);
            upload(stream, bSegs, b.segs);
            upload(stream, bRuns, b.runs);
            HIPCHECK(hipMemcpyAsync(bHdr.p, &b.hdr, sizeof(BatchHeader), hipMemcpyHostToDevice, stream));
            if (plan.rounds) {
                HIPCHECK(hipGraphLaunch(bGraph, stream));
                camRounds += plan.rounds;
            } else {
                launchBatchInit(stream, bd, b.hdr.nTiles, pxa());
                batchRound(plan.maxSlots, plan.fraction, true);
                ++camRounds;
            }
            for (uint32_t extra = 0;; ++extra) {
                HIPCHECK(hipMemcpyAsync(flagHost.p, stField(offsetof(DevState, overflow)), 4,
                                        hipMemcpyDeviceToHost, stream));
                HIPCHECK(hipStreamSynchronize(stream));
                if (*flagHost.as<uint32_t>() == 0u) break;
                if (extra > 4096) throw StatusError(PALE_ERR_INTERNAL, "batched camera gather: no progress");
                batchRound(plan.maxSlots, plan.fraction, true);
                ++camRounds;
            }
            launchBatchResolve(stream, bd, b.hdr.nResolve, pxa(), viewsFb.as<float4>());
        }
        launchPostprocess(stream, viewsFb.as<float4>(), viewsLdr.as<float4>(), total);
        viewOffsets = plan.fbOff;
        batchCount = (uint32_t)plan.batches.size();
    }

    void renderViews(const std::vector<HostCamera>& vc, const std::vector<uint32_t>& spp,
                     const std::vector<uint32_t>& seeds) {
        if (!mapValid) throw StatusError(PALE_ERR_INVALID_ARG, "render_views: no photon map; build it first");
        ForwardPlan plan = planForward(vc, spp, [&](uint32_t v, uint32_t p) { return seeds[v] + p; });
        tReset();
        camRounds = 0;
        HIPCHECK(hipEventRecord(ev0, stream));
        launchBeginView(stream, st(), 0u);
        runForward(plan);
        finishForward();
    }
    std::vector<size_t> viewOffsets;
    uint32_t batchCount = 0;

    // One view over the current photon map: spp passes seeded by seed + k.
    void renderView(const HostCamera& c, uint32_t spp, uint32_t seed) {
        if (!mapValid) throw StatusError(PALE_ERR_INVALID_ARG, "render_view: no photon map; build it first");
        if (spp == 0) throw StatusError(PALE_ERR_INVALID_ARG, "render_view: spp must be >= 1");
        if (S.batch_passes) {
            renderViews({c}, {spp}, {seed});
            outFb = viewsFb.as<float4>();
            outLdr = viewsLdr.as<float4>();
            return;
        }
        outFb = viewFb.as<float4>();
        outLdr = viewLdr.as<float4>();
        const uint32_t npx = c.W * c.H;
        ensureCamera(npx);
        viewFb.reserve((size_t)npx * sizeof(float4));
        viewLdr.reserve((size_t)npx * sizeof(float4));
        tReset();
        camRounds = 0;
        HIPCHECK(hipEventRecord(ev0, stream));
        DevState* st = this->st();
        launchBeginView(stream, st, seed);
        launchFillF(stream, viewFb.as<float>(), npx * 4, 0.0f);
        const DevCamera dc = devCam(c);
        const float invSpp = 1.0f / (float)spp;
        cursors.reserve(kTraceRegions * sizeof(uint32_t));
        for (uint32_t pass = 0; pass < spp; ++pass) cameraPass(dc, pass, invSpp, viewFb.as<float4>(), 1);
        launchPostprocess(stream, viewFb.as<float4>(), viewLdr.as<float4>(), npx);
        finishForward();
    }

    // submitAdjointKernel (SyclBridge.cpp:169-294).
    // Adjoint prologue: keep the forward's deposit count (the gather
    // normalisation T) and reset only the adjoint counters; dL/dI of every
    // view to the device (off = float offsets); zero gradients; start timing.
    void beginBackward(const std::vector<size_t>& off, const std::vector<const float*>& src, bool srcOnDevice) {
        const uint32_t N = (uint32_t)bp.points.size();
        DevState h{};
        HIPCHECK(hipMemcpyAsync(&h, state.p, sizeof(DevState), hipMemcpyDeviceToHost, stream));
        HIPCHECK(hipStreamSynchronize(stream));
        h.adjointQ = 0;
        h.events = 0;
        h.qCount = 0;
        h.evPool = 0;
        HIPCHECK(hipMemcpyAsync(state.p, &h, sizeof(DevState), hipMemcpyHostToDevice, stream));
        HIPCHECK(hipStreamSynchronize(stream));
        dloss.reserve(std::max<size_t>(off.back(), 1) * sizeof(float));
        debugOff.clear();
        if (S.debug_images) {
            // The reference clears its debug framebuffers per call
            // (SyclBridge.cpp:183-205).
            for (size_t o : off) debugOff.push_back(o / 3);
            debugImg.reserve(std::max<size_t>(debugOff.back(), 1) * 4 * sizeof(float));
            launchFillF(stream, debugImg.as<float>(), (uint32_t)(debugOff.back() * 4), 0.0f);
        }
        for (size_t k = 0; k + 1 < off.size(); ++k)
            HIPCHECK(hipMemcpyAsync(dloss.as<float>() + off[k], src[k], (off[k + 1] - off[k]) * sizeof(float),
                                    srcOnDevice ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream));
        launchFillF(stream, gradOpacity.as<float>(), N, 0.0f);
        if (S.full_gradients) {
            gradFull.reserve(std::max<size_t>(1, (size_t)N * 16) * sizeof(float));
            launchFillF(stream, gradFull.as<float>(), N * 16, 0.0f);
        }
        HIPCHECK(hipStreamSynchronize(stream));
        tReset();
        HIPCHECK(hipEventRecord(ev0, stream));
    }

    // Transport of the pooled events in grid-cell order (one launch).
    void flushEvents() {
        const size_t tt = tBegin(kTTransport);
        DevState* st = this->st();
        launchEventSort(stream, grid, st, poolCap, ea(evCmp), evCell.as<uint32_t>(), evSorted.as<uint32_t>(),
                        cellCount.as<uint32_t>(), entStart.as<uint32_t>(), gridTiles.as<uint32_t>(),
                        gridTotal.as<uint32_t>());
        launchAdjTransport(stream, scene, devGrid(), st, poolCap, ea(evCmp), evSorted.as<uint32_t>(),
                           gradOpacity.as<float>(), S.full_gradients ? gradFull.as<float>() : nullptr,
                           S.debug_images ? debugImg.as<float>() : nullptr, countWork);
        tEnd(tt);
        launchFill(stream, stField(offsetof(DevState, evPool)), 1, 0u);
    }

    void finishBackward() {
        HIPCHECK(hipEventRecord(ev1, stream));
        HIPCHECK(hipStreamSynchronize(stream));
        HIPCHECK(hipGetLastError());
        float ms = 0.0f;
        HIPCHECK(hipEventElapsedTime(&ms, ev0, ev1));
        stats.backward_ms = ms;
        readStats(false);
    }

    struct BackwardPlan {
        std::vector<HostBatch> batches;
        std::vector<size_t> pixOff, off;  // pixel and dL/dI float offsets of the views
        uint32_t maxTiles = 0, maxSlots = 1;
        size_t maxSegs = 1;
    };

    // Buffers of a backward plan (A = Allocate or Probe, as sizeForward).
    template <typename A>
    void sizeBackward(const BackwardPlan& plan, size_t nViews, A& a) {
        sizeAdjoint(plan.maxSlots, a);
        a(bTiles, (size_t)plan.maxTiles * sizeof(uint4));
        a(bSegs, plan.maxSegs * sizeof(BatchSeg));
        a(bRuns, sizeof(BatchRun));
        a(bHdr, sizeof(BatchHeader));
        a(bSegFirst, plan.maxSegs * sizeof(uint32_t));
        a(bSegEnd, plan.maxSegs * sizeof(uint32_t));
        a(bGroups, std::max<size_t>(nViews, 1) * sizeof(uint4));
        a(bGroupSegs, plan.maxSegs * sizeof(uint32_t));
        a(bCams, std::max<size_t>(nViews, 1) * sizeof(DevCamera));
        a(dloss, std::max<size_t>(plan.off.back(), 1) * sizeof(float));
        if (S.full_gradients) a(gradFull, std::max<size_t>(1, bp.points.size() * 16) * sizeof(float));
    }

    // (view, sample) segments in sample-major order cut into batches like
    // the forward's.
    BackwardPlan cutBackward(const std::vector<HostCamera>& vc, const std::vector<uint32_t>& spp) const {
        BackwardPlan plan;
        const uint32_t n = (uint32_t)vc.size();
        std::vector<size_t>& pixOff = plan.pixOff;
        pixOff.assign(n + 1, 0);
        plan.off.assign(n + 1, 0);
        uint32_t maxSpp = 0;
        for (uint32_t v = 0; v < n; ++v) {
            if (spp[v] == 0) throw StatusError(PALE_ERR_INVALID_ARG, "render_views_backward: spp must be >= 1");
            pixOff[v + 1] = pixOff[v] + (size_t)vc[v].W * vc[v].H;
            plan.off[v + 1] = pixOff[v + 1] * 3;
            maxSpp = std::max(maxSpp, spp[v]);
        }
        if (pixOff[n] > 0x7FFFFFFFull) throw StatusError(PALE_ERR_INVALID_ARG, "render_views_backward: too many pixels");
        const uint64_t budget = tune.slots_per_batch ? tune.slots_per_batch : (4u << 20);
        std::vector<HostBatch>& batches = plan.batches;
        uint64_t cur = budget + 1;
        for (uint32_t p = 0; p < maxSpp; ++p) {
            for (uint32_t v = 0; v < n; ++v) {
                if (spp[v] <= p) continue;
                const HostCamera& c = vc[v];
                const uint32_t npx = c.W * c.H;
                if (cur > 0 && cur + npx > budget) {
                    batches.emplace_back();
                    cur = 0;
                }
                HostBatch& b = batches.back();
                const uint32_t tx = (c.W + 7u) / 8u, ty = (c.H + 7u) / 8u;
                const uint32_t si = (uint32_t)b.segs.size();
                BatchSeg sg{v, p, c.W, c.H, tx, (uint32_t)cur, adjointSeed(p), (uint32_t)pixOff[v], 1.0f / (float)spp[v],
                            {0, 0, 0}};
                b.segs.push_back(sg);
                for (uint32_t t = 0; t < tx * ty; ++t)
                    b.tiles.push_back(make_uint4(si, t, (uint32_t)b.tiles.size(), 1u));
                cur += npx;
            }
        }
        for (HostBatch& b : batches) {
            uint32_t slots = 0;
            for (const BatchSeg& s : b.segs) slots += s.W * s.H;
            b.hdr = BatchHeader{(uint32_t)b.tiles.size(), slots, (uint32_t)b.segs.size(), 0u, 0u, {0, 0, 0}};
            plan.maxTiles = std::max(plan.maxTiles, b.hdr.nTiles);
            plan.maxSlots = std::max(plan.maxSlots, slots);
            plan.maxSegs = std::max(plan.maxSegs, b.segs.size());
        }
        return plan;
    }

    // Adjoint of several views (gradients summed): each batch of the plan is
    // one adjoint wavefront over all its rays.  A ray seeds its draws with
    // its index among its segment's rays and its events weigh its view's
    // 1/spp, so each view contributes what renderBackwardCams gives it.
    void renderViewsBackward(const std::vector<HostCamera>& vc, const std::vector<uint32_t>& spp,
                             const std::vector<const float*>& src, bool srcOnDevice) {
        if (!mapValid) throw StatusError(PALE_ERR_INVALID_ARG, "render_views_backward: no photon map");
        BackwardPlan plan = cutBackward(vc, spp);
        std::vector<HostBatch>& batches = plan.batches;
        const uint32_t n = (uint32_t)vc.size();
        Allocate a;
        sizeBackward(plan, n, a);
        std::vector<DevCamera> dcs(n);
        for (uint32_t v = 0; v < n; ++v) dcs[v] = devCam(vc[v]);
        upload(stream, bCams, dcs);
        beginBackward(plan.off, src, srcOnDevice);
        DevState* st = this->st();
        uint32_t* qCount = stField(offsetof(DevState, qCount));
        uint32_t* extTotal = stField(offsetof(DevState, extTotal));
        const BatchDesc bd = batchDesc();
        TraceParams tpb = tp(1.0f, bSegs.as<BatchSeg>(), bSegFirst.as<uint32_t>());
        tpb.segEnd = bSegEnd.as<uint32_t>();
        tpb.groups = bGroups.as<uint4>();
        tpb.groupSegs = bGroupSegs.as<uint32_t>();
        const PendingArrays pend{pend0.as<float4>(), pend1.as<float4>(), pend2.as<uint2>(), pend3.as<float4>(),
                                 pend4.as<float4>()};
        uint64_t pooledBound = 0;
        for (HostBatch& b : batches) {
            // Events carry their weight (1/spp) and surfel, not their
            // segment: the pool spans batches.
            upload(stream, bTiles, b.tiles);
            upload(stream, bSegs, b.segs);
            HIPCHECK(hipMemcpyAsync(bHdr.p, &b.hdr, sizeof(BatchHeader), hipMemcpyHostToDevice, stream));
            // The views' groups of segments (their samples in this batch),
            // for the interleaved lanes of the adjoint trace.
            std::vector<uint4> groups;
            std::vector<uint32_t> groupSegs;
            for (uint32_t v = 0; v < n; ++v) {
                const uint32_t first = (uint32_t)groupSegs.size();
                for (uint32_t s2 = 0; s2 < b.segs.size(); ++s2)
                    if (b.segs[s2].view == v) groupSegs.push_back(s2);
                if (groupSegs.size() > first)
                    groups.push_back(make_uint4(first, (uint32_t)groupSegs.size() - first, 0u, 0u));
            }
            upload(stream, bGroups, groups);
            upload(stream, bGroupSegs, groupSegs);
            HIPCHECK(hipStreamSynchronize(stream));  // the host vectors go out of scope
            tpb.nGroups = (uint32_t)groups.size();
            const uint32_t nr = b.hdr.nSlots;
            DevMem* in = qA;
            DevMem* out = qB;
            launchAdjRaygenBatch(stream, bd, b.hdr.nTiles, dloss.as<float>(), rq(in), pend, st);
            for (uint32_t k = 0; k < S.adjoint_bounces; ++k) {
                if (pooledBound + nr > poolCap) {
                    flushEvents();
                    pooledBound = 0;
                }
                launchSegFirst(stream, st, nr, rq(in), bSegFirst.as<uint32_t>(), bSegEnd.as<uint32_t>(), b.hdr.nSegs,
                               bGroups.as<uint4>(), tpb.nGroups, bGroupSegs.as<uint32_t>());
                const size_t ta = tBegin(kTAdjTrace);
                launchAdjTrace(stream, scene, st, nr, k, rq(in), rq(stg), extFlags.as<uint8_t>(), ea(evStg),
                               evFlags.as<uint8_t>(), pend, withTraceSpill(tpb, nr), countWork);
                tEnd(ta);
                launchCompactEvents(stream, evFlags.as<uint8_t>(), qCount, nr, tiles.as<uint32_t>(), st, ea(evStg),
                                    ea(evCmp), poolCap);
                pooledBound += nr;
                launchCompactRays(stream, extFlags.as<uint8_t>(), qCount, nr, tiles.as<uint32_t>(), extTotal,
                                  rq(stg), rq(out), queueCap);
                launchSetQueue(stream, st, extTotal);
                std::swap(in, out);
            }
        }
        if (pooledBound) flushEvents();
        batchCount = (uint32_t)batches.size();
        finishBackward();
    }

    void renderBackward(const std::vector<uint32_t>& camIdx, const std::vector<const float*>& src,
                        bool srcOnDevice = false) {
        debugCams = camIdx;
        if (S.batch_passes && S.adjoint_passes > 0) {
            std::vector<HostCamera> vc;
            for (uint32_t c : camIdx) vc.push_back(cams[c]);
            renderViewsBackward(vc, std::vector<uint32_t>(vc.size(), S.adjoint_passes), src, srcOnDevice);
            return;
        }
        std::vector<const HostCamera*> cp;
        for (uint32_t c : camIdx) cp.push_back(&cams[c]);
        renderBackwardCams(cp, src, srcOnDevice, S.adjoint_passes);
    }

    void renderBackwardCams(const std::vector<const HostCamera*>& camList, const std::vector<const float*>& src,
                            bool srcOnDevice, uint32_t samples) {
        uint32_t maxNpx = 1;
        for (const HostCamera* c : camList) maxNpx = std::max(maxNpx, c->W * c->H);
        ensureAdjoint(maxNpx);
        std::vector<size_t> off(camList.size() + 1, 0);
        for (size_t k = 0; k < camList.size(); ++k) off[k + 1] = off[k] + (size_t)camList[k]->W * camList[k]->H * 3;
        beginBackward(off, src, srcOnDevice);
        DevState* st = this->st();
        uint32_t* qCount = stField(offsetof(DevState, qCount));
        uint32_t* extTotal = stField(offsetof(DevState, extTotal));
        const float invSpp = 1.0f / (float)samples;
        // launchAdjointTransportKernel (AdjointKernels.cpp:389-509) over the
        // pooled events of several bounces, in grid-cell order.  Gradients
        // are sums of atomics, so the grouping does not change them beyond
        // the order of the float additions.
        uint64_t pooledBound = 0;
        auto flush = [&]() {
            if (!pooledBound) return;
            flushEvents();
            pooledBound = 0;
        };
        for (size_t k = 0; k < camList.size(); ++k) {
            const HostCamera& c = *camList[k];
            const uint32_t npx = c.W * c.H;
            const DevCamera dc = devCam(c);
            const PendingArrays pend{pend0.as<float4>(), pend1.as<float4>(), pend2.as<uint2>(), pend3.as<float4>(),
                                     pend4.as<float4>()};
            for (uint32_t spp = 0; spp < samples; ++spp) {
                DevMem* in = qA;
                DevMem* out = qB;
                launchAdjRaygen(stream, dc, dloss.as<float>() + off[k], rq(in), pend, st, adjointSeed(spp));
                for (uint32_t b = 0; b < S.adjoint_bounces; ++b) {
                    if (pooledBound + npx > poolCap) flush();
                    const size_t ta = tBegin(kTAdjTrace);
                    TraceParams tpc = tp(invSpp, nullptr, nullptr, adjointSeed(spp));
                    tpc.debugBase = (uint32_t)(off[k] / 3);
                    launchAdjTrace(stream, scene, st, npx, b, rq(in), rq(stg), extFlags.as<uint8_t>(), ea(evStg),
                                   evFlags.as<uint8_t>(), pend, withTraceSpill(tpc, npx), countWork);
                    tEnd(ta);
                    launchCompactEvents(stream, evFlags.as<uint8_t>(), qCount, npx, tiles.as<uint32_t>(), st,
                                        ea(evStg), ea(evCmp), poolCap);
                    pooledBound += npx;
                    launchCompactRays(stream, extFlags.as<uint8_t>(), qCount, npx, tiles.as<uint32_t>(), extTotal,
                                      rq(stg), rq(out), queueCap);
                    launchSetQueue(stream, st, extTotal);
                    std::swap(in, out);
                }
            }
        }
        flush();
        finishBackward();
    }
    // The adjoint of one view (opacity gradient) with every intermediate in
    // the static scratch array: the kernels of renderBackwardCams, with the
    // events of each bounce transported right away, in compaction order (the
    // pool holds at most one event per ray), so no buffer is allocated.
    // gradDev (N floats, device) receives the gradient.
    void renderViewBackwardNoAlloc(const HostCamera& c, uint32_t samples, const float* dlossDev, float* gradDev) {
        if (!mapValid) throw StatusError(PALE_ERR_INVALID_ARG, "render_view_backward: no photon map; build it first");
        const uint32_t npx = c.W * c.H, N = (uint32_t)bp.points.size();
        const size_t M = kNoAllocMaxPixels;
        char* p = scratchBase;
        auto take = [&](size_t bytes) {
            char* r = p;
            p += (bytes + 255) & ~(size_t)255;
            return r;
        };
        float4* q[9];
        for (float4*& x : q) x = reinterpret_cast<float4*>(take(M * sizeof(float4)));
        uint8_t* ext = reinterpret_cast<uint8_t*>(take(M));
        uint8_t* evf = reinterpret_cast<uint8_t*>(take(M));
        const PendingArrays pend{reinterpret_cast<float4*>(take(M * sizeof(float4))),
                                 reinterpret_cast<float4*>(take(M * sizeof(float4))),
                                 reinterpret_cast<uint2*>(take(M * sizeof(uint2))), nullptr, nullptr};
        float4* e[8];
        for (float4*& x : e) x = reinterpret_cast<float4*>(take(M * sizeof(float4)));
        const EventArrays evS{e[0], e[1], e[2], e[3], nullptr, nullptr};
        const EventArrays evP{e[4], e[5], e[6], e[7], nullptr, nullptr};
        uint32_t* order = reinterpret_cast<uint32_t*>(take(M * sizeof(uint32_t)));
        uint32_t* tl = reinterpret_cast<uint32_t*>(take((M / kTile + 4) * sizeof(uint32_t)));
        if ((size_t)(p - scratchBase) > kNoAllocScratchBytes)
            throw StatusError(PALE_ERR_INTERNAL, "no-allocation scratch layout exceeds the array");
        const RayQueue A{q[0], q[1], q[2]}, B{q[3], q[4], q[5]}, stgq{q[6], q[7], q[8]};
        DevState* st = this->st();
        uint32_t* qCount = stField(offsetof(DevState, qCount));
        uint32_t* extTotal = stField(offsetof(DevState, extTotal));
        uint32_t* evPool = stField(offsetof(DevState, evPool));
        // Fill kernels, not hipMemsetAsync: the calls stay capturable.
        launchFillF(stream, gradDev, N, 0.0f);
        launchFill(stream, qCount, 1, 0u);
        launchFill(stream, evPool, 1, 0u);
        launchIota(stream, order, npx);
        const DevCamera dc = devCam(c);
        const float invSpp = 1.0f / (float)samples;
        for (uint32_t spp = 0; spp < samples; ++spp) {
            RayQueue in = A, out = B;
            launchAdjRaygen(stream, dc, dlossDev, in, pend, st, adjointSeed(spp));
            TraceParams t = tp(invSpp, nullptr, nullptr, adjointSeed(spp));
            t.fullGradients = 0;
            t.debug = 0;
            for (uint32_t b = 0; b < S.adjoint_bounces; ++b) {
                launchAdjTrace(stream, scene, st, npx, b, in, stgq, ext, evS, evf, pend, t, false);
                launchCompactEvents(stream, evf, qCount, npx, tl, st, evS, evP, npx);
                launchAdjTransport(stream, scene, devGrid(), st, npx, evP, order, gradDev, nullptr, nullptr, false);
                launchFill(stream, evPool, 1, 0u);
                launchCompactRays(stream, ext, qCount, npx, tl, extTotal, stgq, out, (uint32_t)M);
                launchSetQueue(stream, st, extTotal);
                std::swap(in, out);
            }
        }
        HIPCHECK(hipStreamSynchronize(stream));
        HIPCHECK(hipGetLastError());
    }
};

namespace {

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

template <typename F>
int guarded(F&& f) {
    try {
        g_err.clear();
        f();
        return PALE_OK;
    } catch (const StatusError& e) {
        return fail(e.code, e.what());
    } catch (const std::exception& e) {
        return fail(PALE_ERR_INTERNAL, e.what());
    }
}

std::vector<AssetPoint> toAsset(const pale_points_desc& p) {
    if (p.count && (!p.position || !p.tangent_u || !p.tangent_v || !p.scale || !p.albedo || !p.opacity || !p.beta))
        throw StatusError(PALE_ERR_INVALID_ARG, "points: missing array");
    std::vector<AssetPoint> out(p.count);
    for (uint32_t i = 0; i < p.count; ++i) {
        AssetPoint a{};
        a.p = ld3h(p.position + 3 * i);
        a.tu = ld3h(p.tangent_u + 3 * i);
        a.tv = ld3h(p.tangent_v + 3 * i);
        a.su = p.scale[2 * i];
        a.sv = p.scale[2 * i + 1];
        a.albedo = ld3h(p.albedo + 3 * i);
        a.opacity = p.opacity[i];
        a.beta = p.beta[i];
        a.shape = p.shape ? p.shape[i] : 0.0f;
        out[i] = a;
    }
    return out;
}

std::vector<uint32_t> selectCameras(const pale_handle* h, const char* name) {
    std::vector<uint32_t> idx;
    const std::string nm = name ? name : "";
    for (uint32_t i = 0; i < h->cams.size(); ++i)
        if (nm.empty() || h->cams[i].name == nm) idx.push_back(i);
    return idx;
}

// glm's quat * vec3 (v + ((uv * w) + uuv) * 2 with uv = q x v, uuv = q x uv).
F3 rotate(const float q[4], F3 v) {
    const F3 qv = mk3(q[0], q[1], q[2]);
    const F3 uv = cross(qv, v);
    const F3 uuv = cross(qv, uv);
    return v + ((uv * q[3]) + uuv) * 2.0f;
}

}  // namespace

extern "C" {

const char* pale_last_error(void) { return g_err.c_str(); }

void pale_default_settings(pale_settings* out) {
    if (!out) return;
    pale_settings s{};
    s.photons = 1000000u;
    s.bounces = 6;
    s.forward_passes = 6;
    s.gather_passes = 6;
    s.adjoint_bounces = 6;
    s.adjoint_passes = 6;
    s.seed = 42;
    s.gather_radius = 0.01f;
    s.rr_start = 6;
    s.q_null = 0.5f;
    s.q_reflect = 0.5f;
    s.q_transmit = 0.0f;
    s.spp_decorrelate = 0;
    s.debug_images = 0;
    s.gather_entry_budget = 0;
    s.photon_map_capacity = 0;
    s.full_gradients = 0;
    s.batch_passes = 1;
    s.adjoint_decorrelate = 0;
    s.adjoint_seed = 42;
    s.bvh_build = 0;
    *out = s;
}

static HostCamera hostCamera(const pale_camera_desc& cd) {
    HostCamera c;
    c.name = std::string(cd.name, strnlen(cd.name, PALE_NAME_LEN));
    c.W = cd.width;
    c.H = cd.height;
    if (!c.W || !c.H) throw StatusError(PALE_ERR_INVALID_ARG, "camera " + c.name + ": empty image");
    c.fovy = cd.fovy_deg;
    std::copy(cd.view, cd.view + 16, c.view);
    std::copy(cd.inv_view, cd.inv_view + 16, c.invView);
    c.pos = ld3h(cd.position);
    c.fwd = ld3h(cd.forward);
    c.adjoint = cd.use_for_adjoint != 0;
    cameraFocal(c);
    return c;
}

int pale_create(const pale_scene_desc* scene, const pale_settings* settings, pale_handle** out) {
    if (!scene || !settings || !out) return fail(PALE_ERR_INVALID_ARG, "pale_create: null argument");
    *out = nullptr;
    pale_handle* h = new pale_handle();
    const int rc = guarded([&] {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
            throw StatusError(PALE_ERR_DEVICE, "no HIP device available");
        // Zero passes are valid, as in the reference: nothing is traced for
        // that stage (render_target.py asks for adjoint_passes = 0).
        h->S = *settings;
        if (!(h->S.gather_radius > 0.0f)) throw StatusError(PALE_ERR_INVALID_ARG, "gather_radius must be > 0");
        if (h->S.bvh_build > 2) throw StatusError(PALE_ERR_INVALID_ARG, "bvh_build must be 0, 1 or 2");
        for (uint32_t i = 0; i < scene->mesh_count; ++i) {
            const pale_mesh_desc& md = scene->meshes[i];
            HostMesh m;
            for (uint32_t v = 0; v < md.vertex_count; ++v) m.verts.push_back(ld3h(md.positions + 3 * v));
            m.idx.assign(md.indices, md.indices + 3 * (size_t)md.triangle_count);
            for (uint32_t x : m.idx)
                if (x >= md.vertex_count) throw StatusError(PALE_ERR_INVALID_ARG, "mesh index out of range");
            std::copy(md.object_to_world, md.object_to_world + 16, m.o2w);
            std::copy(md.world_to_object, md.world_to_object + 16, m.w2o);
            m.base = ld3h(md.base_color);
            m.power = md.power;
            m.name = std::string(md.name, strnlen(md.name, PALE_NAME_LEN));
            h->meshes.push_back(std::move(m));
        }
        for (uint32_t i = 0; i < scene->camera_count; ++i) h->cams.push_back(hostCamera(scene->cameras[i]));
        h->asset = toAsset(scene->points);
        HIPCHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        // The no-allocation scratch of this device, resolved at set-up; a
        // failure here is reported by the no-allocation entry point (which
        // resolves it again), not by every renderer.
        try {
            h->scratchBase = noAllocScratch();
        } catch (const StatusError&) {
            h->scratchBase = nullptr;
        }
        HIPCHECK(hipEventCreate(&h->ev0));
        HIPCHECK(hipEventCreate(&h->ev1));
        h->state.reserve(sizeof(DevState));
        h->resetState();
        for (const HostCamera& c : h->cams) {
            DevMem f, l;
            f.reserve((size_t)c.W * c.H * sizeof(float4));
            l.reserve((size_t)c.W * c.H * sizeof(float4));
            h->fb.push_back(std::move(f));
            h->ldr.push_back(std::move(l));
        }
        h->rebuild();
    });
    if (rc != PALE_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return PALE_OK;
}

int pale_destroy(pale_handle* h) {
    delete h;
    return PALE_OK;
}

int pale_camera_count(const pale_handle* h, uint32_t* count) {
    if (!h || !count) return fail(PALE_ERR_INVALID_ARG, "null argument");
    *count = (uint32_t)h->cams.size();
    return PALE_OK;
}

int pale_camera_info(const pale_handle* h, uint32_t index, char name_out[PALE_NAME_LEN], uint32_t* width,
                     uint32_t* height, int32_t* use_for_adjoint) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    if (index >= h->cams.size()) return fail(PALE_ERR_RANGE, "camera index out of range");
    const HostCamera& c = h->cams[index];
    if (name_out) {
        std::memset(name_out, 0, PALE_NAME_LEN);
        std::memcpy(name_out, c.name.data(), std::min<size_t>(c.name.size(), PALE_NAME_LEN - 1));
    }
    if (width) *width = c.W;
    if (height) *height = c.H;
    if (use_for_adjoint) *use_for_adjoint = c.adjoint ? 1 : 0;
    return PALE_OK;
}

// Adjoint cameras in camera order that appear in the input
// (pale_bindings.cpp:301-376): index, image pointer, input position.
static void selectAdjoint(const pale_handle* h, uint32_t n, const char* const* names, const float* const* images,
                          std::vector<uint32_t>& idx, std::vector<const float*>& src, std::vector<uint32_t>& which) {
    for (uint32_t c = 0; c < h->cams.size(); ++c) {
        if (!h->cams[c].adjoint) continue;
        for (uint32_t k = 0; k < n; ++k) {
            if (names[k] && h->cams[c].name == names[k]) {
                if (!images[k]) throw StatusError(PALE_ERR_INVALID_ARG, "null image for " + h->cams[c].name);
                idx.push_back(c);
                src.push_back(images[k]);
                which.push_back(k);
                break;
            }
        }
    }
}

// Gradient arrays in ABI order: position, tangent_u, tangent_v, scale,
// albedo, opacity, beta, shape (pale_bindings.cpp:410-450).
static const size_t kGradWidth[8] = {3, 3, 3, 2, 3, 1, 1, 1};

static void copyGradsHost(pale_handle* h, float* const* grads) {
    const size_t N = h->bp.points.size();
    if (!grads || !N) return;
    // The arrays are split on the device (as for device outputs), come back
    // in one copy into pinned staging memory, and are copied into the
    // caller's arrays by host threads.  Staging layout: position, tangent_u,
    // tangent_v, scale, albedo, opacity, beta, one after another.
    size_t off[8], total = 0;
    for (int g = 0; g < 8; ++g) {
        off[g] = total;
        if (g != 7) total += N * kGradWidth[g];
    }
    h->gradStage.reserve(total * sizeof(float));
    float* d = h->gradStage.as<float>();
    const bool full = h->S.full_gradients != 0;
    if (full)
        launchSplitGradients(h->stream, h->gradFull.as<float>(), (uint32_t)N, d + off[0], d + off[1], d + off[2],
                             d + off[3], d + off[4], d + off[6]);
    HIPCHECK(hipMemcpyAsync(d + off[5], h->gradOpacity.p, N * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    if (!h->gradHost.p || h->gradHost.bytes < total * sizeof(float)) {
        h->gradHost.release();
        HIPCHECK(hipHostMalloc(&h->gradHost.p, total * sizeof(float), hipHostMallocDefault));
        h->gradHost.bytes = total * sizeof(float);
        h->gradHost.pinned = true;
    }
    float* st = h->gradHost.as<float>();
    // without full gradients only the opacity column is computed
    const size_t from = full ? 0 : off[5], to = full ? total : off[6];
    HIPCHECK(hipMemcpyAsync(st + from, d + from, (to - from) * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    HIPCHECK(hipGetLastError());
    constexpr size_t kChunk = 1u << 15;
    parallelFor((N + kChunk - 1) / kChunk, [&](size_t c) {
        const size_t a = c * kChunk, b = std::min(N, a + kChunk);
        for (int g = 0; g < 8; ++g) {
            if (!grads[g]) continue;
            const size_t w = kGradWidth[g];
            if (g == 5 || (full && g != 7))
                std::memcpy(grads[g] + a * w, st + off[g] + a * w, (b - a) * w * sizeof(float));
            else
                std::memset(grads[g] + a * w, 0, (b - a) * w * sizeof(float));
        }
    });
}

static void copyGradsDevice(pale_handle* h, float* const* grads_dev) {
    const uint32_t N = (uint32_t)h->bp.points.size();
    if (grads_dev && N) {
        const bool full = h->S.full_gradients != 0;
        for (int g = 0; g < 8; ++g) {
            if (!grads_dev[g]) continue;
            if (g == 5)
                HIPCHECK(hipMemcpyAsync(grads_dev[g], h->gradOpacity.p, (size_t)N * sizeof(float),
                                        hipMemcpyDeviceToDevice, h->stream));
            else if (!full || g == 7)
                launchFillF(h->stream, grads_dev[g], N * (uint32_t)kGradWidth[g], 0.0f);
        }
        if (full)
            launchSplitGradients(h->stream, h->gradFull.as<float>(), N, grads_dev[0], grads_dev[1], grads_dev[2],
                                 grads_dev[3], grads_dev[4], grads_dev[6]);
    }
    HIPCHECK(hipStreamSynchronize(h->stream));
    HIPCHECK(hipGetLastError());
}

int pale_render_forward(pale_handle* h, const char* camera_name, float* const* ldr_out, float* const* raw_out,
                        uint32_t* rendered) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    return guarded([&] {
        const std::vector<uint32_t> idx = selectCameras(h, camera_name);
        h->renderForward(idx);
        for (size_t k = 0; k < idx.size(); ++k) {
            const HostCamera& c = h->cams[idx[k]];
            const size_t bytes = (size_t)c.W * c.H * 4 * sizeof(float);
            if (raw_out && raw_out[k]) HIPCHECK(hipMemcpy(raw_out[k], h->fb[idx[k]].p, bytes, hipMemcpyDeviceToHost));
            if (ldr_out && ldr_out[k]) HIPCHECK(hipMemcpy(ldr_out[k], h->ldr[idx[k]].p, bytes, hipMemcpyDeviceToHost));
        }
        if (rendered) *rendered = (uint32_t)idx.size();
    });
}

int pale_render_backward(pale_handle* h, uint32_t n, const char* const* camera_names, const float* const* dloss_rgb,
                         float* const* grads, float* const* adjoint_source_out) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    if (n && (!camera_names || !dloss_rgb)) return fail(PALE_ERR_INVALID_ARG, "null camera names or images");
    return guarded([&] {
        // Adjoint cameras in camera order that appear in the input
        // (pale_bindings.cpp:301-376).
        std::vector<uint32_t> idx, which;
        std::vector<const float*> src;
        selectAdjoint(h, n, camera_names, dloss_rgb, idx, src, which);
        h->renderBackward(idx, src);
        copyGradsHost(h, grads);
        if (adjoint_source_out) {
            // The adjoint-source image is the uploaded dL/dI with alpha 1
            // (pale_bindings.cpp:352-392, :470-482), in input order.
            for (size_t k = 0; k < idx.size(); ++k) {
                float* dst = adjoint_source_out[which[k]];
                if (!dst) continue;
                const HostCamera& c = h->cams[idx[k]];
                const size_t npx = (size_t)c.W * c.H;
                for (size_t i = 0; i < npx; ++i) {
                    dst[4 * i + 0] = src[k][3 * i + 0];
                    dst[4 * i + 1] = src[k][3 * i + 1];
                    dst[4 * i + 2] = src[k][3 * i + 2];
                    dst[4 * i + 3] = 1.0f;
                }
            }
        }
    });
}

int pale_get_debug_image(const pale_handle* h, const char* camera_name, float* rgba_out) {
    if (!h || !camera_name || !rgba_out) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        if (!h->S.debug_images) throw StatusError(PALE_ERR_INVALID_ARG, "debug images: setting debug_images is off");
        for (size_t k = 0; k < h->debugCams.size() && k + 1 < h->debugOff.size(); ++k) {
            const HostCamera& c = h->cams[h->debugCams[k]];
            if (c.name != camera_name) continue;
            const size_t npx = h->debugOff[k + 1] - h->debugOff[k];
            HIPCHECK(hipMemcpyAsync(rgba_out, h->debugImg.as<float>() + 4 * h->debugOff[k], npx * 4 * sizeof(float),
                                    hipMemcpyDeviceToHost, h->stream));
            HIPCHECK(hipStreamSynchronize(h->stream));
            return;
        }
        throw StatusError(PALE_ERR_INVALID_ARG,
                          std::string("debug images: camera '") + camera_name + "' was not in the last render_backward");
    });
}

int pale_render_forward_device(pale_handle* h, const char* camera_name, float* const* ldr_dev, float* const* raw_dev,
                               uint32_t* rendered) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    return guarded([&] {
        const std::vector<uint32_t> idx = selectCameras(h, camera_name);
        h->renderForward(idx);
        for (size_t k = 0; k < idx.size(); ++k) {
            const HostCamera& c = h->cams[idx[k]];
            const size_t bytes = (size_t)c.W * c.H * 4 * sizeof(float);
            if (raw_dev && raw_dev[k])
                HIPCHECK(hipMemcpyAsync(raw_dev[k], h->fb[idx[k]].p, bytes, hipMemcpyDeviceToDevice, h->stream));
            if (ldr_dev && ldr_dev[k])
                HIPCHECK(hipMemcpyAsync(ldr_dev[k], h->ldr[idx[k]].p, bytes, hipMemcpyDeviceToDevice, h->stream));
        }
        HIPCHECK(hipStreamSynchronize(h->stream));
        if (rendered) *rendered = (uint32_t)idx.size();
    });
}

int pale_render_backward_device(pale_handle* h, uint32_t n, const char* const* camera_names,
                                const float* const* dloss_rgb_dev, float* const* grads_dev) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    if (n && (!camera_names || !dloss_rgb_dev)) return fail(PALE_ERR_INVALID_ARG, "null camera names or images");
    return guarded([&] {
        std::vector<uint32_t> idx, which;
        std::vector<const float*> src;
        selectAdjoint(h, n, camera_names, dloss_rgb_dev, idx, src, which);
        h->renderBackward(idx, src, true);
        copyGradsDevice(h, grads_dev);
    });
}

int pale_set_opacity_device(pale_handle* h, const float* opacity_dev, uint32_t n) {
    if (!h || (n && !opacity_dev)) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        if (n != h->bp.points.size())
            throw StatusError(PALE_ERR_SIZE, "opacity: expected " + std::to_string(h->bp.points.size()) + " values");
        if (!n) return;
        h->syncHostPoints();
        launchSetOpacity(h->stream, h->shade.as<SurfelShade>(), opacity_dev, n);
        std::vector<float> v(n);
        HIPCHECK(hipMemcpyAsync(v.data(), opacity_dev, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        HIPCHECK(hipStreamSynchronize(h->stream));
        HIPCHECK(hipGetLastError());
        // (the host image is rebuilt before any upload; after a device BVH
        // build it is not kept)
        const bool img = h->img.surfShade.size() == n;
        for (uint32_t i = 0; i < n; ++i) {
            h->bp.points[i].opacity = v[i];
            h->asset[i].opacity = v[i];
            if (img) h->img.surfShade[i].opacity = v[i];
        }
    });
}

int pale_apply_points_device(pale_handle* h, uint32_t n, const float* position, const float* tangent_u,
                             const float* tangent_v, const float* scale, const float* albedo, const float* opacity,
                             const float* beta) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    if (n && (!position || !tangent_u || !tangent_v || !scale || !albedo || !opacity || !beta))
        return fail(PALE_ERR_INVALID_ARG, "points: missing array");
    return guarded([&] {
        if (n == 0) return;  // ignored, as pale_apply_points ignores an empty dict
        if (n != h->bp.points.size())
            throw StatusError(PALE_ERR_SIZE,
                              "apply_point_optimization expects consistent point count; use densification API for "
                              "adding/removing points.");
        // The host mirror is left stale (refreshed from rawPts when host code
        // next reads it): the fields the device arrays do not carry (shape,
        // alpha_r, alpha_t) are never changed here, and rawPts always holds
        // the last applied parameters, so applies in a row need no download.
        h->rawPts.reserve((size_t)n * 16 * sizeof(float));
        launchApplyPoints(h->stream, h->geo.as<SurfelGeo>(), h->shade.as<SurfelShade>(), h->rawPts.as<float>(),
                          DevPointArrays{position, tangent_u, tangent_v, scale, albedo, opacity, beta}, n);
        HIPCHECK(hipGetLastError());
        h->hostStale = true;
        HIPCHECK(hipStreamSynchronize(h->stream));
    });
}

int pale_set_batch_tuning(pale_handle* h, const pale_batch_tuning* t) {
    if (!h || !t) return fail(PALE_ERR_INVALID_ARG, "null argument");
    if (!(t->grid_fraction >= 0.0f && t->grid_fraction <= 1.0f))
        return fail(PALE_ERR_INVALID_ARG, "grid_fraction must be in [0, 1]");
    if (t->graph_rounds > 64) return fail(PALE_ERR_INVALID_ARG, "graph_rounds must be <= 64");
    if (t->walk_lds_entries > 64) return fail(PALE_ERR_INVALID_ARG, "walk_lds_entries must be <= 64");
    return guarded([&] {
        h->tune = *t;
        h->ensureWalkSpill();
    });
}

int pale_get_batch_tuning(const pale_handle* h, pale_batch_tuning* t) {
    if (!h || !t) return fail(PALE_ERR_INVALID_ARG, "null argument");
    *t = h->tune;
    return PALE_OK;
}

int pale_render_views(pale_handle* h, uint32_t n, const pale_camera_desc* cameras, const uint32_t* spp,
                      const uint32_t* seeds, float* const* ldr_out, float* const* raw_out, uint32_t* batches_out) {
    if (!h || (n && (!cameras || !spp || !seeds))) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        std::vector<HostCamera> vc;
        for (uint32_t i = 0; i < n; ++i) vc.push_back(hostCamera(cameras[i]));
        h->renderViews(vc, std::vector<uint32_t>(spp, spp + n), std::vector<uint32_t>(seeds, seeds + n));
        for (uint32_t i = 0; i < n; ++i) {
            const size_t off = h->viewOffsets[i], bytes = (h->viewOffsets[i + 1] - off) * sizeof(float4);
            if (raw_out && raw_out[i])
                HIPCHECK(hipMemcpy(raw_out[i], h->viewsFb.as<float4>() + off, bytes, hipMemcpyDeviceToHost));
            if (ldr_out && ldr_out[i])
                HIPCHECK(hipMemcpy(ldr_out[i], h->viewsLdr.as<float4>() + off, bytes, hipMemcpyDeviceToHost));
        }
        if (batches_out) *batches_out = h->batchCount;
    });
}

int pale_render_views_backward(pale_handle* h, uint32_t n, const pale_camera_desc* cameras, const uint32_t* spp,
                               const float* const* dloss_rgb, float* const* grads, uint32_t* batches_out) {
    if (!h || (n && (!cameras || !spp || !dloss_rgb))) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        std::vector<HostCamera> vc;
        std::vector<const float*> src;
        for (uint32_t i = 0; i < n; ++i) {
            vc.push_back(hostCamera(cameras[i]));
            if (!dloss_rgb[i]) throw StatusError(PALE_ERR_INVALID_ARG, "null dL/dI image");
            src.push_back(dloss_rgb[i]);
        }
        h->renderViewsBackward(vc, std::vector<uint32_t>(spp, spp + n), src, false);
        copyGradsHost(h, grads);
        if (batches_out) *batches_out = h->batchCount;
    });
}

int pale_device_bytes(uint64_t* out) {
    if (!out) return fail(PALE_ERR_INVALID_ARG, "null argument");
    *out = g_deviceBytes.load();
    return PALE_OK;
}

int pale_view_bytes(pale_handle* h, uint32_t width, uint32_t height, uint32_t spp, int32_t adjoint, uint64_t* out) {
    if (!h || !out) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] { *out = h->viewGrowth(width, height, spp, adjoint != 0); });
}

int pale_views_bytes(pale_handle* h, uint32_t n, const uint32_t* widths, const uint32_t* heights,
                     const uint32_t* spp, const int32_t* adjoint, uint64_t* out) {
    if (!h || !out || (n && (!widths || !heights || !spp || !adjoint)))
        return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        if (!h->S.batch_passes) throw StatusError(PALE_ERR_INVALID_ARG, "views_bytes: needs batch_passes = 1");
        std::vector<HostCamera> vc(n);
        std::vector<uint32_t> s(n);
        std::vector<bool> adj(n);
        for (uint32_t i = 0; i < n; ++i) {
            vc[i].W = widths[i];
            vc[i].H = heights[i];
            s[i] = std::max(spp[i], 1u);
            adj[i] = adjoint[i] != 0;
        }
        *out = h->viewsGrowth(vc, s, adj);
    });
}

int pale_view_buffer_bytes(pale_handle* h, uint64_t* out) {
    if (!h || !out) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] { *out = h->viewBufferBytes(); });
}

int pale_release_view_buffers(pale_handle* h) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    return guarded([&] {
        HIPCHECK(hipStreamSynchronize(h->stream));
        h->releaseViewBuffers();
    });
}

int pale_build_photon_map(pale_handle* h) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    return guarded([&] { h->buildPhotonMap(); });
}

int pale_render_view(pale_handle* h, const pale_camera_desc* camera, uint32_t spp, uint32_t seed, float* ldr_out,
                     float* raw_out) {
    if (!h || !camera) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        const HostCamera c = hostCamera(*camera);
        h->renderView(c, spp, seed);
        const size_t bytes = (size_t)c.W * c.H * 4 * sizeof(float);
        if (raw_out) HIPCHECK(hipMemcpy(raw_out, h->outFb, bytes, hipMemcpyDeviceToHost));
        if (ldr_out) HIPCHECK(hipMemcpy(ldr_out, h->outLdr, bytes, hipMemcpyDeviceToHost));
    });
}

int pale_render_view_device(pale_handle* h, const pale_camera_desc* camera, uint32_t spp, uint32_t seed,
                            float* ldr_dev, float* raw_dev) {
    if (!h || !camera) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        const HostCamera c = hostCamera(*camera);
        h->renderView(c, spp, seed);
        const size_t bytes = (size_t)c.W * c.H * 4 * sizeof(float);
        if (raw_dev) HIPCHECK(hipMemcpyAsync(raw_dev, h->outFb, bytes, hipMemcpyDeviceToDevice, h->stream));
        if (ldr_dev) HIPCHECK(hipMemcpyAsync(ldr_dev, h->outLdr, bytes, hipMemcpyDeviceToDevice, h->stream));
        HIPCHECK(hipStreamSynchronize(h->stream));
    });
}

int pale_render_view_backward(pale_handle* h, const pale_camera_desc* camera, uint32_t spp, const float* dloss_rgb,
                              float* const* grads) {
    if (!h || !camera || !dloss_rgb) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        if (spp == 0) throw StatusError(PALE_ERR_INVALID_ARG, "render_view_backward: spp must be >= 1");
        const HostCamera c = hostCamera(*camera);
        if (h->S.batch_passes)
            h->renderViewsBackward({c}, {spp}, {dloss_rgb}, false);
        else
            h->renderBackwardCams({&c}, {dloss_rgb}, false, spp);
        copyGradsHost(h, grads);
    });
}

int pale_render_view_backward_device(pale_handle* h, const pale_camera_desc* camera, uint32_t spp,
                                     const float* dloss_rgb_dev, float* const* grads_dev) {
    if (!h || !camera || !dloss_rgb_dev) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        if (spp == 0) throw StatusError(PALE_ERR_INVALID_ARG, "render_view_backward: spp must be >= 1");
        const HostCamera c = hostCamera(*camera);
        if (h->S.batch_passes)
            h->renderViewsBackward({c}, {spp}, {dloss_rgb_dev}, true);
        else
            h->renderBackwardCams({&c}, {dloss_rgb_dev}, true, spp);
        copyGradsDevice(h, grads_dev);
    });
}

int pale_render_view_backward_noalloc(pale_handle* h, const pale_camera_desc* camera, uint32_t spp,
                                      const float* dloss_rgb_dev, float* opacity_grad_dev) {
    if (!h || !camera || !dloss_rgb_dev || !opacity_grad_dev) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        if (spp == 0) throw StatusError(PALE_ERR_INVALID_ARG, "render_view_backward_noalloc: spp must be >= 1");
        const HostCamera c = hostCamera(*camera);
        const uint64_t npx = (uint64_t)c.W * c.H;
        if (npx == 0 || npx > kNoAllocMaxPixels)
            throw StatusError(PALE_ERR_SIZE, "render_view_backward_noalloc: the view has " + std::to_string(npx) +
                                                 " pixels; at most " + std::to_string(kNoAllocMaxPixels) +
                                                 " (256 x 256) fit the scratch array");
        if (!h->scratchBase) h->scratchBase = noAllocScratch();
        h->renderViewBackwardNoAlloc(c, spp, dloss_rgb_dev, opacity_grad_dev);
    });
}

int pale_point_count(const pale_handle* h, uint32_t* count) {
    if (!h || !count) return fail(PALE_ERR_INVALID_ARG, "null argument");
    *count = (uint32_t)h->bp.points.size();
    return PALE_OK;
}

int pale_get_points(const pale_handle* h, float* position, float* tangent_u, float* tangent_v, float* scale,
                    float* albedo, float* opacity, float* beta, float* shape) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    if (h->hostStale) {
        const int rc = guarded([&] { const_cast<pale_handle*>(h)->syncHostPoints(); });
        if (rc != PALE_OK) return rc;
    }
    const auto& P = h->bp.points;
    for (size_t i = 0; i < P.size(); ++i) {
        const HostPoint& p = P[i];
        if (position) { position[3 * i] = p.p.x; position[3 * i + 1] = p.p.y; position[3 * i + 2] = p.p.z; }
        if (tangent_u) { tangent_u[3 * i] = p.tu.x; tangent_u[3 * i + 1] = p.tu.y; tangent_u[3 * i + 2] = p.tu.z; }
        if (tangent_v) { tangent_v[3 * i] = p.tv.x; tangent_v[3 * i + 1] = p.tv.y; tangent_v[3 * i + 2] = p.tv.z; }
        if (scale) { scale[2 * i] = p.su; scale[2 * i + 1] = p.sv; }
        if (albedo) { albedo[3 * i] = p.albedo.x; albedo[3 * i + 1] = p.albedo.y; albedo[3 * i + 2] = p.albedo.z; }
        if (opacity) opacity[i] = p.opacity;
        if (beta) beta[i] = p.beta;
        if (shape) shape[i] = p.shape;
    }
    return PALE_OK;
}

int pale_apply_points(pale_handle* h, const pale_points_desc* points) {
    if (!h || !points) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        if (points->count == 0) return;  // nothing to render; ignored as in the reference
        if (points->count != h->bp.points.size())
            throw StatusError(PALE_ERR_SIZE,
                              "apply_point_optimization expects consistent point count; use densification API for "
                              "adding/removing points.");
        const pale_points_desc& p = *points;
        if (!p.position || !p.tangent_u || !p.tangent_v || !p.scale || !p.albedo || !p.opacity || !p.beta)
            throw StatusError(PALE_ERR_INVALID_ARG, "points: missing array");
        h->syncHostPoints();
        constexpr size_t kChunk = 1u << 14;
        parallelFor((p.count + kChunk - 1) / kChunk, [&](size_t c) {
            for (size_t i = c * kChunk; i < std::min<size_t>(p.count, (c + 1) * kChunk); ++i) {
                HostPoint& q = h->bp.points[i];
                q.p = ld3h(p.position + 3 * i);
                q.tu = ld3h(p.tangent_u + 3 * i);
                q.tv = ld3h(p.tangent_v + 3 * i);
                q.su = p.scale[2 * i];
                q.sv = p.scale[2 * i + 1];
                q.albedo = ld3h(p.albedo + 3 * i);
                q.opacity = p.opacity[i];
                q.beta = p.beta[i];
                // Mirror into the asset (shape is not mirrored, as in the
                // reference, pale_bindings.cpp:1025-1040).
                AssetPoint& a = h->asset[i];
                a.p = q.p;
                a.tu = q.tu;
                a.tv = q.tv;
                a.su = q.su;
                a.sv = q.sv;
                a.albedo = q.albedo;
                a.opacity = q.opacity;
                a.beta = q.beta;
            }
        });
        refreshSurfels(h->bp.points, h->img);
        upload(h->stream, h->geo, h->img.surfGeo);
        upload(h->stream, h->shade, h->img.surfShade);
        h->scene.geo = h->geo.as<SurfelGeo>();
        h->scene.shade = h->shade.as<SurfelShade>();
        HIPCHECK(hipStreamSynchronize(h->stream));
        h->setScene();
    });
}

int pale_add_points(pale_handle* h, const pale_points_desc* points) {
    if (!h || !points) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        if (points->count == 0) return;
        std::vector<AssetPoint> in = toAsset(*points);
        h->syncHostPoints();
        for (AssetPoint& a : in) a.shape = 0.0f;  // new points get shape 0
        h->asset.insert(h->asset.end(), in.begin(), in.end());
        h->rebuild();
    });
}

int pale_remove_points(pale_handle* h, const int64_t* indices, size_t n) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    if (n && !indices) return fail(PALE_ERR_INVALID_ARG, "null indices");
    return guarded([&] {
        h->syncHostPoints();
        if (n == 0 || h->asset.empty()) return;
        std::vector<char> keep(h->asset.size(), 1);
        for (size_t k = 0; k < n; ++k) {
            if (indices[k] < 0) throw StatusError(PALE_ERR_RANGE, "remove_points: negative index is not allowed");
            if ((uint64_t)indices[k] >= h->asset.size())
                throw StatusError(PALE_ERR_RANGE, "remove_points: index out of range");
            keep[(size_t)indices[k]] = 0;
        }
        std::vector<AssetPoint> kept;
        for (size_t i = 0; i < h->asset.size(); ++i)
            if (keep[i]) kept.push_back(h->asset[i]);
        h->asset.swap(kept);
        h->rebuild();
    });
}

int pale_rebuild_bvh(pale_handle* h) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    return guarded([&] {
        h->syncHostPoints();
        h->rebuild();
    });
}

int pale_set_point_properties(pale_handle* h, const float translation[3], const float rotation_xyzw[4],
                              const float scale[3], const float albedo[3], float opacity, float beta, int32_t index) {
    if (!h || !translation || !rotation_xyzw || !scale || !albedo) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        h->syncHostPoints();
        auto& A = h->asset;
        const F3 t = ld3h(translation), c = ld3h(albedo);
        if (index < 0) {
            // The reference advances i twice per iteration here
            // (pale_bindings.cpp:1544-1561): every second point is changed.
            for (size_t i = 0; i < A.size(); ++i) {
                A[i].p = A[i].p + t;
                A[i].tu = rotate(rotation_xyzw, A[i].tu);
                A[i].tv = rotate(rotation_xyzw, A[i].tv);
                A[i].albedo = A[i].albedo + c;
                A[i].opacity += opacity;
                A[i].beta += beta;
                A[i].su *= scale[0];
                A[i].sv *= scale[1];
                ++i;
            }
        } else {
            if ((size_t)index > A.size() - 1 || A.empty())
                throw StatusError(PALE_ERR_INVALID_ARG, "set_point_properties: index out of range");
            AssetPoint& a = A[(size_t)index];
            a.p = a.p + t;
            a.tu = rotate(rotation_xyzw, a.tu);
            a.tv = rotate(rotation_xyzw, a.tv);
            a.albedo = c;
            a.opacity = opacity;
            a.beta = beta;
            a.su = scale[0];
            a.sv = scale[1];
        }
        h->rebuild();
    });
}

int pale_set_point_opacity(pale_handle* h, float opacity, int32_t index) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    return guarded([&] {
        h->syncHostPoints();
        if (index < 0 || (size_t)index >= h->asset.size())
            throw StatusError(PALE_ERR_INVALID_ARG, "set_gaussian_opacity: index out of range");
        h->asset[(size_t)index].opacity = opacity;
        h->rebuild();
    });
}

int pale_set_point_beta(pale_handle* h, float beta, int32_t index) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    return guarded([&] {
        h->syncHostPoints();
        if (index < 0 || (size_t)index >= h->asset.size())
            throw StatusError(PALE_ERR_INVALID_ARG, "set_point_beta: index out of range");
        h->asset[(size_t)index].beta = beta;
        h->rebuild();
    });
}

int pale_trace_rays(pale_handle* h, uint32_t n, const float* origins, const float* directions, int32_t* instance_out,
                    int32_t* primitive_out, float* t_out, float* position_out, float* alpha_out) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    if (n == 0) return PALE_OK;
    if (!origins || !directions || !instance_out || !primitive_out || !t_out || !position_out || !alpha_out)
        return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        DevMem o, d, ii, pp, tt, ps, al;
        o.reserve((size_t)n * 12);
        d.reserve((size_t)n * 12);
        ii.reserve((size_t)n * 4);
        pp.reserve((size_t)n * 4);
        tt.reserve((size_t)n * 4);
        ps.reserve((size_t)n * 12);
        al.reserve((size_t)n * 4);
        HIPCHECK(hipMemcpy(o.p, origins, (size_t)n * 12, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(d.p, directions, (size_t)n * 12, hipMemcpyHostToDevice));
        launchTraceRays(h->stream, h->scene, n, o.as<float>(), d.as<float>(), ii.as<int32_t>(), pp.as<int32_t>(),
                        tt.as<float>(), ps.as<float>(), al.as<float>());
        HIPCHECK(hipStreamSynchronize(h->stream));
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpy(instance_out, ii.p, (size_t)n * 4, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(primitive_out, pp.p, (size_t)n * 4, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(t_out, tt.p, (size_t)n * 4, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(position_out, ps.p, (size_t)n * 12, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(alpha_out, al.p, (size_t)n * 4, hipMemcpyDeviceToHost));
    });
}

int pale_gather_irradiance(pale_handle* h, uint32_t n, const float* positions, const float* normals,
                           float* irradiance_out) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    if (n == 0) return PALE_OK;
    if (!positions || !normals || !irradiance_out) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        DevMem q, nr, e;
        q.reserve((size_t)n * 12);
        nr.reserve((size_t)n * 12);
        e.reserve((size_t)n * 12);
        HIPCHECK(hipMemcpy(q.p, positions, (size_t)n * 12, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(nr.p, normals, (size_t)n * 12, hipMemcpyHostToDevice));
        launchGatherPoints(h->stream, h->devGrid(), n, q.as<float>(), nr.as<float>(), e.as<float>());
        HIPCHECK(hipStreamSynchronize(h->stream));
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpy(irradiance_out, e.p, (size_t)n * 12, hipMemcpyDeviceToHost));
    });
}

int pale_get_surfel_bvh(const pale_handle* h, uint32_t* node_count, float* bounds_out, uint32_t* left_first_out,
                        uint32_t* count_out, uint32_t* permutation_out) {
    if (!h || !node_count) return fail(PALE_ERR_INVALID_ARG, "null argument");
    const Range r = h->bp.blasRanges.back();
    *node_count = r.count;
    for (uint32_t i = 0; i < r.count; ++i) {
        const RefNode& n = h->bp.blas[r.first + i];
        if (bounds_out) {
            float* b = bounds_out + 6 * i;
            b[0] = n.mn.x; b[1] = n.mn.y; b[2] = n.mn.z;
            b[3] = n.mx.x; b[4] = n.mx.y; b[5] = n.mx.z;
        }
        if (left_first_out) left_first_out[i] = n.leftFirst;
        if (count_out) count_out[i] = n.count;
    }
    if (permutation_out && !h->bp.perm.empty())
        std::memcpy(permutation_out, h->bp.perm.data(), h->bp.perm.size() * sizeof(uint32_t));
    return PALE_OK;
}

int pale_get_photons(const pale_handle* h, uint32_t* count, uint64_t* attempted, float* position, float* power,
                     float* direction) {
    if (!h || !count || !attempted) return fail(PALE_ERR_INVALID_ARG, "null argument");
    return guarded([&] {
        DevState s{};
        HIPCHECK(hipMemcpy(&s, h->state.p, sizeof(DevState), hipMemcpyDeviceToHost));
        const uint32_t n = h->mapValid ? std::min(s.photonCount, h->photonCap) : 0u;
        *count = n;
        *attempted = h->mapValid ? s.photonCount : 0u;
        if (!n || !(position || power || direction)) return;
        std::vector<float4> a(n), b(n);
        std::vector<float> c(n);
        HIPCHECK(hipMemcpy(a.data(), h->rawPosPow.p, n * sizeof(float4), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(b.data(), h->rawDirPow.p, n * sizeof(float4), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(c.data(), h->rawPowZ.p, n * sizeof(float), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) {
            if (position) { position[3 * i] = a[i].x; position[3 * i + 1] = a[i].y; position[3 * i + 2] = a[i].z; }
            if (power) { power[3 * i] = a[i].w; power[3 * i + 1] = b[i].w; power[3 * i + 2] = c[i]; }
            if (direction) { direction[3 * i] = b[i].x; direction[3 * i + 1] = b[i].y; direction[3 * i + 2] = b[i].z; }
        }
    });
}

int pale_set_count_work(pale_handle* h, int32_t on) {
    if (!h) return fail(PALE_ERR_INVALID_ARG, "null handle");
    h->countWork = on != 0;
    return PALE_OK;
}

int pale_get_stats(const pale_handle* h, pale_stats* out) {
    if (!h || !out) return fail(PALE_ERR_INVALID_ARG, "null argument");
    *out = h->stats;
    return PALE_OK;
}

}  // extern "C"
