// pale_scene.hip -- host scene build: surfel collection, median-split BVHs,
// TLAS, light CDF, and conversion to the device layouts.
//
// The BVH topology and the point permutation must equal the reference's
// (BVH.ixx:49-133 with std::nth_element at count/2 on the centroid of the
// longest node axis, children allocated as an adjacent pair, depth-first
// left-then-right), so this build uses the same libstdc++ selection
// algorithm and recursion order.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <functional>
#include <limits>
#include <stdexcept>
#include <thread>

#include "pale_scene.h"

namespace pale {
namespace {

const float kInf = std::numeric_limits<float>::infinity();
const float kFltMax = std::numeric_limits<float>::max();

struct Box {
    F3 mn, mx;
};

inline float clampf(float v, float lo, float hi) { return std::min(std::max(v, lo), hi); }

// surfelObjectAabbBeta (SceneBuild.cpp:395-444)
Box surfelBox(const HostPoint& s) {
    Box b;
    surfelBoxOf(s, b.mn, b.mx);
    return b;
}

// Generic median-split builder (BVH.ixx:49-133 for boxes, :26-45/:141-220
// for triangles).  `order` holds the primitives being partitioned (an index,
// or a record that carries its index with its box and centroid); `bounds(e)`
// is the box of element e, `key(e, axis)` its centroid coordinate.
// std::nth_element's moves depend only on its comparisons, so partitioning
// records compared on the same keys leaves their indices in the same order as
// partitioning the indices themselves.
template <typename Elem, typename BoundsF, typename KeyF>
struct MedianSplit {
    std::vector<RefNode>& nodes;
    std::vector<Elem>& order;
    BoundsF bounds;
    KeyF key;
    uint32_t maxLeaf;

    // Bounds of a node's primitives.  min / max are exact in any order, so
    // large ranges are reduced in parallel chunks.
    void fit(RefNode& nd) {
        auto range = [&](uint32_t lo, uint32_t hi, F3& bmin, F3& bmax) {
            bmin = mk3(kInf, kInf, kInf);
            bmax = mk3(-kInf, -kInf, -kInf);
            for (uint32_t i = lo; i < hi; ++i) {
                const Box b = bounds(order[nd.leftFirst + i]);
                bmin = vmin(bmin, b.mn);
                bmax = vmax(bmax, b.mx);
            }
        };
        constexpr uint32_t kChunk = 1u << 16;
        if (nd.count < 4 * kChunk) {
            range(0, nd.count, nd.mn, nd.mx);
            return;
        }
        const uint32_t nc = (nd.count + kChunk - 1) / kChunk;
        std::vector<F3> mins(nc), maxs(nc);
        parallelFor(nc, [&](size_t c) {
            range((uint32_t)c * kChunk, std::min(nd.count, (uint32_t)(c + 1) * kChunk), mins[c], maxs[c]);
        });
        F3 bmin = mins[0], bmax = maxs[0];
        for (uint32_t c = 1; c < nc; ++c) {
            bmin = vmin(bmin, mins[c]);
            bmax = vmax(bmax, maxs[c]);
        }
        nd.mn = bmin;
        nd.mx = bmax;
    }

    // One split of the reference's recursion: false if nd stays a leaf,
    // else the two children (fitted) over the halves of nd's range.
    bool split(const RefNode& nd, RefNode& l, RefNode& r) {
        if (nd.count <= maxLeaf) return false;
        const F3 e = nd.mx - nd.mn;
        const int axis = (e.y > e.x && e.y > e.z) ? 1 : (e.z > e.x && e.z > e.y) ? 2 : 0;
        auto b = order.begin() + nd.leftFirst;
        auto en = b + nd.count;
        auto mid = b + (nd.count >> 1);
        std::nth_element(b, mid, en, [&](const Elem& p, const Elem& q) { return key(p, axis) < key(q, axis); });
        const uint32_t leftCount = (uint32_t)(mid - b);
        if (leftCount == 0 || leftCount == nd.count) return false;
        l = RefNode{};
        r = RefNode{};
        l.leftFirst = nd.leftFirst;
        l.count = leftCount;
        r.leftFirst = nd.leftFirst + leftCount;
        r.count = nd.count - leftCount;
        fit(l);
        fit(r);
        return true;
    }

    // Depth first, left before right; the children of a node are appended
    // as a pair when it splits.
    void subdivide(uint32_t ni) {
        RefNode l, r;
        if (!split(nodes[ni], l, r)) return;
        const uint32_t L = (uint32_t)nodes.size();
        nodes.push_back(l);
        nodes.push_back(r);
        nodes[ni].leftFirst = L;
        nodes[ni].count = 0;
        subdivide(L);
        subdivide(L + 1);
    }

    // order[] must hold the n primitives in index order.
    void build(uint32_t n) {
        nodes.clear();
        nodes.reserve(std::max<uint32_t>(1u, n * 2));
        nodes.push_back(RefNode{});
        nodes[0].leftFirst = 0;
        nodes[0].count = n;
        fit(nodes[0]);
        if (n >= kParallelBuild)
            buildParallel();
        else
            subdivide(0);
    }

    // The same tree built by several threads.  Every split works on its own
    // range of `order` and its subtree only depends on that range, so the
    // subtrees below the first kTaskDepth levels are built independently,
    // each into a local node list in the recursion's order, and then copied
    // into place by a walk that appends the pairs in the sequential
    // recursion's order: nodes, bounds and permutation come out identical.
    static constexpr uint32_t kParallelBuild = 1u << 16;
    static constexpr int kTaskDepth = 6;
    struct Part {
        RefNode nd;
        int32_t left = -1;              // index of the left child Part (right = left + 1)
        std::vector<RefNode> subtree;  // local tree below a task node (nodes[0] = the node)
        bool task = false;
    };

    void buildParallel() {
        std::vector<Part> parts;
        parts.reserve(size_t(2) << kTaskDepth);
        parts.push_back(Part{nodes[0]});
        // The top levels, breadth first (their splits touch disjoint ranges).
        std::vector<uint32_t> level{0u};
        for (int depth = 0; depth < kTaskDepth && !level.empty(); ++depth) {
            std::vector<std::pair<RefNode, RefNode>> kids(level.size());
            std::vector<char> did(level.size(), 0);
            parallelFor(level.size(), [&](size_t k) { did[k] = split(parts[level[k]].nd, kids[k].first, kids[k].second); });
            std::vector<uint32_t> next;
            for (size_t k = 0; k < level.size(); ++k) {
                if (!did[k]) continue;
                const uint32_t li = (uint32_t)parts.size();
                parts[level[k]].left = (int32_t)li;
                parts.push_back(Part{kids[k].first});
                parts.push_back(Part{kids[k].second});
                next.push_back(li);
                next.push_back(li + 1);
            }
            level.swap(next);
        }
        for (uint32_t pi : level) parts[pi].task = true;
        parallelFor(level.size(), [&](size_t k) {
            Part& p = parts[level[k]];
            MedianSplit sub{p.subtree, order, bounds, key, maxLeaf};
            p.subtree.reserve(std::max<uint32_t>(1u, p.nd.count / 2));
            p.subtree.push_back(p.nd);
            sub.subdivide(0);
        });
        nodes.clear();
        nodes.push_back(RefNode{});
        place(parts, 0, 0);
    }

    void place(const std::vector<Part>& parts, uint32_t pi, uint32_t gi) {
        const Part& p = parts[pi];
        if (p.task) {
            placeLocal(p.subtree, 0, gi);
            return;
        }
        RefNode nd = p.nd;
        if (p.left < 0) {
            nodes[gi] = nd;
            return;
        }
        const uint32_t L = (uint32_t)nodes.size();
        nodes.push_back(RefNode{});
        nodes.push_back(RefNode{});
        nd.leftFirst = L;
        nd.count = 0;
        nodes[gi] = nd;
        place(parts, (uint32_t)p.left, L);
        place(parts, (uint32_t)p.left + 1, L + 1);
    }

    void placeLocal(const std::vector<RefNode>& loc, uint32_t li, uint32_t gi) {
        RefNode nd = loc[li];
        if (nd.count != 0) {
            nodes[gi] = nd;
            return;
        }
        const uint32_t L = (uint32_t)nodes.size();
        nodes.push_back(RefNode{});
        nodes.push_back(RefNode{});
        const uint32_t lc = nd.leftFirst;
        nd.leftFirst = L;
        nodes[gi] = nd;
        placeLocal(loc, lc, L);
        placeLocal(loc, lc + 1, L + 1);
    }
};

template <typename Elem, typename BoundsF, typename KeyF>
MedianSplit<Elem, BoundsF, KeyF> medianSplit(std::vector<RefNode>& nodes, std::vector<Elem>& order, BoundsF bounds,
                                             KeyF key, uint32_t maxLeaf) {
    return MedianSplit<Elem, BoundsF, KeyF>{nodes, order, bounds, key, maxLeaf};
}

uint32_t maxDepth(const std::vector<RefNode>& nodes, uint32_t first, uint32_t count) {
    if (count == 0) return 0;
    uint32_t best = 0;
    std::vector<std::pair<uint32_t, uint32_t>> st{{0u, 0u}};
    while (!st.empty()) {
        auto [ni, d] = st.back();
        st.pop_back();
        best = std::max(best, d);
        const RefNode& n = nodes[first + ni];
        if (n.count == 0 && count > 1) {
            st.push_back({n.leftFirst, d + 1});
            st.push_back({n.leftFirst + 1, d + 1});
        }
    }
    return best;
}

}  // namespace

CloudShape cloudShape(uint32_t n, uint32_t subtree, uint32_t fitChunk) {
    struct Info {
        uint32_t first, count, depth;
        bool parentBig;
    };
    CloudShape sh;
    std::vector<RefNode>& topo = sh.topo;
    topo.reserve(n);
    std::vector<Info> info;
    info.reserve(n);
    topo.push_back(RefNode{mk3(0, 0, 0), mk3(0, 0, 0), 0, n});
    info.push_back(Info{0, n, 0, true});
    // depth first, left before right; children appended as a pair when a
    // node splits (MedianSplit::subdivide)
    std::vector<uint32_t> st{0u};
    while (!st.empty()) {
        const uint32_t ni = st.back();
        st.pop_back();
        const Info in = info[ni];
        if (in.count <= kMaxLeafPrims) continue;
        const uint32_t L = (uint32_t)topo.size(), lc = in.count >> 1;
        const bool isBig = in.count > subtree;
        topo.push_back(RefNode{mk3(0, 0, 0), mk3(0, 0, 0), in.first, lc});
        topo.push_back(RefNode{mk3(0, 0, 0), mk3(0, 0, 0), in.first + lc, in.count - lc});
        info.push_back(Info{in.first, lc, in.depth + 1, isBig});
        info.push_back(Info{in.first + lc, in.count - lc, in.depth + 1, isBig});
        topo[ni].leftFirst = L;
        topo[ni].count = 0;
        st.push_back(L + 1);
        st.push_back(L);
    }
    uint32_t levels = 0;
    for (const Info& in : info) {
        sh.depth = std::max(sh.depth, in.depth);
        if (in.count > subtree) levels = std::max(levels, in.depth + 1);
    }
    std::vector<std::vector<uint32_t>> byLevel(levels);
    for (uint32_t r = 0; r < (uint32_t)info.size(); ++r) {
        const Info& in = info[r];
        if (in.count > subtree) {
            byLevel[in.depth].push_back(r);
        } else if (in.parentBig) {
            const uint32_t rank = topo[r].count == 0 ? (topo[r].leftFirst - 1) / 2 : 0;
            sh.roots.push_back(uint4{in.first, in.count, r, rank});
        }
    }
    sh.levelBig.assign(1, 0);
    sh.levelChunk.assign(1, 0);
    for (const auto& lv : byLevel) {
        for (uint32_t r : lv) {
            const uint32_t slot = (uint32_t)sh.big.size(), a0 = info[r].first, a1 = a0 + info[r].count;
            sh.big.push_back(uint4{a0, info[r].count, r, slot});
            for (uint32_t a = a0; a < a1; a += fitChunk) sh.chunks.push_back(uint4{slot, a, std::min(a + fitChunk, a1), 0});
        }
        sh.levelBig.push_back((uint32_t)sh.big.size());
        sh.levelChunk.push_back((uint32_t)sh.chunks.size());
    }
    sh.internalOf.assign(subtree + 1, 0);
    for (uint32_t c = kMaxLeafPrims + 1; c <= subtree; ++c)
        sh.internalOf[c] = 1 + sh.internalOf[c >> 1] + sh.internalOf[c - (c >> 1)];
    // wide order = internal nodes by reference index
    std::vector<uint32_t> wideOf(topo.size(), kInvalid);
    uint32_t w = 0;
    for (uint32_t r = 0; r < (uint32_t)topo.size(); ++r)
        if (topo[r].count == 0) wideOf[r] = w++;
    sh.wide.reserve(w);
    auto code = [&](uint32_t c) { return topo[c].count == 0 ? wideOf[c] : leafRef(topo[c].leftFirst, topo[c].count); };
    for (uint32_t r = 0; r < (uint32_t)topo.size(); ++r)
        if (topo[r].count == 0) {
            const uint32_t L = topo[r].leftFirst;
            sh.wide.push_back(uint4{L, code(L), code(L + 1), 0});
        }
    return sh;
}

void cameraFocal(HostCamera& c) {
    // makePrimaryRayFromPixelJitteredFov (KernelHelpers.h:764-796): the focal
    // lengths only depend on the camera, so they are host constants.
    const float W = (float)c.W, H = (float)c.H;
    const float rad = c.fovy * 0.01745329251994329576923690768489f;  // glm::radians
    c.fy = 0.5f * H / (float)std::tan((double)(0.5f * rad));
    c.fx = c.fy * (W / H);
}

void buildScene(const std::vector<HostMesh>& meshes, const std::vector<AssetPoint>& asset, BuildProducts& bp,
                CloudBuild* cloud) {
    // The products of the last build are rebuilt in place: the large arrays
    // keep their memory (a rebuild of the same size writes into pages that
    // are already mapped).
    bp.verts.clear();
    bp.tris.clear();
    bp.meshRanges.clear();
    bp.instances.clear();
    bp.blas.clear();
    bp.blasRanges.clear();
    bp.tlas.clear();
    bp.lights.clear();
    bp.emTris.clear();
    bp.blasMaxDepth = bp.tlasMaxDepth = 0;
    bp.cloudOnDevice = false;
    // collectGeometry (SceneBuild.cpp:72-127)
    const float oneThird = 1.0f / 3.0f;
    std::vector<uint32_t> vertBase;
    for (const HostMesh& m : meshes) {
        const uint32_t vb = (uint32_t)bp.verts.size(), tb = (uint32_t)bp.tris.size();
        for (const F3& v : m.verts) bp.verts.push_back(v);
        for (size_t i = 0; i + 2 < m.idx.size(); i += 3) {
            HostTri t{m.idx[i] + vb, m.idx[i + 1] + vb, m.idx[i + 2] + vb, {0, 0, 0}};
            t.centroid = (bp.verts[t.v0] + bp.verts[t.v1] + bp.verts[t.v2]) * oneThird;
            bp.tris.push_back(t);
        }
        bp.meshRanges.push_back(Range{tb, (uint32_t)(m.idx.size() / 3)});
        vertBase.push_back(vb);
    }
    // collectInstances (SceneBuild.cpp:129-177): one instance per mesh.
    for (uint32_t i = 0; i < meshes.size(); ++i) {
        HostInstance in{};
        in.type = kGeoMesh;
        in.geom = i;
        in.material = i;
        std::copy(meshes[i].o2w, meshes[i].o2w + 16, in.o2w);
        std::copy(meshes[i].w2o, meshes[i].w2o + 16, in.w2o);
        bp.instances.push_back(in);
    }
    // collectPointCloudGeometry (SceneBuild.cpp:181-227), into the device
    // build's staging memory when it takes the scene
    const uint32_t n = (uint32_t)asset.size();
    HostPoint* pts = cloud ? cloud->staging(n) : nullptr;
    const bool staged = pts != nullptr;
    if (!staged) {
        bp.collected.resize(n);
        pts = bp.collected.data();
    }
    parallelFor((asset.size() + 16383) / 16384, [&](size_t c) {
      for (size_t i = c * 16384; i < std::min(asset.size(), (c + 1) * 16384); ++i) {
        const AssetPoint& a = asset[i];
        HostPoint p{};
        p.p = a.p;
        p.tu = normalize(a.tu);
        p.tv = normalize(a.tv);
        p.su = a.su;
        p.sv = a.sv;
        p.albedo = mk3(clampf(a.albedo.x, 0, 1), clampf(a.albedo.y, 0, 1), clampf(a.albedo.z, 0, 1));
        p.opacity = clampf(a.opacity, 0, 1);
        p.alpha_r = 1.0f;
        p.alpha_t = 0.0f;
        p.beta = a.beta;
        p.shape = clampf(a.shape, -5.0f, 5.0f);
        pts[i] = p;
      }
    });
    // The "Gaussian" point-cloud instance with identity transform
    // (pale_bindings.cpp:90-101, SceneBuild.cpp:229-251).
    {
        HostInstance in{};
        in.type = kGeoCloud;
        in.geom = 0;
        in.material = kInvalid;
        for (int k = 0; k < 16; ++k) in.o2w[k] = in.w2o[k] = (k % 5 == 0) ? 1.0f : 0.0f;
        bp.instances.push_back(in);
    }
    // collectLights (SceneBuild.cpp:263-343), before the BLAS build reorders
    // triangles: emissive-triangle indices refer to the pre-build order.
    for (uint32_t ii = 0; ii < bp.instances.size(); ++ii) {
        const HostInstance& in = bp.instances[ii];
        if (in.type == kGeoCloud) continue;
        const HostMesh& m = meshes[in.geom];
        if (!(m.power > 0.0f)) continue;
        const Range mr = bp.meshRanges[in.geom];
        if (mr.count == 0) continue;
        const uint32_t off = (uint32_t)bp.emTris.size();
        float total = 0.0f;
        for (uint32_t t = 0; t < mr.count; ++t) {
            const HostTri& T = bp.tris[mr.first + t];
            const F3 p0 = toWorldPoint(in.o2w, bp.verts[T.v0]);
            const F3 p1 = toWorldPoint(in.o2w, bp.verts[T.v1]);
            const F3 p2 = toWorldPoint(in.o2w, bp.verts[T.v2]);
            const float a = 0.5f * length(cross(p1 - p0, p2 - p0));
            total += a;
            bp.emTris.push_back(HostEmTri{mr.first + t, a, 0.0f});
        }
        if (total <= 0.0f) {
            bp.emTris.resize(off);
            continue;
        }
        float run = 0.0f;
        for (uint32_t t = 0; t < mr.count; ++t) {
            run += bp.emTris[off + t].area;
            bp.emTris[off + t].cdf = run / total;
        }
        bp.emTris[off + mr.count - 1].cdf = 1.0f;
        bp.lights.push_back(HostLight{ii, off, mr.count, m.base, m.power, total});
    }
    // buildBottomLevelBVHs (SceneBuild.cpp:642-772): meshes, then the cloud.
    std::vector<uint32_t> meshToBlas(meshes.size());
    for (uint32_t mi = 0; mi < meshes.size(); ++mi) {
        const Range mr = bp.meshRanges[mi];
        const uint32_t vb = vertBase[mi];
        std::vector<HostTri> lt(bp.tris.begin() + mr.first, bp.tris.begin() + mr.first + mr.count);
        std::vector<RefNode> nodes;
        std::vector<uint32_t> order(mr.count);
        for (uint32_t i = 0; i < mr.count; ++i) order[i] = i;
        auto ms = medianSplit(
            nodes, order,
            [&](uint32_t i) {
                const HostTri& T = lt[i];
                Box b{mk3(kInf, kInf, kInf), mk3(-kInf, -kInf, -kInf)};
                const F3 v[3] = {bp.verts[T.v0], bp.verts[T.v1], bp.verts[T.v2]};
                for (const F3& p : v) {
                    b.mn = vmin(b.mn, p);
                    b.mx = vmax(b.mx, p);
                }
                return b;
            },
            [&](uint32_t i, int a) { return comp(lt[i].centroid, a); }, kMaxLeafPrims);
        (void)vb;
        ms.build(mr.count);
        for (uint32_t k = 0; k < order.size(); ++k) bp.tris[mr.first + k] = lt[order[k]];
        const uint32_t first = (uint32_t)bp.blas.size();
        for (RefNode& nd : nodes)
            if (nd.count > 0) nd.leftFirst += mr.first;
        bp.blas.insert(bp.blas.end(), nodes.begin(), nodes.end());
        meshToBlas[mi] = (uint32_t)bp.blasRanges.size();
        bp.blasRanges.push_back(Range{first, (uint32_t)nodes.size()});
    }
    uint32_t cloudDepth = 0;
    {
        const size_t chunk = 1u << 14;
        uint32_t wideBase = 0;
        for (const Range& r : bp.blasRanges)
            for (uint32_t i = 0; i < r.count; ++i)
                wideBase += (bp.blas[r.first + i].count == 0 && r.count > 1) ? 1u : 0u;
        const uint32_t first = (uint32_t)bp.blas.size();
        if (staged && cloud->build(pts, n, wideBase, bp.blas, bp.perm, cloudDepth)) {
            bp.cloudOnDevice = true;
        } else {
            std::vector<RefNode> nodes;
            // Each surfel as one record (centroid, box, index), so that the
            // selections and the box fits stream through memory instead of
            // gathering keys and boxes by index.
            struct Prim {
                float key[3];
                uint32_t idx;
                Box box;
            };
            std::vector<Prim> prims(n);
            parallelFor((n + chunk - 1) / chunk, [&](size_t c) {
                for (size_t i = c * chunk; i < std::min<size_t>(n, (c + 1) * chunk); ++i)
                    prims[i] = Prim{{pts[i].p.x, pts[i].p.y, pts[i].p.z}, (uint32_t)i, surfelBox(pts[i])};
            });
            auto ms = medianSplit(
                nodes, prims, [](const Prim& e) { return e.box; }, [](const Prim& e, int a) { return e.key[a]; },
                kMaxLeafPrims);
            ms.build(n);
            bp.perm.resize(n);
            parallelFor((n + chunk - 1) / chunk, [&](size_t c) {
                for (size_t k = c * chunk; k < std::min<size_t>(n, (c + 1) * chunk); ++k) bp.perm[k] = prims[k].idx;
            });
            cloudDepth = maxDepth(nodes, 0, (uint32_t)nodes.size());
            bp.blas.insert(bp.blas.end(), nodes.begin(), nodes.end());
        }
        bp.points.resize(n);
        parallelFor((n + chunk - 1) / chunk, [&](size_t c) {
            for (size_t k = c * chunk; k < std::min<size_t>(n, (c + 1) * chunk); ++k) bp.points[k] = pts[bp.perm[k]];
        });
        const uint32_t bi = (uint32_t)bp.blasRanges.size();
        bp.blasRanges.push_back(Range{first, (uint32_t)bp.blas.size() - first});
        for (HostInstance& in : bp.instances) in.blas = (in.type == kGeoMesh) ? meshToBlas[in.geom] : bi;
    }
    // the surfel BLAS's depth comes with its build
    for (size_t i = 0; i + 1 < bp.blasRanges.size(); ++i)
        bp.blasMaxDepth = std::max(bp.blasMaxDepth, maxDepth(bp.blas, bp.blasRanges[i].first, bp.blasRanges[i].count));
    bp.blasMaxDepth = std::max(bp.blasMaxDepth, cloudDepth);
    // buildTLAS (SceneBuild.cpp:555-640)
    struct IBox {
        F3 mn, mx;
        uint32_t inst;
    };
    std::vector<IBox> boxes;
    for (uint32_t i = 0; i < bp.instances.size(); ++i) {
        const HostInstance& in = bp.instances[i];
        const RefNode& root = bp.blas[bp.blasRanges[in.blas].first];
        F3 wmin = mk3(kFltMax, kFltMax, kFltMax), wmax = mk3(-kFltMax, -kFltMax, -kFltMax);
        for (int c = 0; c < 8; ++c) {
            const bool bx = c & 4, by = c & 2, bz = c & 1;
            const F3 po = mk3(bx ? root.mx.x : root.mn.x, by ? root.mx.y : root.mn.y, bz ? root.mx.z : root.mn.z);
            const F3 pw = toWorldPoint(in.o2w, po);
            wmin = vmin(wmin, pw);
            wmax = vmax(wmax, pw);
        }
        boxes.push_back(IBox{wmin, wmax, i});
    }
    uint32_t tdepth = 0;
    std::function<uint32_t(uint32_t, uint32_t, uint32_t)> build = [&](uint32_t start, uint32_t end,
                                                                      uint32_t depth) -> uint32_t {
        tdepth = std::max(tdepth, depth);
        const uint32_t n = (uint32_t)bp.tlas.size();
        bp.tlas.push_back(TlasNode{});
        F3 bmin = mk3(kFltMax, kFltMax, kFltMax), bmax = mk3(-kFltMax, -kFltMax, -kFltMax);
        for (uint32_t i = start; i < end; ++i) {
            bmin = vmin(bmin, boxes[i].mn);
            bmax = vmax(bmax, boxes[i].mx);
        }
        TlasNode& N0 = bp.tlas[n];
        N0.mnx = bmin.x; N0.mny = bmin.y; N0.mnz = bmin.z;
        N0.mxx = bmax.x; N0.mxy = bmax.y; N0.mxz = bmax.z;
        const uint32_t count = end - start;
        if (count == 1) {
            bp.tlas[n].count = 1;
            bp.tlas[n].left = boxes[start].inst;
            bp.tlas[n].right = 0;
        } else {
            bp.tlas[n].count = 0;
            F3 cmin = mk3(kFltMax, kFltMax, kFltMax), cmax = mk3(-kFltMax, -kFltMax, -kFltMax);
            for (uint32_t i = start; i < end; ++i) {
                const F3 c = (boxes[i].mn + boxes[i].mx) * 0.5f;
                cmin = vmin(cmin, c);
                cmax = vmax(cmax, c);
            }
            const F3 ext = cmax - cmin;
            const int axis = (ext.x > ext.y && ext.x > ext.z) ? 0 : (ext.y > ext.z) ? 1 : 2;
            const float pivot = (comp(cmin, axis) + comp(cmax, axis)) * 0.5f;
            auto it = std::partition(boxes.begin() + start, boxes.begin() + end, [&](const IBox& b) {
                const F3 c = (b.mn + b.mx) * 0.5f;
                return comp(c, axis) < pivot;
            });
            uint32_t mid = (uint32_t)(it - boxes.begin());
            if (mid == start || mid == end) mid = start + count / 2;
            const uint32_t l = build(start, mid, depth + 1);
            const uint32_t r = build(mid, end, depth + 1);
            bp.tlas[n].left = l;
            bp.tlas[n].right = r;
        }
        return n;
    };
    if (!boxes.empty()) build(0, (uint32_t)boxes.size(), 0);
    bp.tlasMaxDepth = tdepth;
}

void refreshSurfels(const std::vector<HostPoint>& pts, DeviceImage& img) {
    img.surfGeo.resize(pts.size());
    img.surfShade.resize(pts.size());
    constexpr size_t kChunk = 1u << 14;
    parallelFor((pts.size() + kChunk - 1) / kChunk, [&](size_t c) {
      for (size_t i = c * kChunk; i < std::min(pts.size(), (c + 1) * kChunk); ++i) {
        img.surfGeo[i] = surfelGeoOf(pts[i]);
        img.surfShade[i] = surfelShadeOf(pts[i]);
      }
    });
}

void makeDeviceImage(const BuildProducts& bp, const std::vector<HostMesh>& meshes, DeviceImage& img) {
    if (bp.points.size() >= (1u << 29)) throw std::runtime_error("more than 2^29 surfels");
    img.nodes.clear();
    img.tris.clear();
    img.instances.clear();
    img.tlas.clear();
    img.materials.clear();
    img.lights.clear();
    img.emTris.clear();
    if (!bp.cloudOnDevice) refreshSurfels(bp.points, img);
    img.tris.resize(bp.tris.size());
    for (size_t i = 0; i < bp.tris.size(); ++i) {
        const HostTri& t = bp.tris[i];
        const F3 a = bp.verts[t.v0], b = bp.verts[t.v1], c = bp.verts[t.v2];
        img.tris[i] = TriGeo{a.x, a.y, a.z, 0.f, b.x, b.y, b.z, 0.f, c.x, c.y, c.z, 0.f};
    }
    // Wide nodes: every internal reference node becomes one WideNode holding
    // both child boxes; leaves become references.
    std::vector<uint32_t> compact(bp.blas.size(), kInvalid);
    uint32_t next = 0;
    for (const Range& r : bp.blasRanges) {
        for (uint32_t i = 0; i < r.count; ++i) {
            const RefNode& n = bp.blas[r.first + i];
            const bool internal = n.count == 0 && r.count > 1;
            if (internal) compact[r.first + i] = next++;
        }
    }
    // the cloud's wide nodes come last; a device build has written them
    const Range& cloudRange = bp.blasRanges.back();
    uint32_t meshWide = next;
    for (uint32_t i = 0; i < cloudRange.count; ++i)
        if (compact[cloudRange.first + i] != kInvalid) meshWide = std::min(meshWide, compact[cloudRange.first + i]);
    img.meshWide = meshWide;
    img.nodes.resize(bp.cloudOnDevice ? meshWide : next);
    auto refOf = [&](const Range& r, uint32_t local) -> uint32_t {
        const RefNode& n = bp.blas[r.first + local];
        if (n.count > 0) return leafRef(n.leftFirst, n.count);
        return compact[r.first + local];
    };
    for (const Range& r : bp.blasRanges) {
      if (bp.cloudOnDevice && &r == &cloudRange) continue;
      parallelFor((r.count + 16383u) / 16384u, [&](size_t c) {
        for (uint32_t i = (uint32_t)c * 16384u; i < std::min(r.count, (uint32_t)(c + 1) * 16384u); ++i) {
            const uint32_t ci = compact[r.first + i];
            if (ci == kInvalid) continue;
            const RefNode& n = bp.blas[r.first + i];
            const RefNode& L = bp.blas[r.first + n.leftFirst];
            const RefNode& R = bp.blas[r.first + n.leftFirst + 1];
            WideNode w{};
            w.lminx = L.mn.x; w.lminy = L.mn.y; w.lminz = L.mn.z; w.lmaxx = L.mx.x;
            w.lmaxy = L.mx.y; w.lmaxz = L.mx.z; w.rminx = R.mn.x; w.rminy = R.mn.y;
            w.rminz = R.mn.z; w.rmaxx = R.mx.x; w.rmaxy = R.mx.y; w.rmaxz = R.mx.z;
            w.left = refOf(r, n.leftFirst);
            w.right = refOf(r, n.leftFirst + 1);
            img.nodes[ci] = w;
        }
      });
    }
    for (const HostInstance& in : bp.instances) {
        const Range& r = bp.blasRanges[in.blas];
        const RefNode& root = bp.blas[r.first];
        DevInstance d{};
        d.type = in.type;
        d.empty = root.count == 0 && r.count <= 1 ? 1u : 0u;
        d.rootRef = d.empty ? kInvalid : refOf(r, 0);
        d.material = in.material;
        d.rmnx = root.mn.x; d.rmny = root.mn.y; d.rmnz = root.mn.z;
        d.rmxx = root.mx.x; d.rmxy = root.mx.y; d.rmxz = root.mx.z;
        std::copy(in.o2w, in.o2w + 16, d.o2w);
        std::copy(in.w2o, in.w2o + 16, d.w2o);
        img.instances.push_back(d);
    }
    img.tlas = bp.tlas;
    for (const HostMesh& m : meshes) img.materials.push_back(DevMaterial{m.base.x, m.base.y, m.base.z, m.power});
    for (const HostLight& l : bp.lights) {
        DevLight d{};
        d.instance = l.instance;
        d.triOffset = l.triOffset;
        d.triCount = l.triCount;
        d.cr = l.color.x; d.cg = l.color.y; d.cb = l.color.z;
        d.power = l.power;
        d.area = l.area;
        img.lights.push_back(d);
    }
    for (const HostEmTri& e : bp.emTris) img.emTris.push_back(DevEmTri{e.tri, e.cdf});
    // Near-far depth-first traversal keeps at most one pending sibling per
    // level: depth + 1 entries per BLAS, tlasDepth + 1 below it.
    img.stackDepth = (bp.tlasMaxDepth + 1) + (bp.blasMaxDepth + 1);
}

}  // namespace pale
