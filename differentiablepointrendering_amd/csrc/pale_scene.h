// pale_scene.h -- host-side scene model and build products.
#pragma once

#include <cstdint>
#include <string>
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
};

// Builds everything from the meshes and the asset points.
BuildProducts buildScene(const std::vector<HostMesh>& meshes, const std::vector<AssetPoint>& asset);

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
};

DeviceImage makeDeviceImage(const BuildProducts& bp, const std::vector<HostMesh>& meshes);

// Refresh only the surfel records of an image (apply_point_optimization).
void refreshSurfels(const std::vector<HostPoint>& pts, DeviceImage& img);

void cameraFocal(HostCamera& c);

}  // namespace pale
