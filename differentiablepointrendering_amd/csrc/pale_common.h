// pale_common.h -- binary32 math, RNG and device data layouts shared by the
// host scene builder and the gfx950 kernels.
//
// Every expression keeps the reference's operand order and rounding points
// (IEEE binary32, no contraction: the library is built with
// -ffp-contract=off).  Reference semantics restated here:
//   float3 ops / dot / cross / normalize  GPUDataTypes.h:238-356
//   matrix x vec4 as row dots             GPUDataTypes.h:48-59
//   SplitMix64 / Xorshift128 / seeding    KernelHelpers.h:12-115
// normalize() is v * (1/sqrt(|v|^2)) with correctly rounded sqrt and divide
// (hipcc's default for binary32), standing in for sycl::rsqrt.
// sin/cos/exp/pow are evaluated in binary64 and rounded once.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define PALE_HD __host__ __device__ __forceinline__

namespace pale {

constexpr float kPi = 3.14159265358979323846f;      // M_PIf
constexpr float kInvPi = 0.318309886183790671538f;  // M_1_PIf
constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kMaxLeafPrims = 4u;  // BuildOptions (SceneBuild.ixx:72-75)

struct F3 {
    float x, y, z;
};

PALE_HD F3 mk3(float x, float y, float z) { return F3{x, y, z}; }
PALE_HD F3 operator+(F3 a, F3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
PALE_HD F3 operator-(F3 a, F3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
PALE_HD F3 operator*(F3 a, F3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
PALE_HD F3 operator*(F3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
PALE_HD F3 operator*(float s, F3 a) { return a * s; }
PALE_HD F3 operator-(F3 a) { return mk3(-a.x, -a.y, -a.z); }
PALE_HD F3 divs(F3 a, float s) {
    const float inv = 1.0f / s;
    return mk3(a.x * inv, a.y * inv, a.z * inv);
}
PALE_HD F3 divv(F3 a, F3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
PALE_HD float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PALE_HD F3 cross(F3 a, F3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
PALE_HD F3 normalize(F3 v) {
    const float sum = v.x * v.x + v.y * v.y + v.z * v.z;
    const float inv = 1.0f / sqrtf(sum);
    return v * inv;
}
PALE_HD float length(F3 v) { return sqrtf(dot(v, v)); }
PALE_HD F3 vmin(F3 a, F3 b) { return mk3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
PALE_HD F3 vmax(F3 a, F3 b) { return mk3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
PALE_HD float comp(F3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
PALE_HD float fsum(F3 v) { return v.x + v.y + v.z; }

PALE_HD float sin_r(float x) { return (float)sin((double)x); }
PALE_HD float cos_r(float x) { return (float)cos((double)x); }
PALE_HD float exp_r(float x) { return (float)exp((double)x); }
PALE_HD float pow_r(float b, float e) { return (float)pow((double)b, (double)e); }

// Row-major 4x4 times (x,y,z,w): out[i] = ((r0*x + r1*y) + r2*z) + r3*w.
PALE_HD void mulv4(const float* M, float x, float y, float z, float w, float out[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float* r = M + 4 * i;
        out[i] = r[0] * x + r[1] * y + r[2] * z + r[3] * w;
    }
}
// toWorldPoint (KernelHelpers.h:240-246)
PALE_HD F3 toWorldPoint(const float* o2w, F3 p) {
    float h[4];
    mulv4(o2w, p.x, p.y, p.z, 1.0f, h);
    return divs(mk3(h[0], h[1], h[2]), h[3]);
}
// transformPoint (KernelHelpers.h:229-234)
PALE_HD F3 transformPoint(const float* M, F3 p) {
    float h[4];
    mulv4(M, p.x, p.y, p.z, 1.0f, h);
    const float invW = 1.0f / h[3];
    return mk3(h[0] * invW, h[1] * invW, h[2] * invW);
}
// transformDirection (KernelHelpers.h:235-239)
PALE_HD F3 transformDirection(const float* M, F3 d) {
    float h[4];
    mulv4(M, d.x, d.y, d.z, 0.0f, h);
    return normalize(mk3(h[0], h[1], h[2]));
}

// ---------------------------------------------------------------- RNG --
PALE_HD uint64_t splitmix64_next(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint32_t x, y, z, w;
    PALE_HD explicit Rng(uint64_t seed) {
        uint64_t s = seed;
        x = (uint32_t)splitmix64_next(s);
        y = (uint32_t)splitmix64_next(s);
        z = (uint32_t)splitmix64_next(s);
        w = (uint32_t)splitmix64_next(s);
        if (!(x | y | z | w)) w = 0xA511E9B3u;
    }
    PALE_HD uint32_t nextUint() {
        const uint32_t t = x ^ (x << 11);
        x = y;
        y = z;
        z = w;
        w = (w ^ (w >> 19)) ^ (t ^ (t >> 8));
        return w;
    }
    PALE_HD float nextFloat() { return (float)(nextUint() >> 8) * 0x1.0p-24f; }
};

PALE_HD uint64_t perItemSeed(uint64_t base, uint64_t id) {
    uint64_t s = base ^ (0x9E3779B97F4A7C15ull + id);
    return splitmix64_next(s);
}

// ----------------------------------------------------- device layouts --
// Surfel, intersection record (64 B, BVH order): the fields a closest-hit
// test reads, with the unit normal of the orthonormalised tangent frame
// precomputed exactly as intersectSurfel builds it (KernelHelpers.h:646-648)
// and 4*exp(beta) for the final alpha.
struct alignas(16) SurfelGeo {
    float px, py, pz, su;
    float nx, ny, nz, sv;
    float tux, tuy, tuz, e4b;
    float tvx, tvy, tvz, pad;
};

// Surfel, shading record (32 B, BVH order).  ns is the canonical shading
// normal normalize(cross(tanU, tanV)) (KernelHelpers.h:497-501).
struct alignas(16) SurfelShade {
    float nsx, nsy, nsz, opacity;
    float ar, ag, ab, alpha_r;
};

// Internal BVH node with both child boxes (64 B).  A child reference with
// kLeafBit set is a leaf: (first << 2) | (count - 1) in the low 31 bits.
struct alignas(16) WideNode {
    float lminx, lminy, lminz, lmaxx;
    float lmaxy, lmaxz, rminx, rminy;
    float rminz, rmaxx, rmaxy, rmaxz;
    uint32_t left, right, pad0, pad1;
};

PALE_HD uint32_t leafRef(uint32_t first, uint32_t count) { return kLeafBit | (first << 2) | (count - 1u); }
PALE_HD uint32_t leafFirst(uint32_t ref) { return (ref & ~kLeafBit) >> 2; }
PALE_HD uint32_t leafCount(uint32_t ref) { return (ref & 3u) + 1u; }

// Triangle, object space, BVH order (48 B).
struct alignas(16) TriGeo {
    float v0x, v0y, v0z, pad0;
    float v1x, v1y, v1z, pad1;
    float v2x, v2y, v2z, pad2;
};

enum : uint32_t { kGeoMesh = 0u, kGeoCloud = 1u };

// TLAS node (TLASNode, GPUDataStructures.h:72-80).
struct alignas(16) TlasNode {
    float mnx, mny, mnz;
    uint32_t left;
    float mxx, mxy, mxz;
    uint32_t right;
    uint32_t count, pad0, pad1, pad2;
};

struct alignas(16) DevInstance {
    uint32_t type;       // kGeoMesh / kGeoCloud
    uint32_t rootRef;    // root of the BLAS (internal index or leaf ref)
    uint32_t material;   // mesh material, kInvalid for the cloud
    uint32_t empty;      // 1 if the BLAS has no primitives
    float rmnx, rmny, rmnz, pad0;  // BLAS root box (object space)
    float rmxx, rmxy, rmxz, pad1;
    float o2w[16];
    float w2o[16];
};

struct alignas(16) DevMaterial {
    float r, g, b, power;
};

struct alignas(16) DevLight {
    uint32_t instance, triOffset, triCount, pad;
    float cr, cg, cb, power;
    float area, pad1, pad2, pad3;
};

struct alignas(8) DevEmTri {
    uint32_t tri;
    float cdf;
};

}  // namespace pale
