// pale_device.h -- device functions of the gfx950 renderer: closest-hit
// traversal, surfel and triangle tests, sampling, Russian roulette, light
// sampling and the photon-grid irradiance gather.
//
// Traversal keeps the reference's visiting order exactly (near-far child
// push with ties to the left child, IntersectionKernels.h:21-32 and
// :585-675), so the closest hit -- including which surfel wins an exact tie
// in t -- is the reference's.  The stack lives in LDS: entry k of a lane is
// at stk[k * stride], stride = blockDim.x (bank-conflict free per row).  A
// stack entry is (node reference, entry distance); re-testing a popped box
// against the current best distance is done with the stored entry distance,
// which equals the reference's full slab re-test because the other two
// rejection conditions do not depend on the best distance.
#pragma once

#include "pale_common.h"

namespace pale {

struct DevScene {
    const SurfelGeo* geo;
    const SurfelShade* shade;
    const WideNode* nodes;
    const TriGeo* tris;
    const DevInstance* inst;
    const TlasNode* tlas;
    const DevMaterial* mats;
    const DevLight* lights;
    const DevEmTri* emTris;
    uint32_t nInst;
    uint32_t nTlas;
    uint32_t nLights;
    uint32_t stackDepth;
    uint32_t cloudInst;  // instance of the point cloud (kInvalid if none)
};

struct DevGrid {
    const float4* posPow;  // sorted photons: position, power.x
    const float4* dirPow;  // incoming direction, power.y
    const float* powZ;     // power.z
    const uint32_t* cellStart;  // exclusive prefix over cells, total+1 entries
    float ox, oy, oz;
    float cx, cy, cz;
    int rx, ry, rz;
    float radius;
    const uint32_t* attempted;  // deposits attempted (T), device memory
    uint32_t capacity;          // photon capacity
    uint32_t valid;      // a photon map exists
};

struct Hit {
    F3 pos;
    float t;
    float alpha;
    uint32_t inst;  // kInvalid: miss
    uint32_t prim;
};

// Per-thread work counters of the counting mode (the per-query units of the
// algorithmic byte count: popped nodes, AABB slab tests, primitive tests,
// gather cells / non-empty cells / photons).
struct Work {
    uint32_t pops, aabb, prims, cells, cellsNE, photons, gathers;
};

__device__ __forceinline__ F3 ld3(const float4& v) { return mk3(v.x, v.y, v.z); }

// safeInvDir (KernelHelpers.h:140-147)
__device__ __forceinline__ F3 safeInvDir(F3 d) {
    return mk3(fabsf(d.x) < 1e-6f ? 1e30f : 1.0f / d.x, fabsf(d.y) < 1e-6f ? 1e30f : 1.0f / d.y,
               fabsf(d.z) < 1e-6f ? 1e30f : 1.0f / d.z);
}

// slabIntersectAABB (KernelHelpers.h:149-214); clampLo is 0 for TLAS nodes
// and 1e-6 for BVH nodes.
__device__ __forceinline__ bool slab(F3 o, F3 inv, float mnx, float mny, float mnz, float mxx, float mxy,
                                     float mxz, float limit, float clampLo, float& tEntry) {
    const float t0x = (mnx - o.x) * inv.x, t0y = (mny - o.y) * inv.y, t0z = (mnz - o.z) * inv.z;
    const float t1x = (mxx - o.x) * inv.x, t1y = (mxy - o.y) * inv.y, t1z = (mxz - o.z) * inv.z;
    const float tmin = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    const float tmax = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    if (tmin > tmax) return false;
    if (tmax < 0.0f) return false;
    if (tmin > limit) return false;
    tEntry = fmaxf(tmin, clampLo);
    return true;
}

// Entries below inLds live in LDS (entry k at base[k * stride]); deeper
// ones, if any, in a per-lane global area (entry k at
// spill[(k - inLds) * spillStride]).
// An LDS word pair.  The device pass types it as LDS memory so that stack
// accesses compile to LDS instructions; the host pass only parses the code.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) uint2 LdsUint2;
typedef __attribute__((address_space(3))) uint64_t LdsU64;
#else
typedef uint2 LdsUint2;
typedef uint64_t LdsU64;
#endif

struct LaneStack {
    LdsUint2* base;
    uint32_t stride;
    uint32_t cap;
    uint2* spill;
    uint32_t spillStride;
    uint32_t inLds;
    // SmallStack::push (GPUDataTypes.h:398-402): a push beyond capacity is
    // dropped, so sp never exceeds cap and every pop reads a written slot.
    __device__ __forceinline__ void push(uint32_t& sp, uint32_t ref, float t) const {
        if (sp < cap) {
            const uint2 v = make_uint2(ref, __float_as_uint(t));
            if (sp < inLds)
                base[sp * stride] = v;
            else
                spill[(sp - inLds) * spillStride] = v;
            ++sp;
        }
    }
    // The LDS read is unconditional and a relaxed atomic load, so that it
    // stays an LDS instruction: two plain loads would be merged into one
    // generic (flat) load of a selected address, which goes through the
    // vector-memory path even when it reads LDS.  A spilled entry is then
    // read again from the spill area.
    __device__ __forceinline__ uint2 get(uint32_t k) const {
        const LdsU64* p = reinterpret_cast<const LdsU64*>(base + min(k, inLds - 1u) * stride);
        const uint64_t bits = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        uint2 v = make_uint2((uint32_t)bits, (uint32_t)(bits >> 32));
        if (k >= inLds) v = spill[(k - inLds) * spillStride];
        return v;
    }
};

struct LocalHit {
    F3 pos;
    float t;
    float alpha;
    uint32_t prim;
};

// intersectBLASPointCloudFirstHit (IntersectionKernels.h:585-675) with
// intersectSurfel (KernelHelpers.h:637-669); alpha_geom is evaluated once for
// the final hit (it is a pure function of the hit).
template <bool kCount>
__device__ bool traverseCloud(const DevScene& S, const DevInstance& I, F3 o, F3 d, const LaneStack& st,
                              uint32_t base, LocalHit& out, Work& w) {
    const F3 inv = safeInvDir(d);
    float best = __builtin_huge_valf();
    float te;
    if (kCount) w.aabb++;  // the reference slab-tests the root when it pops it
    if (!slab(o, inv, I.rmnx, I.rmny, I.rmnz, I.rmxx, I.rmxy, I.rmxz, best, 1e-6f, te)) {
        if (kCount) w.pops++;
        return false;
    }
    bool firstPop = true;
    uint32_t sp = base;
    st.push(sp, I.rootRef, te);
    uint32_t bestIdx = 0;
    float bestR2 = 0.0f;
    F3 bestX = mk3(0, 0, 0);
    bool any = false;
    while (sp > base) {
        const uint2 e = st.get(--sp);
        if (kCount) {
            w.pops++;
            if (!firstPop) w.aabb++;  // re-test of a popped box (root already counted)
            firstPop = false;
        }
        if (__uint_as_float(e.y) > best) continue;
        const uint32_t ref = e.x;
        if (ref & kLeafBit) {
            const uint32_t first = leafFirst(ref), cnt = leafCount(ref);
            if (kCount) w.prims += cnt;
            for (uint32_t k = 0; k < cnt; ++k) {
                const uint32_t si = first + k;
                const float4* g = reinterpret_cast<const float4*>(S.geo + si);
                const float4 g0 = g[0], g1 = g[1];
                const F3 n = mk3(g1.x, g1.y, g1.z);
                const float nd = dot(n, d);
                if (fabsf(nd) < 1e-6f) continue;
                const F3 p = mk3(g0.x, g0.y, g0.z);
                const float t = dot(n, p - o) / nd;
                if (t <= 1e-5f || t >= best) continue;
                const float4 g2 = g[2], g3 = g[3];
                const F3 x = o + d * t;
                const F3 r = x - p;
                const float uu = dot(mk3(g2.x, g2.y, g2.z), r) / g0.w;
                const float vv = dot(mk3(g3.x, g3.y, g3.z), r) / g1.w;
                const float r2 = uu * uu + vv * vv;
                if (r2 >= 1.0f) continue;
                any = true;
                best = t;
                bestIdx = si;
                bestX = x;
                bestR2 = r2;
            }
            continue;
        }
        const float4* nd4 = reinterpret_cast<const float4*>(S.nodes + ref);
        const float4 a = nd4[0], b = nd4[1], c = nd4[2];
        const uint2 refs = reinterpret_cast<const uint2*>(S.nodes + ref)[6];  // left, right (not the padding)
        if (kCount) w.aabb += 2;
        float tl = __builtin_huge_valf(), tr = __builtin_huge_valf();
        const bool hl = slab(o, inv, a.x, a.y, a.z, a.w, b.x, b.y, best, 1e-6f, tl);
        const bool hr = slab(o, inv, b.z, b.w, c.x, c.y, c.z, c.w, best, 1e-6f, tr);
        if (hl && hr) {
            if (tl <= tr) {
                st.push(sp, refs.y, tr);
                st.push(sp, refs.x, tl);
            } else {
                st.push(sp, refs.x, tl);
                st.push(sp, refs.y, tr);
            }
        } else if (hl) {
            st.push(sp, refs.x, tl);
        } else if (hr) {
            st.push(sp, refs.y, tr);
        }
    }
    if (!any) return false;
    out.t = best;
    out.prim = bestIdx;
    out.pos = bestX;
    out.alpha = pow_r(1.0f - bestR2, S.geo[bestIdx].e4b);
    return true;
}

// Words of a leaf's surfels loaded before its tests (kLeafWords): 4 = the
// whole 64-byte records, which suits queues whose neighbouring lanes trace
// unrelated rays (photon and adjoint bounces); 2 = the plane words only, the
// tangents of a surfel loaded when its plane hit is in range, which suits
// the camera walk (fewer bytes per lane when neighbouring rays diverge).
// Same tests in the same order either way.  A node that every lane of the
// wave is at is read with scalar loads: in the camera walk, whose waves hold
// neighbouring pixels (and with shared seeds the same ray in several
// passes), 57 % of the wave-wide node reads of the plain workload with
// shared seeds, 31 % with independent samples; fewer in the photon and
// adjoint queues, where it still saves 3-4 %.

// Node and surfel records read through the constant address space (scalar
// loads when the address is the same in every lane).
typedef __attribute__((address_space(4))) const WideNode ConstWideNode;

// The closest hit's alpha, (1 - r^2)^(4 e^beta) in binary64, as a call: its
// 64-bit constants then stay inside the call instead of being hoisted into
// registers that live across the traversal loop (inlined, they pushed the
// camera walk into scratch spills; as a call the walk has none, and the
// walk, photon and adjoint traces run 2-10 % faster).
__device__ __attribute__((noinline)) float hitAlpha(float oneMinusR2, float e4b) { return pow_r(oneMinusR2, e4b); }

// The same traversal without the counters, and without a stack round trip
// for the nearer child: the reference pushes the farther child, then the
// nearer, and pops the nearer at once with an unchanged best, so descending
// into the nearer child directly visits the same nodes in the same order.
template <int kLeafWords>
__device__ bool traverseCloudFast(const DevScene& S, const DevInstance& I, F3 o, F3 d, const LaneStack& st,
                                  uint32_t base, LocalHit& out) {
    const F3 inv = safeInvDir(d);
    float best = __builtin_huge_valf();
    float te;
    if (!slab(o, inv, I.rmnx, I.rmny, I.rmnz, I.rmxx, I.rmxy, I.rmxz, best, 1e-6f, te)) return false;
    uint32_t sp = base;
    uint32_t ref = I.rootRef;
    uint32_t bestIdx = 0;
    float bestR2 = 0.0f;
    F3 bestX = mk3(0, 0, 0);
    bool any = false;
    while (true) {
        if (ref & kLeafBit) {
            // A leaf holds at most four surfels: their plane words (center,
            // normal, scales) are loaded together before the first test, the
            // tangents of a surfel only when its plane hit is in range.  The
            // tests run in leaf order as before.
            const uint32_t first = leafFirst(ref), cnt = leafCount(ref);
            const float4* g = reinterpret_cast<const float4*>(S.geo + first);
            float4 w0[4], w1[4], w2[4], w3[4];
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                if (k < cnt) {
                    w0[k] = g[4 * k];
                    w1[k] = g[4 * k + 1];
                    if (kLeafWords == 4) {
                        w2[k] = g[4 * k + 2];
                        w3[k] = g[4 * k + 3];
                    }
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                if (k >= cnt) break;
                const float4 g0 = w0[k], g1 = w1[k];
                const F3 n = mk3(g1.x, g1.y, g1.z);
                const float nd = dot(n, d);
                if (fabsf(nd) < 1e-6f) continue;
                const F3 p = mk3(g0.x, g0.y, g0.z);
                const float t = dot(n, p - o) / nd;
                if (t <= 1e-5f || t >= best) continue;
                float4 g2, g3;
                if (kLeafWords == 4) {
                    g2 = w2[k];
                    g3 = w3[k];
                } else {
                    g2 = g[4 * k + 2];
                    g3 = g[4 * k + 3];
                }
                const F3 x = o + d * t;
                const F3 r = x - p;
                const float uu = dot(mk3(g2.x, g2.y, g2.z), r) / g0.w;
                const float vv = dot(mk3(g3.x, g3.y, g3.z), r) / g1.w;
                const float r2 = uu * uu + vv * vv;
                if (r2 >= 1.0f) continue;
                any = true;
                best = t;
                bestIdx = first + k;
                bestX = x;
                bestR2 = r2;
            }
        } else {
            float4 a, b, c;
            uint2 refs;
            const uint32_t r0 = __builtin_amdgcn_readfirstlane(ref);
            if (__builtin_amdgcn_ballot_w64(ref != r0) == 0ull) {
                // Every lane of the wave here is at the same node: read it
                // with scalar loads, off the vector-memory path.  (The empty
                // asm keeps the compiler from folding this read into the
                // vector one below.)
                __asm__ volatile("");
                const ConstWideNode* cn = (const ConstWideNode*)(S.nodes + r0);
                a = make_float4(cn->lminx, cn->lminy, cn->lminz, cn->lmaxx);
                b = make_float4(cn->lmaxy, cn->lmaxz, cn->rminx, cn->rminy);
                c = make_float4(cn->rminz, cn->rmaxx, cn->rmaxy, cn->rmaxz);
                refs = make_uint2(cn->left, cn->right);
            } else {
                const float4* nd4 = reinterpret_cast<const float4*>(S.nodes + ref);
                a = nd4[0];
                b = nd4[1];
                c = nd4[2];
                refs = reinterpret_cast<const uint2*>(S.nodes + ref)[6];  // left, right (not the padding)
            }
            float tl = __builtin_huge_valf(), tr = __builtin_huge_valf();
            const bool hl = slab(o, inv, a.x, a.y, a.z, a.w, b.x, b.y, best, 1e-6f, tl);
            const bool hr = slab(o, inv, b.z, b.w, c.x, c.y, c.z, c.w, best, 1e-6f, tr);
            if (hl && hr) {
                if (tl <= tr) {
                    st.push(sp, refs.y, tr);
                    ref = refs.x;
                } else {
                    st.push(sp, refs.x, tl);
                    ref = refs.y;
                }
                continue;
            }
            if (hl) {
                ref = refs.x;
                continue;
            }
            if (hr) {
                ref = refs.y;
                continue;
            }
        }
        // pop, skipping entries that start beyond the current best
        bool found = false;
        while (sp > base) {
            const uint2 e = st.get(--sp);
            if (__uint_as_float(e.y) > best) continue;
            ref = e.x;
            found = true;
            break;
        }
        if (!found) break;
    }
    if (!any) return false;
    out.t = best;
    out.prim = bestIdx;
    out.pos = bestX;
    out.alpha = hitAlpha(1.0f - bestR2, S.geo[bestIdx].e4b);
    return true;
}

// intersectTriangle (KernelHelpers.h:248-277)
__device__ __forceinline__ bool intersectTri(F3 o, F3 d, F3 v0, F3 v1, F3 v2, float tMin, float& tOut) {
    const F3 e1 = v1 - v0, e2 = v2 - v0;
    const F3 h = cross(d, e2);
    const float a = dot(e1, h);
    if (fabsf(a) < 1.0e-4f) return false;
    const float f = 1.0f / a;
    const F3 s = o - v0;
    const float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    const F3 q = cross(s, e1);
    const float v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    const float t = f * dot(e2, q);
    if (t <= tMin) return false;
    tOut = t;
    return true;
}

// intersectBLASMesh (IntersectionKernels.h:37-105)
template <bool kCount>
__device__ bool traverseMesh(const DevScene& S, const DevInstance& I, F3 o, F3 d, const LaneStack& st,
                             uint32_t base, LocalHit& out, Work& w) {
    const F3 inv = safeInvDir(d);
    float best = __builtin_huge_valf();
    float te;
    if (kCount) w.aabb++;  // the reference slab-tests the root when it pops it
    if (!slab(o, inv, I.rmnx, I.rmny, I.rmnz, I.rmxx, I.rmxy, I.rmxz, best, 1e-6f, te)) {
        if (kCount) w.pops++;
        return false;
    }
    bool firstPop = true;
    uint32_t sp = base;
    st.push(sp, I.rootRef, te);
    bool any = false;
    uint32_t bestTri = 0;
    while (sp > base) {
        const uint2 e = st.get(--sp);
        if (kCount) {
            w.pops++;
            if (!firstPop) w.aabb++;  // re-test of a popped box (root already counted)
            firstPop = false;
        }
        if (__uint_as_float(e.y) > best) continue;
        const uint32_t ref = e.x;
        if (ref & kLeafBit) {
            const uint32_t first = leafFirst(ref), cnt = leafCount(ref);
            if (kCount) w.prims += cnt;
            for (uint32_t k = 0; k < cnt; ++k) {
                const TriGeo tg = S.tris[first + k];
                float t;
                if (intersectTri(o, d, mk3(tg.v0x, tg.v0y, tg.v0z), mk3(tg.v1x, tg.v1y, tg.v1z),
                                 mk3(tg.v2x, tg.v2y, tg.v2z), 1e-4f, t) &&
                    t < best && t > 0.0f) {
                    best = t;
                    any = true;
                    bestTri = first + k;
                }
            }
            continue;
        }
        const float4* nd4 = reinterpret_cast<const float4*>(S.nodes + ref);
        const float4 a = nd4[0], b = nd4[1], c = nd4[2];
        const uint2 refs = reinterpret_cast<const uint2*>(S.nodes + ref)[6];  // left, right (not the padding)
        if (kCount) w.aabb += 2;
        float tl = __builtin_huge_valf(), tr = __builtin_huge_valf();
        const bool hl = slab(o, inv, a.x, a.y, a.z, a.w, b.x, b.y, best, 1e-6f, tl);
        const bool hr = slab(o, inv, b.z, b.w, c.x, c.y, c.z, c.w, best, 1e-6f, tr);
        if (hl && hr) {
            if (tl <= tr) {
                st.push(sp, refs.y, tr);
                st.push(sp, refs.x, tl);
            } else {
                st.push(sp, refs.x, tl);
                st.push(sp, refs.y, tr);
            }
        } else if (hl) {
            st.push(sp, refs.x, tl);
        } else if (hr) {
            st.push(sp, refs.y, tr);
        }
    }
    if (!any) return false;
    out.t = best;
    out.prim = bestTri;
    out.pos = toWorldPoint(I.o2w, o + d * best);
    out.alpha = 0.0f;
    return true;
}

// intersectScene, FirstHit mode (IntersectionKernels.h:954-1097).  The TLAS
// stack shares the lane's LDS stack below the BLAS entries.
// The scene's top level (TLAS nodes and instances), read where it lies in
// global memory (TopGlobal) or from a copy in LDS that a workgroup makes when
// it starts (TopLds: every lane of a wave reads the same few nodes and
// instances, which the LDS broadcasts without using the vector-memory path).
struct TopGlobal {
    const TlasNode* tlas;
    const DevInstance* inst;
};
constexpr uint32_t kTopLdsTlas = 15, kTopLdsInst = 8;
struct TopLds {
    TlasNode tlas[kTopLdsTlas];
    DevInstance inst[kTopLdsInst];
};
__host__ __device__ inline bool topFitsLds(const DevScene& S) {
    return S.nTlas <= kTopLdsTlas && S.nInst <= kTopLdsInst;
}
// Copy of the top level into LDS by the whole workgroup (then a barrier).
__device__ inline void loadTopLds(const DevScene& S, TopLds& t) {
    const uint32_t nt = S.nTlas * (uint32_t)(sizeof(TlasNode) / 4), ni = S.nInst * (uint32_t)(sizeof(DevInstance) / 4);
    const uint32_t* gt = reinterpret_cast<const uint32_t*>(S.tlas);
    const uint32_t* gi = reinterpret_cast<const uint32_t*>(S.inst);
    uint32_t* lt = reinterpret_cast<uint32_t*>(t.tlas);
    uint32_t* li = reinterpret_cast<uint32_t*>(t.inst);
    for (uint32_t k = threadIdx.x; k < nt; k += blockDim.x) lt[k] = gt[k];
    for (uint32_t k = threadIdx.x; k < ni; k += blockDim.x) li[k] = gi[k];
    __syncthreads();
}

template <bool kCount, int kLeafWords = 4>
__device__ bool intersectScene(const DevScene& S, F3 o, F3 d, const LaneStack& st, Hit& h, Work& w);

template <bool kCount, int kLeafWords, typename Top>
__device__ bool intersectSceneTop(const DevScene& S, const Top& top, F3 o, F3 d, const LaneStack& st, Hit& h,
                                  Work& w) {
    h.inst = kInvalid;
    if (S.nTlas == 0) return false;
    const F3 inv = safeInvDir(d);
    float bestW = __builtin_huge_valf();
    bool found = false;
    uint32_t sp = 0;
    st.push(sp, 0u, 0.0f);
    while (sp > 0) {
        const uint32_t ni = st.get(--sp).x;
        const TlasNode nd = top.tlas[ni];
        if (kCount) {
            w.pops++;
            w.aabb++;
        }
        float te;
        if (!slab(o, inv, nd.mnx, nd.mny, nd.mnz, nd.mxx, nd.mxy, nd.mxz, bestW, 0.0f, te)) continue;
        if (nd.count == 0) {
            const TlasNode L = top.tlas[nd.left], R = top.tlas[nd.right];
            if (kCount) w.aabb += 2;
            float tl = __builtin_huge_valf(), tr = __builtin_huge_valf();
            const bool hl = slab(o, inv, L.mnx, L.mny, L.mnz, L.mxx, L.mxy, L.mxz, bestW, 0.0f, tl);
            const bool hr = slab(o, inv, R.mnx, R.mny, R.mnz, R.mxx, R.mxy, R.mxz, bestW, 0.0f, tr);
            if (hl && hr) {
                if (tl <= tr) {
                    st.push(sp, nd.right, 0.0f);
                    st.push(sp, nd.left, 0.0f);
                } else {
                    st.push(sp, nd.left, 0.0f);
                    st.push(sp, nd.right, 0.0f);
                }
            } else if (hl) {
                st.push(sp, nd.left, 0.0f);
            } else if (hr) {
                st.push(sp, nd.right, 0.0f);
            }
            continue;
        }
        const uint32_t ii = nd.left;
        const DevInstance& I = top.inst[ii];
        if (I.empty) continue;
        // toObjectSpace (KernelHelpers.h:217-227)
        float hv[4];
        mulv4(I.w2o, o.x, o.y, o.z, 1.0f, hv);
        const F3 oo = divs(mk3(hv[0], hv[1], hv[2]), hv[3]);
        mulv4(I.w2o, d.x, d.y, d.z, 0.0f, hv);
        const F3 od = normalize(mk3(hv[0], hv[1], hv[2]));
        LocalHit lh;
        bool acc;
        if (I.type == kGeoMesh)
            acc = traverseMesh<kCount>(S, I, oo, od, st, sp, lh, w);
        else if (kCount)
            acc = traverseCloud<kCount>(S, I, oo, od, st, sp, lh, w);
        else
            acc = traverseCloudFast<kLeafWords>(S, I, oo, od, st, sp, lh);
        if (acc) {
            const float tw = dot(lh.pos - o, d);
            if (tw > 0.0f && tw < bestW) {
                bestW = tw;
                found = true;
                h.t = tw;
                h.pos = lh.pos;
                h.inst = ii;
                h.prim = lh.prim;
                h.alpha = lh.alpha;
            }
        }
    }
    return found;
}

template <bool kCount, int kLeafWords>
__device__ bool intersectScene(const DevScene& S, F3 o, F3 d, const LaneStack& st, Hit& h, Work& w) {
    return intersectSceneTop<kCount, kLeafWords>(S, TopGlobal{S.tlas, S.inst}, o, d, st, h, w);
}

// Gradient of alpha_geom (intersectSurfel, KernelHelpers.h:637-669) with
// respect to the surfel's center, raw tangents, scales and beta, for the
// object-space ray (o, d): a reverse pass through the test's expressions
// (orthonormal frame from normalize(t_u) and Gram-Schmidt t_v, plane hit,
// (u, v) from the raw tangents, alpha = (1 - r^2)^(4 e^beta)).  g[0..2]
// center, g[3..5] t_u, g[6..8] t_v, g[9] s_u, g[10] s_v, g[11] beta.  False
// if the ray misses the support.
__device__ bool alphaGradient(const SurfelGeo& s, F3 o, F3 d, float g[12]) {
    const F3 p = mk3(s.px, s.py, s.pz), tu = mk3(s.tux, s.tuy, s.tuz), tv = mk3(s.tvx, s.tvy, s.tvz);
    const float lu = sqrtf(dot(tu, tu));
    const F3 uh = tu * (1.0f / lu);
    const float sdot = dot(uh, tv);
    const F3 wv = tv - uh * sdot;
    const float lw = sqrtf(dot(wv, wv));
    const F3 vh = wv * (1.0f / lw);
    const F3 m = cross(uh, vh);
    const float lm = sqrtf(dot(m, m));
    const F3 n = m * (1.0f / lm);
    const float nd = dot(n, d);
    if (fabsf(nd) < 1e-6f) return false;
    const F3 po = p - o;
    const float t = dot(n, po) / nd;
    const F3 r = (o + d * t) - p;
    const float U = dot(tu, r) / s.su, V = dot(tv, r) / s.sv;
    const float R = U * U + V * V;
    if (!(R < 1.0f)) return false;
    const float k = s.e4b;
    const float a = pow_r(1.0f - R, k);
    const float dadR = -k * a / (1.0f - R);
    const float gU = dadR * 2.0f * U, gV = dadR * 2.0f * V;
    const F3 gr = tu * (gU / s.su) + tv * (gV / s.sv);
    const float gt = dot(gr, d);
    const float gnp = gt / nd, gnd = -gt * t / nd;
    const F3 gp = n * gnp - gr;
    const F3 gn = po * gnp + d * gnd;
    const F3 gm = (gn - n * dot(n, gn)) * (1.0f / lm);
    F3 guh = cross(vh, gm);
    const F3 gvh = cross(gm, uh);
    const F3 gw = (gvh - vh * dot(vh, gvh)) * (1.0f / lw);
    F3 gtv = r * (gV / s.sv) + (gw - uh * dot(uh, gw));
    guh = guh - gw * sdot - tv * dot(uh, gw);
    const F3 gtu = r * (gU / s.su) + (guh - uh * dot(uh, guh)) * (1.0f / lu);
    g[0] = gp.x; g[1] = gp.y; g[2] = gp.z;
    g[3] = gtu.x; g[4] = gtu.y; g[5] = gtu.z;
    g[6] = gtv.x; g[7] = gtv.y; g[8] = gtv.z;
    g[9] = -gU * U / s.su;
    g[10] = -gV * V / s.sv;
    g[11] = a * (float)log((double)(1.0f - R)) * k;
    return true;
}

// buildIntersectionNormal (KernelHelpers.h:480-503)
__device__ __forceinline__ F3 hitNormal(const DevScene& S, const Hit& h) {
    const DevInstance& I = S.inst[h.inst];
    if (I.type == kGeoMesh) {
        const TriGeo tg = S.tris[h.prim];
        const F3 p0 = toWorldPoint(I.o2w, mk3(tg.v0x, tg.v0y, tg.v0z));
        const F3 p1 = toWorldPoint(I.o2w, mk3(tg.v1x, tg.v1y, tg.v1z));
        const F3 p2 = toWorldPoint(I.o2w, mk3(tg.v2x, tg.v2y, tg.v2z));
        return normalize(cross(p1 - p0, p2 - p0));
    }
    const SurfelShade sh = S.shade[h.prim];
    return mk3(sh.nsx, sh.nsy, sh.nsz);
}

// sampleCosineHemisphere(rng, n, dir, pdf) (KernelHelpers.h:317-337)
__device__ __forceinline__ void sampleCosine(Rng& rng, F3 n, F3& dir, float& pdf) {
    const float u1 = rng.nextFloat();
    const float u2 = rng.nextFloat();
    const float z = sqrtf(1.0f - u1);
    const float r = sqrtf(1.0f - z * z);
    const float phi = 2.0f * kPi * u2;
    const float x = r * cos_r(phi);
    const float y = r * sin_r(phi);
    const F3 up = fabsf(n.z) < 0.999f ? mk3(0, 0, 1) : mk3(1, 0, 0);
    const F3 tang = normalize(cross(up, n));
    const F3 bit = cross(n, tang);
    dir = normalize(tang * x + bit * y + n * z);
    pdf = fmaxf(0.0f, dot(dir, n)) / kPi;
}

// sampleUniformHemisphereAroundNormal (KernelHelpers.h:551-596)
__device__ __forceinline__ void sampleUniformHemi(Rng& rng, F3 normal, F3& dir, float& pdf) {
    const F3 n = normalize(normal);
    const float u1 = rng.nextFloat();
    const float u2 = rng.nextFloat();
    const float z = u1;
    const float phi = 2.0f * kPi * u2;
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    const float x = r * cos_r(phi);
    const float y = r * sin_r(phi);
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    const F3 t = mk3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    const F3 bt = mk3(b, sign + n.y * n.y * a, -n.y);
    dir = normalize(t * x + bt * y + n * z);
    pdf = 1.0f / (2.0f * kPi);
}

// applyRussianRoulette (KernelHelpers.h:1612-1657)
__device__ __forceinline__ bool russianRoulette(Rng& rng, uint32_t bounce, F3& T, uint32_t start) {
    if (bounce < start) return true;
    if (!isfinite(T.x) || !isfinite(T.y) || !isfinite(T.z)) return false;
    const float lum = 0.2126f * T.x + 0.7152f * T.y + 0.0722f * T.z;
    const float p = fminf(fmaxf(lum, 0.0f), 0.99f);
    if (rng.nextFloat() > p) return false;
    float c = 1.0f / p;
    c = fminf(c, 10.0f);
    T = T * c;
    return true;
}

// sampleMeshAreaLight (KernelHelpers.h:340-449)
__device__ bool sampleLight(const DevScene& S, Rng& rng, F3& pos, F3& dir, F3& power) {
    const uint32_t lc = S.nLights;
    if (lc == 0) return false;
    const float ul = rng.nextFloat();
    const uint32_t li = min((uint32_t)(ul * (float)lc), lc - 1u);
    const DevLight L = S.lights[li];
    if (L.triCount == 0u || L.area <= 0.0f) return false;
    const float ut = rng.nextFloat();
    uint32_t lo = 0u, hi = L.triCount - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1u;
        if (ut <= S.emTris[L.triOffset + mid].cdf) hi = mid;
        else lo = mid + 1u;
    }
    const TriGeo tg = S.tris[S.emTris[L.triOffset + lo].tri];
    const F3 v0 = mk3(tg.v0x, tg.v0y, tg.v0z), v1 = mk3(tg.v1x, tg.v1y, tg.v1z), v2 = mk3(tg.v2x, tg.v2y, tg.v2z);
    const float u1 = rng.nextFloat();
    const float u2 = rng.nextFloat();
    const float s1 = sqrtf(u1);
    const float b0 = 1.0f - s1, b1 = s1 * (1.0f - u2), b2 = s1 * u2;
    const F3 xo = v0 * b0 + v1 * b1 + v2 * b2;
    const float* o2w = S.inst[L.instance].o2w;
    const F3 p0 = toWorldPoint(o2w, v0), p1 = toWorldPoint(o2w, v1), p2 = toWorldPoint(o2w, v2);
    const F3 e0 = p1 - p0, e1 = p2 - p0;
    F3 n = mk3(e0.y * e1.z - e0.z * e1.y, e0.z * e1.x - e0.x * e1.z, e0.x * e1.y - e0.y * e1.x);
    const float nl = sqrtf(dot(n, n));
    if (nl <= 0.0f) return false;
    n = divs(n, nl);
    float pdf;
    sampleCosine(rng, n, dir, pdf);
    pos = toWorldPoint(o2w, xo);
    power = mk3(L.cr, L.cg, L.cb) * L.power;
    return true;
}

// worldToCellClamped (KernelHelpers.h:920-934)
__device__ __forceinline__ void worldToCell(const DevGrid& g, F3 p, int& x, int& y, int& z) {
    const F3 f = divv(p - mk3(g.ox, g.oy, g.oz), mk3(g.cx, g.cy, g.cz));
    x = min(max((int)floorf(f.x), 0), g.rx - 1);
    y = min(max((int)floorf(f.y), 0), g.ry - 1);
    z = min(max((int)floorf(f.z), 0), g.rz - 1);
}

// Cells of one grid row (cy, cz) that can hold a photon within radius r of
// q: the row is skipped when its y-z extent lies farther than r from q, and
// otherwise the x range shrinks to the chord of the sphere.  Every bound is
// widened by a margin (1e-3 of a cell plus 2e-6 of the coordinates) that
// exceeds the binary32 rounding of worldToCell and of the distance test by
// orders of magnitude, and boundary cells, which also hold the photons that
// worldToCell clamps into them, extend to infinity.  A photon in a skipped
// cell therefore fails the reference's d2 > r2 test, so skipping it leaves
// the sum unchanged.
__device__ __forceinline__ bool rowSpan(const DevGrid& g, F3 q, float r2, int cy, int cz, int lx, int hx, int& ax,
                                        int& bx) {
    auto axisGap = [](float qc, float o, float c, int i, int res, float m) {
        const float lo = (i == 0) ? -INFINITY : o + (float)i * c - m;
        const float hi = (i == res - 1) ? INFINITY : o + (float)(i + 1) * c + m;
        return fmaxf(fmaxf(lo - qc, qc - hi), 0.0f);
    };
    const float my = 1e-3f * g.cy + 2e-6f * (fabsf(q.y) + fabsf(g.oy) + g.radius);
    const float mz = 1e-3f * g.cz + 2e-6f * (fabsf(q.z) + fabsf(g.oz) + g.radius);
    const float dy = axisGap(q.y, g.oy, g.cy, cy, g.ry, my);
    const float dz = axisGap(q.z, g.oz, g.cz, cz, g.rz, mz);
    const float room = r2 * 1.001f - (dy * dy + dz * dz);
    if (!(room >= 0.0f)) return false;
    const float hw = sqrtf(room);
    const float mx = 1e-3f * g.cx + 2e-6f * (fabsf(q.x) + fabsf(g.ox) + g.radius);
    const float fa = floorf((q.x - hw - mx - g.ox) / g.cx);
    const float fb = floorf((q.x + hw + mx - g.ox) / g.cx);
    ax = (fa > (float)lx) ? (int)fa : lx;
    bx = (fb < (float)hx) ? (int)fb : hx;
    return ax <= bx;
}

typedef __attribute__((address_space(4))) const float4 ConstF4;
typedef __attribute__((address_space(4))) const float ConstF;
typedef __attribute__((address_space(4))) const uint32_t ConstU32;

__device__ __forceinline__ float waveMinF(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ float waveMaxF(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

// rowSpan for every point of the box [qlo, qhi]: the row is skipped only
// when its extent lies farther than r from the whole box, and the x range
// covers the chords of all its points (every bound and margin is monotone in
// the point, so the span contains each point's own span).
__device__ __forceinline__ bool rowSpanBox(const DevGrid& g, F3 qlo, F3 qhi, float r2, int cy, int cz, int lx,
                                           int hx, int& ax, int& bx) {
    auto axisGap = [](float ql, float qh, float o, float c, int i, int res, float m) {
        const float lo = (i == 0) ? -INFINITY : o + (float)i * c - m;
        const float hi = (i == res - 1) ? INFINITY : o + (float)(i + 1) * c + m;
        return fmaxf(fmaxf(lo - qh, ql - hi), 0.0f);
    };
    const float my = 1e-3f * g.cy + 2e-6f * (fmaxf(fabsf(qlo.y), fabsf(qhi.y)) + fabsf(g.oy) + g.radius);
    const float mz = 1e-3f * g.cz + 2e-6f * (fmaxf(fabsf(qlo.z), fabsf(qhi.z)) + fabsf(g.oz) + g.radius);
    const float dy = axisGap(qlo.y, qhi.y, g.oy, g.cy, cy, g.ry, my);
    const float dz = axisGap(qlo.z, qhi.z, g.oz, g.cz, cz, g.rz, mz);
    const float room = r2 * 1.001f - (dy * dy + dz * dz);
    if (!(room >= 0.0f)) return false;
    const float hw = sqrtf(room);
    const float mx = 1e-3f * g.cx + 2e-6f * (fmaxf(fabsf(qlo.x), fabsf(qhi.x)) + fabsf(g.ox) + g.radius);
    const float fa = floorf((qlo.x - hw - mx - g.ox) / g.cx);
    const float fb = floorf((qhi.x + hw + mx - g.ox) / g.cx);
    ax = (fa > (float)lx) ? (int)fa : lx;
    bx = (fb < (float)hx) ? (int)fb : hx;
    return ax <= bx;
}

// The gather of a full wave whose points share one cell box and lie close
// together (qlo..qhi, within kWaveSpread of the radius on every axis): the
// rows and photons are the same for every lane, so they are walked once for
// the wave with scalar loads, and each lane applies its own tests to every
// photon of the rows' spans for the whole wave.  A lane adds exactly the
// photons of its own box that pass its tests, in the same order as the
// per-lane walk, so the sums are the same to the bit.
constexpr float kWaveSpread = 0.5f;
__device__ F3 gatherWave(const DevGrid& g, F3 q, F3 n, F3 qlo, F3 qhi, float r2, float invArea, int lx, int ly,
                         int lz, int hx, int hy, int hz) {
    F3 E = mk3(0, 0, 0);
    auto addPhoton = [&](const float4& pp, const float4& dp, float pz) {
        const F3 dd = mk3(pp.x, pp.y, pp.z) - q;
        const float cosine = dot(n, -mk3(dp.x, dp.y, dp.z));
        if (cosine < 0.0f) return;
        const float d2 = dot(dd, dd);
        if (d2 > r2) return;
        E = E + mk3(pp.w, dp.w, pz) * invArea;
    };
    const ConstF4* pos = (const ConstF4*)g.posPow;
    const ConstF4* dir = (const ConstF4*)g.dirPow;
    const ConstF* pzs = (const ConstF*)g.powZ;
    const ConstU32* cs = (const ConstU32*)g.cellStart;
    for (int cz = lz; cz <= hz; ++cz) {
        for (int cy = ly; cy <= hy; ++cy) {
            int ax, bx;
            if (!rowSpanBox(g, qlo, qhi, r2, cy, cz, lx, hx, ax, bx)) continue;
            ax = __builtin_amdgcn_readfirstlane(ax);
            bx = __builtin_amdgcn_readfirstlane(bx);
            const uint32_t row = (uint32_t)((cz * g.ry + cy) * g.rx);
            __asm__ volatile("");
            uint32_t j = cs[row + ax];
            const uint32_t b = cs[row + bx + 1];
            for (; j + 4 <= b; j += 4) {
                const float4 p0 = make_float4(pos[j].x, pos[j].y, pos[j].z, pos[j].w);
                const float4 p1 = make_float4(pos[j + 1].x, pos[j + 1].y, pos[j + 1].z, pos[j + 1].w);
                const float4 p2 = make_float4(pos[j + 2].x, pos[j + 2].y, pos[j + 2].z, pos[j + 2].w);
                const float4 p3 = make_float4(pos[j + 3].x, pos[j + 3].y, pos[j + 3].z, pos[j + 3].w);
                const float4 d0 = make_float4(dir[j].x, dir[j].y, dir[j].z, dir[j].w);
                const float4 d1 = make_float4(dir[j + 1].x, dir[j + 1].y, dir[j + 1].z, dir[j + 1].w);
                const float4 d2 = make_float4(dir[j + 2].x, dir[j + 2].y, dir[j + 2].z, dir[j + 2].w);
                const float4 d3 = make_float4(dir[j + 3].x, dir[j + 3].y, dir[j + 3].z, dir[j + 3].w);
                const float z0 = pzs[j], z1 = pzs[j + 1], z2 = pzs[j + 2], z3 = pzs[j + 3];
                addPhoton(p0, d0, z0);
                addPhoton(p1, d1, z1);
                addPhoton(p2, d2, z2);
                addPhoton(p3, d3, z3);
            }
            for (; j < b; ++j)
                addPhoton(make_float4(pos[j].x, pos[j].y, pos[j].z, pos[j].w),
                          make_float4(dir[j].x, dir[j].y, dir[j].z, dir[j].w), pzs[j]);
        }
    }
    return E;
}

// gatherDiffuseIrradianceAtPoint (KernelHelpers.h:990-1037).  Photons are
// stored sorted by (cell, deposit index), so the cells of one grid row form
// one contiguous range; summation order equals the reference's cell-by-cell,
// photon-by-photon order.  Cells that cannot hold an accepted photon are not
// scanned (rowSpan).  Memory-level parallelism: the row bounds of a whole
// z-slice are loaded before any photon of that slice, and photons are loaded
// four at a time (added in order).  Counting mode counts the reference's
// cells and photons (the whole box), not the scanned ones.  kWave (the
// camera gather, whose waves hold cell-sorted points, with shared seeds
// often the same point several times): a wave whose points share a cell box
// and lie close together gathers through gatherWave.
constexpr int kSliceRows = 6;  // cells per axis: 4 + 1, +1 for rounding

template <bool kCount, bool kWave = false>
__device__ F3 gatherIrradiance(const DevGrid& g, F3 q, F3 n, Work& w) {
    if (kCount) w.gathers++;
    if (!g.valid) return mk3(0, 0, 0);
    const float r = g.radius;
    const float r2 = r * r;
    const float invArea = 1.0f / (kPi * r2);
    int lx, ly, lz, hx, hy, hz;
    worldToCell(g, q - mk3(r, r, r), lx, ly, lz);
    worldToCell(g, q + mk3(r, r, r), hx, hy, hz);
    F3 E = mk3(0, 0, 0);
    bool waveDone = false;
    if (kWave && !kCount && __builtin_amdgcn_ballot_w64(true) == ~0ull) {
        const int ulx = __builtin_amdgcn_readfirstlane(lx), uly = __builtin_amdgcn_readfirstlane(ly),
                  ulz = __builtin_amdgcn_readfirstlane(lz), uhx = __builtin_amdgcn_readfirstlane(hx),
                  uhy = __builtin_amdgcn_readfirstlane(hy), uhz = __builtin_amdgcn_readfirstlane(hz);
        if (__builtin_amdgcn_ballot_w64(lx != ulx || ly != uly || lz != ulz || hx != uhx || hy != uhy || hz != uhz) ==
            0ull) {
            const F3 qlo = mk3(waveMinF(q.x), waveMinF(q.y), waveMinF(q.z));
            const F3 qhi = mk3(waveMaxF(q.x), waveMaxF(q.y), waveMaxF(q.z));
            const F3 ext = qhi - qlo;
            if (fmaxf(fmaxf(ext.x, ext.y), ext.z) <= kWaveSpread * r) {
                E = gatherWave(g, q, n, qlo, qhi, r2, invArea, ulx, uly, ulz, uhx, uhy, uhz);
                waveDone = true;
            }
        }
    }
    if (!waveDone) {
    auto addPhoton = [&](const float4& pp, const float4& dp, float pz) {
        const F3 dd = mk3(pp.x, pp.y, pp.z) - q;
        // dot(n, -dir) == dot(-n, dir) exactly (negation is exact).
        const float cosine = dot(n, -mk3(dp.x, dp.y, dp.z));
        if (cosine < 0.0f) return;
        const float d2 = dot(dd, dd);
        if (d2 > r2) return;
        E = E + mk3(pp.w, dp.w, pz) * invArea;
    };
    for (int cz = lz; cz <= hz; ++cz) {
        for (int cy0 = ly; cy0 <= hy; cy0 += kSliceRows) {
            uint32_t ra[kSliceRows], rb[kSliceRows];
#pragma unroll
            for (int k = 0; k < kSliceRows; ++k) {
                const int cy = cy0 + k;
                ra[k] = rb[k] = 0u;
                int ax, bx;
                if (cy <= hy && rowSpan(g, q, r2, cy, cz, lx, hx, ax, bx)) {
                    const uint32_t row = (uint32_t)((cz * g.ry + cy) * g.rx);
                    ra[k] = g.cellStart[row + ax];
                    rb[k] = g.cellStart[row + bx + 1];
                }
            }
#pragma unroll
            for (int k = 0; k < kSliceRows; ++k) {
                if (kCount && cy0 + k <= hy) {
                    const uint32_t row = (uint32_t)((cz * g.ry + cy0 + k) * g.rx);
                    w.cells += (uint32_t)(hx - lx + 1);
                    w.photons += g.cellStart[row + hx + 1] - g.cellStart[row + lx];
                    for (int cx = lx; cx <= hx; ++cx) w.cellsNE += g.cellStart[row + cx + 1] > g.cellStart[row + cx];
                }
                uint32_t j = ra[k];
                const uint32_t b = rb[k];
                for (; j + 4 <= b; j += 4) {
                    const float4 p0 = g.posPow[j], p1 = g.posPow[j + 1], p2 = g.posPow[j + 2], p3 = g.posPow[j + 3];
                    const float4 d0 = g.dirPow[j], d1 = g.dirPow[j + 1], d2 = g.dirPow[j + 2], d3 = g.dirPow[j + 3];
                    const float z0 = g.powZ[j], z1 = g.powZ[j + 1], z2 = g.powZ[j + 2], z3 = g.powZ[j + 3];
                    addPhoton(p0, d0, z0);
                    addPhoton(p1, d1, z1);
                    addPhoton(p2, d2, z2);
                    addPhoton(p3, d3, z3);
                }
                for (; j < b; ++j) addPhoton(g.posPow[j], g.dirPow[j], g.powZ[j]);
            }
        }
    }
    }
    const uint32_t T = *g.attempted;
    const uint32_t D = min(T, g.capacity);
    if (D == 0u) return mk3(0, 0, 0);
    return E * ((float)T / (float)D);
}

// makePrimaryRayFromPixelJitteredFov (KernelHelpers.h:764-796)
struct DevCamera {
    float invView[16];
    uint32_t W, H;
    float fx, fy;
};

__device__ __forceinline__ void primaryRay(const DevCamera& c, float px, float py, float jx, float jy, F3& o,
                                           F3& d) {
    const float W = (float)c.W, H = (float)c.H;
    const float u = px + 0.5f + jx;
    const float v = py + 0.5f + jy;
    const float vf = H - v;
    const float ndcX = 2.0f * u / W - 1.0f;
    const float ndcY = 2.0f * vf / H - 1.0f;
    const F3 dc = normalize(mk3(ndcX * (0.5f * W) / c.fx, ndcY * (0.5f * H) / c.fy, -1.0f));
    d = transformDirection(c.invView, dc);
    o = transformPoint(c.invView, mk3(0, 0, 0));
}

}  // namespace pale
