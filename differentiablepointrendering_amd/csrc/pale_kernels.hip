// pale_kernels.hip -- gfx950 kernels of the beta-surfel photon-mapping
// renderer and its opacity adjoint.
//
// Wavefront structure: every bounce is one traversal kernel over a ray queue
// followed by a stable (order-preserving) compaction of its outputs.  Queue
// slot = rank in the previous queue, so each ray meets the per-item random
// stream the reference's sequential order gives it (the reference seeds every
// work item from its queue slot, PrimalKernels.cpp:80-82,
// AdjointKernels.cpp:99-101).  Counts stay on the device; no kernel waits for
// the host.
#include <mutex>
#include <algorithm>
#include <cstddef>

#include "pale_kernels.h"

namespace pale {

namespace {

constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kTileSpan = kTile / kScanThreads;  // items per thread

enum : uint32_t { kNone = 0, kNullT = 1, kReflectS = 2, kTransmitS = 3, kProjection = 4 };

__device__ __forceinline__ uint32_t laneId() { return __lane_id(); }

// Wave-level count of a predicate, one atomic per wave.
__device__ __forceinline__ void waveCount(bool pred, unsigned long long* counter) {
    const unsigned long long m = __ballot(pred);
    if (laneId() == __ffsll((long long)__ballot(1)) - 1 && m) atomicAdd(counter, (unsigned long long)__popcll(m));
}

__device__ __forceinline__ uint32_t waveSum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Flush a thread's work counters: one atomic per counter per wave.  Every
// lane of the wave must call this.
__device__ __forceinline__ void flushWork(const Work& w, DevState* st, int cls) {
    const uint32_t v[7] = {waveSum(w.pops), waveSum(w.aabb), waveSum(w.prims), waveSum(w.gathers),
                           waveSum(w.cells), waveSum(w.cellsNE), waveSum(w.photons)};
    if (__lane_id() == 0) {
#pragma unroll
        for (int k = 0; k < 7; ++k)
            if (v[k]) atomicAdd(&st->work[cls][k], (unsigned long long)v[k]);
    }
}

// Exclusive scan over a 256-thread block; `total` receives the block sum.
__device__ __forceinline__ uint32_t blockExclusiveScan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t waveTotals[kScanThreads / 64 + 1];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63u) waveTotals[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (uint32_t w = 0; w < kScanThreads / 64; ++w) {
            const uint32_t t = waveTotals[w];
            waveTotals[w] = s;
            s += t;
        }
        waveTotals[kScanThreads / 64] = s;
    }
    __syncthreads();
    const uint32_t excl = x - v + waveTotals[wave];
    total = waveTotals[kScanThreads / 64];
    __syncthreads();
    return excl;
}

__device__ __forceinline__ uint32_t loadCount(const uint32_t* count, uint32_t maxItems) {
    const uint32_t n = *count;
    return n < maxItems ? n : maxItems;
}

__device__ __forceinline__ LaneStack laneStack(uint2* lds, uint32_t depth) {
    return LaneStack{(LdsUint2*)(lds + threadIdx.x), blockDim.x, depth, nullptr, 0u, depth};
}

// The layer walk's stack: its first inLds entries in LDS, the rest in the
// spill area (entry k of lane l at spill[(k - inLds) * lanes + l]), so that
// more waves fit on a CU; spill == nullptr: all in LDS.
__device__ __forceinline__ LaneStack walkStack(uint2* lds, uint32_t depth, uint2* spill, uint32_t inLds) {
    if (!spill) return laneStack(lds, depth);
    const uint32_t lanes = gridDim.x * blockDim.x;
    return LaneStack{(LdsUint2*)(lds + threadIdx.x), blockDim.x, depth, spill + blockIdx.x * blockDim.x + threadIdx.x,
                     lanes, inLds};
}

__device__ __forceinline__ float4 f4(F3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }

}  // namespace

// ---------------------------------------------------------------- small --
__global__ void k_fill(uint32_t* p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fillf(float* p, uint32_t n, float v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_iota(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = i;
}

// Seed schedule of submitPhotonMappingKernel (SyclBridge.cpp:83-143): one
// draw per emitter launch, per executed bounce and per camera.  `bounce` < 0
// always draws; a bounce draws only if its queue is not empty.
__global__ void k_seed_next(DevState* st, const uint32_t* table, uint32_t len, int bounce) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (bounce >= 0 && st->qCount == 0) {
        st->live = 0;
        return;
    }
    const uint32_t c = st->cursor;
    st->number = c < len ? table[c] : 0u;
    st->cursor = c + 1;
    st->live = 1;
}

// Start of a served view: its passes draw from `number` (+ pass); the
// camera-gather counters restart, the photon map's stay.
__global__ void k_begin_view(DevState* st, uint32_t number) {
    if (blockIdx.x != 0) return;
    const uint32_t t = threadIdx.x;
    if (t < 8) st->work[kWorkGather][t] = 0ull;
    if (t == 0) {
        st->number = number;
        st->gatherQ = 0ull;
        st->entries = 0u;
        st->overflow = 0u;
    }
}

// photonCount counts deposit attempts (the reference's counter T); only the
// first `cap` are stored.
__global__ void k_finish_photon_bounce(DevState* st, uint32_t cap) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    (void)cap;
    if (st->live) {
        st->qCount = st->extTotal;
        st->photonCount += st->depTotal;
    } else {
        st->qCount = 0;
    }
}

// Events of one bounce were appended to the pool.
__global__ void k_event_pool_add(DevState* st, uint32_t cap) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const uint32_t n = st->evPool + st->evTotal;
    st->evPool = n < cap ? n : cap;
}

__global__ void k_set_queue(DevState* st, const uint32_t* src) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->qCount = *src;
}

// ---------------------------------------------------------- emission --
// launchRayGenEmitterKernel (PrimalKernels.cpp:13-62)
__global__ __launch_bounds__(256) void k_emit(DevScene S, const DevState* st, uint32_t P, float invEmitted,
                                              RayQueue out, uint8_t* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    Rng rng(perItemSeed(st->number, i));
    F3 pos, dir, power;
    const bool ok = sampleLight(S, rng, pos, dir, power);
    flags[i] = ok ? 1 : 0;
    if (!ok) return;
    const F3 T = (power * (float)S.nLights) * invEmitted;
    out.o[i] = f4(pos, u2f(kInvalid));
    out.d[i] = f4(dir, u2f(0u));
    out.T[i] = f4(T, 0.0f);
}

// ---------------------------------------------------- photon bounce --
// launchIntersectKernel (PrimalKernels.cpp:65-321) for one queue.
template <bool kCount>
__global__ __launch_bounds__(kTraceBlock) void k_photon_trace(DevScene S, DevState* st, RayQueue in,
                                                              RayQueue out, uint8_t* extFlags, PhotonArrays dep,
                                                              uint8_t* depFlags, TraceParams tp, Reservoir rs) {
    extern __shared__ uint2 s_stack[];
    const LaneStack stk = walkStack(s_stack, S.stackDepth, tp.spill, tp.inLds);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = st->live ? st->qCount : 0u;
    const bool active = i < n;
    const bool resv = st->resv != 0u;
    Work w{};
    if (active) {
        uint8_t ef = 0, df = 0;
        Rng rng(perItemSeed(st->number, i));
        // depositPhotonSurface past capacity (KernelHelpers.h:1705-1718): the
        // attempt index comes from the probe pass; accept with probability
        // cap / (attempt + 1), then pick the slot (two draws of this item).
        auto reservoirDraws = [&]() {
            if (!resv) return;
            const uint32_t a = rs.attempt[i];
            uint32_t sl = kNoSlot;
            if (a >= rs.cap) {
                const float u = rng.nextFloat();
                const float accept = (float)rs.cap / (float)(a + 1u);
                if (u < accept) sl = rng.nextUint() % rs.cap;
            }
            rs.slot[i] = sl;
        };
        if (resv) rs.slot[i] = kNoSlot;
        const float4 o4 = in.o[i], d4 = in.d[i], T4 = in.T[i];
        const F3 o = ld3(o4), d = ld3(d4), T = ld3(T4);
        const uint32_t bounce = f2u(d4.w);
        Hit h;
        if (intersectScene<kCount>(S, o, d, stk, h, w)) {
            const DevInstance& I = S.inst[h.inst];
            F3 no = mk3(0, 0, 0), nT = mk3(0, 0, 0), ndir = d;
            bool spawn = false;
            if (I.type == kGeoMesh) {
                const F3 ng = hitNormal(S, h);
                df = 1;
                dep.posPow[i] = f4(h.pos, T.x);
                dep.dirPow[i] = f4(d, T.y);
                dep.powZ[i] = T.z;
                reservoirDraws();
                const DevMaterial m = S.mats[I.material];
                float pdf;
                sampleCosine(rng, ng, ndir, pdf);
                no = h.pos + ng * 1e-6f;
                nT = T * mk3(m.r, m.g, m.b);
                spawn = true;
            } else {
                const float u = rng.nextFloat();
                const SurfelShade sh = S.shade[h.prim];
                if (u < tp.qNull) {
                    const float att = 1.0f - h.alpha * sh.opacity;
                    const float w = att / tp.qNull;
                    no = h.pos + d * 1e-5f;
                    nT = T * w;
                    spawn = true;
                } else if (u < tp.qNull + tp.qReflect) {
                    const F3 nc = mk3(sh.nsx, sh.nsy, sh.nsz);
                    const float sc = dot(nc, -d);
                    const int sign = sc >= 0.0f ? 1 : -1;
                    const F3 on = (float)sign * nc;
                    const F3 flux = divs(T, tp.qReflect);
                    df = 1;
                    dep.posPow[i] = f4(h.pos, flux.x);
                    dep.dirPow[i] = f4(d, flux.y);
                    dep.powZ[i] = flux.z;
                    reservoirDraws();
                    float pdf;
                    sampleUniformHemi(rng, on, ndir, pdf);
                    const float cosT = fmaxf(0.0f, dot(ndir, on));
                    const F3 fs = (sh.alpha_r * mk3(sh.ar, sh.ag, sh.ab)) * kInvPi;
                    const float alpha = h.alpha * sh.opacity;
                    const F3 w = divs(((alpha / tp.qReflect) * fs) * cosT, pdf);
                    no = h.pos + on * 1e-5f;
                    nT = T * w;
                    spawn = true;
                } else if (u < tp.qNull + tp.qReflect + tp.qTransmit) {
                    const F3 nc = mk3(sh.nsx, sh.nsy, sh.nsz);
                    const float sc = dot(nc, -d);
                    const int sign = sc >= 0.0f ? 1 : -1;
                    const F3 on = (float)sign * nc;
                    float pdf;
                    sampleUniformHemi(rng, on, ndir, pdf);
                    const float cosT = fmaxf(0.0f, dot(ndir, on));
                    const F3 fs = (0.0f * mk3(sh.ar, sh.ag, sh.ab)) * kInvPi;  // alpha_t = 0
                    const float alpha = h.alpha * sh.opacity;
                    const F3 w = divs(((alpha / tp.qReflect) * fs) * cosT, pdf);
                    no = h.pos + (-on) * 1e-5f;
                    nT = T * w;
                    spawn = true;
                }
            }
            if (spawn && russianRoulette(rng, bounce + 1, nT, tp.rrStart)) {
                ef = 1;
                out.o[i] = f4(no, o4.w);
                out.d[i] = f4(ndir, u2f(bounce + 1));
                out.T[i] = f4(nT, 0.0f);
            }
        }
        extFlags[i] = ef;
        depFlags[i] = df;
    }
    waveCount(active, &st->photonQ);
    if (kCount) flushWork(w, st, kWorkPhoton);
}

// Reservoir mode of one bounce: on when the deposits so far plus one per
// queued photon could pass the capacity.
__global__ void k_resv_check(DevState* st, uint32_t cap) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    st->resv = (st->live && (uint64_t)st->photonCount + (uint64_t)st->qCount > (uint64_t)cap) ? 1u : 0u;
}

// Probe pass of a reservoir bounce: which items deposit (the same hit and
// the same first draw as k_photon_trace), so that every deposit learns its
// attempt index in queue order before its random stream is used.
__global__ __launch_bounds__(kTraceBlock) void k_photon_probe(DevScene S, DevState* st, RayQueue in,
                                                              uint8_t* depFlags, TraceParams tp) {
    if (!st->resv) return;
    extern __shared__ uint2 s_stack[];
    const LaneStack stk = laneStack(s_stack, S.stackDepth);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = st->live ? st->qCount : 0u;
    if (i >= n) return;
    Rng rng(perItemSeed(st->number, i));
    Work w{};
    Hit h;
    uint8_t df = 0;
    if (intersectScene<false>(S, ld3(in.o[i]), ld3(in.d[i]), stk, h, w)) {
        if (S.inst[h.inst].type == kGeoMesh) {
            df = 1;
        } else {
            const float u = rng.nextFloat();
            df = (u >= tp.qNull && u < tp.qNull + tp.qReflect) ? 1 : 0;
        }
    }
    depFlags[i] = df;
}

// attempt[i] = photonCount + rank of item i among the depositing items.
__global__ __launch_bounds__(kScanThreads) void k_resv_rank(const uint8_t* flags, const DevState* st,
                                                            const uint32_t* tiles, uint32_t* attempt) {
    if (!st->resv) return;
    const uint32_t n = st->live ? st->qCount : 0u;
    const uint32_t first = blockIdx.x * kTile + threadIdx.x * kTileSpan;
    uint8_t f[kTileSpan];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kTileSpan; ++k) {
        f[k] = (first + k < n) ? flags[first + k] : 0;
        s += f[k];
    }
    uint32_t total;
    uint32_t off = blockExclusiveScan(s, total) + tiles[blockIdx.x] + st->photonCount;
#pragma unroll
    for (uint32_t k = 0; k < kTileSpan; ++k) {
        if (f[k]) attempt[first + k] = off++;
    }
}

__global__ __launch_bounds__(256) void k_resv_claim(const DevState* st, Reservoir rs) {
    if (!st->resv) return;
    const uint32_t n = st->live ? st->qCount : 0u;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t sl = rs.slot[i];
    if (sl != kNoSlot) atomicMax(&rs.owner[sl], rs.attempt[i]);
}

__global__ __launch_bounds__(256) void k_resv_write(const DevState* st, Reservoir rs, PhotonArrays dep,
                                                    PhotonArrays raw) {
    if (!st->resv) return;
    const uint32_t n = st->live ? st->qCount : 0u;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t sl = rs.slot[i];
    if (sl == kNoSlot || rs.owner[sl] != rs.attempt[i]) return;
    raw.posPow[sl] = dep.posPow[i];
    raw.dirPow[sl] = dep.dirPow[i];
    raw.powZ[sl] = dep.powZ[i];
}

// -------------------------------------------------- stable compaction --
__global__ __launch_bounds__(kScanThreads) void k_tile_count(const uint8_t* flags, const uint32_t* count,
                                                             uint32_t maxItems, uint32_t* tiles) {
    const uint32_t n = loadCount(count, maxItems);
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * kTileSpan;
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kTileSpan; ++k)
        if (base + k < n) s += flags[base + k];
    uint32_t total;
    blockExclusiveScan(s, total);
    if (threadIdx.x == 0) tiles[blockIdx.x] = total;
}

// Exclusive scan of the tile sums (one workgroup); *total = sum, clamped to
// `cap`.  The number of tiles is ceil(n / kTile) for n = *count (or
// `fixedN` if count is null).
__global__ __launch_bounds__(1024) void k_tile_scan(uint32_t* tiles, const uint32_t* count, uint32_t fixedN,
                                                    uint32_t maxItems, uint32_t* total, uint32_t cap) {
    __shared__ uint32_t sums[1024 / 64 + 1];
    const uint32_t n = count ? loadCount(count, maxItems) : fixedN;
    const uint32_t numTiles = (n + kTile - 1) / kTile;
    uint32_t carry = 0;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t base = 0; base < numTiles; base += 1024) {
        const uint32_t idx = base + threadIdx.x;
        const uint32_t v = idx < numTiles ? tiles[idx] : 0u;
        uint32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= (uint32_t)o) x += y;
        }
        if (lane == 63u) sums[wave] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t s = 0;
            for (uint32_t w = 0; w < 1024 / 64; ++w) {
                const uint32_t t = sums[w];
                sums[w] = s;
                s += t;
            }
            sums[1024 / 64] = s;
        }
        __syncthreads();
        if (idx < numTiles) tiles[idx] = carry + x - v + sums[wave];
        carry += sums[1024 / 64];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry < cap ? carry : cap;
}

template <typename Copy>
__device__ __forceinline__ void scatterTile(const uint8_t* flags, uint32_t n, const uint32_t* tiles, uint32_t base,
                                            uint32_t cap, Copy copy) {
    const uint32_t first = blockIdx.x * kTile + threadIdx.x * kTileSpan;
    uint8_t f[kTileSpan];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kTileSpan; ++k) {
        f[k] = (first + k < n) ? flags[first + k] : 0;
        s += f[k];
    }
    uint32_t total;
    uint32_t off = blockExclusiveScan(s, total) + tiles[blockIdx.x] + base;
#pragma unroll
    for (uint32_t k = 0; k < kTileSpan; ++k) {
        if (f[k]) {
            if (off < cap) copy(first + k, off);
            ++off;
        }
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scatter_rays(const uint8_t* flags, const uint32_t* count,
                                                               uint32_t maxItems, const uint32_t* tiles,
                                                               RayQueue src, RayQueue dst, uint32_t cap) {
    const uint32_t n = loadCount(count, maxItems);
    scatterTile(flags, n, tiles, 0u, cap, [&](uint32_t i, uint32_t j) {
        dst.o[j] = src.o[i];
        dst.d[j] = src.d[i];
        dst.T[j] = src.T[i];
    });
}

__global__ __launch_bounds__(kScanThreads) void k_scatter_photons(const uint8_t* flags, const uint32_t* count,
                                                                  uint32_t maxItems, const uint32_t* tiles,
                                                                  const uint32_t* base, PhotonArrays src,
                                                                  PhotonArrays dst, uint32_t cap) {
    const uint32_t n = loadCount(count, maxItems);
    scatterTile(flags, n, tiles, *base, cap, [&](uint32_t i, uint32_t j) {
        dst.posPow[j] = src.posPow[i];
        dst.dirPow[j] = src.dirPow[i];
        dst.powZ[j] = src.powZ[i];
    });
}

__global__ __launch_bounds__(kScanThreads) void k_scatter_events(const uint8_t* flags, const uint32_t* count,
                                                                 uint32_t maxItems, const uint32_t* tiles,
                                                                 const uint32_t* base, EventArrays src,
                                                                 EventArrays dst, uint32_t cap) {
    const uint32_t n = loadCount(count, maxItems);
    const bool rays = src.e4 != nullptr;
    scatterTile(flags, n, tiles, *base, cap, [&](uint32_t i, uint32_t j) {
        dst.e0[j] = src.e0[i];
        dst.e1[j] = src.e1[i];
        dst.e2[j] = src.e2[i];
        dst.e3[j] = src.e3[i];
        if (rays) {
            dst.e4[j] = src.e4[i];
            dst.e5[j] = src.e5[i];
        }
    });
}

// ------------------------------------------------------ photon grid --
// computePhotonCellIdsAndPermutation + countPhotonsPerCell
// (PrimalKernels.cpp:796-858).
__global__ __launch_bounds__(256) void k_cell_count(GridParams g, const DevState* st, uint32_t cap,
                                                    const float4* posPow, uint32_t* cellId, uint32_t* cellCount) {
    const uint32_t n = min(st->photonCount, cap);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = posPow[i];
    const F3 f = divv(mk3(p.x, p.y, p.z) - mk3(g.ox, g.oy, g.oz), mk3(g.cell, g.cell, g.cell));
    const int x = min(max((int)floorf(f.x), 0), g.rx - 1);
    const int y = min(max((int)floorf(f.y), 0), g.ry - 1);
    const int z = min(max((int)floorf(f.z), 0), g.rz - 1);
    const uint32_t c = (uint32_t)((z * g.ry + y) * g.rx + x);
    cellId[i] = c;
    atomicAdd(&cellCount[c], 1u);
}

// Exclusive scan, pass 1: per-tile scan of `in` into `out`, tile sums.
__global__ __launch_bounds__(kScanThreads) void k_scan_local(const uint32_t* in, uint32_t* out, uint32_t n,
                                                             uint32_t* tiles) {
    const uint32_t first = blockIdx.x * kTile + threadIdx.x * kTileSpan;
    uint32_t v[kTileSpan];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kTileSpan; ++k) {
        v[k] = (first + k < n) ? in[first + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t off = blockExclusiveScan(s, total);
#pragma unroll
    for (uint32_t k = 0; k < kTileSpan; ++k) {
        if (first + k < n) out[first + k] = off;
        off += v[k];
    }
    if (threadIdx.x == 0) tiles[blockIdx.x] = total;
}

// Pass 3: add the scanned tile offsets; out[n] = grand total.
__global__ __launch_bounds__(256) void k_scan_add(uint32_t* out, uint32_t n, const uint32_t* tiles,
                                                  const uint32_t* grand) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += tiles[i / kTile];
    if (i == 0) out[n] = *grand;
}

// scatterPhotonsIntoCells (PrimalKernels.cpp:861-893).  The decrement leaves
// cellCount zeroed for the next build.
__global__ __launch_bounds__(256) void k_cell_scatter(const DevState* st, uint32_t cap, const uint32_t* cellId,
                                                      const uint32_t* cellStart, uint32_t* cellCount,
                                                      uint32_t* sortedIdx) {
    const uint32_t n = min(st->photonCount, cap);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = cellId[i];
    const uint32_t slot = cellStart[c] + atomicSub(&cellCount[c], 1u) - 1u;
    if (slot < n) sortedIdx[slot] = i;
}

// Order each cell by deposit index: the reference's sequential scatter
// order, which makes every gather sum in the same order.  A wave takes 64
// consecutive cells and sorts each one of up to 64 photons together: lane j
// holds photon j of the cell and its place is the number of smaller indices
// (indices are distinct).  Larger cells go to a list for k_cell_sort_big.
constexpr uint32_t kWaveCell = 64;
constexpr uint32_t kBlockCell = 4096;

__global__ __launch_bounds__(256) void k_cell_sort(uint32_t total, const uint32_t* cellStart, uint32_t* sortedIdx,
                                                   DevState* st, uint32_t* bigList) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = laneId();
    uint32_t a = 0, n = 0;
    if (c < total) {
        a = cellStart[c];
        n = cellStart[c + 1] - a;
    }
    const unsigned long long big = __ballot(n > kWaveCell);
    if (big) {
        const uint32_t leader = (uint32_t)__ffsll((long long)big) - 1u;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&st->bigCells, (uint32_t)__popcll(big));
        base = __shfl(base, (int)leader, 64);
        if (n > kWaveCell) bigList[base + (uint32_t)__popcll(big & ((1ull << lane) - 1ull))] = c;
    }
    unsigned long long todo = __ballot(n >= 2u && n <= kWaveCell);
    while (todo) {
        const uint32_t j = (uint32_t)__ffsll((long long)todo) - 1u;
        todo &= todo - 1ull;
        const uint32_t ca = __shfl(a, (int)j, 64), cn = __shfl(n, (int)j, 64);
        const uint32_t x = lane < cn ? sortedIdx[ca + lane] : 0xFFFFFFFFu;
        uint32_t r = 0;
        for (uint32_t m = 0; m < cn; ++m) r += __shfl(x, (int)m, 64) < x ? 1u : 0u;
        if (lane < cn) sortedIdx[ca + r] = x;
    }
}

// Cells of more than 64 photons, one workgroup each: the cell in LDS, each
// photon's place = the number of smaller indices.  A cell beyond the LDS
// tile (more than 4096 photons, which clustered scenes of millions of
// photons do not reach) is sorted by one thread.
__global__ __launch_bounds__(256) void k_cell_sort_big(const DevState* st, const uint32_t* bigList,
                                                       const uint32_t* cellStart, uint32_t* sortedIdx) {
    __shared__ uint32_t v[kBlockCell];
    const uint32_t nb = st->bigCells;
    for (uint32_t q = blockIdx.x; q < nb; q += gridDim.x) {
        const uint32_t c = bigList[q];
        const uint32_t a = cellStart[c], n = cellStart[c + 1] - a;
        if (n <= kBlockCell) {
            for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) v[j] = sortedIdx[a + j];
            __syncthreads();
            for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
                const uint32_t x = v[j];
                uint32_t r = 0;
                for (uint32_t m = 0; m < n; ++m) r += v[m] < x ? 1u : 0u;
                sortedIdx[a + r] = x;
            }
            __syncthreads();
        } else if (threadIdx.x == 0) {
            for (uint32_t j = a + 1; j < a + n; ++j) {
                const uint32_t x = sortedIdx[j];
                uint32_t k = j;
                while (k > a && sortedIdx[k - 1] > x) {
                    sortedIdx[k] = sortedIdx[k - 1];
                    --k;
                }
                sortedIdx[k] = x;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_cell_payload(const DevState* st, uint32_t cap, const uint32_t* sortedIdx,
                                                      PhotonArrays raw, PhotonArrays sorted) {
    const uint32_t n = min(st->photonCount, cap);
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t k = sortedIdx[j];
    sorted.posPow[j] = raw.posPow[k];
    sorted.dirPow[j] = raw.dirPow[k];
    sorted.powZ[j] = raw.powZ[k];
}

// ---------------------------------------------------- camera gather --
// launchCameraGatherKernel (PrimalKernels.cpp:500-653), split in stages so
// that the irradiance gathers, which dominate the pass, run in grid-cell
// order instead of pixel order:
//
//   k_cam_init      primary ray of every pixel (the reference's jitter);
//   k_cam_trace     persistent lanes walk each pixel's layers (closest-hit
//                   queries only) and append one entry per layer: the gather
//                   point and normal, the layer's weight w and colour factor
//                   k, in layer order within the pixel;
//   k_entry_key     entry -> place in its pixel's layer list, and a
//                   counting sort of the gather entries by photon-grid cell;
//   k_entry_gather  one irradiance gather per entry, neighbouring lanes on
//                   neighbouring points (shared photon rows, similar loops);
//                   w * (E * k) goes to the entry's place in the layer list;
//   k_cam_composite per pixel, C = C + w * (E * k) over the layers in order,
//                   then the reference's framebuffer update.
//
// Every value is computed by the same expressions as in the reference's
// single loop, and the per-pixel sums run in layer order, so the image is
// bit-identical.  If the entry buffer fills, a pixel stops before the layer
// that did not fit and resumes from there in the next round.
enum : uint32_t { kPixActive = 0u, kPixFinished = 1u, kPixFlushed = 2u };
constexpr uint32_t kEntryNoGather = 0x80000000u;

__global__ __launch_bounds__(256) void k_cam_init(DevCamera cam, const DevState* st, uint32_t pass, int decorrelate,
                                                  PixelArrays px) {
    const uint32_t npx = cam.W * cam.H;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    const uint32_t seed = decorrelate ? st->number + pass : st->number;
    Rng rng(perItemSeed(seed, i));
    const float jx = rng.nextFloat() - 0.5f;
    const float jy = rng.nextFloat() - 0.5f;
    F3 o, d;
    primaryRay(cam, (float)(i % cam.W), (float)(i / cam.W), jx, jy, o, d);
    px.ot[i] = f4(o, 1.0f);
    px.dl[i] = f4(d, u2f(0u));
    px.cs[i] = make_float4(0.0f, 0.0f, 0.0f, u2f(kPixActive));
    px.cnt[i] = 0u;
}

// Slot -> pixel maps of the layer walk.  Pixels are handed out in 8x8 tiles
// (a wave's first rays are neighbours); the order does not change any
// pixel's result.  ImageMap: the tiles of one image.  BatchMap: a list of
// tiles of several images (views x passes).
struct ImageMap {
    uint32_t width, height, tilesX;
    __device__ __forceinline__ uint32_t slots() const { return tilesX * ((height + 7u) / 8u) * 64u; }
    __device__ __forceinline__ bool locate(uint32_t slot, uint32_t& p) const {
        const uint32_t tile = slot >> 6, within = slot & 63u;
        const uint32_t pxx = (tile % tilesX) * 8u + (within & 7u);
        const uint32_t pxy = (tile / tilesX) * 8u + (within >> 3);
        p = pxy * width + pxx;
        return pxx < width && pxy < height;
    }
};

// Segment and pixel of a slot of a batch (see BatchDesc::tiles); false if
// the pixel lies outside the image (right and bottom tile edges).
__device__ __forceinline__ bool batchSlot(const uint4* tiles, uint32_t slot, uint32_t& seg, uint32_t& pxx,
                                          uint32_t& pxy, const BatchSeg*& s, const BatchSeg* segs) {
    const uint4 e = tiles[slot >> 6];
    const uint32_t j = slot - (e.z << 6);
    const uint32_t q = j / e.w;
    seg = e.x + (j - q * e.w);
    s = segs + seg;
    const uint32_t tilesX = s->tilesX;
    pxx = (e.y % tilesX) * 8u + (q & 7u);
    pxy = (e.y / tilesX) * 8u + (q >> 3);
    return pxx < s->W && pxy < s->H;
}

struct BatchMap {
    const uint4* tiles;
    const BatchSeg* segs;
    const BatchHeader* hdr;
    __device__ __forceinline__ uint32_t slots() const { return hdr->nTiles * 64u; }
    __device__ __forceinline__ bool locate(uint32_t slot, uint32_t& p) const {
        uint32_t seg, pxx, pxy;
        const BatchSeg* s;
        if (!batchSlot(tiles, slot, seg, pxx, pxy, s, segs)) return false;
        p = s->slotBase + pxy * s->W + pxx;
        return true;
    }
};

// Slots a wave of the layer walk claims at a time from its region.
constexpr uint32_t kWalkUnit = 64;

template <bool kCount, typename Map, bool kTopLds>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_cam_trace(
    DevScene S, DevState* st, Map map, PixelArrays px, EntryArrays ent, uint32_t cap, uint32_t* cursor,
    uint2* spill, uint32_t inLds) {
    extern __shared__ uint2 s_stack[];
    __shared__ TopLds sTop;
    if (kTopLds) loadTopLds(S, sTop);
    const LaneStack stk = walkStack(s_stack, S.stackDepth, spill, inLds);
    const uint32_t lane = laneId();
    const uint32_t slots = map.slots();
    const uint32_t tiles = slots >> 6;
    uint32_t region = blockIdx.x % kTraceRegions, emptyRegions = 0;
    uint32_t chunkNext = 0, chunkEnd = 0;  // the wave's claimed, untaken slots
    uint32_t layers = 0;
    Work w{};
    bool active = false, exhausted = false;
    uint32_t pixel = 0, r = 0, layer = 0;
    F3 o = mk3(0, 0, 0), d = mk3(0, 0, 0);
    float tr = 1.0f;
    while (true) {
        if (!exhausted) {
            const unsigned long long need = __ballot(!active);
            if (need) {
                // Slots [lo, hi) of the current region; cursor[region] counts
                // the slots drawn from it this round.
                const uint32_t lo = (uint32_t)(((uint64_t)region * tiles) / kTraceRegions) * 64u;
                const uint32_t hi = (uint32_t)(((uint64_t)(region + 1u) * tiles) / kTraceRegions) * 64u;
                // The wave claims a whole 64-slot unit (a few pixels in all
                // their passes) and keeps what its lanes have not taken for
                // its next draws: its lanes stay on neighbouring rays.
                // Claiming only what the lanes need interleaves the waves'
                // draws and scatters each wave over many pixels (the walk
                // took 17 % longer, 28 % with decorrelated passes).
                if (chunkNext >= chunkEnd) {
                    const uint32_t leader = (uint32_t)__ffsll((long long)need) - 1u;
                    uint32_t base = 0;
                    if (lane == leader) base = lo + atomicAdd(cursor + region, kWalkUnit);
                    base = __shfl(base, (int)leader, 64);
                    if (base >= hi) {
                        // Used up (it stays so): the next region, done after all.
                        if (++emptyRegions == kTraceRegions) exhausted = true;
                        region = (region + 1u) % kTraceRegions;
                    } else {
                        chunkNext = base;
                        chunkEnd = min(base + kWalkUnit, hi);
                        if (*reinterpret_cast<volatile const uint32_t*>(&st->overflow) != 0u) exhausted = true;
                    }
                }
                if (chunkNext < chunkEnd) {
                    if (!active) {
                        const uint32_t slot = chunkNext + (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
                        uint32_t p = 0;
                        if (slot < chunkEnd && map.locate(slot, p) && f2u(px.cs[p].w) == kPixActive) {
                            pixel = p;
                            active = true;
                            const float4 a = px.ot[p], b = px.dl[p];
                            o = mk3(a.x, a.y, a.z);
                            tr = a.w;
                            d = mk3(b.x, b.y, b.z);
                            layer = f2u(b.w);
                            r = 0;
                        }
                    }
                    chunkNext = min(chunkNext + (uint32_t)__popcll(need), chunkEnd);
                }
            }
        }
        if (!__any(active)) {
            // Every lane drew a finished pixel (or none): draw again until the
            // slots run out or the entry buffer has filled.  Leaving here
            // earlier would strand the pixels no other wave draws.
            if (exhausted) break;
            continue;
        }
        // One layer: closest hit, then the entry it contributes.
        bool want = false, terminal = true, gather = false;
        F3 q = mk3(0, 0, 0), n = mk3(0, 0, 0), k = mk3(0, 0, 0), nextO = o;
        float wgt = 0.0f, nextTr = tr;
        Work wl{};
        if (active) {
            Hit h;
            const bool hit = kTopLds ? intersectSceneTop<kCount, 2>(S, sTop, o, d, stk, h, wl)
                                     : intersectSceneTop<kCount, 2>(S, TopGlobal{S.tlas, S.inst}, o, d, stk, h, wl);
            if (hit) {
                const DevInstance& I = kTopLds ? sTop.inst[h.inst] : S.inst[h.inst];
                want = true;
                if (I.type == kGeoCloud) {
                    const SurfelShade sh = S.shade[h.prim];
                    n = mk3(sh.nsx, sh.nsy, sh.nsz);
                    if (dot(n, -d) < 0.0f) n = -n;
                    q = h.pos;
                    gather = true;
                    k = (sh.alpha_r * mk3(sh.ar, sh.ag, sh.ab)) * kInvPi;
                    const float ae = sh.opacity * h.alpha;
                    wgt = tr * ae;
                    nextTr = tr * (1.0f - ae);
                    nextO = h.pos + d * 1e-4f;
                    terminal = false;
                } else {
                    const DevMaterial m = S.mats[I.material];
                    const F3 base = mk3(m.r, m.g, m.b);
                    wgt = tr;
                    if (m.power > 0.0f) {
                        k = vmin(base * m.power, mk3(1.0f, 1.0f, 1.0f));
                    } else {
                        q = h.pos;
                        n = hitNormal(S, h);
                        gather = true;
                        k = base * kInvPi;
                    }
                }
            }
        }
        // Wave-aggregated append of the entries of this iteration.
        const unsigned long long m = __ballot(want);
        uint32_t e = 0xFFFFFFFFu;
        if (m) {
            const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1u;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&st->entries, (uint32_t)__popcll(m));
            base = __shfl(base, (int)leader, 64);
            if (want) e = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        }
        bool full = false;
        if (active) {
            bool finished = false;
            if (want && e >= cap) {
                // No room: stop before this layer, resume next round.
                px.ot[pixel] = f4(o, tr);
                px.dl[pixel] = f4(d, u2f(layer));
                st->overflow = 1u;
                px.cnt[pixel] = r;
                active = false;
                full = true;
            } else {
                ++layers;
                if (kCount) {
                    w.pops += wl.pops;
                    w.aabb += wl.aabb;
                    w.prims += wl.prims;
                }
                if (want) {
                    ent.qa[e] = f4(q, n.x);
                    ent.qb[e] = make_float4(n.y, n.z, u2f(pixel), u2f(r | (gather ? 0u : kEntryNoGather)));
                    ent.comp[e] = f4(k, wgt);
                    ++r;
                    ++layer;
                }
                if (terminal) {
                    finished = true;
                } else {
                    o = nextO;
                    tr = nextTr;
                }
            }
            if (finished) {
                px.cnt[pixel] = r;
                px.cs[pixel].w = u2f(kPixFinished);
                active = false;
            }
        }
        if (__any(full)) exhausted = true;
    }
    const uint32_t ws = waveSum(layers);
    if (lane == 0 && ws) atomicAdd(&st->gatherQ, (unsigned long long)ws);
    if (kCount) flushWork(w, st, kWorkGather);
}

// The lanes of a wave that hold the same key, wherever they sit: the lowest
// of them (the group's head), their number and this lane's rank among them.
// One pass per distinct key of the wave; every lane of the wave calls it.
struct KeyGroup {
    uint32_t head, len, rank;
};
__device__ __forceinline__ KeyGroup keyGroup(uint32_t key) {
    const uint32_t lane = laneId();
    unsigned long long todo = __ballot(1);
    KeyGroup g{lane, 1u, 0u};
    while (todo) {
        const uint32_t leader = (uint32_t)__ffsll((long long)todo) - 1u;
        const uint32_t lk = __shfl(key, (int)leader, 64);
        const unsigned long long peers = __ballot(key == lk);
        if (key == lk) {
            g.head = leader;
            g.len = (uint32_t)__popcll(peers);
            g.rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
        }
        todo &= ~peers;
    }
    return g;
}

constexpr uint32_t kNoCell = 0xFFFFFFFFu;

// Place of every entry in its slot's layer-ordered list; grid cell and
// per-cell count of the gather entries (the cell of worldToCellClamped,
// KernelHelpers.h:920-934).  An entry that gathers nothing (an emitter) has
// its contribution w * k written here (the gather's E is one).  Neighbouring
// entries come from neighbouring lanes of the layer walk, so a wave holds
// few distinct cells: one atomic per distinct cell of the wave.
// The entry kernels stride over the entries with a bounded grid: the count
// is known only on the device and the capacity is far above it.  The loop is
// uniform over the block, so whole waves take part in the run detection.
__global__ __launch_bounds__(256) void k_entry_key(GridParams g, const DevState* st, uint32_t cap, PixelArrays px,
                                                   EntryArrays ent, uint32_t* cellCount) {
    const uint32_t n = min(st->entries, cap);
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        uint32_t c = kNoCell;
        if (i < n) {
            const float4 b = ent.qb[i];
            const uint32_t tag = f2u(b.w);
            const uint32_t pos = px.off[f2u(b.z)] + (tag & ~kEntryNoGather);
            ent.pos[i] = pos;
            if (tag & kEntryNoGather) {
                const float4 k = ent.comp[i];
                ent.contrib[pos] = f4(k.w * mk3(k.x, k.y, k.z), 0.0f);
            } else {
                const float4 a = ent.qa[i];
                const F3 f = divv(mk3(a.x, a.y, a.z) - mk3(g.ox, g.oy, g.oz), mk3(g.cell, g.cell, g.cell));
                const int x = min(max((int)floorf(f.x), 0), g.rx - 1);
                const int y = min(max((int)floorf(f.y), 0), g.ry - 1);
                const int z = min(max((int)floorf(f.z), 0), g.rz - 1);
                c = (uint32_t)((z * g.ry + y) * g.rx + x);
            }
            ent.cell[i] = c;
        }
        const KeyGroup r = keyGroup(c);
        if (laneId() == r.head && c != kNoCell) atomicAdd(&cellCount[c], r.len);
    }
}

// Counting-sort scatter of the gather entries; leaves cellCount zeroed.  The
// lanes of a wave with the same cell take their places with one atomic.
__global__ __launch_bounds__(256) void k_entry_scatter(const DevState* st, uint32_t cap, const uint32_t* cellStart,
                                                       uint32_t* cellCount, EntryArrays ent) {
    const uint32_t n = min(st->entries, cap);
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        const uint32_t c = i < n ? ent.cell[i] : kNoCell;
        const KeyGroup r = keyGroup(c);
        const uint32_t lane = laneId();
        uint32_t base = 0;
        if (lane == r.head && c != kNoCell) base = cellStart[c] + atomicSub(&cellCount[c], r.len) - r.len;
        base = __shfl(base, (int)r.head, 64);
        if (c != kNoCell) ent.sorted[base + r.rank] = i;
    }
}

// One irradiance gather per entry in cell order; the entry's contribution
// w * (E * k) goes to its place in the slot's layer list (the products the
// reference's loop forms, PrimalKernels.cpp:583-585, :620, :626-628).
template <bool kCount>
__global__ __launch_bounds__(256) void k_entry_gather(DevGrid G, DevState* st, const uint32_t* nGather,
                                                      EntryArrays ent) {
    // One lane per entry (a one-shot grid over the capacity): gathers vary
    // in cost, and fresh waves take the place of finished ones, which a
    // strided loop would not.
    const uint32_t n = *nGather;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    Work w{};
    if (i < n) {
        const uint32_t e = ent.sorted[i];
        const float4 a = ent.qa[e], b = ent.qb[e], c = ent.comp[e];
        const uint32_t pos = ent.pos[e];
        const F3 E = gatherIrradiance<kCount, true>(G, mk3(a.x, a.y, a.z), mk3(a.w, b.x, b.y), w);
        const F3 Lo = E * mk3(c.x, c.y, c.z);
        ent.contrib[pos] = f4(c.w * Lo, 0.0f);
    }
    if (kCount) flushWork(w, st, kWorkGather);
}

// fb == nullptr (batches): a finished pixel keeps its colour in cs and the
// framebuffer update is left to k_batch_resolve.  nDev: device-side count.
// C = C + w * (E * k) over the pixel's layers in order.
__global__ __launch_bounds__(256) void k_cam_composite(uint32_t npx, const uint32_t* nDev, PixelArrays px,
                                                       EntryArrays ent, float invSpp, float4* fb) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (nDev ? min(*nDev, npx) : npx)) return;
    const float4 cs = px.cs[p];
    const uint32_t status = f2u(cs.w);
    if (status == kPixFlushed) return;
    F3 C = mk3(cs.x, cs.y, cs.z);
    const uint32_t first = px.off[p], cnt = px.cnt[p];
    for (uint32_t j = 0; j < cnt; ++j) C = C + ld3(ent.contrib[first + j]);
    // The next round's slot scan must see zero for every pixel it did not
    // touch, finished ones included.
    px.cnt[p] = 0u;
    if (status == kPixFinished) {
        if (fb) {
            const float4 f = fb[p];
            fb[p] = make_float4(C.x * invSpp + f.x, C.y * invSpp + f.y, C.z * invSpp + f.z, 1.0f * invSpp + f.w);
        }
        px.cs[p] = make_float4(C.x, C.y, C.z, u2f(kPixFlushed));
    } else {
        px.cs[p] = make_float4(C.x, C.y, C.z, u2f(kPixActive));
    }
}

// Primary rays of a batch: one thread per tile pixel; the same jitter and
// ray as k_cam_init for (view, pass) with seed = view seed + pass.
__global__ __launch_bounds__(256) void k_batch_init(BatchDesc b, PixelArrays px) {
    const uint32_t t = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (t >= b.hdr->nTiles) return;
    uint32_t seg, pxx, pxy;
    const BatchSeg* sp;
    if (!batchSlot(b.tiles, t * 64u + (threadIdx.x & 63u), seg, pxx, pxy, sp, b.segs)) return;
    const BatchSeg s = *sp;
    const uint32_t pixel = pxy * s.W + pxx, slot = s.slotBase + pixel;
    Rng rng(perItemSeed(s.seed, pixel));
    const float jx = rng.nextFloat() - 0.5f;
    const float jy = rng.nextFloat() - 0.5f;
    F3 o, d;
    primaryRay(b.cams[s.view], (float)pxx, (float)pxy, jx, jy, o, d);
    px.ot[slot] = f4(o, 1.0f);
    px.dl[slot] = f4(d, u2f(0u));
    px.cs[slot] = make_float4(0.0f, 0.0f, 0.0f, u2f(kPixActive));
    px.cnt[slot] = 0u;
}

// Framebuffer update of a batch, per view pixel over the view's passes in
// pass order: the sequence of k_cam_composite's updates.
__global__ __launch_bounds__(256) void k_batch_resolve(BatchDesc b, PixelArrays px, float4* fb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.hdr->nResolve) return;
    uint32_t lo = 0, hi = b.hdr->nRuns - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (b.runs[mid].resolveBase <= i)
            lo = mid;
        else
            hi = mid - 1u;
    }
    const BatchRun r = b.runs[lo];
    const uint32_t pixel = i - r.resolveBase;
    float4 f = fb[r.fbOffset + pixel];
    for (uint32_t k = 0; k < r.numSegs; ++k) {
        const float4 c = px.cs[b.segs[r.firstSeg + k].slotBase + pixel];
        f = make_float4(c.x * r.invSpp + f.x, c.y * r.invSpp + f.y, c.z * r.invSpp + f.z, 1.0f * r.invSpp + f.w);
    }
    fb[r.fbOffset + pixel] = f;
}

// launchPostProcessKernel (PostprocessKernel.cpp:11-106), float LDR output.
__global__ __launch_bounds__(256) void k_post(const float4* fb, float4* ldr, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float invGamma = 1.0f / 2.2f;
    const float4 v = fb[i];
    float c[3] = {v.x, v.y, v.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float x = c[k] * 1.0f;
        x = fmaxf(x, 0.0f);
        x = pow_r(x, invGamma);
        x = fminf(fmaxf(x, 0.0f), 1.0f);
        c[k] = x;
    }
    ldr[i] = make_float4(c[0], c[1], c[2], 1.0f);
}

// ---------------------------------------------------------- adjoint --
// launchRayGenAdjointKernel (AdjointKernels.cpp:15-84).  The adjoint seeds
// from PathTracerSettings::random.number (default 42, setting adjoint_seed);
// with adjoint_decorrelate the host passes adjoint_seed + sample instead.
__global__ __launch_bounds__(256) void k_adj_raygen(DevCamera cam, const float* dloss, RayQueue q,
                                                    PendingArrays pend, DevState* st, uint32_t seed) {
    const uint32_t npx = cam.W * cam.H;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) st->qCount = npx;
    if (i >= npx) return;
    Rng rng(perItemSeed((uint64_t)seed, i));
    const float jx = rng.nextFloat() - 0.5f;
    const float jy = rng.nextFloat() - 0.5f;
    F3 o, d;
    primaryRay(cam, (float)(i % cam.W), (float)(i / cam.W), jx, jy, o, d);
    q.o[i] = f4(o, u2f(i));
    q.d[i] = f4(d, u2f(0u));
    q.T[i] = make_float4(dloss[3 * (size_t)i], dloss[3 * (size_t)i + 1], dloss[3 * (size_t)i + 2], 0.0f);
    pend.p0[i] = make_float4(0.0f, 0.0f, 0.0f, u2f(kNone));
}

// Adjoint primary rays of a batch: one thread per tile pixel, the jitter and
// ray of k_adj_raygen for the pixel of its view; queue index = slot.
__global__ __launch_bounds__(256) void k_adj_raygen_batch(BatchDesc b, const float* dloss, RayQueue q,
                                                          PendingArrays pend, DevState* st) {
    const uint32_t t = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (blockIdx.x == 0 && threadIdx.x == 0) st->qCount = b.hdr->nSlots;
    if (t >= b.hdr->nTiles) return;
    uint32_t seg, pxx, pxy;
    const BatchSeg* sp;
    if (!batchSlot(b.tiles, t * 64u + (threadIdx.x & 63u), seg, pxx, pxy, sp, b.segs)) return;
    const BatchSeg s = *sp;
    const uint32_t pixel = pxy * s.W + pxx, slot = s.slotBase + pixel;
    Rng rng(perItemSeed((uint64_t)s.seed, pixel));
    const float jx = rng.nextFloat() - 0.5f;
    const float jy = rng.nextFloat() - 0.5f;
    F3 o, d;
    primaryRay(b.cams[s.view], (float)pxx, (float)pxy, jx, jy, o, d);
    const size_t src = 3 * ((size_t)s.pixOffset + pixel);
    q.o[slot] = f4(o, u2f(slot));
    q.d[slot] = f4(d, u2f(0u));
    q.T[slot] = make_float4(dloss[src], dloss[src + 1], dloss[src + 2], u2f(seg));
    pend.p0[slot] = make_float4(0.0f, 0.0f, 0.0f, u2f(kNone));
}

// First and end queue index of every segment present in a batch's queue
// (the queue stays in slot order: raygen writes slot order, compaction is
// stable); absent segments keep first = end = 0 from k_seg_clear.
__global__ __launch_bounds__(256) void k_seg_clear(uint32_t* segFirst, uint32_t* segEnd, uint32_t nSegs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nSegs) segFirst[i] = segEnd[i] = 0u;
}

__global__ __launch_bounds__(256) void k_seg_first(const DevState* st, RayQueue q, uint32_t* segFirst,
                                                   uint32_t* segEnd) {
    const uint32_t n = st->qCount;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t seg = f2u(q.T[i].w);
    if (i == 0 || f2u(q.T[i - 1].w) != seg) segFirst[seg] = i;
    if (i + 1 == n || f2u(q.T[i + 1].w) != seg) segEnd[seg] = i + 1;
}

// Lanes of each view group: K segments x the most rays of one of them.
__global__ __launch_bounds__(64) void k_group_lanes(const uint32_t* segFirst, const uint32_t* segEnd, uint4* groups,
                                                    uint32_t nGroups, const uint32_t* groupSegs) {
    if (threadIdx.x != 0) return;
    uint32_t base = 0;
    for (uint32_t g = 0; g < nGroups; ++g) {
        uint4 gr = groups[g];
        uint32_t most = 0;
        for (uint32_t k = 0; k < gr.y; ++k) {
            const uint32_t s = groupSegs[gr.x + k];
            most = max(most, segEnd[s] - segFirst[s]);
        }
        gr.z = base;
        gr.w = most;
        groups[g] = gr;
        base += gr.y * most;
    }
}

// launchAdjointIntersectKernel (AdjointKernels.cpp:87-386) for one queue.
// Capped at 128 VGPRs (4 waves per SIMD): it needs 125 without spills since
// the hit's alpha is a call (uncapped it took 129, and 3 waves per SIMD ran
// 5-8 % slower).
template <bool kCount>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_adj_trace(DevScene S, DevState* st, uint32_t bounceIndex,
                                                           RayQueue in, RayQueue out, uint8_t* extFlags,
                                                           EventArrays ev, uint8_t* evFlags, PendingArrays pend,
                                                           uint32_t pendCap, TraceParams tp) {
    extern __shared__ uint2 s_stack[];
    const LaneStack stk = walkStack(s_stack, S.stackDepth, tp.spill, tp.inLds);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = st->qCount;
    uint32_t i = t;
    bool active = t < n;
    if (tp.groups) {
        // Lane t -> (view group, segment, rank) -> queue index.
        active = false;
        uint32_t lo = 0, hi = tp.nGroups;  // last group whose first lane <= t
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (tp.groups[mid].z <= t)
                lo = mid;
            else
                hi = mid;
        }
        const uint4 gr = tp.nGroups ? tp.groups[lo] : make_uint4(0u, 0u, 0u, 0u);
        const uint32_t local = t - gr.z;
        if (t >= gr.z && local < gr.y * gr.w) {
            const uint32_t s = tp.groupSegs[gr.x + local % gr.y], r = local / gr.y;
            if (r < tp.segEnd[s] - tp.segFirst[s]) {
                i = tp.segFirst[s] + r;
                active = true;
            }
        }
    }
    Work w{};
    if (active) {
        uint8_t ef = 0, vf = 0;
        const float4 o4 = in.o[i], d4 = in.d[i], T4 = in.T[i];
        const uint32_t seg = f2u(T4.w);
        const uint32_t rank = tp.segFirst ? i - tp.segFirst[seg] : i;
        const float invS = tp.segs ? tp.segs[seg].invSpp : tp.invSpp;
        Rng rng(perItemSeed((uint64_t)(tp.segs ? tp.segs[seg].seed : tp.seed), rank));
        const F3 o = ld3(o4), d = ld3(d4), T = ld3(T4);
        const uint32_t path = f2u(o4.w);
        const uint32_t bounce = f2u(d4.w);
        const uint32_t dpx = !tp.debug ? 0u
                             : tp.segs ? tp.segs[seg].pixOffset + (path - tp.segs[seg].slotBase)
                                       : tp.debugBase + path;
        Hit h;
        if (intersectScene<kCount>(S, o, d, stk, h, w)) {
            const DevInstance& I = S.inst[h.inst];
            const bool cloud = I.type == kGeoCloud;
            F3 on, ng = mk3(0, 0, 0), dir;
            float pdf;
            SurfelShade sh{};
            if (!cloud) {
                ng = hitNormal(S, h);
                sampleCosine(rng, ng, dir, pdf);
                on = ng;
            } else {
                sh = S.shade[h.prim];
                const F3 nc = mk3(sh.nsx, sh.nsy, sh.nsz);
                const float sc = dot(nc, -d);
                const int sign = sc >= 0.0f ? 1 : -1;
                on = (float)sign * nc;
                sampleUniformHemi(rng, on, dir, pdf);
            }
            // Complete the pending interaction with this hit as its endpoint.
            if (path < pendCap) {
                const float4 p0 = pend.p0[path];
                const uint32_t kind = f2u(p0.w);
                const bool reflectDone = kind == kReflectS && cloud;
                const bool nullDone = kind == kNullT;
                if (reflectDone || nullDone) {
                    const float4 p1 = pend.p1[path];
                    const uint2 p2 = pend.p2[path];
                    const F3 Tp = ld3(p1) * (reflectDone ? tp.qReflect : tp.qNull);
                    const float endCos = fmaxf(0.0f, dot(dir, on));
                    ev.e0[i] = f4(h.pos, u2f(kind));
                    if (reflectDone) {
                        ev.e1[i] = f4(on, u2f(h.prim));
                        ev.e2[i] = f4(Tp, h.alpha);
                        ev.e3[i] = make_float4(u2f(S.inst[p2.y].material), pdf, endCos, invS);
                        if (tp.fullGradients || tp.debug) {  // this ray hit the endpoint surfel
                            ev.e4[i] = f4(o, 0.0f);
                            ev.e5[i] = f4(d, u2f(dpx));
                        }
                    } else {
                        ev.e1[i] = f4(on, u2f(p2.x));
                        ev.e2[i] = f4(Tp, p1.w);
                        ev.e3[i] = make_float4(u2f(cloud ? kGeoCloud : kGeoMesh), u2f(cloud ? h.prim : I.material),
                                               endCos, invS);
                        if (tp.fullGradients) {  // the ray that crossed the pending surfel
                            ev.e4[i] = pend.p3[path];
                            float4 e5 = pend.p4[path];
                            e5.w = u2f(dpx);
                            ev.e5[i] = e5;
                        } else if (tp.debug) {
                            ev.e5[i] = make_float4(0.0f, 0.0f, 0.0f, u2f(dpx));
                        }
                    }
                    vf = 1;
                    pend.p0[path] = make_float4(0.0f, 0.0f, 0.0f, u2f(kNone));
                }
            }
            F3 no, nT;
            bool spawn = false;
            uint32_t newKind = kNone;
            F3 pT = mk3(0, 0, 0);
            float pAlpha = 0.0f;
            uint2 pIds = make_uint2(kInvalid, kInvalid);
            F3 ndir = dir;
            bool projection = false;
            if (!cloud) {
                const DevMaterial m = S.mats[I.material];
                const F3 base = mk3(m.r, m.g, m.b);
                no = h.pos + ng * 1e-6f;
                nT = T * base;
                spawn = true;
                newKind = kReflectS;
                pT = T * base;
                pIds = make_uint2(kInvalid, h.inst);
            } else {
                const float u = rng.nextFloat();
                if (u < tp.qNull) {
                    const float att = 1.0f - h.alpha * sh.opacity;
                    const float w = att / tp.qNull;
                    no = h.pos + d * 1e-5f;
                    ndir = d;
                    nT = T * w;
                    spawn = true;
                    newKind = kNullT;
                    pT = divs(T, tp.qNull);
                    pAlpha = h.alpha;
                    pIds = make_uint2(h.prim, kInvalid);
                } else if (u < tp.qNull + tp.qReflect) {
                    const F3 fs = (sh.alpha_r * mk3(sh.ar, sh.ag, sh.ab)) * kInvPi;
                    const float cosT = fmaxf(0.0f, dot(dir, on));
                    const float alpha = h.alpha * sh.opacity;
                    const F3 mult = divs((alpha / tp.qReflect) * (fs * cosT), pdf);
                    no = h.pos + on * 1e-5f;
                    nT = T * mult;
                    spawn = true;
                    projection = bounceIndex == 0;
                } else if (u < tp.qNull + tp.qReflect + tp.qTransmit) {
                    const float alpha = h.alpha * sh.opacity;
                    const float w = (alpha * 0.0f) / tp.qTransmit;  // alpha_t = 0
                    const F3 nc = mk3(sh.nsx, sh.nsy, sh.nsz);
                    const float sc = dot(nc, -d);
                    const int sign = sc >= 0.0f ? 1 : -1;
                    const F3 ot = (float)sign * nc;
                    float p2df;
                    sampleUniformHemi(rng, ot, ndir, p2df);
                    no = h.pos + (-ot) * 1e-5f;
                    nT = (T * w) * mk3(sh.ar, sh.ag, sh.ab);
                    spawn = true;
                }
            }
            if (spawn && russianRoulette(rng, bounce + 1, nT, tp.rrStart)) {
                if (newKind != kNone && path < pendCap) {
                    pend.p0[path] = f4(h.pos, u2f(newKind));
                    pend.p1[path] = f4(pT, pAlpha);
                    pend.p2[path] = pIds;
                    if (tp.fullGradients && newKind == kNullT) {
                        pend.p3[path] = f4(o, 0.0f);
                        pend.p4[path] = f4(d, 0.0f);
                    }
                }
                if (projection && path < pendCap) {
                    ev.e0[i] = f4(h.pos, u2f(kProjection));
                    ev.e1[i] = f4(on, u2f(h.prim));
                    ev.e2[i] = f4(divs(T, tp.qReflect), h.alpha);
                    ev.e3[i] = make_float4(0.0f, 0.0f, 0.0f, invS);
                    if (tp.fullGradients || tp.debug) {
                        ev.e4[i] = f4(o, 0.0f);
                        ev.e5[i] = f4(d, u2f(dpx));
                    }
                    vf = 1;
                }
                ef = 1;
                out.o[i] = f4(no, o4.w);
                out.d[i] = f4(ndir, u2f(bounce + 1));
                out.T[i] = f4(nT, T4.w);
            }
        }
        extFlags[i] = ef;
        evFlags[i] = vf;
    }
    waveCount(active, &st->adjointQ);
    if (kCount) flushWork(w, st, kWorkAdjoint);
}

// Counting sort of the pooled events by the photon-grid cell of their
// gather point, so the transport's irradiance gathers run in cell order.
// Cell of every pooled event and per-cell counts; neighbouring events come
// from neighbouring pixels, so equal cells arrive in runs: one atomic per
// run (the loop is uniform over the block, as in k_entry_key).
__global__ __launch_bounds__(256) void k_event_key(GridParams g, const DevState* st, uint32_t cap, EventArrays ev,
                                                   uint32_t* cell, uint32_t* cellCount) {
    const uint32_t n = min(st->evPool, cap);
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        uint32_t c = kNoCell;
        if (i < n) {
            const float4 a = ev.e0[i];
            const F3 f = divv(mk3(a.x, a.y, a.z) - mk3(g.ox, g.oy, g.oz), mk3(g.cell, g.cell, g.cell));
            const int x = min(max((int)floorf(f.x), 0), g.rx - 1);
            const int y = min(max((int)floorf(f.y), 0), g.ry - 1);
            const int z = min(max((int)floorf(f.z), 0), g.rz - 1);
            c = (uint32_t)((z * g.ry + y) * g.rx + x);
            cell[i] = c;
        }
        const KeyGroup r = keyGroup(c);
        if (laneId() == r.head && c != kNoCell) atomicAdd(&cellCount[c], r.len);
    }
}

// Counting-sort scatter of the events (order within a cell: any; the
// transport's FP32 atomics sum in device order anyway); leaves cellCount
// zeroed.
__global__ __launch_bounds__(256) void k_event_scatter(const DevState* st, uint32_t cap, const uint32_t* cell,
                                                       const uint32_t* cellStart, uint32_t* cellCount,
                                                       uint32_t* sorted) {
    const uint32_t n = min(st->evPool, cap);
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        const uint32_t c = i < n ? cell[i] : kNoCell;
        const KeyGroup r = keyGroup(c);
        const uint32_t lane = laneId();
        uint32_t base = 0;
        if (lane == r.head && c != kNoCell) base = cellStart[c] + atomicSub(&cellCount[c], r.len) - r.len;
        base = __shfl(base, (int)r.head, 64);
        if (c != kNoCell) sorted[base + r.rank] = i;
    }
}

// Full-gradient contribution of one event: d(alpha_geom)/d(geometry) of
// surfel `prim` for the recorded ray, times eta * w (w = the event's weight
// sum_rgb(T * Lo) * invSpp with its sign).
// c[] = the surfel's 16 accumulator columns (position 0-2, t_u 3-5, t_v 6-8,
// scale 9-10, albedo 11-13, beta 15).
__device__ void addGeometryGradient(const DevScene& S, uint32_t prim, float4 ro, float4 rd, float w, float c[16]) {
    const DevInstance& I = S.inst[S.cloudInst];
    float hv[4];
    mulv4(I.w2o, ro.x, ro.y, ro.z, 1.0f, hv);
    const F3 oo = divs(mk3(hv[0], hv[1], hv[2]), hv[3]);
    mulv4(I.w2o, rd.x, rd.y, rd.z, 0.0f, hv);
    const F3 od = normalize(mk3(hv[0], hv[1], hv[2]));
    float g[12];
    if (!alphaGradient(S.geo[prim], oo, od, g)) return;
    const float f = S.shade[prim].opacity * w;
#pragma unroll
    for (int k = 0; k < 11; ++k) c[k] += g[k] * f;
    c[15] += g[11] * f;
}

// Sums v over each run of equal keys among consecutive lanes of the wave
// into the run's first lane (a segmented suffix reduction by shuffles);
// returns whether this lane heads its run.  Every lane of the wave calls it.
template <int kN>
__device__ __forceinline__ bool runReduce(uint32_t key, float (&v)[kN]) {
    const uint32_t lane = laneId();
    const uint32_t next = __shfl_down(key, 1, 64);
    const uint32_t prev = __shfl_up(key, 1, 64);
    const unsigned long long ends = __ballot(lane == 63u || next != key);
    const uint32_t end = lane + (uint32_t)__ffsll((long long)(ends >> lane)) - 1u;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
#pragma unroll
        for (int k = 0; k < kN; ++k) {
            const float t = __shfl_down(v[k], o, 64);
            if (lane + o <= end) v[k] += t;
        }
    }
    return lane == 0u || prev != key;
}

// launchAdjointTransportKernel (AdjointKernels.cpp:389-509), one lane per
// pooled event.  Each event carries the 1/spp of the view whose adjoint ray
// made it (e3.w).
template <bool kCount>
__global__ __launch_bounds__(256) void k_adj_transport(DevScene S, DevGrid G, DevState* st, uint32_t poolCap,
                                                       EventArrays ev, const uint32_t* sorted, float* grad,
                                                       float* full, float* debugImg) {
    const uint32_t n = min(st->evPool, poolCap);
    if (blockIdx.x * blockDim.x >= n) return;  // the grid covers the pool's capacity
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < n;
    Work w{};
    // Contributions of this lane's event: opacity in c[0], the full
    // gradient's columns in c[1 + column]; lanes of a wave that follow one
    // another with the same target surfel (events are cell-sorted, so runs
    // are common) add theirs up before one lane issues the atomics.
    float c[17];
#pragma unroll
    for (int k = 0; k < 17; ++k) c[k] = 0.0f;
    uint32_t target = 0xFFFFFFFFu;
    if (active) {
        const uint32_t j = sorted[i];
        const float4 e0 = ev.e0[j], e1 = ev.e1[j], e2 = ev.e2[j], e3 = ev.e3[j];
        const uint32_t kind = f2u(e0.w);
        target = f2u(e1.w);
        const float invSpp = e3.w;
        const F3 T = ld3(e2);
        const float a = e2.w;
        const F3 E = gatherIrradiance<kCount>(G, ld3(e0), ld3(e1), w);
        float val = 0.0f;
        if (kind == kProjection) {
            const SurfelShade sh = S.shade[target];
            const F3 fr = (sh.alpha_r * mk3(sh.ar, sh.ag, sh.ab)) * kInvPi;
            const F3 Lr = fr * E;
            const F3 g = (a * T) * Lr;
            val = fsum(g) * invSpp;
            if (full) {
                addGeometryGradient(S, target, ev.e4[j], ev.e5[j], fsum(T * Lr) * invSpp, c + 1);
                const F3 ga = ((a * sh.opacity * sh.alpha_r * kInvPi) * (T * E)) * invSpp;
                c[1 + 11] += ga.x;
                c[1 + 12] += ga.y;
                c[1 + 13] += ga.z;
            }
        } else if (kind == kNullT) {
            F3 base;
            float power;
            if (f2u(e3.x) == kGeoMesh) {
                const DevMaterial m = S.mats[f2u(e3.y)];
                base = mk3(m.r, m.g, m.b);
                power = m.power;
            } else {
                const SurfelShade es = S.shade[f2u(e3.y)];
                base = es.alpha_r * mk3(es.ar, es.ag, es.ab);
                power = 0.0f;
            }
            const F3 Lr = (E * base) * kInvPi;
            F3 Le = mk3(0, 0, 0);
            if (power > 0.0f && S.nLights > 0) Le = (base * (power / (kPi * S.lights[0].area))) * e3.z;
            const F3 Lo = Le + Lr;
            const F3 g = ((-a) * T) * Lo;
            val = fsum(g) * invSpp;
            if (full) addGeometryGradient(S, target, ev.e4[j], ev.e5[j], -(fsum(T * Lo) * invSpp), c + 1);
        } else if (kind == kReflectS) {
            const DevMaterial m = S.mats[f2u(e3.x)];
            const F3 Lo = (E * mk3(m.r, m.g, m.b)) * kInvPi;
            const F3 p = divs(T, e3.y);
            const F3 g = (p * Lo) * a;
            val = fsum(g) * invSpp;
            if (full) addGeometryGradient(S, target, ev.e4[j], ev.e5[j], fsum(p * Lo) * invSpp, c + 1);
        }
        c[0] = val;
        if (debugImg) {
            float* px = debugImg + 4 * (size_t)f2u(ev.e5[j].w);
            atomicAdd(px, val);
            atomicAdd(px + 1, val);
            atomicAdd(px + 2, val);
            __hip_atomic_store(px + 3, 1.0f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    bool head;
    if (full) {
        head = runReduce(target, c);
    } else {
        float v[1] = {c[0]};
        head = runReduce(target, v);
        c[0] = v[0];
    }
    if (head && target != 0xFFFFFFFFu) {
        atomicAdd(&grad[target], c[0]);
        if (full) {
            float* out = full + (size_t)target * 16;
#pragma unroll
            for (int k = 0; k < 16; ++k)
                if (k != 14 && c[1 + k] != 0.0f) atomicAdd(out + k, c[1 + k]);
        }
    }
    waveCount(active, &st->events);
    if (kCount) flushWork(w, st, kWorkAdjoint);
}

// ------------------------------------------------------ diagnostics --
__global__ __launch_bounds__(kTraceBlock) void k_trace_rays(DevScene S, uint32_t n, const float* o, const float* d,
                                                            int32_t* inst, int32_t* prim, float* t, float* pos,
                                                            float* alpha) {
    extern __shared__ uint2 s_stack[];
    const LaneStack stk = laneStack(s_stack, S.stackDepth);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Hit h;
    Work w{};
    const bool hit = intersectScene<false>(S, mk3(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                                           mk3(d[3 * i], d[3 * i + 1], d[3 * i + 2]), stk, h, w);
    inst[i] = hit ? (int32_t)h.inst : -1;
    prim[i] = hit ? (int32_t)h.prim : -1;
    t[i] = hit ? h.t : 0.0f;
    pos[3 * i] = hit ? h.pos.x : 0.0f;
    pos[3 * i + 1] = hit ? h.pos.y : 0.0f;
    pos[3 * i + 2] = hit ? h.pos.z : 0.0f;
    alpha[i] = hit ? h.alpha : 0.0f;
}

// Full-gradient array (16 floats per surfel) into the ABI's arrays.
__global__ __launch_bounds__(256) void k_split_gradients(const float* full, uint32_t n, float* position,
                                                         float* tangentU, float* tangentV, float* scale,
                                                         float* albedo, float* beta) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* g = full + (size_t)i * 16;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (position) position[3 * (size_t)i + k] = g[k];
        if (tangentU) tangentU[3 * (size_t)i + k] = g[3 + k];
        if (tangentV) tangentV[3 * (size_t)i + k] = g[6 + k];
        if (albedo) albedo[3 * (size_t)i + k] = g[11 + k];
    }
    if (scale) {
        scale[2 * (size_t)i] = g[9];
        scale[2 * (size_t)i + 1] = g[10];
    }
    if (beta) beta[i] = g[15];
}

// Opacity of every surfel from device memory (BVH order).
__global__ __launch_bounds__(256) void k_set_opacity(SurfelShade* shade, const float* opacity, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) shade[i].opacity = opacity[i];
}

// apply_point_optimization with the parameters already in device memory:
// each surfel's intersection and shading records from the same binary32
// expressions as the host's refreshSurfels (orthonormal frame of
// intersectSurfel, KernelHelpers.h:646-648; shading normal
// normalize(cross(t_u, t_v)), :497-501; 4 e^beta), so that both paths give
// identical records.  The values are written as given (no clamp or
// normalisation), as the reference's apply_point_optimization writes them
// (pale_bindings.cpp:853-1056); alpha_r is kept.  `raw` receives the
// parameters for the host mirror.
__global__ __launch_bounds__(256) void k_apply_points(SurfelGeo* geo, SurfelShade* shade, float* raw,
                                                      DevPointArrays a, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const F3 p = mk3(a.position[3 * i], a.position[3 * i + 1], a.position[3 * i + 2]);
    const F3 tu = mk3(a.tangent_u[3 * i], a.tangent_u[3 * i + 1], a.tangent_u[3 * i + 2]);
    const F3 tv = mk3(a.tangent_v[3 * i], a.tangent_v[3 * i + 1], a.tangent_v[3 * i + 2]);
    const float su = a.scale[2 * i], sv = a.scale[2 * i + 1];
    const F3 alb = mk3(a.albedo[3 * i], a.albedo[3 * i + 1], a.albedo[3 * i + 2]);
    const float op = a.opacity[i], beta = a.beta[i];
    const F3 u = normalize(tu);
    const F3 v = normalize(tv - u * dot(u, tv));
    const F3 nrm = normalize(cross(u, v));
    SurfelGeo g;
    g.px = p.x; g.py = p.y; g.pz = p.z; g.su = su;
    g.nx = nrm.x; g.ny = nrm.y; g.nz = nrm.z; g.sv = sv;
    g.tux = tu.x; g.tuy = tu.y; g.tuz = tu.z; g.e4b = 4.0f * exp_r(beta);
    g.tvx = tv.x; g.tvy = tv.y; g.tvz = tv.z; g.pad = 0.0f;
    geo[i] = g;
    const F3 ns = normalize(cross(tu, tv));
    SurfelShade h;
    h.nsx = ns.x; h.nsy = ns.y; h.nsz = ns.z; h.opacity = op;
    h.ar = alb.x; h.ag = alb.y; h.ab = alb.z; h.alpha_r = shade[i].alpha_r;
    shade[i] = h;
    float4* r = reinterpret_cast<float4*>(raw) + 4 * (size_t)i;
    r[0] = make_float4(p.x, p.y, p.z, tu.x);
    r[1] = make_float4(tu.y, tu.z, tv.x, tv.y);
    r[2] = make_float4(tv.z, su, sv, alb.x);
    r[3] = make_float4(alb.y, alb.z, op, beta);
}

// Irradiance estimate at given points (gatherDiffuseIrradianceAtPoint,
// KernelHelpers.h:990-1037) for tests of the gather on its own.
__global__ __launch_bounds__(256) void k_gather_points(DevGrid G, uint32_t n, const float* q, const float* nrm,
                                                       float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Work w{};
    const F3 E = gatherIrradiance<false>(G, mk3(q[3 * i], q[3 * i + 1], q[3 * i + 2]),
                                         mk3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]), w);
    out[3 * i] = E.x;
    out[3 * i + 1] = E.y;
    out[3 * i + 2] = E.z;
}

// ------------------------------------------------------------ launchers --
static inline uint32_t blocks(uint32_t n, uint32_t b) { return (n + b - 1) / b; }
// Grid of the light per-item kernels (keys and scatters of the entry and
// event sorts) that stride over buffers whose count is known only on the
// device: 256 CUs x 32 workgroups of 256 lanes.
constexpr uint32_t kEntryBlocks = 8192;
static inline size_t stackBytes(const DevScene& S) { return (size_t)S.stackDepth * sizeof(uint2) * kTraceBlock; }

void launchFill(hipStream_t s, uint32_t* p, uint32_t n, uint32_t v) {
    if (n) k_fill<<<min(blocks(n, 256), 65535u), 256, 0, s>>>(p, n, v);
}
void launchFillF(hipStream_t s, float* p, uint32_t n, float v) {
    if (n) k_fillf<<<min(blocks(n, 256), 65535u), 256, 0, s>>>(p, n, v);
}
void launchIota(hipStream_t s, uint32_t* p, uint32_t n) {
    if (n) k_iota<<<min(blocks(n, 256), 65535u), 256, 0, s>>>(p, n);
}
void launchSeedNext(hipStream_t s, DevState* st, const uint32_t* table, uint32_t len, int bounce) {
    k_seed_next<<<1, 64, 0, s>>>(st, table, len, bounce);
}
void launchBeginView(hipStream_t s, DevState* st, uint32_t number) { k_begin_view<<<1, 64, 0, s>>>(st, number); }
void launchEmit(hipStream_t s, const DevScene& S, DevState* st, uint32_t P, float invEmitted, RayQueue staged,
                uint8_t* flags) {
    if (P) k_emit<<<blocks(P, 256), 256, 0, s>>>(S, st, P, invEmitted, staged, flags);
}
void launchPhotonTrace(hipStream_t s, const DevScene& S, DevState* st, uint32_t maxRays, RayQueue in,
                       RayQueue staged, uint8_t* extFlags, PhotonArrays dep, uint8_t* depFlags, TraceParams tp,
                       Reservoir rs, uint32_t* tiles, bool count) {
    if (!maxRays) return;
    const dim3 g(blocks(maxRays, kTraceBlock));
    if (rs.attempt) {
        // The map can fill during this render: probe, rank, then trace.
        k_resv_check<<<1, 64, 0, s>>>(st, rs.cap);
        k_photon_probe<<<g, kTraceBlock, stackBytes(S), s>>>(S, st, in, depFlags, tp);
        const uint32_t* qCount =
            reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(st) + offsetof(DevState, qCount));
        const uint32_t nt = blocks(maxRays, kTile);
        k_tile_count<<<nt, kScanThreads, 0, s>>>(depFlags, qCount, maxRays, tiles);
        k_tile_scan<<<1, 1024, 0, s>>>(tiles, qCount, 0u, maxRays, tiles + nt, 0xFFFFFFFFu);
        k_resv_rank<<<nt, kScanThreads, 0, s>>>(depFlags, st, tiles, rs.attempt);
    }
    const size_t lds = tp.spill ? (size_t)tp.inLds * sizeof(uint2) * kTraceBlock : stackBytes(S);
    if (count)
        k_photon_trace<true><<<g, kTraceBlock, lds, s>>>(S, st, in, staged, extFlags, dep, depFlags, tp, rs);
    else
        k_photon_trace<false><<<g, kTraceBlock, lds, s>>>(S, st, in, staged, extFlags, dep, depFlags, tp, rs);
}
void launchReservoirWrite(hipStream_t s, DevState* st, uint32_t maxRays, Reservoir rs, PhotonArrays dep,
                          PhotonArrays raw) {
    if (!maxRays || !rs.attempt) return;
    k_resv_claim<<<blocks(maxRays, 256), 256, 0, s>>>(st, rs);
    k_resv_write<<<blocks(maxRays, 256), 256, 0, s>>>(st, rs, dep, raw);
}
void launchCompactRays(hipStream_t s, const uint8_t* flags, const uint32_t* count, uint32_t maxItems,
                       uint32_t* tiles, uint32_t* total, RayQueue src, RayQueue dst, uint32_t dstCap) {
    const uint32_t nt = blocks(maxItems, kTile);
    k_tile_count<<<nt, kScanThreads, 0, s>>>(flags, count, maxItems, tiles);
    k_tile_scan<<<1, 1024, 0, s>>>(tiles, count, 0u, maxItems, total, dstCap);
    k_scatter_rays<<<nt, kScanThreads, 0, s>>>(flags, count, maxItems, tiles, src, dst, dstCap);
}
void launchCompactPhotons(hipStream_t s, const uint8_t* flags, const uint32_t* count, uint32_t maxItems,
                          uint32_t* tiles, uint32_t* total, const uint32_t* base, PhotonArrays src,
                          PhotonArrays dst, uint32_t dstCap) {
    const uint32_t nt = blocks(maxItems, kTile);
    k_tile_count<<<nt, kScanThreads, 0, s>>>(flags, count, maxItems, tiles);
    k_tile_scan<<<1, 1024, 0, s>>>(tiles, count, 0u, maxItems, total, 0xFFFFFFFFu);
    k_scatter_photons<<<nt, kScanThreads, 0, s>>>(flags, count, maxItems, tiles, base, src, dst, dstCap);
}
void launchCompactEvents(hipStream_t s, const uint8_t* flags, const uint32_t* count, uint32_t maxItems,
                         uint32_t* tiles, DevState* st, EventArrays src, EventArrays pool, uint32_t poolCap) {
    const uint32_t nt = blocks(maxItems, kTile);
    uint32_t* total = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(st) + offsetof(DevState, evTotal));
    const uint32_t* base = reinterpret_cast<const uint32_t*>(reinterpret_cast<char*>(st) + offsetof(DevState, evPool));
    k_tile_count<<<nt, kScanThreads, 0, s>>>(flags, count, maxItems, tiles);
    k_tile_scan<<<1, 1024, 0, s>>>(tiles, count, 0u, maxItems, total, 0xFFFFFFFFu);
    k_scatter_events<<<nt, kScanThreads, 0, s>>>(flags, count, maxItems, tiles, base, src, pool, poolCap);
    k_event_pool_add<<<1, 64, 0, s>>>(st, poolCap);
}
void launchFinishPhotonBounce(hipStream_t s, DevState* st, uint32_t cap) {
    k_finish_photon_bounce<<<1, 64, 0, s>>>(st, cap);
}
void launchSetQueue(hipStream_t s, DevState* st, const uint32_t* src) { k_set_queue<<<1, 64, 0, s>>>(st, src); }

void launchBuildGrid(hipStream_t s, const GridParams& g, DevState* st, uint32_t cap, PhotonArrays raw,
                     PhotonArrays sorted, uint32_t* cellId, uint32_t* sortedIdx, uint32_t* cellCount,
                     uint32_t* cellStart, uint32_t* tiles, uint32_t* scratchTotal) {
    const uint32_t pb = blocks(cap, 256);
    k_cell_count<<<pb, 256, 0, s>>>(g, st, cap, raw.posPow, cellId, cellCount);
    const uint32_t nt = blocks(g.total, kTile);
    k_scan_local<<<nt, kScanThreads, 0, s>>>(cellCount, cellStart, g.total, tiles);
    k_tile_scan<<<1, 1024, 0, s>>>(tiles, nullptr, g.total, g.total, scratchTotal, 0xFFFFFFFFu);
    k_scan_add<<<blocks(g.total, 256), 256, 0, s>>>(cellStart, g.total, tiles, scratchTotal);
    k_cell_scatter<<<pb, 256, 0, s>>>(st, cap, cellId, cellStart, cellCount, sortedIdx);
    // cellId is free once the photons are scattered: it holds the list of
    // large cells.
    launchFill(s, reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(st) + offsetof(DevState, bigCells)), 1, 0u);
    k_cell_sort<<<blocks(g.total, 256), 256, 0, s>>>(g.total, cellStart, sortedIdx, st, cellId);
    k_cell_sort_big<<<min(pb, 2048u), 256, 0, s>>>(st, cellId, cellStart, sortedIdx);
    k_cell_payload<<<pb, 256, 0, s>>>(st, cap, sortedIdx, raw, sorted);
}

// Persistent grid of the camera trace: as many workgroups as fit on the device.
// Renderers on several host threads share the cache.
template <bool kCount, typename Map, bool kTopLds>
static uint32_t camTraceGrid(size_t lds) {
    static std::mutex lock;
    static uint32_t cached[2] = {0, 0};
    static size_t cachedLds[2] = {0, 0};
    const int k = kCount ? 1 : 0;
    std::lock_guard<std::mutex> guard(lock);
    if (cached[k] && cachedLds[k] == lds) return cached[k];
    int dev = 0, cus = 0, perCu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, k_cam_trace<kCount, Map, kTopLds>, kTraceBlock, lds);
    cached[k] = (uint32_t)std::max(1, cus) * (uint32_t)std::max(1, perCu);
    cachedLds[k] = lds;
    return cached[k];
}

void launchExclusiveScan(hipStream_t s, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tiles,
                         uint32_t* scratchTotal) {
    if (!n) return;
    const uint32_t nt = blocks(n, kTile);
    k_scan_local<<<nt, kScanThreads, 0, s>>>(in, out, n, tiles);
    k_tile_scan<<<1, 1024, 0, s>>>(tiles, nullptr, n, n, scratchTotal, 0xFFFFFFFFu);
    k_scan_add<<<blocks(n, 256), 256, 0, s>>>(out, n, tiles, scratchTotal);
}

void launchCamInit(hipStream_t s, const DevCamera& cam, const DevState* st, uint32_t pass, int decorrelate,
                   PixelArrays px) {
    const uint32_t npx = cam.W * cam.H;
    if (npx) k_cam_init<<<blocks(npx, 256), 256, 0, s>>>(cam, st, pass, decorrelate, px);
}

// LDS of one layer-walk workgroup: the in-LDS part of its lanes' stacks.
static size_t walkLdsBytes(const DevScene& S, const uint2* spill, uint32_t inLds) {
    return spill ? (size_t)inLds * sizeof(uint2) * kTraceBlock : stackBytes(S);
}

size_t walkSpillBytes(const DevScene& S, uint32_t inLds) {
    if (S.stackDepth <= inLds) return 0;
    const size_t lds = (size_t)inLds * sizeof(uint2) * kTraceBlock;
    const uint32_t grid =
        std::max(std::max(std::max(camTraceGrid<false, BatchMap, false>(lds), camTraceGrid<true, BatchMap, false>(lds)),
                          std::max(camTraceGrid<false, ImageMap, false>(lds), camTraceGrid<true, ImageMap, false>(lds))),
                 std::max(std::max(camTraceGrid<false, BatchMap, true>(lds), camTraceGrid<true, BatchMap, true>(lds)),
                          std::max(camTraceGrid<false, ImageMap, true>(lds), camTraceGrid<true, ImageMap, true>(lds))));
    return (size_t)(S.stackDepth - inLds) * grid * kTraceBlock * sizeof(uint2);
}

void launchCamTrace(hipStream_t s, const DevScene& S, DevState* st, uint32_t npx, uint32_t width, PixelArrays px,
                    EntryArrays ent, uint32_t cap, uint32_t* cursor, uint2* spill, uint32_t inLds, bool count) {
    if (!npx) return;
    inLds = std::max(1u, inLds);
    if (S.stackDepth <= inLds) spill = nullptr;
    const size_t lds = walkLdsBytes(S, spill, inLds);
    const ImageMap map{width, npx / width, (width + 7u) / 8u};
    const bool top = topFitsLds(S);
    if (count) {
        const uint32_t g = min(top ? camTraceGrid<true, ImageMap, true>(lds) : camTraceGrid<true, ImageMap, false>(lds),
                               blocks(npx, kTraceBlock));
        if (top)
            k_cam_trace<true, ImageMap, true><<<g, kTraceBlock, lds, s>>>(S, st, map, px, ent, cap, cursor, spill, inLds);
        else
            k_cam_trace<true, ImageMap, false><<<g, kTraceBlock, lds, s>>>(S, st, map, px, ent, cap, cursor, spill, inLds);
    } else {
        const uint32_t g = min(top ? camTraceGrid<false, ImageMap, true>(lds) : camTraceGrid<false, ImageMap, false>(lds),
                               blocks(npx, kTraceBlock));
        if (top)
            k_cam_trace<false, ImageMap, true><<<g, kTraceBlock, lds, s>>>(S, st, map, px, ent, cap, cursor, spill, inLds);
        else
            k_cam_trace<false, ImageMap, false><<<g, kTraceBlock, lds, s>>>(S, st, map, px, ent, cap, cursor, spill, inLds);
    }
}

void launchBatchInit(hipStream_t s, BatchDesc b, uint32_t maxTiles, PixelArrays px) {
    if (maxTiles) k_batch_init<<<blocks(maxTiles, 4), 256, 0, s>>>(b, px);
}

void launchCamTraceBatch(hipStream_t s, const DevScene& S, DevState* st, BatchDesc b, uint32_t maxSlots,
                         PixelArrays px, EntryArrays ent, uint32_t cap, uint32_t* cursor, uint2* spill,
                         uint32_t inLds, float gridFraction, bool count) {
    if (!maxSlots) return;
    inLds = std::max(1u, inLds);
    if (S.stackDepth <= inLds) spill = nullptr;
    const size_t lds = walkLdsBytes(S, spill, inLds);
    const BatchMap map{b.tiles, b.segs, b.hdr};
    const bool top = topFitsLds(S);
    if (count) {
        const uint32_t full = top ? camTraceGrid<true, BatchMap, true>(lds) : camTraceGrid<true, BatchMap, false>(lds);
        const uint32_t g = min(max(1u, (uint32_t)((float)full * gridFraction)), blocks(maxSlots, kTraceBlock));
        if (top)
            k_cam_trace<true, BatchMap, true><<<g, kTraceBlock, lds, s>>>(S, st, map, px, ent, cap, cursor, spill, inLds);
        else
            k_cam_trace<true, BatchMap, false><<<g, kTraceBlock, lds, s>>>(S, st, map, px, ent, cap, cursor, spill, inLds);
    } else {
        const uint32_t full = top ? camTraceGrid<false, BatchMap, true>(lds) : camTraceGrid<false, BatchMap, false>(lds);
        const uint32_t g = min(max(1u, (uint32_t)((float)full * gridFraction)), blocks(maxSlots, kTraceBlock));
        if (top)
            k_cam_trace<false, BatchMap, true><<<g, kTraceBlock, lds, s>>>(S, st, map, px, ent, cap, cursor, spill, inLds);
        else
            k_cam_trace<false, BatchMap, false><<<g, kTraceBlock, lds, s>>>(S, st, map, px, ent, cap, cursor, spill, inLds);
    }
}

void launchCamCompositeBatch(hipStream_t s, uint32_t maxSlots, const BatchHeader* hdr, PixelArrays px,
                             EntryArrays ent) {
    if (maxSlots)
        k_cam_composite<<<blocks(maxSlots, 256), 256, 0, s>>>(maxSlots, &hdr->nSlots, px, ent, 0.0f, nullptr);
}

void launchBatchResolve(hipStream_t s, BatchDesc b, uint32_t maxResolve, PixelArrays px, float4* fb) {
    if (maxResolve) k_batch_resolve<<<blocks(maxResolve, 256), 256, 0, s>>>(b, px, fb);
}

void launchEntrySort(hipStream_t s, const GridParams& g, DevState* st, uint32_t cap, PixelArrays px,
                     EntryArrays ent, uint32_t* cellCount, uint32_t* cellStart, uint32_t* tiles,
                     uint32_t* scratchTotal) {
    if (!cap) return;
    const uint32_t gb = min(blocks(cap, 256), kEntryBlocks);
    k_entry_key<<<gb, 256, 0, s>>>(g, st, cap, px, ent, cellCount);
    launchExclusiveScan(s, cellCount, cellStart, g.total, tiles, scratchTotal);
    k_entry_scatter<<<gb, 256, 0, s>>>(st, cap, cellStart, cellCount, ent);
}

void launchEntryGather(hipStream_t s, const DevGrid& G, DevState* st, uint32_t cap, const uint32_t* nGather,
                       EntryArrays ent, bool count) {
    if (!cap) return;
    if (count)
        k_entry_gather<true><<<blocks(cap, 256), 256, 0, s>>>(G, st, nGather, ent);
    else
        k_entry_gather<false><<<blocks(cap, 256), 256, 0, s>>>(G, st, nGather, ent);
}

void launchCamComposite(hipStream_t s, uint32_t npx, PixelArrays px, EntryArrays ent, float invSpp, float4* fb) {
    if (npx) k_cam_composite<<<blocks(npx, 256), 256, 0, s>>>(npx, nullptr, px, ent, invSpp, fb);
}

void launchPostprocess(hipStream_t s, const float4* fb, float4* ldr, uint32_t n) {
    if (n) k_post<<<blocks(n, 256), 256, 0, s>>>(fb, ldr, n);
}
void launchAdjRaygenBatch(hipStream_t s, BatchDesc b, uint32_t maxTiles, const float* dloss, RayQueue q,
                          PendingArrays pend, DevState* st) {
    k_adj_raygen_batch<<<blocks(max(maxTiles, 1u), 4), 256, 0, s>>>(b, dloss, q, pend, st);
}
void launchSegFirst(hipStream_t s, const DevState* st, uint32_t maxRays, RayQueue q, uint32_t* segFirst,
                    uint32_t* segEnd, uint32_t nSegs, uint4* groups, uint32_t nGroups, const uint32_t* groupSegs) {
    if (!maxRays) return;
    k_seg_clear<<<blocks(max(nSegs, 1u), 256), 256, 0, s>>>(segFirst, segEnd, nSegs);
    k_seg_first<<<blocks(maxRays, 256), 256, 0, s>>>(st, q, segFirst, segEnd);
    k_group_lanes<<<1, 64, 0, s>>>(segFirst, segEnd, groups, nGroups, groupSegs);
}
void launchAdjRaygen(hipStream_t s, const DevCamera& cam, const float* dloss, RayQueue q, PendingArrays pend,
                     DevState* st, uint32_t seed) {
    const uint32_t npx = cam.W * cam.H;
    k_adj_raygen<<<blocks(npx > 0 ? npx : 1, 256), 256, 0, s>>>(cam, dloss, q, pend, st, seed);
}
void launchAdjTrace(hipStream_t s, const DevScene& S, DevState* st, uint32_t maxRays, uint32_t bounce, RayQueue in,
                    RayQueue staged, uint8_t* extFlags, EventArrays ev, uint8_t* evFlags, PendingArrays pend,
                    TraceParams tp, bool count) {
    if (!maxRays) return;
    const dim3 g(blocks(maxRays, kTraceBlock));
    const size_t lds = tp.spill ? (size_t)tp.inLds * sizeof(uint2) * kTraceBlock : stackBytes(S);
    if (count)
        k_adj_trace<true><<<g, kTraceBlock, lds, s>>>(S, st, bounce, in, staged, extFlags, ev, evFlags, pend,
                                                      maxRays, tp);
    else
        k_adj_trace<false><<<g, kTraceBlock, lds, s>>>(S, st, bounce, in, staged, extFlags, ev, evFlags, pend,
                                                       maxRays, tp);
}
void launchEventSort(hipStream_t s, const GridParams& g, DevState* st, uint32_t poolCap, EventArrays ev,
                     uint32_t* cell, uint32_t* sorted, uint32_t* cellCount, uint32_t* cellStart, uint32_t* tiles,
                     uint32_t* scratchTotal) {
    if (!poolCap) return;
    k_event_key<<<min(blocks(poolCap, 256), kEntryBlocks), 256, 0, s>>>(g, st, poolCap, ev, cell, cellCount);
    launchExclusiveScan(s, cellCount, cellStart, g.total, tiles, scratchTotal);
    k_event_scatter<<<min(blocks(poolCap, 256), kEntryBlocks), 256, 0, s>>>(st, poolCap, cell, cellStart, cellCount, sorted);
}
void launchAdjTransport(hipStream_t s, const DevScene& S, const DevGrid& G, DevState* st, uint32_t poolCap,
                        EventArrays ev, const uint32_t* sorted, float* gradOpacity, float* gradFull, float* debugImg,
                        bool count) {
    if (!poolCap) return;
    if (count)
        k_adj_transport<true><<<blocks(poolCap, 256), 256, 0, s>>>(S, G, st, poolCap, ev, sorted, gradOpacity,
                                                                  gradFull, debugImg);
    else
        k_adj_transport<false><<<blocks(poolCap, 256), 256, 0, s>>>(S, G, st, poolCap, ev, sorted,
                                                                   gradOpacity, gradFull, debugImg);
}
void launchSplitGradients(hipStream_t s, const float* gradFull, uint32_t n, float* position, float* tangentU,
                          float* tangentV, float* scale, float* albedo, float* beta) {
    if (n) k_split_gradients<<<blocks(n, 256), 256, 0, s>>>(gradFull, n, position, tangentU, tangentV, scale, albedo,
                                                            beta);
}
void launchTraceRays(hipStream_t s, const DevScene& S, uint32_t n, const float* o, const float* d, int32_t* inst,
                     int32_t* prim, float* t, float* pos, float* alpha) {
    if (n)
        k_trace_rays<<<blocks(n, kTraceBlock), kTraceBlock, stackBytes(S), s>>>(S, n, o, d, inst, prim, t, pos,
                                                                               alpha);
}

void launchGatherPoints(hipStream_t s, const DevGrid& G, uint32_t n, const float* q, const float* nrm, float* out) {
    if (n) k_gather_points<<<blocks(n, 256), 256, 0, s>>>(G, n, q, nrm, out);
}

void launchSetOpacity(hipStream_t s, SurfelShade* shade, const float* opacity, uint32_t n) {
    if (n) k_set_opacity<<<blocks(n, 256), 256, 0, s>>>(shade, opacity, n);
}

void launchApplyPoints(hipStream_t s, SurfelGeo* geo, SurfelShade* shade, float* raw, DevPointArrays a, uint32_t n) {
    if (n) k_apply_points<<<blocks(n, 256), 256, 0, s>>>(geo, shade, raw, a, n);
}

}  // namespace pale
