// pale_bvh.hip -- the surfel BVH built on the device, node for node and
// permutation for permutation equal to the host build (BVH.ixx:49-133 with
// std::nth_element at count/2 on the centroid of the longest node axis,
// children allocated as an adjacent pair, depth-first left-then-right).
//
// The tree's shape depends only on the surfel count: a node of c > 4
// surfels splits into c/2 (rounded down) and the rest.  So the node
// numbering, every node's range and the wide-node layout are computed once
// per count on the host; the device computes what depends on the surfels:
// the node boxes and the element order that std::nth_element leaves.
//
// Kernels, in launch order:
//   k_bvh_prims     surfel boxes (surfelObjectAabbBeta), keys, identity order
//   per level of nodes with more than kSubtree surfels:
//     k_bvh_fit     node boxes: chunked min/max, combined with atomics on
//                   order-preserving integer images of the floats
//     k_bvh_select  one workgroup per node: the node box, its longest axis,
//                   the keys on that axis, then libstdc++'s introselect with
//                   each partition done by the whole workgroup (ranks of the
//                   Hoare scan pairs, pale_select.h)
//   k_bvh_subtrees  one thread per node of at most kSubtree surfels: its
//                   whole subtree, sequentially (pale_select.h selectSeq)
//   k_bvh_wide      the 64-byte wide nodes (both child boxes + references)
//   k_bvh_records   the surfel records in the new order
#include "pale_kernels.h"
#include "pale_scene.h"
#include "pale_select.h"

namespace pale {

namespace {

// A selecting workgroup has 1024, 256 or 64 threads, by the size of the
// level's nodes (more threads for the few large nodes of the top levels).
constexpr uint32_t kSelItems = 8;       // consecutive elements per thread and chunk
constexpr uint32_t kSelSeq = 32;        // ranges this short finish on one thread in LDS
constexpr uint32_t kSwapIlp = 4;        // swaps per thread and step (independent loads in flight)
constexpr uint32_t kFitBlock = 256;

// Order-preserving image of a binary32 value in uint32 (for finite values
// and infinities; -0 sorts below +0, which the build rules out).
__device__ __forceinline__ uint32_t ordOf(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fromOrd(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

__device__ __forceinline__ int longestAxis(F3 mn, F3 mx) {
    const F3 e = mx - mn;
    return (e.y > e.x && e.y > e.z) ? 1 : (e.z > e.x && e.z > e.y) ? 2 : 0;
}

__device__ __forceinline__ const float* keysOf(const float* kx, const float* ky, const float* kz, int axis) {
    return axis == 0 ? kx : axis == 1 ? ky : kz;
}

__device__ __forceinline__ uint32_t waveInclusiveScan(uint32_t x) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    return x;
}

__device__ __forceinline__ uint32_t waveSum(uint32_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// Exclusive prefix of x over the workgroup (thread order) and the total.
// sw: kBlock / 64 + 1 words of LDS.
template <uint32_t kBlock>
__device__ uint32_t blockExclusiveScan(uint32_t x, uint32_t& total, uint32_t* sw) {
    constexpr uint32_t nw = kBlock / 64;
    const uint32_t incl = waveInclusiveScan(x);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 63u) sw[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t i = 0; i < nw; ++i) {
            const uint32_t t = sw[i];
            sw[i] = run;
            run += t;
        }
        sw[nw] = run;
    }
    __syncthreads();
    const uint32_t r = sw[w] + incl - x;
    total = sw[nw];
    __syncthreads();
    return r;
}

template <uint32_t kBlock>
__device__ uint32_t blockSum(uint32_t x, uint32_t* sw) {
    uint32_t total;
    (void)blockExclusiveScan<kBlock>(x, total, sw);
    return total;
}

__global__ __launch_bounds__(256) void k_bvh_prims(const HostPoint* pts, uint32_t n, float* kx, float* ky, float* kz,
                                                   float* boxes, uint32_t* ord, uint32_t* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const HostPoint s = pts[i];
    F3 mn, mx;
    surfelBoxOf(s, mn, mx);
    const float b[6] = {mn.x, mn.y, mn.z, mx.x, mx.y, mx.z};
    bool odd = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        boxes[6 * (size_t)i + c] = b[c];
        // non-finite values or -0 would make min/max order-dependent or the
        // selection ill-defined: such a scene is built on the host
        odd |= !isfinite(b[c]) || __float_as_uint(b[c]) == 0x80000000u;
    }
    if (odd) atomicOr(flags, 1u);
    kx[i] = s.p.x;
    ky[i] = s.p.y;
    kz[i] = s.p.z;
    ord[i] = i;
}

__global__ __launch_bounds__(256) void k_bvh_init_fit(uint32_t* acc, uint32_t nNodes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 6 * nNodes) return;
    acc[i] = (i % 6u) < 3u ? ordOf(__builtin_inff()) : ordOf(-__builtin_inff());
}

// chunks[k] = {node slot in acc, first, end, -}: min / max of the boxes of
// the elements ord[first..end) folded with fminf / fmaxf (as the host fit),
// then combined per node.
__global__ __launch_bounds__(kFitBlock) void k_bvh_fit(const uint4* chunks, const uint32_t* ord, const float* boxes,
                                                      uint32_t* acc) {
    const uint4 c = chunks[blockIdx.x];
    float v[6] = {__builtin_inff(), __builtin_inff(), __builtin_inff(), -__builtin_inff(), -__builtin_inff(),
                  -__builtin_inff()};
    for (uint32_t p = c.y + threadIdx.x; p < c.z; p += kFitBlock) {
        const float* b = boxes + 6 * (size_t)ord[p];
        v[0] = fminf(v[0], b[0]);
        v[1] = fminf(v[1], b[1]);
        v[2] = fminf(v[2], b[2]);
        v[3] = fmaxf(v[3], b[3]);
        v[4] = fmaxf(v[4], b[4]);
        v[5] = fmaxf(v[5], b[5]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v[0] = fminf(v[0], __shfl_xor(v[0], o, 64));
        v[1] = fminf(v[1], __shfl_xor(v[1], o, 64));
        v[2] = fminf(v[2], __shfl_xor(v[2], o, 64));
        v[3] = fmaxf(v[3], __shfl_xor(v[3], o, 64));
        v[4] = fmaxf(v[4], __shfl_xor(v[4], o, 64));
        v[5] = fmaxf(v[5], __shfl_xor(v[5], o, 64));
    }
    if ((threadIdx.x & 63u) == 0) {
        uint32_t* a = acc + 6 * (size_t)c.x;
        atomicMin(a + 0, ordOf(v[0]));
        atomicMin(a + 1, ordOf(v[1]));
        atomicMin(a + 2, ordOf(v[2]));
        atomicMax(a + 3, ordOf(v[3]));
        atomicMax(a + 4, ordOf(v[4]));
        atomicMax(a + 5, ordOf(v[5]));
    }
}

// Swaps m0 <= m < m1 exchange positions lp[m] and rp[m] (keys and
// indices); each thread takes kSwapIlp swaps at a time, loads before stores.
template <uint32_t kBlock>
__device__ void swapPairs(float* kv, uint32_t* ord, const uint32_t* lp, const uint32_t* rp, uint32_t m0, uint32_t m1) {
    for (uint32_t m = m0 + threadIdx.x * kSwapIlp; m < m1; m += kBlock * kSwapIlp) {
        uint32_t pa[kSwapIlp], pb[kSwapIlp], ia[kSwapIlp], ib[kSwapIlp];
        float ka[kSwapIlp], kb[kSwapIlp];
#pragma unroll
        for (uint32_t j = 0; j < kSwapIlp; ++j)
            if (m + j < m1) {
                pa[j] = lp[m + j];
                pb[j] = rp[m + j];
            }
#pragma unroll
        for (uint32_t j = 0; j < kSwapIlp; ++j)
            if (m + j < m1) {
                ka[j] = kv[pa[j]];
                kb[j] = kv[pb[j]];
                ia[j] = ord[pa[j]];
                ib[j] = ord[pb[j]];
            }
#pragma unroll
        for (uint32_t j = 0; j < kSwapIlp; ++j)
            if (m + j < m1) {
                kv[pa[j]] = kb[j];
                kv[pb[j]] = ka[j];
                ord[pa[j]] = ib[j];
                ord[pb[j]] = ia[j];
            }
    }
}

// nodes[k] = {first, count, ref, slot in acc}: one workgroup per node.
//
// One introselect step's partition of [lo + 1, hi) around P, by ranks
// (pale_select.h): a position p with key >= P and G(p) such positions left
// of it is the left partner of swap G(p) iff E(p) > G(p), E(p) = keys <= P
// right of p; a position q with key <= P is the right partner of swap E(q)
// iff G(q) > E(q).  The first position with key >= P that is not swapped is
// L_K, and the cut is min(L_K, R_{K-1}).  Pass A counts the keys <= P; pass
// B walks the range in chunks of kSelChunk (one barrier each: wave scans,
// wave totals double-buffered in LDS) and records the partners of each swap;
// pass C swaps, kSwapIlp swaps per thread at a time.
template <uint32_t kSelBlock>
__global__ __launch_bounds__(kSelBlock) void k_bvh_select(const uint4* nodes, const uint32_t* acc, const float* kx,
                                                         const float* ky, const float* kz, float* kv, uint32_t* ord,
                                                         uint32_t* lpos, uint32_t* rpos, float* refBox,
                                                         const BvhWideState* resume) {
    constexpr uint32_t kSelWaves = kSelBlock / 64;
    constexpr uint32_t kSelChunk = kSelBlock * kSelItems;
    __shared__ uint32_t sw[kSelWaves + 1];
    __shared__ uint32_t sTot[2][kSelWaves];
    __shared__ uint32_t sLK, sK;
    __shared__ float sP;
    __shared__ float sK32[kSelSeq];
    __shared__ uint32_t sI32[kSelSeq];
    const uint4 nd = nodes[blockIdx.x];
    const uint32_t first = nd.x, count = nd.y, tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t lo = first, hi = first + count;
    const uint32_t nth = first + (count >> 1);
    int depth = 2 * sel::floorLog2(count);
    if (resume) {
        // the state the workgroups of k_wsel_* left (box, keys and the first
        // introselect steps done)
        const BvhWideState w = resume[blockIdx.x];
        lo = w.lo;
        hi = w.hi;
        depth = (int)w.depth;
    } else {
        const uint32_t* a = acc + 6 * (size_t)nd.w;
        const F3 mn = mk3(fromOrd(a[0]), fromOrd(a[1]), fromOrd(a[2]));
        const F3 mx = mk3(fromOrd(a[3]), fromOrd(a[4]), fromOrd(a[5]));
        if (tid < 6) refBox[6 * (size_t)nd.z + tid] = tid < 3 ? (&mn.x)[tid] : (&mx.x)[tid - 3];
        const float* key = keysOf(kx, ky, kz, longestAxis(mn, mx));
        for (uint32_t p = first + tid; p < first + count; p += kSelBlock) kv[p] = key[ord[p]];
        __syncthreads();
    }
    while (true) {
        const uint32_t len = hi - lo;
        if (len <= kSelSeq) {
            // the rest on one thread, in LDS (introselect is translation
            // invariant: positions relative to lo)
            if (tid < len) {
                sK32[tid] = kv[lo + tid];
                sI32[tid] = ord[lo + tid];
            }
            __syncthreads();
            if (tid == 0) sel::selectFrom(sK32, sI32, 0, nth - lo, len, depth);
            __syncthreads();
            if (tid < len) {
                kv[lo + tid] = sK32[tid];
                ord[lo + tid] = sI32[tid];
            }
            break;
        }
        if (depth == 0) {
            if (tid == 0) sel::selectFrom(kv, ord, lo, nth, hi, 0);  // heap selection
            break;
        }
        --depth;
        if (tid == 0) {
            sel::medianToFirst(kv, ord, lo, lo + 1, lo + len / 2, hi - 1);
            sP = kv[lo];
            sLK = hi;
            sK = 0;
        }
        __syncthreads();
        const float P = sP;
        // pass A
        uint32_t cnt = 0;
        {
            uint32_t p = lo + 1 + tid;
            for (; p + 3 * kSelBlock < hi; p += 4 * kSelBlock) {
                const float k0 = kv[p], k1 = kv[p + kSelBlock], k2 = kv[p + 2 * kSelBlock], k3 = kv[p + 3 * kSelBlock];
                cnt += (!(P < k0) ? 1u : 0u) + (!(P < k1) ? 1u : 0u) + (!(P < k2) ? 1u : 0u) + (!(P < k3) ? 1u : 0u);
            }
            for (; p < hi; p += kSelBlock) cnt += !(P < kv[p]) ? 1u : 0u;
        }
        const uint32_t nLE = blockSum<kSelBlock>(cnt, sw);
        // pass B
        uint32_t carryG = 0, carryL = 0, kacc = 0, lk = hi, par = 0;
        for (uint32_t base = lo + 1; base < hi; base += kSelChunk, par ^= 1u) {
            const uint32_t b0 = base + tid * kSelItems;
            float k[kSelItems];
            uint32_t g = 0, l = 0;
#pragma unroll
            for (uint32_t j = 0; j < kSelItems; ++j) {
                const bool in = b0 + j < hi;
                k[j] = in ? kv[b0 + j] : 0.0f;
                g += (in && !(k[j] < P)) ? 1u : 0u;
                l += (in && !(P < k[j])) ? 1u : 0u;
            }
            const uint32_t x = g | (l << 16);
            const uint32_t incl = waveInclusiveScan(x);
            if (lane == 63u) sTot[par][wave] = incl;
            __syncthreads();
            uint32_t before = 0, tot = 0;
#pragma unroll
            for (uint32_t w = 0; w < kSelWaves; ++w) {
                const uint32_t t = sTot[par][w];
                before += w < wave ? t : 0u;
                tot += t;
            }
            const uint32_t pre = before + incl - x;
            uint32_t G = carryG + (pre & 0xFFFFu), Lc = carryL + (pre >> 16);
#pragma unroll
            for (uint32_t j = 0; j < kSelItems; ++j) {
                if (b0 + j >= hi) break;
                const bool ge = !(k[j] < P), le = !(P < k[j]);
                if (le) ++Lc;
                const uint32_t E = nLE - Lc;  // keys <= P right of this position
                if (ge) {
                    if (E > G) {
                        lpos[first + G] = b0 + j;
                        ++kacc;
                    } else {
                        lk = min(lk, b0 + j);
                    }
                }
                if (le && G > E) rpos[first + E] = b0 + j;
                if (ge) ++G;
            }
            carryG += tot & 0xFFFFu;
            carryL += tot >> 16;
        }
        if (lk < hi) atomicMin(&sLK, lk);
        if (kacc) atomicAdd(&sK, kacc);
        __syncthreads();  // also orders the lpos / rpos writes before their reads
        const uint32_t K = sK;
        uint32_t cut = sLK;
        if (K > 0) cut = min(cut, rpos[first + K - 1]);
        // pass C
        swapPairs<kSelBlock>(kv, ord, lpos + first, rpos + first, 0, K);
        __syncthreads();
        if (cut <= nth)
            lo = cut;
        else
            hi = cut;
    }
}

// The introselect steps of large nodes spread over many workgroups (levels
// whose nodes exceed wideMin surfels, where one workgroup per node would
// leave most of the chip idle).  The node's range is cut into tiles of
// tileSize positions; per step: k_wsel_next (one thread per node: the cut
// of the step before, then the median of three moved to the front as the
// pivot), k_wsel_count (keys >= P and <= P per tile), k_wsel_rank (each
// tile adds up the counts of the node's earlier tiles, ranks its positions
// and records the swap partners, as k_bvh_select does over the whole
// range), k_wsel_swap.  A node whose range has come down to wideMin, or
// whose depth limit is reached, stops; k_bvh_select then continues from its
// state.  tiles[t] = {node (index in the level), first, end, tile index in
// node}; a node's tiles are consecutive.
constexpr uint32_t kWideBlock = 256;
constexpr uint32_t kWideChunk = kWideBlock * kSelItems;
constexpr uint32_t kSwapSplit = 4;  // swap workgroups per tile

__global__ __launch_bounds__(kWideBlock) void k_wsel_init(const uint4* nodes, const uint4* tiles, const uint32_t* acc,
                                                         const float* kx, const float* ky, const float* kz,
                                                         float* kv, const uint32_t* ord, float* refBox,
                                                         BvhWideState* ws) {
    const uint4 t = tiles[blockIdx.x];
    const uint4 nd = nodes[t.x];
    const uint32_t* a = acc + 6 * (size_t)nd.w;
    const F3 mn = mk3(fromOrd(a[0]), fromOrd(a[1]), fromOrd(a[2]));
    const F3 mx = mk3(fromOrd(a[3]), fromOrd(a[4]), fromOrd(a[5]));
    const float* key = keysOf(kx, ky, kz, longestAxis(mn, mx));
    for (uint32_t p = t.y + threadIdx.x; p < t.z; p += kWideBlock) kv[p] = key[ord[p]];
    if (t.w == 0) {
        if (threadIdx.x < 6) refBox[6 * (size_t)nd.z + threadIdx.x] = threadIdx.x < 3 ? (&mn.x)[threadIdx.x]
                                                                                         : (&mx.x)[threadIdx.x - 3];
        if (threadIdx.x == 0) {
            BvhWideState w{};
            w.lo = nd.x;
            w.hi = nd.x + nd.y;
            w.depth = (uint32_t)(2 * sel::floorLog2(nd.y));
            w.active = 1u;
            ws[t.x] = w;
        }
    }
}

__global__ __launch_bounds__(64) void k_wsel_next(const uint4* nodes, uint32_t nNodes, float* kv, uint32_t* ord,
                                                 BvhWideState* ws, const uint32_t* rpos, uint32_t wideMin,
                                                 uint32_t cut, uint32_t pivot) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nNodes) return;
    BvhWideState s = ws[w];
    if (!s.active) return;
    if (cut) {
        const uint4 nd = nodes[w];
        const uint32_t nth = nd.x + (nd.y >> 1);
        uint32_t c = s.LK;
        if (s.K > 0) c = min(c, rpos[nd.x + s.K - 1]);
        if (c <= nth)
            s.lo = c;
        else
            s.hi = c;
    }
    if (pivot) {
        if (s.hi - s.lo <= wideMin || s.depth == 0) {
            s.active = 0u;
        } else {
            --s.depth;
            sel::medianToFirst(kv, ord, s.lo, s.lo + 1, s.lo + (s.hi - s.lo) / 2, s.hi - 1);
            s.P = kv[s.lo];
            s.K = 0u;
            s.LK = s.hi;
        }
    }
    ws[w] = s;
}

__global__ __launch_bounds__(kWideBlock) void k_wsel_count(const uint4* tiles, const float* kv,
                                                          const BvhWideState* ws, uint2* tileCnt) {
    __shared__ uint32_t sw[kWideBlock / 64 + 1];
    const uint4 t = tiles[blockIdx.x];
    const BvhWideState s = ws[t.x];
    if (!s.active) return;
    const uint32_t a = max(t.y, s.lo + 1), b = min(t.z, s.hi);
    const float P = s.P;
    uint32_t g = 0, l = 0, p = a + threadIdx.x;
    for (; p + 7 * kWideBlock < b; p += 8 * kWideBlock) {
        float k[8];
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) k[j] = kv[p + j * kWideBlock];
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
            g += !(k[j] < P) ? 1u : 0u;
            l += !(P < k[j]) ? 1u : 0u;
        }
    }
    for (; p < b; p += kWideBlock) {
        const float k = kv[p];
        g += !(k < P) ? 1u : 0u;
        l += !(P < k) ? 1u : 0u;
    }
    const uint32_t tot = blockSum<kWideBlock>(g | (l << 16), sw);  // a tile holds fewer than 2^16 keys
    if (threadIdx.x == 0) tileCnt[blockIdx.x] = make_uint2(tot & 0xFFFFu, tot >> 16);
}

__global__ __launch_bounds__(kWideBlock) void k_wsel_rank(const uint4* tiles, const uint4* nodes, const float* kv,
                                                         BvhWideState* ws, const uint2* tileCnt, uint32_t tileSize,
                                                         uint32_t* lpos, uint32_t* rpos) {
    __shared__ uint32_t sTot[2][kWideBlock / 64];
    __shared__ uint32_t sRed[3][kWideBlock / 64];
    const uint4 t = tiles[blockIdx.x];
    const BvhWideState s = ws[t.x];
    if (!s.active) return;
    const uint4 nd = nodes[t.x];
    const uint32_t first = nd.x, lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    // keys >= P and <= P in the node's tiles before this one, and keys <= P
    // in all of them
    uint32_t pg = 0, pl = 0, al = 0;
    {
        const uint32_t t0 = blockIdx.x - t.w, nt = (nd.y + tileSize - 1) / tileSize;
        for (uint32_t i = threadIdx.x; i < nt; i += kWideBlock) {
            const uint2 c = tileCnt[t0 + i];
            pg += i < t.w ? c.x : 0u;
            pl += i < t.w ? c.y : 0u;
            al += c.y;
        }
        pg = waveSum(pg);
        pl = waveSum(pl);
        al = waveSum(al);
        if (lane == 0) {
            sRed[0][wave] = pg;
            sRed[1][wave] = pl;
            sRed[2][wave] = al;
        }
        __syncthreads();
        pg = pl = al = 0;
#pragma unroll
        for (uint32_t q = 0; q < kWideBlock / 64; ++q) {
            pg += sRed[0][q];
            pl += sRed[1][q];
            al += sRed[2][q];
        }
    }
    const uint32_t a = max(t.y, s.lo + 1), b = min(t.z, s.hi);
    const float P = s.P;
    const uint32_t nLE = al;
    uint32_t carryG = pg, carryL = pl, kacc = 0, lk = 0xFFFFFFFFu, par = 0;
    for (uint32_t base = a; base < b; base += kWideChunk, par ^= 1u) {
        const uint32_t b0 = base + threadIdx.x * kSelItems;
        float k[kSelItems];
        uint32_t g = 0, l = 0;
#pragma unroll
        for (uint32_t j = 0; j < kSelItems; ++j) {
            const bool in = b0 + j < b;
            k[j] = in ? kv[b0 + j] : 0.0f;
            g += (in && !(k[j] < P)) ? 1u : 0u;
            l += (in && !(P < k[j])) ? 1u : 0u;
        }
        const uint32_t x = g | (l << 16);
        const uint32_t incl = waveInclusiveScan(x);
        if (lane == 63u) sTot[par][wave] = incl;
        __syncthreads();
        uint32_t before = 0, tot = 0;
#pragma unroll
        for (uint32_t q = 0; q < kWideBlock / 64; ++q) {
            const uint32_t v = sTot[par][q];
            before += q < wave ? v : 0u;
            tot += v;
        }
        const uint32_t pre = before + incl - x;
        uint32_t G = carryG + (pre & 0xFFFFu), Lc = carryL + (pre >> 16);
#pragma unroll
        for (uint32_t j = 0; j < kSelItems; ++j) {
            if (b0 + j >= b) break;
            const bool ge = !(k[j] < P), le = !(P < k[j]);
            if (le) ++Lc;
            const uint32_t E = nLE - Lc;
            if (ge) {
                if (E > G) {
                    lpos[first + G] = b0 + j;
                    ++kacc;
                } else {
                    lk = min(lk, b0 + j);
                }
            }
            if (le && G > E) rpos[first + E] = b0 + j;
            if (ge) ++G;
        }
        carryG += tot & 0xFFFFu;
        carryL += tot >> 16;
    }
    if (lk != 0xFFFFFFFFu) atomicMin(&ws[t.x].LK, lk);
    if (kacc) atomicAdd(&ws[t.x].K, kacc);
}

// kSwapSplit workgroups per tile; the swaps of a node are split over its
// tiles (a node has more tile positions than swaps)
__global__ __launch_bounds__(kWideBlock) void k_wsel_swap(const uint4* tiles, const uint4* nodes,
                                                         const BvhWideState* ws, float* kv, uint32_t* ord,
                                                         const uint32_t* lpos, const uint32_t* rpos,
                                                         uint32_t tileSize) {
    const uint4 t = tiles[blockIdx.x / kSwapSplit];
    const BvhWideState s = ws[t.x];
    if (!s.active) return;
    const uint32_t first = nodes[t.x].x, part = (tileSize + kSwapSplit - 1) / kSwapSplit;
    const uint32_t tileEnd = (t.w + 1) * tileSize;
    const uint32_t m0 = t.w * tileSize + (blockIdx.x % kSwapSplit) * part;
    const uint32_t m1 = min(min(m0 + part, tileEnd), s.K);
    swapPairs<kWideBlock>(kv, ord, lpos + first, rpos + first, m0, m1);
}

// roots[k] = {first, count, ref, preorder rank among internal nodes}: one
// thread builds the whole subtree below a node of at most kSubtree
// surfels.  The children of the internal node of rank r are the pair at
// 1 + 2r; the left child has rank r + 1, the right r + 1 + I(left count),
// I(c) = internal nodes of a subtree over c surfels.
__global__ __launch_bounds__(256) void k_bvh_subtrees(const uint4* roots, uint32_t nRoots, const uint32_t* internalOf,
                                                     const float* kx, const float* ky, const float* kz,
                                                     const float* boxes, float* kv, uint32_t* ord, float* refBox) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nRoots) return;
    uint4 st[24];
    int sp = 0;
    st[sp++] = roots[t];
    while (sp > 0) {
        const uint4 nd = st[--sp];
        const uint32_t first = nd.x, count = nd.y;
        F3 mn = mk3(__builtin_inff(), __builtin_inff(), __builtin_inff());
        F3 mx = mk3(-__builtin_inff(), -__builtin_inff(), -__builtin_inff());
        for (uint32_t p = first; p < first + count; ++p) {
            const float* b = boxes + 6 * (size_t)ord[p];
            mn = vmin(mn, mk3(b[0], b[1], b[2]));
            mx = vmax(mx, mk3(b[3], b[4], b[5]));
        }
        float* rb = refBox + 6 * (size_t)nd.z;
        rb[0] = mn.x; rb[1] = mn.y; rb[2] = mn.z;
        rb[3] = mx.x; rb[4] = mx.y; rb[5] = mx.z;
        if (count <= kMaxLeafPrims) continue;
        const float* key = keysOf(kx, ky, kz, longestAxis(mn, mx));
        for (uint32_t p = first; p < first + count; ++p) kv[p] = key[ord[p]];
        const uint32_t lc = count >> 1;
        sel::selectSeq(kv, ord, first, first + lc, first + count);
        const uint32_t L = 1 + 2 * nd.w;
        st[sp++] = make_uint4(first + lc, count - lc, L + 1, nd.w + 1 + internalOf[lc]);
        st[sp++] = make_uint4(first, lc, L, nd.w + 1);
    }
}

// wide[k] = {pair index of the children, left reference, right reference}
// for the internal node of wide index base + k; a reference without the
// leaf bit is a wide index relative to the cloud's first wide node.
__global__ __launch_bounds__(256) void k_bvh_wide(const uint4* wide, uint32_t nWide, const float* refBox,
                                                 uint32_t base, WideNode* out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nWide) return;
    const uint4 e = wide[k];
    const float* l = refBox + 6 * (size_t)e.x;
    const float* r = l + 6;
    WideNode w;
    w.lminx = l[0]; w.lminy = l[1]; w.lminz = l[2]; w.lmaxx = l[3];
    w.lmaxy = l[4]; w.lmaxz = l[5]; w.rminx = r[0]; w.rminy = r[1];
    w.rminz = r[2]; w.rmaxx = r[3]; w.rmaxy = r[4]; w.rmaxz = r[5];
    w.left = (e.y & kLeafBit) ? e.y : e.y + base;
    w.right = (e.z & kLeafBit) ? e.z : e.z + base;
    out[base + k] = w;
}

__global__ __launch_bounds__(256) void k_bvh_records(const HostPoint* pts, const uint32_t* ord, uint32_t n,
                                                    SurfelGeo* geo, SurfelShade* shade) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const HostPoint s = pts[ord[k]];
    geo[k] = surfelGeoOf(s);
    shade[k] = surfelShadeOf(s);
}

inline uint32_t nblocks(size_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

void launchBvhPrims(hipStream_t s, const HostPoint* pts, uint32_t n, float* kx, float* ky, float* kz, float* boxes,
                    uint32_t* ord, uint32_t* flags) {
    if (n) k_bvh_prims<<<nblocks(n, 256), 256, 0, s>>>(pts, n, kx, ky, kz, boxes, ord, flags);
}

void launchBvhInitFit(hipStream_t s, uint32_t* acc, uint32_t nNodes) {
    if (nNodes) k_bvh_init_fit<<<nblocks(6 * (size_t)nNodes, 256), 256, 0, s>>>(acc, nNodes);
}

void launchBvhFit(hipStream_t s, const uint4* chunks, uint32_t nChunks, const uint32_t* ord, const float* boxes,
                  uint32_t* acc) {
    if (nChunks) k_bvh_fit<<<nChunks, kFitBlock, 0, s>>>(chunks, ord, boxes, acc);
}

void launchBvhSelect(hipStream_t s, const uint4* nodes, uint32_t nNodes, uint32_t count, const uint32_t* acc,
                     const float* kx, const float* ky, const float* kz, float* kv, uint32_t* ord, uint32_t* lpos,
                     uint32_t* rpos, float* refBox, const BvhWideState* resume) {
    if (!nNodes) return;
    if (count > 32768)
        k_bvh_select<1024><<<nNodes, 1024, 0, s>>>(nodes, acc, kx, ky, kz, kv, ord, lpos, rpos, refBox, resume);
    else if (count > 2048)
        k_bvh_select<256><<<nNodes, 256, 0, s>>>(nodes, acc, kx, ky, kz, kv, ord, lpos, rpos, refBox, resume);
    else
        k_bvh_select<64><<<nNodes, 64, 0, s>>>(nodes, acc, kx, ky, kz, kv, ord, lpos, rpos, refBox, resume);
}

void launchBvhWideSteps(hipStream_t s, const uint4* nodes, uint32_t nNodes, const uint4* tiles, uint32_t nTiles,
                        uint32_t tileSize, uint32_t steps, uint32_t wideMin, const uint32_t* acc, const float* kx,
                        const float* ky, const float* kz, float* kv, uint32_t* ord, uint32_t* lpos, uint32_t* rpos,
                        float* refBox, BvhWideState* ws, uint2* tileCnt) {
    if (!nNodes || !nTiles) return;
    const uint32_t nb = nblocks(nNodes, 64);
    k_wsel_init<<<nTiles, kWideBlock, 0, s>>>(nodes, tiles, acc, kx, ky, kz, kv, ord, refBox, ws);
    for (uint32_t k = 0; k < steps; ++k) {
        k_wsel_next<<<nb, 64, 0, s>>>(nodes, nNodes, kv, ord, ws, rpos, wideMin, k > 0 ? 1u : 0u, 1u);
        k_wsel_count<<<nTiles, kWideBlock, 0, s>>>(tiles, kv, ws, tileCnt);
        k_wsel_rank<<<nTiles, kWideBlock, 0, s>>>(tiles, nodes, kv, ws, tileCnt, tileSize, lpos, rpos);
        k_wsel_swap<<<nTiles * kSwapSplit, kWideBlock, 0, s>>>(tiles, nodes, ws, kv, ord, lpos, rpos, tileSize);
    }
    if (steps) k_wsel_next<<<nb, 64, 0, s>>>(nodes, nNodes, kv, ord, ws, rpos, wideMin, 1u, 0u);
}

void launchBvhSubtrees(hipStream_t s, const uint4* roots, uint32_t nRoots, const uint32_t* internalOf, const float* kx,
                       const float* ky, const float* kz, const float* boxes, float* kv, uint32_t* ord, float* refBox) {
    if (nRoots)
        k_bvh_subtrees<<<nblocks(nRoots, 256), 256, 0, s>>>(roots, nRoots, internalOf, kx, ky, kz, boxes, kv, ord,
                                                            refBox);
}

void launchBvhWide(hipStream_t s, const uint4* wide, uint32_t nWide, const float* refBox, uint32_t base,
                   WideNode* out) {
    if (nWide) k_bvh_wide<<<nblocks(nWide, 256), 256, 0, s>>>(wide, nWide, refBox, base, out);
}

void launchBvhRecords(hipStream_t s, const HostPoint* pts, const uint32_t* ord, uint32_t n, SurfelGeo* geo,
                      SurfelShade* shade) {
    if (n) k_bvh_records<<<nblocks(n, 256), 256, 0, s>>>(pts, ord, n, geo, shade);
}

}  // namespace pale
