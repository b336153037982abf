// pale_kernels.h -- kernel launch interface between the orchestration
// (pale_render.hip) and the gfx950 kernels (pale_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "pale_device.h"

namespace pale {

constexpr uint32_t kTile = 1024;       // compaction / scan tile (256 threads x 4)
constexpr uint32_t kTraceBlock = 128;  // threads per traversal workgroup
// The layer walk's tiles are split into this many contiguous regions, one
// cursor each; a workgroup starts in region blockIdx.x % kTraceRegions (its
// XCD under round-robin dispatch) so each XCD's L2 serves one part of the
// image, and moves on when its region is used up.
constexpr uint32_t kTraceRegions = 8;

// Device-side progress of one render call.  Counts live on the device so the
// whole forward or backward pass runs without host round trips.
struct DevState {
    uint32_t cursor;       // next entry of the seed table
    uint32_t number;       // seed of the current launch
    uint32_t live;         // current bounce runs (queue not empty)
    uint32_t qCount;       // rays in the current queue
    uint32_t photonCount;  // photon deposits so far (attempted = stored)
    uint32_t extTotal;     // survivors of the last bounce
    uint32_t depTotal;     // deposits of the last bounce
    uint32_t evTotal;      // gradient events of the last bounce
    uint32_t entries;      // camera-gather entries appended this round
    uint32_t overflow;     // the entry buffer filled this round
    uint32_t evPool;       // adjoint events waiting for the transport
    uint32_t resv;         // this bounce may pass the photon capacity
    uint32_t bigCells;     // photon cells too large for the per-wave sort
    uint32_t pad32;
    unsigned long long photonQ, gatherQ, adjointQ, deposits, events, pad;
    // Counting mode, per kernel class (photon, gather, adjoint): popped
    // nodes, AABB tests, primitive tests, gathers, gather cells, non-empty
    // cells, photons scanned.
    unsigned long long work[3][8];
};

enum : int { kWorkPhoton = 0, kWorkGather = 1, kWorkAdjoint = 2 };

// Ray queue, structure of arrays: origin (w = pixel / path id bits),
// direction (w = bounce bits), throughput (w unused).
struct RayQueue {
    float4* o;
    float4* d;
    float4* T;
};

// Photon deposits: position + power.x, direction + power.y, power.z.
struct PhotonArrays {
    float4* posPow;
    float4* dirPow;
    float* powZ;
};

// Adjoint gradient event: e0 = gather position + kind, e1 = gather normal +
// target surfel, e2 = throughput + alpha, e3 = kind-specific words; with
// full gradients also e4 / e5 = origin / direction of the ray that hit the
// surfel whose alpha the event weighs (null pointers otherwise).
struct EventArrays {
    float4* e0;
    float4* e1;
    float4* e2;
    float4* e3;
    float4* e4;
    float4* e5;
};

// Pending adjoint state per path: p0 = position + kind, p1 = throughput +
// alpha, p2 = (primitive, instance) bits; p3 / p4 = the ray of a pending
// NullTransmittance (full gradients only).
struct PendingArrays {
    float4* p0;
    float4* p1;
    uint2* p2;
    float4* p3;
    float4* p4;
};

// Camera gather, per pixel: ot = ray origin + throughput tr, dl = direction +
// next layer (bits), cs = accumulated colour + status (bits), cnt = entries
// of this round, off = their first slot in the ordered list.
struct PixelArrays {
    float4* ot;
    float4* dl;
    float4* cs;
    uint32_t* cnt;
    uint32_t* off;
};

// A batch of camera-gather images (views x passes) walked by one persistent
// launch: the slot space is a list of 8x8 pixel tiles, each naming its
// image (seg) and its tile within it.  Segment = one pass of one view; its
// pixels occupy slots slotBase .. slotBase + W*H - 1 of the pixel arrays.
struct BatchSeg {
    uint32_t view, pass, W, H, tilesX, slotBase, seed, pixOffset;  // pixOffset: the view's first pixel
    float invSpp;                                                   // 1 / the view's spp
    uint32_t pad[3];
};
// The passes of one view in a batch (consecutive segments, pass order): the
// framebuffer update of its pixels, fb = C * invSpp + fb per pass in order.
struct BatchRun {
    uint32_t view, firstSeg, numSegs, npx, fbOffset, resolveBase;
    float invSpp;
    uint32_t pad;
};
struct BatchHeader {
    uint32_t nTiles, nSlots, nSegs, nRuns, nResolve, pad[3];
};
struct BatchDesc {
    // Units of 64 slots: (first segment, tile within the image, first unit
    // of the group, K).  A group is the K units of one tile of one view's K
    // segments of the batch; its 64 K slots take the tile's pixels K at a
    // time, segment fastest (slot j of the group: segment first + j % K,
    // pixel j / K of the 8x8 tile).
    const uint4* tiles;
    const BatchSeg* segs;
    const BatchRun* runs;
    const BatchHeader* hdr;
    const DevCamera* cams;  // per view
};

// Camera-gather entries (one per layer): qa = gather point + normal.x,
// qb = normal.y, normal.z, pixel (bits), index in the pixel's round (bits,
// top bit = no gather); comp = colour factor k + weight w; E = irradiance;
// ord = entries in (pixel, layer) order; cell / sorted = counting sort of
// the gather entries by grid cell.
// Camera-gather entries of one round, in append order: qa = gather point +
// normal.x, qb = normal.yz, slot, layer tag; comp = colour factor k and
// weight w.  pos[i] = place of entry i in its slot's layer-ordered list
// (slot offset + layer); contrib holds each entry's w * (E * k) at that place,
// so a slot's layers are contiguous for the in-order sum.  cell / sorted:
// the counting sort of the gather entries by photon-grid cell.
struct EntryArrays {
    float4* qa;
    float4* qb;
    float4* comp;
    float4* contrib;
    uint32_t* pos;
    uint32_t* cell;
    uint32_t* sorted;
};

// Adjoint trace parameters.  Outside batches every ray weighs invSpp and
// seeds its draws with its queue index.  In a batch (segs != null) a ray
// carries its segment in T.w, weighs that segment's invSpp, and seeds with
// its index among its segment's rays (queue index - segFirst[segment]), the
// index it has when its view is traced alone.
//
// In a batch the lanes also take the rays in another order than the queue's:
// the rays of one view's segments (its samples) interleave, lane t of the
// view's range taking rank t / K of segment t % K (K = the view's segments in
// the batch), so neighbouring lanes trace the same pixel of different
// samples.  groups[g] = (first entry in groupSegs, K, first lane, rays per
// segment at most) for the views of the batch; segEnd[s] ends segment s.
struct TraceParams {
    uint32_t rrStart;
    float qNull, qReflect, qTransmit;
    uint32_t fullGradients;  // record hit rays for the full gradients
    float invSpp;
    const BatchSeg* segs;
    const uint32_t* segFirst;
    const uint32_t* segEnd;
    const uint4* groups;
    const uint32_t* groupSegs;
    uint32_t nGroups;
    uint32_t seed;  // adjoint base seed of an unbatched queue (batched: BatchSeg::seed)
    // Debug gradient images (setting debug_images): events also record their
    // image pixel, the index over all adjoint cameras' pixels (camera
    // offset + pixel) in e5.w; debugBase is the unbatched queue's camera
    // offset (batched: BatchSeg::pixOffset).
    uint32_t debug;
    uint32_t debugBase;
    // The traversal stack: with spill null the whole stack of every lane is
    // in LDS; else its first inLds entries, the rest in this per-lane area
    // (stackDepth - inLds entries for each of the launch's lanes).
    uint2* spill;
    uint32_t inLds;
};

// Photon-map reservoir (depositPhotonSurface past capacity): per queue item
// its deposit attempt index (valid when DevState::resv) and the slot it
// replaces (kNoSlot if none); per stored slot the highest attempt that
// claimed it.
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
struct Reservoir {
    uint32_t* attempt;
    uint32_t* slot;
    uint32_t* owner;
    uint32_t cap;
};

struct GridParams {
    float ox, oy, oz, cell;
    int rx, ry, rz;
    uint32_t total;
};

// Launchers (all asynchronous on `s`).
void launchFill(hipStream_t s, uint32_t* p, uint32_t n, uint32_t v);
void launchFillF(hipStream_t s, float* p, uint32_t n, float v);
void launchIota(hipStream_t s, uint32_t* p, uint32_t n);  // p[i] = i
void launchSeedNext(hipStream_t s, DevState* st, const uint32_t* table, uint32_t tableLen, int bounce);
void launchBeginView(hipStream_t s, DevState* st, uint32_t number);
void launchEmit(hipStream_t s, const DevScene& S, DevState* st, uint32_t photons, float invEmitted,
                RayQueue staged, uint8_t* flags);
void launchPhotonTrace(hipStream_t s, const DevScene& S, DevState* st, uint32_t maxRays, RayQueue in,
                       RayQueue staged, uint8_t* extFlags, PhotonArrays dep, uint8_t* depFlags, TraceParams tp,
                       Reservoir rs, uint32_t* tiles, bool count);
// After the bounce's deposits are compacted: deposits past capacity that the
// reservoir accepted overwrite their slots, the highest attempt winning.
void launchReservoirWrite(hipStream_t s, DevState* st, uint32_t maxRays, Reservoir rs, PhotonArrays dep,
                          PhotonArrays raw);
// Stable compaction of flagged items i < *count (items >= *count are ignored).
void launchCompactRays(hipStream_t s, const uint8_t* flags, const uint32_t* count, uint32_t maxItems,
                       uint32_t* tiles, uint32_t* total, RayQueue src, RayQueue dst, uint32_t dstCap);
void launchCompactPhotons(hipStream_t s, const uint8_t* flags, const uint32_t* count, uint32_t maxItems,
                          uint32_t* tiles, uint32_t* total, const uint32_t* base, PhotonArrays src,
                          PhotonArrays dst, uint32_t dstCap);
// Stable compaction of one bounce's events, appended to the event pool
// (DevState::evPool items; this bounce's count goes to DevState::evTotal).
void launchCompactEvents(hipStream_t s, const uint8_t* flags, const uint32_t* count, uint32_t maxItems,
                         uint32_t* tiles, DevState* st, EventArrays src, EventArrays pool, uint32_t poolCap);
void launchFinishPhotonBounce(hipStream_t s, DevState* st, uint32_t cap);
void launchSetQueue(hipStream_t s, DevState* st, const uint32_t* src);
// Photon grid (counting sort by cell, stable in deposit order).
void launchBuildGrid(hipStream_t s, const GridParams& g, DevState* st, uint32_t photonCap, PhotonArrays raw,
                     PhotonArrays sorted, uint32_t* cellId, uint32_t* sortedIdx, uint32_t* cellCount,
                     uint32_t* cellStart, uint32_t* tiles, uint32_t* scratchTotal);
// Exclusive scan of n values into out[0..n] (out[n] = total).
void launchExclusiveScan(hipStream_t s, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tiles,
                         uint32_t* scratchTotal);
// Camera gather of one pass (see pale_kernels.hip for the stages).
void launchCamInit(hipStream_t s, const DevCamera& cam, const DevState* st, uint32_t pass, int decorrelate,
                   PixelArrays px);
// Layer walk: the first inLds stack entries of a lane (default
// kWalkLdsEntries) in LDS, the rest in a spill area of
// walkSpillBytes(S, inLds) (null: the whole stack in LDS).
constexpr uint32_t kWalkLdsEntries = 16;
size_t walkSpillBytes(const DevScene& S, uint32_t inLds);
void launchCamTrace(hipStream_t s, const DevScene& S, DevState* st, uint32_t npx, uint32_t width, PixelArrays px,
                    EntryArrays ent, uint32_t cap, uint32_t* cursor, uint2* spill, uint32_t inLds, bool count);
void launchEntrySort(hipStream_t s, const GridParams& g, DevState* st, uint32_t cap, PixelArrays px,
                     EntryArrays ent, uint32_t* cellCount, uint32_t* cellStart, uint32_t* tiles,
                     uint32_t* scratchTotal);
void launchEntryGather(hipStream_t s, const DevGrid& G, DevState* st, uint32_t cap, const uint32_t* nGather,
                       EntryArrays ent, bool count);
void launchCamComposite(hipStream_t s, uint32_t npx, PixelArrays px, EntryArrays ent, float invSpp, float4* fb);
// Batched camera gather: primary rays of every slot of the batch; the layer
// walk over the batch's tiles (persistent grid scaled by gridFraction);
// composite of up to maxSlots slots (count read from the header) without the
// framebuffer update; the pass-ordered framebuffer update per view.
void launchBatchInit(hipStream_t s, BatchDesc b, uint32_t maxTiles, PixelArrays px);
void launchCamTraceBatch(hipStream_t s, const DevScene& S, DevState* st, BatchDesc b, uint32_t maxSlots,
                         PixelArrays px, EntryArrays ent, uint32_t cap, uint32_t* cursor, uint2* spill,
                         uint32_t inLds, float gridFraction, bool count);
void launchCamCompositeBatch(hipStream_t s, uint32_t maxSlots, const BatchHeader* hdr, PixelArrays px,
                             EntryArrays ent);
void launchBatchResolve(hipStream_t s, BatchDesc b, uint32_t maxResolve, PixelArrays px, float4* fb);
void launchPostprocess(hipStream_t s, const float4* fb, float4* ldr, uint32_t n);
// Batched adjoint: primary rays of every slot of the batch (queue index =
// slot, T.w = segment); first queue index of every segment present.
void launchAdjRaygenBatch(hipStream_t s, BatchDesc b, uint32_t maxTiles, const float* dloss, RayQueue q,
                          PendingArrays pend, DevState* st);
void launchSegFirst(hipStream_t s, const DevState* st, uint32_t maxRays, RayQueue q, uint32_t* segFirst,
                    uint32_t* segEnd, uint32_t nSegs, uint4* groups, uint32_t nGroups, const uint32_t* groupSegs);
void launchAdjRaygen(hipStream_t s, const DevCamera& cam, const float* dloss, RayQueue q, PendingArrays pend,
                     DevState* st, uint32_t seed);
void launchAdjTrace(hipStream_t s, const DevScene& S, DevState* st, uint32_t maxRays, uint32_t bounce, RayQueue in,
                    RayQueue staged, uint8_t* extFlags, EventArrays ev, uint8_t* evFlags, PendingArrays pend,
                    TraceParams tp, bool count);
// Transport of the pooled events (in cell order after launchEventSort).
void launchEventSort(hipStream_t s, const GridParams& g, DevState* st, uint32_t poolCap, EventArrays ev,
                     uint32_t* cell, uint32_t* sorted, uint32_t* cellCount, uint32_t* cellStart, uint32_t* tiles,
                     uint32_t* scratchTotal);
// gradFull (16 floats per surfel: position 3, tangent_u 3, tangent_v 3,
// scale 2, albedo 3, unused, beta) is null unless full gradients are on.
// debugImg (nullable): RGBA per pixel over all adjoint cameras; each event
// adds its opacity-gradient contribution to RGB and sets alpha to 1
// (atomicAddFloat4ToImage, KernelHelpers.h:1489-1502).
void launchAdjTransport(hipStream_t s, const DevScene& S, const DevGrid& G, DevState* st, uint32_t poolCap,
                        EventArrays ev, const uint32_t* sorted, float* gradOpacity, float* gradFull, float* debugImg,
                        bool count);
// Split gradFull into the ABI's arrays (any may be null).
void launchSplitGradients(hipStream_t s, const float* gradFull, uint32_t n, float* position, float* tangentU,
                          float* tangentV, float* scale, float* albedo, float* beta);
void launchSetOpacity(hipStream_t s, SurfelShade* shade, const float* opacity, uint32_t n);
// The seven parameter arrays of apply_point_optimization in device memory
// (BVH order): position, tangent_u, tangent_v (3 floats each), scale (2),
// albedo (3), opacity, beta.
struct DevPointArrays {
    const float *position, *tangent_u, *tangent_v, *scale, *albedo, *opacity, *beta;
};
// Surfel records from device parameter arrays, and their raw copy (16 floats
// per surfel: p, t_u, t_v, s_u, s_v, albedo, opacity, beta).
void launchApplyPoints(hipStream_t s, SurfelGeo* geo, SurfelShade* shade, float* raw, DevPointArrays a, uint32_t n);

// Device build of the surfel BVH (pale_bvh.hip).  boxes / refBox: 6 floats
// (min xyz, max xyz) per surfel / per reference node; acc: 6 order-
// preserving integers per node of a fitted level; uint4 tables as described
// at each kernel.
struct HostPoint;
void launchBvhPrims(hipStream_t s, const HostPoint* pts, uint32_t n, float* kx, float* ky, float* kz, float* boxes,
                    uint32_t* ord, uint32_t* flags);
void launchBvhInitFit(hipStream_t s, uint32_t* acc, uint32_t nNodes);
void launchBvhFit(hipStream_t s, const uint4* chunks, uint32_t nChunks, const uint32_t* ord, const float* boxes,
                  uint32_t* acc);
// State of a large node's introselect between the multi-workgroup steps:
// the current range and depth limit, the step's pivot, its counts and its
// cut candidates.
struct BvhWideState {
    uint32_t lo, hi, depth, active;
    float P;
    uint32_t nLE, K, LK;
};
// count: the level's largest node (picks the workgroup size); resume: the
// state each node's introselect continues from (null: from the start)
void launchBvhSelect(hipStream_t s, const uint4* nodes, uint32_t nNodes, uint32_t count, const uint32_t* acc,
                     const float* kx, const float* ky, const float* kz, float* kv, uint32_t* ord, uint32_t* lpos,
                     uint32_t* rpos, float* refBox, const BvhWideState* resume);
// `steps` introselect steps of a level's nodes over many workgroups (ranges
// cut into tiles of tileSize positions, tiles[t] = {node, first, end, tile
// index in node}); a node stops at wideMin positions
void launchBvhWideSteps(hipStream_t s, const uint4* nodes, uint32_t nNodes, const uint4* tiles, uint32_t nTiles,
                        uint32_t tileSize, uint32_t steps, uint32_t wideMin, const uint32_t* acc, const float* kx,
                        const float* ky, const float* kz, float* kv, uint32_t* ord, uint32_t* lpos, uint32_t* rpos,
                        float* refBox, BvhWideState* ws, uint2* tileCnt);
void launchBvhSubtrees(hipStream_t s, const uint4* roots, uint32_t nRoots, const uint32_t* internalOf, const float* kx,
                       const float* ky, const float* kz, const float* boxes, float* kv, uint32_t* ord, float* refBox);
void launchBvhWide(hipStream_t s, const uint4* wide, uint32_t nWide, const float* refBox, uint32_t base,
                   WideNode* out);
void launchBvhRecords(hipStream_t s, const HostPoint* pts, const uint32_t* ord, uint32_t n, SurfelGeo* geo,
                      SurfelShade* shade);
void launchGatherPoints(hipStream_t s, const DevGrid& G, uint32_t n, const float* q, const float* nrm, float* out);
void launchTraceRays(hipStream_t s, const DevScene& S, uint32_t n, const float* o, const float* d, int32_t* inst,
                     int32_t* prim, float* t, float* pos, float* alpha);

}  // namespace pale
