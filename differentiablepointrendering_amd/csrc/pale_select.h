// pale_select.h -- the element order of libstdc++'s std::nth_element,
// restated for a device build of the median-split BVH.
//
// The surfel BVH's leaves and its point permutation are the order that
// std::nth_element leaves the primitives in (BVH.ixx:49-133 partitions the
// primitive list with it at count/2, comparing centroids along one axis).
// That order is a deterministic function of the comparisons, so a build
// that makes the same comparisons and the same moves reproduces it: this
// file restates the selection algorithm of GCC's libstdc++ (introselect:
// median-of-three pivot moved to the front, Hoare's unguarded partition,
// heap selection once 2 floor(log2 n) partitions have not narrowed the
// range to 3 elements, insertion sort of the last 3).
//
// Elements are (key, id) pairs held in two arrays; only keys are compared
// (key[p] < key[q], the reference's centroid comparator).
//
// Two forms:
//  - selectSeq: the algorithm as one sequential loop, for one thread
//    (small ranges, and every heap selection);
//  - the partition of one introselect step restated by ranks (used by the
//    workgroup kernel for large ranges): with pivot key P, let L_m be the
//    m-th position from the left whose key is not < P and R_m the m-th from
//    the right whose key is not > P.  Hoare's scan swaps L_m with R_m for
//    every m < K, K = #{m : L_m < R_m}, and returns min(L_K, R_{K-1})
//    (R_{-1} = the end, L_K = the end when there are only K such
//    positions).  A position p with key >= P is L_m for m = #(keys >= P
//    left of p); it is swapped iff #(keys <= P right of p) > m.
#pragma once

#include <cstdint>

#ifndef PALE_HD
#define PALE_HD inline
#endif

namespace pale {
namespace sel {

PALE_HD int floorLog2(uint32_t n) {
    int r = 0;
    while (n >>= 1) ++r;
    return r;
}

PALE_HD void swapAt(float* k, uint32_t* id, uint32_t p, uint32_t q) {
    const float tk = k[p];
    k[p] = k[q];
    k[q] = tk;
    const uint32_t ti = id[p];
    id[p] = id[q];
    id[q] = ti;
}

// Sift `value` up from `hole` toward `top` (the push step of a max-heap on
// keys, heap rooted at base).
PALE_HD void pushHeap(float* k, uint32_t* id, uint32_t base, uint32_t hole, uint32_t top, float vk, uint32_t vi) {
    uint32_t parent = (hole - 1) / 2;
    while (hole > top && k[base + parent] < vk) {
        k[base + hole] = k[base + parent];
        id[base + hole] = id[base + parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    k[base + hole] = vk;
    id[base + hole] = vi;
}

// Move the larger child up into the hole until the hole is a leaf, then
// push `value` up from there.
PALE_HD void adjustHeap(float* k, uint32_t* id, uint32_t base, uint32_t hole, uint32_t len, float vk, uint32_t vi) {
    const uint32_t top = hole;
    uint32_t child = hole;
    while (len >= 2 && child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (k[base + child] < k[base + child - 1]) --child;
        k[base + hole] = k[base + child];
        id[base + hole] = id[base + child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        k[base + hole] = k[base + child - 1];
        id[base + hole] = id[base + child - 1];
        hole = child - 1;
    }
    pushHeap(k, id, base, hole, top, vk, vi);
}

// Heap selection of [first, middle) from [first, last): a max-heap of the
// first part, then every later element smaller than the top replaces it.
PALE_HD void heapSelect(float* k, uint32_t* id, uint32_t first, uint32_t middle, uint32_t last) {
    const uint32_t len = middle - first;
    if (len >= 2) {
        uint32_t parent = (len - 2) / 2;
        while (true) {
            adjustHeap(k, id, first, parent, len, k[first + parent], id[first + parent]);
            if (parent == 0) break;
            --parent;
        }
    }
    for (uint32_t i = middle; i < last; ++i) {
        if (k[i] < k[first]) {
            // pop the top into position i, re-heap with i's old element
            const float vk = k[i];
            const uint32_t vi = id[i];
            k[i] = k[first];
            id[i] = id[first];
            adjustHeap(k, id, first, 0, len, vk, vi);
        }
    }
}

// Median of the keys at a, b, c swapped into `result`.
PALE_HD void medianToFirst(float* k, uint32_t* id, uint32_t result, uint32_t a, uint32_t b, uint32_t c) {
    uint32_t m;
    if (k[a] < k[b]) {
        if (k[b] < k[c])
            m = b;
        else if (k[a] < k[c])
            m = c;
        else
            m = a;
    } else if (k[a] < k[c]) {
        m = a;
    } else if (k[b] < k[c]) {
        m = c;
    } else {
        m = b;
    }
    swapAt(k, id, result, m);
}

// Hoare's partition of [lo, hi) around the key at `pivot` (which lies
// outside the range).  The reference's scans are unguarded: the median-of-
// three leaves an element >= and an element <= the pivot inside the range,
// so they stop there.  The bounds tests here never stop a scan earlier on
// such input; they keep keys that are not ordered (NaN) inside the range.
PALE_HD uint32_t unguardedPartition(float* k, uint32_t* id, uint32_t lo, uint32_t hi, uint32_t pivot) {
    const float P = k[pivot];
    const uint32_t lo0 = lo, hi0 = hi;
    while (true) {
        while (lo + 1 < hi0 && k[lo] < P) ++lo;
        --hi;
        while (hi > lo0 && P < k[hi]) --hi;
        if (!(lo < hi)) return lo;
        swapAt(k, id, lo, hi);
        ++lo;
    }
}

// Insertion sort of [first, last) (stable).
PALE_HD void insertionSort(float* k, uint32_t* id, uint32_t first, uint32_t last) {
    if (first == last) return;
    for (uint32_t i = first + 1; i < last; ++i) {
        const float vk = k[i];
        const uint32_t vi = id[i];
        if (vk < k[first]) {
            for (uint32_t j = i; j > first; --j) {
                k[j] = k[j - 1];
                id[j] = id[j - 1];
            }
            k[first] = vk;
            id[first] = vi;
        } else {
            uint32_t j = i;
            while (vk < k[j - 1]) {
                k[j] = k[j - 1];
                id[j] = id[j - 1];
                --j;
            }
            k[j] = vk;
            id[j] = vi;
        }
    }
}

// The introselect loop from a given state (first, last, depth): the state
// of std::nth_element(first0, nth, last0) after some of its steps.
PALE_HD void selectFrom(float* k, uint32_t* id, uint32_t first, uint32_t nth, uint32_t last, int depth) {
    while (last - first > 3) {
        if (depth == 0) {
            heapSelect(k, id, first, nth + 1, last);
            swapAt(k, id, first, nth);
            return;
        }
        --depth;
        const uint32_t mid = first + (last - first) / 2;
        medianToFirst(k, id, first, first + 1, mid, last - 1);
        const uint32_t cut = unguardedPartition(k, id, first + 1, last, first);
        if (cut <= nth)
            first = cut;
        else
            last = cut;
    }
    insertionSort(k, id, first, last);
}

// std::nth_element(first, nth, last) on (key, id) pairs compared by key.
PALE_HD void selectSeq(float* k, uint32_t* id, uint32_t first, uint32_t nth, uint32_t last) {
    if (first == last || nth == last) return;
    selectFrom(k, id, first, nth, last, 2 * floorLog2(last - first));
}

}  // namespace sel
}  // namespace pale
