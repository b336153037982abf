// The selection the GPU BVH build makes (pale_select.h) against libstdc++'s
// std::nth_element, element for element, on (key, id) pairs:
//  - selectSeq, the sequential restatement (one thread per small range and
//    every heap selection), on random, tied, sorted and reversed keys;
//  - selectFrom at depth 0 (heap selection) against std::__heap_select;
//  - the rank form of one Hoare partition (the workgroup kernel's), run
//    here as a sequential loop over positions, inside the introselect loop.
// Prints "ok ..." and exits 0 when everything matches.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "pale_select.h"

using namespace pale::sel;

namespace {
struct E {
    float k;
    uint32_t id;
};
bool byKey(const E& p, const E& q) { return p.k < q.k; }

float keyOf(std::mt19937_64& g, int mode, uint32_t i, uint32_t n) {
    switch (mode) {
        case 0: return (float)(g() % 1000000) * 1e-3f;
        case 1: return (float)(g() % 3);
        case 2: return (float)i;
        case 3: return (float)(n - i);
        case 4: return (float)(g() % 2);
        default: return (float)((i * 7919u) % 13u) + (float)(g() % 3) * 0.5f;
    }
}

// One introselect step's partition by ranks, as k_bvh_select computes it:
// a position with key >= P is the left partner of swap G iff E > G (else the
// first such position is L_K); a position with key <= P is the right partner
// of swap E iff G > E; only the partners are recorded.
uint32_t rankPartition(float* k, uint32_t* id, uint32_t lo, uint32_t hi, float P, std::vector<uint32_t>& L,
                       std::vector<uint32_t>& R) {
    uint32_t nLE = 0;
    for (uint32_t p = lo; p < hi; ++p) nLE += !(P < k[p]);
    std::fill(L.begin(), L.end(), 0xFFFFFFFFu);
    std::fill(R.begin(), R.end(), 0xFFFFFFFFu);
    uint32_t G = 0, Lc = 0, K = 0, lk = hi;
    for (uint32_t p = lo; p < hi; ++p) {
        const bool ge = !(k[p] < P), le = !(P < k[p]);
        if (le) ++Lc;
        const uint32_t Ep = nLE - Lc;
        if (ge) {
            if (Ep > G) {
                L[G] = p;
                ++K;
            } else {
                lk = std::min(lk, p);
            }
        }
        if (le && G > Ep) R[Ep] = p;
        if (ge) ++G;
    }
    for (uint32_t m = 0; m < K; ++m) {
        if (L[m] == 0xFFFFFFFFu || R[m] == 0xFFFFFFFFu) return 0;  // a partner not recorded
        swapAt(k, id, L[m], R[m]);
    }
    uint32_t cut = lk;
    if (K > 0) cut = std::min(cut, R[K - 1]);
    return cut;
}

void selectRanks(float* k, uint32_t* id, uint32_t first, uint32_t nth, uint32_t last) {
    int depth = 2 * floorLog2(last - first);
    std::vector<uint32_t> L(last), R(last);
    while (last - first > 3) {
        if (depth == 0) {
            selectFrom(k, id, first, nth, last, 0);
            return;
        }
        --depth;
        medianToFirst(k, id, first, first + 1, first + (last - first) / 2, last - 1);
        const uint32_t cut = rankPartition(k, id, first + 1, last, k[first], L, R);
        if (cut == 0) {
            printf("rank partition: a swap partner was not recorded\n");
            std::exit(1);
        }
        if (cut <= nth)
            first = cut;
        else
            last = cut;
    }
    insertionSort(k, id, first, last);
}
}  // namespace

int main() {
    std::mt19937_64 g(7);
    long sel = 0, heaps = 0, ranks = 0;
    for (int t = 0; t < 200000; ++t) {
        const uint32_t n = 1 + (uint32_t)(g() % (t < 150000 ? 12 : (t < 198000 ? 300 : 20000)));
        const int mode = (int)(g() % 6);
        std::vector<float> k(n);
        std::vector<uint32_t> id(n);
        std::vector<E> ref(n);
        for (uint32_t i = 0; i < n; ++i) {
            k[i] = keyOf(g, mode, i, n);
            id[i] = i;
            ref[i] = E{k[i], i};
        }
        const uint32_t nth = n >> 1;
        std::nth_element(ref.begin(), ref.begin() + nth, ref.end(), byKey);
        std::vector<float> k2 = k;
        std::vector<uint32_t> id2 = id;
        selectSeq(k.data(), id.data(), 0, nth, n);
        ++sel;
        for (uint32_t i = 0; i < n; ++i)
            if (ref[i].id != id[i]) {
                printf("selectSeq differs: n=%u mode=%d at %u\n", n, mode, i);
                return 1;
            }
        if (n >= 1) {
            selectRanks(k2.data(), id2.data(), 0, nth, n);
            ++ranks;
            for (uint32_t i = 0; i < n; ++i)
                if (ref[i].id != id2[i]) {
                    printf("rank partition differs: n=%u mode=%d at %u\n", n, mode, i);
                    return 1;
                }
        }
    }
    for (int t = 0; t < 50000; ++t) {
        const uint32_t n = 4 + (uint32_t)(g() % 40);
        std::vector<float> k(n);
        std::vector<uint32_t> id(n);
        std::vector<E> ref(n);
        for (uint32_t i = 0; i < n; ++i) {
            k[i] = (float)(g() % 7);
            id[i] = i;
            ref[i] = E{k[i], i};
        }
        const uint32_t nth = (uint32_t)(g() % n);
        std::__heap_select(ref.begin(), ref.begin() + nth + 1, ref.end(), __gnu_cxx::__ops::__iter_comp_iter(byKey));
        std::iter_swap(ref.begin(), ref.begin() + nth);
        selectFrom(k.data(), id.data(), 0, nth, n, 0);
        ++heaps;
        for (uint32_t i = 0; i < n; ++i)
            if (ref[i].id != id[i]) {
                printf("heap selection differs: n=%u at %u\n", n, i);
                return 1;
            }
    }
    printf("ok %ld selections, %ld rank-form selections, %ld heap selections\n", sel, ranks, heaps);
    return 0;
}
