// Exact emulation of std::priority_queue<IFPair, vector<IFPair>, Cmp> as the
// reference uses it (src/Header.h:39-66 ordering by value only), usable from
// host and device code.  The reference resolves exact float ties through the
// internal steps of libstdc++'s heap algorithms, so the engine replays those
// steps for the rare (point, table), bucket or query whose result depends on
// a tie:
//   push  = append, then sift the new element up while its parent compares
//           "less" (comp(parent, value)) than it;
//   pop   = take the last element out, move the top to the end, walk the hole
//           from the root down to a leaf always taking the child that does not
//           compare less (the right one unless comp(right, left)), handle the
//           lone last child of an even-sized heap, then sift the taken element
//           up from the hole.
// These are the steps of push_heap / pop_heap / __adjust_heap in libstdc++
// (bits/stl_heap.h); they are restated here because the outcome on ties has
// to be the same, and are checked against std::priority_queue in
// tests/test_heap_emulation.py.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define FPP_HD __host__ __device__ __forceinline__
#else
#define FPP_HD inline
#endif

namespace fpp {

struct IFPair {
    int idx;
    float val;
};

// greater<>: min-heap (top is the smallest value).
struct CmpGreater {
    FPP_HD bool operator()(const IFPair& a, const IFPair& b) const { return a.val > b.val; }
};
// less<>: max-heap (top is the largest value).
struct CmpLess {
    FPP_HD bool operator()(const IFPair& a, const IFPair& b) const { return a.val < b.val; }
};

template <class Cmp>
struct HeapQ {
    IFPair* c;  // storage, capacity owned by the caller
    int n;
    Cmp comp;

    FPP_HD void init(IFPair* storage) {
        c = storage;
        n = 0;
    }
    FPP_HD int size() const { return n; }
    FPP_HD bool empty() const { return n == 0; }
    FPP_HD const IFPair& top() const { return c[0]; }

    FPP_HD void sift_up(int hole, int topIndex, IFPair value) {
        int parent = (hole - 1) / 2;
        while (hole > topIndex && comp(c[parent], value)) {
            c[hole] = c[parent];
            hole = parent;
            parent = (hole - 1) / 2;
        }
        c[hole] = value;
    }

    FPP_HD void push(IFPair v) {
        c[n] = v;
        n++;
        sift_up(n - 1, 0, v);
    }

    FPP_HD void adjust(int hole, int len, IFPair value) {
        const int topIndex = hole;
        int second = hole;
        while (second < (len - 1) / 2) {
            second = 2 * (second + 1);
            if (comp(c[second], c[second - 1])) second--;
            c[hole] = c[second];
            hole = second;
        }
        if ((len & 1) == 0 && second == (len - 2) / 2) {
            second = 2 * (second + 1);
            c[hole] = c[second - 1];
            hole = second - 1;
        }
        sift_up(hole, topIndex, value);
    }

    FPP_HD void pop() {
        if (n > 1) {
            const int last = n - 1;
            IFPair value = c[last];
            c[last] = c[0];
            adjust(0, last, value);
        }
        n--;
    }
};

using MinHeap = HeapQ<CmpGreater>;
using MaxHeap = HeapQ<CmpLess>;

#if defined(__HIPCC__)
// Order-preserving integer key of a finite float value: a.val < b.val exactly
// when key(a) < key(b).  -0 maps like +0 (they compare equal).
__device__ __forceinline__ uint32_t heap_key(float v) {
    uint32_t u = __float_as_uint(v);
    if (u == 0x80000000u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
#endif

}  // namespace fpp
