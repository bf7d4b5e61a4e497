// Compares the engine's heap emulation (falconnpp_amd/csrc/heap.h) with
// libstdc++'s std::priority_queue on tie-heavy random push/pop sequences:
// every top() must agree in both index and value.  Prints "OK <n>" on success.
#include <cstdio>
#include <functional>
#include <queue>
#include <random>
#include <vector>
#include "../../falconnpp_amd/csrc/heap.h"

struct P {
    int i;
    float v;
    P(int a, double b) : i(a), v((float)b) {}
    bool operator<(const P& o) const { return v < o.v; }
    bool operator>(const P& o) const { return v > o.v; }
};

template <class StdQ, class EmuQ>
static bool run(std::mt19937& g, int cap, int ops, int distinct) {
    StdQ s;
    std::vector<fpp::IFPair> store(ops + 1);
    EmuQ e;
    e.init(store.data());
    std::uniform_int_distribution<int> val(0, distinct - 1), coin(0, 9);
    int id = 0;
    for (int t = 0; t < ops; ++t) {
        bool do_pop = !s.empty() && ((int)s.size() >= cap || coin(g) < 3);
        if (do_pop) {
            s.pop();
            e.pop();
        } else {
            float v = (float)val(g) * 0.5f;
            s.emplace(id, v);
            e.push(fpp::IFPair{id, v});
            id++;
        }
        if ((int)s.size() != e.size()) return false;
        if (!s.empty() && (s.top().i != e.top().idx || s.top().v != e.top().val)) return false;
    }
    while (!s.empty()) {
        if (s.top().i != e.top().idx) return false;
        s.pop();
        e.pop();
    }
    return true;
}

int main() {
    std::mt19937 g(12345);
    int n = 0;
    for (int rep = 0; rep < 3000; ++rep) {
        int cap = 1 + rep % 37, distinct = 1 + rep % 7, ops = 50 + rep % 400;
        if (!run<std::priority_queue<P, std::vector<P>, std::greater<>>, fpp::MinHeap>(g, cap, ops, distinct)) {
            printf("MISMATCH min rep=%d\n", rep);
            return 1;
        }
        if (!run<std::priority_queue<P, std::vector<P>>, fpp::MaxHeap>(g, cap, ops, distinct)) {
            printf("MISMATCH max rep=%d\n", rep);
            return 1;
        }
        n += 2;
    }
    printf("OK %d\n", n);
    return 0;
}
