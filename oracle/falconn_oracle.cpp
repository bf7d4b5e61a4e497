// Falconn++ CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// A line-by-line restatement, in plain C++17 + OpenMP, of the reference's
// 1D hot path: src/FalconnPP.h:52-100 (sign bits, parameters),
// src/FalconnPP.cpp:21-303 (build2Layers_1D) and :310-707 (query2Layers_1D),
// src/Header.h:39-66 (IFPair ordering), src/Utilities.h:26-31 (sgn) and
// src/BF.cpp:17-87 (brute force, here in FP64).  Eigen and Boost are not
// available, so matrices are plain col-major float arrays and the bitset is a
// vector<uint64_t>.  Every selection uses std::priority_queue<IFPair> from
// libstdc++, exactly as the reference, so exact float ties resolve the same way.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, and only to check results or as the CPU baseline.  The product never
// links it.
//
// Parity status of each function (how it is pinned):
//   orc_sign_bits      closed form of minstd_rand0 + uniform_int(0,1) (tests)
//   orc_fwht           golden vectors from the reference's fht_float + H*H=nI
//   orc_hash_*         brute-force top-p + invariants on tiny inputs
//   orc_build          invariants (iProbes distinct buckets/point, iLimit rule,
//                      kept = top-iLimit by value) + brute force on tiny inputs
//   orc_query          with no filtering and every bucket probed, the top-k
//                      equals brute force; probe set = global top-qProbes
//   centre (X mean)    parity unpinned against the reference bits (Eigen's
//                      reduction order); the oracle takes the centre as input.
//   rerank dot         parity unpinned bitwise (Eigen's dot order); tolerance.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <queue>
#include <random>
#include <string>
#include <vector>
#include <omp.h>

namespace {

// src/Header.h:39-66 -- ordering is on the value only.
struct IFPair {
    int m_iIndex;
    float m_fValue;
    IFPair() : m_iIndex(0), m_fValue(0.0f) {}
    IFPair(int i, double v) : m_iIndex(i), m_fValue((float)v) {}
    bool operator<(const IFPair& p) const { return m_fValue < p.m_fValue; }
    bool operator>(const IFPair& p) const { return m_fValue > p.m_fValue; }
};

using MinQ = std::priority_queue<IFPair, std::vector<IFPair>, std::greater<>>;
using MaxQ = std::priority_queue<IFPair, std::vector<IFPair>>;

// src/Utilities.h:26-31
inline int sgn(float x) { return x >= 0 ? 1 : -1; }

// Unnormalised in-place Walsh-Hadamard transform with ascending strides; the
// arithmetic of the reference's fht_float (src/fht_avx.c:374-786,
// src/fht_sse.c:630-1158) as the sequence of butterflies (u,v)->(u+v,u-v).
void fwht(float* a, int log_n) {
    const int n = 1 << log_n;
    for (int h = 1; h < n; h <<= 1)
        for (int i = 0; i < n; i += 2 * h)
            for (int j = i; j < i + h; ++j) {
                float u = a[j], v = a[j + h];
                a[j] = u + v;
                a[j + h] = u - v;
            }
}

struct Index {
    int n_points, n_features;
    int n_tables = 10, n_proj = 512, n_rotate = 3, n_threads = 1;
    int bucket_minSize = 50;
    float bucket_scale = 1.0f;
    int iProbes = 1, qProbes = 1, seed = -1;
    int fhtDim = 0;
    std::vector<float> X;        // centred data, col-major D x N
    std::vector<float> center;   // D
    std::vector<uint8_t> bit1, bit2;
    std::vector<uint64_t> bucketPos;   // global bucket -> first id
    std::vector<uint32_t> bucketSize;  // global bucket -> kept size
    std::vector<int32_t> ids;          // kept ids, bucket after bucket (value-descending inside a bucket)
    bool built = false;
    double build_seconds = 0;
    int64_t limited_buckets = 0;
    std::string err;

    int log2_D() const { return (int)std::log2((double)n_proj); }
    int log2_F() const { return (int)std::log2((double)fhtDim); }
    int nb() const { return 4 * n_proj * n_proj; }
};

// src/FalconnPP.h:52-75: interleaved draws bitHD1[d], bitHD2[d].
void sign_bits(unsigned seed, int nbits, uint8_t* b1, uint8_t* b2) {
    std::default_random_engine gen(seed);
    std::uniform_int_distribution<uint32_t> unif(0, 1);
    for (int d = 0; d < nbits; ++d) {
        b1[d] = (uint8_t)(unif(gen) & 1);
        b2[d] = (uint8_t)(unif(gen) & 1);
    }
}

// One table's two-layer rotation of vector v (length n_features), following
// src/FalconnPP.cpp:77-93 (build) and :360-376 (query).
void rotate2(const Index& I, const float* v, int l, float* r1, float* r2) {
    const int F = I.fhtDim, lf = I.log2_F();
    std::fill(r1, r1 + F, 0.0f);
    std::copy(v, v + I.n_features, r1);
    std::copy(r1, r1 + F, r2);
    for (int r = 0; r < I.n_rotate; ++r) {
        const size_t base = (size_t)l * I.n_rotate * F + (size_t)r * F;
        for (int d = 0; d < F; ++d) {
            r1[d] *= (2 * static_cast<float>(I.bit1[base + d]) - 1);
            r2[d] *= (2 * static_cast<float>(I.bit2[base + d]) - 1);
        }
        fwht(r1, lf);
        fwht(r2, lf);
    }
}

// Top-p (code, |value|) per layer with the reference's min-heap and strict '>'
// replacement (src/FalconnPP.cpp:101-140 / :390-433); returned largest first
// (:145-157 / :438-453).
void layer_top(const Index& I, const float* rot, int p, std::vector<IFPair>& out) {
    const int log2_D = I.log2_D();
    MinQ q;
    for (int r = 0; r < I.n_proj; ++r) {
        int s = sgn(rot[r]);
        float v = s * rot[r];
        int code = r;
        if (s < 0) code |= 1UL << log2_D;
        if ((int)q.size() < p) q.emplace(code, v);
        else if (v > q.top().m_fValue) { q.pop(); q.emplace(code, v); }
    }
    out.assign(p, IFPair());
    for (int i = p - 1; i >= 0; --i) { out[i] = q.top(); q.pop(); }
}

// Build-side pairs (src/FalconnPP.cpp:159-201): top-iProbes pair sums, in the
// order the min-heap pops them (ascending value).
void build_pairs(const Index& I, const std::vector<IFPair>& v1, const std::vector<IFPair>& v2,
                 std::vector<IFPair>& popped) {
    MinQ q;
    for (const auto& a : v1)
        for (const auto& b : v2) {
            float s = b.m_fValue + a.m_fValue;
            int bucket = a.m_iIndex * (2 * I.n_proj) + b.m_iIndex;
            if ((int)q.size() < I.iProbes) q.emplace(bucket, s);
            else if (s > q.top().m_fValue) { q.pop(); q.emplace(bucket, s); }
        }
    popped.clear();
    while (!q.empty()) { popped.push_back(q.top()); q.pop(); }
}

// src/FalconnPP.cpp:21-303
void build(Index& I, const float* Xraw, int64_t ld, const float* center_in) {
    const int N = I.n_points, D = I.n_features, L = I.n_tables;
    I.center.assign(D, 0.0f);
    if (center_in) {
        std::copy(center_in, center_in + D, I.center.begin());
    } else {
        // The reference uses Eigen's rowwise().mean() (float, Eigen's order);
        // that order is not reproducible here, so the oracle uses a sequential
        // double sum.  Parity tests pass the engine's centre explicitly.
        for (int d = 0; d < D; ++d) {
            double s = 0;
            for (int n = 0; n < N; ++n) s += Xraw[(size_t)n * ld + d];
            I.center[d] = (float)(s / N);
        }
    }
    I.X.assign((size_t)D * N, 0.0f);
    for (int n = 0; n < N; ++n)
        for (int d = 0; d < D; ++d)
            I.X[(size_t)n * D + d] = Xraw[(size_t)n * ld + d] - I.center[d];

    auto t0 = std::chrono::high_resolution_clock::now();
    const int nbits = I.fhtDim * L * I.n_rotate;
    I.bit1.assign(nbits, 0);
    I.bit2.assign(nbits, 0);
    unsigned seed = (unsigned)std::chrono::system_clock::now().time_since_epoch().count();
    if (I.seed > -1) seed = (unsigned)I.seed;
    sign_bits(seed, nbits, I.bit1.data(), I.bit2.data());

    const int NB = I.nb();
    std::vector<std::vector<uint32_t>> tSize(L), tOff(L);
    std::vector<std::vector<int32_t>> tIds(L);
    int64_t limited = 0;
    omp_set_num_threads(I.n_threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : limited)
    for (int l = 0; l < L; ++l) {
        std::vector<MaxQ> bq(NB);
        std::vector<float> r1(I.fhtDim), r2(I.fhtDim);
        std::vector<IFPair> v1, v2, popped;
        for (int n = 0; n < N; ++n) {
            rotate2(I, &I.X[(size_t)n * D], l, r1.data(), r2.data());
            layer_top(I, r1.data(), I.iProbes, v1);
            layer_top(I, r2.data(), I.iProbes, v2);
            build_pairs(I, v1, v2, popped);
            for (const auto& pr : popped) bq[pr.m_iIndex].emplace(n, pr.m_fValue);
        }
        auto& sz = tSize[l];
        auto& off = tOff[l];
        auto& out = tIds[l];
        sz.assign(NB, 0);
        off.assign(NB, 0);
        std::vector<int32_t> bucket;
        for (int h = 0; h < NB; ++h) {
            if (bq[h].empty()) continue;
            int B = (int)bq[h].size();
            bucket.assign(B, -1);
            for (int i = 0; i < B; ++i) { bucket[i] = bq[h].top().m_iIndex; bq[h].pop(); }
            // src/FalconnPP.cpp:233: float arithmetic, then ceil.
            int iLimit = (int)std::ceil(I.bucket_scale * B / I.iProbes);
            if (iLimit < I.bucket_minSize) {
                iLimit = std::min(I.bucket_minSize, B);
                limited++;
            }
            if (iLimit > B) iLimit = B;  // guard on the reference's overrun
            off[h] = (uint32_t)out.size();
            sz[h] = (uint32_t)iLimit;
            out.insert(out.end(), bucket.begin(), bucket.begin() + iLimit);
        }
    }
    I.limited_buckets = limited;
    // Concatenate tables in order (the reference appends under a critical
    // section, so only bucket contents, not positions, are comparable).
    I.bucketPos.assign((size_t)L * NB, 0);
    I.bucketSize.assign((size_t)L * NB, 0);
    size_t total = 0;
    for (int l = 0; l < L; ++l) total += tIds[l].size();
    I.ids.clear();
    I.ids.reserve(total);
    for (int l = 0; l < L; ++l) {
        uint64_t base = I.ids.size();
        for (int h = 0; h < NB; ++h) {
            I.bucketPos[(size_t)l * NB + h] = base + tOff[l][h];
            I.bucketSize[(size_t)l * NB + h] = tSize[l][h];
        }
        I.ids.insert(I.ids.end(), tIds[l].begin(), tIds[l].end());
        std::vector<int32_t>().swap(tIds[l]);
    }
    I.built = true;
    I.build_seconds = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
}

struct QueryDebug {
    std::vector<int32_t> probes;     // global bucket ids in probing order
    std::vector<float> probeVals;
    std::vector<int32_t> cands;      // unique candidates, ascending
    int64_t ctot = 0;
    int64_t empty = 0;
};

// src/FalconnPP.cpp:336-337
inline int per_table(const Index& I) { return (int)std::ceil(2.0 * I.qProbes / I.n_tables); }
inline int per_layer(const Index& I) { return (int)std::ceil(std::sqrt(1.0 * per_table(I))); }

// The probing sequence of one query (src/FalconnPP.cpp:349-582).
void probe_sequence(const Index& I, const float* q, QueryDebug& dbg, std::vector<IFPair>& seq) {
    const int L = I.n_tables, NB = I.nb();
    const int perTable = per_table(I), perLayer = per_layer(I);
    std::vector<float> r1(I.fhtDim), r2(I.fhtDim);
    std::vector<IFPair> v1, v2;
    std::vector<IFPair> probes((size_t)L * perTable);
    for (int l = 0; l < L; ++l) {
        rotate2(I, q, l, r1.data(), r2.data());
        layer_top(I, r1.data(), perLayer, v1);
        layer_top(I, r2.data(), perLayer, v2);
        MinQ tq;
        for (const auto& a : v1)
            for (const auto& b : v2) {
                int bucket = a.m_iIndex * (2 * I.n_proj) + b.m_iIndex;
                float s = a.m_fValue + b.m_fValue;
                if ((int)tq.size() < perTable) tq.emplace(bucket, s);
                else if (s > tq.top().m_fValue) { tq.pop(); tq.emplace(bucket, s); }
            }
        int idx = perTable - 1;
        while (!tq.empty()) {
            IFPair p = tq.top();
            tq.pop();
            p.m_iIndex = l * NB + p.m_iIndex;
            probes[(size_t)l * perTable + idx] = p;
            idx--;
        }
    }
    MaxQ mq;
    for (int l = 0; l < L; ++l) mq.push(probes[(size_t)l * perTable]);
    std::vector<int> track(L, 0);
    seq.clear();
    for (int c = 0; c < I.qProbes; ++c) {
        IFPair p = mq.top();
        mq.pop();
        seq.push_back(p);
        int t = p.m_iIndex / NB;
        track[t]++;
        if (track[t] < perTable) mq.push(probes[(size_t)t * perTable + track[t]]);
    }
}

void query_one(const Index& I, const float* q, int k, int32_t* topk, float* scores, QueryDebug* dbgp) {
    QueryDebug local;
    QueryDebug& dbg = dbgp ? *dbgp : local;
    std::vector<IFPair> seq;
    probe_sequence(I, q, dbg, seq);
    const int N = I.n_points, D = I.n_features;
    std::vector<uint64_t> bits(((size_t)N + 63) / 64, 0);
    int64_t nuniq = 0;
    dbg.probes.clear();
    dbg.probeVals.clear();
    for (const auto& p : seq) {
        dbg.probes.push_back(p.m_iIndex);
        dbg.probeVals.push_back(p.m_fValue);
        uint64_t pos = I.bucketPos[p.m_iIndex];
        uint32_t sz = I.bucketSize[p.m_iIndex];
        if (sz == 0) { dbg.empty++; continue; }
        dbg.ctot += sz;
        for (uint32_t i = 0; i < sz; ++i) {
            int id = I.ids[pos + i];
            uint64_t m = 1ull << (id & 63);
            if (!(bits[id >> 6] & m)) { nuniq++; bits[id >> 6] |= m; }
        }
    }
    for (int i = 0; i < k; ++i) { topk[i] = 0; scores[i] = 0.0f; }
    if (dbgp) {
        dbg.cands.clear();
        for (size_t w = 0; w < bits.size(); ++w)
            for (uint64_t b = bits[w]; b; b &= b - 1) dbg.cands.push_back((int)(w * 64 + __builtin_ctzll(b)));
    }
    if (nuniq == 0) return;
    MinQ tk;
    for (size_t w = 0; w < bits.size(); ++w)
        for (uint64_t b = bits[w]; b; b &= b - 1) {
            int id = (int)(w * 64 + __builtin_ctzll(b));
            const float* x = &I.X[(size_t)id * D];
            float s = 0.0f;
            for (int d = 0; d < D; ++d) s += q[d] * x[d];
            if ((int)tk.size() < k) tk.emplace(id, s);
            else if (s > tk.top().m_fValue) { tk.pop(); tk.emplace(id, s); }
        }
    for (int i = (int)tk.size() - 1; i >= 0; --i) {
        topk[i] = tk.top().m_iIndex;
        scores[i] = tk.top().m_fValue;
        tk.pop();
    }
}

}  // namespace

extern "C" {

// ---- primitives -----------------------------------------------------------
void orc_sign_bits(unsigned seed, int nbits, uint8_t* b1, uint8_t* b2) { sign_bits(seed, nbits, b1, b2); }

void orc_fwht(float* buf, int log_n) { fwht(buf, log_n); }

// ---- index ----------------------------------------------------------------
void* orc_create(int n, int d) {
    Index* I = new Index();
    I->n_points = n;
    I->n_features = d;
    return I;
}

void orc_destroy(void* h) { delete (Index*)h; }

// src/FalconnPP.h:85-100
int orc_set_index_param(void* h, int L, int D, int bucketLimit, float alpha, int p, int t, int s) {
    Index& I = *(Index*)h;
    I.n_tables = L;
    I.n_proj = D;
    I.bucket_minSize = bucketLimit;
    I.bucket_scale = alpha;
    I.iProbes = p;
    I.n_threads = t <= 0 ? omp_get_max_threads() : t;
    I.seed = s;
    if (I.n_proj < I.n_features) I.fhtDim = 1 << (int)std::ceil(std::log2((double)I.n_features));
    else I.fhtDim = 1 << (int)std::ceil(std::log2((double)I.n_proj));
    return 0;
}

void orc_set_qprobes(void* h, int p) { ((Index*)h)->qProbes = p; }
void orc_set_threads(void* h, int t) { ((Index*)h)->n_threads = t <= 0 ? omp_get_max_threads() : t; }
int orc_fht_dim(void* h) { return ((Index*)h)->fhtDim; }

// X col-major D x N with leading dimension ld (>= D); centre may be null.
int orc_build(void* h, const float* X, int64_t ld, const float* center) {
    build(*(Index*)h, X, ld, center);
    return 0;
}

double orc_build_seconds(void* h) { return ((Index*)h)->build_seconds; }
int64_t orc_limited_buckets(void* h) { return ((Index*)h)->limited_buckets; }
int64_t orc_num_ids(void* h) { return (int64_t)((Index*)h)->ids.size(); }
int64_t orc_num_buckets(void* h) { return (int64_t)((Index*)h)->bucketSize.size(); }

void orc_get_center(void* h, float* c) {
    Index& I = *(Index*)h;
    std::copy(I.center.begin(), I.center.end(), c);
}

void orc_get_sign_bits(void* h, uint8_t* b1, uint8_t* b2) {
    Index& I = *(Index*)h;
    std::copy(I.bit1.begin(), I.bit1.end(), b1);
    std::copy(I.bit2.begin(), I.bit2.end(), b2);
}

void orc_get_buckets(void* h, uint64_t* pos, uint32_t* size, int32_t* ids) {
    Index& I = *(Index*)h;
    std::copy(I.bucketPos.begin(), I.bucketPos.end(), pos);
    std::copy(I.bucketSize.begin(), I.bucketSize.end(), size);
    std::copy(I.ids.begin(), I.ids.end(), ids);
}

// Build-side hash of m vectors (col-major D x m, already centred by the
// caller) for every table: codes/vals[(l*m + j)*iProbes + i] in the order the
// pair min-heap pops them (ascending value), buckets within the table.
// Requires the sign bits (i.e. a prior orc_build or orc_set_sign_bits).
void orc_hash_build(void* h, const float* V, int m, int32_t* codes, float* vals) {
    Index& I = *(Index*)h;
    const int L = I.n_tables, p = I.iProbes;
#pragma omp parallel for schedule(dynamic, 1)
    for (int l = 0; l < L; ++l) {
        std::vector<float> r1(I.fhtDim), r2(I.fhtDim);
        std::vector<IFPair> v1, v2, popped;
        for (int j = 0; j < m; ++j) {
            rotate2(I, V + (size_t)j * I.n_features, l, r1.data(), r2.data());
            layer_top(I, r1.data(), p, v1);
            layer_top(I, r2.data(), p, v2);
            build_pairs(I, v1, v2, popped);
            for (int i = 0; i < p; ++i) {
                codes[((size_t)l * m + j) * p + i] = popped[i].m_iIndex;
                vals[((size_t)l * m + j) * p + i] = popped[i].m_fValue;
            }
        }
    }
}

// Adopt an index built elsewhere (the engine's CSR: off[L*NB+1], ids) so the
// oracle's query path can be timed or checked on identical tables.  X raw
// (col-major D x N, leading dimension ld), centre given; sign bits from seed.
int orc_load_index(void* h, const float* X, int64_t ld, const float* center, const uint64_t* off,
                   const int32_t* ids) {
    Index& I = *(Index*)h;
    const int N = I.n_points, D = I.n_features;
    I.center.assign(center, center + D);
    I.X.assign((size_t)D * N, 0.0f);
    for (int n = 0; n < N; ++n)
        for (int d = 0; d < D; ++d) I.X[(size_t)n * D + d] = X[(size_t)n * ld + d] - I.center[d];
    const int nbits = I.fhtDim * I.n_tables * I.n_rotate;
    I.bit1.assign(nbits, 0);
    I.bit2.assign(nbits, 0);
    sign_bits((unsigned)I.seed, nbits, I.bit1.data(), I.bit2.data());
    const size_t nbk = (size_t)I.n_tables * I.nb();
    I.bucketPos.assign(off, off + nbk);
    I.bucketSize.resize(nbk);
    for (size_t b = 0; b < nbk; ++b) I.bucketSize[b] = (uint32_t)(off[b + 1] - off[b]);
    I.ids.assign(ids, ids + off[nbk]);
    I.built = true;
    return 0;
}

// Install sign bits without building (for hashing tests).
void orc_set_sign_bits(void* h, unsigned seed) {
    Index& I = *(Index*)h;
    const int nbits = I.fhtDim * I.n_tables * I.n_rotate;
    I.bit1.assign(nbits, 0);
    I.bit2.assign(nbits, 0);
    sign_bits(seed, nbits, I.bit1.data(), I.bit2.data());
}

// Query-side per-layer tops for one vector and table: codes/vals of length
// perLayer each, largest first (src/FalconnPP.cpp:438-453).
void orc_hash_query_layers(void* h, const float* q, int l, int perLayer, int32_t* c1, float* v1o, int32_t* c2,
                           float* v2o) {
    Index& I = *(Index*)h;
    std::vector<float> r1(I.fhtDim), r2(I.fhtDim);
    std::vector<IFPair> v1, v2;
    rotate2(I, q, l, r1.data(), r2.data());
    layer_top(I, r1.data(), perLayer, v1);
    layer_top(I, r2.data(), perLayer, v2);
    for (int i = 0; i < perLayer; ++i) {
        c1[i] = v1[i].m_iIndex;
        v1o[i] = v1[i].m_fValue;
        c2[i] = v2[i].m_iIndex;
        v2o[i] = v2[i].m_fValue;
    }
}

int orc_per_table(void* h) { return per_table(*(Index*)h); }
int orc_per_layer(void* h) { return per_layer(*(Index*)h); }

// Whole batch (src/FalconnPP.cpp:310-707): Q col-major D x nq (raw, not
// centred).  topk/scores are nq x k row-major; stats[0]=sum C_tot,
// stats[1]=sum C_uniq, stats[2]=empty buckets probed.
int orc_query(void* h, const float* Q, int nq, int k, int32_t* topk, float* scores, int64_t* stats) {
    Index& I = *(Index*)h;
    int64_t ctot = 0, cuniq = 0, emp = 0;
    omp_set_num_threads(I.n_threads);
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : ctot, cuniq, emp)
    for (int q = 0; q < nq; ++q) {
        QueryDebug dbg;
        query_one(I, Q + (size_t)q * I.n_features, k, topk + (size_t)q * k, scores + (size_t)q * k, &dbg);
        ctot += dbg.ctot;
        cuniq += (int64_t)dbg.cands.size();
        emp += dbg.empty;
    }
    if (stats) { stats[0] = ctot; stats[1] = cuniq; stats[2] = emp; }
    return 0;
}

// One query with full debug output.  probes/probeVals: qProbes entries in
// probing order; cands: up to cap unique candidates ascending; returns the
// number of unique candidates (may exceed cap); *ctot = sum of probed sizes.
int64_t orc_query_debug(void* h, const float* q, int k, int32_t* topk, float* scores, int32_t* probes,
                        float* probeVals, int32_t* cands, int64_t cap, int64_t* ctot) {
    Index& I = *(Index*)h;
    QueryDebug dbg;
    query_one(I, q, k, topk, scores, &dbg);
    for (size_t i = 0; i < dbg.probes.size(); ++i) {
        probes[i] = dbg.probes[i];
        probeVals[i] = dbg.probeVals[i];
    }
    for (int64_t i = 0; i < (int64_t)dbg.cands.size() && i < cap; ++i) cands[i] = dbg.cands[i];
    *ctot = dbg.ctot;
    return (int64_t)dbg.cands.size();
}

// Exact FP64 top-k inner product over all N (src/BF.cpp:17-87 in double):
// X col-major D x N (uncentred), Q col-major D x nq; outputs nq x k row-major,
// largest first.  Same min-heap with strict '>' over ascending ids.
void orc_bruteforce_f64(const float* X, int N, int D, const float* Q, int nq, int k, int32_t* ids,
                        double* scores, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
    for (int q = 0; q < nq; ++q) {
        std::priority_queue<std::pair<double, int>, std::vector<std::pair<double, int>>, std::greater<>> tk;
        const float* qv = Q + (size_t)q * D;
        for (int n = 0; n < N; ++n) {
            const float* x = X + (size_t)n * D;
            double s = 0;
            for (int d = 0; d < D; ++d) s += (double)qv[d] * (double)x[d];
            if ((int)tk.size() < k) tk.emplace(s, -n);
            else if (s > tk.top().first) { tk.pop(); tk.emplace(s, -n); }
        }
        for (int i = k - 1; i >= 0; --i) {
            if (tk.empty()) { ids[(size_t)q * k + i] = -1; scores[(size_t)q * k + i] = 0; continue; }
            ids[(size_t)q * k + i] = -tk.top().second;
            scores[(size_t)q * k + i] = tk.top().first;
            tk.pop();
        }
    }
}

}  // extern "C"
