// Host side of the gfx950 Falconn++ engine: index state, the C ABI and the
// orchestration of the kernels (hash.hip, build.hip, query.hip, bf.hip).
//
// Reference mapping: FalconnPP class state src/FalconnPP.h:10-132,
// build2Layers_1D src/FalconnPP.cpp:21-303, query2Layers_1D :310-707.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <functional>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/falconnpp_amd.h"
#include "common.h"
#include "kernels.h"
#include "textio.h"

namespace fpp {
namespace {

thread_local std::string g_last_error;

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // (re)allocate to at least m elements; contents are not preserved
    void reserve(size_t m) {
        if (m <= n && p) return;
        reset();
        const size_t bytes = std::max<size_t>(m, 1) * sizeof(T);
        FPP_HIP(hipMalloc(&p, bytes));
        n = std::max<size_t>(m, 1);
    }
    T* get() const { return p; }
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
};

struct Index {
    int n_points = 0, n_features = 0;
    // src/FalconnPP.h:17-28 defaults
    int n_tables = 10, n_proj = 512, n_rotate = 3, n_threads = -1;
    int bucket_minSize = 50;
    float bucket_scale = 1.0f;
    int iProbes = 1, qProbes = 1, seed = -1;
    int fhtDim = 0;
    bool params = false, built = false;

    hipStream_t stream = nullptr;
    int64_t ldx = 0;
    DevBuf<float> Xc;
    DevBuf<float> center;
    DevBuf<uint32_t> masks;
    DevBuf<uint64_t> bucket_off;
    DevBuf<int32_t> ids;
    int64_t n_ids = 0;
    DevBuf<uint32_t> pre_cnt;  // bucket sizes before the alpha / bucketLimit filter (built indexes)
    bool have_pre_cnt = false;
    // insertable index: every point's pre-filter entries, [l][n*p + i]
    bool insertable = false, entries_kept = false;
    DevBuf<uint32_t> entry_codes;
    DevBuf<float> entry_vals;
    std::vector<uint8_t> bit1, bit2;

    // bucket sizes of the built index, largest first above SIZE_BINS: they
    // bound any query's C_tot by the sum of its qProbes largest buckets
    uint32_t max_bucket = 0;
    std::vector<uint32_t> size_hist, size_big;
    bool size_big_complete = false;
    DevBuf<uint32_t> scratch_word;

    // query calls: one device workspace carved per call (query copy, hash
    // lists, flags, probes, counts, ordering, replay work, and the gather's
    // per-workgroup candidate scratch, at most ws_budget bytes of it)
    DevBuf<unsigned char> ws;
    size_t ws_budget = size_t(8) << 30;
    DevBuf<int32_t> out_ids;  // host-pointer query calls
    std::vector<QueryParams> hqp;  // per-query probe counts and k (batched serving)
    int gather_threads = 0;        // 0: chosen per call (gather_shape)
    int order_mode = 1;            // ORDER_* below
    bool keep_candidates = false;  // every query's candidates kept after the call (inspection)
    int replay_mode = 0;           // tie replay: 0 per call, 1 separate kernels, 2 inside the gather
    bool small_queries = true;     // queries of at most SMALL_MAX_CTOT bucket ids served one per wave
    DevBuf<int32_t> keep_cand, keep_part;
    DevBuf<uint64_t> keep_off;
    int wall_khz = 0;
    // the whole query pipeline of a call captured in a hipGraph, replayed
    // while the call's shape and buffers stay the same
    bool use_graph = false;
    bool tune_verbose = false;  // autotuner prints each timed shape
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_key;
    uint64_t build_gen = 0;
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_key.clear();
    }
    // the last query call: its device arrays (statistics are summed from them
    // when asked for) and, in keep-candidates mode, every query's segment
    struct LastCall {
        int nq = 0;
        uint64_t nprobes = 0;
        const uint32_t* probes = nullptr;
        const uint64_t* ctot = nullptr;
        const uint64_t* ncand = nullptr;
        const uint32_t* qflag = nullptr;
        const uint32_t* nsmall = nullptr;  // small queries (device word), or null
        hipStream_t st = nullptr;
        bool summed = true;
        std::vector<uint64_t> seg;  // keep-candidates: segment offsets, nq + 1
    } last;
    std::vector<uint64_t> last_ts;  // per-query phase stamps of the last stamped call

    int64_t stat_hash_ties = 0, stat_bucket_ties = 0, stat_ctot = 0, stat_cuniq = 0, stat_qties = 0, stat_small = 0;
    int stat_probe_path = -1;  // probe selection of the last query call (PROBE_*)

    // reduced-precision scoring: 0 = FP32 (reference semantics), 1 = BF16,
    // 2 = FP8-e4m3 + per-row scale; survivors rescored exactly in FP32
    int precision = 0;
    int refine = 0;  // 0: 4 * n_neighbors
    int lp_ready = 0;  // format of the current low-precision copy
    bool fp8_per_row = false;      // FP8 rows keep their own scale even when one could be shared
    bool lp_per_row_made = false;  // the FP8 copy was made with that setting
    int64_t ldl = 0;
    DevBuf<uint16_t> Xb;
    DevBuf<uint8_t> X8;
    DevBuf<float> s8;
    float s8_shared = 0.0f;  // FP8 scale shared by every row, 0 when rows keep their own
    int refine_count(int k) const { return refine > 0 ? refine : std::min(256, 4 * k); }

    // optional per-stage timing of query calls (HIP events on the call's stream)
    bool profiling = false;   // HIP events around the query stages
    bool phase_stamps = false;  // plus per-query phase stamps in the gather kernel
    hipEvent_t ev[8] = {};
    // hash, probe select, tie replay, gather+rerank (ms); then mean per-query
    // microseconds of the gather kernel's phases: bucket gather + dedup,
    // candidate compaction, scoring, top-k
    double stage_ms[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    void mark(int i, hipStream_t st) {
        if (!profiling) return;
        if (!ev[i]) FPP_HIP(hipEventCreate(&ev[i]));
        FPP_HIP(hipEventRecord(ev[i], st));
    }

    int log2_D() const { return ilog2_exact(n_proj); }
    int NB() const { return 4 * n_proj * n_proj; }
    HashGeom hgeom() const { return HashGeom{fhtDim, n_features, n_proj, log2_D(), n_tables}; }
    ~Index() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        drop_graph();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

int fail(const Error& e) {
    g_last_error = e.what();
    return e.code;
}

template <class F>
int guard(F&& f) {
    try {
        f();
        g_last_error.clear();
        return FPP_OK;
    } catch (const Error& e) {
        return fail(e);
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return FPP_ENOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return FPP_EINTERNAL;
    }
}

hipStream_t pick_stream(Index& I, void* s) {
    if (s) return reinterpret_cast<hipStream_t>(s);
    if (!I.stream) FPP_HIP(hipStreamCreateWithFlags(&I.stream, hipStreamNonBlocking));
    return I.stream;
}

int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// rows x D floats from host src (row stride ld floats) to dense device dst:
// one plain copy when the rows are contiguous (a 2D copy of many short rows
// from pageable memory goes row by row)
void copy_rows_h2d(float* dst, const float* src, int64_t ld, int64_t rows, int D, hipStream_t st) {
    if (rows <= 0) return;
    if (ld == D)
        FPP_HIP(hipMemcpyAsync(dst, src, (size_t)rows * D * 4, hipMemcpyHostToDevice, st));
    else
        FPP_HIP(hipMemcpy2DAsync(dst, (size_t)D * 4, src, (size_t)ld * 4, (size_t)D * 4, rows, hipMemcpyHostToDevice,
                                 st));
}

// src/FalconnPP.h:52-75 with libstdc++'s own engine and distribution.
void make_sign_bits(Index& I) {
    const int nbits = I.fhtDim * I.n_tables * I.n_rotate;
    unsigned seed = (unsigned)std::chrono::system_clock::now().time_since_epoch().count();
    if (I.seed > -1) seed = (unsigned)I.seed;
    std::default_random_engine gen(seed);
    std::uniform_int_distribution<uint32_t> unif(0, 1);
    I.bit1.assign(nbits, 0);
    I.bit2.assign(nbits, 0);
    for (int d = 0; d < nbits; ++d) {
        I.bit1[d] = (uint8_t)(unif(gen) & 1);
        I.bit2[d] = (uint8_t)(unif(gen) & 1);
    }
    // packed flip bits: word ((l*2 + layer)*3 + r)*MW + d/32, bit d%32 set when
    // the reference's sign bit is 0, i.e. factor -1 (src/FalconnPP.cpp:87-88)
    const int F = I.fhtDim, MW = (F + 31) / 32;
    std::vector<uint32_t> m((size_t)I.n_tables * 2 * 3 * MW, 0u);
    for (int l = 0; l < I.n_tables; ++l)
        for (int layer = 0; layer < 2; ++layer)
            for (int r = 0; r < 3; ++r)
                for (int d = 0; d < F; ++d) {
                    const size_t bit = (size_t)l * 3 * F + (size_t)r * F + d;
                    const uint8_t b = layer == 0 ? I.bit1[bit] : I.bit2[bit];
                    if (!b) m[(((size_t)l * 2 + layer) * 3 + r) * MW + d / 32] |= 1u << (d % 32);
                }
    I.masks.reserve(m.size());
    FPP_HIP(hipMemcpy(I.masks.get(), m.data(), m.size() * 4, hipMemcpyHostToDevice));
}

void check_params(const Index& I) {
    FPP_REQUIRE(I.params, FPP_ESTATE, "setIndexParam must be called before build");
    FPP_REQUIRE(I.n_points >= 1 && I.n_features >= 1, FPP_EINVAL, "n_points and n_features must be >= 1");
    FPP_REQUIRE(I.n_tables >= 1, FPP_EINVAL, "n_tables must be >= 1");
    FPP_REQUIRE(I.log2_D() >= 0, FPP_EINVAL, "n_proj must be a power of two (src/FalconnPP.cpp:46 truncates log2)");
    FPP_REQUIRE(I.fhtDim >= 2 && I.fhtDim <= 1024, FPP_EINVAL, "fhtDim = 2^ceil(log2(max(n_proj, n_features))) must be in [2, 1024]");
    FPP_REQUIRE(I.iProbes >= 1 && I.iProbes <= std::min(15, I.n_proj), FPP_EINVAL,
                "iProbes must be in [1, min(15, n_proj)]");
    FPP_REQUIRE((int64_t)I.n_points * I.iProbes < (int64_t)UINT32_MAX, FPP_EINVAL, "n_points * iProbes must be < 2^32");
    FPP_REQUIRE((int64_t)I.n_tables * I.NB() < (int64_t)INT32_MAX, FPP_EINVAL, "n_tables * 4 * n_proj^2 must be < 2^31");
}

// Bucket sizes of a built or loaded index: the largest, and the histogram
// that bounds C_tot per query.
void index_sizes(Index& I, hipStream_t st) {
    const int64_t nbk = (int64_t)I.n_tables * I.NB();
    I.scratch_word.reserve(1);
    I.max_bucket = max_bucket_size(I.bucket_off.get(), nbk, I.scratch_word.get(), st);
    I.size_big_complete = bucket_size_histogram(I.bucket_off.get(), nbk, I.size_hist, I.size_big, st);
    std::sort(I.size_big.begin(), I.size_big.end(), std::greater<uint32_t>());
}

// Build-side hash of the N centred rows at Xc (row stride I.ldx) into
// codes/vals[l*N*p + n*p + i], the items whose result depends on an exact
// float tie replayed with the reference's heaps; returns how many were.
int64_t hash_rows(Index& I, const float* Xc, int64_t N, uint32_t* codes, float* vals, hipStream_t st) {
    const HashGeom g = I.hgeom();
    const int p = I.iProbes, L = I.n_tables;
    int64_t ties = 0;
    // hashing + exact replay of tie-dependent items
    {
        const uint32_t cap = (uint32_t)std::min<int64_t>(N * L, std::max<int64_t>(1 << 16, N * L / 16));
        DevBuf<uint64_t> flags;
        DevBuf<uint32_t> fcount;
        flags.reserve(cap);
        fcount.reserve(1);
        FPP_HIP(hipMemsetAsync(fcount.get(), 0, 4, st));
        launch_hash_build(g, Xc, I.ldx, N, p, I.masks.get(), codes, vals, flags.get(),
                          fcount.get(), cap, st);
        uint32_t nflag = 0;
        FPP_HIP(hipMemcpyAsync(&nflag, fcount.get(), 4, hipMemcpyDeviceToHost, st));
        FPP_HIP(hipStreamSynchronize(st));
        const bool all = nflag > cap;
        const uint64_t count = all ? (uint64_t)N * L : nflag;
        ties = (int64_t)nflag;
        if (count) {
            // bounded workspace: replay in slices
            const int64_t per_item = 2 * (int64_t)I.fhtDim + 8 * p + 8;
            const int64_t slice = std::max<int64_t>(1, std::min<int64_t>(count, (int64_t)(1 << 28) / per_item));
            DevBuf<float> work;
            work.reserve((size_t)slice * per_item);
            for (uint64_t s = 0; s < count; s += slice) {
                const uint64_t c = std::min<uint64_t>(slice, count - s);
                if (all) {
                    // all_items mode walks items 0..c-1: offset via a temporary list
                    std::vector<uint64_t> lst(c);
                    for (uint64_t i = 0; i < c; ++i) lst[i] = s + i;
                    DevBuf<uint64_t> dl;
                    dl.reserve(c);
                    FPP_HIP(hipMemcpyAsync(dl.get(), lst.data(), c * 8, hipMemcpyHostToDevice, st));
                    launch_hash_build_exact(g, Xc, I.ldx, N, p, I.masks.get(), codes, vals,
                                            dl.get(), c, 0, work.get(), st);
                    FPP_HIP(hipStreamSynchronize(st));
                } else {
                    launch_hash_build_exact(g, Xc, I.ldx, N, p, I.masks.get(), codes, vals,
                                            flags.get() + s, c, 0, work.get(), st);
                }
            }
            FPP_HIP(hipStreamSynchronize(st));
        }
    }
    return ties;
}

// Bucket fill and filter over the pre-filter entries of N points, laid out
// as hash_rows writes them (entry n*p + i of table l belongs to point n, so
// a stable sort by bucket keeps the reference's insertion order): bucket
// sizes, the alpha / bucketLimit rule, CSR offsets, then per table the copy
// of whole buckets, the selection of the iLimit largest values and the exact
// replay of buckets whose cut falls between equal values.
void fill_tables(Index& I, const uint32_t* codes, const float* vals, int64_t N, hipStream_t st) {
    const int L = I.n_tables, p = I.iProbes, NB = I.NB();
    const int64_t per_table = N * p;
    // bucket sizes, limits and CSR offsets over all tables
    const int64_t nbk = (int64_t)L * NB;
    // pre-filter bucket sizes stay with the index (inspection)
    DevBuf<uint32_t>& cnt = I.pre_cnt;
    DevBuf<uint32_t> ksz;
    cnt.reserve(nbk);
    ksz.reserve(nbk);
    FPP_HIP(hipMemsetAsync(cnt.get(), 0, nbk * 4, st));
    launch_bucket_count(codes, per_table, L, NB, cnt.get(), st);
    launch_bucket_limit(cnt.get(), nbk, I.bucket_scale, p, I.bucket_minSize, ksz.get(), st);
    const int64_t radix_hist = radix_hist_elems(per_table);
    DevBuf<uint64_t> scan_tmp;
    scan_tmp.reserve(std::max(scan_tmp_elems(nbk), std::max(scan_tmp_elems(radix_hist), scan_tmp_elems(NB))));
    I.bucket_off.reserve(nbk + 1);
    scan_u32_to_u64(ksz.get(), nbk, I.bucket_off.get(), scan_tmp.get(), st);
    uint64_t kept = 0;
    FPP_HIP(hipMemcpyAsync(&kept, I.bucket_off.get() + nbk, 8, hipMemcpyDeviceToHost, st));
    FPP_HIP(hipStreamSynchronize(st));
    I.ids.reserve(kept + IDS_PAD);
    I.n_ids = (int64_t)kept;
    // per table: stable sort by bucket, copy / select / exact replay
    DevBuf<uint32_t> k0, v0, k1, v1, hist, sel, selc, ties, tiec;
    DevBuf<uint64_t> hist_off, tstart;
    k0.reserve(per_table);
    v0.reserve(per_table);
    k1.reserve(per_table);
    v1.reserve(per_table);
    hist.reserve(radix_hist);
    hist_off.reserve(radix_hist + 1);
    tstart.reserve(NB + 1);
    sel.reserve(NB);
    selc.reserve(2);
    ties.reserve(NB);
    tiec.reserve(1);
    const int key_bits = ilog2_exact(NB);
    int64_t bucket_ties = 0;
    for (int l = 0; l < L; ++l) {
        const uint32_t* cnt_l = cnt.get() + (size_t)l * NB;
        const uint32_t* ksz_l = ksz.get() + (size_t)l * NB;
        const uint64_t* off_l = I.bucket_off.get() + (size_t)l * NB;
        const float* vals_l = vals + (size_t)l * per_table;
        FPP_HIP(hipMemcpyAsync(k0.get(), codes + (size_t)l * per_table, per_table * 4, hipMemcpyDeviceToDevice,
                               st));
        uint32_t *sk = nullptr, *sv = nullptr;
        radix_sort_pairs(k0.get(), v0.get(), k1.get(), v1.get(), per_table, key_bits, true, hist.get(),
                         hist_off.get(), scan_tmp.get(), &sk, &sv, st);
        scan_u32_to_u64(cnt_l, NB, tstart.get(), scan_tmp.get(), st);
        launch_seg_copy(sk, sv, per_table, p, cnt_l, ksz_l, tstart.get(), off_l, I.ids.get(), st);
        FPP_HIP(hipMemsetAsync(selc.get(), 0, 8, st));
        launch_make_select_list(cnt_l, ksz_l, NB, sel.get(), selc.get(), st);
        uint32_t nsel = 0;
        FPP_HIP(hipMemcpyAsync(&nsel, selc.get(), 4, hipMemcpyDeviceToHost, st));
        FPP_HIP(hipStreamSynchronize(st));
        if (nsel) {
            FPP_HIP(hipMemsetAsync(tiec.get(), 0, 4, st));
            launch_seg_select(sel.get(), nsel, sv, vals_l, p, cnt_l, ksz_l, tstart.get(), off_l, I.ids.get(),
                              ties.get(), tiec.get(), st);
            uint32_t ntie = 0;
            FPP_HIP(hipMemcpyAsync(&ntie, tiec.get(), 4, hipMemcpyDeviceToHost, st));
            FPP_HIP(hipStreamSynchronize(st));
            if (ntie) {
                bucket_ties += ntie;
                std::vector<uint32_t> tl(ntie), bsz(ntie);
                FPP_HIP(hipMemcpy(tl.data(), ties.get(), ntie * 4, hipMemcpyDeviceToHost));
                std::vector<uint32_t> cnt_h(NB);
                FPP_HIP(hipMemcpy(cnt_h.data(), cnt_l, (size_t)NB * 4, hipMemcpyDeviceToHost));
                std::vector<uint64_t> woff(ntie);
                uint64_t run = 0;
                for (uint32_t i = 0; i < ntie; ++i) {
                    woff[i] = run;
                    run += round_up((int64_t)cnt_h[tl[i]] * 9, 16);
                }
                DevBuf<uint64_t> dwoff;
                DevBuf<unsigned char> work;
                dwoff.reserve(ntie);
                work.reserve(run);
                FPP_HIP(hipMemcpy(dwoff.get(), woff.data(), ntie * 8, hipMemcpyHostToDevice));
                launch_seg_exact(ties.get(), ntie, dwoff.get(), sv, vals_l, p, cnt_l, ksz_l, tstart.get(), off_l,
                                 I.ids.get(), work.get(), st);
                FPP_HIP(hipStreamSynchronize(st));
            }
        }
    }
    FPP_HIP(hipStreamSynchronize(st));
    index_sizes(I, st);
    I.stat_bucket_ties = bucket_ties;
    I.have_pre_cnt = true;
}

void do_build(Index& I, const float* dX, int64_t ld, hipStream_t st) {
    check_params(I);
    const int64_t N = I.n_points;
    const int D = I.n_features, L = I.n_tables, p = I.iProbes, NB = I.NB();
    FPP_REQUIRE(ld >= D, FPP_EINVAL, "leading dimension of X must be >= n_features");
    I.built = false;
    I.lp_ready = 0;
    make_sign_bits(I);
    // centre (deterministic FP64 chunk sums) and the centred copy
    I.ldx = round_up(D, 4);
    const int64_t chunk = 4096;
    const int nchunks = (int)ceil_div(N, chunk);
    {
        DevBuf<double> partial;
        partial.reserve((size_t)nchunks * D);
        I.center.reserve(D);
        launch_center_partial(dX, ld, N, D, partial.get(), chunk, nchunks, st);
        launch_center_final(partial.get(), nchunks, D, N, I.center.get(), st);
        I.Xc.reserve((size_t)N * I.ldx);
        launch_center_apply(dX, ld, N, D, I.center.get(), I.Xc.get(), I.ldx, st);
        FPP_HIP(hipStreamSynchronize(st));
    }
    const int64_t total = N * p * L;
    DevBuf<uint32_t> codes;
    DevBuf<float> vals;
    codes.reserve(total);
    vals.reserve(total);
    I.stat_hash_ties = hash_rows(I, I.Xc.get(), N, codes.get(), vals.get(), st);
    fill_tables(I, codes.get(), vals.get(), N, st);
    // an insertable index keeps every point's pre-filter entries
    I.entry_codes.reset();
    I.entry_vals.reset();
    if (I.insertable) {
        I.entry_codes.swap(codes);
        I.entry_vals.swap(vals);
    }
    I.entries_kept = I.insertable;
    I.build_gen++;
    I.drop_graph();
    I.built = true;
}

// Appends m points (dX: D x m col-major, leading dimension ld) as ids
// n_points .. n_points+m-1: centred with the index's centre, hashed with its
// sign bits, their entries placed after the old points' in every table, and
// every bucket filtered again over old and new entries.  The index then
// equals a build of all points with that centre.
void do_insert(Index& I, const float* dX, int64_t ld, int m, hipStream_t st) {
    FPP_REQUIRE(I.built, FPP_ESTATE, "insert called before build (or after clear)");
    FPP_REQUIRE(I.entries_kept, FPP_ESTATE,
                "the index keeps no pre-filter entries: build it after fpp_set_insertable(h, 1)");
    FPP_REQUIRE(m >= 1, FPP_EINVAL, "number of inserted points must be >= 1");
    FPP_REQUIRE(ld >= I.n_features, FPP_EINVAL, "leading dimension of X must be >= n_features");
    const int64_t N0 = I.n_points, N1 = N0 + m;
    const int D = I.n_features, L = I.n_tables, p = I.iProbes;
    FPP_REQUIRE(N1 <= INT32_MAX && N1 * p < (int64_t)UINT32_MAX, FPP_EINVAL, "n_points * iProbes must be < 2^32");
    // centred rows: the old ones, then the new ones with the same centre
    {
        DevBuf<float> Xn;
        Xn.reserve((size_t)N1 * I.ldx);
        FPP_HIP(hipMemcpyAsync(Xn.get(), I.Xc.get(), (size_t)N0 * I.ldx * 4, hipMemcpyDeviceToDevice, st));
        launch_center_apply(dX, ld, m, D, I.center.get(), Xn.get() + (size_t)N0 * I.ldx, I.ldx, st);
        FPP_HIP(hipStreamSynchronize(st));
        I.Xc.swap(Xn);
    }
    DevBuf<uint32_t> nc;
    DevBuf<float> nv;
    nc.reserve((size_t)m * p * L);
    nv.reserve((size_t)m * p * L);
    const int64_t ties = hash_rows(I, I.Xc.get() + (size_t)N0 * I.ldx, m, nc.get(), nv.get(), st);
    DevBuf<uint32_t> codes;
    DevBuf<float> vals;
    codes.reserve((size_t)N1 * p * L);
    vals.reserve((size_t)N1 * p * L);
    const size_t o0 = (size_t)N0 * p, o1 = (size_t)N1 * p, om = (size_t)m * p;
    for (int l = 0; l < L; ++l) {
        FPP_HIP(hipMemcpyAsync(codes.get() + l * o1, I.entry_codes.get() + l * o0, o0 * 4, hipMemcpyDeviceToDevice, st));
        FPP_HIP(hipMemcpyAsync(codes.get() + l * o1 + o0, nc.get() + l * om, om * 4, hipMemcpyDeviceToDevice, st));
        FPP_HIP(hipMemcpyAsync(vals.get() + l * o1, I.entry_vals.get() + l * o0, o0 * 4, hipMemcpyDeviceToDevice, st));
        FPP_HIP(hipMemcpyAsync(vals.get() + l * o1 + o0, nv.get() + l * om, om * 4, hipMemcpyDeviceToDevice, st));
    }
    nc.reset();
    nv.reset();
    I.entry_codes.reset();
    I.entry_vals.reset();
    I.built = false;
    I.n_points = (int)N1;
    fill_tables(I, codes.get(), vals.get(), N1, st);
    I.entry_codes.swap(codes);
    I.entry_vals.swap(vals);
    I.stat_hash_ties = ties;
    I.lp_ready = 0;
    I.Xb.reset();
    I.X8.reset();
    I.s8.reset();
    I.build_gen++;
    I.drop_graph();
    I.built = true;
}

// Low-precision copy of the centred data for the selected format, made once
// per build.
void ensure_lowprec(Index& I, hipStream_t st) {
    if (I.lp_ready == I.precision && (I.precision != 2 || I.lp_per_row_made == I.fp8_per_row)) return;
    const int64_t N = I.n_points;
    // rows padded to whole 128-byte lines: a scoring load instruction reads
    // 8 rows' lines whole (FP8: 128 elements, BF16: 64 per line)
    I.ldl = round_up(I.n_features, I.precision == 2 ? 128 : 64);
    if (I.precision == 1) {
        I.X8.reset();
        I.s8.reset();
        I.Xb.reserve((size_t)N * I.ldl);
        launch_quant_bf16(I.Xc.get(), I.ldx, N, I.n_features, I.ldl, I.Xb.get(), st);
    } else {
        I.Xb.reset();
        I.X8.reserve((size_t)N * I.ldl);
        I.s8.reserve(N);
        I.scratch_word.reserve(2);
        I.s8_shared = launch_quant_fp8(I.Xc.get(), I.ldx, N, I.n_features, I.ldl, I.X8.get(), I.s8.get(),
                                       I.scratch_word.get(), I.fp8_per_row, st);
        I.lp_per_row_made = I.fp8_per_row;
    }
    I.lp_ready = I.precision;
}

struct QueryPlan {
    int perTable, perLayer;
};

// src/FalconnPP.cpp:336-337 in double arithmetic
QueryPlan plan_for(const Index& I, int qProbes) {
    QueryPlan pl;
    pl.perTable = (int)std::ceil(2.0 * qProbes / I.n_tables);
    pl.perLayer = (int)std::ceil(std::sqrt(1.0 * pl.perTable));
    return pl;
}

// Probes per layer that the fast hashing and probe-selection kernels handle
// (top lists of at most 32 entries, PL^2 pair sums in registers).  A plain
// query call with more replays every query exactly instead.
constexpr int FAST_MAX_PL = 31;

void check_plan(const Index& I, const QueryPlan& pl, bool batched) {
    FPP_REQUIRE(pl.perLayer <= I.n_proj, FPP_EINVAL,
                "qProbes too large: ceil(sqrt(ceil(2*qProbes/n_tables))) exceeds n_proj (the reference pops an empty heap)");
    FPP_REQUIRE(!batched || pl.perLayer <= FAST_MAX_PL, FPP_EINVAL,
                "qProbes too large for a batched call: more than 31 probes per layer");
}

// Order in which the persistent gather takes the queries.
enum { ORDER_IDENTITY = 0, ORDER_BUCKET = 1, ORDER_LONGEST = 2 };

// Upper bound of C_tot over queries with at most qp probes: the sum of the qp
// largest buckets of the index (a query's probes are distinct buckets).
uint64_t ctot_bound(const Index& I, uint64_t qp) {
    if (!I.size_big_complete) return qp * (uint64_t)I.max_bucket;
    uint64_t s = 0, left = qp;
    for (uint32_t v : I.size_big) {
        if (!left) return s;
        s += v;
        --left;
    }
    for (int b = (int)I.size_hist.size() - 1; b > 0 && left; --b) {
        const uint64_t c = std::min<uint64_t>(left, I.size_hist[b]);
        s += c * (uint64_t)b;
        left -= c;
    }
    return s;
}

// Byte offsets of one call's arrays in its device workspace.
struct QueryLayout {
    size_t Qpad, qc, qv, qflag, probes, ctot, ncand, qparams, okey, oval, okey2, oval2, ohist, ohist_off, oscan, qlist,
        counters, rwork, phase_ts, cand, part, total;
};
// [0, 8) queue heads, [8] replay claim, [9] flagged queries, [10] small
// queries
constexpr int N_COUNTERS = 16;

QueryLayout query_layout(int nq, int L, int PL, uint64_t nprobes, bool batched, unsigned rareas, int64_t wpq,
                         bool stamps, unsigned slots, uint64_t slot_cand, uint64_t slot_part) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        off = (off + 255) & ~size_t(255);
        const size_t o = off;
        off += std::max<size_t>(bytes, 1);
        return o;
    };
    QueryLayout w{};
    const size_t n = (size_t)nq;
    const int64_t rh = radix_hist_elems(nq);
    w.Qpad = 0;  // filled in by the caller's ldx
    w.qc = take(n * L * 2 * PL * 4);
    w.qv = take(n * L * 2 * PL * 4);
    w.qflag = take(n * 4);
    w.probes = take(nprobes * 4);
    w.ctot = take(n * 8);
    w.ncand = take(n * 8);
    w.qparams = take(batched ? n * sizeof(QueryParams) : 0);
    w.okey = take(n * 4);
    w.oval = take(n * 4);
    w.okey2 = take(n * 4);
    w.oval2 = take(n * 4);
    w.ohist = take((size_t)rh * 4);
    w.ohist_off = take(((size_t)rh + 1) * 8);
    w.oscan = take((size_t)scan_tmp_elems(rh) * 8);
    w.qlist = take(n * 4);
    w.counters = take(N_COUNTERS * 4);
    w.rwork = take((size_t)rareas * (size_t)wpq);
    w.phase_ts = take(stamps ? n * PHASE_STAMPS * 8 : 0);
    w.cand = take((size_t)slots * slot_cand * 4);
    w.part = take((size_t)slots * slot_part * 4);
    w.total = (off + 255) & ~size_t(255);
    return w;
}

// Launch shape of the gather: workgroup size and workgroups per CU.  The
// tuned size, or automatically: when even a workgroup alone on its CU needs
// several dedup windows, two workgroups per CU (one query's candidate
// collection then overlaps another's scoring, which outweighs their smaller
// windows): 256 threads for the FP32 gather, 512 with half the rows in
// flight per wave for BF16/FP8; else one 512-thread workgroup per CU.
// 256-thread workgroups always run two per CU.
struct GatherShape {
    int threads;
    int per_cu;
};
GatherShape gather_shape(const Index& I, bool lowp, int refine) {
    if (I.gather_threads == 256) return {256, 2};
    if (I.gather_threads) return {lowp ? 512 : I.gather_threads, 1};
    const int64_t extra = lowp ? lowprec_extra_words(refine, 512) : 512 + 256 + 2048;
    const bool windows = I.n_points > gather_window_bits(I.n_points, extra, 1);
    if (!windows) return {512, 1};
    return lowp ? GatherShape{512, 2} : GatherShape{256, 2};
}

// One batch of queries.  k is the number of neighbours, or with per-query
// counts (h_qprobes, h_k: host arrays of nq) the largest k and the row stride
// of the outputs.  Everything runs on st in one sequence with no host read:
//  1. queries copied (padded), hashed (k_hash_query), probes selected
//     (k_probe_select), the tie-flagged queries listed on the device and
//     their probes replayed exactly (k_probe_replay), so C_tot is exact;
//  2. the queries ordered on the device (by their best table-0 bucket, in 8
//     runs for the 8 XCDs, or longest first) for the persistent gather, whose
//     workgroups each own a scratch slot bounded by the C_tot bound above;
//  3. one persistent gather (dedup, scoring, top-k).
// With the graph on, the whole sequence is captured once and replayed.
// fixed_ws (no allocation at all): the workspace is fixed_bytes there.
// sync_end: wait for the results before returning.
void do_query(Index& I, const float* dQ, int64_t ldq, int nq, int k, int32_t* d_out, float* d_scores,
              hipStream_t st, const int32_t* h_qprobes = nullptr, const int32_t* h_k = nullptr,
              unsigned char* fixed_ws = nullptr, size_t fixed_bytes = 0, bool sync_end = false) {
    FPP_REQUIRE(I.built, FPP_ESTATE, "query called before build (or after clear)");
    FPP_REQUIRE(nq >= 0, FPP_EINVAL, "number of queries must be >= 0");
    FPP_REQUIRE(k >= 1 && k <= 1024, FPP_EINVAL, "n_neighbors must be in [1, 1024]");
    FPP_REQUIRE(I.precision == 0 || k <= I.refine_count(k), FPP_EINVAL,
                "n_neighbors exceeds the refine count of the reduced-precision path");
    FPP_REQUIRE(ldq >= I.n_features, FPP_EINVAL, "leading dimension of Q must be >= n_features");
    FPP_REQUIRE((h_qprobes == nullptr) == (h_k == nullptr), FPP_EINVAL, "per-query qProbes and k go together");
    const bool batched = h_qprobes != nullptr;
    const bool fixed = fixed_ws != nullptr;
    QueryPlan pl;
    int qp_max = I.qProbes;
    uint64_t nprobes = 0;
    if (batched) {
        // per query: its qProbes, perTable, PL and k; sizes from the largest
        I.hqp.resize(nq);
        pl = QueryPlan{0, 0};
        qp_max = 0;
        for (int q = 0; q < nq; ++q) {
            FPP_REQUIRE(h_qprobes[q] >= 1, FPP_EINVAL, "qProbes must be >= 1");
            FPP_REQUIRE(h_k[q] >= 1 && h_k[q] <= k, FPP_EINVAL, "per-query n_neighbors must be in [1, kmax]");
            const QueryPlan p = plan_for(I, h_qprobes[q]);
            check_plan(I, p, true);
            I.hqp[q] = QueryParams{h_qprobes[q], p.perTable, p.perLayer, h_k[q], nprobes};
            nprobes += (uint64_t)h_qprobes[q];
            pl.perTable = std::max(pl.perTable, p.perTable);
            pl.perLayer = std::max(pl.perLayer, p.perLayer);
            qp_max = std::max(qp_max, h_qprobes[q]);
        }
    } else {
        FPP_REQUIRE(I.qProbes >= 1, FPP_EINVAL, "qProbes must be >= 1");
        pl = plan_for(I, I.qProbes);
        check_plan(I, pl, false);
        nprobes = (uint64_t)nq * I.qProbes;
    }
    // more probes per layer than the fast kernels take: every query's probes
    // come from the exact replay (its per-table lists are rebuilt from the
    // rotated query, qflag bit 0)
    const bool exact_only = pl.perLayer > FAST_MAX_PL;
    const int L = I.n_tables, PL = pl.perLayer;
    const HashGeom hg = I.hgeom();
    ProbeGeom pg{L, I.NB(), PL, pl.perTable, qp_max, 2 * I.n_proj};
    // a pool that fits no probe-selection kernel: every query's probes come
    // from the exact replay, which reuses the fast per-layer lists
    const int probe_path = exact_only ? PROBE_EXACT : probe_select_path(pg);
    I.stat_probe_path = probe_path;
    I.last = Index::LastCall{};
    I.last.nq = nq;
    I.last.nprobes = nprobes;
    if (nq == 0) return;

    // gather configuration
    const bool lowp = I.precision != 0;
    if (lowp) {
        if (fixed)
            FPP_REQUIRE(I.lp_ready == I.precision && (I.precision != 2 || I.lp_per_row_made == I.fp8_per_row),
                        FPP_ESTATE, "the reduced-precision copy is made by a first query through fpp_query_device");
        else
            ensure_lowprec(I, st);
    }
    // (two workgroups per CU size their dedup windows for half the LDS)
    const GatherShape shape = gather_shape(I, lowp, lowp ? I.refine_count(k) : 0);
    const int gthreads = shape.threads;
    GatherArgs ga;
    ga.N = I.n_points;
    ga.D = I.n_features;
    ga.ldx = I.ldx;
    ga.qProbes = I.qProbes;
    ga.k = k;
    if (batched) ga.ldo = k;
    ga.window_bits = gather_window_bits(I.n_points,
                                        lowp ? lowprec_extra_words(I.refine_count(k), gthreads)
                                             : (int64_t)gthreads + 256 + 2048,
                                        shape.per_cu);
    LowPrecArgs lp;
    if (lowp) {
        lp.format = I.precision;
        lp.refine = I.refine_count(k);
        lp.refine_per_k = batched && I.refine == 0;
        lp.ldl = I.ldl;
        lp.xb = I.Xb.get();
        lp.x8 = I.X8.get();
        lp.scale = I.s8.get();
        lp.uscale = I.precision == 2 ? I.s8_shared : 0.0f;
        lp.scale_in_row = I.precision == 2 && fp8_scale_in_row(I.ldl, I.n_features, I.s8_shared) ? 1 : 0;
        lp.threads = gthreads;
    }
    const bool keep = I.keep_candidates && !fixed;
    // exact replay of the tie-flagged queries: inside the gather (the
    // workgroup that claims such a query replays it first, in a work area of
    // its own; those queries head the work queue), or, when every query's
    // C_tot must be known before the gather (kept candidates) or a table's
    // replay does not fit a workgroup's LDS, by separate kernels before it
    ReplayArgs rp;
    rp.hg = hg;
    rp.pg = pg;
    const int64_t wpq = probe_replay_work_bytes(pg);
    const size_t rlds = replay_lds_bytes(rp, gthreads, shape.per_cu);
    const bool in_gather = !keep && rlds > 0 && I.replay_mode != 1;
    const size_t glds =
        std::max(lowp ? gather_lp_lds(ga, lp) : gather_f32_lds(ga, gthreads), in_gather ? rlds : size_t(0));
    unsigned grid = persistent_grid(nq, glds, shape.per_cu);
    // per-workgroup scratch: the unique candidates (at most min(C_tot, N)) and
    // the window scatter or score area (C_tot).  With several dedup windows
    // the scatter is a bin per window, filled by one walk over the buckets; a
    // query whose ids overflow a bin takes two walks instead.  The area is
    // grown towards a bin of C_tot per window (no overflow possible) as far
    // as the budget allows for every workgroup.
    const uint64_t bound = std::max<uint64_t>(1, ctot_bound(I, (uint64_t)qp_max));
    const bool one_window = I.n_points <= ga.window_bits;
    const uint64_t slot_cand = std::min<uint64_t>(bound, (uint64_t)I.n_points);
    uint64_t slot_part = one_window ? slot_cand : bound;
    if (!one_window && !fixed && !keep) {
        const uint64_t nwin = (uint64_t)window_count(I.n_points, ga.window_bits);
        const size_t fixed_part = in_gather ? (size_t)grid * wpq : 0;
        const size_t room = I.ws_budget > fixed_part ? (I.ws_budget - fixed_part) / ((size_t)grid * 4) : 0;
        if (room > slot_cand) slot_part = std::max<uint64_t>(bound, std::min<uint64_t>(nwin * bound, room - slot_cand));
    }
    // separate replay kernels: one work area for each of the first rcap
    // flagged queries, and rover persistent workgroups for any beyond
    unsigned rcap = in_gather ? 0u : (unsigned)std::min(nq, 512);
    unsigned rover = !in_gather && nq > (int)rcap ? (unsigned)std::min(nq - (int)rcap, 256) : 0u;
    const bool stamps = I.phase_stamps && !fixed;
    const size_t qbytes = (size_t)nq * I.ldx * 4;
    auto layout = [&](unsigned slots, unsigned rareas) {
        QueryLayout w = query_layout(nq, L, PL, nprobes, batched, rareas, wpq, stamps, keep ? 0 : slots,
                                     keep ? 0 : slot_cand, keep ? 0 : slot_part);
        // the padded queries go first
        const size_t shift = (qbytes + 255) & ~size_t(255);
        for (size_t* f : {&w.qc, &w.qv, &w.qflag, &w.probes, &w.ctot, &w.ncand, &w.qparams, &w.okey, &w.oval,
                          &w.okey2, &w.oval2, &w.ohist, &w.ohist_off, &w.oscan, &w.qlist, &w.counters, &w.rwork,
                          &w.phase_ts, &w.cand, &w.part})
            *f += shift;
        w.total += shift;
        return w;
    };
    // replay work areas: one per gather workgroup, or those of the kernels
    auto areas = [&](unsigned g) { return in_gather ? g : rcap + rover; };
    QueryLayout lo = layout(grid, areas(grid));
    const size_t per_slot = (size_t)(slot_cand + slot_part) * 4 + (in_gather ? (size_t)wpq : 0);
    if (fixed) {
        // fit the fixed workspace: fewer replay work areas, then fewer
        // gather workgroups (each step halves one of them, so the loop ends)
        while (!in_gather && lo.total > fixed_bytes && (rover > 1 || rcap > 16)) {
            if (rover > 1)
                rover /= 2;
            else
                rcap = std::max(16u, rcap / 2);
            if (nq > (int)rcap) rover = std::max(rover, 1u);
            lo = layout(grid, areas(grid));
        }
        if (lo.total > fixed_bytes) {
            const size_t base = layout(0, areas(0)).total;
            grid = base + per_slot <= fixed_bytes ? (unsigned)std::min<size_t>(grid, (fixed_bytes - base) / per_slot) : 0;
            FPP_REQUIRE(grid >= 1, FPP_EINVAL,
                        "query batch too large for the fixed workspace of " + std::to_string(fixed_bytes >> 20) +
                            " MiB (fewer queries or probes per call)");
            lo = layout(grid, areas(grid));
        }
    } else if (!keep && (size_t)grid * per_slot > I.ws_budget) {
        grid = (unsigned)std::max<size_t>(1, I.ws_budget / per_slot);
        lo = layout(grid, areas(grid));
    }
    unsigned char* W;
    if (fixed) {
        W = fixed_ws;
    } else {
        I.ws.reserve(lo.total);
        W = I.ws.get();
    }
    float* Qpad = reinterpret_cast<float*>(W);
    int32_t* qc = reinterpret_cast<int32_t*>(W + lo.qc);
    float* qv = reinterpret_cast<float*>(W + lo.qv);
    uint32_t* qflag = reinterpret_cast<uint32_t*>(W + lo.qflag);
    uint32_t* probes = reinterpret_cast<uint32_t*>(W + lo.probes);
    uint64_t* ctot = reinterpret_cast<uint64_t*>(W + lo.ctot);
    uint64_t* ncand = reinterpret_cast<uint64_t*>(W + lo.ncand);
    QueryParams* qparams = reinterpret_cast<QueryParams*>(W + lo.qparams);
    uint32_t* counters = reinterpret_cast<uint32_t*>(W + lo.counters);
    uint32_t* qlist = reinterpret_cast<uint32_t*>(W + lo.qlist);
    if (batched) {
        pg.qp = qparams;
        rp.pg.qp = qparams;
        ga.qp = qparams;
        FPP_HIP(hipMemcpyAsync(qparams, I.hqp.data(), (size_t)nq * sizeof(QueryParams), hipMemcpyHostToDevice, st));
    }
    ga.phase_ts = stamps ? reinterpret_cast<uint64_t*>(W + lo.phase_ts) : nullptr;
    ga.heads = counters;
    ga.ncand = ncand;
    ga.ctot = ctot;
    ga.nitems = (uint32_t)nq;
    ga.cand = reinterpret_cast<int32_t*>(W + lo.cand);
    ga.part = reinterpret_cast<int32_t*>(W + lo.part);
    ga.slot_cand = slot_cand;
    ga.slot_part = slot_part;
    rp.Q = Qpad;
    rp.ldq = I.ldx;
    rp.masks = I.masks.get();
    rp.qflag = in_gather ? qflag : nullptr;
    rp.qc = qc;
    rp.qv = qv;
    rp.bucket_off = I.bucket_off.get();
    rp.probes = probes;
    rp.ctot = ctot;
    rp.work = W + lo.rwork;
    rp.wpq = wpq;
    const int order_mode = exact_only ? ORDER_LONGEST : I.order_mode;
    const bool marks = I.profiling && !fixed;
    // small queries (at most SMALL_MAX_CTOT bucket ids, not tie-flagged,
    // k <= 64) go to the end of the sorted order and are served one per wave
    // by launch_gather_small; the workgroup gather takes the rest
    const bool small = I.small_queries && !exact_only && order_mode != ORDER_IDENTITY && nq > 1;
    SmallPick sp;
    sp.ctot = ctot;
    sp.qp = batched ? qparams : nullptr;
    sp.k = k;
    sp.max_ctot = small ? SMALL_MAX_CTOT : 0u;
    sp.count = counters + 10;
    // 1. hashing, probe selection, exact replay of the flagged queries
    auto front = [&](bool ev) {
        // padded queries, the flags (bit 0: per-layer lists rebuilt by the
        // replay; bit 1: probes replayed) and the counters, in one kernel
        launch_pad_rows_init(dQ, ldq, nq, I.n_features, Qpad, I.ldx, qflag,
                             exact_only ? 3u : (probe_path == PROBE_EXACT ? 2u : 0u), nq, counters, N_COUNTERS, st);
        if (ev) I.mark(0, st);
        if (!exact_only)
            launch_hash_query(hg, Qpad, I.ldx, nq, PL, I.masks.get(), qc, qv, qflag, st);
        if (ev) I.mark(1, st);
        if (probe_path != PROBE_EXACT)
            launch_probe_select(pg, nq, qc, qv, I.bucket_off.get(), probes, ctot, qflag, st);
        if (ev) I.mark(2, st);
        if (ev) I.mark(3, st);
        if (!in_gather) {
            launch_compact_flagged(nq, qflag, qlist, counters + 9, st);
            launch_probe_replay(hg, pg, Qpad, I.ldx, I.masks.get(), qlist, counters + 9, counters + 8, qflag, qc, qv,
                                I.bucket_off.get(), probes, ctot, W + lo.rwork, wpq, rcap, rover, st);
        }
        if (ev) I.mark(4, st);
    };
    // 2-3. order, gather
    auto back = [&](bool ev) {
        const uint32_t* order = nullptr;
        int nqueues = 1;
        if (order_mode != ORDER_IDENTITY && nq > 1) {
            uint32_t* okey = reinterpret_cast<uint32_t*>(W + lo.okey);
            // key bits, and the smallest key of a small query (see SmallPick)
            int bits = 32;
            uint32_t small_min = 0xffffffffu - SMALL_MAX_CTOT;
            if (order_mode == ORDER_BUCKET) {
                launch_order_keys(qc, qflag, nq, L, PL, 2 * I.n_proj, okey, sp, st);
                bits = ilog2_exact(I.NB()) + 1;
                small_min = 1u << bits;
                nqueues = nq >= 64 ? 8 : 1;
            } else {
                launch_ctot_keys(ctot, qflag, nq, okey, sp, st);
            }
            if (nq <= SORT1_MAX) {
                // one workgroup, keys in LDS (one launch instead of five per
                // pass); the small queries are only moved to the end
                uint32_t* sv = reinterpret_cast<uint32_t*>(W + lo.oval);
                launch_sort1(okey, nq, bits, small, small_min, sv, st);
                order = sv;
            } else {
                // every key sorted; small queries' bucket keys carry the bit
                // above the bucket's
                if (small && order_mode == ORDER_BUCKET) bits += 1;
                uint32_t *sk = nullptr, *sv = nullptr;
                radix_sort_pairs(okey, reinterpret_cast<uint32_t*>(W + lo.oval),
                                 reinterpret_cast<uint32_t*>(W + lo.okey2), reinterpret_cast<uint32_t*>(W + lo.oval2),
                                 nq, bits, true, reinterpret_cast<uint32_t*>(W + lo.ohist),
                                 reinterpret_cast<uint64_t*>(W + lo.ohist_off), reinterpret_cast<uint64_t*>(W + lo.oscan),
                                 &sk, &sv, st);
                order = sv;
            }
        }
        GatherArgs a = ga;
        a.order = order;
        a.nqueues = nqueues;
        a.small_at = small ? 10 : -1;
        if (ev) I.mark(5, st);
        if (lowp)
            launch_gather_score_lp(a, lp, rp, glds, grid, probes, I.bucket_off.get(), I.ids.get(), I.Xc.get(), Qpad,
                                   I.ldx, d_out, d_scores, shape.per_cu, st);
        else
            launch_gather_score_f32(a, rp, glds, grid, probes, I.bucket_off.get(), I.ids.get(), I.Xc.get(), Qpad,
                                    I.ldx, d_out, d_scores, gthreads, st);
        if (small)
            launch_gather_small(a, probes, I.bucket_off.get(), I.ids.get(), I.Xc.get(), Qpad, I.ldx, d_out, d_scores,
                                st);
        if (ev) I.mark(6, st);
    };
    if (keep) {
        // every query's candidates stay in a segment of its own, sized by
        // its C_tot (read back after the replay)
        front(marks);
        std::vector<uint64_t> ct(nq);
        FPP_HIP(hipMemcpyAsync(ct.data(), ctot, (size_t)nq * 8, hipMemcpyDeviceToHost, st));
        FPP_HIP(hipStreamSynchronize(st));
        std::vector<uint64_t>& seg = I.last.seg;
        seg.assign(nq + 1, 0);
        for (int q = 0; q < nq; ++q)
            seg[q + 1] = seg[q] + std::max<uint64_t>(1, one_window ? std::min<uint64_t>(ct[q], I.n_points) : ct[q]);
        I.keep_cand.reserve(seg[nq]);
        I.keep_part.reserve(seg[nq]);
        I.keep_off.reserve(nq + 1);
        FPP_HIP(hipMemcpyAsync(I.keep_off.get(), seg.data(), (size_t)(nq + 1) * 8, hipMemcpyHostToDevice, st));
        ga.cand = I.keep_cand.get();
        ga.part = I.keep_part.get();
        ga.seg_off = I.keep_off.get();
        back(marks);
    } else if (I.use_graph && !fixed) {
        const std::vector<uint64_t> key = {
            (uint64_t)nq, (uint64_t)ldq, (uint64_t)(uintptr_t)dQ, (uint64_t)(uintptr_t)st, I.build_gen,
            (uint64_t)(uintptr_t)W, lo.total, (uint64_t)(uintptr_t)d_out, (uint64_t)(uintptr_t)d_scores,
            (uint64_t)k, (uint64_t)pg.PL, (uint64_t)pg.perTable, (uint64_t)pg.qProbes, (uint64_t)batched,
            (uint64_t)I.precision, (uint64_t)lp.refine, (uint64_t)lp.refine_per_k,
            (uint64_t)(uintptr_t)(lowp ? (I.precision == 1 ? (const void*)lp.xb : (const void*)lp.x8) : nullptr),
            (uint64_t)__builtin_bit_cast(uint32_t, lp.uscale), (uint64_t)gthreads, (uint64_t)order_mode, grid,
            slot_cand, slot_part, (uint64_t)stamps, (uint64_t)probe_path, (uint64_t)ga.window_bits, rcap, rover,
            (uint64_t)in_gather, glds, (uint64_t)small};
        if (!I.graph_exec || key != I.graph_key) {
            I.drop_graph();
            FPP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                front(false);
                back(false);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            FPP_HIP(hipStreamEndCapture(st, &I.graph));
            FPP_HIP(hipGraphInstantiate(&I.graph_exec, I.graph, nullptr, nullptr, 0));
            I.graph_key = key;
        }
        // the whole call is timed as one stage
        for (int i = 0; i < 6; ++i)
            if (marks) I.mark(i, st);
        FPP_HIP(hipGraphLaunch(I.graph_exec, st));
        if (marks) I.mark(6, st);
    } else {
        front(marks);
        back(marks);
    }
    I.last.probes = probes;
    I.last.ctot = ctot;
    I.last.ncand = ncand;
    I.last.qflag = qflag;
    I.last.nsmall = small ? counters + 10 : nullptr;
    I.last.st = st;
    I.last.summed = false;
    if (marks) {
        float a = 0, b = 0, c = 0, d = 0;
        FPP_HIP(hipEventSynchronize(I.ev[6]));
        FPP_HIP(hipEventElapsedTime(&a, I.ev[0], I.ev[1]));
        FPP_HIP(hipEventElapsedTime(&b, I.ev[1], I.ev[2]));
        FPP_HIP(hipEventElapsedTime(&c, I.ev[3], I.ev[4]));
        FPP_HIP(hipEventElapsedTime(&d, I.ev[5], I.ev[6]));
        I.stage_ms[0] = a;
        I.stage_ms[1] = b;
        I.stage_ms[2] = c;
        I.stage_ms[3] = d;
    }
    if (stamps) {
        if (!I.wall_khz) {
            int dev = 0;
            FPP_HIP(hipGetDevice(&dev));
            FPP_HIP(hipDeviceGetAttribute(&I.wall_khz, hipDeviceAttributeWallClockRate, dev));
        }
        std::vector<uint64_t>& ts = I.last_ts;
        ts.resize((size_t)nq * PHASE_STAMPS);
        FPP_HIP(hipMemcpyAsync(ts.data(), ga.phase_ts, ts.size() * 8, hipMemcpyDeviceToHost, st));
        FPP_HIP(hipStreamSynchronize(st));
        double acc[4] = {0, 0, 0, 0};
        uint64_t longest = 0, longest_flagged = 0;
        std::vector<uint32_t> fl;
        fl.resize(nq);
        FPP_HIP(hipMemcpy(fl.data(), qflag, (size_t)nq * 4, hipMemcpyDeviceToHost));
        for (int q = 0; q < nq; ++q) {
            const uint64_t* t = &ts[(size_t)q * PHASE_STAMPS];
            for (int i = 0; i < 4; ++i) acc[i] += (double)(t[i + 1] - t[i]);
            longest = std::max(longest, t[4] - t[0]);
            if (fl[q]) longest_flagged = std::max(longest_flagged, t[1] - t[0]);
        }
        const double us = 1000.0 / std::max(1, I.wall_khz);
        for (int i = 0; i < 4; ++i) I.stage_ms[4 + i] = acc[i] / nq * us;
        I.stage_ms[8] = (double)longest * us;
        I.stage_ms[9] = (double)longest_flagged * us;
    }
    if (sync_end) FPP_HIP(hipStreamSynchronize(st));
}

// Statistics of the last query call, summed from its device arrays on the
// first request (this waits for the call).
void sum_last_call(Index& I) {
    Index::LastCall& c = I.last;
    if (c.summed) return;
    c.summed = true;
    I.stat_ctot = I.stat_cuniq = I.stat_qties = I.stat_small = 0;
    if (c.nq == 0 || !c.ctot) return;
    std::vector<uint64_t> ct(c.nq), nc(c.nq);
    std::vector<uint32_t> fl(c.nq);
    FPP_HIP(hipStreamSynchronize(c.st));
    FPP_HIP(hipMemcpy(ct.data(), c.ctot, (size_t)c.nq * 8, hipMemcpyDeviceToHost));
    FPP_HIP(hipMemcpy(nc.data(), c.ncand, (size_t)c.nq * 8, hipMemcpyDeviceToHost));
    FPP_HIP(hipMemcpy(fl.data(), c.qflag, (size_t)c.nq * 4, hipMemcpyDeviceToHost));
    for (int q = 0; q < c.nq; ++q) {
        I.stat_ctot += (int64_t)ct[q];
        I.stat_cuniq += (int64_t)nc[q];
        I.stat_qties += fl[q] ? 1 : 0;
    }
    if (c.nsmall) {
        uint32_t ns = 0;
        FPP_HIP(hipMemcpy(&ns, c.nsmall, 4, hipMemcpyDeviceToHost));
        I.stat_small = ns;
    }
}

// ---- index files ------------------------------------------------------------
// A built index as a little-endian file: a fixed header (magic, version, the
// parameters and the section sizes), then the sections in order: sign bits of
// layer 1 and 2 (one byte each), flip masks (u32), centre (f32), CSR offsets
// (u64), CSR ids (i32), centred data (f32, rows of ldx).  Loading restores the
// built state without hashing again.
constexpr char kIndexMagic[8] = {'F', 'P', 'P', 'A', 'M', 'D', 'I', 'X'};
constexpr int64_t kIndexVersion = 1;

struct IndexFileHeader {
    char magic[8];
    int64_t version;
    int64_t n_points, n_features, n_tables, n_proj, fhtDim, n_rotate, bucket_minSize, iProbes, seed;
    int64_t ldx, n_ids, max_bucket, hash_ties, bucket_ties;
    int64_t n_bits, n_mask_words, n_offsets;
    float bucket_scale;
    int32_t reserved;
};

struct File {
    FILE* f = nullptr;
    std::string path;
    File(const char* p, const char* mode) : f(std::fopen(p, mode)), path(p) {
        FPP_REQUIRE(f != nullptr, FPP_EIO, std::string("cannot open index file ") + p);
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void write(const void* src, size_t bytes) {
        FPP_REQUIRE(std::fwrite(src, 1, bytes, f) == bytes, FPP_EIO, "cannot write index file " + path);
    }
    void read(void* dst, size_t bytes) {
        FPP_REQUIRE(std::fread(dst, 1, bytes, f) == bytes, FPP_EIO, "index file " + path + " is truncated");
    }
};

// device <-> file through a host staging buffer of at most 64 MiB
void write_device(File& out, const void* d, size_t bytes) {
    std::vector<unsigned char> buf(std::min<size_t>(bytes, size_t(64) << 20));
    for (size_t o = 0; o < bytes; o += buf.size()) {
        const size_t c = std::min(buf.size(), bytes - o);
        FPP_HIP(hipMemcpy(buf.data(), static_cast<const unsigned char*>(d) + o, c, hipMemcpyDeviceToHost));
        out.write(buf.data(), c);
    }
}
// check(ptr, count) sees each staged piece of T values before it is copied
template <class T, class Check>
void read_device(File& in, T* d, size_t count, Check&& check) {
    const size_t per = std::max<size_t>(1, (size_t(64) << 20) / sizeof(T));
    std::vector<T> buf(std::min(count, per));
    for (size_t o = 0; o < count; o += buf.size()) {
        const size_t c = std::min(buf.size(), count - o);
        in.read(buf.data(), c * sizeof(T));
        check(buf.data(), c);
        FPP_HIP(hipMemcpy(d + o, buf.data(), c * sizeof(T), hipMemcpyHostToDevice));
    }
}
template <class T>
void read_device(File& in, T* d, size_t count) {
    read_device(in, d, count, [](const T*, size_t) {});
}

void save_index(const Index& I, const char* path) {
    FPP_REQUIRE(I.built, FPP_ESTATE, "save called before build (or after clear)");
    IndexFileHeader h{};
    std::memcpy(h.magic, kIndexMagic, 8);
    h.version = kIndexVersion;
    h.n_points = I.n_points;
    h.n_features = I.n_features;
    h.n_tables = I.n_tables;
    h.n_proj = I.n_proj;
    h.fhtDim = I.fhtDim;
    h.n_rotate = I.n_rotate;
    h.bucket_minSize = I.bucket_minSize;
    h.iProbes = I.iProbes;
    h.seed = I.seed;
    h.ldx = I.ldx;
    h.n_ids = I.n_ids;
    h.max_bucket = I.max_bucket;
    h.hash_ties = I.stat_hash_ties;
    h.bucket_ties = I.stat_bucket_ties;
    h.n_bits = (int64_t)I.bit1.size();
    h.n_mask_words = (int64_t)I.masks.n;
    h.n_offsets = (int64_t)I.n_tables * I.NB() + 1;
    h.bucket_scale = I.bucket_scale;
    File out(path, "wb");
    out.write(&h, sizeof h);
    out.write(I.bit1.data(), I.bit1.size());
    out.write(I.bit2.data(), I.bit2.size());
    write_device(out, I.masks.get(), (size_t)h.n_mask_words * 4);
    write_device(out, I.center.get(), (size_t)I.n_features * 4);
    write_device(out, I.bucket_off.get(), (size_t)h.n_offsets * 8);
    write_device(out, I.ids.get(), (size_t)I.n_ids * 4);
    write_device(out, I.Xc.get(), (size_t)I.n_points * I.ldx * 4);
}

void load_index(Index& I, const char* path) {
    File in(path, "rb");
    IndexFileHeader h{};
    in.read(&h, sizeof h);
    FPP_REQUIRE(std::memcmp(h.magic, kIndexMagic, 8) == 0, FPP_EINVAL, std::string(path) + " is not an index file");
    FPP_REQUIRE(h.version == kIndexVersion, FPP_EINVAL, "unsupported index file version");
    FPP_REQUIRE(h.n_points == I.n_points && h.n_features == I.n_features, FPP_EINVAL,
                "index file holds n_points = " + std::to_string(h.n_points) + ", n_features = " +
                    std::to_string(h.n_features) + ", not those of this index");
    Index T;  // parameters checked on a scratch index first
    T.n_points = (int)h.n_points;
    T.n_features = (int)h.n_features;
    T.n_tables = (int)h.n_tables;
    T.n_proj = (int)h.n_proj;
    T.fhtDim = (int)h.fhtDim;
    T.iProbes = (int)h.iProbes;
    T.params = true;
    check_params(T);
    FPP_REQUIRE(h.n_rotate == I.n_rotate, FPP_EINVAL, "index file: unsupported number of rotations");
    FPP_REQUIRE(h.n_bits == h.fhtDim * h.n_tables * h.n_rotate &&
                    h.n_mask_words == h.n_tables * 2 * 3 * ((h.fhtDim + 31) / 32) &&
                    h.n_offsets == h.n_tables * T.NB() + 1 && h.ldx == round_up(h.n_features, 4) && h.n_ids >= 0 &&
                    h.max_bucket >= 0,
                FPP_EINVAL, "index file: inconsistent sizes");
    // section bytes against the file length
    const int64_t need = (int64_t)sizeof h + 2 * h.n_bits + 4 * h.n_mask_words + 4 * h.n_features + 8 * h.n_offsets +
                         4 * h.n_ids + 4 * h.n_points * h.ldx;
    const long here = std::ftell(in.f);
    FPP_REQUIRE(std::fseek(in.f, 0, SEEK_END) == 0, FPP_EIO, "cannot seek in index file");
    const int64_t len = (int64_t)std::ftell(in.f);
    FPP_REQUIRE(std::fseek(in.f, here, SEEK_SET) == 0, FPP_EIO, "cannot seek in index file");
    FPP_REQUIRE(len == need, FPP_EINVAL, "index file " + std::string(path) + " has " + std::to_string(len) +
                                             " bytes, its header describes " + std::to_string(need));
    I.built = false;
    I.lp_ready = 0;
    I.Xb.reset();
    I.X8.reset();
    I.s8.reset();
    I.pre_cnt.reset();
    I.have_pre_cnt = false;
    I.entry_codes.reset();
    I.entry_vals.reset();
    I.entries_kept = false;
    I.n_tables = T.n_tables;
    I.n_proj = T.n_proj;
    I.fhtDim = T.fhtDim;
    I.iProbes = T.iProbes;
    I.bucket_minSize = (int)h.bucket_minSize;
    I.bucket_scale = h.bucket_scale;
    I.seed = (int)h.seed;
    I.params = true;
    I.bit1.resize((size_t)h.n_bits);
    I.bit2.resize((size_t)h.n_bits);
    in.read(I.bit1.data(), I.bit1.size());
    in.read(I.bit2.data(), I.bit2.size());
    I.masks.reserve((size_t)h.n_mask_words);
    read_device(in, I.masks.get(), (size_t)h.n_mask_words);
    I.center.reserve((size_t)h.n_features);
    read_device(in, I.center.get(), (size_t)h.n_features);
    // offsets: from 0, non-decreasing, ending at the id count; the largest
    // bucket is recomputed from them (the kernels size segments by it)
    I.bucket_off.reserve((size_t)h.n_offsets);
    uint64_t prev = 0, largest = 0;
    bool first = true;
    read_device(in, I.bucket_off.get(), (size_t)h.n_offsets, [&](const uint64_t* o, size_t c) {
        for (size_t i = 0; i < c; ++i) {
            FPP_REQUIRE(first ? o[i] == 0 : o[i] >= prev, FPP_EINVAL,
                        "index file: bucket offsets do not start at 0 and increase");
            FPP_REQUIRE(o[i] <= (uint64_t)h.n_ids, FPP_EINVAL, "index file: bucket offset past the id count");
            if (!first) largest = std::max<uint64_t>(largest, o[i] - prev);
            prev = o[i];
            first = false;
        }
    });
    FPP_REQUIRE((int64_t)prev == h.n_ids, FPP_EINVAL, "index file: bucket offsets do not end at the id count");
    FPP_REQUIRE(largest <= (uint64_t)h.n_points, FPP_EINVAL, "index file: a bucket holds more ids than points");
    I.ids.reserve((size_t)(h.n_ids + IDS_PAD));
    read_device(in, I.ids.get(), (size_t)h.n_ids, [&](const int32_t* v, size_t c) {
        for (size_t i = 0; i < c; ++i)
            FPP_REQUIRE(v[i] >= 0 && (int64_t)v[i] < h.n_points, FPP_EINVAL, "index file: point id out of range");
    });
    I.ldx = h.ldx;
    I.Xc.reserve((size_t)h.n_points * h.ldx);
    read_device(in, I.Xc.get(), (size_t)h.n_points * h.ldx);
    I.n_ids = h.n_ids;
    index_sizes(I, nullptr);
    FPP_REQUIRE(I.max_bucket == (uint32_t)largest, FPP_EINTERNAL, "largest bucket differs from the offsets read");
    I.stat_hash_ties = h.hash_ties;
    I.stat_bucket_ties = h.bucket_ties;
    I.build_gen++;
    I.drop_graph();
    I.built = true;
}

// Times the gather launch shapes with run(st) (one query call on stream st)
// and keeps the fastest: workgroup size and the order the persistent gather
// takes the queries in (by best table-0 bucket in per-XCD runs, or longest
// first).  chosen receives {threads, order, longest first, microseconds}.
template <class Run>
void autotune(Index& I, hipStream_t st, Run run, int32_t* chosen) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    FPP_HIP(hipEventCreate(&e0));
    FPP_HIP(hipEventCreate(&e1));
    const bool profiling = I.profiling, stamps = I.phase_stamps;
    I.profiling = I.phase_stamps = false;
    // workgroup size applies to the FP32 gather; the reduced-precision
    // gather runs 256 or 512; 0 is the automatic launch shape (for the
    // reduced-precision gather the only way to two 512-thread workgroups
    // per CU)
    const std::vector<int> threads =
        I.precision == 0 ? std::vector<int>{0, 256, 512, 1024} : std::vector<int>{0, 256, 512};
    float best = 0.0f;
    int bt = I.gather_threads, bo = I.order_mode;
    try {
        for (int t : threads)
            for (int o : {ORDER_BUCKET, ORDER_LONGEST}) {
                I.gather_threads = t;
                I.order_mode = o;
                run(st);
                FPP_HIP(hipEventRecord(e0, st));
                for (int r = 0; r < 2; ++r) run(st);
                FPP_HIP(hipEventRecord(e1, st));
                FPP_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                FPP_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (I.tune_verbose) {
                    fprintf(stderr, "[autotune] threads=%d order=%s: %.3f ms per call\n", t,
                            o == ORDER_BUCKET ? "bucket" : "longest-first", ms / 2.0f);
                    fflush(stderr);
                }
                if (best == 0.0f || ms < best) {
                    best = ms;
                    bt = t;
                    bo = o;
                }
            }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        I.profiling = profiling;
        I.phase_stamps = stamps;
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    I.profiling = profiling;
    I.phase_stamps = stamps;
    I.gather_threads = bt;
    I.order_mode = bo;
    if (chosen) {
        chosen[0] = bt ? bt : gather_shape(I, I.precision != 0, I.refine_count(1)).threads;
        chosen[1] = bo;
        chosen[2] = bo == ORDER_LONGEST ? 1 : 0;
        chosen[3] = (int32_t)(best / 2.0f * 1000.0f);  // microseconds per batch
    }
}

}  // namespace
}  // namespace fpp

using namespace fpp;

// Workspace of fpp_query_device_noalloc: a device array of fixed size that
// each GPU the library runs on holds until the process ends.  On the host the
// array's name is not a device address; the address of each device's copy is
// taken once with hipGetSymbolAddress (with that device current) and kept.
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

unsigned char* noalloc_scratch() {
    int dev = 0;
    FPP_HIP(hipGetDevice(&dev));
    FPP_REQUIRE(dev >= 0 && dev < kMaxDevices, FPP_EINTERNAL, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        FPP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        FPP_REQUIRE(p != nullptr, FPP_EHIP, "no device address for the fixed workspace");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

struct fpp_index {
    Index I;
};

extern "C" {

const char* fpp_last_error(void) { return g_last_error.c_str(); }

int fpp_create(int n_points, int n_features, fpp_index** out) {
    return guard([&] {
        FPP_REQUIRE(out != nullptr, FPP_EINVAL, "null output handle");
        FPP_REQUIRE(n_points >= 1 && n_features >= 1, FPP_EINVAL, "n_points and n_features must be >= 1");
        int ndev = 0;
        FPP_HIP(hipGetDeviceCount(&ndev));
        FPP_REQUIRE(ndev > 0, FPP_EHIP, "no HIP device visible");
        auto* h = new fpp_index();
        h->I.n_points = n_points;
        h->I.n_features = n_features;
        *out = h;
    });
}

void fpp_destroy(fpp_index* h) { delete h; }

int fpp_set_index_param(fpp_index* h, int n_tables, int n_proj, int bucket_minSize, float bucket_scale, int iProbes,
                        int n_threads, int seed) {
    return guard([&] {
        FPP_REQUIRE(h, FPP_EINVAL, "null index");
        Index& I = h->I;
        I.n_tables = n_tables;
        I.n_proj = n_proj;
        I.bucket_minSize = bucket_minSize;
        I.bucket_scale = bucket_scale;
        I.iProbes = iProbes;
        I.n_threads = n_threads;
        I.seed = seed;
        FPP_REQUIRE(n_proj >= 1, FPP_EINVAL, "n_proj must be >= 1");
        // src/FalconnPP.h:94-99
        if (I.n_proj < I.n_features)
            I.fhtDim = 1 << (int)std::ceil(std::log2((double)I.n_features));
        else
            I.fhtDim = 1 << (int)std::ceil(std::log2((double)I.n_proj));
        I.params = true;
    });
}

int fpp_set_qprobes(fpp_index* h, int qProbes) {
    return guard([&] {
        FPP_REQUIRE(h, FPP_EINVAL, "null index");
        h->I.qProbes = qProbes;
    });
}

int fpp_set_threads(fpp_index* h, int n_threads) {
    return guard([&] {
        FPP_REQUIRE(h, FPP_EINVAL, "null index");
        h->I.n_threads = n_threads;
    });
}

int fpp_clear(fpp_index* h) {
    return guard([&] {
        FPP_REQUIRE(h, FPP_EINVAL, "null index");
        Index& I = h->I;
        I.Xc.reset();
        I.bucket_off.reset();
        I.ids.reset();
        I.masks.reset();
        I.center.reset();
        I.bit1.clear();
        I.bit2.clear();
        I.Xb.reset();
        I.X8.reset();
        I.s8.reset();
        I.pre_cnt.reset();
        I.have_pre_cnt = false;
        I.entry_codes.reset();
        I.entry_vals.reset();
        I.entries_kept = false;
        I.lp_ready = 0;
        I.n_ids = 0;
        I.built = false;
    });
}

int fpp_build_device(fpp_index* h, const float* dX, int64_t ld, void* stream) {
    return guard([&] {
        FPP_REQUIRE(h && dX, FPP_EINVAL, "null argument");
        do_build(h->I, dX, ld, pick_stream(h->I, stream));
    });
}

int fpp_set_insertable(fpp_index* h, int on) {
    return guard([&] {
        FPP_REQUIRE(h, FPP_EINVAL, "null index");
        h->I.insertable = on != 0;
    });
}

int fpp_insert_device(fpp_index* h, const float* dX, int64_t ld, int m, void* stream) {
    return guard([&] {
        FPP_REQUIRE(h && dX, FPP_EINVAL, "null argument");
        do_insert(h->I, dX, ld, m, pick_stream(h->I, stream));
    });
}

int fpp_insert(fpp_index* h, const float* X, int64_t ld, int m) {
    return guard([&] {
        FPP_REQUIRE(h && X, FPP_EINVAL, "null argument");
        Index& I = h->I;
        FPP_REQUIRE(m >= 1 && ld >= I.n_features, FPP_EINVAL, "bad shape of the inserted points");
        hipStream_t st = pick_stream(I, nullptr);
        DevBuf<float> dX;
        dX.reserve((size_t)m * I.n_features);
        copy_rows_h2d(dX.get(), X, ld, m, I.n_features, st);
        do_insert(I, dX.get(), I.n_features, m, st);
    });
}

int fpp_build(fpp_index* h, const float* X, int64_t ld) {
    return guard([&] {
        FPP_REQUIRE(h && X, FPP_EINVAL, "null argument");
        Index& I = h->I;
        FPP_REQUIRE(ld >= I.n_features, FPP_EINVAL, "leading dimension of X must be >= n_features");
        hipStream_t st = pick_stream(I, nullptr);
        DevBuf<float> dX;
        dX.reserve((size_t)I.n_points * I.n_features);
        copy_rows_h2d(dX.get(), X, ld, I.n_points, I.n_features, st);
        do_build(I, dX.get(), I.n_features, st);
    });
}

int fpp_query_device(fpp_index* h, const float* dQ, int64_t ldq, int nq, int k, int32_t* d_out, float* d_scores,
                     void* stream) {
    return guard([&] {
        FPP_REQUIRE(h && (nq == 0 || (dQ && d_out)), FPP_EINVAL, "null argument");
        // on the library's own stream the call returns with the results
        // written; on the caller's stream it is asynchronous
        do_query(h->I, dQ, ldq, nq, k, d_out, d_scores, pick_stream(h->I, stream), nullptr, nullptr, nullptr, 0,
                 stream == nullptr);
    });
}

int fpp_query_device_noalloc(fpp_index* h, const float* dQ, int64_t ldq, int nq, int k, int32_t* d_out,
                             float* d_scores, void* stream) {
    return guard([&] {
        FPP_REQUIRE(h && (nq == 0 || (dQ && d_out)), FPP_EINVAL, "null argument");
        Index& I = h->I;
        // no stream of the caller's and none made yet: the null stream
        // (creating one is left to the allocating entry points)
        hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : I.stream;
        unsigned char* ws = noalloc_scratch();
        do_query(I, dQ, ldq, nq, k, d_out, d_scores, st, nullptr, nullptr, ws, kNoAllocScratchBytes,
                 stream == nullptr);
    });
}

int fpp_query_batch_device(fpp_index* h, const float* dQ, int64_t ldq, int nq, const int32_t* qprobes,
                           const int32_t* ks, int kmax, int32_t* d_out, float* d_scores, void* stream) {
    return guard([&] {
        FPP_REQUIRE(h && (nq == 0 || (dQ && d_out && qprobes && ks)), FPP_EINVAL, "null argument");
        do_query(h->I, dQ, ldq, nq, kmax, d_out, d_scores, pick_stream(h->I, stream), qprobes, ks, nullptr, 0,
                 stream == nullptr);
    });
}

int fpp_autotune(fpp_index* h, const float* dQ, int64_t ldq, int nq, int k, void* stream, int32_t* chosen) {
    return guard([&] {
        FPP_REQUIRE(h && dQ && nq >= 1, FPP_EINVAL, "null argument or empty sample");
        Index& I = h->I;
        DevBuf<int32_t> out;
        out.reserve((size_t)nq * k);
        autotune(I, pick_stream(I, stream), [&](hipStream_t st) { do_query(I, dQ, ldq, nq, k, out.get(), nullptr, st); },
                 chosen);
    });
}

int fpp_autotune_batch(fpp_index* h, const float* dQ, int64_t ldq, int nq, const int32_t* qprobes,
                       const int32_t* ks, int kmax, void* stream, int32_t* chosen) {
    return guard([&] {
        FPP_REQUIRE(h && dQ && qprobes && ks && nq >= 1, FPP_EINVAL, "null argument or empty sample");
        Index& I = h->I;
        FPP_REQUIRE(kmax >= 1, FPP_EINVAL, "kmax must be >= 1");
        DevBuf<int32_t> out;
        out.reserve((size_t)nq * kmax);
        autotune(
            I, pick_stream(I, stream),
            [&](hipStream_t st) { do_query(I, dQ, ldq, nq, kmax, out.get(), nullptr, st, qprobes, ks); }, chosen);
    });
}

}  // extern "C"

namespace {
// query2Layers_1D / query2Layers on host buffers; a verbose call prints the
// reference's statistics and writes <layout>_Top_...txt.
int query_host(fpp_index* h, const float* Q, int64_t ldq, int nq, int k, int32_t* out, int verbose,
               const char* layout) {
    return guard([&] {
        FPP_REQUIRE(h && (nq == 0 || (Q && out)), FPP_EINVAL, "null argument");
        Index& I = h->I;
        FPP_REQUIRE(ldq >= I.n_features, FPP_EINVAL, "leading dimension of Q must be >= n_features");
        FPP_REQUIRE(nq >= 0, FPP_EINVAL, "number of queries must be >= 0");
        hipStream_t st = pick_stream(I, nullptr);
        auto t0 = std::chrono::high_resolution_clock::now();
        // a verbose call times its stages and phases like the reference's
        const bool prof = I.profiling, stamps = I.phase_stamps;
        if (verbose) I.profiling = I.phase_stamps = true;
        struct Restore {
            Index& I;
            bool p, s;
            ~Restore() {
                I.profiling = p;
                I.phase_stamps = s;
            }
        } restore{I, prof, stamps};
        DevBuf<float> dQ;
        dQ.reserve((size_t)std::max(nq, 1) * I.n_features);
        if (nq > 0)
            copy_rows_h2d(dQ.get(), Q, ldq, nq, I.n_features, st);
        I.out_ids.reserve((size_t)std::max(nq, 1) * k);
        do_query(I, dQ.get(), I.n_features, nq, k, I.out_ids.get(), nullptr, st);
        if (nq > 0)
            FPP_HIP(hipMemcpyAsync(out, I.out_ids.get(), (size_t)nq * k * 4, hipMemcpyDeviceToHost, st));
        FPP_HIP(hipStreamSynchronize(st));
        if (verbose) {
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
            // src/FalconnPP.cpp:682-704.  The gather kernel fuses the bucket
            // lookup with the distances: its time is split by the mean
            // per-query phase times (walk + dedup vs scoring + top-k).
            I.scratch_word.reserve(2);
            sum_last_call(I);
            const uint64_t empty = count_empty_probes(I.last.probes, (int64_t)I.last.nprobes, I.bucket_off.get(),
                                                      I.scratch_word.get(), st);
            const double ph_look = I.stage_ms[4] + I.stage_ms[5], ph_dist = I.stage_ms[6] + I.stage_ms[7];
            const double share = ph_look + ph_dist > 0 ? ph_look / (ph_look + ph_dist) : 0.0;
            const double dq = nq ? (double)nq : 1.0;
            printf("Average number of empty buckets per query: %g\n", (double)empty / dq);
            printf("Average number of probes per query: %g\n", (double)I.qProbes);
            printf("Average number of unique candidates per query: %g\n", (double)I.stat_cuniq / dq);
            printf("Average number of candidates per query: %g\n", (double)I.stat_ctot / dq);
            printf("Hash and Probing Time: %g ms\n", I.stage_ms[0] + I.stage_ms[1] + I.stage_ms[2]);
            printf("Lookup Time: %g ms\n", I.stage_ms[3] * share);
            printf("Distance Time: %g ms\n", I.stage_ms[3] * (1.0 - share));
            printf("Querying Time: %g ms\n", ms);
            fflush(stdout);
            const std::string name = topk_file_name(k, I.n_proj, I.n_tables, I.iProbes, I.bucket_minSize,
                                                    I.bucket_scale, I.qProbes, layout);
            write_topk_txt(name.c_str(), out, nq, k);
        }
    });
}
}  // namespace

extern "C" {

int fpp_query(fpp_index* h, const float* Q, int64_t ldq, int nq, int k, int32_t* out, int verbose) {
    return query_host(h, Q, ldq, nq, k, out, verbose, "1D");
}

int fpp_query_2d(fpp_index* h, const float* Q, int64_t ldq, int nq, int k, int32_t* out, int verbose) {
    return query_host(h, Q, ldq, nq, k, out, verbose, "2D");
}

int fpp_save(fpp_index* h, const char* path) {
    return guard([&] {
        FPP_REQUIRE(h && path, FPP_EINVAL, "null argument");
        save_index(h->I, path);
    });
}

int fpp_load(fpp_index* h, const char* path) {
    return guard([&] {
        FPP_REQUIRE(h && path, FPP_EINVAL, "null argument");
        load_index(h->I, path);
    });
}

int fpp_load_txt(const char* path, int n_points, int n_features, float* out, int64_t* n_read) {
    return guard([&] {
        FPP_REQUIRE(path && out, FPP_EINVAL, "null argument");
        FPP_REQUIRE(n_points >= 0 && n_features >= 1, FPP_EINVAL, "n_points must be >= 0 and n_features >= 1");
        const int64_t n = load_txt(path, (int64_t)n_points * n_features, out);
        if (n_read) *n_read = n;
    });
}

int fpp_write_topk_txt(const char* path, const int32_t* ids, int nq, int k) {
    return guard([&] {
        FPP_REQUIRE(path && (nq == 0 || ids), FPP_EINVAL, "null argument");
        FPP_REQUIRE(nq >= 0 && k >= 1, FPP_EINVAL, "nq must be >= 0 and k >= 1");
        write_topk_txt(path, ids, nq, k);
    });
}

}  // extern "C"

namespace {
int file_name(int k, int n_proj, int n_tables, int iProbes, int bucket_minSize, float bucket_scale, int qProbes,
              char* buf, int buflen, const char* layout) {
    return guard([&] {
        FPP_REQUIRE(buf && buflen >= 1, FPP_EINVAL, "null argument");
        const std::string s =
            topk_file_name(k, n_proj, n_tables, iProbes, bucket_minSize, bucket_scale, qProbes, layout);
        FPP_REQUIRE((int64_t)s.size() < (int64_t)buflen, FPP_EINVAL, "buffer too short for the file name");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}
}  // namespace

extern "C" {

int fpp_topk_file_name(int k, int n_proj, int n_tables, int iProbes, int bucket_minSize, float bucket_scale,
                       int qProbes, char* buf, int buflen) {
    return file_name(k, n_proj, n_tables, iProbes, bucket_minSize, bucket_scale, qProbes, buf, buflen, "1D");
}

int fpp_topk_file_name_2d(int k, int n_proj, int n_tables, int iProbes, int bucket_minSize, float bucket_scale,
                          int qProbes, char* buf, int buflen) {
    return file_name(k, n_proj, n_tables, iProbes, bucket_minSize, bucket_scale, qProbes, buf, buflen, "2D");
}

int fpp_get_shape(fpp_index* h, int64_t* out) {
    return guard([&] {
        FPP_REQUIRE(h && out, FPP_EINVAL, "null argument");
        Index& I = h->I;
        out[0] = I.fhtDim;
        out[1] = I.params ? I.NB() : 0;
        out[2] = I.n_tables;
        out[3] = I.n_ids;
        out[4] = I.ldx;
        out[5] = I.n_points;
    });
}

int fpp_get_center(fpp_index* h, float* out) {
    return guard([&] {
        FPP_REQUIRE(h && out, FPP_EINVAL, "null argument");
        FPP_REQUIRE(h->I.built, FPP_ESTATE, "index not built");
        FPP_HIP(hipMemcpy(out, h->I.center.get(), (size_t)h->I.n_features * 4, hipMemcpyDeviceToHost));
    });
}

int fpp_get_sign_bits(fpp_index* h, uint8_t* b1, uint8_t* b2) {
    return guard([&] {
        FPP_REQUIRE(h && b1 && b2, FPP_EINVAL, "null argument");
        FPP_REQUIRE(!h->I.bit1.empty(), FPP_ESTATE, "index not built");
        std::copy(h->I.bit1.begin(), h->I.bit1.end(), b1);
        std::copy(h->I.bit2.begin(), h->I.bit2.end(), b2);
    });
}

int fpp_get_buckets(fpp_index* h, uint64_t* off, int32_t* ids) {
    return guard([&] {
        FPP_REQUIRE(h && off && ids, FPP_EINVAL, "null argument");
        Index& I = h->I;
        FPP_REQUIRE(I.built, FPP_ESTATE, "index not built");
        const size_t nbk = (size_t)I.n_tables * I.NB();
        FPP_HIP(hipMemcpy(off, I.bucket_off.get(), (nbk + 1) * 8, hipMemcpyDeviceToHost));
        if (I.n_ids) FPP_HIP(hipMemcpy(ids, I.ids.get(), (size_t)I.n_ids * 4, hipMemcpyDeviceToHost));
    });
}

int fpp_get_bucket_counts(fpp_index* h, uint32_t* cnt) {
    return guard([&] {
        FPP_REQUIRE(h && cnt, FPP_EINVAL, "null argument");
        Index& I = h->I;
        FPP_REQUIRE(I.built && I.have_pre_cnt, FPP_ESTATE,
                    "no pre-filter bucket sizes: the index was not built in this process");
        FPP_HIP(hipMemcpy(cnt, I.pre_cnt.get(), (size_t)I.n_tables * I.NB() * 4, hipMemcpyDeviceToHost));
    });
}

int fpp_hash_points(fpp_index* h, const float* V, int64_t ld, int m, uint32_t* codes, float* vals) {
    return guard([&] {
        FPP_REQUIRE(h && V && codes && vals, FPP_EINVAL, "null argument");
        Index& I = h->I;
        FPP_REQUIRE(!I.bit1.empty(), FPP_ESTATE, "sign bits not drawn: build first");
        FPP_REQUIRE(ld >= I.n_features && m >= 1, FPP_EINVAL, "bad shape");
        hipStream_t st = pick_stream(I, nullptr);
        const int L = I.n_tables, p = I.iProbes;
        const int64_t ldv = round_up(I.n_features, 4);
        DevBuf<float> dV;
        dV.reserve((size_t)m * ldv);
        FPP_HIP(hipMemsetAsync(dV.get(), 0, (size_t)m * ldv * 4, st));
        FPP_HIP(hipMemcpy2DAsync(dV.get(), ldv * 4, V, ld * 4, (size_t)I.n_features * 4, m, hipMemcpyHostToDevice,
                                 st));
        const int64_t total = (int64_t)m * p * L;
        DevBuf<uint32_t> dc, fcount;
        DevBuf<float> dv;
        DevBuf<uint64_t> flags;
        dc.reserve(total);
        dv.reserve(total);
        const uint32_t cap = (uint32_t)((int64_t)m * L);
        flags.reserve(cap);
        fcount.reserve(1);
        FPP_HIP(hipMemsetAsync(fcount.get(), 0, 4, st));
        const HashGeom g = I.hgeom();
        launch_hash_build(g, dV.get(), ldv, m, p, I.masks.get(), dc.get(), dv.get(), flags.get(), fcount.get(), cap,
                          st);
        uint32_t nflag = 0;
        FPP_HIP(hipMemcpyAsync(&nflag, fcount.get(), 4, hipMemcpyDeviceToHost, st));
        FPP_HIP(hipStreamSynchronize(st));
        if (nflag) {
            DevBuf<float> work;
            work.reserve((size_t)nflag * (2 * I.fhtDim + 8 * p + 8));
            launch_hash_build_exact(g, dV.get(), ldv, m, p, I.masks.get(), dc.get(), dv.get(), flags.get(), nflag, 0,
                                    work.get(), st);
        }
        FPP_HIP(hipMemcpyAsync(codes, dc.get(), total * 4, hipMemcpyDeviceToHost, st));
        FPP_HIP(hipMemcpyAsync(vals, dv.get(), total * 4, hipMemcpyDeviceToHost, st));
        FPP_HIP(hipStreamSynchronize(st));
        I.stat_hash_ties = nflag;
    });
}

int fpp_get_stats(fpp_index* h, int64_t* out) {
    return guard([&] {
        FPP_REQUIRE(h && out, FPP_EINVAL, "null argument");
        Index& I = h->I;
        sum_last_call(I);
        out[0] = I.stat_hash_ties;
        out[1] = I.stat_bucket_ties;
        out[2] = I.stat_ctot;
        out[3] = I.stat_cuniq;
        out[4] = I.stat_qties;
        out[5] = I.last.nq;
        out[6] = I.stat_probe_path;
        out[7] = I.stat_small;
    });
}

int fpp_set_profiling(fpp_index* h, int on) {
    return guard([&] {
        FPP_REQUIRE(h, FPP_EINVAL, "null argument");
        h->I.profiling = on != 0;
        h->I.phase_stamps = on >= 2;
    });
}

int fpp_get_timings(fpp_index* h, double* ms) {
    return guard([&] {
        FPP_REQUIRE(h && ms, FPP_EINVAL, "null argument");
        for (int i = 0; i < 10; ++i) ms[i] = h->I.stage_ms[i];
    });
}

int fpp_set_precision(fpp_index* h, int mode, int refine) {
    return guard([&] {
        FPP_REQUIRE(h, FPP_EINVAL, "null argument");
        FPP_REQUIRE(mode >= 0 && mode <= 2, FPP_EINVAL, "precision mode must be 0 (FP32), 1 (BF16) or 2 (FP8)");
        FPP_REQUIRE(refine >= 0 && refine <= 256, FPP_EINVAL, "refine count must be in [0, 256]");
        h->I.precision = mode;
        h->I.refine = refine;
    });
}

int fpp_get_fp8_scale(fpp_index* h, float* shared) {
    return guard([&] {
        FPP_REQUIRE(h && shared, FPP_EINVAL, "null argument");
        FPP_REQUIRE(h->I.built && h->I.lp_ready == 2, FPP_ESTATE, "no FP8 copy: make an FP8 query first");
        *shared = h->I.s8_shared;
    });
}

int fpp_set_tuning(fpp_index* h, int which, int value) {
    return guard([&] {
        FPP_REQUIRE(h, FPP_EINVAL, "null argument");
        switch (which) {
            case FPP_TUNE_QUERY_ORDER:
                FPP_REQUIRE(value >= 0 && value <= 2, FPP_EINVAL, "query order must be 0, 1 or 2");
                h->I.order_mode = value;
                break;
            case FPP_TUNE_GATHER_THREADS:
                FPP_REQUIRE(value == 0 || value == 256 || value == 512 || value == 1024, FPP_EINVAL,
                            "gather threads must be 0 (automatic), 256, 512 or 1024");
                h->I.gather_threads = value;
                break;
            case FPP_TUNE_PERSISTENT:
                // the gather is always persistent; 1 asks for longest first
                h->I.order_mode = value ? ORDER_LONGEST : ORDER_BUCKET;
                break;
            case FPP_TUNE_KEEP_CANDIDATES:
                h->I.keep_candidates = value != 0;
                break;
            case FPP_TUNE_WORKSPACE_MB:
                FPP_REQUIRE(value >= 1, FPP_EINVAL, "workspace bound must be at least 1 MiB");
                h->I.ws_budget = size_t(value) << 20;
                break;
            case FPP_TUNE_REPLAY:
                FPP_REQUIRE(value >= 0 && value <= 2, FPP_EINVAL, "replay placement must be 0, 1 or 2");
                h->I.replay_mode = value;
                break;
            case FPP_TUNE_FP8_PER_ROW:
                h->I.fp8_per_row = value != 0;
                break;
            case FPP_TUNE_VERBOSE:
                h->I.tune_verbose = value != 0;
                break;
            case FPP_TUNE_GRAPH:
                h->I.use_graph = value != 0;
                if (!h->I.use_graph) h->I.drop_graph();
                break;
            case FPP_TUNE_SMALL_QUERIES:
                h->I.small_queries = value != 0;
                break;
            default:
                throw Error(FPP_EINVAL, "unknown tuning knob");
        }
    });
}

int fpp_get_query_times(fpp_index* h, double* service_us, double* done_us) {
    return guard([&] {
        FPP_REQUIRE(h && service_us && done_us, FPP_EINVAL, "null argument");
        Index& I = h->I;
        const size_t nq = I.last_ts.size() / PHASE_STAMPS;
        FPP_REQUIRE(nq > 0 && nq == (size_t)I.last.nq, FPP_ESTATE,
                    "no per-query stamps: make the last query call with profiling level 2");
        uint64_t t0 = ~0ull;
        for (size_t q = 0; q < nq; ++q) t0 = std::min(t0, I.last_ts[q * PHASE_STAMPS]);
        const double us = 1000.0 / std::max(1, I.wall_khz);
        for (size_t q = 0; q < nq; ++q) {
            const uint64_t* t = &I.last_ts[q * PHASE_STAMPS];
            service_us[q] = (double)(t[4] - t[0]) * us;
            done_us[q] = (double)(t[4] - t0) * us;
        }
    });
}

int fpp_get_last_probes(fpp_index* h, uint32_t* out) {
    return guard([&] {
        FPP_REQUIRE(h && out, FPP_EINVAL, "null argument");
        Index& I = h->I;
        if (I.last.nq && I.last.probes) {
            FPP_HIP(hipStreamSynchronize(I.last.st));
            FPP_HIP(hipMemcpy(out, I.last.probes, (size_t)I.last.nprobes * 4, hipMemcpyDeviceToHost));
        }
    });
}

int64_t fpp_get_last_candidates(fpp_index* h, int j, int32_t* out, int64_t cap) {
    int64_t n = -1;
    const int rc = guard([&] {
        FPP_REQUIRE(h, FPP_EINVAL, "null argument");
        Index& I = h->I;
        FPP_REQUIRE(j >= 0 && j < I.last.nq, FPP_EINVAL, "query index out of range");
        FPP_REQUIRE(!I.last.seg.empty(), FPP_ESTATE,
                    "candidates are kept only by calls made with FPP_TUNE_KEEP_CANDIDATES on");
        FPP_HIP(hipStreamSynchronize(I.last.st));
        uint64_t nc = 0;
        FPP_HIP(hipMemcpy(&nc, I.last.ncand + j, 8, hipMemcpyDeviceToHost));
        n = (int64_t)nc;
        const int64_t c = std::min<int64_t>(cap, n);
        if (c > 0 && out) {
            // above one dedup window the list is unordered inside each
            // window; report it ascending
            std::vector<int32_t> all((size_t)n);
            FPP_HIP(hipMemcpy(all.data(), I.keep_cand.get() + I.last.seg[j], (size_t)n * 4, hipMemcpyDeviceToHost));
            std::sort(all.begin(), all.end());
            std::copy(all.begin(), all.begin() + c, out);
        }
    });
    return rc == FPP_OK ? n : -(int64_t)rc;
}

int fpp_bruteforce_f64(const float* dX, int64_t ldx, int64_t N, int D, const float* dQ, int64_t ldq, int nq, int k,
                       const int32_t* seed_ids, int32_t* out_ids, double* out_scores) {
    return guard([&] {
        FPP_REQUIRE(dX && dQ && out_ids && out_scores, FPP_EINVAL, "null argument");
        FPP_REQUIRE(k >= 1 && N >= k && nq >= 0 && ldx >= D && ldq >= D, FPP_EINVAL, "bad shape");
        if (nq == 0) return;
        hipStream_t st;
        FPP_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct SG {
            hipStream_t s;
            ~SG() { (void)hipStreamDestroy(s); }
        } sg{st};
        // seed rows: given ids (deduplicated per query) or 0..k-1
        std::vector<int32_t> seeds((size_t)nq * k);
        for (int q = 0; q < nq; ++q) {
            std::vector<int32_t> s;
            if (seed_ids)
                for (int j = 0; j < k; ++j) {
                    const int32_t id = seed_ids[(size_t)q * k + j];
                    if (id >= 0 && id < N && std::find(s.begin(), s.end(), id) == s.end()) s.push_back(id);
                }
            for (int32_t id = 0; (int)s.size() < k; ++id)
                if (std::find(s.begin(), s.end(), id) == s.end()) s.push_back(id);
            std::copy(s.begin(), s.end(), seeds.begin() + (size_t)q * k);
        }
        DevBuf<int32_t> dseeds;
        DevBuf<double> dsc, dth;
        dseeds.reserve(seeds.size());
        dsc.reserve(seeds.size());
        dth.reserve(nq);
        FPP_HIP(hipMemcpyAsync(dseeds.get(), seeds.data(), seeds.size() * 4, hipMemcpyHostToDevice, st));
        launch_dot64_rows(dX, ldx, D, dQ, ldq, nq, k, dseeds.get(), dsc.get(), st);
        std::vector<double> sc(seeds.size()), th(nq);
        FPP_HIP(hipMemcpyAsync(sc.data(), dsc.get(), sc.size() * 8, hipMemcpyDeviceToHost, st));
        FPP_HIP(hipStreamSynchronize(st));
        for (int q = 0; q < nq; ++q) th[q] = *std::min_element(sc.begin() + (size_t)q * k, sc.begin() + (size_t)(q + 1) * k);
        FPP_HIP(hipMemcpyAsync(dth.get(), th.data(), nq * 8, hipMemcpyHostToDevice, st));
        uint32_t cap = (uint32_t)std::min<int64_t>((int64_t)nq * k * 16 + (1 << 16), (int64_t)1 << 28);
        for (int attempt = 0; attempt < 3; ++attempt) {
            DevBuf<uint32_t> hq, hid, hc;
            DevBuf<double> hs;
            hq.reserve(cap);
            hid.reserve(cap);
            hs.reserve(cap);
            hc.reserve(1);
            FPP_HIP(hipMemsetAsync(hc.get(), 0, 4, st));
            launch_bf64(dX, ldx, N, D, dQ, ldq, nq, k, dth.get(), hq.get(), hid.get(), hs.get(), hc.get(), cap, st);
            uint32_t nh = 0;
            FPP_HIP(hipMemcpyAsync(&nh, hc.get(), 4, hipMemcpyDeviceToHost, st));
            FPP_HIP(hipStreamSynchronize(st));
            if (nh > cap) {
                cap = (uint32_t)std::min<int64_t>((int64_t)nh + 1024, (int64_t)UINT32_MAX / 2);
                continue;
            }
            std::vector<uint32_t> vq(nh), vid(nh);
            std::vector<double> vs(nh);
            FPP_HIP(hipMemcpy(vq.data(), hq.get(), (size_t)nh * 4, hipMemcpyDeviceToHost));
            FPP_HIP(hipMemcpy(vid.data(), hid.get(), (size_t)nh * 4, hipMemcpyDeviceToHost));
            FPP_HIP(hipMemcpy(vs.data(), hs.get(), (size_t)nh * 8, hipMemcpyDeviceToHost));
            std::vector<std::vector<std::pair<double, int32_t>>> per(nq);
            for (uint32_t i = 0; i < nh; ++i) per[vq[i]].emplace_back(vs[i], (int32_t)vid[i]);
            for (int q = 0; q < nq; ++q) {
                auto& v = per[q];
                std::sort(v.begin(), v.end(), [](const std::pair<double, int32_t>& a, const std::pair<double, int32_t>& b) {
                    return a.first > b.first || (a.first == b.first && a.second < b.second);
                });
                FPP_REQUIRE((int)v.size() >= k, FPP_EINTERNAL, "brute force lost rows above the seed threshold");
                for (int j = 0; j < k; ++j) {
                    out_ids[(size_t)q * k + j] = v[j].second;
                    out_scores[(size_t)q * k + j] = v[j].first;
                }
            }
            return;
        }
        throw Error(FPP_ENOMEM, "brute-force hit list overflow");
    });
}

}  // extern "C"
