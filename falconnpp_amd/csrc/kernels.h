// Launchers of the gfx950 kernels, called by the host engine (engine.cpp).
// All pointers are device pointers; every launcher is asynchronous on `st`.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

namespace fpp {

// ---- hashing (hash.hip) ---------------------------------------------------
// Sign flips as packed bits: masks[((l*2 + layer)*3 + r)*MW + d/32] with
// MW = ceil(F/32); bit d%32 is set when the reference's sign bit
// (l*3*F + r*F + d) is 0, i.e. the factor is -1.
struct HashGeom {
    int F;          // fhtDim (power of two, 2..1024)
    int D;          // n_features
    int n_proj;     // power of two, <= F
    int log2_nproj;
    int L;          // n_tables
};

// Build side: codes/vals[l*N*p + n*p + i] (i in pop order: ascending value);
// (n, l) pairs whose result depends on an exact float tie are appended to
// flag_list (as n*L + l) for the exact replay.
void launch_hash_build(const HashGeom& g, const float* Xc, int64_t ldx, int64_t N, int p, const uint32_t* masks,
                       uint32_t* codes, float* vals, uint64_t* flag_list, uint32_t* flag_count, uint32_t flag_cap,
                       hipStream_t st);
// Exact replay of the build hash for `count` flagged items (or all N*L
// items when all_items != 0), emulating the reference's heaps.
void launch_hash_build_exact(const HashGeom& g, const float* Xc, int64_t ldx, int64_t N, int p,
                             const uint32_t* masks, uint32_t* codes, float* vals, const uint64_t* flag_list,
                             uint64_t count, int all_items, float* work, hipStream_t st);

// Query side: per (q, l, layer) the top-PL (code, |value|) largest first:
// qc/qv[((q*L + l)*2 + layer)*PL + j]; qflag[q] |= 1 when a tie at or above
// the PL boundary makes the selection depend on heap internals.
void launch_hash_query(const HashGeom& g, const float* Q, int64_t ldq, int nq, int PL, const uint32_t* masks,
                       int32_t* qc, float* qv, uint32_t* qflag, hipStream_t st);

// ---- build (build.hip) ----------------------------------------------------
void launch_center_partial(const float* X, int64_t ld, int64_t N, int D, double* partial, int64_t chunk,
                           int nchunks, hipStream_t st);
void launch_center_final(const double* partial, int nchunks, int D, int64_t N, float* center, hipStream_t st);
void launch_center_apply(const float* X, int64_t ld, int64_t N, int D, const float* center, float* Xc,
                         int64_t ldx, hipStream_t st);
void launch_bucket_count(const uint32_t* codes, int64_t per_table, int L, int NB, uint32_t* cnt, hipStream_t st);
void launch_bucket_limit(const uint32_t* cnt, int64_t nbk, float alpha, int p, int bucket_min, uint32_t* ksz,
                         hipStream_t st);
// exclusive scan of n u32 values into n+1 u64 offsets (out[n] = total)
void scan_u32_to_u64(const uint32_t* in, int64_t n, uint64_t* out, uint64_t* tmp, hipStream_t st);
int64_t scan_tmp_elems(int64_t n);
// stable LSD radix sort of (key, val) by the low `bits` key bits; if vals_in
// is null the values are 0..n-1.  Returns through (*keys_out, *vals_out)
// which of the two buffer pairs holds the result.
void radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, int64_t n, int bits, bool iota_vals,
                      uint32_t* hist, uint64_t* hist_off, uint64_t* scan_tmp, uint32_t** keys_out,
                      uint32_t** vals_out, hipStream_t st);
int64_t radix_hist_elems(int64_t n);
void launch_make_select_list(const uint32_t* cnt, const uint32_t* ksz, int NB, uint32_t* list, uint32_t* count,
                             hipStream_t st);
void launch_seg_copy(const uint32_t* skeys, const uint32_t* svals, int64_t n, int p, const uint32_t* cnt,
                     const uint32_t* ksz, const uint64_t* tstart, const uint64_t* bucket_off, int32_t* ids,
                     hipStream_t st);
void launch_seg_select(const uint32_t* list, uint32_t nlist, const uint32_t* svals, const float* vals, int p,
                       const uint32_t* cnt, const uint32_t* ksz, const uint64_t* tstart,
                       const uint64_t* bucket_off, int32_t* ids, uint32_t* tie_list, uint32_t* tie_count,
                       hipStream_t st);
void launch_seg_exact(const uint32_t* tie_list, uint32_t ntie, const uint64_t* work_off, const uint32_t* svals,
                      const float* vals, int p, const uint32_t* cnt, const uint32_t* ksz, const uint64_t* tstart,
                      const uint64_t* bucket_off, int32_t* ids, void* work, hipStream_t st);

// ---- query (query.hip) ----------------------------------------------------
// Per-query parameters of a batch that mixes probe counts and k (perTable
// and PL follow from qProbes by the reference's formulas).
struct QueryParams {
    int32_t qProbes, perTable, PL, k;
    uint64_t probe_off;  // first probe of the query in the probe array
};

struct ProbeGeom {
    int L, NB, PL, perTable, qProbes;  // PL: stride of the per-layer top lists
    int two_np;  // 2 * n_proj: codes per layer
    const QueryParams* qp = nullptr;  // per query (null: the uniform values above)
};

// Parameters of query q: its own when the batch has per-query parameters.
struct QueryView {
    int qProbes, perTable, PL, k;
    uint64_t probe_off;
};
__device__ __forceinline__ QueryView query_view(const QueryParams* qp, int q, int qProbes, int perTable, int PL,
                                                int k) {
    // q is the same for the whole workgroup: keep the fields scalar
    QueryView v{qProbes, perTable, PL, k, (uint64_t)q * (uint64_t)qProbes};
    if (qp) {
        const QueryParams p = qp[q];
        v = QueryView{p.qProbes, p.perTable, p.PL, p.k, p.probe_off};
    }
    v.qProbes = __builtin_amdgcn_readfirstlane(v.qProbes);
    v.perTable = __builtin_amdgcn_readfirstlane(v.perTable);
    v.PL = __builtin_amdgcn_readfirstlane(v.PL);
    v.k = __builtin_amdgcn_readfirstlane(v.k);
    v.probe_off = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(v.probe_off >> 32)) << 32) |
                  (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v.probe_off);
    return v;
}
// Global top-qProbes bucket set per query; ctot[q] = sum of probed sizes;
// qflag[q] |= 2 when a tie straddles a per-table or the global boundary.
// probe_select_path() tells which selection a launch takes: the pool of every
// pair, the per-table top lists (two-stage), or none because neither fits the
// LDS (the caller then flags every query for the exact replay).
enum { PROBE_FULL_POOL = 0, PROBE_TWO_STAGE = 1, PROBE_EXACT = 2 };
int probe_select_path(const ProbeGeom& g);
void launch_probe_select(const ProbeGeom& g, int nq, const int32_t* qc, const float* qv,
                         const uint64_t* bucket_off, uint32_t* probes, uint64_t* ctot, uint32_t* qflag,
                         hipStream_t st);
// The queries with a nonzero qflag, in query order, into list; *count =
// their number (device side, so no host read is needed).
void launch_compact_flagged(int nq, const uint32_t* qflag, uint32_t* list, uint32_t* count, hipStream_t st);
// Exact replay of the probing sequence of the *nflag queries in qlist, as
// separate kernels (used when every query's C_tot is needed before the
// gather): per-table lists, the global cut by radix select, and the
// reference's heap merge only where equal values straddle the cut.  Writes
// their probes and ctot.  The first `cap` flagged queries get one wave per
// (query, table) and a work area each; the rest go to `over` persistent
// workgroups (claim zeroed beforehand) with an area each.  work holds
// (cap + over) * wpq bytes, wpq = probe_replay_work_bytes(g).
void launch_probe_replay(const HashGeom& hg, const ProbeGeom& g, const float* Q, int64_t ldq, const uint32_t* masks,
                         const uint32_t* qlist, const uint32_t* nflag, uint32_t* claim, const uint32_t* qflag,
                         const int32_t* qc, const float* qv, const uint64_t* bucket_off, uint32_t* probes,
                         uint64_t* ctot, void* work, int64_t wpq, unsigned cap, unsigned over, hipStream_t st);

// Exact replay of a flagged query's probing sequence done by the gather
// workgroup that claims it, before its gather: the workgroup's waves rebuild
// the per-table lists into work + blockIdx.x * wpq, then the cut or merge.
// qflag null: no query is replayed there.
struct ReplayArgs {
    HashGeom hg;
    ProbeGeom pg;
    const float* Q = nullptr;  // padded queries, row stride ldq
    int64_t ldq = 0;
    const uint32_t* masks = nullptr;
    const uint32_t* qflag = nullptr;
    const int32_t* qc = nullptr;
    const float* qv = nullptr;
    const uint64_t* bucket_off = nullptr;
    uint32_t* probes = nullptr;
    uint64_t* ctot = nullptr;
    unsigned char* work = nullptr;
    int64_t wpq = 0;
    int table_lds = 0;     // LDS bytes of one table's wave
    int twaves = 0;        // waves that rebuild tables
    int merge_staged = 0;  // lists staged in LDS for the merge
};
// LDS the replay needs inside a gather workgroup of `threads`, per_cu of
// them sharing a CU (0 when a workgroup cannot hold one table's wave); fills
// r's LDS fields.
size_t replay_lds_bytes(ReplayArgs& r, int threads, int per_cu);
int64_t probe_replay_work_bytes(const ProbeGeom& g);

// One persistent gather launch: every workgroup takes queries from a work
// queue (claim_query) until it is drained.
// The bucket id array is allocated with IDS_PAD entries past its end: the
// gather reads a bucket's ids four at a time and may run up to three past
// the last id.
constexpr int64_t IDS_PAD = 4;

// Wall-clock stamps per query of a gather: 0 start, 1 bucket walk done,
// 2 candidates collected, 3 scored, 4 done.
constexpr int PHASE_STAMPS = 5;

struct GatherArgs {
    int64_t N;
    int D;
    int64_t ldx;          // row stride of the scored copy (elements)
    int qProbes;
    int k;
    int64_t window_bits;  // LDS bitset window
    uint64_t* phase_ts = nullptr;     // per query PHASE_STAMPS wall-clock stamps (null: off)
    const QueryParams* qp = nullptr;  // per-query probe counts and k (null: qProbes, k above)
    int ldo = 0;                      // row stride of the outputs (0: k)
    int small_at = -1;                // see the work queue below
    // work queue: order[0 .. nitems) (identity when null) in nqueues runs;
    // heads[nqueues] zeroed beforehand
    // small_at >= 0: the last heads[small_at] items of order are small
    // queries, served one per wave by launch_gather_small; the workgroup
    // kernels take the items before them
    const uint32_t* order = nullptr;
    uint32_t nitems = 0;
    int nqueues = 1;
    uint32_t* heads = nullptr;
    // scratch: workgroup b uses cand[b*slot_cand ...] (the unique candidates,
    // at least min(C_tot, N) per query) and part[b*slot_part ...] (the scatter
    // by dedup window, at least C_tot with several windows; FP32 scores or
    // approximate score keys).  With seg_off, query q uses both at seg_off[q]
    // instead (its candidates stay there after the call).
    int32_t* cand = nullptr;
    int32_t* part = nullptr;
    uint64_t slot_cand = 0, slot_part = 0;
    const uint64_t* seg_off = nullptr;
    uint64_t* ncand = nullptr;  // unique candidates per query
    const uint64_t* ctot = nullptr;  // C_tot per query (sum of its probed bucket sizes)
};
// Small queries: at most SMALL_MAX_CTOT bucket ids (C_tot), not tie-flagged,
// k <= 64.  One wave gathers, dedups and scores such a query (FP32, every
// candidate, in every precision mode).  SmallPick: the test (max_ctot 0: no
// query is small); count receives the number of small queries.
constexpr uint32_t SMALL_MAX_CTOT = 512;
struct SmallPick {
    const uint64_t* ctot = nullptr;
    const QueryParams* qp = nullptr;  // per-query k (null: k)
    int k = 0;
    uint32_t max_ctot = 0;
    uint32_t* count = nullptr;
};
// Query order sort keys: the best bucket of table 0 (queries with equal keys
// share most candidates), or C_tot descending (longest first).
// Queries whose qflag is set (replayed inside the gather) get key 0, ahead of
// the others; small queries (sp) get keys above every other, so they end the
// order: for the bucket keys the bit above the bucket's (sort one more bit),
// for the C_tot keys every other query's key is held below them.
void launch_order_keys(const int32_t* qc, const uint32_t* qflag, int nq, int L, int PL, int two_np, uint32_t* key,
                       const SmallPick& sp, hipStream_t st);
void launch_ctot_keys(const uint64_t* ctot, const uint32_t* qflag, int nq, uint32_t* key, const SmallPick& sp,
                      hipStream_t st);
// out[0 .. n) = 0 .. n-1 stably ordered by the low `bits` bits of keys, in
// one workgroup (n <= SORT1_MAX; the order radix_sort_pairs gives).  With
// split, the positions whose key is >= small_min (small queries) come last in
// position order, and only the others are sorted.
constexpr int SORT1_MAX = 11264;
void launch_sort1(const uint32_t* keys, int n, int bits, bool split, uint32_t small_min, uint32_t* out,
                  hipStream_t st);
// dst[r*ldd + c] = src[r*lds + c] for c < D, 0 up to ldd (queries padded to
// the scored row stride).
void launch_pad_rows(const float* src, int64_t lds, int64_t rows, int D, float* dst, int64_t ldd, hipStream_t st);
// launch_pad_rows, and in the same kernel flags[0 .. nflags) = fv and
// counters[0 .. ncounters) = 0.
void launch_pad_rows_init(const float* src, int64_t lds, int64_t rows, int D, float* dst, int64_t ldd,
                          uint32_t* flags, uint32_t fv, int nflags, uint32_t* counters, int ncounters,
                          hipStream_t st);
// p[0 .. n) = v (a kernel, so it can be captured in a graph).
void launch_fill_u32(uint32_t* p, uint32_t v, int n, hipStream_t st);
// Workgroups of a persistent launch over nitems: per_cu on every CU, fewer
// when their LDS does not fit, no more than there are items.
unsigned persistent_grid(int nitems, size_t lds_bytes, int per_cu);
// Sizes of the buckets of a built index (synchronises st): hist[s] = number
// of buckets holding s < SIZE_BINS ids; big = the sizes of the others.
// Returns false when there were more than 65536 of those (big incomplete).
constexpr int SIZE_BINS = 4096;
bool bucket_size_histogram(const uint64_t* off, int64_t nbk, std::vector<uint32_t>& hist, std::vector<uint32_t>& big,
                           hipStream_t st);
// Largest bucket of a built index (synchronises st); scratch: one word.
uint32_t max_bucket_size(const uint64_t* off, int64_t nbk, uint32_t* scratch, hipStream_t st);
// Dedup windows over the ids [0, N) and the LDS words of their counters
// (count, segment offset, fill per window) when there is more than one.
__host__ __device__ inline int64_t window_count(int64_t N, int64_t W) { return (N + W - 1) / W; }
__host__ __device__ inline int64_t window_counter_words(int64_t N, int64_t W) {
    return N > W ? (3 * window_count(N, W) + 1) & ~int64_t(1) : 0;
}
// Dedup window (ids): all of N when its bitset fits in LDS beside
// extra_words of scratch, else the largest power of two that fits together
// with the window counters.
// number of probes[0 .. n) with an empty bucket; scratch2: two device words
uint64_t count_empty_probes(const uint32_t* probes, int64_t n, const uint64_t* bucket_off, uint32_t* scratch2,
                            hipStream_t st);
int64_t gather_window_bits(int64_t N, int64_t extra_words, int per_cu);
// Dedup (LDS bitset windows), FP32 rerank and top-k of each query, one
// persistent launch of `grid` workgroups (persistent_grid with the LDS of
// gather_f32_lds); a.ncand[q] = unique candidates.
size_t gather_f32_lds(const GatherArgs& a, int threads);
// The small queries of a call (the end of a.order, a.small_at), one wave
// each, wave g of the launch taking items g, g + waves, ...: candidates
// deduped in the wave's LDS, FP32 scores in the gather kernel's order of
// addition, top-k in the wave.  a.ncand[q] = unique candidates; with
// a.seg_off the candidates are also kept at a.cand + seg_off[q].
void launch_gather_small(const GatherArgs& a, const uint32_t* probes, const uint64_t* bucket_off, const int32_t* ids,
                         const float* Xc, const float* Q, int64_t ldq, int32_t* out_ids, float* out_scores,
                         hipStream_t st);
void launch_gather_score_f32(const GatherArgs& a, const ReplayArgs& rp, size_t lds, unsigned grid,
                             const uint32_t* probes, const uint64_t* bucket_off, const int32_t* ids, const float* Xc,
                             const float* Q, int64_t ldq, int32_t* out_ids, float* out_scores, int threads,
                             hipStream_t st);

// ---- reduced-precision scoring + FP32 refine (lowprec.hip) ----------------
struct LowPrecArgs {
    int format;              // 1 = BF16, 2 = FP8-e4m3 with per-row scale
    int refine;              // survivors rescored in FP32 (R), 1..256; with per-query k the largest
    int refine_per_k = 0;    // 1: query q keeps min(refine, 4 * k_q) survivors
    int64_t ldl;             // row stride of the low-precision copy (elements; 128 bytes a multiple)
    const uint16_t* xb;      // BF16 copy (bit patterns)
    const uint8_t* x8;       // FP8 copy
    const float* scale;      // FP8 per-row scale
    float uscale = 0.0f;     // FP8 scale shared by every row (0: per-row scales)
    int threads = 512;       // workgroup size, 256 or 512
    int scale_in_row = 0;    // FP8 per-row scale in the row's last four bytes (fp8_scale_in_row)
};
// Whether an FP8 copy keeps each row's scale in the row's last four bytes
// (per-row scales and at least four bytes of padding): scoring then reads
// one line per candidate instead of the line and a separate scale.
bool fp8_scale_in_row(int64_t ldl, int D, float uscale);
int64_t lowprec_extra_words(int refine, int threads);
void launch_quant_bf16(const float* Xc, int64_t ldx, int64_t N, int D, int64_t ldl, uint16_t* Xb, hipStream_t st);
// returns the scale shared by every row, or 0 when rows keep their own
float launch_quant_fp8(const float* Xc, int64_t ldx, int64_t N, int D, int64_t ldl, uint8_t* X8, float* scale,
                       uint32_t* range, bool per_row, hipStream_t st);
size_t gather_lp_lds(const GatherArgs& a, const LowPrecArgs& lp);
void launch_gather_score_lp(const GatherArgs& a, const LowPrecArgs& lp, const ReplayArgs& rp, size_t lds,
                            unsigned grid, const uint32_t* probes, const uint64_t* bucket_off, const int32_t* ids,
                            const float* Xc, const float* Q, int64_t ldq, int32_t* out_ids, float* out_scores,
                            int per_cu, hipStream_t st);

// ---- FP64 brute force (bf.hip), verification only -------------------------
void launch_bf64(const float* X, int64_t ldx, int64_t N, int D, const float* Q, int64_t ldq, int nq, int k,
                 const double* thresh, uint32_t* hit_q, uint32_t* hit_id, double* hit_s, uint32_t* hit_count,
                 uint32_t hit_cap, hipStream_t st);
void launch_dot64_rows(const float* X, int64_t ldx, int D, const float* Q, int64_t ldq, int nq, int k,
                       const int32_t* ids, double* out, hipStream_t st);

}  // namespace fpp
