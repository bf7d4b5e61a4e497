/*
 * C ABI of the MI355X (gfx950) Falconn++ engine.
 *
 * The entry points mirror the reference's FalconnPP class one to one: its
 * only foreign-function boundary is the pybind11 module that binds that class
 * (python/python_wrapper.cpp:14-30).  Each declaration names the reference
 * member it replaces.  Plain pointers and sizes only; every call returns 0 on
 * success or an FPP_E* status, with the message in fpp_last_error().
 *
 * Matrices follow the reference layout: X is D x N column-major (each point's
 * D floats contiguous, consecutive points `ld` floats apart); Q is D x nq the
 * same way.  Results are nq x k row-major int32 ids, best first; slots with no
 * candidate stay 0 (src/FalconnPP.cpp:615-673).
 *
 * An index is not thread-safe: like the reference (mutable qProbes, global
 * OpenMP thread count) one index serves one caller at a time.
 */
#ifndef FALCONNPP_AMD_H
#define FALCONNPP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FPP_OK = 0,
    FPP_EINVAL = 1,    /* bad argument */
    FPP_ESTATE = 2,    /* call out of order (query before build, ...) */
    FPP_EHIP = 3,      /* HIP runtime error */
    FPP_ENOMEM = 4,    /* allocation failed */
    FPP_EINTERNAL = 5, /* engine invariant broken */
    FPP_EIO = 6        /* file cannot be opened, read or written */
};

typedef struct fpp_index fpp_index;

/* Message of the last failed call on this thread ("" if none). */
const char* fpp_last_error(void);

/* FalconnPP(int n, int d) -- src/FalconnPP.h:80-83 */
int fpp_create(int n_points, int n_features, fpp_index** out);
/* ~FalconnPP() -- src/FalconnPP.h:131 */
void fpp_destroy(fpp_index* h);

/* Index2Layers(L, D, bucketLimit, alpha, p, t, s) -- src/FalconnPP.h:85-100,
 * bound as setIndexParam(n_tables, n_proj, bucket_minSize, bucket_scale,
 * iProbes, n_threads=8, random_seed=-1) at python/python_wrapper.cpp:19-22.
 * n_threads sizes host-side work only.  seed <= -1 draws from the clock. */
int fpp_set_index_param(fpp_index* h, int n_tables, int n_proj, int bucket_minSize, float bucket_scale,
                        int iProbes, int n_threads, int seed);
/* set_qProbes(p) -- src/FalconnPP.h:113-115 */
int fpp_set_qprobes(fpp_index* h, int qProbes);
/* set_threads(t) -- src/FalconnPP.h:117-123 */
int fpp_set_threads(fpp_index* h, int n_threads);
/* clear() -- src/FalconnPP.h:102-111: drops data, tables and sign bits. */
int fpp_clear(fpp_index* h);

/* build2Layers_1D(X) -- src/FalconnPP.cpp:21-303, bound as build(dataset) at
 * python/python_wrapper.cpp:28.  X: host D x N col-major, leading dim ld >= D.
 * The index keeps its own centred copy in HBM (src/FalconnPP.cpp:33-34). */
int fpp_build(fpp_index* h, const float* X, int64_t ld);
/* Same with X already in device memory (HBM); stream may be NULL. */
int fpp_build_device(fpp_index* h, const float* dX, int64_t ld, void* stream);

/* query2Layers_1D(Q, k, verbose) -- src/FalconnPP.cpp:310-707, bound as
 * query(queries, n_neighbors, verbose=False) at python/python_wrapper.cpp:29-30.
 * verbose != 0 prints the reference's per-call statistics and writes the ids
 * to fpp_topk_file_name(...) in the working directory, as the reference does
 * (src/FalconnPP.cpp:682-704).
 * Q: host D x nq col-major (ldq >= D); out: host nq x k. */
int fpp_query(fpp_index* h, const float* Q, int64_t ldq, int nq, int k, int32_t* out, int verbose);
/* query2Layers(Q, k, verbose) -- src/FalconnPP.cpp:1013-1413, bound as
 * query2D(queries, n_neighbors, verbose=False) at python/python_wrapper.cpp:27.
 * The 2D-storage index holds the same bucket contents as the 1D one, so this
 * answers from the same tables with the same ids; verbose output goes to
 * fpp_topk_file_name_2d(...) (2D_Top_..., src/FalconnPP.cpp:1400-1406). */
int fpp_query_2d(fpp_index* h, const float* Q, int64_t ldq, int nq, int k, int32_t* out, int verbose);
/* Same with Q and outputs in device memory; scores (nq x k, FP32 q.(x-c)) may
 * be NULL. */
int fpp_query_device(fpp_index* h, const float* dQ, int64_t ldq, int nq, int k, int32_t* d_out, float* d_scores,
                     void* stream);

/* ---- streaming insert (no reference counterpart: the reference keeps its 2D
 * layout "to support delete and insert new points in future",
 * src/FalconnPP.cpp:711-713, and implements none) ---------------------------
 * fpp_set_insertable(h, 1) before build keeps every point's bucket entries
 * from before the alpha / bucketLimit filter (n_points * n_tables * iProbes
 * codes and values, 8 bytes each).  fpp_insert then appends m points (host,
 * D x m col-major, leading dimension ld) as ids n_points .. n_points+m-1:
 * centred with the index's centre (not recomputed), hashed with its sign
 * bits, and every bucket filtered again over old and new entries in point
 * order, so the index equals build2Layers_1D (src/FalconnPP.cpp:21-303) of
 * all points with that centre.  Cost: that of a build without the centre
 * and the old points' hashing.  FPP_ESTATE when the index keeps no entries. */
int fpp_set_insertable(fpp_index* h, int on);
int fpp_insert(fpp_index* h, const float* X, int64_t ld, int m);
/* Same with the points in device memory; stream may be NULL. */
int fpp_insert_device(fpp_index* h, const float* dX, int64_t ld, int m, void* stream);

/* No counterpart in the reference.  fpp_query_device that allocates nothing,
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, serving loops and harnesses that forbid allocation.  Every
 * scratch buffer of the call (query copy, hash lists, probes, counts,
 * ordering, replay work, candidate slots) lives in one fixed device array of
 * 256 MiB that the library holds per GPU until the process ends.  A batch
 * that does not fit (its per-query arrays, about nq * (n_tables * perLayer
 * * 16 + qProbes * 4) bytes, plus one candidate slot) fails with FPP_EINVAL
 * before any work; a batch that fits with fewer candidate slots than the
 * allocating call uses runs on fewer gather workgroups.  FP32 scoring needs
 * nothing else; BF16 / FP8 need the reduced-precision copy, made by one
 * fpp_query_device call first (FPP_ESTATE otherwise).  The results equal
 * fpp_query_device's.  Calls on one GPU share the array, which keeps the
 * last call's contents: calls that overlap on one GPU (two streams, two
 * threads) race and must be serialised by the caller.  stream may be NULL
 * (the index's stream if it has one, else the null stream; the call then
 * waits for its results). */
int fpp_query_device_noalloc(fpp_index* h, const float* dQ, int64_t ldq, int nq, int k, int32_t* d_out,
                             float* d_scores, void* stream);

/* ---- inspection (parity tests and tools; no reference counterpart) ------ */
/* index shape: out[0]=fhtDim, [1]=4*n_proj^2 buckets per table, [2]=n_tables,
 * [3]=number of stored ids, [4]=row stride of the centred copy,
 * [5]=n_points (grows with fpp_insert).  out must hold 6 values. */
int fpp_get_shape(fpp_index* h, int64_t* out);
/* centre of the data (n_features floats) */
int fpp_get_center(fpp_index* h, float* out);
/* the two sign-bit vectors of src/FalconnPP.h:52-75 (fhtDim*n_tables*3 each) */
int fpp_get_sign_bits(fpp_index* h, uint8_t* bit1, uint8_t* bit2);
/* CSR of the tables: off[n_tables*4*n_proj^2 + 1] and ids[off[last]];
 * bucket b (global id = table*4*n_proj^2 + code) holds ids[off[b]..off[b+1])
 * in ascending id order. */
int fpp_get_buckets(fpp_index* h, uint64_t* off, int32_t* ids);
/* bucket sizes before the alpha / bucketLimit filter, the B of
 * src/FalconnPP.cpp:226-242: cnt[n_tables*4*n_proj^2].  FPP_ESTATE for an
 * index restored from a file. */
int fpp_get_bucket_counts(fpp_index* h, uint32_t* cnt);
/* build-side hash of m host vectors (D x m col-major, taken as already
 * centred): codes/vals[(l*m + j)*iProbes + i], i in ascending value order.
 * Needs the sign bits of a previous build. */
int fpp_hash_points(fpp_index* h, const float* V, int64_t ld, int m, uint32_t* codes, float* vals);
/* statistics: out[0]=build hash items replayed for ties, [1]=buckets replayed
 * for ties, [2]=sum C_tot of the last query call, [3]=sum C_uniq, [4]=queries
 * replayed for ties, [5]=queries in the last call, [6]=probe selection of the
 * last call (0: pool of every pair, 1: per-table top lists, 2: exact replay of
 * every query because neither pool fits the LDS), [7]=queries of the last
 * call served one per wave (FPP_TUNE_SMALL_QUERIES).  out must hold 8
 * values. */
int fpp_get_stats(fpp_index* h, int64_t* out);
/* per-stage device time of the last query call when profiling is on
 * (on >= 1: HIP events on the call's stream): ms[0]=query hashing,
 * [1]=probe selection, [2]=tie replay, [3]=gather + dedup + rerank + top-k;
 * with on >= 2 also the mean per-query microseconds of that last kernel's
 * phases (device wall clock): [4]=bucket gather + dedup, [5]=candidate
 * compaction, [6]=scoring, [7]=top-k, [8]=the longest query's time in that
 * kernel, [9]=the longest first phase of a tie-flagged query (its exact probe
 * replay when done inside the gather, then its bucket gather).  ms must hold
 * 10 doubles. */
int fpp_set_profiling(fpp_index* h, int on);
int fpp_get_timings(fpp_index* h, double* ms);
/* Scoring precision (no reference counterpart; the reference scores in FP32):
 * mode 0 = FP32 (default, reference results), 1 = BF16 copy scored on MFMA,
 * 2 = FP8-e4m3 copy with an FP32 scale per row (one scale for all rows when
 * their largest magnitudes lie within 2^8 of each other) scored on MFMA.  With
 * per-query k (fpp_query_batch_device) and refine 0, query q keeps
 * min(256, 4 * k_q).  Modes 1-2 keep
 * the `refine` best approximate candidates (0: 4 * n_neighbors, at most 256)
 * and rescore them exactly in FP32, so returned scores are FP32 scores; the
 * ids can differ from mode 0 when a true top-k row falls outside the refine
 * set.  The low-precision copy is made on the first query after a build. */
int fpp_set_precision(fpp_index* h, int mode, int refine);
/* The FP8 copy's scale (no reference counterpart): *shared receives the scale
 * every row uses, or 0 when rows keep their own.  FPP_ESTATE before an FP8
 * query has made the copy. */
int fpp_get_fp8_scale(fpp_index* h, float* shared);
/* launch-shape knobs (no reference counterpart).  The gather is one
 * persistent launch whose workgroups take queries from a work queue.
 * FPP_TUNE_QUERY_ORDER: the queue's order, 0 = query order, 1 = by the best
 * table-0 bucket (queries sharing candidates together) in 8 runs, one per
 * XCD (default), 2 = longest first (by C_tot).  FPP_TUNE_GATHER_THREADS
 * 0/256/512/1024 (workgroup of the gather kernel; 256 runs two workgroups per
 * CU, the reduced-precision kernels use 256 or 512; 0, the default, runs
 * two workgroups per CU when n_points needs several dedup windows even for
 * one 512-thread workgroup per CU: 256 threads for the FP32 gather, 512 with
 * half the rows in flight per wave for BF16/FP8; else one of 512).  FPP_TUNE_PERSISTENT 0/1
 * kept for callers of earlier versions: 1 sets order 2, 0 sets order 1.
 * FPP_TUNE_GRAPH 0/1: the whole query call (copy, hashing, probe selection,
 * exact replay, ordering, gather) captured once in a hipGraph and replayed
 * while the call's shape and buffers stay the same.  FPP_TUNE_VERBOSE 0/1:
 * the autotuner prints one line per launch shape it times to stderr.
 * FPP_TUNE_KEEP_CANDIDATES 0/1: every query's candidate list is kept after
 * the call for fpp_get_last_candidates (costs one host read per call and
 * memory for all lists).  None of these changes any result.  FPP_TUNE_FP8_PER_ROW 0/1: the FP8 copy keeps one
 * scale per row even where one shared scale would do (the copy is remade on
 * the next FP8 query); it changes the approximate scores, so it can change
 * which candidates reach the FP32 refine.  FPP_TUNE_REPLAY: where the exact
 * probe replay of the tie-flagged queries runs, 0 = chosen per call (default),
 * 1 = separate kernels before the gather (each (query, table) on a wave of its
 * own across the device), 2 = inside the gather, by the workgroup that claims
 * the query, before it gathers it (flagged queries head the queue).
 * FPP_TUNE_WORKSPACE_MB: bound, in MiB, of the per-workgroup candidate
 * scratch of a query call (default 8192); a tighter bound first gives the
 * dedup windows smaller bins (a query whose ids overflow one takes two walks
 * over its buckets instead of one), then runs the gather on fewer
 * workgroups, with the same results.  FPP_TUNE_SMALL_QUERIES 0/1
 * (default 1): a query whose probed buckets hold at most 512 ids in all, that
 * needs no exact tie replay and asks for at most 64 neighbours is gathered
 * and scored by one wave instead of a workgroup (with a sorted query order,
 * FPP_TUNE_QUERY_ORDER 1 or 2, and more than one query).  FP32 results are
 * the same either way; in the reduced-precision modes such a query's
 * candidates are all scored exactly in FP32. */
enum {
    FPP_TUNE_QUERY_ORDER = 1,
    FPP_TUNE_GATHER_THREADS = 2,
    FPP_TUNE_PERSISTENT = 3,
    FPP_TUNE_GRAPH = 4,
    FPP_TUNE_VERBOSE = 5,
    FPP_TUNE_FP8_PER_ROW = 6,
    FPP_TUNE_KEEP_CANDIDATES = 7,
    FPP_TUNE_REPLAY = 8,
    FPP_TUNE_WORKSPACE_MB = 9,
    FPP_TUNE_SMALL_QUERIES = 10
};
int fpp_set_tuning(fpp_index* h, int which, int value);
/* Autotuner (no reference counterpart): times the gather workgroup sizes,
 * with one workgroup per query (query order off and on) and with the
 * persistent launch, on the given device queries, and keeps the fastest.
 * chosen (may be NULL) receives {threads, order, persistent, microseconds
 * per batch}.  fpp_autotune_batch times batched calls (qprobes, ks: HOST
 * arrays of nq, as fpp_query_batch_device takes), so that a sample of a
 * mixed batch tunes for that mix. */
int fpp_autotune(fpp_index* h, const float* dQ, int64_t ldq, int nq, int k, void* stream, int32_t* chosen);
int fpp_autotune_batch(fpp_index* h, const float* dQ, int64_t ldq, int nq, const int32_t* qprobes,
                       const int32_t* ks, int kmax, void* stream, int32_t* chosen);
/* Batched serving (no reference counterpart; each query is
 * query2Layers_1D, src/FalconnPP.cpp:310-707, with its own qProbes and k):
 * qprobes and ks are HOST arrays of nq, dQ/d_out/d_scores device pointers.
 * Row q of d_out (nq x kmax int32) holds the ks[q] best ids, then zeros;
 * d_scores (may be NULL) likewise. */
int fpp_query_batch_device(fpp_index* h, const float* dQ, int64_t ldq, int nq, const int32_t* qprobes,
                           const int32_t* ks, int kmax, int32_t* d_out, float* d_scores, void* stream);
/* Per-query times of the last query call made with profiling level 2 (no
 * reference counterpart): service_us[q] = time query q spent in the gather
 * kernel, done_us[q] = its completion after the first query of the call
 * entered the gather.  Both arrays hold nq doubles. */
int fpp_get_query_times(fpp_index* h, double* service_us, double* done_us);
/* probes of the last query call (global bucket ids, qProbes per query in
 * query order, as a set per query) and, for query j, its unique candidates
 * (ascending; returns count) */
int fpp_get_last_probes(fpp_index* h, uint32_t* out);
int64_t fpp_get_last_candidates(fpp_index* h, int j, int32_t* out, int64_t cap);

/* FP64 brute-force top-k over all N (src/BF.cpp:17-87 in FP64), device
 * inputs: X raw (not centred) D x N col-major, Q D x nq.  seed_ids (host,
 * nq x k, may be NULL) are known rows used to bound each query's k-th score.
 * Outputs on the host, best first.  For recall checks only. */
int fpp_bruteforce_f64(const float* dX, int64_t ldx, int64_t N, int D, const float* dQ, int64_t ldq, int nq, int k,
                       const int32_t* seed_ids, int32_t* out_ids, double* out_scores);

/* ---- index files (no reference counterpart: the reference rebuilds its
 * index in every process) */
/* Writes the built index (parameters, sign bits, centre, CSR tables, centred
 * data) to path.  FPP_ESTATE before build; FPP_EIO when the file cannot be
 * written. */
int fpp_save(fpp_index* h, const char* path);
/* Restores an index written by fpp_save into h, which must have been created
 * with the file's n_points and n_features; afterwards h answers queries as the
 * saved index did, without building.  FPP_EIO when the file cannot be read,
 * FPP_EINVAL when it is not an index file, is truncated or does not match.  On
 * a failure after the checks, h is left unbuilt. */
int fpp_load(fpp_index* h, const char* path);

/* ---- text formats of the reference's command-line driver (no device use) */
/* loadtxtData(dataset, n, d, X) -- src/Utilities.cpp:41-60: n_points *
 * n_features floats, whitespace separated (one point per line in the
 * reference's files), read in file order into out, i.e. n_points x n_features
 * row-major = the D x N column-major layout fpp_build takes.  Values the file
 * lacks, or values past the first token that is not a number, stay 0, as with
 * the reference's fscanf into a zeroed matrix.  FPP_EIO when the file cannot
 * be opened.  *n_read (may be NULL) receives the number of values read. */
int fpp_load_txt(const char* path, int n_points, int n_features, float* out, int64_t* n_read);
/* outputFile(matKNN, file) -- src/Utilities.cpp:13-31: one line per query,
 * each of its k ids followed by a space.  ids: nq x k row-major. */
int fpp_write_topk_txt(const char* path, const int32_t* ids, int nq, int k);
/* Name of the file query2Layers_1D writes when verbose
 * (src/FalconnPP.cpp:694-700): 1D_Top_<k>_D_<n_proj>_L_<n_tables>_iProbe_<iProbes>
 * _bucketMinSize_<bucket_minSize>_bucketScale_<(int)(bucket_scale*100)>_qProbe_<qProbes>.txt.
 * buf receives at most buflen bytes including the terminating 0 (FPP_EINVAL
 * when it is too short). */
int fpp_topk_file_name(int k, int n_proj, int n_tables, int iProbes, int bucket_minSize, float bucket_scale,
                       int qProbes, char* buf, int buflen);
/* The same with the 2D_Top_ prefix of query2Layers (src/FalconnPP.cpp:1400-1406). */
int fpp_topk_file_name_2d(int k, int n_proj, int n_tables, int iProbes, int bucket_minSize, float bucket_scale,
                          int qProbes, char* buf, int buflen);

#ifdef __cplusplus
}
#endif

#endif /* FALCONNPP_AMD_H */
