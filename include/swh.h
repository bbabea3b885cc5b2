/*
 * swh.h -- C ABI of the MI355X Smith-Waterman search library (libswh.so).
 *
 * Plain pointers and sizes only.  Every call returns an int status
 * (SWH_OK = 0, negative on failure) and leaves a message for swh_last_error()
 * on the calling thread; the library never exits the process.  Calls on one
 * context must be serialised by the caller; distinct contexts are independent.
 *
 * Conventions
 *   residue codes   0..23 (A0 B1 C2 D3 E4 F5 G6 H7 I8 K9 L10 M11 N12 P13 Q14
 *                   R15 S16 T17 V18 W19 X20 Y21 Z22, 23 = pad), as produced by
 *                   the reference's translation table (src/DataPool.cpp:33-47).
 *                   Database codes above 23 are treated as 23, as the
 *                   reference's packDB does (src/Worker.cpp:150-154); query
 *                   codes above 23 are rejected.
 *   matrix          32x32 int32, row = query code, column = database code
 *                   (src/Param.h:110-115).  Entries must lie in [-128, 127].
 *   gaps            gap_open, gap_extend as the reference's -o / -e flags; the
 *                   first gap residue costs gap_open + gap_extend
 *                   (src/Param.h:109).
 *   scores          exact int32 local-alignment scores, one per database
 *                   sequence, in the caller's sequence order.
 */
#ifndef SWH_H_
#define SWH_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWH_OK 0
#define SWH_ERR_ARG (-1)
#define SWH_ERR_HIP (-2)
#define SWH_ERR_NOMEM (-3)
#define SWH_ERR_STATE (-4)
#define SWH_ERR_CAPACITY (-5)

typedef struct swh_ctx swh_ctx;
typedef struct swh_db swh_db;

/* Timing and counters of the last search on a context. */
typedef struct swh_stats {
    double kernel_ms;         /* int16 scan kernel (HIP events on the search stream) */
    double rescore_ms;        /* int32 rescore kernel */
    double total_ms;          /* whole call, host wall clock */
    uint64_t cells;           /* sum over sequences of len(db seq) * len(query) */
    uint64_t cells_group;     /*   of which in group-mode (long) sequences */
    uint64_t padded_cells;    /* cells actually computed, padding included */
    uint64_t n_overflow;      /* sequences rescored in int32 */
    uint64_t n_launches;      /* kernel launches issued by the call */
} swh_stats;

/* Where the lane scan reads the query profile from (swh_search_opts.profile_mode). */
#define SWH_PROFILE_AUTO 0
#define SWH_PROFILE_SINGLE 1   /* per-residue profile rows staged in LDS; a row of the
                                  database needs two of them and a byte permute */
#define SWH_PROFILE_PAIR 2     /* residue-pair table in LDS, built per strip
                                  (strip widths 16 and 24 only) */

/* Per-call tuning.  Zero-initialise for defaults. */
typedef struct swh_search_opts {
    int strip_width;          /* query columns held in registers per lane:
                                 0 = auto, else 16, 24 or 32 */
    int overflow_threshold;   /* 0 = exact default (31744: the int16 scan is exact
                                 below it); a lower value forces more int32
                                 rescoring */
    int group_lanes;          /* lanes cooperating on one long sequence pair
                                 (1..64); 0 = auto */
    int group_width;          /* query columns per lane in group mode: 0 = auto,
                                 else 8, 16 or 24 (24 with 24-column lane
                                 strips only) */
    int use_graph;            /* capture the launch sequence in a hipGraph and
                                 replay it while the query lengths, database and
                                 options stay the same */
    int profile_mode;         /* SWH_PROFILE_*; auto = tuning table, else pair
                                 table where the width allows */
} swh_search_opts;

/* Database layout options.  Zero-initialise for defaults. */
typedef struct swh_db_opts {
    uint32_t group_min_len;   /* sequences at least this long are stored as
                                 pairs for group mode; 0 = default (total
                                 residues / (2560 x CUs), at least 2048; below
                                 1024 sequences per CU a second layout with
                                 pairs from 128 residues serves launches of
                                 few queries), 0xffffffff = never */
    uint32_t queries_per_launch;   /* with the default layout: queries the caller
                                      will search per call, 0 = not known.  When
                                      known, a small database gets only the one
                                      layout those calls run on */
} swh_db_opts;

const char* swh_last_error(void);
int swh_version(void);
int swh_device_count(int* count);

/* Page-locked host memory: buffers handed to swh_db_create from it are
 * uploaded by DMA without a staging copy.  No reference counterpart (the
 * reference's CUDAWorker pins its chunk buffers, src/CUDAWorker.cu:53-98). */
int swh_pinned_alloc(size_t bytes, void** out);
int swh_pinned_free(void* p);

/* Context: device, scoring scheme, search stream and scratch.
 * Replaces CUDAWorker's constructor/alloc (src/CUDAWorker.cu:29-98) and the
 * Param scoring fields it reads (src/Param.h:59-117). */
int swh_ctx_create(int device, const int32_t* matrix, int gap_open, int gap_extend,
                   swh_ctx** out);
/* swh_ctx_destroy returns SWH_ERR_STATE while databases made on the context
 * still exist (swh_db_destroy them first).
 * While a server is attached to a context (swh_server_create), its loop
 * thread owns the context's stream and launch buffers: swh_ctx_destroy,
 * swh_ctx_set_scoring, swh_search, swh_search_batch, swh_search_noalloc,
 * swh_autotune and swh_set_tuning on that context, and swh_db_destroy of the
 * served database, return SWH_ERR_STATE until the server is destroyed. */
int swh_ctx_destroy(swh_ctx* ctx);
int swh_ctx_set_scoring(swh_ctx* ctx, const int32_t* matrix, int gap_open, int gap_extend);

/* Database: sort by length, interleave into the HBM layout and upload once.
 * Replaces CUDAWorker::packDB + asyncCopy (src/CUDAWorker.cu:144-223), which
 * re-packed and re-copied every 256 MB chunk.  residues holds every sequence
 * back to back; sequence i is residues[offsets[i] .. offsets[i]+lengths[i]). */
int swh_db_create(swh_ctx* ctx, const uint8_t* residues, const uint64_t* offsets,
                  const uint32_t* lengths, uint64_t nseq, const swh_db_opts* opts,
                  swh_db** out);
/* The database's device buffers stay with its context (at most 32 buffers,
 * 8 GiB) for the next swh_db_create on it and are freed by swh_ctx_destroy:
 * freeing device memory waits for every search running on the device, and
 * the worker pipeline builds one chunk's database while the previous chunk is
 * searched.  No search on db may still be running. */
int swh_db_destroy(swh_db* db);
/* nseq, total residues, device bytes the packed layout(s) use: the packed
 * residues and the block, order and length tables (the buffers holding them
 * may be larger: they are reused from destroyed databases). */
int swh_db_info(const swh_db* db, uint64_t* nseq, uint64_t* residues, uint64_t* device_bytes);

/* One query, one search: int16 scan + on-device int32 rescore; scores[i] for
 * every database sequence in caller order.  A page-locked scores buffer
 * (swh_pinned_alloc) is written by the device copy itself, with no host copy
 * after it.
 * Replaces kernelLaunch + bindQueryPrf (src/CUDAKernel.h:9-22,
 * src/CUDAKernel.cu:194-226) and the score copy-back of CUDAWorker::search
 * (src/CUDAWorker.cu:224-246). */
int swh_search(swh_ctx* ctx, swh_db* db, const uint8_t* query, int qlen,
               const swh_search_opts* opts, int32_t* scores, swh_stats* stats);

/* swh_search without allocation: the same scores for the same arguments,
 * with every device buffer of the call taken from a static scratch array of
 * 256 MiB per device (a __device__ array of libswh.so, which each GPU the
 * library runs on holds until the process ends) and its descriptor tables
 * staged through 1 MiB of page-locked memory that swh_ctx_create allocated.
 * The call itself makes no hipMalloc, hipFree or other allocating call.
 * Limits: query length <= 8192; the scratch must hold the database's score
 * vector and overflow list (8 bytes per sequence) plus the strip borders of
 * at least one workgroup; the work descriptors must fit 1 MiB.  Beyond them
 * the call fails with SWH_ERR_CAPACITY.  use_graph must be 0.  Calls on one
 * device share its array and are serialised by a process-wide lock per
 * device; the array keeps the last call's contents (every buffer the kernels
 * read before writing is cleared or written by the call itself).
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under graph capture, tight serving loops, harnesses
 * that forbid allocation). */
int swh_search_noalloc(swh_ctx* ctx, swh_db* db, const uint8_t* query, int qlen,
                       const swh_search_opts* opts, int32_t* scores, swh_stats* stats);

/* Many queries in one persistent launch with a device work queue; scores is
 * nq rows of nseq int32 in query order.  No reference counterpart: the
 * reference searches one query per process (src/main.cpp:7-16). */
int swh_search_batch(swh_ctx* ctx, swh_db* db, const uint8_t* const* queries,
                     const int* qlens, int nq, const swh_search_opts* opts,
                     int32_t* scores, swh_stats* stats);

/* Top-k of a score vector: (score desc, index asc), exactly k entries, padded
 * with (0, 0) -- the ResultList order (src/ResultList.h:37-79) made
 * deterministic for ties. */
int swh_topk(const int32_t* scores, uint64_t n, int k, int32_t* out_score,
             int64_t* out_index);

/* ---- Autotuning -------------------------------------------------------------
 * The launch shape is chosen per query-length bucket from measurements on a
 * length-stratified sample of the database: lanes per sequence (the database
 * length from which sequences are searched by a group of lanes, and the group
 * size per query), the query-profile placement and the register tile depth
 * (query columns per lane).  No reference counterpart: the reference's shapes
 * are compile-time constants (src/CUDAKernel.cu:152-192, src/AVXKernel.cpp). */
#define SWH_TUNE_MAX 16
/* Lane-scan configurations measured per bucket, in lane_cost order:
 * (16, pair) (24, pair) (16, single) (24, single) (32, single). */
#define SWH_TUNE_LANE_CONFIGS 5

/* One bucket: queries of length <= qlen_max (rows ascend; the last row also
 * takes longer queries). */
typedef struct swh_tune_entry {
    int qlen_max;
    int strip_width;          /* best lane-scan tile depth (16, 24, 32) */
    int profile_mode;         /* best profile placement (SWH_PROFILE_SINGLE / _PAIR) */
    int group_lanes;          /* best lanes per long sequence pair (1..64), or 0:
                                 chosen per launch within its pass budget */
    int group_width;          /* best group-mode tile depth (8, 16, 24), or 0: the
                                 built-in choice per launch (then group_lanes 0) */
    float lane_cost[SWH_TUNE_LANE_CONFIGS];   /* measured launch ms per lane column
                                                 (query length rounded up to 8, a
                                                 narrower last strip weighted by
                                                 its per-row work), per
                                                 configuration */
} swh_tune_entry;

typedef struct swh_tune_result {
    uint32_t group_min_len;   /* best swh_db_opts.group_min_len among those tried */
    int n_entries;
    swh_tune_entry entries[SWH_TUNE_MAX];
    double seconds;           /* wall time of the tuning */
    uint64_t sample_seqs, sample_residues;
} swh_tune_result;

/* Measure on every sample_every-th sequence in length order (0 = 8), with
 * queries_per_bucket random queries of each bucket length (0 = 4) and reps
 * timed repetitions (0 = 1), then install the table on ctx.  bucket_qlens
 * ascending, 1..SWH_TUNE_MAX of them. */
int swh_autotune(swh_ctx* ctx, const uint8_t* residues, const uint64_t* offsets,
                 const uint32_t* lengths, uint64_t nseq, int sample_every,
                 const int* bucket_qlens, int n_buckets, int queries_per_bucket, int reps,
                 swh_tune_result* out);
/* Install a table (n = 0 clears it: built-in shape choice). */
int swh_set_tuning(swh_ctx* ctx, const swh_tune_entry* entries, int n);
int swh_get_tuning(const swh_ctx* ctx, swh_tune_entry* entries, int cap, int* n);

/* ---- Serving: continuous batching on a resident database ------------------
 * A server owns a loop thread.  Requests are admitted, in arrival order, when
 * the device memory their query profile, score vector and overflow list need
 * fits the budget.  With the live queue (the default on a database of one
 * layout) ONE persistent scan launch stays on the GPU while there is work:
 * a request admitted while others run has its work items appended to that
 * launch's device queue at once, between theirs (shortest remaining work
 * first), and its scores are copied out as soon as its own items are done.
 * With rounds, each round runs one launch over the next database slice of
 * every admitted request, and an arriving request joins at the next round.
 * Scores are exact and independent of which requests share the device.  No
 * reference counterpart (the reference runs one query per process). */
typedef struct swh_server swh_server;

typedef struct swh_server_opts {
    int max_inflight;          /* requests on the device at once (0 = 64) */
    uint64_t memory_budget;    /* device bytes for request buffers (0 = 4 GiB) */
    int slices;                /* 0 = budgeted rounds: each round holds about
                                  round_ms of device work, requests served
                                  shortest remaining work first (with aging), a
                                  long request advancing one slice per round;
                                  n >= 1 = every request in n fixed slices, each a
                                  round (1: the whole database per round).  A
                                  small database (two layouts) always runs whole
                                  rounds */
    int max_qlen;              /* longest admissible query (0 = 8192) */
    int max_k;                 /* largest top-k a request may ask for (0 = 100) */
    swh_search_opts search;    /* strip widths, overflow threshold */
    double round_ms;           /* budgeted rounds: device work per round in ms
                                  (0 = 160) */
    double wait_weight;        /* live queue and budgeted rounds: a request's
                                  priority is its remaining work less
                                  wait_weight x the work the device could have
                                  done while it waited (0 = 0.5; larger favours
                                  older requests) */
    int queue;                 /* SWH_QUEUE_AUTO (0): the live queue on a
                                  database of one layout with slices = 0, rounds
                                  otherwise; SWH_QUEUE_LIVE: one persistent scan
                                  launch stays on the device while there is
                                  work and requests join it as they arrive
                                  (slices and round_ms unused; a small database
                                  is searched on its first layout);
                                  SWH_QUEUE_ROUNDS: one launch per round */
} swh_server_opts;

#define SWH_QUEUE_AUTO 0
#define SWH_QUEUE_LIVE 1
#define SWH_QUEUE_ROUNDS 2

typedef struct swh_request_info {
    double submit_ms, admit_ms, done_ms;   /* server clock: ms since creation */
    double latency_ms;                     /* done - submit */
    uint64_t cells;
    int rounds;                            /* rounds (live queue: launches) it took part in */
    int status;                            /* SWH_OK or the failure code */
} swh_request_info;

typedef struct swh_server_stats {
    uint64_t rounds, completed, cells;     /* rounds: rounds run (live queue: launches begun) */
    double busy_ms;                        /* wall time spent in rounds (launches) */
    int peak_inflight;
    int waiting;                           /* submitted, not yet admitted */
} swh_server_stats;

int swh_server_create(swh_ctx* ctx, swh_db* db, const swh_server_opts* opts, swh_server** out);
/* Copies the query; returns at once with the request id. */
int swh_server_submit(swh_server* srv, const uint8_t* query, int qlen, int k, int want_scores,
                      uint64_t* id);
/* Blocks until request id is done (SWH_OK) or timeout_ms passes
 * (SWH_ERR_STATE; timeout_ms < 0 waits forever).  On success writes k
 * (score, index) pairs, the full score vector if it was asked for at submit
 * (scores may then be NULL to discard it), and the timing; the request is
 * then forgotten. */
int swh_server_wait(swh_server* srv, uint64_t id, int timeout_ms, int32_t* top_score,
                    int64_t* top_index, int32_t* scores, swh_request_info* info);
int swh_server_get_stats(swh_server* srv, swh_server_stats* out);
/* Finishes the requests already admitted, fails the waiting ones with
 * SWH_ERR_STATE (a thread blocked in swh_server_wait on one of them returns
 * that status before destroy frees the server), stops, and detaches the
 * server from its context.  A request whose device buffers alone exceed
 * memory_budget is refused at submit (SWH_ERR_CAPACITY). */
int swh_server_destroy(swh_server* srv);

#ifdef __cplusplus
}
#endif
#endif
