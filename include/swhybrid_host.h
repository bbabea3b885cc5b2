/*
 * swhybrid_host.h -- C entry points of the host side (libswhybrid_host.so):
 * a DataPool (in memory, or over the reference's database files) feeds
 * HIPWorker threads (the Worker plugin that reaches libswh.so), a
 * RecalcWorker collects anything a worker could not score, and the
 * per-worker top-k lists are merged; plus the DBmaker and FASTA input.
 *
 * Replaces the reference's process entry (src/main.cpp:7-16: Param, DataPool,
 * WorkerPool constructed from argv, results printed by ~WorkerPool,
 * src/WorkerPool.cpp:84-115) with a call that returns the results.
 */
#ifndef SWHYBRID_HOST_H_
#define SWHYBRID_HOST_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* swhy_last_error(void);

/* Search one query through the worker pipeline.
 *   chunk_bytes    DataPool chunk budget per fill (0 = the default schedule
 *                  of swhy_pool_opts.chunk_bytes; the reference's CUDAWorker
 *                  uses 256 MB, src/CUDAWorker.cu:44)
 *   n_workers      HIPWorkers on devices 0..n_workers-1
 *   out_top_*      topk entries: score, caller index, pool index; ties are
 *                  ordered by pool index (the order sequences are handed out),
 *                  padded with (0, 0)
 *   out_scores     optional: every sequence's exact score, caller order
 * Returns 0 on success, negative with swhy_last_error() set otherwise. */
int swhy_pool_search(const uint8_t* residues, const uint64_t* offsets, const uint32_t* lengths,
                     uint64_t nseq, const uint8_t* query, int qlen, const int32_t* matrix,
                     int gap_open, int gap_extend, int topk, uint64_t chunk_bytes,
                     int n_workers, int32_t* out_top_score, int64_t* out_top_index,
                     int64_t* out_top_pool_index, int32_t* out_scores, double* out_seconds,
                     double* out_gcups);

/* The host's AVX2 scan alone, over a whole database, with a team of nthreads
 * threads (0 = every core): the CPU worker's kernel without the pipeline.
 * Sequences longest first, 32 per 256-bit vector in biased int8, the lanes
 * that may have saturated again 16 per vector in biased int16; a score at or
 * above 65536 - max(matrix) comes back as -1 (the reference's int16 pass is
 * exact only below that too).  query_segment: query columns per pass of the
 * scan (0 = default, -1 = whole query).  The database is first packed, 32
 * sequences per group interleaved row by row (out_pack_seconds, as the
 * reference's packDB prepares a chunk before its search), then scanned
 * (out_seconds).
 * Counterpart: avxCompute under the reference's OpenMP team with the
 * avxCompute16 rescore of its overflowed lanes (src/AVXKernel.cpp:408-554,
 * 866-1016; src/AVXWorker.cpp:397-427; src/RecalcWorker.cpp:184-257). */
int swhy_cpu_search(const uint8_t* residues, const uint64_t* offsets, const uint32_t* lengths,
                    uint64_t nseq, const uint8_t* query, int qlen, const int32_t* matrix,
                    int gap_open, int gap_extend, int nthreads, int query_segment,
                    int32_t* out_scores, double* out_seconds, double* out_pack_seconds);

/* ---- Hybrid pool: HIPWorkers and a CPU worker on one DataPool ----------------
 * The reference's hybrid scheduling (src/WorkerPool.cpp:25-78): every worker
 * pulls chunks from the same DataPool, so the faster device takes more of
 * them.  The CPU worker scans with the host's AVX2 units (int8, then int16
 * for overflowed lanes); sequences beyond int16 go to the RecalcWorker and
 * are rescored exactly on device 0. */
typedef struct swhy_pool_opts {
    int topk;
    int hip_workers;           /* HIPWorkers on devices 0..hip_workers-1 */
    int cpu_threads;           /* threads of the CPU worker; 0 = no CPU worker */
    uint64_t chunk_bytes;      /* HIPWorker chunk (0 = the default schedule: 512 MB
                                  below query length 600, 256 MB below 2,500, from
                                  there a 256 MB first chunk, then 1 GB) */
    uint64_t cpu_chunk_bytes;  /* CPU worker chunk (0 = 4 MB); the CPU worker takes no
                                  chunk once only hip_workers x chunk_bytes are left */
    uint64_t first_chunk_bytes; /* a HIPWorker's first chunk; each later one doubles, up
                                  to chunk_bytes (0 = chunk_bytes when that is set,
                                  else the default schedule) */
} swhy_pool_opts;

typedef struct swhy_pool_report {
    double seconds;            /* first worker start to last stop (recalc included) */
    double gcups;              /* padded residues x query length / seconds */
    uint64_t hip_residues, cpu_residues;   /* padded residues each kind searched */
    double hip_busy_s, cpu_busy_s;         /* time each kind spent computing */
    uint64_t recalc;           /* sequences rescored after the workers */
    double wall_seconds;       /* the whole call: worker set-up, the first chunk's
                                  fill and pack (outside `seconds`, as in the
                                  reference's timing), search, merge */
    /* the HIPWorkers' time, summed over them: */
    double dev_first_fill_s;   /* first DataPool fill + pack, before their loop */
    double dev_aux_fill_s;     /* later fills + packs, on their aux threads */
    double dev_pack_s;         /* of the fills, swh_db_create (upload, device packing) */
    double dev_search_s;       /* swh_search calls */
    double dev_aux_wait_s;     /* waits for the aux thread after a search */
    double dev_kernel_s;       /* scan + rescore kernels of the searches */
    int dev_chunks;            /* chunks searched */
    uint64_t chunk_bytes;      /* the HIPWorkers' chunk schedule that ran: */
    uint64_t first_chunk_bytes; /*   first chunk, doubling up to chunk_bytes */
} swhy_pool_report;

/* swhy_pool_search with a CPU worker beside the HIPWorkers.  out_top_* get
 * opts->topk entries, caller indices; out_scores (optional) every score. */
int swhy_pool_search_ex(const uint8_t* residues, const uint64_t* offsets,
                        const uint32_t* lengths, uint64_t nseq, const uint8_t* query, int qlen,
                        const int32_t* matrix, int gap_open, int gap_extend,
                        const swhy_pool_opts* opts, int32_t* out_top_score,
                        int64_t* out_top_index, int32_t* out_scores, swhy_pool_report* report);

/* ---- The reference's on-disk database -------------------------------------
 * <base>.seq / .map / .title as the reference's DBmaker writes them
 * (src/FastaFile.cpp:72-359) and its DataPool reads them
 * (src/DataPool.cpp:145-303).  A sequence's index is its position in .seq,
 * the all-pad fillers of the final batch included. */
typedef struct swhy_dbfile swhy_dbfile;

/* The DBmaker (src/DBmaker.cpp, src/FastaFile.cpp, src/Bucket.cpp): FASTA ->
 * <out_base>.seq/.map/.title in the same format, with the mixed-length final
 * batches padded per sequence (the reference pads them at shifted positions,
 * src/Bucket.cpp:123-130).  flush_bytes 0 = 256 MB. */
int swhy_dbfile_make(const char* fasta_path, const char* out_base, uint64_t flush_bytes,
                     uint64_t* out_records, uint64_t* out_slots);
/* Maps <base>.seq read-only and reads the map; replaces DataPool::run /
 * initRead (src/DataPool.cpp:79-160). */
int swhy_dbfile_open(const char* base, swhy_dbfile** out);
int swhy_dbfile_info(const swhy_dbfile* f, uint64_t* records, uint64_t* slots,
                     uint64_t* residues, uint64_t* batches);
/* Per index: residues (codes, valid while f is open) and length without the
 * trailing pad codes (0 for fillers): ready for swh_db_create. */
int swhy_dbfile_arrays(const swhy_dbfile* f, const uint8_t** residues, const uint64_t** offsets,
                       const uint32_t** lengths);
/* The '>' line of sequence index, NUL-terminated and cut to cap bytes; *len
 * gets its full length; "" for fillers.  Replaces DataPool::getTitle
 * (src/DataPool.cpp:289-310), without its dangling pointer. */
int swhy_dbfile_title(swhy_dbfile* f, uint64_t index, char* buf, uint64_t cap, uint64_t* len);
int swhy_dbfile_close(swhy_dbfile* f);
/* swhy_pool_search over the file's batches in file order, as the reference's
 * DataPool::fillChunk hands them out (src/DataPool.cpp:163-286); indices are
 * file indices. */
int swhy_dbfile_pool_search(swhy_dbfile* f, const uint8_t* query, int qlen,
                            const int32_t* matrix, int gap_open, int gap_extend, int topk,
                            uint64_t chunk_bytes, int n_workers, int32_t* out_top_score,
                            int64_t* out_top_index, int32_t* out_scores, double* out_seconds,
                            double* out_gcups);
/* The same with a CPU worker beside the HIPWorkers (swhy_pool_search_ex). */
int swhy_dbfile_pool_search_ex(swhy_dbfile* f, const uint8_t* query, int qlen,
                               const int32_t* matrix, int gap_open, int gap_extend,
                               const swhy_pool_opts* opts, int32_t* out_top_score,
                               int64_t* out_top_index, int32_t* out_scores,
                               swhy_pool_report* report);

/* ---- FASTA queries ------------------------------------------------------------
 * Every record of a FASTA file (the reference reads the first only,
 * src/DataPool.cpp:312-344), letters translated with its table: 0..22,
 * '@' = 23, other letters 0xFF (which the search rejects in a query). */
typedef struct swhy_fasta swhy_fasta;
int swhy_fasta_open(const char* path, swhy_fasta** out);
int swhy_fasta_count(const swhy_fasta* f, uint64_t* n);
int swhy_fasta_record(const swhy_fasta* f, uint64_t i, const uint8_t** codes, uint64_t* len,
                      const char** title);
int swhy_fasta_close(swhy_fasta* f);

#ifdef __cplusplus
}
#endif
#endif
