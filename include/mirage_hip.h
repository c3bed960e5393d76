/*
 * mirage_hip.h - C ABI of the MI355X (gfx950) read-classification library.
 *
 * Drop-in for MirAge's matrix path (mirdesign/mirage): the reference's
 * Analyser calls
 *   vector<MatrixResult> CUDAGetHighestScore(...)      MirAge/Analyser.h:228,
 *                                                      MirAge/MatrixGPUCU.cu:571
 *   void   InitialiseCuda(int numDevices)              MirAge/Analyser.h:231
 *   size_t Get_LOCAL_DATA_rowhit_SIZE()                MirAge/Analyser.h:232
 *   size_t CUDAGetMaxBlockSize(int)                    MirAge/Hardware.h:16
 * and its CPU sibling MatrixCPU::CPUGetHighestScore (MirAge/MatrixCPU.h:41).
 * The C++ shim in mirage_amd/csrc/matrix_hip.hpp keeps that call surface and
 * returns vector<MatrixResult>; this header is the plain-pointer layer below
 * it.  Unlike the reference GPU form, the whole path runs on the device:
 * k-mer probe, candidate filtering, diagonal scoring, best hit, strand merge.
 *
 * Conventions: every entry point returns mh_status (0 = success, negative =
 * error; no exceptions cross this boundary, cf. MirAge/MatrixGPUCU.cu:490-530
 * which throws).  "d_" arguments are device pointers, "h_" host pointers.
 * A handle owns one HIP stream; calls on one handle are serialised by the
 * caller (as mutexCUDARunning does, MirAge/Analyser.cpp:718,989); different
 * handles may run concurrently.
 */
#ifndef MIRAGE_HIP_H
#define MIRAGE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t mh_status;
#define MH_OK 0
#define MH_ERR_ARG (-1)        /* bad argument */
#define MH_ERR_HIP (-2)        /* HIP runtime error (message via mh_last_error) */
#define MH_ERR_NOMEM (-3)      /* device allocation failed (caller may shrink the batch) */
#define MH_ERR_RANGE (-4)      /* read longer than the handle's maximum */
#define MH_ERR_OVERFLOW (-5)   /* an output pool was too small; *_needed says how big */
#define MH_ERR_BADBASE (-6)    /* a read holds a byte with no complement */
#define MH_ERR_STATE (-7)      /* wrong order of calls */

typedef struct mh_handle mh_handle;
typedef struct mh_index mh_index;

/* Index-build flags. */
#define MH_INDEX_LAST_BUCKET_EMPTY 1u /* keep the reference loader's behaviour:
                                         the highest-hash bucket is present but
                                         empty (MirAge/Hashmap.cpp:540-554) */

/* Classification flags. */
#define MH_CLASSIFY_ULTRAFAST 1u   /* no head/tail extension (MatrixCPU.cpp:69) */
#define MH_CLASSIFY_NO_DIRECT 2u   /* disable the direct-hit fast path */
#define MH_CLASSIFY_BATCHED 4u     /* length-bucketed persistent kernels */
#define MH_CLASSIFY_TIMING 8u      /* record kernel timings (mh_get_timing) */
#define MH_CLASSIFY_GRAPH 16u      /* batched form replayed from a captured hipGraph;
                                      captured again when any argument changes */
#define MH_CLASSIFY_PREPROBE 32u   /* batched form: hash and look up the probes of the
                                      shortest length tile in a separate pass, whose
                                      records the LDS overflow tiers also read */

/* Per-read result after the strand merge (MirAge/Analyser.cpp:592-611).
 * Top-tie reference ids (ascending) are top_pool[top_off, top_off+n_top).
 * n_top is 0 when the reference would print nothing for the read
 * (MirAge/Analyser.cpp:388). */
typedef struct {
  float highest;     /* HighestScore of the surviving strand, -1 if none */
  uint32_t strand;   /* 0 forward, 1 reverse complement */
  uint32_t n_top;
  uint32_t top_off;
  uint32_t n_cand;   /* candidates with Score > -1 on the surviving strand */
  uint32_t status;   /* 0, or MH_READ_* bits */
} mh_read_result;

#define MH_READ_BADBASE 1u
#define MH_READ_TOO_LONG 2u
#define MH_READ_CAPACITY 4u  /* more bucket entries than the largest workspace */

/* One seeded candidate (detail output for parity checking). */
typedef struct {
  uint32_t read;
  uint32_t strand;
  uint32_t ref;
  uint32_t hits;     /* popcount of the row-hit bitmap */
  uint64_t sum_r;    /* diagonalRepeatedTotalScore (MatrixCPU.cpp:183) */
  uint64_t sum_c;    /* diagonalTotalConcidered (MatrixCPU.cpp:181) */
  float score;       /* MatrixCPU Score, -2 when no seed reached T */
  float rel;         /* MatrixResult::GetRelativeScore, -1 when score <= -1 */
  uint32_t n_seeds;
  uint32_t pad;
} mh_cand;

typedef struct {
  uint64_t n_entries;      /* (ref, pos) entries in HBM */
  uint64_t n_buckets;      /* non-empty buckets */
  uint64_t device_bytes;   /* HBM held by the index */
  uint32_t n_refs;
  int32_t threshold;       /* T; stride T/2, k-mer length T - T/2 + 1 */
  uint32_t empty_hash;     /* present-but-empty bucket, 0xFFFFFFFF if none */
  uint32_t pad;
} mh_index_info;

/* Work counters of the last classify call on a handle. */
typedef struct {
  uint64_t queries;        /* read strands */
  uint64_t probes;         /* k-mer lookups */
  uint64_t hits;           /* lookups that found a bucket */
  uint64_t entries;        /* bucket entries gathered */
  uint64_t seeds;          /* seeds scored */
  uint64_t cands;          /* seeded candidates */
  uint64_t direct;         /* strands resolved by the direct-hit fast path */
  uint64_t overflow;       /* tier escalations: each pass of a strand from its length
                              tile to the overflow tiers, and from one overflow tier to
                              the next (second chance, middle, HBM workspaces, last
                              resort) */
  uint64_t ext_compares;   /* base comparisons of head/tail extension */
} mh_counters;

/* Kernel timings of the last call made with MH_CLASSIFY_TIMING. */
typedef struct {
  float total_ms;          /* first to last kernel of the call */
  float bucket_ms[8];      /* LDS kernel per length bucket (plain form: [0]) */
  float big_ms;            /* overflow tiers: second chance, middle, HBM workspaces */
  float merge_ms;          /* strand merge */
  uint32_t n_buckets;
  float probe_ms;          /* probe pre-pass (MH_CLASSIFY_PREPROBE), else 0 */
  float tier_ms[4];        /* big_ms per overflow tier: second chance, middle, HBM
                              workspaces, last resort */
} mh_timing;

/* ---- handle ------------------------------------------------------------ */

/* Opens device `device` (replaces InitialiseCuda, MatrixGPUCU.cu:481-488).
 * max_read_len bounds read length (replaces Get_LOCAL_DATA_rowhit_SIZE,
 * MatrixGPUCU.cu:39-42); 0 selects 32768. */
mh_status mh_open(int32_t device, uint32_t max_read_len, mh_handle **out);
mh_status mh_close(mh_handle *h);
const char *mh_last_error(const mh_handle *h);
/* The handle's HIP stream (a hipStream_t). */
void *mh_stream(mh_handle *h);
/* Largest read length the handle accepts (Get_LOCAL_DATA_rowhit_SIZE). */
uint32_t mh_max_read_len(const mh_handle *h);
/* Preferred reads per batch for the device (CUDAGetMaxBlockSize,
 * MatrixGPUCU.cu:549-556). */
uint64_t mh_max_batch_reads(const mh_handle *h);

/* ---- index ------------------------------------------------------------- */

/* Builds the bucketed hash index on the device from host reference bytes
 * (ref_off: n_refs+1 byte offsets).  Bucket contents follow
 * Hashmap::TBuildHashmapForReference (MirAge/Hashmap.cpp:2771-2976) and
 * SaveHashmapTableV3's fragment rule (:1759-1833). */
mh_status mh_index_build(mh_handle *h, const uint8_t *h_ref_bytes, const uint64_t *h_ref_off,
                         uint32_t n_refs, int32_t threshold, uint32_t flags, mh_index **out);
mh_status mh_index_free(mh_index *idx);
mh_status mh_index_get_info(const mh_index *idx, mh_index_info *out);
/* Copies the entries of bucket `hash` (sorted by ref, pos) to host arrays.
 * *count receives the bucket size, or -1 when the hash is absent. */
mh_status mh_index_bucket(mh_handle *h, const mh_index *idx, uint32_t hash, uint32_t *h_refs,
                          uint32_t *h_pos, uint64_t cap, int64_t *count);

/* ---- .mhm hashmap files (version 4) -------------------------------------- */

/* An opened .mhm file, memory-mapped on the host.  Replaces the reference's
 * Hashmap::ReadHashmapFile / LoadHashMapFromHashmapDB
 * (MirAge/Hashmap.cpp:329-580) and the bucket decode of
 * Analyser.cpp:1596-1720.  The functions below that take no handle run on
 * the host only; their error text is mh_last_error(NULL) (per thread). */
typedef struct mh_mhm mh_mhm;

typedef struct {
  uint64_t n_buckets;       /* TOC entries (Hashmap.cpp:428-430) */
  uint64_t file_bytes;
  uint64_t record_bytes;    /* bucket records: first TOC position to end of file */
  uint64_t ref_bytes;       /* reference sequence bytes */
  uint64_t header_bytes;    /* reference header bytes */
  uint32_t n_refs;
  int32_t threshold;        /* T the file was built with (Hashmap.cpp:456) */
  uint32_t hash_bytes;      /* HASH_Supported_Bytes: 4 (32-bit build) or 8 */
  uint32_t streampos_bytes; /* sizeof(HASH_STREAMPOS) of the writing build */
} mh_mhm_info;

/* Opens and validates the file's header and tables of contents (the
 * indicator string, version 4, hash and position widths, ascending 28-bit
 * hashes, in-file extents).  A file that fails a check returns MH_ERR_ARG. */
mh_status mh_mhm_open(const char *path, mh_mhm **out);
mh_status mh_mhm_close(mh_mhm *m);
mh_status mh_mhm_get_info(const mh_mhm *m, mh_mhm_info *out);
/* Copies the reference sequences as stored (h_off: n_refs+1 offsets), the
 * headers as stored ('>' name NUL, Hashmap.cpp:2176-2240) and the taxonomy
 * ids (Hashmap.cpp:552-558).  Any output may be NULL. */
mh_status mh_mhm_references(const mh_mhm *m, uint8_t *h_bytes, uint64_t *h_off,
                            uint8_t *h_headers, uint64_t *h_header_off, int32_t *h_taxids);
/* Decodes the record of bucket `hash` on the host, entries in file order
 * (ascending position).  *count receives its size, or -1 when the hash is
 * absent; with MH_INDEX_LAST_BUCKET_EMPTY the last TOC entry decodes empty,
 * as the reference loader leaves it (Hashmap.cpp:525-538). */
mh_status mh_mhm_bucket(const mh_mhm *m, uint32_t hash, uint32_t flags, uint32_t *h_refs,
                        uint32_t *h_pos, uint64_t cap, int64_t *count);
/* Builds the device index from an opened file: the records are copied to
 * HBM as stored and decoded there.  threshold 0 takes the file's; a smaller
 * one is refused as the reference refuses it (Hashmap.cpp:457-463), a
 * larger one is used for the queries (:464-469).  Every decoded entry is
 * checked (reference id below n_refs, position inside its reference); a
 * record that fails returns MH_ERR_ARG naming the bucket. */
mh_status mh_index_from_mhm(mh_handle *h, const mh_mhm *m, int32_t threshold, uint32_t flags,
                            mh_index **out);
/* Writes the index as a .mhm version 4 file, as SaveHashmapTableV3 does
 * (MirAge/Hashmap.cpp:1701-2540): buckets in ascending hash order, each
 * sorted by position (ties by reference id), then the sequences (uppercased,
 * as indexed), the headers and the taxonomy ids.  h_headers/h_header_off:
 * per reference the bytes MirAge stores ('>' name NUL); NULL writes '>' id
 * NUL.  h_taxids NULL writes -1.  hash_bytes and streampos_bytes are 4 or 8
 * (0 selects 4: a 32-bit build).  The index must hold every bucket (built
 * without MH_INDEX_LAST_BUCKET_EMPTY), else MH_ERR_STATE. */
mh_status mh_index_save_mhm(mh_handle *h, const mh_index *idx, const char *path,
                            const uint8_t *h_headers, const uint64_t *h_header_off,
                            const int32_t *h_taxids, uint32_t hash_bytes,
                            uint32_t streampos_bytes);

/* ---- classification ---------------------------------------------------- */

/* Most reads one classify call takes, plus one: a strand's id (2 * read +
 * strand) and a flag share 32 bits.  Larger calls return MH_ERR_ARG. */
#define MH_MAX_CALL_READS 1073741824u /* 2^30 */

/* Full path on device-resident reads: d_reads bytes, d_read_off n_reads+1
 * offsets (n_reads below MH_MAX_CALL_READS).  Results go to d_out[n_reads]
 * and the top-tie pool d_top_pool
 * (capacity top_cap ids).  d_cands (may be NULL) receives every seeded
 * candidate of both strands, in no particular order.  On MH_ERR_OVERFLOW,
 * *top_needed / *cand_needed hold the sizes required.  Asynchronous on the
 * handle's stream unless h_n_top/h_n_cand are non-NULL, which makes the call
 * wait for completion and report the pool fill. */
mh_status mh_classify(mh_handle *h, const mh_index *idx, const uint8_t *d_reads,
                      const uint64_t *d_read_off, uint32_t n_reads, uint32_t flags,
                      mh_read_result *d_out, uint32_t *d_top_pool, uint64_t top_cap,
                      mh_cand *d_cands, uint64_t cand_cap, uint64_t *top_needed,
                      uint64_t *cand_needed);

/* Same path from host memory: uploads reads, classifies, downloads.
 * h_top_pool must hold top_cap ids; h_cands may be NULL. */
mh_status mh_classify_host(mh_handle *h, const mh_index *idx, const uint8_t *h_reads,
                           const uint64_t *h_read_off, uint32_t n_reads, uint32_t flags,
                           mh_read_result *h_out, uint32_t *h_top_pool, uint64_t top_cap,
                           mh_cand *h_cands, uint64_t cand_cap, uint64_t *top_needed,
                           uint64_t *cand_needed);

/* mh_classify without allocation.  No counterpart in the reference: this
 * entry point is for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments, results and errors are mh_classify's; instead of the handle's
 * growing scratch arena the call uses a fixed device array of
 * MH_NOALLOC_SCRATCH_BYTES that the library holds on each device it runs on
 * until the process ends.  At most MH_NOALLOC_MAX_READS reads per call;
 * more, or a batch whose scratch does not fit (long maximum read lengths,
 * MH_CLASSIFY_PREPROBE), returns MH_ERR_RANGE.  The HBM overflow tiers run
 * on as many workspaces as fit (fewer than mh_classify uses: slower, same
 * results), and tie sets larger than the fixed tie pool return
 * MH_ERR_OVERFLOW rather than growing it.  MH_CLASSIFY_GRAPH is refused.
 * Every call on a device shares that device's array: calls that overlap on
 * one device (from different handles or threads) race, so serialise them.
 * The array keeps the last call's contents; the call clears what must start
 * at zero itself. */
#define MH_NOALLOC_SCRATCH_BYTES 268435456ull /* 256 MiB */
#define MH_NOALLOC_MAX_READS 1048576u
mh_status mh_classify_noalloc(mh_handle *h, const mh_index *idx, const uint8_t *d_reads,
                              const uint64_t *d_read_off, uint32_t n_reads, uint32_t flags,
                              mh_read_result *d_out, uint32_t *d_top_pool, uint64_t top_cap,
                              mh_cand *d_cands, uint64_t cand_cap, uint64_t *top_needed,
                              uint64_t *cand_needed);

/* Counters of the last mh_classify / mh_classify_host call (synchronises). */
mh_status mh_get_counters(mh_handle *h, mh_counters *out);
/* The same counters per kernel of the last call (synchronises): rows 0-7
 * the length tiles in table order (the plain form: row 0), rows 8-11 the
 * overflow tiers (second chance, middle, HBM workspaces, last resort).  In
 * each row `queries` counts the strands that kernel finished and the work
 * fields their work; `overflow` counts, for a tier row, the strands that
 * entered the tier.  *n receives the number of rows (12); up to cap are
 * written. */
mh_status mh_get_kernel_counters(mh_handle *h, mh_counters *out, uint32_t cap, uint32_t *n);
/* Kernel timings of the last call made with MH_CLASSIFY_TIMING (synchronises). */
mh_status mh_get_timing(mh_handle *h, mh_timing *out);

/* Scoring only, the role CUDAGetHighestScore plays in the reference
 * (MatrixGPUCU.cu:108-467 with host :571-1969): score given seeds.
 * Pairs p = 0..n_pairs-1 score query pair_query[p] (bytes
 * d_queries[d_query_off[q], d_query_off[q+1])) against reference
 * pair_ref[p] of idx, with seeds d_seeds[4*d_seed_off[p] .. 4*d_seed_off[p+1])
 * as (start.ref, start.query, last.ref, last.query) - the
 * CombiPositions / CombiPositions_LastPositionInRange pairs of
 * Analyser.h:97-98.  Outputs per pair: Score (MatrixCPU.cpp:195-213),
 * relative score (MatrixResult.cpp:21-29, -1 when Score <= -1), and the
 * integer triple. */
mh_status mh_score_seeds(mh_handle *h, const mh_index *idx, const uint8_t *d_queries,
                         const uint64_t *d_query_off, const uint32_t *d_pair_query,
                         const uint32_t *d_pair_ref, const uint64_t *d_seed_off,
                         const uint32_t *d_seeds, uint32_t n_pairs, uint32_t flags,
                         float *d_score, float *d_rel, uint64_t *d_sum_r, uint64_t *d_sum_c,
                         uint32_t *d_hits);

/* ---- FASTA ingest ---------------------------------------------------------- */

/* Parses FASTA text already in device memory into sequences and headers on
 * the device, as Sequences::ReadFastaSequencesFromMem reads a file
 * (MirAge/Sequences.cpp:57-211, :283-562): reading starts at the first line
 * beginning with '>'; headers are kept as MirAge keeps them ('>' name NUL,
 * a final '\r' dropped) in d_hdr with n+1 offsets d_hdr_off; sequence
 * lines, without a final '\r', are concatenated per record and uppercased
 * (:40-52) into d_seq with n+1 offsets d_seq_off, ready for mh_classify.
 * Sizes go to *n_records, *seq_bytes, *hdr_bytes; with any output NULL only
 * the sizes are computed (seq_bytes <= n_bytes; hdr_bytes <= n_bytes + 1).
 * A text with an empty line, a line starting with '@', a header without a
 * sequence line or a NUL byte returns MH_ERR_ARG naming the line (the
 * reference's reading of those is undefined).  Synchronises the handle's
 * stream. */
mh_status mh_fasta_parse(mh_handle *h, const uint8_t *d_text, uint64_t n_bytes, uint8_t *d_seq,
                         uint64_t seq_cap, uint64_t *d_seq_off, uint8_t *d_hdr, uint64_t hdr_cap,
                         uint64_t *d_hdr_off, uint64_t rec_cap, uint64_t *n_records,
                         uint64_t *seq_bytes, uint64_t *hdr_bytes);

/* ---- result logs ---------------------------------------------------------- */

/* Score thresholds of the filtered logs (the LogThreshold_* settings,
 * MirAge/LogManager.cpp:10-12; NULL selects 0.4, 1, 0.7). */
typedef struct {
  float high_sensitivity;  /* results_HighSensitivity.log */
  float high_precision;    /* results_HighPrecision.log */
  float balanced;          /* results_Balanced.log */
} mh_log_thresholds;

#define MH_RESULTS_APPEND 1u      /* append to the files, no header line (a later batch) */
#define MH_RESULTS_ALL_GROUPS 2u  /* a line per score group (output_firstresults_only: 0) */

/* Writes the reference's four result logs (results_Raw.log,
 * results_HighSensitivity.log, results_HighPrecision.log, results_Balanced.log)
 * into directory `dir`, replacing Analyser::LogMatrixResults /
 * LogResultsToFileManager (MirAge/Analyser.cpp:226-491) and LogManager's
 * files (MirAge/LogManager.cpp:294-404).  Host only.  Read r of the batch
 * has InputID first_read_id + r; h_results / h_top_pool are a classify
 * call's outputs; h_cands (may be NULL) its candidate output, which fixes
 * the order of tied references exactly (without it, the top ties stand in
 * for the scored candidates).  Headers are as MirAge keeps them ('>' name
 * NUL; the last byte is not printed); h_taxids NULL means -1 throughout. */
mh_status mh_write_results(const char *dir, uint64_t first_read_id, uint32_t n_reads,
                           const uint8_t *h_read_headers, const uint64_t *h_read_header_off,
                           const mh_read_result *h_results, const uint32_t *h_top_pool,
                           const mh_cand *h_cands, uint64_t n_cands, uint32_t n_refs,
                           const uint8_t *h_ref_headers, const uint64_t *h_ref_header_off,
                           const int32_t *h_taxids, const mh_log_thresholds *thresholds,
                           uint32_t flags);

/* ---- batched form and tuning -------------------------------------------- */

/* Tile choice for the length-bucketed persistent form. */
typedef struct {
  uint32_t max_len;        /* bucket holds reads with length <= max_len */
  uint32_t block;          /* threads per query: 8, 16 or 32 (lane teams, 64 / block
                              queries per one-wavefront workgroup) or 64..1024
                              (a workgroup per query), a power of two */
  uint32_t entries_cap;    /* LDS capacity in bucket entries (length tiles: a power of
                              two >= 32; overflow tiers: a multiple of 16 >= 32) */
  uint32_t grid;           /* persistent workgroups (0: derive from occupancy) */
} mh_tile;

/* Installs the bucket table used by MH_CLASSIFY_BATCHED (at most 8 rows,
 * ascending max_len; the last row must cover mh_max_read_len).  Without one
 * (or after mh_set_tiles(h, NULL, 0)) each call uses defaults sized for its
 * index: the entry capacities follow the index's entries per bucket. */
mh_status mh_set_tiles(mh_handle *h, const mh_tile *tiles, uint32_t n_tiles);
/* The length tiles (and, when tiers is not NULL, the two LDS overflow
 * tiers) that a batched call on idx uses. */
mh_status mh_get_tiles(const mh_handle *h, const mh_index *idx, mh_tile *tiles, uint32_t cap,
                       uint32_t *n_tiles, mh_tile *tiers);
/* The two LDS overflow tiers that strands too large for their length tile
 * pass through before the HBM-workspace tiers: tier 0 (second chance) holds
 * reads up to tiers[0].max_len with tiers[0].entries_cap entries (a strand
 * with more entries skips both LDS tiers), tier 1 the longer reads.  Block
 * 64..1024, entries_cap a multiple of 16 >= 32, max_len 64..max_read_len.
 * NULL restores the defaults: sized from the first length tile, tier 0's
 * capacity then lowered (by at most a quarter) to the largest at which one
 * more of its workgroups is resident per CU.  Results do not depend on
 * them. */
mh_status mh_set_overflow_tiles(mh_handle *h, const mh_tile *tiers);
/* Times candidate tiles on the given device-resident reads (at most the
 * first 2^20) and installs the fastest (the project's autotuner): per length
 * bucket that holds strands, a local search from the current tile over half
 * and double its team size and entry capacity, timed by the bucket's kernel
 * plus the overflow tiers; then the two LDS overflow tiers the same way.
 * iters: calls per candidate (0: 2), the fastest counts. */
mh_status mh_autotune(mh_handle *h, const mh_index *idx, const uint8_t *d_reads,
                      const uint64_t *d_read_off, uint32_t n_reads, uint32_t iters);

/* ---- serving ------------------------------------------------------------ */

typedef struct mh_server mh_server;

typedef struct {
  uint64_t hbm_budget;      /* bytes of HBM for requests in flight (0: 8 GiB) */
  uint32_t max_inflight;    /* batches on the device at once (0: 3) */
  uint32_t max_batch_reads; /* reads per coalesced batch (0: 262144) */
  uint32_t flags;           /* classification flags for every batch */
  uint32_t pad;
} mh_server_config;

/* Completed request. */
typedef struct {
  uint64_t request_id;
  uint32_t n_reads;
  int32_t status;
  double submit_s;          /* monotonic seconds */
  double admit_s;
  double done_s;
} mh_request_info;

/* Starts a serving loop over a resident index: requests are admitted as HBM
 * budget frees up, coalesced into batches, and several batches run on the
 * device at once (continuous batching). */
mh_status mh_server_start(mh_handle *h, const mh_index *idx, const mh_server_config *cfg,
                          mh_server **out);
/* Submits a request (host reads are copied; the call does not block on the
 * device).  *request_id receives its id. */
mh_status mh_server_submit(mh_server *s, const uint8_t *h_reads, const uint64_t *h_read_off,
                           uint32_t n_reads, uint64_t *request_id);
/* Waits for request_id (timeout_s < 0: forever) and copies its results:
 * h_out[n_reads], top-tie ids to h_top (capacity top_cap; *top_needed
 * receives the count, MH_ERR_OVERFLOW when larger). */
mh_status mh_server_wait(mh_server *s, uint64_t request_id, double timeout_s,
                         mh_read_result *h_out, uint32_t *h_top, uint64_t top_cap,
                         uint64_t *top_needed, mh_request_info *info);
mh_status mh_server_stop(mh_server *s);

#ifdef __cplusplus
}
#endif
#endif /* MIRAGE_HIP_H */
