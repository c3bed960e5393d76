/*
 * CPU oracle for MirAge's read-classification path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++ restatement of the
 * reference's CPU algorithm (mirdesign/mirage, MirAge/Hashmap.cpp,
 * MirAge/Analyser.cpp, MirAge/MatrixCPU.cpp, MirAge/MatrixResult.cpp).  Only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it,
 * and only as the checker; the product (mirage_amd/) never links or calls it.
 *
 * Every function cites the reference lines it restates.  Pins for each part
 * are listed at the top of mirage_oracle.cpp.
 */
#ifndef MIRAGE_ORACLE_H
#define MIRAGE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Index-build flags. */
#define MO_FLAG_LAST_BUCKET_EMPTY 1u  /* reproduce the .mhm loader leaving the
                                          highest-hash bucket with size 0 */

/* Classification flags. */
#define MO_FLAG_ULTRAFAST 1u           /* UltraFast speed mode: no extension */

typedef struct mo_index mo_index;

/* One scored candidate (a reference that received seeds). */
typedef struct {
  uint32_t read;     /* read index in the batch */
  uint32_t strand;   /* 0 forward, 1 reverse complement */
  uint32_t ref;      /* reference id */
  uint32_t hits;     /* popcount of the row-hit bitmap (H) */
  uint64_t sum_r;    /* sum of tot*(tot+1)/2 over seeds with tot >= T */
  uint64_t sum_c;    /* sum of tot over seeds with tot >= T */
  float score;       /* MatrixCPU Score, -2 when no seed reached T */
  float rel;         /* relative score, -1 when score <= -1 */
  uint32_t n_seeds;  /* number of seeds (chains) for this reference */
  uint32_t pad;
} mo_cand;

/* Per-read result after the strand merge. */
typedef struct {
  float highest;     /* HighestScore of the surviving strand (-1: none) */
  uint32_t strand;   /* strand whose result survived */
  uint32_t n_top;    /* size of the top-tie set (0 when not reported) */
  uint32_t n_cand;   /* candidate references of the surviving strand */
} mo_read_result;

uint32_t mo_hash32(const uint8_t *p, int32_t len);
uint32_t mo_hash28(const uint8_t *p, int32_t len);

/* Build the index over concatenated reference bytes; ref_off has n_refs+1
 * offsets.  Returns NULL on bad arguments. */
mo_index *mo_index_build(const uint8_t *ref_bytes, const uint64_t *ref_off,
                         uint32_t n_refs, int32_t threshold, uint32_t flags);
/* The index restricted to the buckets that the reads' probe windows (both
 * strands) hash to: those buckets are exactly mo_index_build's, every other
 * hash is absent, so classifying these reads gives the full index's results.
 * For large references, where the full index does not fit host memory.
 * T < 20 (the fragment rule needs every bucket's size) returns NULL. */
mo_index *mo_index_build_for_reads(const uint8_t *ref_bytes, const uint64_t *ref_off,
                                   uint32_t n_refs, int32_t threshold, uint32_t flags,
                                   const uint8_t *reads, const uint64_t *read_off,
                                   uint32_t n_reads);
void mo_index_free(mo_index *idx);
uint64_t mo_index_num_entries(const mo_index *idx);
uint64_t mo_index_num_buckets(const mo_index *idx);
/* Bucket lookup: returns -1 when the hash is absent, else the number of
 * entries (which may be 0 for a present-but-empty bucket).  Up to cap entries
 * are copied, sorted by (ref, pos). */
int64_t mo_index_bucket(const mo_index *idx, uint32_t hash, uint32_t *refs,
                        uint32_t *pos, uint64_t cap);
/* Per-reference window cutoff (first window whose hash equals H("N"*k)). */
uint64_t mo_index_ref_cutoff(const mo_index *idx, uint32_t ref);

/* Full path for n_reads reads: both strands, candidate generation, scoring,
 * best hit, strand merge.  top_refs receives up to top_cap ids per read
 * (sorted ascending) at top_refs[r*top_cap].  cands (may be NULL) receives
 * every seeded candidate of both strands in (read, strand, ref) order; the
 * total count is stored in *n_cands even when it exceeds cand_cap.
 * n_threads <= 0 uses all hardware threads.  Returns 0 on success. */
int32_t mo_classify(const mo_index *idx, const uint8_t *reads,
                    const uint64_t *read_off, uint32_t n_reads,
                    int32_t threshold, uint32_t flags, int32_t n_threads,
                    mo_read_result *out, uint32_t *top_refs, uint32_t top_cap,
                    mo_cand *cands, uint64_t cand_cap, uint64_t *n_cands);

/* Seed scoring only (MatrixCPU::CalculateScore, MatrixCPU.cpp:18-214): one
 * query against one reference with explicit seeds.  seeds holds n_seeds
 * records of 4 uint32: start.ref, start.query, last.ref, last.query. */
float mo_score_seeds(const uint8_t *query, uint64_t qlen, const uint8_t *ref,
                     uint64_t ref_len, const uint32_t *seeds, uint64_t n_seeds,
                     int32_t threshold, uint32_t flags, uint64_t *sum_r,
                     uint64_t *sum_c, uint32_t *hits);

/* MatrixResult::GetRelativeScore (MatrixResult.cpp:21-29). */
float mo_relative_score(float query_size, float ref_size, float score);

/* Work counters for the same inputs (probes, index bytes touched, extension
 * byte comparisons), used to state the algorithmic bytes of a workload. */
typedef struct {
  uint64_t probes;        /* hash lookups */
  uint64_t hits;          /* lookups that found a bucket */
  uint64_t entries;       /* bucket entries read */
  uint64_t ext_compares;  /* base comparisons in head/tail extension */
  uint64_t seeds;         /* seeds scored */
  uint64_t cands;         /* seeded candidates */
} mo_work;
int32_t mo_count_work(const mo_index *idx, const uint8_t *reads,
                      const uint64_t *read_off, uint32_t n_reads,
                      int32_t threshold, uint32_t flags, int32_t n_threads,
                      mo_work *work);

#ifdef __cplusplus
}
#endif
#endif
