/*
 * CPU oracle for the BWA-MEM hot path of sunwookim028/G3SA.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C, the
 * reference's GPU pipeline (bwa/src/bwamem.cu:92-425 and the kernels it
 * launches) so that the HIP implementation can be checked against it.  Only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
 * load it.  It is never part of the product path.
 *
 * Where the reference's behaviour is undefined (data races, uninitialised
 * reads), the oracle implements a documented resolution; each is marked
 * "DR" at the point of use.
 *
 * Parity unpinned: orc_hash_64 (no known-answer vectors; see its
 * declaration).
 */
#ifndef G3A_ORACLE_H
#define G3A_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_index orc_index;

/* BWA default options (bwa/src/fastmap.cpp:77-113). */
typedef struct {
  int a, b, o_del, e_del, o_ins, e_ins, w, T, pen_clip5, pen_clip3;
  int min_seed_len, split_width, max_occ, max_chain_gap, flag;
  uint64_t max_mem_intv;
  float split_factor, mask_level, drop_ratio, mask_level_redun;
  int mapQ_coef_len, mapQ_coef_fac;
  int8_t mat[25];
} orc_opt;

typedef struct {
  int64_t k, l, s;   /* SA interval [k, k+s), reverse-complement start l */
  int32_t m, n;      /* read[m..n] inclusive */
} orc_intv;

typedef struct {
  int64_t rbeg;
  int32_t qbeg, len, rid;
} orc_seed;

/* One output alignment. */
typedef struct {
  int32_t rid;          /* contig index */
  int32_t is_rev;
  int64_t pos;          /* reference-format position (0-based, see DR-12) */
  int32_t flag;
  int32_t score;        /* region (extension) score */
  int32_t nw_score;     /* global-alignment score */
  int32_t NM;           /* edit distance as the reference's traceback counts it */
  int32_t qb, qe;
  int64_t rb, re;
  int32_t n_cigar;      /* reference-format CIGAR (clips added) */
  int32_t n_core;       /* global-alignment CIGAR, read orientation */
  int64_t cigar_off;    /* into orc_batch.cigar: n_cigar words then n_core words */
} orc_aln;

typedef struct {
  int32_t n_reads;
  int32_t *aln_off;     /* n_reads + 1 */
  int64_t n_alns;
  orc_aln *alns;
  uint32_t *cigar;
  int64_t n_cigar_words;
  /* work counters for the roofline (summed over the batch) */
  int64_t fm_steps;     /* single-base FM extension steps the reference performs */
  int64_t lf_steps;     /* inverse-Psi steps of SA resolution (sa_intv 32) */
  int64_t sa_reads;     /* SA samples read */
  int64_t ext_cells;    /* DP cells in the extension cell set */
  int64_t ext_tasks;
  int64_t tb_cells;     /* global-alignment band cells */
} orc_batch;

void orc_opt_init(orc_opt *o);
orc_index *orc_index_load(const char *prefix);
void orc_index_free(orc_index *idx);
int64_t orc_index_l_pac(const orc_index *idx);
int orc_index_n_contigs(const orc_index *idx);
const char *orc_index_contig_name(const orc_index *idx, int rid);
int64_t orc_index_contig_len(const orc_index *idx, int rid);
int64_t orc_index_contig_offset(const orc_index *idx, int rid);
uint64_t orc_index_primary(const orc_index *idx);

/* Occ(p, c): occurrences of base c in BWT rows [0, p], '$' excluded. */
int64_t orc_occ(const orc_index *idx, int64_t p, int c);
/* SA value of BWT row k by the inverse-Psi walk (bwa/src/bwt_CUDA.cu:444-454). */
int64_t orc_sa(const orc_index *idx, int64_t k);

/* Per-read stage results (stage dumps). Return counts; arrays must hold cap entries. */
int orc_seed_intervals(const orc_index *idx, const orc_opt *opt, const uint8_t *read, int len,
                       orc_intv *out, int cap, int64_t *fm_steps);
int orc_read_seeds(const orc_index *idx, const orc_opt *opt, const uint8_t *read, int len,
                   orc_seed *out, int cap);

/* Chaining + chain filter of one read's sorted seeds (stage dump). */
int orc_chain_dump(const orc_index *idx, const orc_opt *opt, const orc_seed *seeds, int n, int32_t *chain_of_seed,
                   int32_t *weight, int32_t *kept, int *n_kept);

/* Region filter + sort of one read's regions (stage dump). */
int orc_region_dump(const orc_index *idx, const orc_opt *opt, int n, const int64_t *rb, const int64_t *re,
                    const int32_t *qb, const int32_t *qe, const int32_t *score, const int32_t *rid, int32_t *out_idx,
                    int32_t *out_sec);

/* Align a batch: reads are 2-bit codes (A0 C1 G2 T3, other 4) concatenated
 * with offsets[n+1].  Returns 0 on success. */
int orc_align_batch(const orc_index *idx, const orc_opt *opt, int n, const uint8_t *seq,
                    const int32_t *offsets, int n_threads, orc_batch *out);
void orc_batch_free(orc_batch *b);

/* Banded local extension of bwa/src/ksw.cu:814-1034, as its cell set and
 * tie-break rule (the product contract), and as a lane-by-lane emulation of
 * the 32-thread warp program (used only to check the former). */
int orc_ksw_extend(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                   const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins, int w,
                   int end_bonus, int h0, int *qle, int *tle, int64_t *cells);
int orc_ksw_extend_warp_emulate(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                                const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins, int w,
                                int end_bonus, int h0, int *qle, int *tle);
/* Global banded alignment with traceback (bwa/src/ksw.cu:706-812). */
int orc_ksw_global2(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                    const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins, int w,
                    int *n_cigar, uint32_t *cigar, int cigar_cap);

/* Paired-end primitives (bwa/src/aux.cu:411-500, 703-714, 1000-1070). */
typedef struct {
  int low, high, failed;
  double avg, std;
} orc_pestat;
int orc_infer_dir(int64_t l_pac, int64_t b1, int64_t b2, int64_t *dist);
int orc_cal_sub(const orc_opt *opt, int n, const int32_t *qb, const int32_t *qe, const int32_t *score);
void orc_mem_pestat(const orc_opt *opt, int64_t l_pac, int max_ins, int n, const int64_t *aln_off,
                    const int32_t *rid, const int64_t *rb, const int32_t *qb, const int32_t *qe,
                    const int32_t *score, orc_pestat pes[4]);
/* Parity unpinned: checked only as a bijection on samples (no known-answer
 * vectors here); constants read against bwa/src/aux.cu:703-714 and
 * bwa/src/utils.h:102-112.  Used only as a tie-break between equal pair
 * scores. */
uint64_t orc_hash_64(uint64_t key);
int orc_mem_pair(const orc_opt *opt, int64_t l_pac, const int64_t *ctg_off, const orc_pestat pes[4], int id,
                 int n0, int n1, const int32_t *rid, const int64_t *rb, const int32_t *score, int *sub,
                 int *n_sub, int z[2]);

/* MAPQ (bwa/src/aux.cu:676-700) with sub = csub = sub_n = 0 and frac_rep = 0. */
int orc_mapq(const orc_opt *opt, int score, int qb, int qe, int64_t rb, int64_t re);

#ifdef __cplusplus
}
#endif
#endif
