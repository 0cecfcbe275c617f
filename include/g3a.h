/*
 * g3a: MI355X-native BWA-MEM short-read alignment hot path (C ABI).
 *
 * Drop-in boundary for the reference's device interface
 * (sunwookim028/G3SA, bwa/src/cuda_wrapper.h and bwa/src/bwamem.h).  The
 * reference splits one GPU batch into five calls on a `process_data_t`; the
 * same five phases are exposed here on an opaque context with plain pointers
 * and sizes:
 *
 *   reference (file:line)                          here
 *   device_alloc       cuda_wrapper.h:22-24        g3a_device_alloc
 *   memcpy_index       cuda_wrapper.h:26-30        g3a_memcpy_index
 *   memcpy_input       cuda_wrapper.h:85-86        g3a_memcpy_input
 *   bwa_align          bwamem.h:10                 g3a_bwa_align
 *   memcpy_output      cuda_wrapper.h:88-90        g3a_memcpy_output
 *   destruct_proc      cuda_wrapper.cu:408-411     g3a_destroy
 *
 * g3a_align_batch runs input copy, alignment and output copy in one call.
 * Every function returns 0 on success or a negative G3A_E* code; the message
 * of the last failure on a context is available from g3a_last_error.  No
 * function exits or aborts the process.
 *
 * Threading: one context per device; calls on one context must not overlap.
 * Contexts on different devices may be used concurrently.
 */
#ifndef G3A_H
#define G3A_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define G3A_OK 0
#define G3A_E_ARG (-1)      /* invalid argument */
#define G3A_E_INDEX (-2)    /* index files missing or malformed */
#define G3A_E_HIP (-3)      /* HIP runtime error (includes device OOM) */
#define G3A_E_STATE (-4)    /* call out of order (e.g. align before input) */
#define G3A_E_INTERNAL (-5)

typedef struct g3a_ctx g3a_ctx;

/* Alignment options: the subset of the reference's mem_opt_t
 * (bwa/src/bwa.h:157-189) read on this path.  g3a_opt_init fills BWA's
 * defaults as mem_opt_init does (bwa/src/fastmap.cpp:77-113). */
typedef struct {
  int a, b, o_del, e_del, o_ins, e_ins, w, T, pen_clip5, pen_clip3;
  int min_seed_len, split_width, max_occ, max_chain_gap, flag;
  int64_t max_mem_intv;
  float split_factor, mask_level, drop_ratio, mask_level_redun;
  int8_t mat[25];
} g3a_opt;

/* Output of one batch (the reference's memcpy_output arrays,
 * cuda_wrapper.cu:427-461, without its 16-operation CIGAR cap).  Alignments
 * of read i are [aln_off[i], aln_off[i+1]); a read without alignments is
 * unmapped.  All arrays are allocated by the library and released with
 * g3a_batch_out_free. */
typedef struct {
  int32_t n_reads;
  int64_t *aln_off;      /* n_reads + 1 */
  int64_t n_alns;
  int32_t *rid;          /* contig index */
  int64_t *pos;          /* 0-based position in the contig (reference record) */
  int32_t *flag;         /* 0x10 reverse strand, 0x800 supplementary, 0x100 secondary */
  int32_t *is_rev;
  int32_t *score;        /* extension score (AS) */
  int32_t *nw_score;     /* global-alignment score */
  int32_t *nm;           /* edit distance as the reference's traceback counts it */
  int32_t *qb, *qe;      /* aligned read interval [qb, qe) */
  int64_t *rb, *re;      /* aligned interval of the fwd+revcomp text [rb, re) */
  int32_t *n_cigar;      /* reference-format CIGAR (with soft clips), BAM encoding */
  int64_t *cigar_off;    /* into cigar[] */
  int32_t *n_core;       /* global-alignment CIGAR in read orientation */
  int64_t *core_off;     /* into cigar[] */
  uint32_t *cigar;
  int64_t n_cigar_words;
  /* device time of the pipeline stages of this batch, milliseconds */
  float ms_seed, ms_sa, ms_chain, ms_extend, ms_traceback, ms_total;
  /* work counters of this batch */
  int64_t n_seeds, n_tasks, ext_cells, fm_steps, n_slow_reads;
  /* MAPQ of each alignment (mem_approx_mapq_se, bwa/src/aux.cu:676-700, with
   * sub = min_seed_len * a and frac_rep = 0), and its MD string (aux.cu:2908-
   * 2941, in genome orientation as written to SAM): alignment i's string is
   * md + md_off[i], zero-terminated; md_off has n_alns + 1 entries. */
  int32_t *mapq;
  int64_t *md_off;
  char *md;
  int64_t n_md_bytes;
} g3a_batch_out;

void g3a_opt_init(g3a_opt *opt);

/* device_alloc: create a context on HIP device `device` for batches of up to
 * max_reads reads and max_bases bases, and load the index at `prefix`
 * (.ann .amb .pac .bwt and .g3sa or .sa) into host memory. */
int g3a_device_alloc(int device, const char *prefix, const g3a_opt *opt, int max_reads, int64_t max_bases,
                     g3a_ctx **out);
/* memcpy_index: upload the index to HBM. */
int g3a_memcpy_index(g3a_ctx *ctx);
/* memcpy_input: copy a batch to the device; seq holds 2-bit codes (A0 C1 G2
 * T3, anything else 4) of n_reads reads concatenated; offsets[n_reads+1]. */
int g3a_memcpy_input(g3a_ctx *ctx, int n_reads, const uint8_t *seq, const int32_t *offsets);
/* bwa_align: run the pipeline on the batch resident on the device. */
int g3a_bwa_align(g3a_ctx *ctx);
/* bwa_align without device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No reference counterpart.  The same pipeline and output as
 * g3a_bwa_align, with its per-batch workspace carved out of a static device
 * array of G3A_NOALLOC_SCRATCH_BYTES per device instead of allocated: it calls
 * no hipMalloc, hipFree or allocating library routine.  The batch (resident
 * from g3a_memcpy_input, on a context set up by g3a_device_alloc and
 * g3a_memcpy_index) may hold at most G3A_NOALLOC_MAX_READS reads; a batch
 * whose intermediate data do not fit the array (many seeds or gapped
 * alignments per read) is refused with G3A_E_ARG and the message names the
 * limit.  Results stay in the array for g3a_memcpy_output and
 * g3a_format_records until the next call on that device.  Every context on
 * one device shares the array: calls that overlap on one device race, and a
 * call overwrites the previous call's results. */
#define G3A_NOALLOC_MAX_READS 4096
#define G3A_NOALLOC_SCRATCH_BYTES (256ull << 20)
int g3a_bwa_align_noalloc(g3a_ctx *ctx);
/* memcpy_output: copy the results of the last g3a_bwa_align to the host. */
int g3a_memcpy_output(g3a_ctx *ctx, g3a_batch_out *out);
/* Names and base qualities of the batch of the last g3a_memcpy_input (same
 * n_reads), for the device record formatter: names concatenated with
 * name_off[n_reads+1], qualities with qual_off[n_reads+1].  No reference
 * counterpart: the reference keeps them on the host and formats there
 * (bwa/src/pipeline.cpp:287-351). */
int g3a_memcpy_text(g3a_ctx *ctx, int n_reads, const char *names, const int64_t *name_off, const char *quals,
                    const int64_t *qual_off);
/* Formats the records of the last g3a_bwa_align on the device, in read
 * order: mode 0 the reference's record lines (NAME RNAME POS0 CIGAR SEQ QUAL,
 * or NAME * 0 * SEQ QUAL when unmapped; bwa/src/pipeline.cpp:287-351), mode 1
 * SAM records without the header (FLAG, 1-based POS, MAPQ, CIGAR and SEQ/QUAL
 * in genome orientation, NM, MD and AS tags).  *bytes = the text's length.
 * The text stays in HBM until the next batch; g3a_copy_text copies it out. */
int g3a_format_records(g3a_ctx *ctx, int mode, int64_t *bytes);
/* Copies the formatted records to dst (cap >= their length). */
int g3a_copy_text(g3a_ctx *ctx, char *dst, int64_t cap);
/* Paired-end primitives over the last g3a_bwa_align, whose batch holds read
 * pairs as reads 2p and 2p+1 (an odd n_reads is G3A_E_ARG).  The reference
 * carries them as device ports that its pipeline never launches (it aligns
 * pairs as single-end reads); its host caller would bind them where BWA calls
 * mem_pestat and mem_pair.
 *
 * Insert-size distribution per orientation (0 FF, 1 FR, 2 RF, 3 RR) from the
 * pairs whose ends have a unique top hit on one contig, with sizes in
 * (0, max_ins]: mem_pestat, bwa/src/aux.cu:435-500 (reference default max_ins
 * 10000, bwa/src/fastmap.cpp:97).  Layout of mem_pestat_t
 * (bwa/src/bwa.h:250-254). */
typedef struct {
  int32_t low, high; /* proper-pair bounds of the insert size */
  int32_t failed;    /* too few pairs in this orientation */
  double avg, std;
} g3a_pestat;
#define G3A_PE_MAX_INS (1 << 20)
int g3a_mem_pestat(g3a_ctx *ctx, int max_ins, g3a_pestat pes[4]);
/* Best pair of hits per pair (mem_pair, bwa/src/aux.cu:1000-1070, with every
 * hit of an end taken as primary) and the proper-pair bit of the two top hits
 * (bwa/src/aux.cu:1288-1295).  out[6 * p + k], k = 0..5: pair score (0 when
 * no consistent pair exists), second-best pair score, number of pairs within
 * the gap-open distance of the second best, index of the chosen hit of read 1
 * and of read 2 in their alignment lists (-1 when none), and 2 when the top
 * hits form a proper pair under pes (else 0).  id0 is the id of the batch's
 * first pair (the reference hashes pair ids into its tie-break). */
int g3a_mem_pair(g3a_ctx *ctx, const g3a_pestat pes[4], int64_t id0, int32_t *out);
/* memcpy_input + bwa_align + memcpy_output */
int g3a_align_batch(g3a_ctx *ctx, int n_reads, const uint8_t *seq, const int32_t *offsets, g3a_batch_out *out);
void g3a_batch_out_free(g3a_batch_out *out);
/* Stage times (ms: seed, sa, chain, extend, traceback, total) and work
 * counters (seeds, tasks, extension cells, FM steps, slow-path reads,
 * alignments) of the last g3a_bwa_align, without copying its output.  No
 * reference counterpart: the reference's per-stage timers are the
 * CUDA-event tprof table (bwa/src/bwamem.cu:31-49). */
int g3a_last_stats(const g3a_ctx *ctx, float ms[6], int64_t work[6]);
/* Device time (ms) of single kernels of the last g3a_bwa_align, from events
 * recorded on the context's stream around each launch: [0] k_seed, [1] the
 * lane extension engine (both sides; 0 in the plain form), [2] one-wavefront
 * extension (both sides, or the plain form's k_extend), [3] k_traceback.
 * No reference counterpart. */
int g3a_kernel_ms(const g3a_ctx *ctx, float ms[4]);
void g3a_destroy(g3a_ctx *ctx);

/* Execution form of the extension stage.  Both forms give identical output.
 *   G3A_FORM_PLAIN    one wavefront per extension task (anti-diagonal DP)
 *   G3A_FORM_BATCHED  tasks binned by query length; one task per lane in
 *                     persistent kernels pulling 64-task chunks from a
 *                     device cursor; tasks outside the lane engine's exact
 *                     domain fall back to one wavefront per task
 * No reference counterpart (the reference has one form, ksw_extend_warp2
 * under localExtending, bwa/src/aux.cu:2289-2386).  Default: BATCHED. */
#define G3A_FORM_PLAIN 0
#define G3A_FORM_BATCHED 1
int g3a_set_form(g3a_ctx *ctx, int form);
/* Run each g3a_bwa_align as one HIP graph (on = 1; default 0).  A batch
 * runs with no host round trip in either mode: its buffers are sized before
 * the first kernel from the seeds, tasks and alignments per read of the
 * batches seen so far, the kernels read the true counts from device memory,
 * and a count that outgrows its buffer stops the kernels after it and makes
 * the call run the batch again with that buffer grown.  With on = 1 the
 * launches are captured into a graph, replayed while the batch's read count,
 * buffers, index, options and tuning stay the same, and captured again when
 * they change (a batch seeded in minimizer order, g3a_set_seed_order(ctx, 1),
 * and the allocation-free entry point run without a graph).  Output is the
 * same in both modes.  Stage times (g3a_last_stats) then give only the
 * total.  No reference counterpart: the reference launches its stages one
 * by one with host synchronizations (bwa/src/bwamem.cu:252, 340). */
int g3a_set_graph(g3a_ctx *ctx, int on);
/* [0] 1 when the last g3a_bwa_align ran as a graph, [1] graph captures so
 * far, [2] batch attempts redone because a count outgrew its buffer.  No
 * reference counterpart. */
int g3a_graph_info(const g3a_ctx *ctx, int64_t info[3]);
/* A second context on src's device that shares src's device index (and its
 * learned index, when src built it before this call): its own stream,
 * staging and workspace for batches of up to max_reads reads and max_bases
 * bases, and src's options, form and tuning.  Batches on src and its
 * siblings can run at the same time, each from its own host thread; while
 * seeding on one waits on memory, extension on the other keeps the vector
 * units busy.  The index buffers are freed with the last context that holds
 * them.  A context whose index is shared refuses g3a_memcpy_index, and a
 * sibling refuses to build a learned index.  No reference counterpart: the
 * reference allocates one process_data_t per GPU (bwa/src/pipeline.cpp:54-67). */
int g3a_device_alloc_sibling(g3a_ctx *src, int max_reads, int64_t max_bases, g3a_ctx **out);
/* The capacities the next batch starts from: v = {seeds, extension tasks,
 * alignments, CIGAR words used, per read; direction bytes, CIGAR slot words,
 * per base; slow-path reads per batch}.  Later batches raise them again
 * from the counts they see.  Small values make batches run again with grown
 * buffers (the path tests exercise); the defaults suit 150 bp reads.  No
 * reference counterpart. */
int g3a_set_batch_caps(g3a_ctx *ctx, const double v[7]);
/* Tuning of the batched form: rows per register strip (16 or 32) and
 * persistent blocks per compute unit for the lane and wavefront kernels.
 * No reference counterpart. */
int g3a_set_tuning(g3a_ctx *ctx, int strip_rows, int lane_blocks_per_cu, int wave_blocks_per_cu);
/* Lane engine arithmetic: 1 (default) runs the plain columns with two strip
 * rows per register in 16-bit lanes (packed integer instructions, the lower
 * half one column behind the upper), 0 with one cell per 32-bit lane.
 * Identical results.  g3a_ext_packed returns the current setting.  No
 * reference counterpart. */
int g3a_set_ext_packed(g3a_ctx *ctx, int packed);
/* Index layout of texts of 2^32 rows or more (32-byte Occ lines of 64 rows
 * with 2^32-row superblocks, 64-bit intervals in seeding, no 12-mer table,
 * no learned form), used on any text when on = 1; called before
 * g3a_memcpy_index.  Output is identical.  For tests of that layout on texts
 * small enough to check.  No reference counterpart. */
int g3a_set_wide_layout(g3a_ctx *ctx, int on);
int g3a_ext_packed(const g3a_ctx *ctx);
/* Autotuner of the batched form: on the batch last aligned (its tasks are
 * still on the device), time the extension stage for each lane arithmetic
 * (packed or not), strip height (16, 32) and persistent lane blocks per
 * compute unit (2, 4, 8, 16), keep the fastest setting on the context and
 * report it (the arithmetic through g3a_ext_packed).  Results are identical
 * for every setting; the batch's output is left as it was.  No reference
 * counterpart. */
int g3a_autotune(g3a_ctx *ctx, int *strip_rows, int *lane_blocks_per_cu, float *best_ms);
/* Persistent seeding blocks per compute unit at the current block size
 * (g3a_set_seed_grid); 0: the occupancy maximum.  Each block's share of the
 * compute unit's LDS holds the first entries of its threads' LEP lists (the
 * rest spill to a global slice per thread), so the grid trades latency
 * hiding against list traffic.  No reference counterpart. */
int g3a_set_seed_blocks(g3a_ctx *ctx, int blocks_per_cu);
/* The seeding grid in full: threads per block (a multiple of 64, at most
 * 1024; default 1024) and blocks per compute unit (0: the occupancy maximum;
 * default 1).  Each block keeps one
 * copy of the Occ superblock counts in LDS, so larger blocks leave more of
 * the compute unit's LDS to the threads' lists.  No reference counterpart. */
int g3a_set_seed_grid(g3a_ctx *ctx, int threads_per_block, int blocks_per_cu);
/* Seeding order: 1 seeds the reads in the order of a minimizer
 * key (the smallest hash over each read's canonical 16-mers), so that
 * overlapping reads run in neighbouring lanes and share cached Occ lines;
 * 0 (default) seeds them in input order.  Results are identical.  No reference
 * counterpart. */
int g3a_set_seed_order(g3a_ctx *ctx, int ordered);
/* Chaining path: 0 (default) chains a read of at most 16 seeds on one
 * thread and a larger one on a wavefront, both keeping the B-tree of chains
 * as a sorted array; 1 runs the B-tree code for every read on one thread.
 * The results are identical.  No reference counterpart. */
int g3a_set_chain_serial(g3a_ctx *ctx, int on);
/* Seeding autotuner: re-aligns the batch last aligned once per seeding grid
 * (threads per block x blocks per compute unit: 1024x1, 512x2, 256x4, 128x6;
 * each block keeps one copy of the superblock counts in LDS and the rest of
 * its share holds the first entries of its threads' LEP lists, so fewer,
 * larger blocks keep more in LDS) and order (input, minimizer), keeps the
 * setting with the fastest seeding stage and reports it with its time.  The
 * batch's output is identical for every setting.  No reference counterpart. */
int g3a_autotune_seed(g3a_ctx *ctx, int *threads_per_block, int *blocks_per_cu, int *ordered, float *best_ms);
/* Learned-index form.  g3a_set_learned(ctx, 1, info) builds, on the device
 * and once per context, the full suffix array, the 32-base key of every
 * suffix-array row and a two-stage FP32 model from a key to its row
 * (piecewise-linear root over the first 6 bases; 4096 leaves, each a
 * one-hidden-layer ReLU network of the key trained by an error-bounded fit
 * to within 250 rows of every distinct key), and switches seeding to it: a
 * read's match that occurs once is extended by comparing it with the text;
 * the last round's intervals are read from the sorted keys by an exact
 * search inside the 12-mer's interval (the k-mer table), narrowed in large
 * buckets to the model's window of at most +-256 rows around its prediction
 * (checked at its edges; a miss searches the whole bucket), with the
 * FM-index as the fallback where no bucket is known.  Output is identical to
 * the FM-index-only path.  0 switches back (the built model is kept); 2 uses
 * the full suffix array (text comparison, SA lookup) without the model's
 * lookups, to measure what they add.  Texts of fewer than 2^32 rows.  No
 * reference counterpart. */
typedef struct g3a_learned_info {
  int64_t rows, distinct_keys;
  int32_t leaves, hidden_units, window_cap, pad; /* hidden_units: the most in one leaf */
  int64_t model_bytes; /* root + leaves + knots + short-suffix table */
  int64_t sa_bytes, key_bytes;
  int64_t max_err;            /* largest |predicted - true row| over the distinct keys */
  int64_t keys_over_window;   /* distinct keys whose error exceeds the window */
  int64_t leaves_over_window; /* leaves whose largest error exceeds it */
  double mean_err;
  float build_ms, pad2;
  int64_t knots; /* knots of all leaves (hidden units + 1 per leaf) */
} g3a_learned_info;
int g3a_set_learned(g3a_ctx *ctx, int on, g3a_learned_info *info);
/* Learned-form work of the last g3a_bwa_align: [0] interval lookups,
 * [1] lookups answered by the FM-index instead, [2] text comparisons of
 * unique matches, [3] bases they covered, [4] lookups whose search the
 * model's window narrowed, [5] model windows that missed (the lookup then
 * searched its whole k-mer bucket).  No reference counterpart. */
int g3a_learned_stats(const g3a_ctx *ctx, int64_t out[6]);
/* The seeding stage's intervals of the last g3a_bwa_align (SMEMs, reseeded
 * and last-round seeds, in each read's order): counts[n_reads] per read, and
 * (k, s, m, n) per interval into kmsn (4 * cap words) unless kmsn is null;
 * *total = their number.  For parity checks of the seeding forms.  No
 * reference counterpart (the reference keeps them in d_aux, bwa/src/seed.cu). */
int g3a_last_intervals(const g3a_ctx *ctx, int32_t *counts, int64_t *kmsn, int64_t cap, int64_t *total);
const char *g3a_last_error(const g3a_ctx *ctx);

/* Index construction (the `g3a index` command): BWA-format files plus the
 * dense suffix-array sample with interval sa_intv (power of two, 1..32). */
int g3a_build_index(const char *fasta, const char *prefix, int sa_intv, char *err, int err_len);
/* The same files, with the suffix array, BWT, samples and count tables
 * built on HIP device `device` (prefix doubling over radix sorts in HBM);
 * byte-identical to g3a_build_index.  The reference builds its index on the
 * host only (bwa/src/FMI_search.cpp:232-420 build_index, bwa/src/bwt.c:62-84);
 * this entry point has no counterpart there. */
int g3a_build_index_dev(int device, const char *fasta, const char *prefix, int sa_intv, char *err, int err_len);

#ifdef __cplusplus
}
#endif
#endif
