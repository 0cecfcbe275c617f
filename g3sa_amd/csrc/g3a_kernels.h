// Kernel argument blocks and launch entry points (g3a_kernels.hip).
#pragma once
#include "g3a_dev.h"
#include "g3a_ext.h"
#include "g3a_text.h"

#include <hip/hip_runtime.h>

namespace g3a {

constexpr int kSlowLep = 400;                  // slow-path LEP list capacity
constexpr int kSlowScratch = 5 * kSlowLep;     // Intv entries of scratch per slow read
constexpr int kSlotCap = 64;                   // fast-path interval slots per read
constexpr int kLepCap = 64;                    // fast-path LEP list capacity
constexpr int kFoundCap = 16;                  // fast-path found-match (and reseed seed) list capacity

struct SeedArgs {
  const int32_t *stop;  // when *stop is set the kernel returns at once (null: never)
  DevIndex idx;
  Opt opt;
  int n_reads;
  const uint8_t *seq;
  const int32_t *off;
  Intv *intv;        // n_reads * slot_cap fast-path slots
  int slot_cap;
  int32_t *n_intv;
  int64_t *n_seed;
  int32_t *slow;
  int64_t *fm_steps;
  Intv *lep_pool;    // kLepCap + kFoundCap entries per k_seed thread
  const int32_t *order;  // reads in processing order, or null for input order
  int lep_lds;           // LEP entries per thread kept in LDS (narrow form; set by launch_seed)
  int64_t *li_stats;     // learned-index form counters (6), or null
  int32_t *next_read;    // zeroed before k_seed: reads are handed out by an atomic counter (null: fixed stride)
};

// Counts known only on the device (totals of a batch's seeds, tasks and
// alignments) reach the kernels as a pointer (`*_dev`) beside the capacity
// their buffers and grids were set up for; a kernel works on
// min(*dev, capacity) items (all `capacity` when the pointer is null).
struct SaArgs {
  const int32_t *stop;  // when *stop is set the kernel returns at once (null: never)
  DevIndex idx;
  Opt opt;
  int n_reads;
  int64_t total;  // capacity
  const int64_t *total_dev;
  const int64_t *seed_off;  // n_reads + 1
  const Intv *intv;
  const int64_t *intv_base;  // per read
  const int32_t *n_intv;
  Seed *seeds;
};


struct ChainArgs {
  const int32_t *stop;  // when *stop is set the kernel returns at once (null: never)
  DevIndex idx;
  Opt opt;
  int n_reads;
  const int32_t *off;
  const int64_t *seed_off;
  Seed *seeds;     // as the SA lookup wrote them; sorted in place only by a read that runs on one thread
  Seed *sort_tmp;  // sort_seeds scratch, one entry per seed
  ChainTmp *ch;
  int32_t *nxt, *order, *tmpi;
  uint16_t *w16;
  uint8_t *kept;
  KbNode *nodes;
  Task *tasks_tmp;
  int32_t *n_task;
  int32_t *error;
  int32_t *big, *n_big;  // reads of more than 128 seeds and their count (zero before the launch)
  int serial;            // 1: every read by the B-tree code on one thread (chain_one)
};

struct CompactArgs {
  const int32_t *stop;  // when *stop is set the kernel returns at once (null: never)
  int n_reads;
  const int64_t *seed_off, *task_off;
  const Task *tasks_tmp;
  Task *tasks;
};

struct ExtArgs {
  const int32_t *stop;  // when *stop is set the kernel returns at once (null: never)
  DevIndex idx;
  Opt opt;
  int64_t n_tasks;  // capacity
  const int64_t *n_dev;
  const uint8_t *seq;
  const int32_t *off;
  const Task *tasks;
  Region *regs;
  int64_t *cells;
};

// Result of one side of one extension task.
struct ExtRes {
  int32_t score, qle, tle, pad;
};

// One side (0 = left, 1 = right) of the extension stage, batched form:
// classify -> sort lane tasks by descending query length -> lane engine
// (persistent, chunks of 64 from a device cursor) -> one wavefront per
// remaining task (persistent) -> finish.
struct ExtSideArgs {
  const int32_t *stop;  // when *stop is set the kernel returns at once (null: never)
  DevIndex idx;
  Opt opt;
  int side;
  int lane_opt_ok;        // scoring admitted by the lane engine
  int64_t n_tasks;        // capacity
  const int64_t *n_dev;
  const uint8_t *seq;
  const int32_t *off;
  const Task *tasks;
  const ExtRes *left;     // left results: h0 of the right side
  ExtRes *res;            // this side's results
  uint32_t *key;          // sort key: query length of lane tasks, 0 otherwise
  const int32_t *sorted;  // lane task indices by descending key
  int32_t *slow_list;     // tasks for the one-wavefront kernel
  int32_t *ctr;           // [0] lane tasks [1] slow tasks [2] lane chunk cursor
  uint32_t *rowbuf;       // kExtQCap * 64 words per lane-kernel block
  int64_t *cells;
};

struct ExtFinishArgs {
  const int32_t *stop;  // when *stop is set the kernel returns at once (null: never)
  int64_t n_tasks;  // capacity
  const int64_t *n_dev;
  const Task *tasks;
  const ExtRes *left, *right;
  Region *regs;
};

struct RegionArgs {
  const int32_t *stop;  // when *stop is set the kernel returns at once (null: never)
  Opt opt;
  int n_reads;
  const int64_t *task_off;
  Region *regs;
  uint8_t *flag;
  int32_t *sec;
  int32_t *n_aln;
};

struct AlnPrepArgs {
  const int32_t *stop;  // when *stop is set the kernel returns at once (null: never)
  Opt opt;
  int n_reads;
  const int64_t *task_off, *aln_off;
  const Region *regs;
  Region *aln_reg;
  int32_t *aln_read, *aln_rank;
  int64_t *z_size, *cig_size;
};

struct TbArgs {
  const int32_t *stop;  // when *stop is set the kernel returns at once (null: never)
  DevIndex idx;
  Opt opt;
  int64_t n_alns;  // capacity
  const int64_t *n_dev;
  const uint8_t *seq;
  const int32_t *off;
  const Region *aln_reg;
  const int32_t *aln_read, *aln_rank;
  const int64_t *z_off, *cig_off;
  uint8_t *z;
  uint32_t *cigar;
  Aln *alns;
  int32_t *eh_pool;  // traceback_pool_words(threads) words
  int32_t *tb_order;  // n_alns: processing order
  int32_t *tb_ctr;    // 8 counters: class sizes, placement cursors
  int32_t *tb_nw;     // n_alns: global-alignment scores from k_tb_wave_dp
};

// threads of k_seed's persistent grid in blocks of `block` threads
// (occupancy-limited, n_cu compute units)
int seed_grid_threads(int n_cu, const DevIndex &x, int block);
// blocks_per_cu: the grid the caller runs (0: the occupancy maximum), which
// with the block size sets the LDS share of each block for the LEP lists
void launch_seed(const SeedArgs &a, int n_threads, int blocks_per_cu, int block, hipStream_t s);
// slow reads list[0, n_list) of the whole list's *n_dev (from entry c0 of it)
void launch_seed_slow(const SeedArgs &a, const int32_t *list, int n_list, const int32_t *n_dev, int c0, Intv *scratch,
                      const int64_t *pool_off, Intv *pool, int pass, int from_scratch, hipStream_t s);
void launch_sa_lookup(const SaArgs &a, hipStream_t s);
void launch_chain(const ChainArgs &a, hipStream_t s);
void launch_task_compact(const CompactArgs &a, hipStream_t s);
void launch_extend(const ExtArgs &a, hipStream_t s);
void launch_ext_classify(const ExtSideArgs &a, hipStream_t s);
// packed: plain columns with two strip rows per register (16-bit lanes)
void launch_ext_lane(const ExtSideArgs &a, int strip_rows, int packed, int n_blocks, hipStream_t s);
void launch_ext_wave(const ExtSideArgs &a, int n_blocks, hipStream_t s);
void launch_ext_finish(const ExtFinishArgs &a, hipStream_t s);
void launch_fill_i32(int32_t *p, int n, int32_t v, hipStream_t s);
void launch_clamp_codes(uint8_t *seq, int64_t n, hipStream_t s);
// fills the 3 * 4^kKmer words of the k-mer interval table
void launch_kmer_table(const DevIndex &x, uint32_t *table, hipStream_t s);
// record formatting (g3a_text.hip): per-read text lengths, then the text at
// txt_off; per-alignment MAPQ and MD-string lengths, then the MD strings
// the batch's records: each read's text length (write = false: len), or
// its text at txt + txt_off[q] (write = true)
void launch_text(const TextArgs &t, int64_t *len, const int64_t *txt_off, char *txt, bool write, hipStream_t s);
// One formatting pass: every read's text into a staging buffer at stage_off
// (offsets from k_text_bound's bounds), its length into len; then
// launch_text_check flags a read longer than its bound and launch_text_pack
// moves the texts to their places.
void launch_text_bound(const TextArgs &t, int64_t *bound, hipStream_t s);
void launch_text_staged(const TextArgs &t, int64_t *len, const int64_t *stage_off, char *stage, hipStream_t s);
void launch_text_check(int n, const int64_t *len, const int64_t *bound, int64_t *flag, hipStream_t s);
void launch_text_pack(int n, const char *stage, const int64_t *stage_off, const int64_t *len, const int64_t *txt_off,
                      char *txt, hipStream_t s);
void launch_aln_mapq_md(const TextArgs &t, int64_t n_alns, const int32_t *aln_read, int32_t *mapq, int64_t *md_len,
                        hipStream_t s);
void launch_aln_md_write(const TextArgs &t, int64_t n_alns, const int32_t *aln_read, const int64_t *md_off,
                         char *md, hipStream_t s);

// Paired-end statistics and pair choice (g3a_pair.hip).  Reads 2p and 2p+1
// are the two ends of pair p.
struct PeStat {  // layout of g3a_pestat (bwa/src/bwa.h:250-254)
  int32_t low, high, failed;
  double avg, std;
};
struct PairArgs {
  int64_t n_pairs;
  const int64_t *aln_off;  // 2 * n_pairs + 1
  const Aln *alns;
  const int64_t *ctg_off;
  int64_t l_pac;
  int a, b, o_del, e_del, o_ins, e_ins, min_seed_len, max_ins;
  float mask_level;
  uint32_t *hist;  // copies * 4 * (max_ins + 1)
  int copies;      // histogram copies that spread the counting atomics
  PeStat *stat;    // 4, written by launch_pe_stat
  PeStat pes[4];   // read by launch_pe_pair
  uint64_t *v;     // two words per alignment
  int32_t *out;    // six words per pair
  int64_t id0;
};
void launch_pe_stat(const PairArgs &a, hipStream_t s);
void launch_pe_pair(const PairArgs &a, hipStream_t s);
// Occ lines (wide, and narrow with superblock counts when `narrow` is not
// null) from the BWA .bwt body, rows 0..n
void launch_occ_lines(const uint32_t *W, uint64_t n, uint64_t primary, uint32_t *wide, uint32_t *narrow,
                      uint32_t *sbc, hipStream_t s);
// The lane tasks (key > 0, key < 512) of key[0, min(*n_dev, cap)) by
// descending key into out[0, number of lane tasks): a counting sort
// (histogram, bin starts, scatter with per-block reservations); equal keys
// in no fixed order.  hist: 1024 words of scratch.
void ext_sort(const uint32_t *key, int64_t cap, const int64_t *n_dev, int32_t *hist, int32_t *out, hipStream_t s);
void launch_region_filter(const RegionArgs &a, hipStream_t s);
void launch_aln_prep(const AlnPrepArgs &a, hipStream_t s);
int traceback_grid_threads(int n_cu);
int64_t traceback_pool_words(int n_threads);
void launch_traceback(const TbArgs &a, int n_threads, int n_cu, hipStream_t s);
void launch_collect_slow(int n, const int32_t *slow, int32_t *list, int32_t *count, int64_t *intv_base, int slot_cap,
                         hipStream_t s);
// over list entries [0, cap) of which *n_dev are set (entries past it: 0)
void launch_gather_slow_counts(int cap, const int32_t *n_dev, const int32_t *list, const int32_t *n_intv, int64_t *out,
                               hipStream_t s);
void launch_set_slow_base(int cap, const int32_t *n_dev, const int32_t *list, const int64_t *off, int64_t base,
                          int64_t *intv_base, hipStream_t s);
// p[i] = 0 for *n_dev <= i < cap
void launch_zero_tail(int64_t *p, int64_t cap, const int64_t *n_dev, hipStream_t s);
// flag |= bit when *total > cap (*total read as int32 when total32)
void launch_cap_check(const int64_t *total, int64_t cap, int32_t *flag, int32_t bit, hipStream_t s);
void launch_cap_check32(const int32_t *total, int64_t cap, int32_t *flag, int32_t bit, hipStream_t s);
// dst[i] = *src[i] for the k pointers in src (k <= 16), one kernel
struct Gather16 {
  const int64_t *src[16];
};
void launch_gather_totals(const Gather16 &g, int k, int64_t *dst, hipStream_t s);

// exclusive scans (rocPRIM) over int32 or int64 counts into int64 offsets;
// out has n+1 entries, out[n] = total
void scan_i32(const int32_t *in, int64_t *out, int64_t n, void *tmp, size_t tmp_bytes, hipStream_t s);
void scan_i64(const int64_t *in, int64_t *out, int64_t n, void *tmp, size_t tmp_bytes, hipStream_t s);
size_t scan_tmp_bytes(int64_t n);
void launch_seed_keys(const uint8_t *seq, const int32_t *off, int n, uint32_t *key, int32_t *idx, hipStream_t s);
size_t seed_sort_tmp_bytes(int64_t n);
int seed_sort(const uint32_t *key_in, uint32_t *key_out, const int32_t *val_in, int32_t *val_out, int64_t n, void *tmp,
              size_t tmp_bytes, hipStream_t s);
// over alignments [0, cap) of which *n_dev are set (words past it: 0)
void launch_cig_words(const Aln *alns, int64_t cap, const int64_t *n_dev, int32_t *words, hipStream_t s);
void launch_cig_gather(Aln *alns, int64_t cap, const int64_t *n_dev, const int32_t *stop, const int64_t *coff,
                       const uint32_t *cig, uint32_t *out, hipStream_t s);

}  // namespace g3a
