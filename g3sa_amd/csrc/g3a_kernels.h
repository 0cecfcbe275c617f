// Kernel argument blocks and launch entry points (g3a_kernels.hip).
#pragma once
#include "g3a_dev.h"

#include <hip/hip_runtime.h>

namespace g3a {

constexpr int kSlowLep = 400;                  // slow-path LEP list capacity
constexpr int kSlowScratch = 5 * kSlowLep;     // Intv entries of scratch per slow read
constexpr int kSlotCap = 64;                   // fast-path interval slots per read

struct SeedArgs {
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
};

struct SaArgs {
  DevIndex idx;
  Opt opt;
  int n_reads;
  int64_t total;
  const int64_t *seed_off;  // n_reads + 1
  const Intv *intv;
  const int64_t *intv_base;  // per read
  const int32_t *n_intv;
  Seed *seeds;
};

struct SortArgs {
  int n_reads;
  const int64_t *seed_off;
  Seed *seeds, *tmp;
};

struct ChainArgs {
  DevIndex idx;
  Opt opt;
  int n_reads;
  const int32_t *off;
  const int64_t *seed_off;
  const Seed *seeds;
  ChainTmp *ch;
  int32_t *nxt, *order, *tmpi;
  uint16_t *w16;
  uint8_t *kept;
  KbNode *nodes;
  Task *tasks_tmp;
  int32_t *n_task;
  int32_t *error;
};

struct CompactArgs {
  int n_reads;
  const int64_t *seed_off, *task_off;
  const Task *tasks_tmp;
  Task *tasks;
};

struct ExtArgs {
  DevIndex idx;
  Opt opt;
  int64_t n_tasks;
  const uint8_t *seq;
  const int32_t *off;
  const Task *tasks;
  Region *regs;
  int64_t *cells;
};

struct RegionArgs {
  Opt opt;
  int n_reads;
  const int64_t *task_off;
  Region *regs;
  uint8_t *flag;
  int32_t *sec;
  int32_t *n_aln;
};

struct AlnPrepArgs {
  Opt opt;
  int n_reads;
  const int64_t *task_off, *aln_off;
  const Region *regs;
  Region *aln_reg;
  int32_t *aln_read, *aln_rank;
  int64_t *z_size, *cig_size;
};

struct TbArgs {
  DevIndex idx;
  Opt opt;
  int64_t n_alns;
  const uint8_t *seq;
  const int32_t *off;
  const Region *aln_reg;
  const int32_t *aln_read, *aln_rank;
  const int64_t *z_off, *cig_off;
  uint8_t *z;
  uint32_t *cigar;
  Aln *alns;
};

void launch_seed(const SeedArgs &a, hipStream_t s);
void launch_seed_slow(const SeedArgs &a, const int32_t *list, int n_list, Intv *scratch, const int64_t *pool_off,
                      Intv *pool, int pass, hipStream_t s);
void launch_sa_lookup(const SaArgs &a, hipStream_t s);
void launch_sort(const SortArgs &a, hipStream_t s);
void launch_chain(const ChainArgs &a, hipStream_t s);
void launch_task_compact(const CompactArgs &a, hipStream_t s);
void launch_extend(const ExtArgs &a, hipStream_t s);
void launch_region_filter(const RegionArgs &a, hipStream_t s);
void launch_aln_prep(const AlnPrepArgs &a, hipStream_t s);
void launch_traceback(const TbArgs &a, hipStream_t s);
void launch_collect_slow(int n, const int32_t *slow, int32_t *list, int32_t *count, int64_t *intv_base, int slot_cap,
                         hipStream_t s);
void launch_gather_slow_counts(int n, const int32_t *list, const int32_t *n_intv, int64_t *out, hipStream_t s);
void launch_set_slow_base(int n, const int32_t *list, const int64_t *off, int64_t base, int64_t *intv_base,
                          hipStream_t s);

// exclusive scans (rocPRIM) over int32 or int64 counts into int64 offsets;
// out has n+1 entries, out[n] = total
void scan_i32(const int32_t *in, int64_t *out, int64_t n, void *tmp, size_t tmp_bytes, hipStream_t s);
void scan_i64(const int64_t *in, int64_t *out, int64_t n, void *tmp, size_t tmp_bytes, hipStream_t s);
size_t scan_tmp_bytes(int64_t n);

}  // namespace g3a
