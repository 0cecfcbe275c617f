// The reference's device interface implemented on top of libg3a.
//
// A maintainer of sunwookim028/G3SA adds this file to the reference's host
// build in place of bwa/src/cuda_wrapper.cu (device_alloc, memcpy_index,
// memcpy_input, memcpy_output, destruct_proc, check_device_count) and
// bwa/src/bwamem.cu (bwa_align), and links -lg3a.  Declarations implemented:
//   device_alloc     bwa/src/cuda_wrapper.h:22-24
//   memcpy_index     bwa/src/cuda_wrapper.h:26-30
//   memcpy_input     bwa/src/cuda_wrapper.h:85-86
//   memcpy_output    bwa/src/cuda_wrapper.h:88-90
//   destruct_proc    bwa/src/cuda_wrapper.h:59
//   check_device_count bwa/src/cuda_wrapper.h:56
//   bwa_align        bwa/src/bwamem.h:10
// The library opens the index files itself; the prefix is the one the
// reference's main passes to bwa_idx_load (bwa/src/fastmap.cpp:511), handed
// over with one added line there: g3a_shim_set_index_prefix(argv[optind]).
// The reference's process_data_t (bwa/src/bwa.h:320-399) is kept as the
// per-GPU handle; the library context travels in its generic pointer field
// CUDA_stream, which no code outside the replaced files reads.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "g3a.h"
#include "bwamem.h"        // bwa_align
#include "cuda_wrapper.h"  // the device interface being implemented
#include "macro.h"         // MAX_BATCH_SIZE, MAX_LEN_READ, MAX_N_CIGAR
#include "pipeline.h"      // pipeline_aux_t, aligned_chunk

static const char *g_index_prefix = nullptr;

extern "C" void g3a_shim_set_index_prefix(const char *prefix) { g_index_prefix = prefix; }

static g3a_ctx *ctx_of(process_data_t *p) { return static_cast<g3a_ctx *>(p->CUDA_stream); }

static void die(g3a_ctx *c, const char *what) {
  std::fprintf(stderr, "[g3a] %s: %s\n", what, c ? g3a_last_error(c) : "failed");
  std::exit(EXIT_FAILURE);  // the reference exits on device errors (errchk.cuh:3-12)
}

extern "C" void check_device_count(int num_requested_gpus) {
  if (num_requested_gpus != 1) {
    std::fprintf(stderr, "[g3a] this library drives one GPU per process (-g 1)\n");
    std::exit(1);
  }
}

extern "C" process_data_t *device_alloc(int gpu_no, pipeline_aux_t *aux) {
  if (!g_index_prefix) die(nullptr, "device_alloc: no index prefix (g3a_shim_set_index_prefix)");
  g3a_opt o;
  g3a_opt_init(&o);
  const mem_opt_t *m = aux->opt;  // same fields, same meaning
  o.a = m->a; o.b = m->b; o.o_del = m->o_del; o.e_del = m->e_del;
  o.o_ins = m->o_ins; o.e_ins = m->e_ins; o.w = m->w; o.T = m->T;
  o.pen_clip5 = m->pen_clip5; o.pen_clip3 = m->pen_clip3;
  o.min_seed_len = m->min_seed_len; o.split_width = m->split_width;
  o.max_occ = m->max_occ; o.max_chain_gap = m->max_chain_gap; o.flag = m->flag;
  o.max_mem_intv = m->max_mem_intv; o.split_factor = m->split_factor;
  o.mask_level = m->mask_level; o.drop_ratio = m->drop_ratio;
  o.mask_level_redun = m->mask_level_redun;
  std::memcpy(o.mat, m->mat, 25);
  process_data_t *p = new process_data_t();
  g3a_ctx *c = nullptr;
  // one batch holds at most MAX_BATCH_SIZE reads of MAX_LEN_READ bases
  if (g3a_device_alloc(gpu_no, g_index_prefix, &o, MAX_BATCH_SIZE, (int64_t)MAX_BATCH_SIZE * MAX_LEN_READ, &c) !=
      G3A_OK)
    die(nullptr, "device_alloc");
  p->CUDA_stream = c;
  p->gpu_no = gpu_no;
  return p;
}

extern "C" void memcpy_index(process_data_t *p, int, pipeline_aux_t *) {
  if (g3a_memcpy_index(ctx_of(p)) != G3A_OK) die(ctx_of(p), "memcpy_index");
}

extern "C" void memcpy_input(int batch_size, process_data_t *p, uint8_t *seq, int *seq_offset) {
  if (g3a_memcpy_input(ctx_of(p), batch_size, seq, seq_offset) != G3A_OK) die(ctx_of(p), "memcpy_input");
  p->batch_size = batch_size;
}

extern "C" int bwa_align(int, process_data_t *p, g3_opt_t *) {
  return g3a_bwa_align(ctx_of(p)) == G3A_OK ? 0 : 1;  // the reference returns 1 on failure
}

extern "C" void memcpy_output(aligned_chunk *ac, process_data_t *p) {
  g3a_batch_out b;
  if (g3a_memcpy_output(ctx_of(p), &b) != G3A_OK) die(ctx_of(p), "memcpy_output");
  const int n = b.n_reads;
  ac->offsets.assign(b.aln_off, b.aln_off + n + 1);  // int64 -> int (at most MAX_BATCH_SIZE reads)
  ac->rids.assign(b.rid, b.rid + b.n_alns);
  ac->positions.assign(b.pos, b.pos + b.n_alns);
  ac->ncigars.assign(b.n_cigar, b.n_cigar + b.n_alns);
  // the reference's line builder reads MAX_N_CIGAR slots per alignment
  ac->cigars.assign((size_t)b.n_alns * MAX_N_CIGAR, 0u);
  for (int64_t i = 0; i < b.n_alns; ++i) {
    const int k = b.n_cigar[i] < MAX_N_CIGAR ? b.n_cigar[i] : MAX_N_CIGAR;
    std::memcpy(&ac->cigars[(size_t)i * MAX_N_CIGAR], b.cigar + b.cigar_off[i], 4 * (size_t)k);
  }
  g3a_batch_out_free(&b);
}

// Paired-end statistics of the last batch (reads 2i, 2i+1 as pairs) in the
// reference's own struct: mem_pestat (bwa/src/aux.cu:435-500), whose host
// caller passes opt->max_ins (bwa/src/bwa.h:185) and the process's pes
// (bwa/src/bwa.h:327).  The reference declares this port but never calls it.
static_assert(sizeof(mem_pestat_t) == sizeof(g3a_pestat), "mem_pestat_t and g3a_pestat share one layout");
extern "C" void g3a_shim_pestat(process_data_t *p, const mem_opt_t *opt, mem_pestat_t pes[4]) {
  if (g3a_mem_pestat(ctx_of(p), opt->max_ins, reinterpret_cast<g3a_pestat *>(pes)) != G3A_OK)
    die(ctx_of(p), "mem_pestat");
}

extern "C" void destruct_proc(process_data_t *p) {
  g3a_destroy(ctx_of(p));
  delete p;
}
