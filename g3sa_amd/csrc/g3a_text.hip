// Device formatting of a batch's output records (emit_read in g3a_text.h):
// thread per read, a length pass, a scan, a write pass.  Also each
// alignment's MAPQ and MD string for the C ABI's output arrays.
#include "g3a_kernels.h"
#include "g3a_text.h"

#include <hip/hip_runtime.h>

namespace g3a {

__global__ void __launch_bounds__(256) k_text_len(TextArgs t, int64_t *len) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= t.n_reads) return;
  CountSink c;
  emit_read(c, t, q);
  len[q] = c.n;
}

__global__ void __launch_bounds__(256) k_text_write(TextArgs t, const int64_t *txt_off, char *txt) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= t.n_reads) return;
  WriteSink16 w(txt + txt_off[q]);
  emit_read(w, t, q);
  w.flush();
}

// SEQ and QUAL of every record slot: one wavefront per slot, lanes on
// consecutive bytes (reads of codes and qualities and the stores coalesce)
__global__ void __launch_bounds__(256) k_text_fill(TextArgs t, const int64_t *txt_off, char *txt, int64_t n_slots) {
  const int64_t slot = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (slot >= n_slots) return;
  const TextSlot d = t.slots[slot];
  if (d.read < 0) return;
  const int q = d.read;
  char *dst = txt + txt_off[q] + d.seq_at;
  const uint8_t *sq = t.seq + t.off[q];
  const int len = t.off[q + 1] - t.off[q];
  const char *qual = t.quals + t.qual_off[q];
  const int64_t ql = t.qual_off[q + 1] - t.qual_off[q];
  char *dq = dst + len + 1;
  if (d.rev) {
    for (int k = lane; k < len; k += 64) {
      const uint8_t c = sq[len - 1 - k];
      dst[k] = "TGCAN"[c < 4 ? c : 4];
    }
    for (int64_t k = lane; k < ql; k += 64) dq[k] = qual[ql - 1 - k];
  } else {
    for (int k = lane; k < len; k += 64) {
      const uint8_t c = sq[k];
      dst[k] = "ACGTN"[c < 4 ? c : 4];
    }
    for (int64_t k = lane; k < ql; k += 64) dq[k] = qual[k];
  }
}

void launch_text_fill(const TextArgs &t, const int64_t *txt_off, char *txt, int64_t n_slots, hipStream_t s) {
  if (n_slots > 0)
    hipLaunchKernelGGL(k_text_fill, dim3((unsigned)((n_slots + 3) / 4)), dim3(256), 0, s, t, txt_off, txt, n_slots);
}

void launch_text_len(const TextArgs &t, int64_t *len, hipStream_t s) {
  if (t.n_reads > 0)
    hipLaunchKernelGGL(k_text_len, dim3((unsigned)((t.n_reads + 255) / 256)), dim3(256), 0, s, t, len);
}

void launch_text_write(const TextArgs &t, const int64_t *txt_off, char *txt, hipStream_t s) {
  if (t.n_reads > 0)
    hipLaunchKernelGGL(k_text_write, dim3((unsigned)((t.n_reads + 255) / 256)), dim3(256), 0, s, t, txt_off, txt);
}

// MAPQ and MD-string length of every alignment (thread per alignment;
// aln_read gives the alignment's read)
__global__ void __launch_bounds__(256) k_aln_mapq_md(TextArgs t, int64_t n_alns, const int32_t *aln_read,
                                                     int32_t *mapq, int64_t *md_len) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_alns) return;
  const Aln &a = t.alns[i];
  const int q = aln_read[i];
  const bool rev = a.rb >= t.l_pac;
  int64_t fpos;
  int beg, endc;
  sam_place(t, a, rev, fpos, beg, endc);
  CountSink c;
  emit_md(c, t, a, rev, t.seq + t.off[q], beg, endc, fpos);
  md_len[i] = c.n + 1;  // with its terminating zero
  mapq[i] = mapq_se(t, a.score, a.qb, a.qe, a.rb, a.re);
}

__global__ void __launch_bounds__(256) k_aln_md_write(TextArgs t, int64_t n_alns, const int32_t *aln_read,
                                                      const int64_t *md_off, char *md) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_alns) return;
  const Aln &a = t.alns[i];
  const int q = aln_read[i];
  const bool rev = a.rb >= t.l_pac;
  int64_t fpos;
  int beg, endc;
  sam_place(t, a, rev, fpos, beg, endc);
  WriteSink w{md + md_off[i]};
  emit_md(w, t, a, rev, t.seq + t.off[q], beg, endc, fpos);
  w.put('\0');
}

void launch_aln_mapq_md(const TextArgs &t, int64_t n_alns, const int32_t *aln_read, int32_t *mapq, int64_t *md_len,
                        hipStream_t s) {
  if (n_alns > 0)
    hipLaunchKernelGGL(k_aln_mapq_md, dim3((unsigned)((n_alns + 255) / 256)), dim3(256), 0, s, t, n_alns, aln_read,
                       mapq, md_len);
}

void launch_aln_md_write(const TextArgs &t, int64_t n_alns, const int32_t *aln_read, const int64_t *md_off,
                         char *md, hipStream_t s) {
  if (n_alns > 0)
    hipLaunchKernelGGL(k_aln_md_write, dim3((unsigned)((n_alns + 255) / 256)), dim3(256), 0, s, t, n_alns, aln_read,
                       md_off, md);
}

}  // namespace g3a
