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
  WriteSink w{txt + txt_off[q]};
  emit_read(w, t, q);
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
