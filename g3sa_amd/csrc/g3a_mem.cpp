// `g3a mem`: host driver keeping the reference's `g3.exe mem` command line
// (bwa/src/fastmap.cpp:416-566).  Reads FASTQ with the reference's field
// rules (bwa/src/pipeline.cpp:111-227), aligns batches through the C ABI and
// writes, formatted on the device, either the reference's records
// (pipeline.cpp:287-351):
//     NAME RNAME POS CIGAR SEQ QUAL        (POS 0-based)
//     NAME * 0 * SEQ QUAL                  (no alignment)
// or, with -S, standard SAM.  Output is in input order.
//
// Options: -o FILE output, -Z N reads per batch, -S standard SAM, -g N GPUs
// (1 only), -t N parsing threads, --learned the learned-index seeding form
// (same output); -b -v -i -z -l -p -F -L -m are accepted for compatibility
// with the reference's command line and ignored.
#include "../../include/g3a.h"
#include "g3a_index.h"

#include <fcntl.h>
#include <malloc.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace {

struct Batch {
  int64_t id = 0;  // position in the input, for ordered output
  std::string names, quals;
  std::vector<int64_t> name_off{0}, qual_off{0};
  std::vector<uint8_t> seq;
  std::vector<int32_t> seq_off{0};
  std::vector<char> text;  // the batch's records, formatted on the device
  int64_t text_bytes = 0;
  int n() const { return (int)seq_off.size() - 1; }
  void clear() {
    names.clear(); quals.clear(); seq.clear();
    name_off.assign(1, 0); qual_off.assign(1, 0); seq_off.assign(1, 0);
  }
};

// A0 C1 G2 T3 (either case), anything else 4
struct Nt4Table {
  uint8_t v[256];
  Nt4Table() {
    std::memset(v, 4, sizeof v);
    v['A'] = v['a'] = 0;
    v['C'] = v['c'] = 1;
    v['G'] = v['g'] = 2;
    v['T'] = v['t'] = 3;
  }
};

uint8_t nt4(int c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
  }
}

// Streaming FASTQ reader: the file is read in chunks of kChunk bytes and
// parsed record by record (NAME runs to the first space or end of line, then
// the rest of the header line, SEQ, the separator line and QUAL); a record
// cut by a chunk boundary is completed from the next chunk.
struct FastqReader {
  size_t kChunk = 64u << 20;  // G3A_FASTQ_CHUNK overrides (bytes) and disables the mapping
  FILE *fp = nullptr;
  std::vector<char> buf;
  const char *map = nullptr;  // whole regular file mapped read-only
  size_t map_len = 0;
  size_t pos = 0, len = 0;
  bool eof = false;
  ~FastqReader() {
    if (fp) std::fclose(fp);
    if (map) munmap((void *)map, map_len);
  }
  const char *data() const { return map ? map : buf.data(); }
  bool open(const char *fn) {
    bool chunked = false;
    if (const char *e = std::getenv("G3A_FASTQ_CHUNK")) {
      const long long v = std::atoll(e);
      if (v >= 16) kChunk = (size_t)v;
      chunked = true;
    }
    if (!chunked) {
      const int fd = ::open(fn, O_RDONLY);
      struct stat st;
      if (fd >= 0 && fstat(fd, &st) == 0 && S_ISREG(st.st_mode) && st.st_size > 0) {
        void *m = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m != MAP_FAILED) {
          madvise(m, (size_t)st.st_size, MADV_SEQUENTIAL);
          map = (const char *)m;
          map_len = len = (size_t)st.st_size;
          eof = true;
        }
      }
      if (fd >= 0) ::close(fd);
      if (map) {
        skip_to_record();
        return true;
      }
    }
    fp = std::fopen(fn, "rb");
    if (!fp) return false;
    fill();
    skip_to_record();
    return true;
  }
  // append more of the file after the unparsed tail
  void fill() {
    if (eof) return;
    if (pos > 0) {
      std::memmove(buf.data(), buf.data() + pos, len - pos);
      len -= pos;
      pos = 0;
    }
    if (buf.size() < len + kChunk) buf.resize(len + kChunk);
    const size_t got = std::fread(buf.data() + len, 1, kChunk, fp);
    len += got;
    if (got < kChunk) eof = true;
  }
  void skip_to_record() {
    for (;;) {
      while (pos < len && data()[pos] != '@') ++pos;
      if (pos < len || eof) return;
      fill();
    }
  }
  // a whole record (four lines) is buffered, or the file ends inside it
  bool have_record() {
    for (;;) {
      const char *p = data() + pos, *e = data() + len;
      int nl = 0;
      while (p < e) {
        const char *q = (const char *)std::memchr(p, '\n', (size_t)(e - p));
        if (!q) break;
        if (++nl == 4) return true;
        p = q + 1;
      }
      if (eof) return pos < len;
      fill();
    }
  }
  // end of the line starting at i (the newline, or len)
  size_t eol(size_t i) const {
    const char *q = (const char *)std::memchr(data() + i, '\n', len > i ? len - i : 0);
    return q ? (size_t)(q - data()) : len;
  }
  bool next(Batch &b) {
    static const Nt4Table tab;
    if (pos >= len && eof) return false;
    if (!have_record()) return false;
    const char *p = data();
    // header: NAME up to the first space or the end of the line
    size_t q = ++pos;
    const size_t h_end = eol(q);
    const char *sp = (const char *)std::memchr(p + q, ' ', h_end - q);
    const size_t n_end = sp ? (size_t)(sp - p) : h_end;
    b.names.append(p + q, n_end - q);
    b.name_off.push_back((int64_t)b.names.size());
    // sequence
    q = h_end + 1 < len ? h_end + 1 : len;
    const size_t s_end = eol(q);
    const size_t old = b.seq.size();
    b.seq.resize(old + (s_end - q));
    uint8_t *d = b.seq.data() + old;
    for (size_t i = q; i < s_end; ++i) *d++ = tab.v[(unsigned char)p[i]];
    b.seq_off.push_back((int32_t)b.seq.size());
    // separator line, then qualities
    q = s_end + 1 < len ? s_end + 1 : len;
    const size_t p_end = eol(q);
    q = p_end + 1 < len ? p_end + 1 : len;
    const size_t q_end = eol(q);
    if (q < len) b.quals.append(p + q, q_end - q);
    b.qual_off.push_back((int64_t)b.quals.size());
    pos = q_end + 1 < len ? q_end + 1 : len;
    skip_to_record();
    return true;
  }

  // Up to max_reads records and max_bases bases into b (a record longer
  // than max_bases is taken alone), parsed by `threads` threads: a pass over
  // the buffer finds the records' line ends, then the records found are split
  // into equal parts whose lengths are summed and whose fields are copied in
  // parallel; the buffer is refilled until the batch is full or the file ends.
  void next_batch(Batch &b, int max_reads, int64_t max_bases, int threads) {
    struct Rec {
      size_t h, s, p, q, e;  // header, sequence, separator, quality starts; end
      size_t name_end;
    };
    std::vector<Rec> recs;
    bool full = false;
    for (;;) {
      recs.clear();
      const int want = max_reads - b.n();
      int64_t bases = (int64_t)b.seq.size();
      while ((int)recs.size() < want && pos < len) {
        size_t ends[4], i = pos;
        int k = 0;
        for (; k < 4; ++k) {
          const char *q = (const char *)std::memchr(data() + i, '\n', len - i);
          if (!q) break;
          ends[k] = (size_t)(q - data());
          i = ends[k] + 1;
        }
        if (k < 4) break;  // cut by the end of the buffer
        const int64_t sl = (int64_t)(ends[1] - ends[0] - 1);
        if (bases + sl > max_bases && b.n() + recs.size() > 0) {
          full = true;
          break;
        }
        bases += sl;
        Rec r;
        r.h = pos + 1;
        r.s = ends[0] + 1;
        r.p = ends[1] + 1;
        r.q = ends[2] + 1;
        r.e = ends[3];
        recs.push_back(r);
        pos = ends[3] + 1;
        while (pos < len && data()[pos] != '@') ++pos;
      }
      copy_records(b, recs, threads);
      if (b.n() >= max_reads || full) return;
      if (eof) break;
      fill();
      skip_to_record();
    }
    // the file's last record without a final newline
    while (b.n() < max_reads && (int64_t)b.seq.size() < max_bases && next(b)) {}
  }

 private:
  template <class Rec>
  void copy_records(Batch &b, std::vector<Rec> &recs, int threads) {
    const size_t n0 = (size_t)b.n(), n = recs.size();
    if (!n) return;
    std::vector<int64_t> nl(n + 1, 0), sl(n + 1, 0), ql(n + 1, 0);
    const char *p = data();
    const int nt = n < 20000 ? 1 : threads;
    auto run = [&](auto fn) {
      std::vector<std::thread> th;
      for (int t = 1; t < nt; ++t) th.emplace_back(fn, t);
      fn(0);
      for (auto &x : th) x.join();
    };
    run([&](int t) {
      const size_t r0 = n * (size_t)t / (size_t)nt, r1 = n * (size_t)(t + 1) / (size_t)nt;
      for (size_t i = r0; i < r1; ++i) {
        Rec &x = recs[i];
        const size_t h_end = x.s - 1;
        const char *sp = (const char *)std::memchr(p + x.h, ' ', h_end - x.h);
        x.name_end = sp ? (size_t)(sp - p) : h_end;
        nl[i + 1] = (int64_t)(x.name_end - x.h);
        sl[i + 1] = (int64_t)(x.p - 1 - x.s);
        ql[i + 1] = (int64_t)(x.e - x.q);
      }
    });
    for (size_t i = 0; i < n; ++i) {
      nl[i + 1] += nl[i];
      sl[i + 1] += sl[i];
      ql[i + 1] += ql[i];
    }
    const size_t nb0 = b.names.size(), sb0 = b.seq.size(), qb0 = b.quals.size();
    b.names.resize(nb0 + (size_t)nl[n]);
    b.seq.resize(sb0 + (size_t)sl[n]);
    b.quals.resize(qb0 + (size_t)ql[n]);
    b.name_off.resize(n0 + n + 1);
    b.seq_off.resize(n0 + n + 1);
    b.qual_off.resize(n0 + n + 1);
    static const Nt4Table tab;
    run([&](int t) {
      const size_t r0 = n * (size_t)t / (size_t)nt, r1 = n * (size_t)(t + 1) / (size_t)nt;
      for (size_t i = r0; i < r1; ++i) {
        const Rec &x = recs[i];
        std::memcpy(&b.names[nb0 + (size_t)nl[i]], p + x.h, x.name_end - x.h);
        uint8_t *d = b.seq.data() + sb0 + sl[i];
        for (size_t k = x.s; k + 1 < x.p; ++k) *d++ = tab.v[(unsigned char)p[k]];
        std::memcpy(&b.quals[qb0 + (size_t)ql[i]], p + x.q, x.e - x.q);
        b.name_off[n0 + i + 1] = (int64_t)(nb0 + nl[i + 1]);
        b.seq_off[n0 + i + 1] = (int32_t)(sb0 + sl[i + 1]);
        b.qual_off[n0 + i + 1] = (int64_t)(qb0 + ql[i + 1]);
      }
    });
  }
};

// Bounded FIFO between pipeline stages.
template <class T>
class Queue {
 public:
  explicit Queue(size_t cap) : cap_(cap) {}
  void push(T v) {
    std::unique_lock<std::mutex> l(m_);
    not_full_.wait(l, [&] { return q_.size() < cap_; });
    q_.push_back(std::move(v));
    not_empty_.notify_one();
  }
  T pop() {
    std::unique_lock<std::mutex> l(m_);
    not_empty_.wait(l, [&] { return !q_.empty(); });
    T v = std::move(q_.front());
    q_.pop_front();
    not_full_.notify_one();
    return v;
  }

 private:
  size_t cap_;
  std::deque<T> q_;
  std::mutex m_;
  std::condition_variable not_full_, not_empty_;
};

void put_int(std::string &s, long long v) {
  char t[24];
  char *e = t + sizeof t, *p = e;
  unsigned long long u = v < 0 ? 0ull - (unsigned long long)v : (unsigned long long)v;
  do {
    *--p = (char)('0' + u % 10);
    u /= 10;
  } while (u);
  if (v < 0) *--p = '-';
  s.append(p, (size_t)(e - p));
}

struct Ref {
  std::vector<std::string> names;
  std::vector<int64_t> offs, lens;
  std::vector<uint8_t> pac;
  int64_t l_pac = 0;
};

bool load_ref(const char *prefix, Ref &ref) {
  g3a::HostIndex h;
  std::string fn = std::string(prefix) + ".ann";
  FILE *fp = std::fopen(fn.c_str(), "r");
  if (!fp) return false;
  long long l_pac = 0;
  int n = 0;
  unsigned seed = 0;
  if (std::fscanf(fp, "%lld %d %u", &l_pac, &n, &seed) != 3) { std::fclose(fp); return false; }
  ref.l_pac = l_pac;
  char line[65536];
  if (!std::fgets(line, sizeof line, fp)) { std::fclose(fp); return false; }
  for (int i = 0; i < n; ++i) {
    if (!std::fgets(line, sizeof line, fp)) { std::fclose(fp); return false; }
    char *p = std::strchr(line, ' ');
    if (!p) { std::fclose(fp); return false; }
    ++p;
    std::string nm;
    while (*p && *p != ' ' && *p != '\n' && *p != '\r') nm.push_back(*p++);
    long long off, ln;
    int namb;
    if (!std::fgets(line, sizeof line, fp) || std::sscanf(line, "%lld %lld %d", &off, &ln, &namb) != 3) {
      std::fclose(fp);
      return false;
    }
    ref.names.push_back(nm);
    ref.offs.push_back(off);
    ref.lens.push_back(ln);
  }
  std::fclose(fp);
  fn = std::string(prefix) + ".pac";
  fp = std::fopen(fn.c_str(), "rb");
  if (!fp) return false;
  std::fseek(fp, 0, SEEK_END);
  long sz = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  ref.pac.resize((size_t)sz);
  bool ok = std::fread(ref.pac.data(), 1, ref.pac.size(), fp) == ref.pac.size();
  std::fclose(fp);
  return ok;
}

}  // namespace

int g3a_mem_main(int argc, char **argv) {
  // Batch-sized buffers (parsed batches, the library's output arrays, text
  // parts) are freed and allocated again every batch: keep them in the heap
  // instead of mapping fresh pages each time, since page faults cost more
  // than the work done on the pages.
  mallopt(M_MMAP_THRESHOLD, 1 << 30);
  mallopt(M_TRIM_THRESHOLD, 1 << 30);
  const char *out_fn = nullptr;
  int batch = 250000, gpus = 1, threads = 8;
  bool sam = false, learned = false, graph = true;
  int i = 1;
  for (; i < argc && argv[i][0] == '-' && argv[i][1]; ++i) {
    const char *a = argv[i];
    auto val = [&](void) -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
    if (!std::strcmp(a, "-o")) out_fn = val();
    else if (!std::strcmp(a, "-Z")) { const char *v = val(); batch = v ? std::atoi(v) : batch; }
    else if (!std::strcmp(a, "-g")) { const char *v = val(); gpus = v ? std::atoi(v) : gpus; }
    else if (!std::strcmp(a, "-t")) { const char *v = val(); threads = v ? std::atoi(v) : threads; }
    else if (!std::strcmp(a, "-S")) sam = true;
    else if (!std::strcmp(a, "--learned")) learned = true;
    else if (!std::strcmp(a, "--no-graph")) graph = false;
    else if (!std::strcmp(a, "-b")) {}
    else if (!std::strcmp(a, "-v") || !std::strcmp(a, "-i") || !std::strcmp(a, "-z") || !std::strcmp(a, "-l") ||
             !std::strcmp(a, "-p") || !std::strcmp(a, "-F") || !std::strcmp(a, "-L") || !std::strcmp(a, "-m"))
      val();
    else {
      std::fprintf(stderr, "g3a mem: unknown option %s\n", a);
      return 1;
    }
  }
  if (i + 2 > argc) {
    std::fprintf(stderr,
                 "usage: g3a mem [-o out] [-Z reads_per_batch] [-S] [-t threads] [--learned] [--no-graph] <prefix> <reads.fq>\n"
                 "  each batch runs as one HIP graph unless --no-graph\n");
    return 1;
  }
  if (gpus != 1) {
    std::fprintf(stderr, "g3a mem: this build drives one GPU (-g 1)\n");
    return 1;
  }
  if (batch < 1) batch = 1;
  if (threads < 1) threads = 1;
  // a batch closes at `batch` reads or 320 bases per read on average; a
  // longer read (left unmapped, as reads over 320 bases are) goes alone
  const int64_t batch_bases = (int64_t)batch * 320, ctx_bases = batch_bases + (1 << 20);
  const char *prefix = argv[i], *fq = argv[i + 1];
  Ref ref;
  if (!load_ref(prefix, ref)) {
    std::fprintf(stderr, "g3a mem: cannot read %s.ann/.pac\n", prefix);
    return 1;
  }
  FastqReader rd;
  if (!rd.open(fq)) {
    std::fprintf(stderr, "g3a mem: cannot read %s\n", fq);
    return 1;
  }
  g3a_opt opt;
  g3a_opt_init(&opt);
  // two contexts on the GPU, each with its own stream and workspace: while
  // one batch's kernels run, the other context moves its batch to or from
  // the host
  const int kCtx = 2;
  std::vector<g3a_ctx *> ctxs;
  auto destroy_all = [&] {
    for (g3a_ctx *c : ctxs) g3a_destroy(c);
  };
  // the first context loads the index (and builds the learned index); the
  // second shares both (g3a_device_alloc_sibling)
  for (int k = 0; k < kCtx; ++k) {
    g3a_ctx *c = nullptr;
    int rc = k == 0 ? g3a_device_alloc(0, prefix, &opt, batch, ctx_bases, &c)
                    : g3a_device_alloc_sibling(ctxs[0], batch, ctx_bases, &c);
    if (rc) {
      if (k == 0) std::fprintf(stderr, "g3a mem: cannot create context (%d)\n", rc);
      else std::fprintf(stderr, "g3a mem: cannot create context: %s\n", g3a_last_error(ctxs[0]));
      destroy_all();
      return 1;
    }
    ctxs.push_back(c);
    if (k > 0) continue;
    if ((rc = g3a_memcpy_index(c)) || (rc = g3a_set_graph(c, graph ? 1 : 0))) {
      std::fprintf(stderr, "g3a mem: %s\n", g3a_last_error(c));
      destroy_all();
      return 1;
    }
    if (learned) {
      g3a_learned_info li;
      if ((rc = g3a_set_learned(c, 1, &li))) {
        std::fprintf(stderr, "g3a mem: %s\n", g3a_last_error(c));
        destroy_all();
        return 1;
      }
      std::fprintf(stderr,
                   "g3a mem: learned index: %d leaves x %d hidden units, model %lld bytes, max error %lld rows "
                   "(mean %.1f), %lld keys beyond the %d-row window, built in %.0f ms\n",
                   li.leaves, li.hidden_units, (long long)li.model_bytes, (long long)li.max_err, li.mean_err,
                   (long long)li.keys_over_window, li.window_cap, li.build_ms);
    }
  }
  FILE *out = out_fn ? std::fopen(out_fn, "wb") : stdout;
  if (!out) {
    std::fprintf(stderr, "g3a mem: cannot write %s\n", out_fn);
    destroy_all();
    return 1;
  }
  if (sam) {
    std::string hdr = "@HD\tVN:1.6\tSO:unsorted\n";
    for (size_t c = 0; c < ref.names.size(); ++c) {
      hdr += "@SQ\tSN:" + ref.names[c] + "\tLN:";
      put_int(hdr, ref.lens[c]);
      hdr += "\n";
    }
    hdr += "@PG\tID:g3a\tPN:g3a\n";
    std::fwrite(hdr.data(), 1, hdr.size(), out);
  }
  // the stages overlap: the loader parses batch k+2 while the aligners
  // (one per context) align batches k+1 and k and the formatter writes
  // batch k-1; the writer puts batches back in input order
  struct Job {
    std::unique_ptr<Batch> b;
  };
  Queue<std::unique_ptr<Batch>> q_in(2);
  Queue<std::unique_ptr<Job>> q_out(4);
  // batches circulate: loader -> aligner -> writer -> loader, so their
  // buffers are allocated (and faulted in) once
  const int kBatches = 7;
  Queue<std::unique_ptr<Batch>> q_free((size_t)kBatches);
  for (int k = 0; k < kBatches; ++k) q_free.push(std::unique_ptr<Batch>(new Batch));
  const auto t0 = std::chrono::steady_clock::now();
  using clk = std::chrono::steady_clock;
  auto since = [](clk::time_point t) { return std::chrono::duration<double>(clk::now() - t).count(); };
  double t_load = 0, t_align = 0, t_write = 0;  // busy time of each stage
  double t_align_first = 0;
  int n_batches = 0;
  int64_t n_loaded = 0;
  std::mutex stat_m;
  std::thread loader([&] {
    for (;;) {
      std::unique_ptr<Batch> b = q_free.pop();
      const auto tl = clk::now();
      b->clear();
      rd.next_batch(*b, batch, batch_bases, threads);
      b->id = n_loaded++;
      t_load += since(tl);
      if (b->n() == 0) {
        q_in.push(nullptr);
        return;
      }
      q_in.push(std::move(b));
    }
  });
  long long total = 0;
  std::thread writer([&] {
    std::map<int64_t, std::unique_ptr<Job>> pending;  // aligned out of order
    int64_t next_id = 0;
    for (;;) {
      std::unique_ptr<Job> j;
      auto it = pending.find(next_id);
      if (it != pending.end()) {
        j = std::move(it->second);
        pending.erase(it);
      } else {
        j = q_out.pop();
        if (!j) return;
        if (j->b->id != next_id) {
          const int64_t id = j->b->id;
          pending[id] = std::move(j);
          continue;
        }
      }
      ++next_id;
      const auto tw = clk::now();
      const Batch &b = *j->b;
      std::fwrite(b.text.data(), 1, (size_t)b.text_bytes, out);
      total += b.n();
      t_write += since(tw);
      q_free.push(std::move(j->b));
    }
  });
  std::atomic<int> status{0};
  auto aligner = [&](g3a_ctx *ctx) {
    for (;;) {
      std::unique_ptr<Batch> b = q_in.pop();
      if (!b) {
        q_in.push(nullptr);  // for the other aligners
        return;
      }
      if (status) {  // drain the loader after a failure
        q_free.push(std::move(b));
        continue;
      }
      std::unique_ptr<Job> j(new Job);
      const auto ta = clk::now();
      // input, names and qualities to the device; align; format the
      // records there; copy the text back
      int rc = g3a_memcpy_input(ctx, b->n(), b->seq.data(), b->seq_off.data());
      if (!rc) rc = g3a_memcpy_text(ctx, b->n(), b->names.data(), b->name_off.data(), b->quals.data(),
                                    b->qual_off.data());
      if (!rc) rc = g3a_bwa_align(ctx);
      if (!rc) rc = g3a_format_records(ctx, sam ? 1 : 0, &b->text_bytes);
      if (!rc) {
        if ((int64_t)b->text.size() < b->text_bytes) b->text.resize((size_t)(b->text_bytes + b->text_bytes / 8));
        rc = g3a_copy_text(ctx, b->text.data(), (int64_t)b->text.size());
      }
      const double dt = since(ta);
      {
        std::lock_guard<std::mutex> l(stat_m);
        t_align += dt;
        if (n_batches++ == 0) t_align_first = dt;
      }
      if (rc) {
        std::fprintf(stderr, "g3a mem: %s\n", g3a_last_error(ctx));
        status = 1;
        q_free.push(std::move(b));
        continue;
      }
      j->b = std::move(b);
      q_out.push(std::move(j));
    }
  };
  std::vector<std::thread> aligners;
  for (g3a_ctx *c : ctxs) aligners.emplace_back(aligner, c);
  for (auto &t : aligners) t.join();
  q_out.push(nullptr);
  loader.join();
  writer.join();
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::fprintf(stderr, "g3a mem: %lld reads in %.2f s (%.0f reads/s end to end)\n", total, sec,
               sec > 0 ? total / sec : 0.0);
  std::fprintf(stderr, "g3a mem: busy time: load %.2f s, align+format %.2f s, write %.2f s\n", t_load, t_align,
               t_write);
  std::fprintf(stderr, "g3a mem: %d batches; align of the first %.3f s, of the others %.3f s each\n", n_batches,
               t_align_first, n_batches > 1 ? (t_align - t_align_first) / (n_batches - 1) : 0.0);
  if (out_fn) std::fclose(out);
  destroy_all();
  return status;
}
