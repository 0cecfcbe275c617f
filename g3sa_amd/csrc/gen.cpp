// g3a-gen: deterministic synthetic references and short reads for the BWA-MEM
// hot path.  Every value is drawn from xoshiro256** seeded through splitmix64,
// so a (command line, seed) pair always produces the same bytes.
//
//   g3a-gen ref   -o ref.fa  --len L --contigs C --seed S [--gc-mean 0.5 --gc-sd 0
//                  --repeat-frac F --families N --fam-min A --fam-max B
//                  --div-min x --div-max y --tandem-frac t]
//   g3a-gen reads -r ref.fa -o reads.fq --n N --seed S [--len 150 | --len-min a --len-max b]
//                 [--paired --insert-mean 400 --insert-sd 50]
//                 [--sub 0.01 | --sub-start a --sub-end b] [--indel 0.0005]
//                 [--hard-frac 0.05 --hard-div 0.08 --tail-min 20 --tail-max 40]
//                 [--err-min a --err-max b]
//
// FASTQ records are written as "@<name> c" so that a parser which stops the
// read name at the first space (bwa/src/pipeline.cpp:191) sees "<name>", and
// qualities are all 'I' so that no quality line starts with '@'
// (bwa/src/pipeline.cpp:166-173 resynchronises chunks on '@').
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto &v : s) v = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform double in [0,1) with 53 random bits
  double uni() { return (next() >> 11) * 0x1.0p-53; }
  // uniform integer in [0, n)
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
  // uniform integer in [a, b]
  int64_t range(int64_t a, int64_t b) { return a + (int64_t)below((uint64_t)(b - a + 1)); }
  double normal(double mu, double sd) {
    double u1 = uni(), u2 = uni();
    if (u1 < 1e-300) u1 = 1e-300;
    return mu + sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

const char kBase[4] = {'A', 'C', 'G', 'T'};

int code_of(char c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
  }
}

struct Args {
  std::vector<std::string> v;
  const char *get(const char *key, const char *def = nullptr) const {
    for (size_t i = 0; i + 1 < v.size(); ++i)
      if (v[i] == key) return v[i + 1].c_str();
    return def;
  }
  bool has(const char *key) const {
    for (auto &x : v) if (x == key) return true;
    return false;
  }
  double num(const char *key, double def) const {
    const char *s = get(key);
    return s ? std::atof(s) : def;
  }
};

// Base with GC probability g: P(C)=P(G)=g/2, P(A)=P(T)=(1-g)/2.
int gc_base(Rng &r, double g) {
  double u = r.uni();
  if (u < g * 0.5) return 1;
  if (u < g) return 2;
  if (u < g + (1.0 - g) * 0.5) return 0;
  return 3;
}

// Copy src with substitutions at rate div and small indels at rate div/10.
std::vector<uint8_t> diverge(Rng &r, const std::vector<uint8_t> &src, double div) {
  std::vector<uint8_t> out;
  out.reserve(src.size() + 16);
  double indel = div * 0.1;
  for (size_t i = 0; i < src.size(); ++i) {
    double u = r.uni();
    if (u < indel * 0.5) {            // deletion of 1..3 bases
      i += (size_t)r.range(0, 2);
      continue;
    } else if (u < indel) {           // insertion of 1..3 random bases
      int k = (int)r.range(1, 3);
      for (int j = 0; j < k; ++j) out.push_back((uint8_t)r.below(4));
    }
    uint8_t b = src[i];
    if (r.uni() < div) b = (uint8_t)((b + 1 + r.below(3)) & 3);
    out.push_back(b);
  }
  return out;
}

int cmd_ref(const Args &a) {
  const char *out = a.get("-o");
  if (!out) { std::fprintf(stderr, "g3a-gen ref: -o required\n"); return 1; }
  int64_t L = (int64_t)a.num("--len", 1000000);
  int C = (int)a.num("--contigs", 1);
  uint64_t seed = (uint64_t)a.num("--seed", 1);
  double gc_mean = a.num("--gc-mean", 0.5), gc_sd = a.num("--gc-sd", 0.0);
  double rep_frac = a.num("--repeat-frac", 0.0);
  int n_fam = (int)a.num("--families", 200);
  int fam_min = (int)a.num("--fam-min", 300), fam_max = (int)a.num("--fam-max", 6000);
  double div_min = a.num("--div-min", 0.05), div_max = a.num("--div-max", 0.15);
  double tandem_frac = a.num("--tandem-frac", 0.0);
  if (L < 1 || C < 1 || L % C) { std::fprintf(stderr, "g3a-gen ref: --len must be a positive multiple of --contigs\n"); return 1; }
  Rng r(seed);
  std::vector<uint8_t> g((size_t)L);
  // Background: i.i.d. bases whose GC fraction is redrawn every 10 kb window.
  for (int64_t w = 0; w < L; w += 10000) {
    double gc = gc_sd > 0 ? r.normal(gc_mean, gc_sd) : gc_mean;
    if (gc < 0.2 && gc_sd > 0) gc = 0.2;
    if (gc > 0.8 && gc_sd > 0) gc = 0.8;
    int64_t e = w + 10000 < L ? w + 10000 : L;
    for (int64_t i = w; i < e; ++i) g[(size_t)i] = (uint8_t)(gc_sd > 0 || gc_mean != 0.5 ? gc_base(r, gc) : r.below(4));
  }
  // Interspersed repeats: copies of family consensus sequences, each diverged.
  if (rep_frac > 0 && n_fam > 0) {
    std::vector<std::vector<uint8_t>> fam((size_t)n_fam);
    for (auto &f : fam) {
      f.resize((size_t)r.range(fam_min, fam_max));
      for (auto &b : f) b = (uint8_t)gc_base(r, gc_mean);
    }
    int64_t target = (int64_t)(rep_frac * (double)L), covered = 0;
    while (covered < target) {
      const auto &f = fam[(size_t)r.below((uint64_t)n_fam)];
      double div = div_min + (div_max - div_min) * r.uni();
      std::vector<uint8_t> cp = diverge(r, f, div);
      if (r.below(2)) {  // reverse-complement copy
        std::vector<uint8_t> rc(cp.rbegin(), cp.rend());
        for (auto &b : rc) b = (uint8_t)(3 - b);
        cp.swap(rc);
      }
      if ((int64_t)cp.size() >= L) continue;
      int64_t pos = (int64_t)r.below((uint64_t)(L - (int64_t)cp.size()));
      std::memcpy(&g[(size_t)pos], cp.data(), cp.size());
      covered += (int64_t)cp.size();
    }
  }
  // Tandem repeats: short units (2..6 bp) repeated with light divergence.
  if (tandem_frac > 0) {
    int64_t target = (int64_t)(tandem_frac * (double)L), covered = 0;
    while (covered < target) {
      int unit = (int)r.range(2, 6);
      int64_t len = r.range(20, 300);
      if (len >= L) break;
      int64_t pos = (int64_t)r.below((uint64_t)(L - len));
      uint8_t u[6];
      for (int i = 0; i < unit; ++i) u[i] = (uint8_t)r.below(4);
      for (int64_t i = 0; i < len; ++i) {
        uint8_t b = u[i % unit];
        if (r.uni() < 0.02) b = (uint8_t)r.below(4);
        g[(size_t)(pos + i)] = b;
      }
      covered += len;
    }
  }
  FILE *fp = std::fopen(out, "w");
  if (!fp) { std::perror(out); return 1; }
  int64_t clen = L / C;
  std::vector<char> line(61);
  for (int c = 0; c < C; ++c) {
    std::fprintf(fp, ">chr%d\n", c + 1);
    for (int64_t i = 0; i < clen; i += 60) {
      int64_t e = i + 60 < clen ? i + 60 : clen;
      int k = 0;
      for (int64_t j = i; j < e; ++j) line[(size_t)k++] = kBase[g[(size_t)(c * clen + j)]];
      line[(size_t)k++] = '\n';
      std::fwrite(line.data(), 1, (size_t)k, fp);
    }
  }
  std::fclose(fp);
  return 0;
}

struct Contig { std::string name; std::vector<uint8_t> s; };

std::vector<Contig> read_fasta(const char *fn) {
  std::vector<Contig> v;
  FILE *fp = std::fopen(fn, "r");
  if (!fp) { std::perror(fn); std::exit(1); }
  char buf[1 << 16];
  while (std::fgets(buf, sizeof buf, fp)) {
    if (buf[0] == '>') {
      v.push_back(Contig());
      char *p = buf + 1;
      while (*p && *p != ' ' && *p != '\t' && *p != '\n') v.back().name.push_back(*p++);
    } else if (!v.empty()) {
      for (char *p = buf; *p && *p != '\n'; ++p) {
        int c = code_of(*p);
        v.back().s.push_back((uint8_t)(c < 4 ? c : 0));
      }
    }
  }
  std::fclose(fp);
  return v;
}

// Emit one read of length len copied from s[start..] with substitution rate
// rising linearly from sub0 to sub1 along the read and indel events at rate
// indel (insertion/deletion 50/50, length U{1,2,3}).  Returns false if the
// source ran out before len bases were produced.
bool sample_read(Rng &r, const std::vector<uint8_t> &s, int64_t start, int len,
                 double sub0, double sub1, double indel, std::vector<uint8_t> &out) {
  out.clear();
  int64_t p = start;
  while ((int)out.size() < len) {
    double u = r.uni();
    if (u < indel * 0.5) {
      p += r.range(1, 3);
    } else if (u < indel) {
      int k = (int)r.range(1, 3);
      for (int j = 0; j < k && (int)out.size() < len; ++j) out.push_back((uint8_t)r.below(4));
      if ((int)out.size() >= len) break;
    }
    if (p >= (int64_t)s.size()) return false;
    double pr = len > 1 ? sub0 + (sub1 - sub0) * (double)out.size() / (double)(len - 1) : sub0;
    uint8_t b = s[(size_t)p++];
    if (r.uni() < pr) b = (uint8_t)((b + 1 + r.below(3)) & 3);
    out.push_back(b);
  }
  return true;
}

void revcomp(std::vector<uint8_t> &v) {
  std::vector<uint8_t> t(v.rbegin(), v.rend());
  for (auto &b : t) b = (uint8_t)(3 - b);
  v.swap(t);
}

void write_fq(FILE *fp, const char *name, const std::vector<uint8_t> &s) {
  std::string rec;
  rec.reserve(s.size() * 2 + 64);
  rec += '@'; rec += name; rec += " c\n";
  for (auto b : s) rec += kBase[b];
  rec += "\n+\n";
  rec.append(s.size(), 'I');
  rec += '\n';
  std::fwrite(rec.data(), 1, rec.size(), fp);
}

int cmd_reads(const Args &a) {
  const char *ref = a.get("-r"), *out = a.get("-o");
  if (!ref || !out) { std::fprintf(stderr, "g3a-gen reads: -r and -o required\n"); return 1; }
  int64_t n = (int64_t)a.num("--n", 10000);
  uint64_t seed = (uint64_t)a.num("--seed", 2);
  int len_fixed = (int)a.num("--len", 150);
  int len_min = (int)a.num("--len-min", len_fixed), len_max = (int)a.num("--len-max", len_fixed);
  bool paired = a.has("--paired");
  double ins_mean = a.num("--insert-mean", 400), ins_sd = a.num("--insert-sd", 50);
  double sub = a.num("--sub", 0.01);
  double sub0 = a.num("--sub-start", sub), sub1 = a.num("--sub-end", sub);
  double indel = a.num("--indel", 0.0005);
  double hard_frac = a.num("--hard-frac", 0.0), hard_div = a.num("--hard-div", 0.08);
  int tail_min = (int)a.num("--tail-min", 20), tail_max = (int)a.num("--tail-max", 40);
  double err_min = a.num("--err-min", -1), err_max = a.num("--err-max", -1);
  const char *truth_fn = a.get("--truth");
  FILE *tfp = truth_fn ? std::fopen(truth_fn, "w") : nullptr;
  std::vector<Contig> ctg = read_fasta(ref);
  if (ctg.empty()) { std::fprintf(stderr, "g3a-gen reads: empty reference\n"); return 1; }
  int64_t total = 0;
  std::vector<int64_t> cum;
  for (auto &c : ctg) { total += (int64_t)c.s.size(); cum.push_back(total); }
  Rng r(seed);
  FILE *fp = std::fopen(out, "w");
  if (!fp) { std::perror(out); return 1; }
  std::vector<uint8_t> rd, rd2, frag;
  char name[64];
  int64_t made = 0;
  while (made < n) {
    // contig chosen proportionally to its length
    int64_t x = (int64_t)r.below((uint64_t)total);
    size_t ci = 0;
    while (cum[ci] <= x) ++ci;
    const std::vector<uint8_t> &s = ctg[ci].s;
    double s0 = sub0, s1 = sub1;
    if (err_min >= 0) { double e = err_min + (err_max - err_min) * r.uni(); s0 = s1 = e; }
    if (!paired) {
      int len = (int)r.range(len_min, len_max);
      bool hard = hard_frac > 0 && r.uni() < hard_frac;
      int tail = hard ? (int)r.range(tail_min, tail_max) : 0;
      int body = len - tail;
      if (body < 1) body = 1;
      if ((int64_t)s.size() <= body + 16) continue;
      int64_t start = (int64_t)r.below((uint64_t)((int64_t)s.size() - body - 15));
      if (!sample_read(r, s, start, body, hard ? hard_div : s0, hard ? hard_div : s1, indel, rd)) continue;
      for (int i = 0; i < tail; ++i) rd.push_back((uint8_t)r.below(4));
      bool rc = r.below(2);
      if (rc) revcomp(rd);
      std::snprintf(name, sizeof name, "r%lld", (long long)made);
      write_fq(fp, name, rd);
      if (tfp) std::fprintf(tfp, "%s\t%s\t%lld\t%c\n", name, ctg[ci].name.c_str(), (long long)start, rc ? '-' : '+');
      ++made;
    } else {
      int len = len_fixed;
      int ins = (int)std::lround(r.normal(ins_mean, ins_sd));
      if (ins < len) ins = len;
      if ((int64_t)s.size() <= ins + 64) continue;
      // fragment s[start, start+ins) with 30 bp of slack on both sides so
      // that deletions near either read end still find reference bases
      int64_t start = 30 + (int64_t)r.below((uint64_t)((int64_t)s.size() - ins - 61));
      frag.assign(s.begin() + (start - 30), s.begin() + (start + ins + 30));
      if (r.below(2)) revcomp(frag);  // fragment from the reverse strand
      if (!sample_read(r, frag, 30, len, s0, s1, indel, rd)) continue;
      std::vector<uint8_t> fr2(frag);
      revcomp(fr2);  // mate reads the other end, opposite strand
      if (!sample_read(r, fr2, 30, len, s0, s1, indel, rd2)) continue;
      std::snprintf(name, sizeof name, "r%lld/1", (long long)(made / 2));
      write_fq(fp, name, rd);
      std::snprintf(name, sizeof name, "r%lld/2", (long long)(made / 2));
      write_fq(fp, name, rd2);
      made += 2;
    }
  }
  std::fclose(fp);
  if (tfp) std::fclose(tfp);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: g3a-gen ref|reads [options]\n");
    return 1;
  }
  Args a;
  for (int i = 2; i < argc; ++i) a.v.push_back(argv[i]);
  std::string cmd = argv[1];
  if (cmd == "ref") return cmd_ref(a);
  if (cmd == "reads") return cmd_reads(a);
  std::fprintf(stderr, "g3a-gen: unknown command %s\n", argv[1]);
  return 1;
}
