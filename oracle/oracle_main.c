/*
 * g3a_oracle: CPU oracle driver (TEST INFRASTRUCTURE ONLY, see oracle.h).
 *
 *   g3a_oracle [-t threads] [-m compat|sam] [-n max_reads] [-o out] <prefix> <reads.fq>
 *
 * Parses FASTQ with the reference's field rules (bwa/src/pipeline.cpp:111-227:
 * the name ends at the first space, codes from encode_nt4 at
 * pipeline.cpp:90-107), aligns with orc_align_batch and writes either the
 * reference's text records (pipeline.cpp:287-351) or standard SAM.  Timing of
 * the alignment alone (index load and parsing excluded) goes to stderr as
 * "oracle_reads_per_s <value> threads <t> reads <n> seconds <s>", and of the
 * record formatting (one thread, output writing excluded) as
 * "oracle_format_seconds <s>".
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
  char *name_buf, *qual_buf;
  uint8_t *seq;
  int64_t *name_off, *qual_off;
  int32_t *seq_off;
  int n;
} reads_t;

static uint8_t nt4(int c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
  }
}

static int load_fastq(const char *fn, int max_reads, reads_t *r) {
  FILE *fp = fopen(fn, "rb");
  if (!fp) return -1;
  fseek(fp, 0, SEEK_END);
  long sz = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  char *buf = (char *)malloc((size_t)sz + 1);
  if (fread(buf, 1, (size_t)sz, fp) != (size_t)sz) { fclose(fp); free(buf); return -1; }
  fclose(fp);
  buf[sz] = 0;
  int cap = 1024;
  memset(r, 0, sizeof *r);
  r->name_buf = (char *)malloc((size_t)sz + 1);
  r->qual_buf = (char *)malloc((size_t)sz + 1);
  r->seq = (uint8_t *)malloc((size_t)sz + 1);
  r->name_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(cap + 1));
  r->qual_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(cap + 1));
  r->seq_off = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cap + 1));
  int64_t no = 0, qo = 0, so = 0;
  char *p = buf, *end = buf + sz;
  while (p < end && *p != '@') ++p;
  while (p < end && (max_reads <= 0 || r->n < max_reads)) {
    if (r->n == cap) {
      cap *= 2;
      r->name_off = (int64_t *)realloc(r->name_off, sizeof(int64_t) * (size_t)(cap + 1));
      r->qual_off = (int64_t *)realloc(r->qual_off, sizeof(int64_t) * (size_t)(cap + 1));
      r->seq_off = (int32_t *)realloc(r->seq_off, sizeof(int32_t) * (size_t)(cap + 1));
    }
    r->name_off[r->n] = no; r->qual_off[r->n] = qo; r->seq_off[r->n] = (int32_t)so;
    char *b;
    /* NAME: up to the first space (a newline also ends it here, DR-13) */
    for (b = ++p; p < end && *p != ' ' && *p != '\n'; ++p) ;
    memcpy(r->name_buf + no, b, (size_t)(p - b)); no += p - b;
    for (; p < end && *p != '\n'; ++p) ;
    for (b = ++p; p < end && *p != '\n'; ++p) ;
    for (char *c = b; c < p; ++c) r->seq[so++] = nt4((unsigned char)*c);
    for (++p; p < end && *p != '\n'; ++p) ;
    for (b = ++p; p < end && *p != '\n'; ++p) ;
    memcpy(r->qual_buf + qo, b, (size_t)(p - b)); qo += p - b;
    ++r->n;
    ++p;
    while (p < end && *p != '@') ++p;
  }
  r->name_off[r->n] = no; r->qual_off[r->n] = qo; r->seq_off[r->n] = (int32_t)so;
  free(buf);
  return 0;
}

typedef struct {
  char *s;
  size_t n, m;
} str_t;
static void sput(str_t *s, const char *p, size_t k) {
  if (s->n + k + 1 > s->m) {
    while (s->n + k + 1 > s->m) s->m = s->m ? s->m * 2 : 1 << 20;
    s->s = (char *)realloc(s->s, s->m);
  }
  memcpy(s->s + s->n, p, k);
  s->n += k;
}
static void sputs(str_t *s, const char *p) { sput(s, p, strlen(p)); }
static void sputi(str_t *s, long long v) { char t[32]; sput(s, t, (size_t)snprintf(t, sizeof t, "%lld", v)); }
static void sputc(str_t *s, char c) { sput(s, &c, 1); }

static int pac_base(const uint8_t *pac, int64_t l) { return (pac[l >> 2] >> ((~l & 3) << 1)) & 3; }

static uint8_t *load_pac(const char *prefix) {
  char fn[4096];
  snprintf(fn, sizeof fn, "%s.pac", prefix);
  FILE *fp = fopen(fn, "rb");
  if (!fp) return NULL;
  fseek(fp, 0, SEEK_END);
  long sz = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  uint8_t *p = (uint8_t *)malloc((size_t)sz);
  if (fread(p, 1, (size_t)sz, fp) != (size_t)sz) { fclose(fp); free(p); return NULL; }
  fclose(fp);
  return p;
}

int main(int argc, char **argv) {
  int threads = 1, max_reads = 0, sam = 0;
  const char *out_fn = NULL;
  int i = 1;
  for (; i < argc && argv[i][0] == '-'; ++i) {
    if (!strcmp(argv[i], "-t") && i + 1 < argc) threads = atoi(argv[++i]);
    else if (!strcmp(argv[i], "-n") && i + 1 < argc) max_reads = atoi(argv[++i]);
    else if (!strcmp(argv[i], "-o") && i + 1 < argc) out_fn = argv[++i];
    else if (!strcmp(argv[i], "-m") && i + 1 < argc) sam = !strcmp(argv[++i], "sam");
    else { fprintf(stderr, "g3a_oracle: unknown option %s\n", argv[i]); return 1; }
  }
  if (i + 2 > argc) {
    fprintf(stderr, "usage: g3a_oracle [-t threads] [-m compat|sam] [-n max_reads] [-o out] <prefix> <reads.fq>\n");
    return 1;
  }
  const char *prefix = argv[i], *fq = argv[i + 1];
  orc_index *idx = orc_index_load(prefix);
  if (!idx) return 1;
  uint8_t *pac = load_pac(prefix);
  reads_t r;
  if (load_fastq(fq, max_reads, &r) != 0) { fprintf(stderr, "g3a_oracle: cannot read %s\n", fq); return 1; }
  orc_opt opt;
  orc_opt_init(&opt);
  orc_batch b;
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  if (orc_align_batch(idx, &opt, r.n, r.seq, r.seq_off, threads, &b) != 0) {
    fprintf(stderr, "g3a_oracle: alignment failed\n");
    return 1;
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  double sec = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  fprintf(stderr, "oracle_reads_per_s %.3f threads %d reads %d seconds %.6f\n", sec > 0 ? r.n / sec : 0.0, threads,
          r.n, sec);
  fprintf(stderr, "oracle_work fm_steps %lld lf_steps %lld sa_reads %lld ext_tasks %lld ext_cells %lld tb_cells %lld\n",
          (long long)b.fm_steps, (long long)b.lf_steps, (long long)b.sa_reads, (long long)b.ext_tasks,
          (long long)b.ext_cells, (long long)b.tb_cells);
  str_t s = {0, 0, 0};
  const int64_t l_pac = orc_index_l_pac(idx);
  clock_gettime(CLOCK_MONOTONIC, &t0);
  if (sam) {
    sputs(&s, "@HD\tVN:1.6\tSO:unsorted\n");
    for (int c = 0; c < orc_index_n_contigs(idx); ++c) {
      sputs(&s, "@SQ\tSN:"); sputs(&s, orc_index_contig_name(idx, c));
      sputs(&s, "\tLN:"); sputi(&s, orc_index_contig_len(idx, c)); sputc(&s, '\n');
    }
    sputs(&s, "@PG\tID:g3a\tPN:g3a\n");
  }
  uint32_t *core2 = NULL;
  size_t core2_m = 0;
  for (int q = 0; q < r.n; ++q) {
    const char *name = r.name_buf + r.name_off[q];
    size_t nl = (size_t)(r.name_off[q + 1] - r.name_off[q]);
    const char *qual = r.qual_buf + r.qual_off[q];
    size_t ql = (size_t)(r.qual_off[q + 1] - r.qual_off[q]);
    const uint8_t *sq = r.seq + r.seq_off[q];
    int len = r.seq_off[q + 1] - r.seq_off[q];
    int a0 = b.aln_off[q], a1 = b.aln_off[q + 1];
    if (!sam) {
      if (a1 <= a0) {
        sput(&s, name, nl); sputs(&s, " * 0 * ");
        for (int k = 0; k < len; ++k) sputc(&s, "ACGTN"[sq[k]]);
        sputc(&s, ' '); sput(&s, qual, ql); sputc(&s, '\n');
        continue;
      }
      for (int a = a0; a < a1; ++a) {
        const orc_aln *al = &b.alns[a];
        sput(&s, name, nl); sputc(&s, ' ');
        sputs(&s, orc_index_contig_name(idx, al->rid)); sputc(&s, ' ');
        sputi(&s, al->pos); sputc(&s, ' ');
        const uint32_t *cg = b.cigar + al->cigar_off;
        for (int k = 0; k < al->n_cigar; ++k) { sputi(&s, cg[k] >> 4); sputc(&s, "MIDSH"[cg[k] & 0xf]); }
        sputc(&s, ' ');
        for (int k = 0; k < len; ++k) sputc(&s, "ACGTN"[sq[k]]);
        sputc(&s, ' '); sput(&s, qual, ql); sputc(&s, '\n');
      }
      continue;
    }
    if (a1 <= a0) {
      sput(&s, name, nl); sputs(&s, "\t4\t*\t0\t0\t*\t*\t0\t0\t");
      for (int k = 0; k < len; ++k) sputc(&s, "ACGTN"[sq[k]]);
      sputc(&s, '\t'); sput(&s, qual, ql); sputc(&s, '\n');
      continue;
    }
    for (int a = a0; a < a1; ++a) {
      const orc_aln *al = &b.alns[a];
      const uint32_t *core = b.cigar + al->cigar_off + al->n_cigar;
      int nc = al->n_core, rev = al->rb >= l_pac;
      if ((size_t)nc > core2_m) { core2_m = (size_t)nc * 2; core2 = (uint32_t *)realloc(core2, core2_m * 4); }
      for (int k = 0; k < nc; ++k) core2[k] = rev ? core[nc - 1 - k] : core[k];
      int64_t fpos = rev ? (l_pac << 1) - al->re : al->rb; /* forward coordinate of leftmost base */
      int beg = 0, endc = nc;
      if (endc > 0 && (core2[0] & 0xf) == 2) { fpos += core2[0] >> 4; beg = 1; }
      if (endc > beg && (core2[endc - 1] & 0xf) == 2) --endc;
      int clip_l = rev ? len - al->qe : al->qb, clip_r = rev ? al->qb : len - al->qe;
      /* oriented query of the aligned part */
      int lq = al->qe - al->qb;
      uint8_t *oq = (uint8_t *)malloc((size_t)lq + 1);
      for (int k = 0; k < lq; ++k) {
        uint8_t c = rev ? sq[al->qe - 1 - k] : sq[al->qb + k];
        oq[k] = rev ? (c < 4 ? (uint8_t)(3 - c) : c) : c;
      }
      str_t md = {0, 0, 0};
      int nm = 0, run = 0, xq = 0;
      int64_t yr = fpos;
      for (int k = beg; k < endc; ++k) {
        int op = core2[k] & 0xf, ln = (int)(core2[k] >> 4);
        if (op == 0) {
          for (int t = 0; t < ln; ++t) {
            int rb = pac_base(pac, yr + t);
            if (oq[xq + t] != rb) { sputi(&md, run); sputc(&md, "ACGT"[rb]); run = 0; ++nm; }
            else ++run;
          }
          xq += ln; yr += ln;
        } else if (op == 1) { xq += ln; nm += ln; }
        else if (op == 2) {
          sputi(&md, run); sputc(&md, '^');
          for (int t = 0; t < ln; ++t) sputc(&md, "ACGT"[pac_base(pac, yr + t)]);
          run = 0; yr += ln; nm += ln;
        }
      }
      sputi(&md, run);
      free(oq);
      int rid = al->rid;
      int flag = (rev ? 0x10 : 0) | (a > a0 ? 0x800 : 0);
      int mapq = orc_mapq(&opt, al->score, al->qb, al->qe, al->rb, al->re);
      sput(&s, name, nl); sputc(&s, '\t'); sputi(&s, flag); sputc(&s, '\t');
      sputs(&s, orc_index_contig_name(idx, rid)); sputc(&s, '\t');
      int64_t contig_off = orc_index_contig_offset(idx, rid);
      sputi(&s, fpos - contig_off + 1); sputc(&s, '\t');
      sputi(&s, mapq); sputc(&s, '\t');
      if (clip_l) { sputi(&s, clip_l); sputc(&s, 'S'); }
      for (int k = beg; k < endc; ++k) { sputi(&s, core2[k] >> 4); sputc(&s, "MIDSH"[core2[k] & 0xf]); }
      if (clip_r) { sputi(&s, clip_r); sputc(&s, 'S'); }
      sputs(&s, "\t*\t0\t0\t");
      for (int k = 0; k < len; ++k) {
        uint8_t c = rev ? sq[len - 1 - k] : sq[k];
        sputc(&s, rev ? "TGCAN"[c] : "ACGTN"[c]);
      }
      sputc(&s, '\t');
      for (size_t k = 0; k < ql; ++k) sputc(&s, rev ? qual[ql - 1 - k] : qual[k]);
      sputs(&s, "\tNM:i:"); sputi(&s, nm);
      sputs(&s, "\tMD:Z:"); sput(&s, md.s, md.n);
      sputs(&s, "\tAS:i:"); sputi(&s, al->score);
      sputc(&s, '\n');
      free(md.s);
    }
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  fprintf(stderr, "oracle_format_seconds %.6f\n",
          (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec));
  FILE *out = out_fn ? fopen(out_fn, "wb") : stdout;
  if (!out) { fprintf(stderr, "g3a_oracle: cannot write %s\n", out_fn); return 1; }
  fwrite(s.s, 1, s.n, out);
  if (out_fn) fclose(out);
  free(s.s);
  free(core2);
  orc_batch_free(&b);
  orc_index_free(idx);
  free(pac);
  return 0;
}
