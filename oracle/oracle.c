/*
 * CPU oracle for the BWA-MEM hot path of sunwookim028/G3SA.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function cites the
 * reference file:line it restates; paths are relative to the reference repo.
 *
 * Parity pins (checked by tests/):
 *   index / Occ / SA ........ E. coli fixtures (.sa, .counts), brute-force SA
 *                             and Occ on small texts
 *   round-1 seeding ......... brute-force longest-match-ending-at-n search
 *   round-2 reseeding ....... the reference's control flow over brute-force
 *                             interval sizes (planted repeats)
 *   round-3 seeding ......... brute-force forward scan
 *   SA resolution ........... full suffix array of small texts
 *   chaining ................ sorted-list restatement (distinct positions);
 *                             the B-tree's duplicate-key order is pinned by
 *                             its own emulation only
 *   chain weight ............ closed form (interval-union sizes)
 *   chain / region filters .. known answers at every threshold + restatement
 *   extension ............... lane-by-lane emulation of the warp program and
 *                             a textbook banded affine DP where the band does
 *                             not bind
 *   global alignment ........ textbook Gotoh score and CIGAR consistency
 *   finalize / MAPQ ......... reads sampled at known places; MAPQ known
 *                             answers
 *   paired-end statistics ... closed-form size sets (quartiles, mean,
 *   and pair choice           population deviation, bounds, failure rules);
 *                             hand-placed hits for the pair scores
 *   hash_64 ................. PARITY UNPINNED (bijection on samples only)
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* options: bwa/src/fastmap.cpp:77-124                                      */
/* ------------------------------------------------------------------------ */
void orc_opt_init(orc_opt *o) {
  memset(o, 0, sizeof *o);
  o->a = 1; o->b = 4;
  o->o_del = o->o_ins = 6;
  o->e_del = o->e_ins = 1;
  o->w = 100;
  o->T = 30;
  o->pen_clip5 = o->pen_clip3 = 5;
  o->max_mem_intv = 20;
  o->min_seed_len = 19;
  o->split_width = 10;
  o->max_occ = 500;
  o->max_chain_gap = 10000;
  o->mask_level = 0.50f;
  o->drop_ratio = 0.50f;
  o->split_factor = 1.5f;
  o->mask_level_redun = 0.95f;
  o->mapQ_coef_len = 50;
  o->mapQ_coef_fac = (int)log(50.0);
  /* bwa_fill_scmat, bwa/src/fastmap.cpp:115-124 */
  int k = 0;
  for (int i = 0; i < 4; ++i) {
    for (int j = 0; j < 4; ++j) o->mat[k++] = (int8_t)(i == j ? o->a : -o->b);
    o->mat[k++] = -1;
  }
  for (int j = 0; j < 5; ++j) o->mat[k++] = -1;
}

/* ------------------------------------------------------------------------ */
/* index                                                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t cnt[4];
  uint64_t onehot[4];
} cp_occ_t; /* bwa/src/bwa.h:75-79; built as bwa/src/FMI_search.cpp:374-425 */

typedef struct {
  char *name;
  int64_t offset, len;
  int is_alt;
} orc_contig;

struct orc_index {
  int64_t l_pac, n;          /* n = 2*l_pac, BWT has n+1 rows */
  int n_contigs;
  orc_contig *contigs;
  uint8_t *pac;              /* forward strand, 2 bits/base */
  uint8_t *B;                /* BWT rows, 4 at the '$' row */
  cp_occ_t *occ;
  int64_t n_occ;
  int64_t count[5];          /* after the +1 load adjustment, FMI_search.cpp:873-878 */
  int64_t L2[5];
  int64_t primary;           /* sentinel_index */
  int64_t sa_intv, n_sa;
  int64_t *sa;               /* sa[0] = -1 as bwt_restore_sa sets it (bwt.c:440) */
  uint64_t onehot_mask[64];  /* FMI_search.cpp:869-880 */
};

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n); abort(); }
  return p;
}
static void *xcalloc(size_t n, size_t m) {
  void *p = calloc(n ? n : 1, m ? m : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

static uint8_t *slurp(const char *fn, size_t *len) {
  FILE *fp = fopen(fn, "rb");
  if (!fp) return NULL;
  fseek(fp, 0, SEEK_END);
  long sz = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  uint8_t *buf = (uint8_t *)xmalloc((size_t)sz + 1);
  size_t got = fread(buf, 1, (size_t)sz, fp);
  fclose(fp);
  if (got != (size_t)sz) { free(buf); return NULL; }
  buf[sz] = 0;
  *len = (size_t)sz;
  return buf;
}

static int get_pac(const uint8_t *pac, int64_t l) { return (pac[l >> 2] >> ((~l & 3) << 1)) & 3; }

orc_index *orc_index_load(const char *prefix) {
  char fn[4096];
  size_t len;
  orc_index *x = (orc_index *)xcalloc(1, sizeof *x);
  /* .ann */
  snprintf(fn, sizeof fn, "%s.ann", prefix);
  uint8_t *ann = slurp(fn, &len);
  if (!ann) { fprintf(stderr, "oracle: cannot read %s\n", fn); free(x); return NULL; }
  {
    char *p = (char *)ann;
    long long l_pac;
    int n_seqs;
    unsigned seed;
    if (sscanf(p, "%lld %d %u", &l_pac, &n_seqs, &seed) != 3) { free(ann); free(x); return NULL; }
    x->l_pac = l_pac;
    x->n_contigs = n_seqs;
    x->contigs = (orc_contig *)xcalloc((size_t)n_seqs, sizeof(orc_contig));
    p = strchr(p, '\n') + 1;
    for (int i = 0; i < n_seqs; ++i) {
      char *e = strchr(p, '\n');
      char *q = strchr(p, ' ');
      if (!e || !q || q > e) { free(ann); free(x); return NULL; }
      ++q;
      char *ne = q;
      while (ne < e && *ne != ' ') ++ne;
      x->contigs[i].name = strndup(q, (size_t)(ne - q));
      p = e + 1;
      long long off, ln;
      int namb;
      if (sscanf(p, "%lld %lld %d", &off, &ln, &namb) != 3) { free(ann); free(x); return NULL; }
      x->contigs[i].offset = off;
      x->contigs[i].len = ln;
      p = strchr(p, '\n') + 1;
    }
    free(ann);
  }
  /* .pac */
  snprintf(fn, sizeof fn, "%s.pac", prefix);
  x->pac = slurp(fn, &len);
  if (!x->pac) { fprintf(stderr, "oracle: cannot read %s\n", fn); return NULL; }
  /* .bwt: primary, L2[1..4], 128-row blocks (bwa/src/bwt.c:442-462) */
  snprintf(fn, sizeof fn, "%s.bwt", prefix);
  uint8_t *bw = slurp(fn, &len);
  if (!bw) { fprintf(stderr, "oracle: cannot read %s\n", fn); return NULL; }
  {
    uint64_t h[5];
    memcpy(h, bw, 40);
    x->primary = (int64_t)h[0];
    x->L2[0] = 0;
    for (int i = 0; i < 4; ++i) x->L2[i + 1] = (int64_t)h[i + 1];
    x->n = x->L2[4];
    const uint32_t *w = (const uint32_t *)(bw + 40);
    x->B = (uint8_t *)xmalloc((size_t)x->n + 1);
    for (int64_t k = 0; k < x->n; ++k) {
      int64_t word = (k >> 7) * 16 + 8 + ((k & 127) >> 4);
      int b = (int)((w[word] >> ((~k & 15) << 1)) & 3);
      x->B[k < x->primary ? k : k + 1] = (uint8_t)b;
    }
    x->B[x->primary] = 4;
    free(bw);
  }
  /* one-hot masks and checkpointed Occ over all n+1 rows */
  {
    uint64_t base = 0x8000000000000000ULL;
    x->onehot_mask[0] = base;
    for (int i = 1; i < 64; ++i) x->onehot_mask[i] = (x->onehot_mask[i - 1] >> 1) | base;
    const int64_t rows = x->n + 1;
    x->n_occ = (rows >> 6) + 1;
    x->occ = (cp_occ_t *)xcalloc((size_t)x->n_occ, sizeof(cp_occ_t));
    int64_t c[4] = {0, 0, 0, 0};
    for (int64_t i = 0; i < x->n_occ * 64; ++i) {
      if ((i & 63) == 0) memcpy(x->occ[i >> 6].cnt, c, sizeof c);
      if (i < rows && x->B[i] < 4) {
        x->occ[i >> 6].onehot[x->B[i]] |= 1ULL << (63 - (i & 63));
        ++c[x->B[i]];
      }
    }
    for (int i = 0; i < 5; ++i) x->count[i] = x->L2[i] + 1;
  }
  /* .sa: primary, L2[1..4], sa_intv, seq_len, sa[1..] (bwa/src/bwt.c:421-441) */
  snprintf(fn, sizeof fn, "%s.sa", prefix);
  uint8_t *sa = slurp(fn, &len);
  if (!sa) { fprintf(stderr, "oracle: cannot read %s\n", fn); return NULL; }
  {
    uint64_t h[7];
    memcpy(h, sa, 56);
    x->sa_intv = (int64_t)h[5];
    x->n_sa = (x->n + x->sa_intv) / x->sa_intv;
    x->sa = (int64_t *)xmalloc((size_t)x->n_sa * 8);
    x->sa[0] = -1;
    memcpy(x->sa + 1, sa + 56, (size_t)(x->n_sa - 1) * 8);
    free(sa);
  }
  return x;
}

void orc_index_free(orc_index *x) {
  if (!x) return;
  for (int i = 0; i < x->n_contigs; ++i) free(x->contigs[i].name);
  free(x->contigs); free(x->pac); free(x->B); free(x->occ); free(x->sa);
  free(x);
}
int64_t orc_index_l_pac(const orc_index *x) { return x->l_pac; }
int orc_index_n_contigs(const orc_index *x) { return x->n_contigs; }
const char *orc_index_contig_name(const orc_index *x, int rid) { return x->contigs[rid].name; }
int64_t orc_index_contig_len(const orc_index *x, int rid) { return x->contigs[rid].len; }
int64_t orc_index_contig_offset(const orc_index *x, int rid) { return x->contigs[rid].offset; }
uint64_t orc_index_primary(const orc_index *x) { return (uint64_t)x->primary; }

/* GET_OCC_GPU, bwa/src/fmindex.cu:29-36 */
int64_t orc_occ(const orc_index *x, int64_t p, int c) {
  const cp_occ_t *o = &x->occ[p >> 6];
  return o->cnt[c] + __builtin_popcountll(o->onehot[c] & x->onehot_mask[p & 63]);
}

/* bwt_invPsi + bwt_sa_gpu, bwa/src/bwt_CUDA.cu:433-454.  LF of row k with
 * BWT char c is L2[c] + (#c in rows [0, k], '$' excluded). */
static int64_t sa_walk(const orc_index *x, int64_t k, int64_t *steps) {
  int64_t sa = 0;
  while (k % x->sa_intv) {
    ++sa;
    if (steps) ++*steps;
    if (k == x->primary) { k = 0; continue; }
    int c = x->B[k];
    k = x->L2[c] + orc_occ(x, k, c);
  }
  return sa + x->sa[k / x->sa_intv];
}
int64_t orc_sa(const orc_index *x, int64_t k) { return sa_walk(x, k, NULL); }

/* ------------------------------------------------------------------------ */
/* FM-index extension: bwa/src/fmindex.cu:107-180, macro.h:50-71             */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t x[3];
  uint64_t info;
} bwtintv_t;
#define INFO(m, n) ((((uint64_t)(m)) << 32) | (uint64_t)((n) + 1))
#define IM(s) ((uint32_t)((s).info >> 32))
#define IN(s) ((uint32_t)((s).info) - 1)
#define ILEN(s) ((uint32_t)((s).info) - (uint32_t)((s).info >> 32))

/* backwardExt: all four extensions, then the bi-directional l update */
static void backward_ext(const orc_index *x, const bwtintv_t *in, int base, bwtintv_t *out) {
  int64_t k[4], l[4], s[4];
  for (int b = 0; b < 4; ++b) {
    int64_t sp = in->x[0] - 1, ep = in->x[0] + in->x[2] - 1;
    int64_t osp = orc_occ(x, sp, b), oep = orc_occ(x, ep, b);
    k[b] = x->count[b] + osp;
    s[b] = oep - osp;
  }
  int64_t so = (in->x[0] <= x->primary && in->x[0] + in->x[2] > x->primary) ? 1 : 0;
  l[3] = in->x[1] + so;
  l[2] = l[3] + s[3];
  l[1] = l[2] + s[2];
  l[0] = l[1] + s[1];
  out->x[0] = k[base];
  out->x[1] = l[base];
  out->x[2] = s[base];
}

/* backwardExtBackward: k and s only (x[1] untouched) */
static void backward_ext_b(const orc_index *x, bwtintv_t *io, int base) {
  int64_t sp = io->x[0] - 1, ep = io->x[0] + io->x[2] - 1;
  int64_t osp = orc_occ(x, sp, base), oep = orc_occ(x, ep, base);
  io->x[0] = x->count[base] + osp;
  io->x[2] = oep - osp;
}

/* FORWARD_EXT, macro.h:64-71 */
static void forward_ext(const orc_index *x, bwtintv_t *io, int base) {
  int64_t t = io->x[0];
  io->x[0] = io->x[1];
  io->x[1] = t;
  bwtintv_t o;
  backward_ext(x, io, 3 - base, &o);
  io->x[0] = o.x[1];
  io->x[1] = o.x[0];
  io->x[2] = o.x[2];
}

static void init_intv(const orc_index *x, bwtintv_t *p, int c) {
  p->x[0] = x->count[c];
  p->x[1] = x->count[3 - c];
  p->x[2] = x->count[c + 1] - x->count[c];
}

/* ------------------------------------------------------------------------ */
/* seeding                                                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
  bwtintv_t *a;
  int n, m;
} intv_v;
static void iv_push(intv_v *v, bwtintv_t p) {
  if (v->n == v->m) {
    v->m = v->m ? v->m * 2 : 16;
    v->a = (bwtintv_t *)realloc(v->a, (size_t)v->m * sizeof(bwtintv_t));
    if (!v->a) abort();
  }
  v->a[v->n++] = p;
}

/* preseedOnePos2Backward, bwa/src/seed.cu:454-531: the longest match
 * read[m..n] found by backward extension.  The reference steps two bases at a
 * time through the dinucleotide table with a one-base fallback; both tables
 * give the exact SA interval of the extended pattern, so the result equals
 * this one-base restatement.  `steps` counts the single-base steps the
 * reference's two-base schedule performs (a two-base step counts 2). */
static bwtintv_t preseed_one(const orc_index *x, const uint8_t *read, int n, int min_seed_len, int64_t *steps) {
  bwtintv_t smem;
  memset(&smem, 0, sizeof smem);
  if (read[n] >= 4) { smem.info = 0; return smem; }
  init_intv(x, &smem, read[n]);
  int m = n;
  while (m > 0 && read[m - 1] < 4) {
    bwtintv_t t = smem;
    backward_ext_b(x, &t, read[m - 1]);
    if (t.x[2] < 1) break;
    smem.x[0] = t.x[0];
    smem.x[2] = t.x[2];
    --m;
  }
  if (steps) {
    /* the reference's schedule (seed.cu:478-520): two-base steps while both
     * bases are valid and the pattern still occurs, a one-base attempt after
     * a failed or N-cut two-base step, and a last one-base attempt at m==1 */
    int mm = n, cost = 0;
    while (mm > 1) {
      int b0 = read[mm - 1], b1 = read[mm - 2];
      if (b1 >= 4) {
        if (b0 < 4) { cost += 1; if (mm - 1 >= m) --mm; }
        break;
      }
      if (b0 >= 4) break;
      cost += 2;
      if (mm - 2 >= m) { mm -= 2; continue; }
      cost += 1;
      if (mm - 1 >= m) --mm;
      break;
    }
    if (mm == 1 && read[0] < 4) cost += 1;
    *steps += cost;
  }
  smem.info = (n - m + 1 >= min_seed_len) ? INFO(m, n) : 0;
  return smem;
}

/* reseedV2, bwa/src/seed.cu:639-913 (per long SMEM: LEPs forward from the
 * pivot, then backward extension collecting at most one seed per step). */
static void reseed_round2(const orc_index *x, const orc_opt *opt, const uint8_t *read, int len,
                          intv_v *mem, int num_smem, int64_t *steps) {
  int split_len = (int)(opt->min_seed_len * opt->split_factor + .499);
  int split_intv = opt->split_width, min_seed_len = opt->min_seed_len;
  intv_v leps = {0, 0, 0}, seeds = {0, 0, 0};
  for (int k = 0; k < num_smem; ++k) {
    bwtintv_t smem = mem->a[k];
    if ((int)ILEN(smem) < split_len || smem.x[2] > split_intv) continue;
    int64_t min_seed_intv = smem.x[2] + 1;
    int pivot = (int)((IM(smem) + IN(smem) + 1) >> 1);
    if (pivot > 77) continue; /* bwa/src/seed.cu:718 */
    if (read[pivot] >= 4) continue;
    leps.n = seeds.n = 0;
    bwtintv_t seed, old;
    init_intv(x, &seed, read[pivot]);
    seed.info = INFO(pivot, pivot);
    int n;
    for (n = pivot; n < len - 1; ++n) {
      old = seed;
      if (read[n + 1] >= 4) break;
      forward_ext(x, &seed, read[n + 1]);
      if (steps) ++*steps;
      seed.info = INFO(pivot, n + 1);
      if (seed.x[2] < old.x[2]) {
        iv_push(&leps, old);
        if (seed.x[2] < min_seed_intv) break;
      }
    }
    if (seed.x[2] >= min_seed_intv) {
      seed.info = INFO(pivot, n);
      iv_push(&leps, seed);
    }
    for (int j = 0; j < leps.n / 2; ++j) {
      bwtintv_t t = leps.a[j];
      leps.a[j] = leps.a[leps.n - 1 - j];
      leps.a[leps.n - 1 - j] = t;
    }
    for (int m = pivot; m > 0; --m) {
      int base = read[m - 1];
      if (base >= 4) break;
      int num_old = leps.n;
      int64_t curr = min_seed_intv;
      int collected = 0;
      leps.n = 0;
      for (int j = 0; j < num_old; ++j) {
        bwtintv_t lep = leps.a[j], olep = leps.a[j], ext;
        backward_ext(x, &lep, base, &ext);
        if (steps) ++*steps;
        ext.info = lep.info;
        if (ext.x[2] >= curr) {
          ext.info = INFO(m - 1, IN(lep));
          leps.a[leps.n++] = ext;
          curr = ext.x[2];
        } else if (!collected && (int)ILEN(olep) >= min_seed_len) {
          iv_push(&seeds, olep);
          collected = 1;
        }
      }
      if (leps.n == 0) break;
    }
    /* DR: the reference appends this last seed without a capacity check
     * (bwa/src/seed.cu:870-873); the oracle simply appends. */
    if (leps.n > 0 && (int)ILEN(leps.a[0]) >= min_seed_len) iv_push(&seeds, leps.a[0]);
    for (int l = 0; l < seeds.n; ++l) iv_push(mem, seeds.a[l]);
  }
  free(leps.a);
  free(seeds.a);
}

/* reseedThirdRound, bwa/src/seed.cu:1143-1190, called from reseedLastRound
 * (seed.cu:1518-1550) with min length min_seed_len+1 and max_mem_intv. */
static void reseed_round3(const orc_index *x, const orc_opt *opt, const uint8_t *read, int len,
                          intv_v *mem, int64_t *steps) {
  int min_len = opt->min_seed_len + 1;
  int64_t max_intv = (int64_t)opt->max_mem_intv;
  int m = 0;
  while (m < len) {
    int next_m = m + 1;
    int base = read[m];
    if (base < 4) {
      bwtintv_t seed;
      memset(&seed, 0, sizeof seed);
      init_intv(x, &seed, base);
      for (int n = m + 1; n < len; ++n) {
        next_m = n + 1;
        base = read[n];
        if (base < 4) {
          forward_ext(x, &seed, base);
          if (steps) ++*steps;
          if (seed.x[2] < max_intv && (n - m + 1) >= min_len) {
            if (seed.x[2] > 0) {
              seed.info = INFO(m, n);
              iv_push(mem, seed);
            }
            break;
          }
        } else break;
      }
    }
    m = next_m;
  }
}

/* All three seeding rounds for one read, in the reference's output order:
 * round-1 SMEMs by end position, round-2 seeds, round-3 seeds. */
static void seed_read(const orc_index *x, const orc_opt *opt, const uint8_t *read, int len, intv_v *mem,
                      int64_t *steps) {
  mem->n = 0;
  if (len <= 0) return;
  /* preseedAndFilterV2, bwa/src/seed.cu:25-104 */
  bwtintv_t *pre = (bwtintv_t *)xcalloc((size_t)len, sizeof(bwtintv_t));
  for (int n = opt->min_seed_len - 1; n < len; ++n) pre[n] = preseed_one(x, read, n, opt->min_seed_len, steps);
  for (int n = 0; n < len; ++n) {
    int is_seed = pre[n].info != 0;
    if (n < len - 1 && (pre[n].info >> 32) == (pre[n + 1].info >> 32) && pre[n + 1].info != 0) is_seed = 0;
    if (is_seed) iv_push(mem, pre[n]);
  }
  free(pre);
  int num_smem = mem->n;
  if (num_smem > 0) reseed_round2(x, opt, read, len, mem, num_smem, steps);
  reseed_round3(x, opt, read, len, mem, steps);
}

int orc_seed_intervals(const orc_index *x, const orc_opt *opt, const uint8_t *read, int len, orc_intv *out,
                       int cap, int64_t *fm_steps) {
  intv_v mem = {0, 0, 0};
  seed_read(x, opt, read, len, &mem, fm_steps);
  int n = mem.n;
  for (int i = 0; i < n && i < cap; ++i) {
    out[i].k = mem.a[i].x[0];
    out[i].l = mem.a[i].x[1];
    out[i].s = mem.a[i].x[2];
    out[i].m = (int32_t)IM(mem.a[i]);
    out[i].n = (int32_t)IN(mem.a[i]);
  }
  free(mem.a);
  return n;
}

/* ------------------------------------------------------------------------ */
/* SA resolution and seed sort                                               */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t rbeg;
  int32_t qbeg, len, score, rid;
} seed_t;

/* bns_pos2rid_gpu / bns_intv2rid_gpu, bwa/src/bntseq.cu:7-37 */
static int pos2rid(const orc_index *x, int64_t pos_f) {
  if (pos_f >= x->l_pac) return -1;
  int left = 0, mid = 0, right = x->n_contigs;
  while (left < right) {
    mid = (left + right) >> 1;
    if (pos_f >= x->contigs[mid].offset) {
      if (mid == x->n_contigs - 1) break;
      if (pos_f < x->contigs[mid + 1].offset) break;
      left = mid + 1;
    } else right = mid;
  }
  return mid;
}
static int64_t depos(const orc_index *x, int64_t pos, int *is_rev) {
  return (*is_rev = (pos >= x->l_pac)) ? (x->l_pac << 1) - 1 - pos : pos;
}
static int intv2rid(const orc_index *x, int64_t rb, int64_t re) {
  int is_rev;
  if (rb < x->l_pac && re > x->l_pac) return -2;
  int rid_b = pos2rid(x, depos(x, rb, &is_rev));
  int rid_e = rb < re ? pos2rid(x, depos(x, re - 1, &is_rev)) : rid_b;
  return rid_b == rid_e ? rid_b : -1;
}

/* saLookup, bwa/src/aux.cu:1365-1425 */
static seed_t *sa_lookup(const orc_index *x, const orc_opt *opt, const intv_v *mem, int *n_out,
                         int64_t *lf_steps, int64_t *sa_reads) {
  int64_t total = 0;
  for (int k = 0; k < mem->n; ++k)
    total += mem->a[k].x[2] > opt->max_occ ? opt->max_occ : mem->a[k].x[2];
  seed_t *a = (seed_t *)xmalloc((size_t)total * sizeof(seed_t));
  int64_t o = 0;
  for (int k = 0; k < mem->n; ++k) {
    const bwtintv_t *p = &mem->a[k];
    int64_t cnt = p->x[2] > opt->max_occ ? opt->max_occ : p->x[2];
    int64_t step = p->x[2] > opt->max_occ ? p->x[2] / opt->max_occ : 1;
    for (int64_t j = 0; j < cnt; ++j) {
      int64_t kk = p->x[0] + step * j;
      seed_t s;
      s.rbeg = sa_walk(x, kk, lf_steps);
      if (sa_reads) ++*sa_reads;
      s.qbeg = (int32_t)IM(*p);
      s.len = s.score = (int32_t)ILEN(*p);
      s.rid = intv2rid(x, s.rbeg, s.rbeg + s.len);
      a[o++] = s;
    }
  }
  *n_out = (int)total;
  return a;
}

/* stable sort by a 64-bit key (restates cub::BlockRadixSort, which is stable) */
static void stable_sort_seeds(seed_t *a, int n, int by_qe) {
  seed_t *tmp = (seed_t *)xmalloc((size_t)n * sizeof(seed_t));
  for (int width = 1; width < n; width *= 2) {
    for (int lo = 0; lo < n; lo += 2 * width) {
      int mid = lo + width < n ? lo + width : n, hi = lo + 2 * width < n ? lo + 2 * width : n;
      int i = lo, j = mid, k = lo;
      while (i < mid && j < hi) {
        int64_t ki = by_qe == 1 ? (int64_t)a[i].qbeg + a[i].len : by_qe == 2 ? a[i].len : a[i].qbeg;
        int64_t kj = by_qe == 1 ? (int64_t)a[j].qbeg + a[j].len : by_qe == 2 ? a[j].len : a[j].qbeg;
        tmp[k++] = kj < ki ? a[j++] : a[i++];
      }
      while (i < mid) tmp[k++] = a[i++];
      while (j < hi) tmp[k++] = a[j++];
    }
    memcpy(a, tmp, (size_t)n * sizeof(seed_t));
  }
  free(tmp);
}

/* sortSeedsLowDim / sortSeedsHighDim, bwa/src/aux.cu:1436-1610: <=64 seeds
 * sorted by len then qbeg; more seeds: only the first 2048 positions are
 * loaded, sorted by qbeg+len then qbeg, and written back. */
static void sort_seeds(seed_t *a, int n) {
  if (n == 0) return;
  if (n <= 64) {
    stable_sort_seeds(a, n, 2);
    stable_sort_seeds(a, n, 0);
  } else {
    int m = n < 2048 ? n : 2048;
    stable_sort_seeds(a, m, 1);
    stable_sort_seeds(a, m, 0);
  }
}

int orc_read_seeds(const orc_index *x, const orc_opt *opt, const uint8_t *read, int len, orc_seed *out, int cap) {
  intv_v mem = {0, 0, 0};
  seed_read(x, opt, read, len, &mem, NULL);
  int n = 0;
  seed_t *a = sa_lookup(x, opt, &mem, &n, NULL, NULL);
  sort_seeds(a, n);
  for (int i = 0; i < n && i < cap; ++i) {
    out[i].rbeg = a[i].rbeg; out[i].qbeg = a[i].qbeg; out[i].len = a[i].len; out[i].rid = a[i].rid;
  }
  free(a);
  free(mem.a);
  return n;
}

/* ------------------------------------------------------------------------ */
/* chaining: BTreeChaining (bwa/src/aux.cu:1755-1825) over klib's B-tree    */
/* (bwa/src/kbtree_CUDA.cu:3-144) with t = 6, as kb_init_chn(512) gives.    */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t pos;
  int n, m, rid, is_alt;
  int w;
  int id; /* B-tree-order index (stage dumps) */
  seed_t *seeds;
} chain_t;

#define KB_T 6
#define KB_N (2 * KB_T - 1)
typedef struct {
  int n, is_internal;
  int key[KB_N];       /* chain ids */
  int child[KB_N + 1]; /* node ids */
} kbnode_t;
typedef struct {
  kbnode_t *nodes;
  int n_nodes, m_nodes, root, n_keys;
  chain_t *chains; /* key storage */
} kbtree_t;

static int kb_new_node(kbtree_t *b, int internal) {
  if (b->n_nodes == b->m_nodes) {
    b->m_nodes = b->m_nodes ? b->m_nodes * 2 : 16;
    b->nodes = (kbnode_t *)realloc(b->nodes, (size_t)b->m_nodes * sizeof(kbnode_t));
    if (!b->nodes) abort();
  }
  kbnode_t *z = &b->nodes[b->n_nodes];
  memset(z, 0, sizeof *z);
  z->is_internal = internal;
  return b->n_nodes++;
}
static int chain_cmp(int64_t a, int64_t b) { return (b < a) - (a < b); }
/* __kb_getp_aux_chn, kbtree_CUDA.cu:22-41 */
static int kb_getp_aux(const kbtree_t *b, const kbnode_t *x, int64_t kpos, int *r) {
  int tr, *rr, begin = 0, end = x->n;
  if (x->n == 0) return -1;
  rr = r ? r : &tr;
  while (begin < end) {
    int mid = (begin + end) >> 1;
    if (chain_cmp(b->chains[x->key[mid]].pos, kpos) < 0) begin = mid + 1;
    else end = mid;
  }
  if (begin == x->n) { *rr = 1; return x->n - 1; }
  if ((*rr = chain_cmp(kpos, b->chains[x->key[begin]].pos)) < 0) --begin;
  return begin;
}
/* kb_intervalp_chn, kbtree_CUDA.cu:44-60: returns the "lower" chain id or -1 */
static int kb_lower(const kbtree_t *b, int64_t kpos) {
  int lower = -1, r = 0, xi = b->root;
  while (xi >= 0) {
    const kbnode_t *x = &b->nodes[xi];
    int i = kb_getp_aux(b, x, kpos, &r);
    if (i >= 0 && r == 0) return x->key[i];
    if (i >= 0) lower = x->key[i];
    if (!x->is_internal) return lower;
    xi = x->child[i + 1];
  }
  return lower;
}
/* __kb_split_chn, kbtree_CUDA.cu:63-78 */
static void kb_split(kbtree_t *b, int xi, int i, int yi) {
  int zi = kb_new_node(b, b->nodes[yi].is_internal);
  kbnode_t *x = &b->nodes[xi], *y = &b->nodes[yi], *z = &b->nodes[zi];
  z->n = KB_T - 1;
  memcpy(z->key, y->key + KB_T, sizeof(int) * (KB_T - 1));
  if (y->is_internal) memcpy(z->child, y->child + KB_T, sizeof(int) * KB_T);
  y->n = KB_T - 1;
  memmove(x->child + i + 2, x->child + i + 1, sizeof(int) * (size_t)(x->n - i));
  x->child[i + 1] = zi;
  memmove(x->key + i + 1, x->key + i, sizeof(int) * (size_t)(x->n - i));
  x->key[i] = y->key[KB_T - 1];
  ++x->n;
}
/* __kb_putp_aux_chn, kbtree_CUDA.cu:81-98 */
static void kb_put_aux(kbtree_t *b, int xi, int cid) {
  int64_t kpos = b->chains[cid].pos;
  kbnode_t *x = &b->nodes[xi];
  if (!x->is_internal) {
    int i = kb_getp_aux(b, x, kpos, NULL);
    if (i != x->n - 1) memmove(x->key + i + 2, x->key + i + 1, (size_t)(x->n - i - 1) * sizeof(int));
    x->key[i + 1] = cid;
    ++x->n;
  } else {
    int i = kb_getp_aux(b, x, kpos, NULL) + 1;
    if (b->nodes[x->child[i]].n == 2 * KB_T - 1) {
      kb_split(b, xi, i, x->child[i]);
      x = &b->nodes[xi];
      if (chain_cmp(kpos, b->chains[x->key[i]].pos) > 0) ++i;
    }
    kb_put_aux(b, x->child[i], cid);
  }
}
/* kb_putp_chn, kbtree_CUDA.cu:100-114 */
static void kb_put(kbtree_t *b, int cid) {
  ++b->n_keys;
  int ri = b->root;
  if (b->nodes[ri].n == 2 * KB_T - 1) {
    int si = kb_new_node(b, 1);
    b->nodes[si].n = 0;
    b->nodes[si].child[0] = ri;
    b->root = si;
    kb_split(b, si, 0, ri);
    ri = si;
  }
  kb_put_aux(b, ri, cid);
}
static void kb_traverse(const kbtree_t *b, int xi, int *out, int *n) {
  const kbnode_t *x = &b->nodes[xi];
  for (int i = 0; i < x->n; ++i) {
    if (x->is_internal) kb_traverse(b, x->child[i], out, n);
    out[(*n)++] = x->key[i];
  }
  if (x->is_internal) kb_traverse(b, x->child[x->n], out, n);
}

static void chain_push(chain_t *c, const seed_t *s) {
  if (c->n == c->m) {
    c->m = c->m ? c->m * 2 : 4;
    c->seeds = (seed_t *)realloc(c->seeds, (size_t)c->m * sizeof(seed_t));
    if (!c->seeds) abort();
  }
  c->seeds[c->n++] = *s;
}

/* test_and_merge, bwa/src/aux.cu:54-84 */
static int test_and_merge(const orc_opt *opt, int64_t l_pac, chain_t *c, const seed_t *p) {
  const seed_t *last = &c->seeds[c->n - 1];
  int64_t qend = last->qbeg + last->len, rend = last->rbeg + last->len;
  if (p->rid != c->rid) return 0;
  if (p->qbeg >= c->seeds[0].qbeg && p->qbeg + p->len <= qend && p->rbeg >= c->seeds[0].rbeg &&
      p->rbeg + p->len <= rend)
    return 1;
  if ((last->rbeg < l_pac || c->seeds[0].rbeg < l_pac) && p->rbeg >= l_pac) return 0;
  int64_t xd = p->qbeg - last->qbeg, yd = p->rbeg - last->rbeg;
  if (yd >= 0 && xd - yd <= opt->w && yd - xd <= opt->w && xd - last->len < opt->max_chain_gap &&
      yd - last->len < opt->max_chain_gap) {
    chain_push(c, p);
    return 2;
  }
  return 0;
}

/* returns chains in B-tree order; if chain_of_seed is given, it receives for
 * each seed the B-tree-order index of the chain that took it, -1 when the
 * seed was contained in a chain (and dropped) or started none (rid rule) */
static chain_t *chain_seeds(const orc_index *x, const orc_opt *opt, const seed_t *a, int n, int *n_chains,
                            int32_t *chain_of_seed) {
  kbtree_t b;
  memset(&b, 0, sizeof b);
  int m_ch = 16, n_ch = 0;
  b.chains = (chain_t *)xmalloc((size_t)m_ch * sizeof(chain_t));
  b.root = kb_new_node(&b, 0);
  for (int i = 0; i < n; ++i) {
    seed_t s = a[i];
    int start_new = 0;
    if (chain_of_seed) chain_of_seed[i] = -1;
    if (b.n_keys) {
      int lower = kb_lower(&b, s.rbeg);
      int merged = lower < 0 ? 0 : test_and_merge(opt, x->l_pac, &b.chains[lower], &s);
      if (!merged) start_new = 1;
      else if (merged == 2 && chain_of_seed) chain_of_seed[i] = lower;
    } else start_new = 1;
    if (start_new) {
      if (s.rid < 0 || s.rid >= 455) continue; /* bwa/src/aux.cu:1799 */
      if (chain_of_seed) chain_of_seed[i] = n_ch;
      if (n_ch == m_ch) {
        m_ch *= 2;
        b.chains = (chain_t *)realloc(b.chains, (size_t)m_ch * sizeof(chain_t));
        if (!b.chains) abort();
      }
      chain_t *c = &b.chains[n_ch];
      memset(c, 0, sizeof *c);
      c->pos = s.rbeg;
      c->rid = s.rid;
      c->is_alt = x->contigs[s.rid].is_alt;
      chain_push(c, &s);
      kb_put(&b, n_ch);
      ++n_ch;
    }
  }
  int *order = (int *)xmalloc((size_t)(n_ch ? n_ch : 1) * sizeof(int));
  int k = 0;
  if (n_ch) kb_traverse(&b, b.root, order, &k);
  chain_t *out = (chain_t *)xmalloc((size_t)(n_ch ? n_ch : 1) * sizeof(chain_t));
  for (int i = 0; i < k; ++i) out[i] = b.chains[order[i]];
  if (chain_of_seed) {
    int *rank = (int *)xmalloc((size_t)(n_ch ? n_ch : 1) * sizeof(int));
    for (int i = 0; i < k; ++i) rank[order[i]] = i;
    for (int i = 0; i < n; ++i)
      if (chain_of_seed[i] >= 0) chain_of_seed[i] = rank[chain_of_seed[i]];
    free(rank);
  }
  free(order);
  free(b.chains);
  free(b.nodes);
  *n_chains = k;
  return out;
}

/* mem_chain_weight, bwa/src/aux.cu:91-110 */
static int chain_weight(const chain_t *c) {
  int64_t end;
  int j, w = 0, tmp;
  for (j = 0, end = 0; j < c->n; ++j) {
    const seed_t *s = &c->seeds[j];
    if (s->qbeg >= end) w += s->len;
    else if (s->qbeg + s->len > end) w += (int)(s->qbeg + s->len - end);
    end = end > s->qbeg + s->len ? end : s->qbeg + s->len;
  }
  tmp = w;
  w = 0;
  for (j = 0, end = 0; j < c->n; ++j) {
    const seed_t *s = &c->seeds[j];
    if (s->rbeg >= end) w += s->len;
    else if (s->rbeg + s->len > end) w += (int)(s->rbeg + s->len - end);
    end = end > s->rbeg + s->len ? end : s->rbeg + s->len;
  }
  w = w < tmp ? w : tmp;
  return w < 1 << 30 ? w : (1 << 30) - 1;
}

/* sortChainsDecreasingWeight (bwa/src/aux.cu:1828-1884) + CHAINFILTERING
 * (aux.cu:1890-1988).  Weights are kept as uint16.  Sorting is stable and
 * descending over the first 2048 chains.
 * DR: more than 3072 chains are left unsorted with their weights set, and
 * chains [2048, n) keep their place (the reference reads uninitialised
 * shared memory there).  The chain filter is the sequential fixed point of
 * the reference's spin-wait: i is dropped iff an earlier kept chain j
 * overlaps it significantly and outweighs it. */
static int filter_chains(const orc_opt *opt, chain_t *a, int n) {
  if (n == 0) return 0;
  for (int i = 0; i < n; ++i) a[i].w = (uint16_t)chain_weight(&a[i]);
  if (n <= 3072) {
    int m = n < 2048 ? n : 2048;
    chain_t *tmp = (chain_t *)xmalloc((size_t)m * sizeof(chain_t));
    for (int width = 1; width < m; width *= 2) {
      for (int lo = 0; lo < m; lo += 2 * width) {
        int mid = lo + width < m ? lo + width : m, hi = lo + 2 * width < m ? lo + 2 * width : m;
        int i = lo, j = mid, k = lo;
        while (i < mid && j < hi) tmp[k++] = a[j].w > a[i].w ? a[j++] : a[i++];
        while (i < mid) tmp[k++] = a[i++];
        while (j < hi) tmp[k++] = a[j++];
      }
      memcpy(a, tmp, (size_t)m * sizeof(chain_t));
    }
    free(tmp);
  }
  if (n > 4096) return n;
  uint8_t *kept = (uint8_t *)xmalloc((size_t)n);
  for (int i = 0; i < n; ++i) {
    uint16_t bi = (uint16_t)a[i].seeds[0].qbeg;
    uint16_t ei = (uint16_t)(a[i].seeds[a[i].n - 1].qbeg + a[i].seeds[a[i].n - 1].len);
    kept[i] = 1;
    if (i == 0) continue;
    for (int j = 0; j < i; ++j) {
      if (!kept[j]) continue;
      uint16_t bj = (uint16_t)a[j].seeds[0].qbeg;
      uint16_t ej = (uint16_t)(a[j].seeds[a[j].n - 1].qbeg + a[j].seeds[a[j].n - 1].len);
      int b_max = bj > bi ? bj : bi, e_min = ej < ei ? ej : ei;
      if (e_min > b_max && (!a[i].is_alt || a[j].is_alt)) {
        int li = ei - bi, lj = ej - bj, min_l = li < lj ? li : lj;
        if (e_min - b_max >= min_l * opt->mask_level && min_l < opt->max_chain_gap) {
          uint16_t wi = (uint16_t)a[i].w, wj = (uint16_t)a[j].w;
          if (wi < wj * opt->drop_ratio && wj - wi >= opt->min_seed_len << 1) {
            kept[i] = 0;
            break;
          }
        }
      }
    }
  }
  int k = 0;
  for (int i = 0; i < n; ++i) {
    if (kept[i]) a[k++] = a[i];
    else free(a[i].seeds);
  }
  free(kept);
  return k;
}

/* Stage dump for tests: chaining and chain filtering of one read's seeds
 * (already in sort_seeds order).  chain_of_seed[i]: chain (B-tree order) that
 * took seed i, or -1.  weight[c]: chain_weight of chain c (B-tree order).
 * kept[k]: B-tree index of the k-th chain after sorting and filtering.
 * Returns the number of chains; *n_kept receives the number kept. */
int orc_chain_dump(const orc_index *x, const orc_opt *opt, const orc_seed *seeds, int n, int32_t *chain_of_seed,
                   int32_t *weight, int32_t *kept, int *n_kept) {
  seed_t *a = (seed_t *)xmalloc((size_t)(n ? n : 1) * sizeof(seed_t));
  for (int i = 0; i < n; ++i) {
    a[i].rbeg = seeds[i].rbeg; a[i].qbeg = seeds[i].qbeg; a[i].len = seeds[i].len; a[i].rid = seeds[i].rid;
    a[i].score = seeds[i].len;
  }
  int n_ch = 0;
  chain_t *ch = n ? chain_seeds(x, opt, a, n, &n_ch, chain_of_seed) : NULL;
  free(a);
  for (int c = 0; c < n_ch; ++c) {
    ch[c].id = c;
    weight[c] = chain_weight(&ch[c]);
  }
  int k = filter_chains(opt, ch, n_ch);
  for (int i = 0; i < k; ++i) kept[i] = ch[i].id;
  for (int i = 0; i < k; ++i) free(ch[i].seeds);
  free(ch);
  *n_kept = k;
  return n_ch;
}

/* ------------------------------------------------------------------------ */
/* banded local extension                                                    */
/* ------------------------------------------------------------------------ */
static int ext_w(int qlen, const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins, int w, int end_bonus) {
  int max_ = 0;
  for (int i = 0; i < 25; ++i) max_ = max_ > mat[i] ? max_ : mat[i];
  int max_ins = (int)((double)(qlen * max_ + end_bonus - o_ins) / e_ins + 1.);
  max_ins = max_ins > 1 ? max_ins : 1;
  w = w < max_ins ? w : max_ins;
  int max_del = (int)((double)(qlen * max_ + end_bonus - o_del) / e_del + 1.);
  max_del = max_del > 1 ? max_del : 1;
  w = w < max_del ? w : max_del;
  return w;
}

#define MAX2(a, b) ((a) > (b) ? (a) : (b))

/* ksw_extend_warp2, bwa/src/ksw.cu:814-1034, as the set of cells it fills
 * and the value it returns:
 *   cell (i,j) is filled iff 0<=i<tlen, 0<=j<qlen, j>=i-w,
 *     j <= 32*(i/32+1)+w+1  (the tile sweep's column limit), and
 *     j <= i+w+1 when i+j <= 30 (the 31 corner anti-diagonals);
 *   unfilled neighbours read as 0, except in rows i = 32k (k >= 1) at
 *     columns j > i+w+1, whose upper neighbours were never filled: there the
 *     diagonal input is the untouched first-row value max(h0-o_ins-e_ins*j, 0)
 *     and the E input is 0 (the warp's shared row buffer still holds them);
 *   recurrence M = H(i-1,j-1) ? H(i-1,j-1)+S : 0, H = max(M,E,F),
 *     E' = max(E-e_del, M-o_del-e_del, 0), F' = max(F-e_ins, M-o_ins-e_ins, 0);
 *   best = max(h0, filled H); its position is the smallest row holding best
 *     and, in it, the largest column, except that when best == h0 the result
 *     is "no cell" (qle = tle = 0) unless every residue class of rows mod 32
 *     holds a cell equal to h0 (per-lane argmax seeded with (-1,-1), then a
 *     warp reduction preferring smaller rows).
 * gscore is never updated by the reference (DR-5) so only the local result is
 * returned. */
int orc_ksw_extend(int qlen, const uint8_t *query, int tlen, const uint8_t *target, const int8_t *mat, int o_del,
                   int e_del, int o_ins, int e_ins, int w, int end_bonus, int h0, int *qle, int *tle,
                   int64_t *cells) {
  if (qlen > 500) return 0;
  const int oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
  w = ext_w(qlen, mat, o_del, e_del, o_ins, e_ins, w, end_bonus);
  int *Hp = (int *)xmalloc(sizeof(int) * (size_t)(qlen + 1)); /* H of the previous row, 0 if unfilled */
  int *Ep = (int *)xmalloc(sizeof(int) * (size_t)(qlen + 1)); /* E flowing into the current row */
  int *Hc = (int *)xmalloc(sizeof(int) * (size_t)(qlen + 1));
  int *Ec = (int *)xmalloc(sizeof(int) * (size_t)(qlen + 1));
  uint8_t *vp = (uint8_t *)xcalloc((size_t)qlen + 1, 1), *vc = (uint8_t *)xcalloc((size_t)qlen + 1, 1);
  int best = h0, bi = -1, bj = -1;
  uint32_t h0_rows = 0; /* residues (mod 32) having a cell equal to h0 */
  int best_row_h0 = -1, best_col_h0 = -1;
  int64_t ncell = 0;
  for (int i = 0; i < tlen; ++i) {
    int limit = 32 * (i / 32 + 1) + w + 1;
    int f = 0, hleft = 0, vleft = 0;
    for (int j = 0; j < qlen; ++j) {
      int valid = j >= i - w && j <= limit && (i + j > 30 || j <= i + w + 1);
      vc[j] = (uint8_t)valid;
      if (!valid) { Hc[j] = 0; Ec[j] = 0; if (j == 0 || !vleft) f = 0; vleft = 0; hleft = 0; continue; }
      ++ncell;
      int hdiag, ein;
      if (i == 0) {
        hdiag = j == 0 ? h0 : MAX2(h0 - oe_ins - (j - 1) * e_ins, 0);
        ein = 0;
      } else if ((i & 31) == 0 && j > i + w + 1) {
        hdiag = MAX2(h0 - oe_ins - (j - 1) * e_ins, 0);
        ein = 0;
      } else {
        if (j == 0) hdiag = MAX2(h0 - (o_del + e_del * i), 0);
        else hdiag = vp[j - 1] ? Hp[j - 1] : 0;
        ein = vp[j] ? Ep[j] : 0;
      }
      int fin = (j == 0 || !vleft) ? 0 : f;
      int M = hdiag ? hdiag + mat[target[i] * 5 + query[j]] : 0;
      int h = MAX2(M, ein);
      h = MAX2(h, fin);
      int t = MAX2(M - oe_del, 0);
      int e = MAX2(ein - e_del, t);
      t = MAX2(M - oe_ins, 0);
      f = MAX2(fin - e_ins, t);
      Hc[j] = h;
      Ec[j] = e;
      vleft = 1;
      hleft = h;
      if (h > best) { best = h; bi = i; bj = j; }
      else if (h == best && best > h0 && i == bi && j > bj) bj = j;
      if (h == h0) {
        h0_rows |= 1u << (i & 31);
        if (best_row_h0 < 0 || i == best_row_h0) { best_row_h0 = i; best_col_h0 = j; }
      }
    }
    (void)hleft;
    int *tp = Hp; Hp = Hc; Hc = tp;
    tp = Ep; Ep = Ec; Ec = tp;
    uint8_t *tv = vp; vp = vc; vc = tv;
  }
  if (best == h0) {
    if (h0_rows == 0xffffffffu) { bi = best_row_h0; bj = best_col_h0; }
    else { bi = -1; bj = -1; }
  }
  free(Hp); free(Ep); free(Hc); free(Ec); free(vp); free(vc);
  if (qle) *qle = bj + 1;
  if (tle) *tle = bi + 1;
  if (cells) *cells += ncell;
  return best;
}

/* Lane-by-lane emulation of ksw_extend_warp2 (bwa/src/ksw.cu:814-1034) on a
 * 32-lane warp executing in lock step: shuffles read the values of the
 * previous step, lane 0 reads the shared row buffers, lane 31 writes them at
 * the end of the step, and registers that the program leaves uninitialised
 * hold an arbitrary value.  Used only to check orc_ksw_extend. */
int orc_ksw_extend_warp_emulate(int qlen, const uint8_t *query, int tlen, const uint8_t *target, const int8_t *mat,
                                int o_del, int e_del, int o_ins, int e_ins, int w, int end_bonus, int h0, int *qle,
                                int *tle) {
  enum { WS = 32, GARBAGE = 7777 };
  if (qlen > 500) return 0;
  int16_t SM_H[512], SM_E[512];
  for (int i = 0; i < 512; ++i) SM_H[i] = SM_E[i] = (int16_t)GARBAGE;
  const int oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
  int e[WS], f[WS], h[WS], e1_[WS], h1_[WS], h_1[WS], h11[WS];
  int max_score[WS], i_m[WS], j_m[WS];
  for (int t = 0; t < WS; ++t) {
    e[t] = f[t] = h[t] = e1_[t] = h1_[t] = h_1[t] = h11[t] = GARBAGE;
    max_score[t] = h0; i_m[t] = -1; j_m[t] = -1;
  }
  SM_H[0] = (int16_t)h0;
  for (int j = 1; j < qlen + 1; ++j) {
    SM_E[j] = 0;
    int hh = h0 - oe_ins - (j - 1) * e_ins;
    SM_H[j] = (int16_t)(hh > 0 ? hh : 0);
  }
  w = ext_w(qlen, mat, o_del, e_del, o_ins, e_ins, w, end_bonus);
  /* corner */
  for (int ad = 0; ad < WS - 1; ++ad) {
    int se[WS], sh[WS], sh1[WS];
    for (int t = 0; t < WS; ++t) {
      int src = t > 0 ? t - 1 : t;
      se[t] = e[src]; sh[t] = h[src]; sh1[t] = h_1[src];
    }
    for (int t = 0; t < WS; ++t) {
      int i = t, j = ad - t;
      e1_[t] = se[t]; h1_[t] = sh[t]; h11[t] = sh1[t];
      h_1[t] = h[t];
      if (j == 0) {
        h_1[t] = h0 - (o_del + e_del * (i + 1));
        if (h_1[t] < 0) h_1[t] = 0;
        f[t] = 0;
      }
      if (t == 0) { e1_[t] = SM_E[j + 1]; h1_[t] = SM_H[j + 1]; h11[t] = SM_H[j]; }
    }
    for (int t = 0; t < WS; ++t) {
      int i = t, j = ad - t;
      if (0 <= i && i < tlen && 0 <= j && j < qlen && i - w <= j && j <= i + w + 1) {
        int M = h11[t];
        M = M ? M + mat[target[i] * 5 + query[j]] : 0;
        h[t] = MAX2(M, e1_[t]);
        h[t] = MAX2(h[t], f[t]);
        if (h[t] > max_score[t]) { max_score[t] = h[t]; i_m[t] = i; j_m[t] = j; }
        else if (h[t] == max_score[t]) {
          if (i == i_m[t] && j > j_m[t]) { i_m[t] = i; j_m[t] = j; }
          else if (i_m[t] < 0 && j_m[t] < 0) { i_m[t] = i; j_m[t] = j; }
        }
        int tt = MAX2(M - oe_del, 0);
        e[t] = MAX2(e1_[t] - e_del, tt);
        tt = MAX2(M - oe_ins, 0);
        f[t] -= e_ins;
        f[t] = MAX2(f[t], tt);
        if (t == WS - 1) { SM_H[j] = (int16_t)h[t]; SM_E[j] = (int16_t)e[t]; }
      } else { h[t] = 0; e[t] = 0; f[t] = 0; }
    }
  }
  /* tiles */
  int Ntile = (int)ceilf((float)tlen / WS);
  int qpad = qlen >= 32 ? qlen : 32;
  for (int T = 0; T < Ntile; ++T) {
    int ii[WS], jj[WS];
    for (int t = 0; t < WS; ++t) { ii[t] = T * WS + t; jj[t] = (WS - 1) - t; }
    for (int ad = 0; ad < (T + 1) * WS + w + 2 && ad < qpad; ++ad) {
      for (int t = 0; t < WS; ++t)
        if (jj[t] >= qpad || jj[t] > (T + 1) * WS + w + 1) { ii[t] += WS; jj[t] = 0; }
      int se[WS], sh[WS], sh1[WS];
      for (int t = 0; t < WS; ++t) {
        int src = t > 0 ? t - 1 : t;
        se[t] = e[src]; sh[t] = h[src]; sh1[t] = h_1[src];
      }
      for (int t = 0; t < WS; ++t) {
        e1_[t] = se[t]; h1_[t] = sh[t]; h11[t] = sh1[t];
        h_1[t] = h[t];
        if (jj[t] == 0) {
          h_1[t] = h0 - (o_del + e_del * (ii[t] + 1));
          if (h_1[t] < 0) h_1[t] = 0;
          f[t] = 0;
        }
        if (t == 0) {
          int j = jj[t];
          e1_[t] = (j >= 0 && j < 512) ? SM_E[j] : GARBAGE;
          h1_[t] = (j + 1 < 512) ? SM_H[j + 1] : GARBAGE;
          h11[t] = (j >= 0 && j < 512) ? SM_H[j] : GARBAGE;
        }
      }
      for (int t = 0; t < WS; ++t) {
        int i = ii[t], j = jj[t];
        if (0 <= i && i < tlen && 0 <= j && j < qlen && i - w <= j) {
          int M = h11[t];
          M = M ? M + mat[target[i] * 5 + query[j]] : 0;
          h[t] = MAX2(M, e1_[t]);
          h[t] = MAX2(h[t], f[t]);
          if (h[t] > max_score[t]) { max_score[t] = h[t]; i_m[t] = i; j_m[t] = j; }
          else if (h[t] == max_score[t]) {
            if (i == i_m[t]) { i_m[t] = i; j_m[t] = j; }
            else if (i_m[t] < 0 && j_m[t] < 0) { i_m[t] = i; j_m[t] = j; }
          }
          int tt = MAX2(M - oe_del, 0);
          e[t] = MAX2(e1_[t] - e_del, tt);
          tt = MAX2(M - oe_ins, 0);
          f[t] -= e_ins;
          f[t] = MAX2(f[t], tt);
          if (t == WS - 1) { SM_H[j] = (int16_t)h_1[t]; SM_E[j] = (int16_t)e[t]; }
        }
        jj[t]++;
      }
    }
  }
  /* shuffle-down reduction, offsets 1,2,4,8,16 (lanes past the end keep their value) */
  for (int r = 0; r < 5; ++r) {
    int off = 1 << r, tm[WS], ti[WS], tj[WS];
    for (int t = 0; t < WS; ++t) {
      int src = t + off < WS ? t + off : t;
      tm[t] = max_score[src]; ti[t] = i_m[src]; tj[t] = j_m[src];
    }
    for (int t = 0; t < WS; ++t) {
      if (max_score[t] < tm[t]) { max_score[t] = tm[t]; i_m[t] = ti[t]; j_m[t] = tj[t]; }
      else if (max_score[t] == tm[t]) {
        if (ti[t] < i_m[t]) { i_m[t] = ti[t]; j_m[t] = tj[t]; }
        else if (ti[t] == i_m[t] && tj[t] > j_m[t]) { i_m[t] = ti[t]; j_m[t] = tj[t]; }
      }
    }
  }
  if (qle) *qle = j_m[0] + 1;
  if (tle) *tle = i_m[0] + 1;
  return max_score[0];
}

/* ------------------------------------------------------------------------ */
/* global alignment with traceback: ksw_global2, bwa/src/ksw.cu:706-812     */
/* ------------------------------------------------------------------------ */
#define MINUS_INF (-0x40000000)
static int push_cigar(uint32_t *cigar, int n, int cap, int op, int len) {
  if (n == 0 || op != (int)(cigar[n - 1] & 0xf)) {
    if (n < cap) cigar[n] = (uint32_t)(len << 4 | op);
    return n + 1;
  }
  cigar[n - 1] += (uint32_t)(len << 4);
  return n;
}

int orc_ksw_global2(int qlen, const uint8_t *query, int tlen, const uint8_t *target, const int8_t *mat, int o_del,
                    int e_del, int o_ins, int e_ins, int w, int *n_cigar_, uint32_t *cigar, int cap) {
  int oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
  int n_col = qlen < 2 * w + 1 ? qlen : 2 * w + 1;
  uint8_t *z = (uint8_t *)xmalloc((size_t)n_col * (size_t)(tlen > 0 ? tlen : 1));
  int8_t *qp = (int8_t *)xmalloc((size_t)qlen * 5 + 1);
  int32_t *eh_h = (int32_t *)xcalloc((size_t)qlen + 1, 4), *eh_e = (int32_t *)xcalloc((size_t)qlen + 1, 4);
  for (int k = 0, i = 0; k < 5; ++k)
    for (int j = 0; j < qlen; ++j) qp[i++] = mat[k * 5 + query[j]];
  eh_h[0] = 0; eh_e[0] = MINUS_INF;
  int j;
  for (j = 1; j <= qlen && j <= w; ++j) eh_h[j] = -(o_ins + e_ins * j), eh_e[j] = MINUS_INF;
  for (; j <= qlen; ++j) eh_h[j] = eh_e[j] = MINUS_INF;
  for (int i = 0; i < tlen; ++i) {
    int32_t f = MINUS_INF, h1, beg, end, t;
    const int8_t *q = &qp[target[i] * qlen];
    beg = i > w ? i - w : 0;
    end = i + w + 1 < qlen ? i + w + 1 : qlen;
    h1 = beg == 0 ? -(o_del + e_del * (i + 1)) : MINUS_INF;
    uint8_t *zi = &z[(size_t)i * (size_t)n_col];
    for (j = beg; j < end; ++j) {
      int32_t h, m = eh_h[j], e = eh_e[j];
      uint8_t d;
      eh_h[j] = h1;
      m += q[j];
      d = m >= e ? 0 : 1;
      h = m >= e ? m : e;
      d = h >= f ? d : 2;
      h = h >= f ? h : f;
      h1 = h;
      t = m - oe_del;
      e -= e_del;
      d |= e > t ? 1 << 2 : 0;
      e = e > t ? e : t;
      eh_e[j] = e;
      t = m - oe_ins;
      f -= e_ins;
      d |= f > t ? 2 << 4 : 0;
      f = f > t ? f : t;
      zi[j - beg] = d;
    }
    eh_h[end] = h1;
    eh_e[end] = MINUS_INF;
  }
  int score = eh_h[qlen];
  int n_cigar = 0, which = 0, i = tlen - 1, k = (i + w + 1 < qlen ? i + w + 1 : qlen) - 1;
  while (i >= 0 && k >= 0) {
    which = z[(size_t)i * (size_t)n_col + (size_t)(k - (i > w ? i - w : 0))] >> (which << 1) & 3;
    if (which == 0) n_cigar = push_cigar(cigar, n_cigar, cap, 0, 1), --i, --k;
    else if (which == 1) n_cigar = push_cigar(cigar, n_cigar, cap, 2, 1), --i;
    else n_cigar = push_cigar(cigar, n_cigar, cap, 1, 1), --k;
  }
  if (i >= 0) n_cigar = push_cigar(cigar, n_cigar, cap, 2, i + 1);
  if (k >= 0) n_cigar = push_cigar(cigar, n_cigar, cap, 1, k + 1);
  if (n_cigar <= cap)
    for (int a = 0; a < n_cigar >> 1; ++a) {
      uint32_t tmp = cigar[a];
      cigar[a] = cigar[n_cigar - 1 - a];
      cigar[n_cigar - 1 - a] = tmp;
    }
  free(z); free(qp); free(eh_h); free(eh_e);
  *n_cigar_ = n_cigar;
  return score;
}

/* mem_approx_mapq_se, bwa/src/aux.cu:676-700, with sub = csub = sub_n = 0 and
 * frac_rep = 0 (DR-3). */
int orc_mapq(const orc_opt *opt, int score, int qb, int qe, int64_t rb, int64_t re) {
  int mapq, l, sub = opt->min_seed_len * opt->a;
  double identity;
  if (sub >= score) return 0;
  l = qe - qb > re - rb ? qe - qb : (int)(re - rb);
  identity = 1. - (double)(l * opt->a - score) / (opt->a + opt->b) / l;
  if (score == 0) mapq = 0;
  else {
    double tmp = l < opt->mapQ_coef_len ? 1. : opt->mapQ_coef_fac / logf((float)l);
    tmp *= identity * identity;
    mapq = (int)(6.02 * (score - sub) / opt->a * tmp * tmp + .499);
  }
  if (mapq > 60) mapq = 60;
  if (mapq < 0) mapq = 0;
  mapq = (int)(mapq * (1. - 0.0) + .499);
  return mapq;
}

/* ------------------------------------------------------------------------ */
/* paired-end primitives (device ports the reference never launches)          */
/* ------------------------------------------------------------------------ */
/* bwa/src/aux.cuh:404-409 */
#define ORC_MIN_RATIO 0.8
#define ORC_MIN_DIR_CNT 10
#define ORC_MIN_DIR_RATIO 0.05
#define ORC_OUTLIER_BOUND 2.0
#define ORC_MAPPING_BOUND 3.0
#define ORC_MAX_STDDEV 4.0

/* mem_infer_dir, bwa/src/aux.cu:411-418 */
int orc_infer_dir(int64_t l_pac, int64_t b1, int64_t b2, int64_t *dist) {
  int64_t p2;
  int r1 = (b1 >= l_pac), r2 = (b2 >= l_pac);
  p2 = r1 == r2 ? b2 : (l_pac << 1) - 1 - b2;
  *dist = p2 > b1 ? p2 - b1 : b1 - p2;
  return (r1 == r2 ? 0 : 1) ^ (p2 > b1 ? 0 : 3);
}

/* cal_sub, bwa/src/aux.cu:420-431, over one read's hits (best first) */
int orc_cal_sub(const orc_opt *opt, int n, const int32_t *qb, const int32_t *qe, const int32_t *score) {
  int j;
  for (j = 1; j < n; ++j) {
    int b_max = qb[j] > qb[0] ? qb[j] : qb[0];
    int e_min = qe[j] < qe[0] ? qe[j] : qe[0];
    if (e_min > b_max) {
      int min_l = qe[j] - qb[j] < qe[0] - qb[0] ? qe[j] - qb[j] : qe[0] - qb[0];
      if (e_min - b_max >= min_l * opt->mask_level) break;
    }
  }
  return j < n ? score[j] : opt->min_seed_len * opt->a;
}

static int cmp_u64(const void *x, const void *y) {
  uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
  return a < b ? -1 : a > b;
}

/* mem_pestat, bwa/src/aux.cu:435-500: reads 2i and 2i+1 are the ends of
 * pair i; each read's hits are aln_off[r]..aln_off[r+1], best first.  The
 * sizes of each orientation are collected, sorted and scanned as there. */
void orc_mem_pestat(const orc_opt *opt, int64_t l_pac, int max_ins, int n, const int64_t *aln_off,
                    const int32_t *rid, const int64_t *rb, const int32_t *qb, const int32_t *qe,
                    const int32_t *score, orc_pestat pes[4]) {
  uint64_t *isize[4];
  size_t cnt[4] = {0, 0, 0, 0}, max = 0;
  int i, d;
  memset(pes, 0, 4 * sizeof(orc_pestat));
  for (d = 0; d < 4; ++d) isize[d] = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n / 2 + 1));
  for (i = 0; i < n >> 1; ++i) {
    int64_t a0 = aln_off[2 * i], a1 = aln_off[2 * i + 1], e1 = aln_off[2 * i + 2], is;
    int dir;
    if (a1 == a0 || e1 == a1) continue;
    if (orc_cal_sub(opt, (int)(a1 - a0), qb + a0, qe + a0, score + a0) > ORC_MIN_RATIO * score[a0]) continue;
    if (orc_cal_sub(opt, (int)(e1 - a1), qb + a1, qe + a1, score + a1) > ORC_MIN_RATIO * score[a1]) continue;
    if (rid[a0] != rid[a1]) continue;
    dir = orc_infer_dir(l_pac, rb[a0], rb[a1], &is);
    if (is && is <= max_ins) isize[dir][cnt[dir]++] = (uint64_t)is;
  }
  for (d = 0; d < 4; ++d) {
    orc_pestat *r = &pes[d];
    uint64_t *q = isize[d];
    size_t qn = cnt[d], k;
    int p25, p50, p75, x;
    if (qn < ORC_MIN_DIR_CNT) {
      r->failed = 1;
      continue;
    }
    qsort(q, qn, sizeof(uint64_t), cmp_u64);
    p25 = (int)q[(int)(.25 * qn + .499)];
    p50 = (int)q[(int)(.50 * qn + .499)];
    p75 = (int)q[(int)(.75 * qn + .499)];
    (void)p50;
    r->low = (int)(p25 - ORC_OUTLIER_BOUND * (p75 - p25) + .499);
    if (r->low < 1) r->low = 1;
    r->high = (int)(p75 + ORC_OUTLIER_BOUND * (p75 - p25) + .499);
    for (k = 0, x = 0, r->avg = 0; k < qn; ++k)
      if (q[k] >= (uint64_t)r->low && q[k] <= (uint64_t)r->high) r->avg += q[k], ++x;
    r->avg /= x;
    for (k = 0, r->std = 0; k < qn; ++k)
      if (q[k] >= (uint64_t)r->low && q[k] <= (uint64_t)r->high) r->std += (q[k] - r->avg) * (q[k] - r->avg);
    r->std = sqrt(r->std / x);
    r->low = (int)(p25 - ORC_MAPPING_BOUND * (p75 - p25) + .499);
    r->high = (int)(p75 + ORC_MAPPING_BOUND * (p75 - p25) + .499);
    if (r->low > r->avg - ORC_MAX_STDDEV * r->std) r->low = (int)(r->avg - ORC_MAX_STDDEV * r->std + .499);
    if (r->high < r->avg + ORC_MAX_STDDEV * r->std) r->high = (int)(r->avg + ORC_MAX_STDDEV * r->std + .499);
    if (r->low < 1) r->low = 1;
  }
  for (d = 0; d < 4; ++d) max = max > cnt[d] ? max : cnt[d];
  for (d = 0; d < 4; ++d)
    if (pes[d].failed == 0 && cnt[d] < (int)max * ORC_MIN_DIR_RATIO) pes[d].failed = 1;
  for (d = 0; d < 4; ++d) free(isize[d]);
}

/* hash_64, bwa/src/aux.cu:703-714 (Thomas Wang's 64-bit integer mix) */
uint64_t orc_hash_64(uint64_t key) {
  key += ~(key << 32);
  key ^= (key >> 22);
  key += ~(key << 13);
  key ^= (key >> 8);
  key += (key << 3);
  key ^= (key >> 15);
  key += ~(key << 27);
  key ^= (key >> 31);
  return key;
}

typedef struct { uint64_t x, y; } orc_pair64;
static int cmp_pair64(const void *p, const void *q) {
  const orc_pair64 *a = (const orc_pair64 *)p, *b = (const orc_pair64 *)q;
  if (a->x != b->x) return a->x < b->x ? -1 : 1;
  return a->y < b->y ? -1 : a->y > b->y;
}

/* mem_pair, bwa/src/aux.cu:1000-1070, for one pair whose ends have n0 and n1
 * hits (all primary), given as arrays of n0 + n1 entries (read 1's first).
 * Returns the best pair score (0 if none) and sets *sub, *n_sub and z[]. */
int orc_mem_pair(const orc_opt *opt, int64_t l_pac, const int64_t *ctg_off, const orc_pestat pes[4], int id,
                 int n0, int n1, const int32_t *rid, const int64_t *rb, const int32_t *score, int *sub,
                 int *n_sub, int z[2]) {
  orc_pair64 *v, *u;
  size_t nv = 0, nu = 0, mu = 16;
  int r, i, k, y[4], ret;
  int nn[2] = {n0, n1};
  v = (orc_pair64 *)malloc(sizeof(orc_pair64) * (size_t)(n0 + n1 + 1));
  u = (orc_pair64 *)malloc(sizeof(orc_pair64) * mu);
  for (r = 0; r < 2; ++r) {
    for (i = 0; i < nn[r]; ++i) {
      int e = r ? n0 + i : i;
      orc_pair64 key;
      key.x = rb[e] < l_pac ? (uint64_t)rb[e] : (uint64_t)((l_pac << 1) - 1 - rb[e]);
      key.x = (uint64_t)rid[e] << 32 | (key.x - (uint64_t)ctg_off[rid[e]]);
      key.y = (uint64_t)score[e] << 32 | (uint64_t)(i << 2 | (rb[e] >= l_pac) << 1 | r);
      v[nv++] = key;
    }
  }
  qsort(v, nv, sizeof(orc_pair64), cmp_pair64);
  y[0] = y[1] = y[2] = y[3] = -1;
  for (i = 0; i < (int)nv; ++i) {
    for (r = 0; r < 2; ++r) {
      int dir = r << 1 | (int)(v[i].y >> 1 & 1), which;
      if (pes[dir].failed) continue;
      which = r << 1 | (int)((v[i].y & 1) ^ 1);
      if (y[which] < 0) continue;
      for (k = y[which]; k >= 0; --k) {
        int64_t dist;
        int q;
        double ns;
        if ((int)(v[k].y & 3) != which) continue;
        dist = (int64_t)v[i].x - (int64_t)v[k].x;
        if (dist > pes[dir].high) break;
        if (dist < pes[dir].low) continue;
        ns = (dist - pes[dir].avg) / pes[dir].std;
        {
          double qd = (v[i].y >> 32) + (v[k].y >> 32) + .721 * log(2. * erfc(fabs(ns) * M_SQRT1_2)) * opt->a + .499;
          q = qd > 0 ? (int)qd : 0; /* the reference's (int) of -inf or NaN, then clamped at 0 */
        }
        if (nu == mu) {
          mu <<= 1;
          u = (orc_pair64 *)realloc(u, sizeof(orc_pair64) * mu);
        }
        u[nu].y = (uint64_t)k << 32 | (uint64_t)(uint32_t)i;
        u[nu].x = (uint64_t)q << 32 |
                  (orc_hash_64(u[nu].y ^ (uint64_t)(int64_t)(int32_t)((uint32_t)id << 8)) & 0xffffffffU);
        ++nu;
      }
    }
    y[v[i].y & 3] = i;
  }
  if (nu) {
    int tmp = opt->a + opt->b;
    tmp = tmp > opt->o_del + opt->e_del ? tmp : opt->o_del + opt->e_del;
    tmp = tmp > opt->o_ins + opt->e_ins ? tmp : opt->o_ins + opt->e_ins;
    qsort(u, nu, sizeof(orc_pair64), cmp_pair64);
    i = (int)(u[nu - 1].y >> 32);
    k = (int)(u[nu - 1].y << 32 >> 32);
    z[v[i].y & 1] = (int)(v[i].y << 32 >> 34);
    z[v[k].y & 1] = (int)(v[k].y << 32 >> 34);
    ret = (int)(u[nu - 1].x >> 32);
    *sub = nu > 1 ? (int)(u[nu - 2].x >> 32) : 0;
    for (i = (int)nu - 2, *n_sub = 0; i >= 0; --i)
      if (*sub - (int)(u[i].x >> 32) <= tmp) ++*n_sub;
  } else ret = 0, *sub = 0, *n_sub = 0;
  free(u);
  free(v);
  return ret;
}

/* ------------------------------------------------------------------------ */
/* per-read pipeline                                                         */
/* ------------------------------------------------------------------------ */
typedef struct {
  int64_t rb, re;
  int qb, qe, rid, score, seedlen0, secondary, is_alt;
  int idx; /* position in the read's task order (stage dumps) */
} reg_t;

static void cigar_append(uint32_t **buf, int64_t *n, int64_t *m, const uint32_t *src, int k) {
  if (*n + k > *m) {
    while (*n + k > *m) *m = *m ? *m * 2 : 1024;
    *buf = (uint32_t *)realloc(*buf, (size_t)*m * 4);
    if (!*buf) abort();
  }
  memcpy(*buf + *n, src, (size_t)k * 4);
  *n += k;
}

/* bns_get_seq_gpu, bwa/src/bntseq.cu:39-60 */
static uint8_t *get_seq(const orc_index *x, int64_t beg, int64_t end, int64_t *len) {
  int64_t l_pac = x->l_pac;
  if (end < beg) { int64_t t = beg; beg = end; end = t; }
  if (end > l_pac << 1) end = l_pac << 1;
  if (beg < 0) beg = 0;
  if (beg >= l_pac || end <= l_pac) {
    uint8_t *s = (uint8_t *)xmalloc((size_t)(end - beg + 1));
    int64_t l = 0;
    *len = end - beg;
    if (beg >= l_pac) {
      int64_t beg_f = (l_pac << 1) - 1 - end, end_f = (l_pac << 1) - 1 - beg;
      for (int64_t k = end_f; k > beg_f; --k) s[l++] = (uint8_t)(3 - get_pac(x->pac, k));
    } else {
      for (int64_t k = beg; k < end; ++k) s[l++] = (uint8_t)get_pac(x->pac, k);
    }
    return s;
  }
  *len = 0;
  return (uint8_t *)xmalloc(1);
}

/* cal_max_gap, bwa/src/aux.cu:168-175 */
static int cal_max_gap(const orc_opt *opt, int qlen) {
  int l_del = (int)((double)(qlen * opt->a - opt->o_del) / opt->e_del + 1.);
  int l_ins = (int)((double)(qlen * opt->a - opt->o_ins) / opt->e_ins + 1.);
  int l = l_del > l_ins ? l_del : l_ins;
  l = l > 1 ? l : 1;
  return l < opt->w << 1 ? l : opt->w << 1;
}

/* infer_bw, bwa/src/aux.cu:506-513 */
static int infer_bw(int l1, int l2, int score, int a, int q, int r) {
  int w;
  if (l1 == l2 && l1 * a - score < (q + r - a) << 1) return 0;
  w = (int)((double)((l1 < l2 ? l1 : l2) * a - score - q) / r + 2.);
  if (w < abs(l1 - l2)) w = abs(l1 - l2);
  return w;
}

typedef struct {
  orc_aln *a;
  int n, m;
  uint32_t *cig;
  int64_t nc, mc;
  int64_t fm_steps, lf_steps, sa_reads, ext_cells, ext_tasks, tb_cells;
} read_out_t;

static void out_push(read_out_t *o, const orc_aln *a) {
  if (o->n == o->m) {
    o->m = o->m ? o->m * 2 : 4;
    o->a = (orc_aln *)realloc(o->a, (size_t)o->m * sizeof(orc_aln));
    if (!o->a) abort();
  }
  o->a[o->n++] = *a;
}

/* filterRegions (aux.cu:2408-2548) + sortRegions (aux.cu:2552-2613) over one
 * read's regions in task order; frees reg and returns the kept regions in
 * output order (their `idx` is the position in the input). */
static reg_t *filter_regions(const orc_index *x, const orc_opt *opt, reg_t *reg, int n_reg, int *n_out) {
  for (int i = 0; i < n_reg; ++i) reg[i].idx = i;
  /* filterRegions.
   * DR: the reference's threads race on the keep/secondary flags; the oracle
   * uses the sequential order of anchors (ascending index), and writes kept
   * regions in ascending index order (DR-8). */
  uint8_t *kept = (uint8_t *)xmalloc((size_t)(n_reg ? n_reg : 1));
  for (int i = 0; i < n_reg; ++i) kept[i] = reg[i].score >= opt->T;
  for (int i = 0; i < n_reg; ++i) {
    if (!kept[i]) continue;
    for (int j = i + 1; j < n_reg; ++j) {
      if (!kept[j]) continue;
      if (reg[i].rid != reg[j].rid) continue;
      int64_t ai_rb = reg[i].rb, ai_re = reg[i].re, aj_rb = reg[j].rb, aj_re = reg[j].re;
      if (ai_rb < aj_rb && ai_re < aj_rb) continue;
      if (aj_rb < ai_rb && aj_re < ai_rb) continue;
      int or_ = (int)(ai_rb < aj_rb ? ai_re - aj_rb : aj_re - ai_rb);
      int oq = reg[i].qb < reg[j].qb ? reg[i].qe - reg[j].qb : reg[j].qe - reg[i].qb;
      int rli = (int)(ai_re - ai_rb), rlj = (int)(aj_re - aj_rb);
      int mr = rli < rlj ? rli : rlj;
      int mq = (reg[i].qe - reg[i].qb < reg[j].qe - reg[j].qb) ? reg[i].qe - reg[i].qb : reg[j].qe - reg[j].qb;
      if (or_ > opt->mask_level_redun * mr && oq > opt->mask_level_redun * mq) {
        if (reg[i].score < reg[j].score) { kept[i] = 0; break; }
        else kept[j] = 0;
      }
    }
  }
  int m = 0;
  for (int i = 0; i < n_reg; ++i)
    if (kept[i]) reg[m++] = reg[i];
  free(kept);
  for (int i = 0; i < m; ++i) reg[i].is_alt = (reg[i].rid >= 0 && x->contigs[reg[i].rid].is_alt) ? 1 : 0;
  int *sec = (int *)xmalloc((size_t)(m ? m : 1) * sizeof(int));
  for (int i = 0; i < m; ++i) sec[i] = -1;
  for (int i = 0; i < m; ++i) {
    for (int j = i + 1; j < m; ++j) {
      if (sec[i] >= 0) continue;
      if (sec[j] >= 0) continue;
      int qb_max = reg[i].qb > reg[j].qb ? reg[i].qb : reg[j].qb;
      int qe_min = reg[i].qe < reg[j].qe ? reg[i].qe : reg[j].qe;
      if (qe_min > qb_max) {
        int li = reg[i].qe - reg[i].qb, lj = reg[j].qe - reg[j].qb, min_l = li < lj ? li : lj;
        if (qe_min - qb_max >= min_l * opt->mask_level) {
          if (reg[i].score > reg[j].score) {
            if (reg[i].is_alt && reg[j].is_alt) sec[j] = i;
            else if (!reg[i].is_alt) sec[j] = i;
          } else {
            if (reg[j].is_alt && reg[i].is_alt) sec[i] = j;
            else if (!reg[j].is_alt) sec[i] = j;
          }
        }
      }
    }
  }
  int nk = 0;
  reg_t *kr = (reg_t *)xmalloc((size_t)(m ? m : 1) * sizeof(reg_t));
  for (int i = 0; i < m; ++i) {
    if (sec[i] >= 0 && (reg[i].is_alt || !(opt->flag & 0x8))) continue;
    if (sec[i] >= 0 && reg[i].score < reg[sec[i]].score * opt->drop_ratio) continue;
    kr[nk] = reg[i];
    kr[nk].secondary = sec[i];
    ++nk;
  }
  free(sec);
  free(reg);
  /* sortRegions, aux.cu:2552-2613 */
  if (nk == 2) {
    if (kr[0].is_alt > kr[1].is_alt || (kr[0].is_alt <= kr[1].is_alt && kr[0].score < kr[1].score)) {
      reg_t t = kr[0]; kr[0] = kr[1]; kr[1] = t;
    }
  } else if (nk > 2) {
    for (int i = 0; i < nk - 1; ++i) {
      int swapped = 0;
      for (int j = 0; j < nk - 1 - i; ++j) {
        if (kr[j].is_alt > kr[j + 1].is_alt || kr[j].score < kr[j + 1].score) {
          reg_t t = kr[j]; kr[j] = kr[j + 1]; kr[j + 1] = t;
          swapped = 1;
        }
      }
      if (!swapped) break;
    }
  }
  *n_out = nk;
  return kr;
}

/* Stage dump for tests: the region filter and sort of one read's regions
 * (rb, re, qb, qe, score, rid in task order).  out_idx[k] = input position of
 * the k-th output region, out_sec[k] = its secondary anchor (an index into
 * the filtered list before sorting, -1 if primary).  Returns the count. */
int orc_region_dump(const orc_index *x, const orc_opt *opt, int n, const int64_t *rb, const int64_t *re,
                    const int32_t *qb, const int32_t *qe, const int32_t *score, const int32_t *rid, int32_t *out_idx,
                    int32_t *out_sec) {
  reg_t *reg = (reg_t *)xcalloc((size_t)(n ? n : 1), sizeof(reg_t));
  for (int i = 0; i < n; ++i) {
    reg[i].rb = rb[i]; reg[i].re = re[i]; reg[i].qb = qb[i]; reg[i].qe = qe[i];
    reg[i].score = score[i]; reg[i].rid = rid[i];
  }
  int nk = 0;
  reg_t *kr = filter_regions(x, opt, reg, n, &nk);
  for (int i = 0; i < nk; ++i) { out_idx[i] = kr[i].idx; out_sec[i] = kr[i].secondary; }
  free(kr);
  return nk;
}

static void align_read(const orc_index *x, const orc_opt *opt, const uint8_t *read, int len, read_out_t *o) {
  const int64_t l_pac = x->l_pac;
  intv_v mem = {0, 0, 0};
  seed_read(x, opt, read, len, &mem, &o->fm_steps);
  int n_seeds = 0;
  seed_t *seeds = sa_lookup(x, opt, &mem, &n_seeds, &o->lf_steps, &o->sa_reads);
  free(mem.a);
  sort_seeds(seeds, n_seeds);
  int n_ch = 0;
  chain_t *ch = n_seeds ? chain_seeds(x, opt, seeds, n_seeds, &n_ch, NULL) : NULL;
  free(seeds);
  n_ch = filter_chains(opt, ch, n_ch);
  /* SWSeed + ExtendingPairGenerate + localExtending, aux.cu:2041-2386 */
  int n_reg = 0;
  for (int i = 0; i < n_ch; ++i) n_reg += ch[i].n;
  reg_t *reg = (reg_t *)xcalloc((size_t)(n_reg ? n_reg : 1), sizeof(reg_t));
  int r = 0;
  for (int ci = 0; ci < n_ch; ++ci) {
    const chain_t *c = &ch[ci];
    int64_t rmax0 = l_pac << 1, rmax1 = 0;
    for (int k = 0; k < c->n; ++k) {
      const seed_t *s = &c->seeds[k];
      int64_t b = s->rbeg - (s->qbeg + cal_max_gap(opt, s->qbeg));
      int64_t e = s->rbeg + s->len + ((len - s->qbeg - s->len) + cal_max_gap(opt, len - s->qbeg - s->len));
      rmax0 = rmax0 < b ? rmax0 : b;
      rmax1 = rmax1 > e ? rmax1 : e;
    }
    rmax0 = rmax0 > 0 ? rmax0 : 0;
    rmax1 = rmax1 < l_pac << 1 ? rmax1 : l_pac << 1;
    if (rmax0 < l_pac && l_pac < rmax1) {
      if (c->seeds[0].rbeg < l_pac) rmax1 = l_pac;
      else rmax0 = l_pac;
    }
    /* bns_fetch_seq_gpu, bntseq.cu:62-87 */
    int is_rev;
    int rid = pos2rid(x, depos(x, c->seeds[0].rbeg, &is_rev));
    int64_t far_beg = x->contigs[rid].offset, far_end = far_beg + x->contigs[rid].len;
    if (is_rev) { int64_t t = far_beg; far_beg = (l_pac << 1) - far_end; far_end = (l_pac << 1) - t; }
    rmax0 = rmax0 > far_beg ? rmax0 : far_beg;
    rmax1 = rmax1 < far_end ? rmax1 : far_end;
    int64_t rlen_all;
    uint8_t *rseq = get_seq(x, rmax0, rmax1, &rlen_all);
    for (int k = 0; k < c->n; ++k, ++r) {
      const seed_t *s = &c->seeds[k];
      reg_t *g = &reg[r];
      g->rid = rid;
      int qle, tle, score;
      int h0 = s->len * opt->a;
      int qlen = s->qbeg, tlen = (int)(s->rbeg - rmax0);
      if (qlen > 0) {
        uint8_t *qs = (uint8_t *)xmalloc((size_t)qlen), *ts = (uint8_t *)xmalloc((size_t)(tlen > 0 ? tlen : 1));
        for (int a = 0; a < qlen; ++a) qs[a] = read[qlen - 1 - a];
        for (int a = 0; a < tlen; ++a) ts[a] = rseq[tlen - 1 - a];
        score = orc_ksw_extend(qlen, qs, tlen, ts, opt->mat, opt->o_del, opt->e_del, opt->o_ins, opt->e_ins, opt->w,
                               opt->pen_clip5, h0, &qle, &tle, &o->ext_cells);
        ++o->ext_tasks;
        g->qb = s->qbeg - qle;
        g->rb = s->rbeg - tle;
        free(qs); free(ts);
      } else {
        score = h0; g->qb = 0; g->rb = s->rbeg;
      }
      h0 = score;
      qlen = len - s->qbeg - s->len;
      tlen = (int)(rmax1 - s->rbeg - s->len);
      if (qlen > 0) {
        score = orc_ksw_extend(qlen, read + s->qbeg + s->len, tlen, rseq + (s->rbeg - rmax0) + s->len, opt->mat,
                               opt->o_del, opt->e_del, opt->o_ins, opt->e_ins, opt->w, opt->pen_clip3, h0, &qle, &tle,
                               &o->ext_cells);
        ++o->ext_tasks;
        g->qe = s->qbeg + s->len + qle;
        g->re = s->rbeg + s->len + tle;
      } else {
        score = h0; g->qe = s->qbeg + s->len; g->re = s->rbeg + s->len;
      }
      g->score = score;
      g->seedlen0 = s->len;
    }
    free(rseq);
  }
  for (int i = 0; i < n_ch; ++i) free(ch[i].seeds);
  free(ch);
  int nk = 0;
  reg_t *kr = filter_regions(x, opt, reg, n_reg, &nk);
  /* FINALIZEALN_preprocessing2 + traceback + finalize, aux.cu:2676-3059 */
  for (int ai = 0; ai < nk; ++ai) {
    const reg_t *g = &kr[ai];
    int lq = g->qe - g->qb;
    const uint8_t *query = read + g->qb;
    int64_t rlen;
    uint8_t *rseq = get_seq(x, g->rb, g->re, &rlen);
    int w;
    if (lq == g->re - g->rb) w = 0;
    else {
      int tmp;
      w = infer_bw(lq, (int)(g->re - g->rb), g->score, opt->a, opt->o_del, opt->e_del);
      tmp = infer_bw(lq, (int)(g->re - g->rb), g->score, opt->a, opt->o_ins, opt->e_ins);
      w = w > tmp ? w : tmp;
      int max_ins = (int)((double)(((lq + 1) >> 1) * opt->a - opt->o_ins) / opt->e_ins + 1.);
      int max_del = (int)((double)(((lq + 1) >> 1) * opt->a - opt->o_del) / opt->e_del + 1.);
      int max_gap = max_ins > max_del ? max_ins : max_del;
      max_gap = max_gap > 1 ? max_gap : 1;
      tmp = (max_gap + abs((int)rlen - lq) + 1) >> 1;
      w = w < tmp ? w : tmp;
      tmp = abs((int)rlen - lq) + 3;
      w = w > tmp ? w : tmp;
    }
    int is_rev_strand = g->rb >= l_pac;
    if (w >= 500) w = 500; /* GLOBALSW_BANDWITH_CUTOFF, aux.cuh:428 */
    int cap = lq + (int)rlen + 4, n_core = 0, nw;
    uint32_t *core = (uint32_t *)xmalloc((size_t)cap * 4);
    if (w == 0) {
      core[0] = (uint32_t)(lq << 4);
      n_core = 1;
      nw = 0; /* DR-11: the reference's score here comes from a shadowed variable */
      for (int i = 0; i < lq; ++i) nw += opt->mat[rseq[i] * 5 + query[i]];
    } else {
      /* DR-9: ksw_global3's int8_t state and shared-memory race are replaced by
       * the defined semantics of ksw_global2 (ksw.cu:706-812) */
      nw = orc_ksw_global2(lq, query, (int)rlen, rseq, opt->mat, opt->o_del, opt->e_del, opt->o_ins, opt->e_ins, w,
                           &n_core, core, cap);
      int64_t band = 0;
      for (int i = 0; i < (int)rlen; ++i) {
        int beg = i > w ? i - w : 0, end = i + w + 1 < lq ? i + w + 1 : lq;
        band += end > beg ? end - beg : 0;
      }
      o->tb_cells += band;
    }
    /* NM as the traceback counts it (aux.cu:2910-2941) */
    int n_mm = 0, n_gap = 0;
    {
      int xq = 0, yr = 0;
      for (int k = 0; k < n_core; ++k) {
        int op = core[k] & 0xf, ln = (int)(core[k] >> 4);
        if (op == 0) {
          for (int i = 0; i < ln; ++i) n_mm += query[xq + i] != rseq[yr + i];
          xq += ln; yr += ln;
        } else if (op == 2) {
          if (k > 0 && k < n_core - 1) n_gap += ln;
          yr += ln;
        } else if (op == 1) { xq += ln; n_gap += ln; }
      }
    }
    free(rseq);
    /* finalize, aux.cu:2952-3059 */
    orc_aln a;
    memset(&a, 0, sizeof a);
    int is_rev;
    int64_t pos = depos(x, !is_rev_strand ? g->rb : g->re - 1, &is_rev);
    const uint32_t *cg = core;
    int nc = n_core;
    if (nc > 0) {
      if ((cg[0] & 0xf) == 2) { pos += cg[0] >> 4; --nc; cg = &cg[1]; }
      else if ((cg[nc - 1] & 0xf) == 2) --nc;
    }
    uint32_t *fin = (uint32_t *)xmalloc((size_t)(n_core + 2) * 4);
    int nf;
    if (g->qb != 0 || g->qe != len) {
      int clip5 = is_rev ? len - g->qe : g->qb, clip3 = is_rev ? g->qb : len - g->qe;
      nf = 0;
      if (clip5) fin[nf++] = (uint32_t)(clip5 << 4 | 3);
      memcpy(fin + nf, cg, (size_t)nc * 4);
      nf += nc;
      if (clip3) fin[nf++] = (uint32_t)(clip3 << 4 | 3);
    } else {
      /* DR-12: without clipping the squeezed CIGAR is not stored back */
      memcpy(fin, core, (size_t)n_core * 4);
      nf = n_core;
    }
    a.rid = pos2rid(x, pos);
    a.pos = pos - x->contigs[a.rid].offset;
    a.is_rev = is_rev;
    int flag = 0;
    if (g->secondary >= 0) flag |= 0x100;
    else if (ai > 0) flag |= 0x800;
    if (g->rid < 0) flag |= 0x4;
    if (is_rev) flag |= 0x10;
    a.flag = flag;
    a.score = g->score;
    a.nw_score = nw;
    a.NM = n_mm + n_gap;
    a.qb = g->qb; a.qe = g->qe; a.rb = g->rb; a.re = g->re;
    a.n_cigar = nf;
    a.n_core = n_core;
    a.cigar_off = o->nc;
    cigar_append(&o->cig, &o->nc, &o->mc, fin, nf);
    cigar_append(&o->cig, &o->nc, &o->mc, core, n_core);
    out_push(o, &a);
    free(fin);
    free(core);
  }
  free(kr);
}

/* ------------------------------------------------------------------------ */
/* batch driver                                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
  const orc_index *x;
  const orc_opt *opt;
  const uint8_t *seq;
  const int32_t *off;
  int n, next;
  pthread_mutex_t mu;
  read_out_t *outs;
} job_t;

static void *worker(void *arg) {
  job_t *j = (job_t *)arg;
  for (;;) {
    pthread_mutex_lock(&j->mu);
    int b = j->next;
    j->next += 64;
    pthread_mutex_unlock(&j->mu);
    if (b >= j->n) break;
    int e = b + 64 < j->n ? b + 64 : j->n;
    for (int i = b; i < e; ++i) {
      int len = j->off[i + 1] - j->off[i];
      /* DR-20: reads longer than MAX_LEN_READ (bwa/src/macro.h:27) overflow
       * the reference's shared arrays; here they are left unaligned. */
      if (len > 320) continue;
      /* codes above 4 (no base) are N, as the reference's encoder maps every
       * non-ACGT character (pipeline.cpp:90-107) */
      uint8_t rd[320];
      for (int k = 0; k < len; ++k) rd[k] = j->seq[j->off[i] + k] > 4 ? 4 : j->seq[j->off[i] + k];
      align_read(j->x, j->opt, rd, len, &j->outs[i]);
    }
  }
  return NULL;
}

int orc_align_batch(const orc_index *x, const orc_opt *opt, int n, const uint8_t *seq, const int32_t *offsets,
                    int n_threads, orc_batch *out) {
  memset(out, 0, sizeof *out);
  for (int i = 0; i < n; ++i)
    if (offsets[i + 1] < offsets[i]) return -1;
  job_t j;
  j.x = x; j.opt = opt; j.seq = seq; j.off = offsets; j.n = n; j.next = 0;
  pthread_mutex_init(&j.mu, NULL);
  j.outs = (read_out_t *)xcalloc((size_t)(n ? n : 1), sizeof(read_out_t));
  if (n_threads < 1) n_threads = 1;
  pthread_t *th = (pthread_t *)xmalloc((size_t)n_threads * sizeof(pthread_t));
  for (int t = 0; t < n_threads; ++t) pthread_create(&th[t], NULL, worker, &j);
  for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
  free(th);
  pthread_mutex_destroy(&j.mu);
  out->n_reads = n;
  out->aln_off = (int32_t *)xmalloc((size_t)(n + 1) * sizeof(int32_t));
  int64_t na = 0, nc = 0;
  for (int i = 0; i < n; ++i) { out->aln_off[i] = (int32_t)na; na += j.outs[i].n; nc += j.outs[i].nc; }
  out->aln_off[n] = (int32_t)na;
  out->n_alns = na;
  out->alns = (orc_aln *)xmalloc((size_t)(na ? na : 1) * sizeof(orc_aln));
  out->cigar = (uint32_t *)xmalloc((size_t)(nc ? nc : 1) * 4);
  int64_t ka = 0, kc = 0;
  for (int i = 0; i < n; ++i) {
    read_out_t *o = &j.outs[i];
    for (int a = 0; a < o->n; ++a) {
      out->alns[ka] = o->a[a];
      out->alns[ka].cigar_off += kc;
      ++ka;
    }
    if (o->nc) memcpy(out->cigar + kc, o->cig, (size_t)o->nc * 4);
    kc += o->nc;
    out->fm_steps += o->fm_steps; out->lf_steps += o->lf_steps; out->sa_reads += o->sa_reads;
    out->ext_cells += o->ext_cells; out->ext_tasks += o->ext_tasks; out->tb_cells += o->tb_cells;
    free(o->a);
    free(o->cig);
  }
  out->n_cigar_words = kc;
  free(j.outs);
  return 0;
}

void orc_batch_free(orc_batch *b) {
  free(b->aln_off); free(b->alns); free(b->cigar);
  memset(b, 0, sizeof *b);
}
