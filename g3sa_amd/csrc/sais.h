// Suffix array construction by induced sorting (SA-IS; Nong, Zhang & Chan,
// "Two Efficient Algorithms for Linear Time Suffix Array Construction", 2011).
// Written for this project.  The input string must end with a unique,
// smallest sentinel symbol 0; all other symbols lie in [1, K).
#pragma once
#include <cstdint>
#include <cstring>
#include <vector>

namespace g3a {

// Progress hook of the index build (null: silent): phase, recursion level,
// length of the string being sorted.
inline void (*sais_progress)(const char *what, int level, int64_t n) = nullptr;

template <typename I>
struct SaisTypes {
  std::vector<uint8_t> bits;  // 1 = S-type
  explicit SaisTypes(size_t n) : bits((n + 7) / 8, 0) {}
  bool s(size_t i) const { return (bits[i >> 3] >> (i & 7)) & 1; }
  void set(size_t i, bool v) {
    if (v) bits[i >> 3] |= (uint8_t)(1u << (i & 7));
    else bits[i >> 3] &= (uint8_t)~(1u << (i & 7));
  }
  bool lms(size_t i) const { return i > 0 && s(i) && !s(i - 1); }
};

template <typename C, typename I>
static void sais_buckets(const C *s, I n, I K, std::vector<I> &bkt, bool end) {
  std::fill(bkt.begin(), bkt.end(), (I)0);
  for (I i = 0; i < n; ++i) ++bkt[(size_t)s[i]];
  I sum = 0;
  for (I c = 0; c < K; ++c) {
    sum += bkt[(size_t)c];
    bkt[(size_t)c] = end ? sum : sum - bkt[(size_t)c];
  }
}

template <typename C, typename I>
static void sais_induce(const C *s, I *SA, I n, I K, const SaisTypes<I> &t, std::vector<I> &bkt, int level) {
  const I EMPTY = (I)~(I)0;
  sais_buckets(s, n, K, bkt, false);
  for (I i = 0; i < n; ++i) {
    I p = SA[i];
    if (p != EMPTY && p > 0 && !t.s((size_t)(p - 1))) SA[bkt[(size_t)s[p - 1]]++] = p - 1;
  }
  if (sais_progress) sais_progress("L-type suffixes placed", level, (int64_t)n);
  sais_buckets(s, n, K, bkt, true);
  for (I i = n; i-- > 0;) {
    I p = SA[i];
    if (p != EMPTY && p > 0 && t.s((size_t)(p - 1))) SA[--bkt[(size_t)s[p - 1]]] = p - 1;
  }
}

// SA of s[0..n); s[n-1] == 0 is the unique minimum.
template <typename C, typename I>
void sais(const C *s, I *SA, I n, I K, int level = 0) {
  const I EMPTY = (I)~(I)0;
  SaisTypes<I> t((size_t)n);
  t.set((size_t)(n - 1), true);
  for (I i = n - 1; i-- > 0;)
    t.set((size_t)i, s[i] < s[i + 1] || (s[i] == s[i + 1] && t.s((size_t)(i + 1))));
  std::vector<I> bkt((size_t)K);
  // stage 1: sort LMS substrings
  sais_buckets(s, n, K, bkt, true);
  for (I i = 0; i < n; ++i) SA[i] = EMPTY;
  for (I i = 1; i < n; ++i)
    if (t.lms((size_t)i)) SA[--bkt[(size_t)s[i]]] = i;
  sais_induce(s, SA, n, K, t, bkt, level);
  if (sais_progress) sais_progress("LMS substrings sorted", level, (int64_t)n);
  I n1 = 0;
  for (I i = 0; i < n; ++i)
    if (t.lms((size_t)SA[i])) SA[n1++] = SA[i];
  for (I i = n1; i < n; ++i) SA[i] = EMPTY;
  // name LMS substrings; equal substrings share a name
  I name = 0, prev = EMPTY;
  for (I i = 0; i < n1; ++i) {
    I pos = SA[i];
    bool diff = false;
    for (I d = 0;; ++d) {
      if (prev == EMPTY || s[pos + d] != s[prev + d] || t.s((size_t)(pos + d)) != t.s((size_t)(prev + d))) {
        diff = true;
        break;
      }
      if (d > 0 && (t.lms((size_t)(pos + d)) || t.lms((size_t)(prev + d)))) break;
    }
    if (diff) { ++name; prev = pos; }
    SA[n1 + pos / 2] = name - 1;
  }
  for (I i = n, j = n; i-- > n1;)
    if (SA[i] != EMPTY) SA[--j] = SA[i];
  // stage 2: sort the reduced string (recursively if names repeat)
  I *s1 = SA + n - n1, *SA1 = SA;
  if (sais_progress) sais_progress("LMS substrings named", level, (int64_t)n);
  if (name < n1) {
    sais<I, I>(s1, SA1, n1, name, level + 1);
  } else {
    for (I i = 0; i < n1; ++i) SA1[s1[i]] = i;
  }
  // stage 3: induce the full SA from the sorted LMS suffixes
  sais_buckets(s, n, K, bkt, true);
  for (I i = 1, j = 0; i < n; ++i)
    if (t.lms((size_t)i)) s1[j++] = i;
  for (I i = 0; i < n1; ++i) SA1[i] = s1[SA1[i]];
  for (I i = n1; i < n; ++i) SA[i] = EMPTY;
  for (I i = n1; i-- > 0;) {
    I j = SA[i];
    SA[i] = EMPTY;
    SA[--bkt[(size_t)s[j]]] = j;
  }
  sais_induce(s, SA, n, K, t, bkt, level);
  if (sais_progress) sais_progress("suffixes induced", level, (int64_t)n);
}

}  // namespace g3a
