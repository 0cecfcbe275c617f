// g3a: command-line driver.
//   g3a index [-s SA_INTV] [-d DEVICE] <ref.fa> [prefix]
//   g3a mem   [options] <prefix> <reads.fq>        (see g3a_mem.cpp)
// `index` writes the BWA-format index plus the dense SA sample; `mem` keeps the
// reference's `g3.exe mem` flags (bwa/src/fastmap.cpp:443-484).
#include "g3a_index.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

int g3a_mem_main(int argc, char **argv);

static std::chrono::steady_clock::time_point g_t0;
static void index_progress(const char *what, int level, int64_t n) {
  const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - g_t0).count();
  std::fprintf(stderr, "g3a index: %7.1f s  %s (level %d, %lld symbols)\n", t, what, level, (long long)n);
  std::fflush(stderr);
}

static int index_main(int argc, char **argv) {
  int intv = 4, device = -1;
  int i = 1;
  for (; i < argc && argv[i][0] == '-'; ++i) {
    if (!std::strcmp(argv[i], "-s") && i + 1 < argc) intv = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "-d") && i + 1 < argc) device = std::atoi(argv[++i]);
    else { std::fprintf(stderr, "g3a index: unknown option %s\n", argv[i]); return 1; }
  }
  if (i >= argc) {
    std::fprintf(stderr, "usage: g3a index [-s SA_INTV] [-d DEVICE] <ref.fa> [prefix]\n"
                         "  -d DEVICE  build the suffix array and BWT on that HIP device (default: host SA-IS)\n");
    return 1;
  }
  const char *fa = argv[i];
  std::string prefix = i + 1 < argc ? argv[i + 1] : fa;
  std::string err;
  g_t0 = std::chrono::steady_clock::now();
  g3a::set_build_progress(index_progress);
  g3a::device_build_progress = index_progress;
  if (!g3a::build_index(fa, prefix.c_str(), intv, err, device)) {
    std::fprintf(stderr, "g3a index: %s\n", err.c_str());
    return 1;
  }
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: g3a index|mem ...\n");
    return 1;
  }
  if (!std::strcmp(argv[1], "index")) return index_main(argc - 1, argv + 1);
  if (!std::strcmp(argv[1], "mem")) return g3a_mem_main(argc - 1, argv + 1);
  std::fprintf(stderr, "g3a: unknown command %s\n", argv[1]);
  return 1;
}
