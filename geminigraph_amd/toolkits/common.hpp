// Helpers shared by the facade toolkits: optional result dumps for tests
// (GG_DUMP=<dir> writes <dir>/<name>.bin) and the run count (GG_RUNS, default
// 1 + 5 as in the reference's main()).
#ifndef GG_TOOLKITS_COMMON_HPP
#define GG_TOOLKITS_COMMON_HPP
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

template <typename T>
inline void dump_array(const char *name, const std::vector<T> &v) {
  const char *dir = getenv("GG_DUMP");
  if (!dir) return;
  std::string p = std::string(dir) + "/" + name + ".bin";
  FILE *f = fopen(p.c_str(), "wb");
  if (!f) {
    fprintf(stderr, "cannot write %s\n", p.c_str());
    exit(-1);
  }
  fwrite(v.data(), sizeof(T), v.size(), f);
  fclose(f);
}

inline int extra_runs() {
  const char *r = getenv("GG_RUNS");
  return r ? atoi(r) - 1 : 5;
}
#endif
