// file_exists / file_size (core/filesystem.hpp:27-36).
#ifndef GG_CORE_FILESYSTEM_HPP
#define GG_CORE_FILESYSTEM_HPP
#include <sys/stat.h>
inline bool file_exists(const char *path) {
  struct stat st;
  return stat(path, &st) == 0;
}
inline long file_size(const char *path) {
  struct stat st;
  if (stat(path, &st) != 0) return -1;
  return st.st_size;
}
#endif
