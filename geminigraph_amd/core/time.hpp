// Wall clock in seconds (core/time.hpp:22-26).
#ifndef GG_CORE_TIME_HPP
#define GG_CORE_TIME_HPP
#include <sys/time.h>
inline double get_time() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec / 1e6;
}
#endif
