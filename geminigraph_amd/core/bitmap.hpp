// Bitmap / VertexSubset in device memory (core/bitmap.hpp:20-61).  The host
// object owns the words; device callbacks use a BitmapView (two words of
// state, copied into the lambda) because a host pointer cannot be followed on
// the device.
#ifndef GG_CORE_BITMAP_HPP
#define GG_CORE_BITMAP_HPP

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

#include "gg.h"

#define WORD_OFFSET(i) ((i) >> 6)
#define BIT_OFFSET(i) ((i) & 0x3f)

struct BitmapView {
  size_t size;
  unsigned long *data;
  __device__ unsigned long get_bit(size_t i) const {
    return data[WORD_OFFSET(i)] & (1ul << BIT_OFFSET(i));
  }
  // Lanes of a wave that set the bits of one word in lane order (a wave over
  // 64 consecutive vertices, as process_vertices runs) take one atomic for
  // the word.  Otherwise the calling lanes are grouped by word, up to four
  // groups (the first pending lane's word each time; sorted vertex runs,
  // such as dense slots over packed messages, span a few words): a group's
  // bits are gathered with LDS atomics in a per-wave word and its first lane
  // takes the one global atomic; lanes left after four groups take one
  // atomic each.  Blocks of at most 1024 threads.
  __device__ void set_bit(size_t i) const {
    __shared__ unsigned long long s_or[16];
    const size_t w = WORD_OFFSET(i);
    const unsigned long long bit = 1ull << BIT_OFFSET(i);
    const unsigned lane = __lane_id();
    unsigned long long todo = __ballot(1);
    const int first = __ffsll((long long)todo) - 1;
    const size_t w0 = ((size_t)(unsigned)__shfl((int)(w >> 32), first) << 32) |
                      (unsigned)__shfl((int)(unsigned)w, first);
    if (__ballot(w != w0 || BIT_OFFSET(i) != lane) == 0) {
      if ((int)lane == first) atomicOr((unsigned long long *)(data + w0), todo);
      return;
    }
    unsigned long long *acc = &s_or[(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y *
                                                                  threadIdx.z)) >> 6 & 15];
    for (int g = 0; g < 4 && todo; g++) {  // todo is the same in every lane
      const int lead = __ffsll((long long)todo) - 1;
      const size_t lw = ((size_t)(unsigned)__shfl((int)(w >> 32), lead) << 32) |
                        (unsigned)__shfl((int)(unsigned)w, lead);
      const bool mine = ((todo >> lane) & 1ull) && w == lw;
      const unsigned long long grp = __ballot(mine);
      if ((int)lane == lead) *acc = 0;
      __builtin_amdgcn_wave_barrier();
      if (mine) atomicOr(acc, bit);
      __builtin_amdgcn_wave_barrier();
      if ((int)lane == lead) atomicOr((unsigned long long *)(data + lw), *acc);
      __builtin_amdgcn_wave_barrier();
      todo &= ~grp;
    }
    if ((todo >> lane) & 1ull) atomicOr((unsigned long long *)(data + w), bit);
  }
};

namespace gg_facade {
inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    fprintf(stderr, "gg: %s: %s\n", what, hipGetErrorString(e));
    abort();
  }
}
template <int = 0>
__global__ void bitmap_fill_words(unsigned long *d, size_t words, size_t size) {
  size_t full = WORD_OFFSET(size);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < words;
       i += (size_t)gridDim.x * blockDim.x) {
    unsigned long x = 0;
    if (i < full)
      x = ~0ul;
    else if (i == full && BIT_OFFSET(size))
      x = (1ul << BIT_OFFSET(size)) - 1ul;
    d[i] = x;
  }
}
template <int = 0>
__global__ void bitmap_set_bit(unsigned long *d, size_t i) {
  d[WORD_OFFSET(i)] |= 1ul << BIT_OFFSET(i);
}
}  // namespace gg_facade

// Host calls are queued on the null stream, in order with the engine's
// launches; get_bit waits for them (a read-back).
class Bitmap {
 public:
  size_t size;
  unsigned long *data;  // device words, WORD_OFFSET(size)+1 of them
  Bitmap() : size(0), data(nullptr) {}
  explicit Bitmap(size_t size) : size(size), data(nullptr) {  // gg_malloc: kept when freed
    if (gg_malloc((void **)&data, sizeof(unsigned long) * (WORD_OFFSET(size) + 1)) != GG_OK) {
      fprintf(stderr, "gg: Bitmap alloc: %s\n", gg_last_error());
      abort();
    }
    clear();
  }
  ~Bitmap() {
    if (data) (void)gg_free(data);
  }
  Bitmap(const Bitmap &) = delete;
  Bitmap &operator=(const Bitmap &) = delete;
  size_t words() const { return WORD_OFFSET(size) + 1; }
  void clear() {
    gg_facade::hip_check(hipMemsetAsync(data, 0, sizeof(unsigned long) * words(), 0),
                         "Bitmap clear");
  }
  void fill() {
    hipLaunchKernelGGL(gg_facade::bitmap_fill_words<0>, dim3((words() + 255) / 256), dim3(256), 0, 0,
                       data, words(), size);
  }
  // host-side single-bit access: a read-back, or one queued store
  unsigned long get_bit(size_t i) const {
    unsigned long w = 0;
    gg_facade::hip_check(hipMemcpy(&w, data + WORD_OFFSET(i), 8, hipMemcpyDeviceToHost), "get_bit");
    return w & (1ul << BIT_OFFSET(i));
  }
  void set_bit(size_t i) {
    hipLaunchKernelGGL(gg_facade::bitmap_set_bit<0>, dim3(1), dim3(1), 0, 0, data, i);
  }
  BitmapView view() const { return BitmapView{size, data}; }
};

typedef Bitmap VertexSubset;

#endif
