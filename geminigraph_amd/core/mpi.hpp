// Single-process stand-in for the reference's MPI_Instance (core/mpi.hpp:49-85):
// one rank, one GPU.  Construction selects the device (GG_DEVICE env var,
// default 0); there is no message passing in a one-partition build.
#ifndef GG_CORE_MPI_HPP
#define GG_CORE_MPI_HPP
#include <stdio.h>
#include <stdlib.h>

#include "gg.h"

class MPI_Instance {
 public:
  int partition_id = 0;
  int partitions = 1;
  MPI_Instance(int *argc, char ***argv) {
    (void)argc;
    (void)argv;
    const char *d = getenv("GG_DEVICE");
    if (gg_init(d ? atoi(d) : 0) != GG_OK) {
      fprintf(stderr, "gg: %s\n", gg_last_error());
      exit(-1);
    }
  }
  ~MPI_Instance() { gg_synchronize(); }
};
#endif
