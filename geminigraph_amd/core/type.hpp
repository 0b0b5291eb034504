// Core types of the engine, host and device visible (core/type.hpp:22-68).
#ifndef GG_CORE_TYPE_HPP
#define GG_CORE_TYPE_HPP

#include <stdint.h>

struct Empty {};

typedef uint32_t VertexId;
typedef uint64_t EdgeId;

// On-disk record: packed <src, dst[, data]> (core/type.hpp:27-41).
template <typename EdgeData>
struct EdgeUnit {
  VertexId src;
  VertexId dst;
  EdgeData edge_data;
} __attribute__((packed));

template <>
struct EdgeUnit<Empty> {
  VertexId src;
  union {
    VertexId dst;
    Empty edge_data;
  };
} __attribute__((packed));

// Adjacency element handed to the toolkit callbacks (core/type.hpp:43-55).
template <typename EdgeData>
struct AdjUnit {
  VertexId neighbour;
  EdgeData edge_data;
} __attribute__((packed));

template <>
struct AdjUnit<Empty> {
  union {
    VertexId neighbour;
    Empty edge_data;
  };
} __attribute__((packed));

// begin/end view of one (sub-)list of a vertex's adjacency (core/type.hpp:62-68).
// Callbacks may receive disjoint sub-lists of one vertex: the reference hands
// one per socket (core/graph.hpp:1600-1633, :1767-1770).
template <typename EdgeData>
struct VertexAdjList {
  AdjUnit<EdgeData> *begin;
  AdjUnit<EdgeData> *end;
  __host__ __device__ VertexAdjList() : begin(nullptr), end(nullptr) {}
  __host__ __device__ VertexAdjList(AdjUnit<EdgeData> *b, AdjUnit<EdgeData> *e) : begin(b), end(e) {}
};

#endif
