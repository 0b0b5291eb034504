/*
Copyright (c) 2014-2015 Xiaowei Zhu, Tsinghua University

   Licensed under the Apache License, Version 2.0 (the "License");
   you may not use this file except in compliance with the License.
   You may obtain a copy of the License at

       http://www.apache.org/licenses/LICENSE-2.0

   Unless required by applicable law or agreed to in writing, software
   distributed under the License is distributed on an "AS IS" BASIS,
   WITHOUT WARRANTIES OR CONDITIONS OF ANY KIND, either express or implied.
   See the License for the specific language governing permissions and
   limitations under the License.
*/
/*
 * Modified from Gemini toolkits/pagerank.cpp (shugo256/GeminiGraph).  Changes:
 * the lambdas are device callables ([=] __device__ captures) over device
 * views of the graph, bitmaps and vertex arrays instead of [&] captures of
 * host pointers; host-side pokes and reads of vertex arrays go through
 * set_value() / to_host(); main() takes the same arguments and runs the same
 * 1 + 5 timed compute() calls, with optional result dumps for tests
 * (common.hpp); code reformatted.
 */
// PageRank on the device facade: toolkits/pagerank.cpp:24-129 (FP64).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "common.hpp"
#include "core/graph.hpp"

const double d = (double)0.85;

void compute(Graph<Empty> *graph, int iterations) {
  double exec_time = 0;
  exec_time -= get_time();

  double *curr = graph->alloc_vertex_array<double>();
  double *next = graph->alloc_vertex_array<double>();
  VertexSubset *active = graph->alloc_vertex_subset();
  active->fill();
  const VertexId *out_degree = graph->out_degree;
  const GraphView<Empty> g = graph->view();

  double delta = graph->process_vertices<double>(
      [=] __device__(VertexId vtx) {
        curr[vtx] = (double)1;
        if (out_degree[vtx] > 0) curr[vtx] /= out_degree[vtx];
        return (double)1;
      },
      active);
  delta /= graph->vertices;

  for (int i_i = 0; i_i < iterations; i_i++) {
    printf("delta(%d)=%lf\n", i_i, delta);
    graph->fill_vertex_array(next, (double)0);
    const double *c = curr;
    double *n = next;
    graph->process_edges<int, double>(
        [=] __device__(VertexId src) { g.emit(src, c[src]); },
        [=] __device__(VertexId src, double msg, VertexAdjList<Empty> outgoing_adj) {
          for (AdjUnit<Empty> *ptr = outgoing_adj.begin; ptr != outgoing_adj.end; ptr++)
            write_add(&n[ptr->neighbour], msg);
          return 0;
        },
        [=] __device__(VertexId dst, VertexAdjList<Empty> incoming_adj) {
          double sum = 0;
          for (AdjUnit<Empty> *ptr = incoming_adj.begin; ptr != incoming_adj.end; ptr++)
            sum += c[ptr->neighbour];
          g.emit(dst, sum);
        },
        [=] __device__(VertexId dst, double msg) {
          write_add(&n[dst], msg);
          return 0;
        },
        active);
    if (i_i == iterations - 1) {
      delta = graph->process_vertices<double>(
          [=] __device__(VertexId vtx) {
            n[vtx] = 1 - d + d * n[vtx];
            return 0.0;
          },
          active);
    } else {
      delta = graph->process_vertices<double>(
          [=] __device__(VertexId vtx) {
            n[vtx] = 1 - d + d * n[vtx];
            if (out_degree[vtx] > 0) {
              n[vtx] /= out_degree[vtx];
              return fabs(n[vtx] - c[vtx]) * out_degree[vtx];
            }
            return fabs(n[vtx] - c[vtx]);
          },
          active);
    }
    delta /= graph->vertices;
    std::swap(curr, next);
  }

  exec_time += get_time();
  printf("exec_time=%lf(s)\n", exec_time);

  const double *c = curr;
  double pr_sum = graph->process_vertices<double>([=] __device__(VertexId vtx) { return c[vtx]; }, active);
  printf("pr_sum=%lf\n", pr_sum);

  graph->gather_vertex_array(curr, 0);
  std::vector<double> hc = graph->to_host(curr);
  VertexId max_v_i = 0;
  for (VertexId v_i = 0; v_i < graph->vertices; v_i++)
    if (hc[v_i] > hc[max_v_i]) max_v_i = v_i;
  printf("pr[%u]=%lf\n", max_v_i, hc[max_v_i]);
  dump_array("pagerank", hc);

  graph->dealloc_vertex_array(curr);
  graph->dealloc_vertex_array(next);
  delete active;
}

int main(int argc, char **argv) {
  MPI_Instance mpi(&argc, &argv);
  if (argc < 4) {
    printf("pagerank [file] [vertices] [iterations]\n");
    exit(-1);
  }
  Graph<Empty> *graph = new Graph<Empty>();
  graph->load_directed(argv[1], std::atoi(argv[2]));
  int iterations = std::atoi(argv[3]);
  compute(graph, iterations);
  for (int run = 0; run < extra_runs(); run++) compute(graph, iterations);
  delete graph;
  return 0;
}
