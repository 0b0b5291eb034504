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
 * Modified from Gemini toolkits/sssp.cpp (shugo256/GeminiGraph).  Changes:
 * the lambdas are device callables ([=] __device__ captures) over device
 * views of the graph, bitmaps and vertex arrays instead of [&] captures of
 * host pointers; host-side pokes and reads of vertex arrays go through
 * set_value() / to_host(); main() takes the same arguments and runs the same
 * 1 + 5 timed compute() calls, with optional result dumps for tests
 * (common.hpp); code reformatted.
 */
// SSSP on the device facade: toolkits/sssp.cpp:22-105 (f32 weights, sentinel 1e9).
#include <stdio.h>
#include <stdlib.h>

#include "common.hpp"
#include "core/graph.hpp"

typedef float Weight;

void compute(Graph<Weight> *graph, VertexId root) {
  double exec_time = 0;
  exec_time -= get_time();

  Weight *distance = graph->alloc_vertex_array<Weight>();
  VertexSubset *active_in = graph->alloc_vertex_subset();
  VertexSubset *active_out = graph->alloc_vertex_subset();
  active_in->clear();
  active_in->set_bit(root);
  graph->fill_vertex_array(distance, (Weight)1e9);
  graph->set_value(distance, root, (Weight)0);
  const GraphView<Weight> g = graph->view();
  VertexId active_vertices = 1;

  for (int i_i = 0; active_vertices > 0; i_i++) {
    printf("active(%d)>=%u\n", i_i, active_vertices);
    active_out->clear();
    const BitmapView out = active_out->view();
    active_vertices = graph->process_edges<VertexId, Weight>(
        [=] __device__(VertexId src) { g.emit(src, distance[src]); },
        [=] __device__(VertexId src, Weight msg, VertexAdjList<Weight> outgoing_adj) {
          VertexId activated = 0;
          for (AdjUnit<Weight> *ptr = outgoing_adj.begin; ptr != outgoing_adj.end; ptr++) {
            VertexId dst = ptr->neighbour;
            Weight relax_dist = msg + ptr->edge_data;
            if (relax_dist < distance[dst]) {
              if (write_min(&distance[dst], relax_dist)) {
                out.set_bit(dst);
                activated += 1;
              }
            }
          }
          return activated;
        },
        [=] __device__(VertexId dst, VertexAdjList<Weight> incoming_adj) {
          Weight msg = 1e9;
          for (AdjUnit<Weight> *ptr = incoming_adj.begin; ptr != incoming_adj.end; ptr++) {
            VertexId src = ptr->neighbour;
            Weight relax_dist = distance[src] + ptr->edge_data;
            if (relax_dist < msg) msg = relax_dist;
          }
          if (msg < 1e9) g.emit(dst, msg);
        },
        [=] __device__(VertexId dst, Weight msg) -> VertexId {
          if (msg < distance[dst]) {
            write_min(&distance[dst], msg);
            out.set_bit(dst);
            return 1;
          }
          return 0;
        },
        active_in);
    std::swap(active_in, active_out);
  }

  exec_time += get_time();
  printf("exec_time=%lf(s)\n", exec_time);

  graph->gather_vertex_array(distance, 0);
  std::vector<Weight> hd = graph->to_host(distance);
  VertexId max_v_i = root;
  for (VertexId v_i = 0; v_i < graph->vertices; v_i++)
    if (hd[v_i] < 1e9 && hd[v_i] > hd[max_v_i]) max_v_i = v_i;
  printf("distance[%u]=%f\n", max_v_i, hd[max_v_i]);
  dump_array("distance", hd);

  graph->dealloc_vertex_array(distance);
  delete active_in;
  delete active_out;
}

int main(int argc, char **argv) {
  MPI_Instance mpi(&argc, &argv);
  if (argc < 4) {
    printf("sssp [file] [vertices] [root]\n");
    exit(-1);
  }
  Graph<Weight> *graph = new Graph<Weight>();
  graph->load_directed(argv[1], std::atoi(argv[2]));
  VertexId root = std::atoi(argv[3]);
  compute(graph, root);
  for (int run = 0; run < extra_runs(); run++) compute(graph, root);
  delete graph;
  return 0;
}
