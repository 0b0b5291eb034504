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
 * Modified from Gemini toolkits/bfs.cpp (shugo256/GeminiGraph).  Changes:
 * the lambdas are device callables ([=] __device__ captures) over device
 * views of the graph, bitmaps and vertex arrays instead of [&] captures of
 * host pointers; host-side pokes and reads of vertex arrays go through
 * set_value() / to_host(); main() takes the same arguments and runs the same
 * 1 + 5 timed compute() calls, with optional result dumps for tests
 * (common.hpp); code reformatted.
 */
// BFS on the device facade: the compute() of toolkits/bfs.cpp:22-109, with
// device callbacks and views in place of host pointers.
#include <stdio.h>
#include <stdlib.h>

#include "common.hpp"
#include "core/graph.hpp"

void compute(Graph<Empty> *graph, VertexId root) {
  double exec_time = 0;
  exec_time -= get_time();

  VertexId *parent = graph->alloc_vertex_array<VertexId>();
  VertexSubset *visited = graph->alloc_vertex_subset();
  VertexSubset *active_in = graph->alloc_vertex_subset();
  VertexSubset *active_out = graph->alloc_vertex_subset();

  visited->clear();
  visited->set_bit(root);
  active_in->clear();
  active_in->set_bit(root);
  graph->fill_vertex_array(parent, graph->vertices);
  graph->set_value(parent, root, root);
  const GraphView<Empty> g = graph->view();

  VertexId active_vertices = 1;
  for (int i_i = 0; active_vertices > 0; i_i++) {
    printf("active(%d)>=%u\n", i_i, active_vertices);
    active_out->clear();
    const BitmapView vis = visited->view(), in = active_in->view(), out = active_out->view();
    active_vertices = graph->process_edges<VertexId, VertexId>(
        [=] __device__(VertexId src) { g.emit(src, src); },
        [=] __device__(VertexId src, VertexId msg, VertexAdjList<Empty> outgoing_adj) {
          VertexId activated = 0;
          for (AdjUnit<Empty> *ptr = outgoing_adj.begin; ptr != outgoing_adj.end; ptr++) {
            VertexId dst = ptr->neighbour;
            if (parent[dst] == g.vertices && cas(&parent[dst], g.vertices, src)) {
              out.set_bit(dst);
              activated += 1;
            }
          }
          return activated;
        },
        [=] __device__(VertexId dst, VertexAdjList<Empty> incoming_adj) {
          if (vis.get_bit(dst)) return;
          for (AdjUnit<Empty> *ptr = incoming_adj.begin; ptr != incoming_adj.end; ptr++) {
            VertexId src = ptr->neighbour;
            if (in.get_bit(src)) {
              g.emit(dst, src);
              break;
            }
          }
        },
        [=] __device__(VertexId dst, VertexId msg) -> VertexId {
          if (cas(&parent[dst], g.vertices, msg)) {
            out.set_bit(dst);
            return 1;
          }
          return 0;
        },
        active_in, visited);
    active_vertices = graph->process_vertices<VertexId>(
        [=] __device__(VertexId vtx) -> VertexId {
          vis.set_bit(vtx);
          return 1;
        },
        active_out);
    std::swap(active_in, active_out);
  }

  exec_time += get_time();
  printf("exec_time=%lf(s)\n", exec_time);

  graph->gather_vertex_array(parent, 0);
  std::vector<VertexId> hp = graph->to_host(parent);
  VertexId found_vertices = 0;
  for (VertexId v_i = 0; v_i < graph->vertices; v_i++)
    if (hp[v_i] < graph->vertices) found_vertices += 1;
  printf("found_vertices = %u\n", found_vertices);
  dump_array("parent", hp);

  graph->dealloc_vertex_array(parent);
  delete active_in;
  delete active_out;
  delete visited;
}

int main(int argc, char **argv) {
  MPI_Instance mpi(&argc, &argv);
  if (argc < 4) {
    printf("bfs [file] [vertices] [root]\n");
    exit(-1);
  }
  Graph<Empty> *graph = new Graph<Empty>();
  VertexId root = std::atoi(argv[3]);
  graph->load_directed(argv[1], std::atoi(argv[2]));
  compute(graph, root);
  for (int run = 0; run < extra_runs(); run++) compute(graph, root);
  delete graph;
  return 0;
}
