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
 * Modified from Gemini toolkits/bc.cpp (shugo256/GeminiGraph).  Changes:
 * the lambdas are device callables ([=] __device__ captures) over device
 * views of the graph, bitmaps and vertex arrays instead of [&] captures of
 * host pointers; host-side pokes and reads of vertex arrays go through
 * set_value() / to_host(); main() takes the same arguments and runs the same
 * 1 + 5 timed compute() calls, with optional result dumps for tests
 * (common.hpp); code reformatted.
 */
// Betweenness dependencies on the device facade: compute() of
// toolkits/bc.cpp:24-196 (FP64; the compact variant is compiled out there too).
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "common.hpp"
#include "core/graph.hpp"

void compute(Graph<Empty> *graph, VertexId root) {
  double exec_time = 0;
  exec_time -= get_time();

  double *num_paths = graph->alloc_vertex_array<double>();
  double *dependencies = graph->alloc_vertex_array<double>();
  VertexSubset *active_all = graph->alloc_vertex_subset();
  active_all->fill();
  VertexSubset *visited = graph->alloc_vertex_subset();
  std::vector<VertexSubset *> levels;
  VertexSubset *active_in = graph->alloc_vertex_subset();

  VertexId active_vertices = 1;
  visited->clear();
  visited->set_bit(root);
  active_in->clear();
  active_in->set_bit(root);
  levels.push_back(active_in);
  graph->fill_vertex_array(num_paths, 0.0);
  graph->set_value(num_paths, root, 1.0);
  const GraphView<Empty> g = graph->view();
  const BitmapView vis = visited->view();
  VertexId i_i;
  printf("forward\n");
  for (i_i = 0; active_vertices > 0; i_i++) {
    printf("active(%d)>=%u\n", i_i, active_vertices);
    VertexSubset *active_out = graph->alloc_vertex_subset();
    active_out->clear();
    const BitmapView in = active_in->view(), out = active_out->view();
    graph->process_edges<VertexId, double>(
        [=] __device__(VertexId src) { g.emit(src, num_paths[src]); },
        [=] __device__(VertexId src, double msg, VertexAdjList<Empty> outgoing_adj) -> VertexId {
          for (AdjUnit<Empty> *ptr = outgoing_adj.begin; ptr != outgoing_adj.end; ptr++) {
            VertexId dst = ptr->neighbour;
            if (!vis.get_bit(dst)) {
              if (num_paths[dst] == 0) out.set_bit(dst);
              write_add(&num_paths[dst], msg);
            }
          }
          return 0;
        },
        [=] __device__(VertexId dst, VertexAdjList<Empty> incoming_adj) {
          if (vis.get_bit(dst)) return;
          double sum = 0;
          for (AdjUnit<Empty> *ptr = incoming_adj.begin; ptr != incoming_adj.end; ptr++) {
            VertexId src = ptr->neighbour;
            if (in.get_bit(src)) sum += num_paths[src];
          }
          if (sum > 0) g.emit(dst, sum);
        },
        [=] __device__(VertexId dst, double msg) -> VertexId {
          if (!vis.get_bit(dst)) {
            out.set_bit(dst);
            write_add(&num_paths[dst], msg);
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
    levels.push_back(active_out);
    active_in = active_out;
  }

  double *inv_num_paths = num_paths;
  graph->process_vertices<VertexId>(
      [=] __device__(VertexId vtx) -> VertexId {
        inv_num_paths[vtx] = 1 / num_paths[vtx];
        dependencies[vtx] = 0;
        return 1;
      },
      active_all);
  visited->clear();
  graph->process_vertices<VertexId>(
      [=] __device__(VertexId vtx) -> VertexId {
        vis.set_bit(vtx);
        dependencies[vtx] += inv_num_paths[vtx];
        return 1;
      },
      levels.back());
  graph->transpose();
  printf("backward\n");
  while (levels.size() > 1) {
    const BitmapView lb = levels.back()->view();
    graph->process_edges<VertexId, double>(
        [=] __device__(VertexId src) { g.emit(src, dependencies[src]); },
        [=] __device__(VertexId src, double msg, VertexAdjList<Empty> outgoing_adj) -> VertexId {
          for (AdjUnit<Empty> *ptr = outgoing_adj.begin; ptr != outgoing_adj.end; ptr++) {
            VertexId dst = ptr->neighbour;
            if (!vis.get_bit(dst)) write_add(&dependencies[dst], msg);
          }
          return 0;
        },
        [=] __device__(VertexId dst, VertexAdjList<Empty> incoming_adj) {
          if (vis.get_bit(dst)) return;
          double sum = 0;
          for (AdjUnit<Empty> *ptr = incoming_adj.begin; ptr != incoming_adj.end; ptr++) {
            VertexId src = ptr->neighbour;
            if (lb.get_bit(src)) sum += dependencies[src];
          }
          g.emit(dst, sum);
        },
        [=] __device__(VertexId dst, double msg) -> VertexId {
          if (!vis.get_bit(dst)) write_add(&dependencies[dst], msg);
          return 0;
        },
        levels.back(), visited);
    delete levels.back();
    levels.pop_back();
    graph->process_vertices<VertexId>(
        [=] __device__(VertexId vtx) -> VertexId {
          vis.set_bit(vtx);
          dependencies[vtx] += inv_num_paths[vtx];
          return 1;
        },
        levels.back());
  }

  graph->process_vertices<VertexId>(
      [=] __device__(VertexId vtx) -> VertexId {
        dependencies[vtx] = (dependencies[vtx] - inv_num_paths[vtx]) / inv_num_paths[vtx];
        return 1;
      },
      active_all);
  graph->transpose();

  exec_time += get_time();
  printf("exec_time=%lf(s)\n", exec_time);

  graph->gather_vertex_array(dependencies, 0);
  graph->gather_vertex_array(inv_num_paths, 0);
  std::vector<double> hd = graph->to_host(dependencies), hi = graph->to_host(inv_num_paths);
  for (VertexId v_i = 0; v_i < 20 && v_i < graph->vertices; v_i++)
    printf("%lf %lf\n", hd[v_i], 1 / hi[v_i]);
  dump_array("dependencies", hd);
  dump_array("inv_num_paths", hi);

  graph->dealloc_vertex_array(dependencies);
  graph->dealloc_vertex_array(inv_num_paths);
  delete levels.back();
  delete visited;
  delete active_all;
}

int main(int argc, char **argv) {
  MPI_Instance mpi(&argc, &argv);
  if (argc < 4) {
    printf("bc [file] [vertices] [root]\n");
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
