// GPU halves of the model set-up (setup.hip), called from the host set-up (model.cpp) and from
// the node-cloud batch (family_batch.hip).  Plain C++ declarations: no HIP types.
#pragma once
#include <string>
#include <vector>

namespace pdx {

struct Model;

// cell grid of a family search: cell (i, j, k) spans lo + [i, i+1) cs ... and has index
// i + nb[0] (j + nb[1] k).  clamp = false: a query's cell is int((p - lo) / cs) per axis (the
// model set-up, setBlock pdsolve.cpp:441-502); clamp = true: floor((p - lo) / cs) clamped to the
// grid on the first dim axes and 0 beyond (the node-cloud batch).
struct CellGrid {
    double lo[3];
    double cs;
    int nb[3];
    int dim;
    bool clamp;
};

// Families of nq query nodes (qnode[q], or node q when qnode is null) among nn nodes x [nn*3]
// with horizons delta [nn], from the cell lists (cell_off [ncell+1], cell_node [nn], each cell's
// nodes in node order).  Writes the offsets off [nq+1], the members into *mem when mem is not
// null, the largest family into max_count, and hands the device copy of the members to
// *dmem_out (the caller frees it) when dmem_out is not null.
int dev_cell_search(int device, int nn, const double *x, const double *delta, const std::vector<int> &cell_off,
                    const std::vector<int> &cell_node, const CellGrid &g, long long nq, const int *qnode,
                    double tiny, std::vector<long long> &off, std::vector<int> *mem, int &max_count,
                    int **dmem_out, std::string &err);

// Node graph of the model (families and FE elements set up; ne_off / ne_item: the FE elements
// of each node in element order): adj_off [nn+1], adj (sorted unique neighbours per node).
int dev_node_graph(int device, const Model &m, const std::vector<int> &ne_off, const std::vector<int> &ne_item,
                   std::vector<int> &adj_off, std::vector<int> &adj, std::string &err);

// device of the model set-up's family search and node graph: -1 (the default) the host
int setup_device();
void set_setup_device(int device);

}  // namespace pdx
