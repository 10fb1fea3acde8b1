// Host-side mesh preparation for libtdt: validation of grid.in / bc.in data,
// bandwidth-reducing renumbering, ghost tables and the node-gather lists that
// replace the reference's cell-to-node scatter (viscous.f90:77-124).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "tdt.h"

namespace tdt {

struct PreparedMesh {
    int n = 0;      // interior cells
    int g = 0;      // ghosts (internal ids n .. n+g-1, in bc.in order)
    int m = 0;      // nodes
    int nmax = 0;   // largest original cell/ghost id
    int bandwidth = 0;
    int ordering_used = TDT_ORDER_INPUT;

    std::vector<int> cell_new2old;   // [n]  original 0-based cell index of new cell
    std::vector<int> cell_old2new;   // [n]
    std::vector<int> node_new2old;   // [m]
    std::vector<int> ghost_orig_id;  // [g]  1-based ghost id from bc.in

    std::vector<double> xn, yn;      // [m] node coordinates in metres, new order
    std::vector<int> nod4;           // [n][4] new node ids
    std::vector<int> nc4;            // [n][4] new neighbour ids (ghost k -> n + k)
    std::vector<int> gparent, gside, gtype, gopp;  // [g] new parent, side 1..4, type,
                                                   // opposite neighbour of the parent

    // Node gather: contributors of node v are nlist[nptr[v] .. nptr[v+1]), in the
    // reference's accumulation order (cells by original id, then ghosts in bc.in
    // order).  novr[v] = {sU, sV, sR, rflag}: U (V, nu~) sums start at position
    // sU (sV, sR) because an override at an earlier ghost reset them; rflag = 1
    // when the nu~ sum restarts from RTOL*RTRED (type 5) instead of 0.
    std::vector<int> nptr, nlist;
    std::vector<int> novr;           // [m][4]

    // BC type 6: for ghost k, entries g6ptr[k] .. g6ptr[k+1] list the cells whose
    // mass flow C1*U*DELY (*YCC) sums to the port flow, in the reference's order,
    // with their DELY (empty for other types)
    std::vector<int> g6ptr, g6cell;
    std::vector<double> g6dely;
};

// Returns "" on success, otherwise an error message; *code receives the
// TDT_ERR_* class of the failure.  order (optional): the cell order to use,
// new -> old 0-based, in place of the one `ordering` would choose.
std::string prepare_mesh(const tdt_mesh &mesh, int ordering, PreparedMesh &out, int *code,
                         const std::vector<int> *order = nullptr);

// Bandwidth of a cell order: max |pos(i) - pos(j)| over interior neighbour pairs.
int order_bandwidth(int n, const std::vector<int> &nc_old, const std::vector<int> &old2new);

// Node lists of the wavefronts (each 64 consecutive cells of the prepared order)
// for kernels that copy their cells' node records into LDS: bnode[w * cap + k],
// k < cap, = the distinct nodes of wave w's cells in ascending order, padded
// with the last (a wave past the last cell lists node 0); nloc[2i], nloc[2i+1] =
// cell i's four nodes' slots in its wave's list, 16 bits each ({slot0 | slot1 <<
// 16, slot2 | slot3 << 16}).  nwaves >= the waves the n cells fill.  Returns cap,
// the longest list rounded up to a multiple of 64 (at most 256).
int wave_node_tables(const std::vector<int> &nod4, int n, int nwaves, std::vector<int> &bnode,
                     std::vector<int> &nloc);

}  // namespace tdt
