// Host-side mesh preparation for libtdt (see mesh_prep.hpp).
#include "mesh_prep.hpp"

#include <algorithm>
#include <cstdio>
#include <numeric>
#include <unordered_map>

namespace tdt {

namespace {

// 4 x (256 CUs x 4 SIMDs x 64 lanes): from here on a kernel runs several
// rounds of wavefronts per SIMD and throughput, not the slowest wavefront,
// sets its time
constexpr int kBoundaryLastMinCells = 4 * 65536;

std::string fmt(const char *f, long a = 0, long b = 0, long c = 0) {
    char buf[256];
    std::snprintf(buf, sizeof buf, f, a, b, c);
    return buf;
}

// Reverse Cuthill-McKee over the interior cell graph.  nc_old: [4][n] 0-based
// neighbour indices, -1 for a boundary face.
std::vector<int> rcm_order(int n, const std::vector<int> &nc_old) {
    std::vector<int> deg(n, 0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 4; ++k) deg[i] += nc_old[(size_t)k * n + i] >= 0;
    std::vector<char> seen(n, 0);
    std::vector<int> order;
    order.reserve(n);
    std::vector<int> level, next;
    auto bfs_last = [&](int s, std::vector<int> &mark, int stamp) {
        // returns the last BFS level from s
        std::vector<int> cur{s}, last{s};
        mark[s] = stamp;
        while (!cur.empty()) {
            last = cur;
            std::vector<int> nx;
            for (int v : cur)
                for (int k = 0; k < 4; ++k) {
                    int w = nc_old[(size_t)k * n + v];
                    if (w >= 0 && mark[w] != stamp) { mark[w] = stamp; nx.push_back(w); }
                }
            cur.swap(nx);
        }
        return last;
    };
    std::vector<int> mark(n, -1);
    int stamp = 0;
    for (int s0 = 0; s0 < n; ++s0) {
        if (seen[s0]) continue;
        // pseudo-peripheral start: minimum degree node of the deepest level, twice
        int s = s0;
        for (int rep = 0; rep < 2; ++rep) {
            std::vector<int> last = bfs_last(s, mark, stamp++);
            int best = last[0];
            for (int v : last)
                if (deg[v] < deg[best] || (deg[v] == deg[best] && v < best)) best = v;
            s = best;
        }
        size_t head = order.size();
        order.push_back(s);
        seen[s] = 1;
        while (head < order.size()) {
            int v = order[head++];
            int nb[4], cnt = 0;
            for (int k = 0; k < 4; ++k) {
                int w = nc_old[(size_t)k * n + v];
                if (w >= 0 && !seen[w]) { seen[w] = 1; nb[cnt++] = w; }
            }
            std::sort(nb, nb + cnt, [&](int a, int b) {
                return deg[a] != deg[b] ? deg[a] < deg[b] : a < b;
            });
            for (int q = 0; q < cnt; ++q) order.push_back(nb[q]);
        }
    }
    std::reverse(order.begin(), order.end());
    return order;  // order[new] = old
}

}  // namespace

int order_bandwidth(int n, const std::vector<int> &nc_old, const std::vector<int> &old2new) {
    int bw = 0;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 4; ++k) {
            int j = nc_old[(size_t)k * n + i];
            if (j >= 0) bw = std::max(bw, std::abs(old2new[i] - old2new[j]));
        }
    return bw;
}

std::string prepare_mesh(const tdt_mesh &ms, int ordering, PreparedMesh &pm, int *code,
                         const std::vector<int> *order) {
    *code = TDT_ERR_MESH;
    const int n = ms.neles, m = ms.nods, g = ms.nghosts;
    if (n <= 0 || m <= 0 || g < 0) return fmt("bad sizes nods=%ld neles=%ld nghosts=%ld", m, n, g);
    if (!ms.x_mm || !ms.y_mm || !ms.nod || !ms.nc || (g > 0 && (!ms.parent || !ms.ghost ||
                                                                 !ms.type || !ms.side))) {
        *code = TDT_ERR_ARG;
        return "null mesh array";
    }
    pm.n = n; pm.g = g; pm.m = m;
    // ---- ghosts (bc.in order) ----
    std::unordered_map<int, int> ghost_index;  // original ghost id -> bc.in position
    pm.nmax = n;
    for (int k = 0; k < g; ++k) {
        int gid = ms.ghost[k], par = ms.parent[k], ty = ms.type[k], sd = ms.side[k];
        if (gid <= n) return fmt("bc.in record %ld: ghost id %ld collides with a cell id", k + 1, gid);
        if (par < 1 || par > n) return fmt("bc.in record %ld: parent %ld out of range", k + 1, par);
        if (sd < 1 || sd > 4) return fmt("bc.in record %ld: side %ld out of range", k + 1, sd);
        if (ty < 2 || ty > 7) return fmt("bc.in record %ld: unknown boundary type %ld", k + 1, ty);
        if (!ghost_index.emplace(gid, k).second)
            return fmt("bc.in record %ld: ghost id %ld listed twice", k + 1, gid);
        pm.nmax = std::max(pm.nmax, gid);
    }
    // ---- connectivity, resolving -1 through bc.in (parent, side) ----
    std::vector<int> nc_old((size_t)4 * n);      // interior: 0-based, ghost: -(2+bcpos)
    std::vector<int> face_ghost((size_t)4 * n, -1);
    for (int k = 0; k < g; ++k) {
        int slot = (ms.side[k] - 1) * n + (ms.parent[k] - 1);
        if (face_ghost[slot] >= 0)
            return fmt("bc.in records %ld and %ld name the same face", face_ghost[slot] + 1, k + 1);
        face_ghost[slot] = k;
    }
    for (int f = 0; f < 4; ++f)
        for (int i = 0; i < n; ++i) {
            size_t s = (size_t)f * n + i;
            int v = ms.nc[s];
            int nd = ms.nod[s];
            if (nd < 1 || nd > m) return fmt("cell %ld: node id %ld out of range", i + 1, nd);
            if (v >= 1 && v <= n) {
                nc_old[s] = v - 1;
                continue;
            }
            int k;
            if (v == -1) {
                k = face_ghost[s];
                if (k < 0) return fmt("cell %ld face %ld: no neighbour and no bc.in record", i + 1, f + 1);
            } else {
                auto it = ghost_index.find(v);
                if (it == ghost_index.end())
                    return fmt("cell %ld face %ld: neighbour %ld is neither a cell nor a ghost",
                               i + 1, f + 1, v);
                k = it->second;
            }
            if (ms.parent[k] != i + 1 || ms.side[k] != f + 1)
                return fmt("cell %ld face %ld: ghost %ld belongs to another face", i + 1, f + 1,
                           ms.ghost[k]);
            nc_old[s] = -(2 + k);
        }
    for (int k = 0; k < g; ++k)
        if (face_ghost[(ms.side[k] - 1) * n + ms.parent[k] - 1] != k ||
            nc_old[(size_t)(ms.side[k] - 1) * n + ms.parent[k] - 1] != -(2 + k))
            return fmt("bc.in record %ld: ghost %ld is not the neighbour across its face", k + 1,
                       ms.ghost[k]);
    // ---- cell order ----
    std::vector<int> nc_int((size_t)4 * n);
    for (size_t s = 0; s < nc_int.size(); ++s) nc_int[s] = nc_old[s] >= 0 ? nc_old[s] : -1;
    std::vector<int> ident(n);
    std::iota(ident.begin(), ident.end(), 0);
    int bw_in = order_bandwidth(n, nc_int, ident);
    std::vector<int> new2old = ident;
    pm.ordering_used = TDT_ORDER_INPUT;
    pm.bandwidth = bw_in;
    const bool boundary_last = !order && (ordering == TDT_ORDER_BOUNDARY_LAST ||
                               (ordering == TDT_ORDER_AUTO && n >= kBoundaryLastMinCells));
    if (order) {
        std::vector<char> seen(n, 0);
        bool perm = (int)order->size() == n;
        for (int q = 0; perm && q < n; ++q) {
            const int o = (*order)[q];
            perm = o >= 0 && o < n && !seen[o];
            if (perm) seen[o] = 1;
        }
        if (!perm) {
            *code = TDT_ERR_ARG;
            return "cell order is not a permutation of the cells";
        }
        new2old = *order;
        std::vector<int> o2n(n);
        for (int q = 0; q < n; ++q) o2n[new2old[q]] = q;
        pm.bandwidth = order_bandwidth(n, nc_int, o2n);
    } else if (ordering == TDT_ORDER_RCM || ordering == TDT_ORDER_AUTO ||
               ordering == TDT_ORDER_BOUNDARY_LAST) {
        std::vector<int> rcm = rcm_order(n, nc_int), o2n(n);
        for (int q = 0; q < n; ++q) o2n[rcm[q]] = q;
        int bw_rcm = order_bandwidth(n, nc_int, o2n);
        if (ordering == TDT_ORDER_RCM || bw_rcm < bw_in) {
            new2old = rcm;
            pm.ordering_used = TDT_ORDER_RCM;
            pm.bandwidth = bw_rcm;
        }
    } else if (ordering != TDT_ORDER_INPUT) {
        *code = TDT_ERR_ARG;
        return fmt("unknown ordering %ld", ordering);
    }
    if (boundary_last) {
        // Cells with a ghost neighbour go last, grouped by which boundary type
        // sits on which face (stable within a group).  A wavefront executes the
        // boundary-condition code of every (face, type) that any of its lanes
        // has: interior wavefronts then skip it all, and a boundary wavefront
        // runs one such branch instead of a mix.  The price is scattered
        // neighbour gathers in the boundary wavefronts, which sets the kernel
        // time on grids too small to fill the GPU (all wavefronts then run at
        // once and the slowest decides), so small grids keep the plain order.
        auto signature = [&](int c) {
            int sig = 0;
            for (int f = 0; f < 4; ++f) {
                int v = nc_old[(size_t)f * n + c];
                sig = sig * 8 + (v < 0 ? ms.type[-v - 2] : 0);
            }
            return sig;
        };
        std::vector<int> sig(n);
        for (int c = 0; c < n; ++c) sig[c] = signature(c);
        std::stable_sort(new2old.begin(), new2old.end(),
                         [&](int a, int b) { return sig[a] < sig[b]; });
        std::vector<int> o2n(n);
        for (int q = 0; q < n; ++q) o2n[new2old[q]] = q;
        pm.bandwidth = order_bandwidth(n, nc_int, o2n);
        pm.ordering_used = TDT_ORDER_BOUNDARY_LAST;
    }
    pm.cell_new2old = new2old;
    pm.cell_old2new.assign(n, 0);
    for (int q = 0; q < n; ++q) pm.cell_old2new[new2old[q]] = q;
    // ---- node order: first touch by cells in the new order ----
    std::vector<int> node_o2n(m, -1);
    pm.node_new2old.clear();
    pm.node_new2old.reserve(m);
    for (int q = 0; q < n; ++q)
        for (int f = 0; f < 4; ++f) {
            int nd = ms.nod[(size_t)f * n + new2old[q]] - 1;
            if (node_o2n[nd] < 0) { node_o2n[nd] = (int)pm.node_new2old.size(); pm.node_new2old.push_back(nd); }
        }
    for (int nd = 0; nd < m; ++nd)
        if (node_o2n[nd] < 0) { node_o2n[nd] = (int)pm.node_new2old.size(); pm.node_new2old.push_back(nd); }
    pm.xn.resize(m);
    pm.yn.resize(m);
    for (int q = 0; q < m; ++q) {
        // X(N) = X(N) * MM_TO_M  (geometry.f90:35-36)
        pm.xn[q] = ms.x_mm[pm.node_new2old[q]] * 0.001;
        pm.yn[q] = ms.y_mm[pm.node_new2old[q]] * 0.001;
    }
    // ---- per-cell tables in the new order ----
    pm.nod4.resize((size_t)4 * n);
    pm.nc4.resize((size_t)4 * n);
    for (int q = 0; q < n; ++q) {
        int o = new2old[q];
        for (int f = 0; f < 4; ++f) {
            pm.nod4[(size_t)4 * q + f] = node_o2n[ms.nod[(size_t)f * n + o] - 1];
            int v = nc_old[(size_t)f * n + o];
            pm.nc4[(size_t)4 * q + f] = v >= 0 ? pm.cell_old2new[v] : n + (-v - 2);
        }
    }
    pm.gparent.resize(g); pm.gside.resize(g); pm.gtype.resize(g); pm.gopp.resize(g);
    pm.ghost_orig_id.resize(g);
    for (int k = 0; k < g; ++k) {
        int par = pm.cell_old2new[ms.parent[k] - 1];
        pm.gparent[k] = par;
        pm.gside[k] = ms.side[k];
        pm.gtype[k] = ms.type[k];
        pm.ghost_orig_id[k] = ms.ghost[k];
        // opposite face: side 1 -> NC3, 2 -> NC4, 3 -> NC1, 4 -> NC2 (boundary.f90:78-82)
        int opp = pm.nc4[(size_t)4 * par + ((ms.side[k] + 1) & 3)];
        pm.gopp[k] = opp;
        if ((ms.type[k] == 3 || ms.type[k] == 6) && opp >= n) {
            *code = TDT_ERR_UNSUPPORTED;
            return fmt("bc.in record %ld: type-%ld ghost whose opposite neighbour is also a ghost",
                       k + 1, ms.type[k]);
        }
    }
    // ---- BC type 6 (boundary.f90:169-305): the column of cells from the parent
    // across to the far boundary (the reference's walk ends at the first ghost),
    // each with the DELY expression the reference uses for it: the parent with
    // its own side's, the first cell with the walking side's, the rest with the
    // opposite one's
    auto dely = [&](int c, int side) {
        const double y1 = pm.yn[pm.nod4[(size_t)4 * c]], y2 = pm.yn[pm.nod4[(size_t)4 * c + 1]];
        const double y3 = pm.yn[pm.nod4[(size_t)4 * c + 2]], y4 = pm.yn[pm.nod4[(size_t)4 * c + 3]];
        switch (side) {
        case 1: return 0.5 * ((y1 + y2) - (y3 + y4));
        case 2: return 0.5 * ((y2 + y3) - (y1 + y4));
        case 3: return 0.5 * ((y3 + y4) - (y1 + y2));
        default: return 0.5 * ((y4 + y1) - (y2 + y3));
        }
    };
    pm.g6ptr.assign(g + 1, 0);
    pm.g6cell.clear();
    pm.g6dely.clear();
    for (int k = 0; k < g; ++k) {
        if (ms.type[k] == 6) {
            const int par = pm.gparent[k], ns = ms.side[k];
            const int side = (ns + 1) % 4 + 1;          // 1 -> 3, 2 -> 4, 3 -> 1, 4 -> 2
            const int lside = (side + 1) % 4 + 1;
            pm.g6cell.push_back(par);
            pm.g6dely.push_back(dely(par, ns));
            int cur = pm.gopp[k];                       // NC(side)(parent), a cell
            pm.g6cell.push_back(cur);
            pm.g6dely.push_back(dely(cur, side));
            cur = pm.nc4[(size_t)4 * cur + side - 1];
            for (int steps = 0; cur < n; ++steps) {
                if (steps >= n) {
                    *code = TDT_ERR_UNSUPPORTED;
                    return fmt("bc.in record %ld: the type-6 column walk reaches no boundary",
                               k + 1);
                }
                pm.g6cell.push_back(cur);
                pm.g6dely.push_back(dely(cur, lside));
                cur = pm.nc4[(size_t)4 * cur + side - 1];
            }
        }
        pm.g6ptr[k + 1] = (int)pm.g6cell.size();
    }
    // ---- node gather lists in the reference accumulation order ----
    std::vector<std::vector<int>> lists(m);
    std::vector<int> sU(m, 0), sV(m, 0), sR(m, 0), rf(m, 0);
    for (int o = 0; o < n; ++o)                       // viscous.f90:79-92
        for (int f = 0; f < 4; ++f)
            lists[node_o2n[ms.nod[(size_t)f * n + o] - 1]].push_back(pm.cell_old2new[o]);
    for (int k = 0; k < g; ++k) {                     // viscous.f90:100-123
        int par_old = ms.parent[k] - 1, sd = ms.side[k];
        int a = node_o2n[ms.nod[(size_t)(sd - 1) * n + par_old] - 1];
        int b = node_o2n[ms.nod[(size_t)(sd & 3) * n + par_old] - 1];
        lists[a].push_back(n + k);
        lists[b].push_back(n + k);
        int ty = ms.type[k];
        for (int v : {a, b}) {
            int after = (int)lists[v].size();   // sums restart after this contribution
            if (ty == 4) sV[v] = after;
            if (ty == 5) { sU[v] = after; sV[v] = after; sR[v] = after; rf[v] = 1; }
        }
    }
    pm.nptr.assign(m + 1, 0);
    for (int v = 0; v < m; ++v) pm.nptr[v + 1] = pm.nptr[v] + (int)lists[v].size();
    pm.nlist.resize(pm.nptr[m]);
    pm.novr.resize((size_t)4 * m);
    for (int v = 0; v < m; ++v) {
        std::copy(lists[v].begin(), lists[v].end(), pm.nlist.begin() + pm.nptr[v]);
        pm.novr[(size_t)4 * v + 0] = pm.nptr[v] + sU[v];
        pm.novr[(size_t)4 * v + 1] = pm.nptr[v] + sV[v];
        pm.novr[(size_t)4 * v + 2] = pm.nptr[v] + sR[v];
        pm.novr[(size_t)4 * v + 3] = rf[v];
    }
    *code = TDT_OK;
    return "";
}

int wave_node_tables(const std::vector<int> &nod4, int n, int nwaves, std::vector<int> &bnode,
                     std::vector<int> &nloc) {
    constexpr int kWaveCells = 64;
    auto wave_nodes = [&](int w, std::vector<int> &v) {
        v.clear();
        for (int i = w * kWaveCells; i < std::min(n, (w + 1) * kWaveCells); ++i)
            for (int q = 0; q < 4; ++q) v.push_back(nod4[(size_t)4 * i + q]);
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        if (v.empty()) v.push_back(0);   // a wave past the last cell: any node
    };
    std::vector<int> v;
    size_t cap = 0;
    for (int w = 0; w < nwaves; ++w) {
        wave_nodes(w, v);
        cap = std::max(cap, v.size());
    }
    cap = (cap + kWaveCells - 1) / kWaveCells * kWaveCells;   // whole 64-lane copy rounds
    bnode.assign((size_t)nwaves * cap, 0);
    nloc.assign((size_t)2 * n, 0);
    for (int w = 0; w < nwaves; ++w) {
        wave_nodes(w, v);
        std::copy(v.begin(), v.end(), bnode.begin() + (size_t)w * cap);
        std::fill(bnode.begin() + (size_t)w * cap + v.size(), bnode.begin() + (size_t)(w + 1) * cap,
                  v.back());
        for (int i = w * kWaveCells; i < std::min(n, (w + 1) * kWaveCells); ++i) {
            int sl[4];
            for (int q = 0; q < 4; ++q)
                sl[q] = (int)(std::lower_bound(v.begin(), v.end(), nod4[(size_t)4 * i + q]) -
                              v.begin());
            nloc[(size_t)2 * i] = sl[0] | sl[1] << 16;
            nloc[(size_t)2 * i + 1] = sl[2] | sl[3] << 16;
        }
    }
    return (int)cap;
}

}  // namespace tdt
