// Host-only driver of the mesh preparation (csrc/mesh_prep.cpp), built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_mesh_prep_sanitized.py.
//
// Input file (little-endian): int32 count of meshes, then per mesh int32 nods,
// neles, nghosts, double x_mm[nods], y_mm[nods], int32 nod[4][neles],
// nc[4][neles], parent[nghosts], ghost[nghosts], type[nghosts], side[nghosts].
// Output, one line per mesh and ordering: "mesh <k> order <o> code <c> n <n> g
// <g> m <m> bw <bandwidth> used <ordering> sum <checksum of the tables> nlcap
// <wave node-list length> nlok <1 if the wave node tables hold every cell's
// nodes at its slots, in sorted, padded lists>".
#include <cstdint>
#include <cstdio>
#include <vector>

#include "mesh_prep.hpp"

namespace {

template <class T>
bool read_n(FILE *f, std::vector<T> &v, size_t n) {
    v.resize(n);
    return n == 0 || fread(v.data(), sizeof(T), n, f) == n;
}

unsigned long long checksum(const tdt::PreparedMesh &pm) {
    unsigned long long s = 0;   // wraps by design
    auto add = [&](const std::vector<int> &v) {
        for (size_t i = 0; i < v.size(); ++i)
            s = s * 1000003ULL + (unsigned long long)(long long)v[i] + i;
    };
    add(pm.cell_new2old); add(pm.node_new2old); add(pm.nod4); add(pm.nc4);
    add(pm.gparent); add(pm.gside); add(pm.gtype); add(pm.gopp);
    add(pm.nptr); add(pm.nlist); add(pm.novr); add(pm.g6ptr); add(pm.g6cell);
    return s;
}

// The wave node tables (wave_node_tables) checked against the cells' nodes.
bool wave_tables_ok(const tdt::PreparedMesh &pm, int *cap_out) {
    const int nwaves = (pm.n + 255) / 256 * 4;   // whole blocks of four waves
    std::vector<int> bnode, nloc;
    const int cap = tdt::wave_node_tables(pm.nod4, pm.n, nwaves, bnode, nloc);
    *cap_out = cap;
    if (cap <= 0 || cap % 64 != 0 || cap > 256) return false;
    if (bnode.size() != (size_t)nwaves * cap || nloc.size() != (size_t)2 * pm.n) return false;
    for (int w = 0; w < nwaves; ++w) {
        const int *row = &bnode[(size_t)w * cap];
        int len = 1;
        while (len < cap && row[len] > row[len - 1]) ++len;   // the ascending part
        for (int k = len; k < cap; ++k)
            if (row[k] != row[len - 1]) return false;         // padded with the last
        for (int k = 0; k < len; ++k)
            if (row[k] < 0 || row[k] >= pm.m) return false;
    }
    for (int i = 0; i < pm.n; ++i) {
        const int w = i / 64;
        const int sl[4] = {nloc[2 * i] & 0xFFFF, (int)((unsigned)nloc[2 * i] >> 16),
                           nloc[2 * i + 1] & 0xFFFF, (int)((unsigned)nloc[2 * i + 1] >> 16)};
        for (int q = 0; q < 4; ++q)
            if (sl[q] >= cap || bnode[(size_t)w * cap + sl[q]] != pm.nod4[4 * i + q]) return false;
    }
    return true;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t count = 0;
    if (fread(&count, 4, 1, f) != 1) return 2;
    for (int k = 0; k < count; ++k) {
        int32_t hdr[3];
        if (fread(hdr, 4, 3, f) != 3) return 2;
        std::vector<double> x, y;
        std::vector<int32_t> nod, nc, parent, ghost, type, side;
        const size_t nn = hdr[0] > 0 ? (size_t)hdr[0] : 0;
        const size_t ne = hdr[1] > 0 ? (size_t)hdr[1] : 0;
        const size_t ng = hdr[2] > 0 ? (size_t)hdr[2] : 0;
        if (!read_n(f, x, nn) || !read_n(f, y, nn) || !read_n(f, nod, 4 * ne) ||
            !read_n(f, nc, 4 * ne) || !read_n(f, parent, ng) || !read_n(f, ghost, ng) ||
            !read_n(f, type, ng) || !read_n(f, side, ng))
            return 2;
        tdt_mesh m{};
        m.nods = hdr[0]; m.neles = hdr[1]; m.nghosts = hdr[2];
        m.x_mm = x.data(); m.y_mm = y.data();
        m.nod = nod.data(); m.nc = nc.data();
        m.parent = parent.data(); m.ghost = ghost.data(); m.type = type.data();
        m.side = side.data();
        for (int o = 0; o < 4; ++o) {
            tdt::PreparedMesh pm;
            int code = 0;
            tdt::prepare_mesh(m, o, pm, &code);
            int cap = 0;
            const bool nlok = code == 0 && wave_tables_ok(pm, &cap);
            printf("mesh %d order %d code %d n %d g %d m %d bw %d used %d sum %llu nlcap %d "
                   "nlok %d\n", k, o, code, pm.n, pm.g, pm.m, pm.bandwidth, pm.ordering_used,
                   code == 0 ? checksum(pm) : 0ULL, cap, nlok ? 1 : 0);
        }
    }
    fclose(f);
    return 0;
}
