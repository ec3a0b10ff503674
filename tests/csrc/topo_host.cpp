// Host build of the island topology builder (exgame_amd/csrc/topology.cpp) for the
// CPU tests: tp_build fills one 320-byte record.
#include "../../exgame_amd/csrc/topology.cpp"

namespace exg {
// island_topology_of routes through these; the tests call build_island_topology
// directly with explicit shapes, so they are never reached here.
int island_class(int, int) { return -1; }
void island_class_shape(int, int& waves, int& per_wave) { waves = per_wave = 1; }
}  // namespace exg

extern "C" int tp_build_md(int m, int md, int J, const int* ja, const int* jb, const int* kind, int waves,
                           int per_wave, unsigned char* out) {
    exg::IslandTopo t;
    const int rc = exg::build_island_topology(m, md, J, ja, jb, kind, waves, per_wave, t);
    std::memcpy(out, &t, sizeof t);
    return rc;
}
// every body dynamic
extern "C" int tp_build(int m, int J, const int* ja, const int* jb, const int* kind, int waves, int per_wave,
                        unsigned char* out) {
    return tp_build_md(m, m, J, ja, jb, kind, waves, per_wave, out);
}
extern "C" int tp_size() { return (int)sizeof(exg::IslandTopo); }

// the mechanism-sized record: tp_build_large fills one LargeTopo; tp_large_layout
// writes the byte offsets of its eight fields and its size
extern "C" int tp_build_large_md(int m, int md, int J, const int* ja, const int* jb, unsigned char* out) {
    exg::LargeTopo t;
    const int rc = exg::build_large_topology(m, md, J, ja, jb, t);
    std::memcpy(out, &t, sizeof t);
    return rc;
}
// every body dynamic
extern "C" int tp_build_large(int m, int J, const int* ja, const int* jb, unsigned char* out) {
    return tp_build_large_md(m, m, J, ja, jb, out);
}
extern "C" void tp_large_layout(long long* out) {
    out[0] = offsetof(exg::LargeTopo, nbr_start);
    out[1] = offsetof(exg::LargeTopo, nbr);
    out[2] = offsetof(exg::LargeTopo, nbr_row);
    out[3] = offsetof(exg::LargeTopo, term_start);
    out[4] = offsetof(exg::LargeTopo, terms);
    out[5] = offsetof(exg::LargeTopo, inc_start);
    out[6] = offsetof(exg::LargeTopo, inc);
    out[7] = offsetof(exg::LargeTopo, kmask);
    out[8] = sizeof(exg::LargeTopo);
}
extern "C" long long tp_large_workspace(int m, int J) { return exg::large_workspace_doubles(m, J); }
