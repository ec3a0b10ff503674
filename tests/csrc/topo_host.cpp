// Host build of the island topology builder (exgame_amd/csrc/topology.cpp) for the
// CPU tests: tp_build fills one 320-byte record.
#include "../../exgame_amd/csrc/topology.cpp"

namespace exg {
// island_topology_of routes through these; the tests call build_island_topology
// directly with explicit shapes, so they are never reached here.
int island_class(int, int) { return -1; }
void island_class_shape(int, int& waves, int& per_wave) { waves = per_wave = 1; }
}  // namespace exg

extern "C" int tp_build(int m, int J, const int* ja, const int* jb, const int* kind, int waves, int per_wave,
                        unsigned char* out) {
    exg::IslandTopo t;
    const int rc = exg::build_island_topology(m, J, ja, jb, kind, waves, per_wave, t);
    std::memcpy(out, &t, sizeof t);
    return rc;
}
extern "C" int tp_size() { return (int)sizeof(exg::IslandTopo); }
