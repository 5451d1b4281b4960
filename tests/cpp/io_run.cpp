// The mesh-file calls of the C++ host classes: reads a Gmsh file, writes a
// Tecplot file with two cell arrays (k * 1000 + cell index), prints the
// mesh's cell and face counts.
#include <cstdio>
#include <string>

#include "dndsmi.hpp"

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: io_run mesh.msh out.plt\n");
        return 2;
    }
    try {
        auto mesh = dndsmi::CompactFacedMeshSerialRW::read_gmsh(argv[1]);
        const char *names[2] = {"A", "B"};
        mesh.PrintSerialPartPltBinaryDataArray(
            argv[2], 2, [&](int k) { return std::string(names[k]); },
            [&](int k, dndsmi::index i) { return double(k * 1000 + i); }, 1.5);
        std::printf("%lld %lld\n", (long long)mesh.NumCell(), (long long)mesh.NumFace());
    } catch (const dndsmi::Error &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
