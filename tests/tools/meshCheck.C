// meshCheck -- the host's polyMesh reader without a device:
//   meshCheck <caseDir>
// prints the block read from <caseDir>/constant/polyMesh:
//   n <nx> <ny> <nz>
//   lo <x> <y> <z>
//   hi <x> <y> <z>
//   side <xmin|xmax|ymin|ymax|zmin|zmax> <patch> <type>     (six lines)
//   patches <name> ...
// Errors print "!! <message>" and exit with status 1.
#include <charconv>
#include <iostream>

#include "blockMesh.H"

namespace {
void real(double x) {
    char b[32];
    auto r = std::to_chars(b, b + sizeof(b), x);
    std::cout << ' ';
    std::cout.write(b, r.ptr - b);
}
}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::cerr << "usage: meshCheck <caseDir>\n";
        return 2;
    }
    try {
        if (!hs::hasPolyMesh(argv[1])) throw hs::FatalError(std::string(argv[1]) + ": no constant/polyMesh");
        const hs::BlockMesh m = hs::readPolyMesh(argv[1]);
        std::cout << "n " << m.n[0] << ' ' << m.n[1] << ' ' << m.n[2] << "\nlo";
        for (double x : m.lo) real(x);
        std::cout << "\nhi";
        for (double x : m.hi) real(x);
        std::cout << '\n';
        const char* side[6] = { "xmin", "xmax", "ymin", "ymax", "zmin", "zmax" };
        for (int s = 0; s < 6; ++s) std::cout << "side " << side[s] << ' ' << m.sidePatch[s] << ' ' << m.sideType[s] << '\n';
        std::cout << "patches";
        for (const auto& p : m.patchNames) std::cout << ' ' << p;
        std::cout << '\n';
        return 0;
    } catch (const hs::FatalError& e) {
        std::cout << "!! " << e.what() << "\n";
        return 1;
    }
}
