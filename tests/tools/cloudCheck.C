// cloudCheck -- exercises the lagrangian cloud reader and writer without a
// device:
//   cloudCheck copy <timeDirIn> <timeDirOut> <timeName> <modes> <nCells> <nSpecies>
//       reads the cloud of <timeDirIn> and writes it to <timeDirOut>
//   cloudCheck field <volScalarField file> <nCells>
//       prints the internalField, one value per line
// Errors print "!! <message>" and exit with status 1.
#include <charconv>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "lagrangianIO.H"

int main(int argc, char** argv) {
    try {
        if (argc == 8 && std::strcmp(argv[1], "copy") == 0) {
            const hs::LagrangianCloud c =
                hs::readLagrangianCloud(argv[2], std::atoi(argv[5]), std::atoll(argv[6]), std::atoi(argv[7]));
            hs::writeLagrangianCloud(argv[3], argv[4], c);
            std::cout << "copied " << c.size() << " parcels\n";
            return 0;
        }
        if (argc == 4 && std::strcmp(argv[1], "field") == 0) {
            for (double x : hs::readVolScalarInternalField(argv[2], std::atoll(argv[3]))) {
                char b[32];
                auto r = std::to_chars(b, b + sizeof(b), x);
                std::cout.write(b, r.ptr - b) << '\n';
            }
            return 0;
        }
        std::cerr << "usage: cloudCheck copy <in> <out> <timeName> <modes> <nCells> <nSpecies>\n"
                     "       cloudCheck field <file> <nCells>\n";
        return 2;
    } catch (const hs::FatalError& e) {
        std::cout << "!! " << e.what() << "\n";
        return 1;
    }
}
