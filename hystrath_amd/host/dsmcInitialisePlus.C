// dsmcInitialisePlus -- dsmcInitialise+ (applications/utilities/
// preProcessing/dsmc/dsmcInitialise+/dsmcInitialise+.C:38-87) on the MI355X
// engine: fill the mesh from system/dsmcInitialiseDict and write the
// initial parcel state at the start time.
//
// usage: dsmcInitialisePlus [-case <dir>]
#include <cstring>
#include <iostream>

#include "dsmcCloud.H"

int main(int argc, char** argv) {
    std::string caseDir = ".";
    for (int i = 1; i < argc; ++i)
        if (!std::strcmp(argv[i], "-case") && i + 1 < argc) caseDir = argv[++i];
    try {
        hs::dsmcCloud dsmc(caseDir, true);
        dsmc.initialise();
        dsmc.write();
        std::cout << "End\n" << std::endl;
    } catch (const hs::FatalError& e) {
        std::cerr << "\n--> FOAM FATAL ERROR:\n" << e.what() << "\n\nFOAM exiting\n" << std::endl;
        return 1;
    }
    return 0;
}
