// dsmcFoamPlus -- the dsmcFoam+ time loop (applications/solvers/
// discreteMethods/dsmc/dsmcFoam+/dsmcFoam+.C:92-180) on the MI355X engine:
// construct the cloud, loop over time, evolve, report, write.
//
// usage: dsmcFoamPlus [-case <dir>]
#include <chrono>
#include <cstring>
#include <iostream>

#include "dsmcCloud.H"

int main(int argc, char** argv) {
    std::string caseDir = ".";
    for (int i = 1; i < argc; ++i)
        if (!std::strcmp(argv[i], "-case") && i + 1 < argc) caseDir = argv[++i];
    try {
        std::cout << "\nConstructing dsmcCloud\n" << std::endl;
        hs::dsmcCloud dsmc(caseDir);
        std::cout << "\nStarting time loop\n" << std::endl;
        const auto t0 = std::chrono::steady_clock::now();
        int infoCounter = 0;
        long long iteration = 0;
        while (dsmc.loop()) {
            ++infoCounter;
            const bool report = infoCounter >= dsmc.nTerminalOutputs();
            if (report) std::cout << "Time = " << dsmc.timeName() << "\n" << std::endl;
            dsmc.evolve();
            if (report) dsmc.info();
            if (dsmc.writeTime()) dsmc.write();
            ++iteration;
            if (report) {
                const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                std::cout << "\nIteration " << iteration << "  ClockTime = " << s << " s\n" << std::endl;
                infoCounter = 0;
            }
        }
        std::cout << "End\n" << std::endl;
    } catch (const hs::FatalError& e) {
        std::cerr << "\n--> FOAM FATAL ERROR:\n" << e.what() << "\n\nFOAM exiting\n" << std::endl;
        return 1;
    }
    return 0;
}
