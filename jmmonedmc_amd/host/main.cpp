// main.cpp -- host driver with the reference's loop (src/Main.cpp:27-192):
// read INPUT, set up, step-0 energy and output, then
// while (incrementStep) { Step; interval-gated output }, then the final
// energy and acceptance counts.  The chain work runs on the GPU through the
// jmmMCState-style facade; the file name of INPUT may be given as argv[1].
#include <omp.h>

#include <cstdio>
#include <string>

#include "jmmMCState.h"
#include "readInput.h"

int main(int argc, char *argv[])
{
    std::printf("#####################################################\n");
    std::printf("#   1D Metropolis Monte Carlo on MI355X (gfx950)    #\n");
    std::printf("#####################################################\n\n");

    // --print-input FILE: parse and print the fields, touch no device
    const bool parse_only = argc > 2 && std::string(argv[1]) == "--print-input";
    const char *path = parse_only ? argv[2] : (argc > 1 ? argv[1] : "INPUT");
    MCInput inp;
    if (!readInputFile(path, inp)) {
        std::printf("FATAL ERROR: %s not found.\n", path);
        return 1;
    }
    if (parse_only) {
        printInputFields(inp);
        return 0;
    }
    struct MCState *mcs = setupMCS(inp);
    if (!mcs) return 1;
    std::printf("Setup completed\n");

    // the reference's team structure: collective calls (fad, relaxVolume,
    // printG, incrementStep, Step) by every thread, output by one
    int failed = 0;
#pragma omp parallel default(shared)
    {
#pragma omp single
        {
            std::printf("Running with %d threads...\n\n", omp_get_num_threads());
            printStep(mcs);
        }
        unsigned long nnn = 0;
        double ddd = 0.5;
        fad(mcs, &nnn, &ddd);
        if (getRelaxFlag(mcs) > 0) relaxVolume(mcs);

        if (!getRestartFlag(mcs)) {
#pragma omp single
            {
                printCoords(mcs);
                printRho(mcs);
                updateThermo(mcs);
                printThermo(mcs);
            }
            printG(mcs);
        }

        while (incrementStep(mcs)) {
            if (Step(mcs)) {
#pragma omp single
                failed = 1;
                break;
            }
#pragma omp single
            {
                if (isCoordPrint(mcs)) printCoords(mcs);
                if (isThermoPrint(mcs)) printThermo(mcs);
                if (isRhoPrint(mcs)) printRho(mcs);
                if (isMaxDisAdjust(mcs)) maxDisAdjust(mcs);
                if (isMaxDVAdjust(mcs)) maxDVAdjust(mcs);
            }
            if (isGPrint(mcs)) printG(mcs);
            if (isRelaxStep(mcs)) relaxVolume(mcs);
            std::fflush(stdout);
            // every thread has read the step counter for the predicates above
            // before incrementStep's single region advances it; without this
            // a late thread could see the next step, call a collective the
            // others skip, and the team would deadlock
#pragma omp barrier
        }
    }
    if (failed) {
        std::fprintf(stderr, "Step failed\n");
        freeMCS(mcs);
        return 1;
    }

    std::printf("\nPROGRAM COMPLETED SUCCESSFULLY!\n");
    printE(mcs);
    printAcc(mcs);
    freeMCS(mcs);
    return 0;
}
