// readInput.h -- the INPUT keyword file of the reference driver
// (src/readInput.h:1-5, struct MCInput src/jmmMCState.h:66-101), with
// explicit defaults and a few additive keys for the engine.
#ifndef JMM_READINPUT_H
#define JMM_READINPUT_H

#include <string>
#include <vector>

struct MCInput {
    // reference keys (src/readInput.cpp:69-253); defaults are explicit here
    bool isRestart = false;            // RESTART
    unsigned long N = 0;               // N
    unsigned long numTrialTypes = 0;   // N+1 (NPT) or N (NLT), set by setupMCS
    double P = 1.0;                    // P
    double L = 0.0;                    // L (NLT length)
    double T = 1.0;                    // T
    unsigned long ns = 0;              // NUMSTEPS (trials)
    int relaxFlag = 0;                 // RELAX
    unsigned long cpi = 0, tpi = 0;    // CPI, TPI (0 = never)
    unsigned long eci = 0;             // ENGCHECK (trials, 0 = never)
    unsigned long mdai = 0, mvai = 0;  // DADJ, VADJ (step-size adaptation intervals)
    unsigned long gpi = 0, rhopi = 0;  // GPI, RHOPI
    unsigned long gnb = 0, rhonb = 0;  // GNB, RHONB
    unsigned long seed = 1;            // SEED
    double maxStep = 0.1;              // MAXSTEP
    double maxdl = 0.1;                // MAXDV
    double potCutOff;                  // second token of POT (INFINITY when absent)
    char potStr[80];                   // POT
    char ensembleStr[80];              // ENSEMBLE (NPT default, src/readInput.cpp:57)
    double rbw = 0.0, gsw = 0.0, gbw = 0.0;
    int gns = 0;
    int nbn = -1;                      // NBN
    int numComputes = 0;               // COMPUTE lines (stored, never run)
    std::vector<std::vector<std::string>> computeStrs;

    // additive keys (unknown to the reference, which only warns on unknown
    // keys, src/readInput.cpp:254-256)
    std::string boundary = "WALLS";    // BOUNDARY WALLS|PERIODIC
    std::string volMove = "LINEAR";    // VOLMOVE LINEAR|LOG
    std::string form = "BATCHED_FP64"; // FORM PLAIN_FP64|BATCHED_FP64|BATCHED_FP32
    double potParams[4] = {0, 0, 0, 0};
    int potParamsSet = 0;              // POTPARAMS p0 [p1 [p2 [p3]]]
    unsigned long chains = 1;          // CHAINS: independent replicas, seeds SEED+k
    unsigned long sweepBlock = 100;    // SWEEPBLOCK: sweeps per kernel launch
    unsigned long histEvery = 1;       // HISTEVERY: sweeps between rho(x)/g(x) samples
    unsigned long bcap = 256;          // BCAP: most particle blocks per chain
    double lStart = 0.0;               // LSTART: NPT initial length (0 = N, the reference's)
    int device = 0;                    // DEVICE
    std::string stepUnit = "TRIAL";    // STEPUNIT TRIAL|SWEEP: what getStepNum counts per Step()
    std::string adjust = "PUBLISHED";  // ADJUST PUBLISHED|REFERENCE: step-adaptation formula

    MCInput();
};

// Parses the keyword file at fstr.  Returns false when the file cannot be
// read (the reference prints and continues with garbage, src/readInput.cpp:51-54).
bool readInputFile(const char *fstr, MCInput &inp);
// The reference's by-value form (src/readInput.cpp:8); prints an error and
// returns the defaults when the file is missing.
MCInput readInput(const char *fstr);
int printInput(const MCInput &inp);
// One "key=value" line per field (machine-readable; used by the tests).
int printInputFields(const MCInput &inp);

#endif
