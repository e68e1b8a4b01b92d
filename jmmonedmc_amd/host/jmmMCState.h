// jmmMCState.h -- the reference's opaque-state chain API
// (src/jmmMCState.h:6-48) on top of the engine's C ABI (include/jmm_engine.h).
//
// Same lifecycle and names: setupMCS -> incrementStep/Step/print* -> freeMCS.
// The reference's collective functions (fad, incrementStep, Step, printG,
// relaxVolume) may be called by every thread of an OpenMP team, as
// src/Main.cpp does, or from one thread: one thread does the work and every
// caller gets its result.  What a "step" is differs: Step() advances
// the chain by one sweep (N displacement trials in the engine's block order,
// then one box-length trial for NPT), i.e. numTrialTypes reference trials, and
// the step counter counts trials as in the reference.  Interval predicates
// (isThermoPrint, ...) fire when the last sweep crossed a multiple of the
// interval.  Sweeps are queued and run on the GPU in blocks; any call that
// needs the state runs the queue first.
#ifndef JMM_MCSTATE_H
#define JMM_MCSTATE_H

#include "readInput.h"

struct MCState;

struct MCState *setupMCS(struct MCInput inp);  // src/jmmMCState.cpp:261
void freeMCS(struct MCState *mcs);              // :792
void printStep(struct MCState *mcs);            // :845
int printMCP(struct MCState *mcs);              // :216
int printThermo(struct MCState *mcs);           // :1896
int printCoords(struct MCState *mcs);           // :1007
int printRho(struct MCState *mcs);              // :1021 rho.dat.mcs
int printG(struct MCState *mcs);                // :1042 g<s>.dat.mcs
int fad(struct MCState *mcs, unsigned long *nm, double *d);  // :853 (initial energy)
unsigned long incrementStep(struct MCState *mcs);            // :1734
int Step(struct MCState *mcs);                  // :1758
int isCoordPrint(struct MCState *mcs);          // :1815
int isThermoPrint(struct MCState *mcs);         // :1825
int isRhoPrint(struct MCState *mcs);            // :1836
int isGPrint(struct MCState *mcs);              // :1847
int isECheck(struct MCState *mcs);              // :1858
int isMaxDisAdjust(struct MCState *mcs);        // :1870
int isMaxDVAdjust(struct MCState *mcs);         // :1883
int updateThermo(struct MCState *mcs);          // :1941
int ECheck(struct MCState *mcs);                // :1965
int maxDisAdjust(struct MCState *mcs);          // :2100 Swendsen step adaptation
int maxDVAdjust(struct MCState *mcs);           // :2120 Swendsen box-step adaptation
int relaxVolume(struct MCState *mcs);           // :2396 Newton relaxation of the box length
unsigned long getStepNum(struct MCState *mcs);  // :2777
int getRelaxFlag(struct MCState *mcs);
bool getRestartFlag(struct MCState *mcs);
int printE(struct MCState *mcs);                // :2143
int printAcc(struct MCState *mcs);              // :2151

// additive: the engine handle behind the facade, and the running energy and
// length of chain k
struct jmm_engine *getEngine(struct MCState *mcs);
double getEnergy(struct MCState *mcs, int chain);
double getLength(struct MCState *mcs, int chain);
int isRelaxStep(struct MCState *mcs);  // src/Main.cpp:173-176 as a predicate

#endif
