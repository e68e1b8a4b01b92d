// readInput.cpp -- INPUT keyword parser.  Same keywords and value rules as
// the reference (src/readInput.cpp:8-262): one "KEY value..." per line, keys
// matched exactly, numbers read with strtod and then cast (so 1E6 is a valid
// integer), an unknown POT falls back to LJ with a message, an unknown key
// prints "Property command ... not understood." and is skipped.  Differences,
// all on inputs the reference mishandles: blank lines and '#' comments are
// skipped (the reference dereferences a NULL token), lines of any length are
// accepted (the reference reads 200 bytes into a 100-byte buffer), and every
// field has a defined default (MCInput's initialisers).
#include "readInput.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

MCInput::MCInput()
{
    potCutOff = INFINITY;
    std::strncpy(potStr, "LJ", sizeof potStr);
    std::strncpy(ensembleStr, "NPT", sizeof ensembleStr);
}

static double num(const std::vector<std::string> &t, size_t k)
{
    if (k >= t.size()) return 0.0;
    return std::strtod(t[k].c_str(), nullptr);
}

static unsigned long unum(const std::vector<std::string> &t, size_t k)
{
    return (unsigned long)num(t, k);
}

static std::string upper(std::string s)
{
    for (auto &c : s) c = (char)std::toupper((unsigned char)c);
    return s;
}

bool readInputFile(const char *fstr, MCInput &inp)
{
    std::ifstream f(fstr);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ss(line);
        std::vector<std::string> t;
        for (std::string w; ss >> w;) t.push_back(w);
        if (t.empty()) continue;
        const std::string &k = t[0];
        if (k == "RESTART") {
            inp.isRestart = true;
            if (t.size() > 1 && (t[1] == "FALSE" || t[1] == "False" || t[1] == "false")) inp.isRestart = false;
        } else if (k == "N") {
            inp.N = unum(t, 1);
        } else if (k == "P") {
            inp.P = num(t, 1);
        } else if (k == "L") {
            inp.L = num(t, 1);
        } else if (k == "T") {
            inp.T = num(t, 1);
        } else if (k == "NBN") {
            inp.nbn = (int)num(t, 1);
        } else if (k == "NUMSTEPS") {
            inp.ns = unum(t, 1);
        } else if (k == "POT") {
            static const char *pots[] = {"LJcut", "LJ", "HARMONIC", "LJshift", "HRSW", "YUKAWA"};
            const std::string name = t.size() > 1 ? t[1] : "";
            bool known = false;
            for (const char *p : pots) known = known || name == p;
            if (known) {
                std::strncpy(inp.potStr, name.c_str(), sizeof inp.potStr - 1);
            } else {
                std::printf("Unknown potential type requested (%s). Using Lennard-Jones potential with no cut-off.\n",
                            name.c_str());
                std::strncpy(inp.potStr, "LJ", sizeof inp.potStr);
            }
            inp.potCutOff = t.size() > 2 ? num(t, 2) : INFINITY;
        } else if (k == "MAXSTEP") {
            inp.maxStep = num(t, 1);
        } else if (k == "MAXDV") {
            inp.maxdl = num(t, 1);
        } else if (k == "CPI") {
            inp.cpi = unum(t, 1);
        } else if (k == "TPI") {
            inp.tpi = unum(t, 1);
        } else if (k == "GPI") {
            inp.gpi = unum(t, 1);
        } else if (k == "RHOPI") {
            inp.rhopi = unum(t, 1);
        } else if (k == "RBW") {
            inp.rbw = num(t, 1);
        } else if (k == "RHONB") {
            inp.rhonb = unum(t, 1);
        } else if (k == "GBW") {
            inp.gbw = num(t, 1);
        } else if (k == "GNB") {
            inp.gnb = unum(t, 1);
        } else if (k == "GSW") {
            inp.gsw = num(t, 1);
        } else if (k == "GNS") {
            inp.gns = (int)num(t, 1);
        } else if (k == "SEED") {
            inp.seed = unum(t, 1);
        } else if (k == "ENGCHECK") {
            inp.eci = unum(t, 1);
        } else if (k == "DADJ") {
            inp.mdai = unum(t, 1);
        } else if (k == "VADJ") {
            inp.mvai = unum(t, 1);
        } else if (k == "RELAX") {
            inp.relaxFlag = 1;
        } else if (k == "ENSEMBLE") {
            const std::string e = t.size() > 1 ? t[1] : "";
            if (e == "NPT" || e == "NLT") {
                std::strncpy(inp.ensembleStr, e.c_str(), sizeof inp.ensembleStr - 1);
            } else {
                std::printf("Unknown ensemble type requested (%s). Using NPT ensemble.\n", e.c_str());
                std::strncpy(inp.ensembleStr, "NPT", sizeof inp.ensembleStr);
            }
        } else if (k == "COMPUTE" || k == "compute") {
            inp.computeStrs.emplace_back(t.begin() + 1, t.end());
            inp.numComputes++;
        } else if (k == "BOUNDARY") {
            inp.boundary = upper(t.size() > 1 ? t[1] : "");
        } else if (k == "VOLMOVE") {
            inp.volMove = upper(t.size() > 1 ? t[1] : "");
        } else if (k == "FORM") {
            inp.form = upper(t.size() > 1 ? t[1] : "");
        } else if (k == "POTPARAMS") {
            for (size_t q = 1; q < t.size() && q <= 4; ++q) inp.potParams[q - 1] = num(t, q);
            inp.potParamsSet = (int)std::min<size_t>(t.size() - 1, 4);
        } else if (k == "CHAINS") {
            inp.chains = unum(t, 1);
        } else if (k == "SWEEPBLOCK") {
            inp.sweepBlock = unum(t, 1);
        } else if (k == "HISTEVERY") {
            inp.histEvery = std::max(1ul, unum(t, 1));
        } else if (k == "BCAP") {
            inp.bcap = unum(t, 1);
        } else if (k == "LSTART") {
            inp.lStart = num(t, 1);
        } else if (k == "DEVICE") {
            inp.device = (int)num(t, 1);
        } else if (k == "STEPUNIT") {
            inp.stepUnit = upper(t.size() > 1 ? t[1] : "");
        } else if (k == "ADJUST") {
            inp.adjust = upper(t.size() > 1 ? t[1] : "");
        } else {
            std::printf("Property command %s not understood.\n", k.c_str());
        }
    }
    return true;
}

MCInput readInput(const char *fstr)
{
    MCInput inp;
    if (!readInputFile(fstr, inp)) std::printf("FATAL ERROR: %s not found.", fstr);
    return inp;
}

int printInput(const MCInput &inp)
{
    std::printf("N = %lu\n", inp.N);
    if (std::strncmp(inp.ensembleStr, "NPT", 80) == 0) std::printf("P = %.5G\n", inp.P);
    else std::printf("L = %.5G\n", inp.L);
    std::printf("T = %.5G\n", inp.T);
    if (inp.nbn > 0) std::printf("NBN = %d\n", inp.nbn);
    else std::printf("NBN = No neighbor number limit.\n");
    std::printf("NUMSTEPS = %lu\n", inp.ns);
    std::printf("POT = %s, CUTOFF = %.5G\n", inp.potStr, inp.potCutOff);
    std::printf("MAXSTEP = %.5G\nMAXDV = %.5G\nENGCHECK = %lu\n", inp.maxStep, inp.maxdl, inp.eci);
    std::printf("DADJ = %lu\nVADJ = %lu\nCPI = %lu\nTPI = %lu\nGPI = %lu\nRHOPI = %lu\n", inp.mdai,
                inp.mvai, inp.cpi, inp.tpi, inp.gpi, inp.rhopi);
    std::printf("RBW = %.5G\nRHONB = %lu\nGBW = %.5G\nGNB = %lu\nGSW = %.5G\nGNS = %d\n", inp.rbw,
                inp.rhonb, inp.gbw, inp.gnb, inp.gsw, inp.gns);
    std::printf("SEED = %lu\nRELAXATION FLAG = %d\n", inp.seed, inp.relaxFlag);
    std::printf("BOUNDARY = %s\nVOLMOVE = %s\nFORM = %s\nCHAINS = %lu\n", inp.boundary.c_str(),
                inp.volMove.c_str(), inp.form.c_str(), inp.chains);
    std::printf("computes = \n");
    for (const auto &c : inp.computeStrs) {
        std::printf("\t");
        for (const auto &w : c) std::printf("%s ", w.c_str());
        std::printf("\n");
    }
    return 0;
}

int printInputFields(const MCInput &i)
{
    std::printf("isRestart=%d\nN=%lu\nP=%.17g\nL=%.17g\nT=%.17g\nns=%lu\nrelaxFlag=%d\n", (int)i.isRestart,
                i.N, i.P, i.L, i.T, i.ns, i.relaxFlag);
    std::printf("cpi=%lu\ntpi=%lu\neci=%lu\nmdai=%lu\nmvai=%lu\ngpi=%lu\nrhopi=%lu\ngnb=%lu\nrhonb=%lu\n", i.cpi,
                i.tpi, i.eci, i.mdai, i.mvai, i.gpi, i.rhopi, i.gnb, i.rhonb);
    std::printf("seed=%lu\nmaxStep=%.17g\nmaxdl=%.17g\npotCutOff=%.17g\npotStr=%s\nensembleStr=%s\n", i.seed,
                i.maxStep, i.maxdl, i.potCutOff, i.potStr, i.ensembleStr);
    std::printf("rbw=%.17g\ngsw=%.17g\ngbw=%.17g\ngns=%d\nnbn=%d\nnumComputes=%d\n", i.rbw, i.gsw, i.gbw, i.gns,
                i.nbn, i.numComputes);
    std::printf("boundary=%s\nvolMove=%s\nform=%s\nchains=%lu\nsweepBlock=%lu\nbcap=%lu\nlStart=%.17g\n",
                i.boundary.c_str(), i.volMove.c_str(), i.form.c_str(), i.chains, i.sweepBlock, i.bcap, i.lStart);
    std::printf("histEvery=%lu\nstepUnit=%s\nadjust=%s\n", i.histEvery, i.stepUnit.c_str(), i.adjust.c_str());
    std::printf("potParams=%.17g,%.17g,%.17g,%.17g\npotParamsSet=%d\n", i.potParams[0], i.potParams[1],
                i.potParams[2], i.potParams[3], i.potParamsSet);
    return 0;
}
