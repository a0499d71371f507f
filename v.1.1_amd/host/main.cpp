// main.cpp -- mps_explicit: command-line driver of the explicit path.
//
// Options follow GeneralOptions::initialise (commandLine.cpp:36-160):
//   -input <file>      input file name
//   -inputDir <dir>    folder of the input file (default "input")
//   -outputDir <dir>   folder of the CSV extractions (default "output")
// plus options of this engine:
//   -device <n>        HIP device ordinal (default 0)
//   -precision FP64|FP32_ELEM   overrides *MPS Precision
//   -check             read and validate the input, print a summary, run nothing
//   -dumpU <file>      write the final displacement (SoA, one "%.17g" per line)
//   -noVtk             skip the VTK series (CSV extractions are still written)
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "inp_reader.h"
#include "mps_gpu.h"
#include "solver_explicit.h"

static void usage() {
    std::printf("mps_explicit: explicit hex8/tet4 neo-Hookean dynamics on an MI355X\n");
    std::printf("Usage: mps_explicit -input file.inp [-inputDir dir] [-outputDir dir] [-device n]\n");
    std::printf("       [-precision FP64|FP32_ELEM] [-check] [-dumpU file] [-noVtk]\n");
}

static int check(const std::string &path) {
    mps::InpModel m;
    std::string err;
    if (!mps::readInp(path, m, err)) {
        std::fprintf(stderr, "ERROR ::: %s\n", err.c_str());
        return 1;
    }
    std::vector<int32_t> matOfElem;
    if (!mps::elementMaterials(m, matOfElem, err)) {
        std::fprintf(stderr, "ERROR ::: %s\n", err.c_str());
        return 1;
    }
    mps::classExtractData x;
    if (!x.init(m, err)) {
        std::fprintf(stderr, "ERROR ::: %s\n", err.c_str());
        return 1;
    }
    // Order-independent exact fingerprints of what was read: XOR of the coordinate bit patterns
    // (SoA order) and the sum of the 0-based connectivity.
    uint64_t xbits = 0;
    for (double x : m.X) {
        uint64_t b;
        std::memcpy(&b, &x, 8);
        xbits ^= b;
    }
    long long csum = 0;
    for (int32_t c : m.conn) csum += c;
    std::printf("{\"nodes\": %d, \"elements\": %d, \"type\": \"%s\", \"x_xor\": \"%016llx\", \"conn_sum\": %lld, "
                "\"materials\": [",
                m.nNodes, m.nElem, m.elemType.c_str(), (unsigned long long)xbits, csum);
    for (size_t k = 0; k < m.materials.size(); k++)
        std::printf("%s{\"name\": \"%s\", \"law\": \"%s\", \"rho\": %.17g, \"E\": %.17g, \"nu\": %.17g}", k ? ", " : "",
                    m.materials[k].name.c_str(), m.materials[k].nameCons.c_str(), m.materials[k].rho,
                    m.materials[k].young, m.materials[k].nu);
    std::printf("], \"solvers\": [");
    for (size_t k = 0; k < m.solvers.size(); k++) {
        const mps::InpSolver &s = m.solvers[k];
        std::printf("%s{\"scale\": %.17g, \"t0\": %.17g, \"tf\": %.17g, \"outputs\": %d, \"bc_mode\": %d, "
                    "\"dirichlet\": %zu, \"fixed_steps\": %lld, \"vel_ramp_steps\": %d, \"precision\": %d, "
                    "\"visc\": %d, \"visc_c1\": %.17g, \"visc_cl\": %.17g, \"out_vars\": [",
                    k ? ", " : "", s.scale, s.t0, s.tf, s.numFiles, s.bcMode, s.dirichlet.size(),
                    (long long)s.fixedSteps, s.velRampSteps, s.precision, s.visc, s.viscC1, s.viscCl);
        for (size_t j = 0; j < s.outVars.size(); j++) std::printf("%s%d", j ? ", " : "", s.outVars[j]);
        std::printf("], \"forces\": ");
        if (!s.forces.on) {
            std::printf("null");
        } else {
            std::printf("{\"period\": %.17g, \"entries\": [", s.forces.period);
            for (size_t j = 0; j < s.forces.nodes.size(); j++)
                std::printf("%s[%d, %d]", j ? ", " : "", s.forces.nodes[j], s.forces.dirs[j]);
            std::printf("]}");
        }
        std::printf("}");
    }
    std::printf("], \"node_extractions\": [");
    bool first = true;
    for (auto &kv : x.nodeData()) {
        std::printf("%s\"%s\"", first ? "" : ", ", kv.first.c_str());
        first = false;
    }
    std::printf("], \"element_extractions\": [");
    first = true;
    for (auto &kv : x.elementData()) {
        std::printf("%s\"%s\"", first ? "" : ", ", kv.first.c_str());
        first = false;
    }
    std::printf("], \"loads\": [");
    const char *names[6] = {"FORCE", "GRAVITY", "PRESSURE RAMP", "PRESSURE INST", "TRACTION RAMP", "TRACTION INST"};
    for (size_t k = 0; k < m.loads.blocks.size(); k++) {
        const mps::InpLoadBlock &b = m.loads.blocks[k];
        std::printf("%s{\"type\": \"%s\", \"t0\": %.17g, \"tf\": %.17g, \"value\": %.17g, ", k ? ", " : "",
                    names[b.type], b.t0, b.tf, b.value);
        if (b.type >= MPS_LOAD_TRACTION_RAMP)
            std::printf("\"traction\": [%.17g, %.17g, %.17g], ", b.force[0], b.force[1], b.force[2]);
        std::printf("\"entries\": [");
        for (size_t j = 0; j < b.index.size(); j++) {
            if (b.type == MPS_LOAD_FORCE)
                std::printf("%s[%d, %.17g, %.17g, %.17g]", j ? ", " : "", b.index[j], b.force[3 * j], b.force[3 * j + 1],
                            b.force[3 * j + 2]);
            else
                std::printf("%s[%d, %d]", j ? ", " : "", b.index[j], b.face[j]);
        }
        std::printf("]}");
    }
    std::printf("], \"warnings\": %zu}\n", m.warnings.size());
    return 0;
}

int main(int argc, char **argv) {
    if (argc == 1) {
        usage();
        return 0;
    }
    std::string input, inputDir = "input";
    bool inputDirSet = false, checkOnly = false;
    std::string dumpU;
    mps::RunOptions opt;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        const char *next = i + 1 < argc ? argv[i + 1] : nullptr;
        if (!std::strcmp(a, "-h") || !std::strcmp(a, "-help")) {
            usage();
            return 0;
        } else if (!std::strcmp(a, "-check")) {
            checkOnly = true;
        } else if (!std::strcmp(a, "-noVtk")) {
            opt.writeVtk = false;
        } else if (!next) {
            std::fprintf(stderr, "ERROR ::: option %s needs a value\n", a);
            return 2;
        } else if (!std::strcmp(a, "-input")) {
            input = next, i++;
        } else if (!std::strcmp(a, "-inputDir")) {
            inputDir = next, inputDirSet = true, i++;
        } else if (!std::strcmp(a, "-outputDir")) {
            opt.outputDir = next, i++;
        } else if (!std::strcmp(a, "-device")) {
            opt.device = std::atoi(next), i++;
        } else if (!std::strcmp(a, "-precision")) {
            if (!std::strcmp(next, "FP64"))
                opt.precision = MPS_PREC_FP64;
            else if (!std::strcmp(next, "FP32_ELEM"))
                opt.precision = MPS_PREC_FP32_ELEM;
            else {
                std::fprintf(stderr, "ERROR ::: -precision is FP64 or FP32_ELEM\n");
                return 2;
            }
            i++;
        } else if (!std::strcmp(a, "-dumpU")) {
            dumpU = next, i++;
        } else {
            std::fprintf(stderr, "ERROR ::: unknown option %s\n", a);
            return 2;
        }
    }
    if (input.empty()) {
        std::fprintf(stderr, "ERROR ::: no -input file\n");
        return 2;
    }
    // The input file is looked up in -inputDir; a path that exists as given
    // is also accepted when no -inputDir was set.
    std::string path = inputDir + "/" + input;
    if (!inputDirSet && input.find('/') != std::string::npos) path = input;
    if (!inputDirSet) {
        FILE *f = std::fopen(path.c_str(), "r");
        if (!f) path = input;
        else std::fclose(f);
    }
    if (checkOnly) return check(path);
    std::printf("INFO ::: Input file: %s\n", path.c_str());
    std::string err;
    std::vector<double> u;
    if (mps::runInputFile(path, opt, err, dumpU.empty() ? nullptr : &u) != 0) {
        std::fprintf(stderr, "ERROR ::: %s\n", err.c_str());
        return 1;
    }
    if (!dumpU.empty()) {
        FILE *f = std::fopen(dumpU.c_str(), "w");
        if (!f) {
            std::fprintf(stderr, "ERROR ::: cannot write %s\n", dumpU.c_str());
            return 1;
        }
        for (double x : u) std::fprintf(f, "%.17g\n", x);
        std::fclose(f);
    }
    return 0;
}
