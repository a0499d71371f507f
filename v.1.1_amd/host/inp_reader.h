// inp_reader.h -- reader of the .inp subset of the explicit solid path.
//
// Grammar and semantics follow the reference readers, keyword by keyword:
//   *Node                       nodesManagement        inputFile.cpp:20-111
//   *Element, type=C3D8|C3D4    elementsManagement     inputFile.cpp:121-330
//   *FEM Nodes / *FEM Elements  (All)                  inputFile.cpp:331-378
//   *MATERIAL ... *END          materialManagement     inputFile.cpp:838-1018, 2010-2054, 2130
//     (*HyperElastic, Neo-Hookean | 3D-Compressible-Neo-Hookean | St-Venant-Kirchhoff; *Linear-Elastic)
//   *SOLVER, EXPLICIT ... *END SOLVER                  solverManagement inputFile.cpp:2253-2762
//   *BOUNDARY, TYPE=DISPLACEMENT RAMP|INST             dispBCManagement inputFile.cpp:2960-3217
//   *Outputs                    outManagement          inputFile.cpp:4769-4829
//   *Bulk viscosity / *Dev viscosity  solverManagement    inputFile.cpp:2462-2510
//   *BOUNDARY, TYPE=FORCE       forceBCManagement      inputFile.cpp:2841-2941  (anywhere in the file,
//   *BOUNDARY, TYPE=PRESSURE RAMP|INST  pressure*BCManagement inputFile.cpp:3580-3790  in order; global
//   *GRAVITY                    NeumannCommonManagement inputFile.cpp:4703-4735   for all solvers)
//   *EXTRACTION, NODE|ELEMENT   dataExtractionManagement inputFile.cpp:2229-2243, 4950-5120
//   *Forces                     forceOutManagement     inputFile.cpp:2537-2539, 4830-4887 (per solver)
// Keywords are matched as prefixes, case-sensitively, as the reference does.
// Where the reference would read garbage, dereference null or silently
// mis-number (duplicate Dirichlet DOFs, *BOUNDARY before *Time, a missing
// *Outputs, an element without a material) the reader reports an error.
//
// Extensions (keywords the reference does not know; a reference build skips
// them as unknown lines of a *SOLVER block):
//   *MPS Fixed Steps            n     run exactly n steps instead of the time window
//   *BOUNDARY, TYPE=VELOCITY    node,dir,V lines: prescribed velocity V*min(step/ramp,1)
//   *MPS Velocity Ramp Steps    k     ramp length of TYPE=VELOCITY (default 0: full V at once)
//   *MPS Precision              FP64 | FP32_ELEM
#ifndef MPS_INP_READER_H
#define MPS_INP_READER_H

#include <cstdint>
#include <string>
#include <vector>

namespace mps {

struct InpMaterial {
    std::string name;     // text after "*MATERIAL, name=" (inputFile.cpp:876)
    std::string nameCons; // "Neo-Hookean" | "3D-Compressible-Neo-Hookean" | "St-Venant-Kirchhoff" | "Linear-Elastic"
    double rho = 0, young = 0, nu = 0;
    bool all = false;           // *List of Elements / All
    std::vector<int32_t> elems; // 0-based element indices otherwise
};

struct InpDirichlet {
    int32_t node; // 0-based
    int32_t dir;  // 0..2
    double value; // U (ramp/inst) or V (velocity)
};

// *Forces of a solver block: output frequency, then the (node, direction) pairs whose internal force the
// solver writes to externalForces<t0>-<tf>.csv (printForcesOut, output.cpp:153-279).
struct InpForces {
    bool on = false;
    double period = 0;            // 1 / frequency
    std::vector<int32_t> nodes;   // 0-based
    std::vector<int32_t> dirs;    // 0..2
};

struct InpSolver {
    std::string type = "EXPLICIT";
    double scale = 1.0;
    double t0 = 0.0, tf = 0.0;
    bool timeSet = false;
    int numFiles = -1;       // *Outputs first line; -1 = no *Outputs block
    std::vector<int> outVars; // *Outputs optional second line: internal variables to print (outManagement)
    int numStepsIgnored = 0; // *Number Of Steps: parsed, not used by EXPLICIT (solverExplicit.cpp:158-171)
    int bcMode = 0;          // MPS_BC_RAMP / MPS_BC_INST / MPS_BC_VELOCITY
    int velRampSteps = 0;
    int64_t fixedSteps = 0;
    int precision = 0;
    int visc = 0;                  // *Bulk viscosity (MPS_VISC_BULK) / *Dev viscosity (MPS_VISC_DEV)
    double viscC1 = 0, viscCl = 0; // their "c1, cl" line (defaults 0.06, 1.2 / 1, 0.5)
    std::vector<InpDirichlet> dirichlet;
    InpForces forces;
};

struct InpNodeExtraction {
    std::string prefix, dataType, operation;
    int comp = 0;
    std::vector<int32_t> nodes; // 0-based
};

struct InpElementExtraction {
    std::string prefix, dataType, operation;
    bool all = false;
    std::vector<int32_t> elems; // 0-based
};

// One Neumann block (an entry of the reference's NeumannBCs list).  NeumannCommonManagement reads the whole
// file; the blocks apply to every solver (solverExplicit.cpp:47).
struct InpLoadBlock {
    int type = 0;                // MPS_LOAD_FORCE, _GRAVITY, _PRESSURE_RAMP, _PRESSURE_INST
    double t0 = 0, tf = 0;       // FORCE, PRESSURE window
    double value = 0;            // GRAVITY g, PRESSURE P
    std::vector<int32_t> index;  // FORCE: node; PRESSURE: element (0-based)
    std::vector<int32_t> face;   // PRESSURE: face S as getSubElement numbers it
    std::vector<double> force;   // FORCE: 3 per entry, F * Ff
};

struct InpLoads {
    std::vector<InpLoadBlock> blocks;  // input order
    bool empty() const { return blocks.empty(); }
};

struct InpModel {
    int ndim = 0;
    int32_t nNodes = 0, nElem = 0, nen = 0;
    std::string elemType;
    std::vector<double> X;     // SoA [3*nNodes]
    std::vector<int32_t> conn; // [nen*nElem] 0-based, file node order
    std::vector<InpMaterial> materials;
    std::vector<InpSolver> solvers;
    std::vector<InpNodeExtraction> nodeExtractions;
    std::vector<InpElementExtraction> elementExtractions;
    InpLoads loads;
    std::vector<std::string> warnings;
};

// Reads `path`.  Returns false and fills `err` on any error.
bool readInp(const std::string &path, InpModel &model, std::string &err);

// Per-element material index into model.materials (the last *MATERIAL that
// lists an element wins, as setConsModel overwrites, inputFile.cpp:2018-2054).
bool elementMaterials(const InpModel &model, std::vector<int32_t> &matOfElem, std::string &err);

} // namespace mps

#endif
