// solver_explicit.h -- solver plugin of the explicit path on the device engine.
//
// Mirrors the reference's solver plugin (class solvers, solvers.h:28-189) and
// classSolverExplicit (solverExplicit.h / solverExplicit.cpp:33-361): one
// object per "*SOLVER, EXPLICIT" block, run in file order, each handing its
// final u, v, a to the next (initValues / getFinalValues, solvers.h:171-187).
// calculate() drives the C-ABI of libmps_gpu (mps_create, mps_run, ...)
// instead of the PETSc vectors and per-Gauss-point objects of the reference,
// and writes the same CSV extractions.
#ifndef MPS_SOLVER_EXPLICIT_H
#define MPS_SOLVER_EXPLICIT_H

#include <memory>
#include <string>
#include <vector>

#include "constitutive.h"
#include "extraction.h"
#include "inp_reader.h"

namespace mps {

class VtkSeriesWriter;

struct RunOptions {
    std::string outputDir = "output";
    bool writeVtk = true;           // VTK series as the reference writes it (outputsManagement)
    VtkSeriesWriter *vtk = nullptr; // set by runInputFile: one writer numbers the files of all solvers
    int device = 0;
    int precision = -1; // -1: as the input says; else MPS_PREC_*
    bool quiet = false;
};

class solvers {
  public:
    virtual ~solvers() = default;
    // Runs the solver; returns 0 on success, -1 on error (message in err).
    virtual int calculate(const InpModel &model, const std::vector<std::unique_ptr<constitutiveModels>> &laws,
                          const std::vector<int32_t> &lawOfElem, classExtractData &extract, const RunOptions &opt,
                          std::string &err) = 0;
    // Hand-over of u, v, a between chained solvers (solvers.h:171-187).
    void initValues(const std::vector<double> &u, const std::vector<double> &v, const std::vector<double> &a) {
        _u0 = u;
        _v0 = v;
        _a0 = a;
    }
    void getFinalValues(std::vector<double> &u, std::vector<double> &v, std::vector<double> &a) const {
        u = _u;
        v = _v;
        a = _a;
    }
    long long stepsTaken() const { return _steps; }
    double finalTime() const { return _time; }

  protected:
    std::vector<double> _u0, _v0, _a0, _u, _v, _a;
    long long _steps = 0;
    double _time = 0;
};

class classSolverExplicit : public solvers {
  public:
    explicit classSolverExplicit(const InpSolver &s) : _s(s) {}
    int calculate(const InpModel &model, const std::vector<std::unique_ptr<constitutiveModels>> &laws,
                  const std::vector<int32_t> &lawOfElem, classExtractData &extract, const RunOptions &opt,
                  std::string &err) override;

  private:
    InpSolver _s;
};

// inputManagement + the solver loop of main() (main.cpp:36-218) for the
// explicit path: read, build laws, run every solver, write extractions.
// Returns 0 on success, -1 on error.
int runInputFile(const std::string &path, const RunOptions &opt, std::string &err,
                 std::vector<double> *finalU = nullptr);

} // namespace mps

#endif
