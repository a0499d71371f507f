// vtk_output.h -- VTK series of the explicit path.
//
// Mirrors vtkPlotting2D3D (output.cpp:296-918) for 3-D FEM meshes of C3D8 / C3D4 elements:
// one legacy ASCII file MuPhiSimOutputs_<rank>_.<n>.vtk per output, numbered across all
// solvers of a run, listed in allOutputs_<rank>.vtk.series (JSON, one entry per file with
// its time).  Sections, order and number format (iostream default: 6 significant digits)
// follow the reference writer:
//   POINTS, CELLS, CELL_TYPES;
//   POINT_DATA: Displacement, Velocity, Acceleration;
//   CELL_DATA: Cauchy, E (GP averages), VMS, VolumetricStrain, EquivalentStrain, Pressure
//   (tr of the averaged Cauchy / 3), IntVars<k> for the internal variables named on the
//   *Outputs line (the laws of this path have none: 0, as the reference prints for a
//   missing variable).
// The nodal and element fields come from the device (mps_snapshot_*, the element fields as
// mps_element_outputs computes them).
#ifndef MPS_VTK_OUTPUT_H
#define MPS_VTK_OUTPUT_H

#include <string>
#include <vector>

#include "inp_reader.h"

namespace mps {

class VtkSeriesWriter {
  public:
    explicit VtkSeriesWriter(std::string dir, int rank = 0) : _dir(std::move(dir)), _rank(rank) {}
    // u, v, a: SoA [3*n_nodes]; eout: [21*n_elem] as mps_element_outputs gives them.  The sections are
    // formatted on up to 16 threads; the geometry (POINTS, CELLS, CELL_TYPES) once per model.
    bool write(const InpModel &m, const double *u, const double *v, const double *a, const double *eout,
               const std::vector<int> &outVars, double timeRun, std::string &err);
    int filesWritten() const { return _cont; }

  private:
    bool updateSeries(double timeRun, std::string &err);
    const std::string &geometry(const InpModel &m);
    std::string _dir;
    int _rank;
    int _cont = 0;  // output counter (static `cont` of output.cpp:28)
    const InpModel *_geomModel = nullptr;
    std::string _geom;
};

} // namespace mps

#endif
