// extraction.h -- CSV time series of the explicit path.
//
// Mirrors classExtractData / classNodeData / classElementData
// (dataExtraction.cpp:150-505, 569-912): one CSV file per extraction in the
// output directory, header "Time,Value" (or "Time,Node<i>Comp<c>,..." for the
// Rough node operation), one row "%.16g,%.16g" per step whose time is larger
// than the last written time.  File names:
//   node:    <prefix>_<Unknown|InternalForce|ExternalForce><comp>_<operation>.csv
//   element: All_<type>_<operation>.csv for "All", else <prefix>_<type>_<operation>.csv
// The first extraction of a given name wins (addClassNodeDataToSave, :469-487).
//
// Values come from the device: nodal values from the per-step DOF series
// (mps_set_dof_series), element sums (DEFO_ENERGY, KIN_ENERGY, VOLUME,
// EXTERNAL_ENERGY with the Sum operation over All elements) from the per-step
// history ring, both read once per chunk of steps.
#ifndef MPS_EXTRACTION_H
#define MPS_EXTRACTION_H

#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "inp_reader.h"

namespace mps {

class classDataBase {
  public:
    classDataBase(std::string name, std::string op) : _name(std::move(name)), _operation(std::move(op)) {}
    virtual ~classDataBase() { closeFile(); }
    const std::string &getName() const { return _name; }
    const std::string &getOperation() const { return _operation; }
    double getLastTime() const { return _lastTime; }
    void closeFile();

  protected:
    bool openFile(const std::string &dir, const std::string &header);
    std::string _name, _operation;
    FILE *_pFile = nullptr;
    double _lastTime = 0.0;
};

class classNodeData : public classDataBase {
  public:
    classNodeData(const InpNodeExtraction &x);
    const std::vector<int32_t> &getNodes() const { return _nodes; }
    const std::string &getDataType() const { return _dataType; }
    int getComp() const { return _comp; }
    // Writes one row for `time` (skipped when time <= last time).
    bool writeData(const std::string &dir, double time, const std::vector<double> &vals);
    double getValue(const std::vector<double> &vals) const; // Mean / Max / Min / Sum (dataExtraction.cpp:247-280)

  private:
    std::vector<int32_t> _nodes;
    std::string _dataType;
    int _comp;
};

class classElementData : public classDataBase {
  public:
    classElementData(const InpElementExtraction &x);
    const std::string &getDataType() const { return _dataType; }
    bool writeData(const std::string &dir, double time, double val);

  private:
    std::string _dataType;
};

// *Forces output of one solver (classPrintForces, output.h:51-101; printForcesOut / writeExternalForces,
// output.cpp:106-135, 153-279): at every step whose time reaches iterations*max(period, dt) + t0 it appends
// one row "<node+1>,<direction>,<force>,<time>," per listed pair (std::to_string formatting) to
// <outputDir>/externalForces<t0>-<tf>.csv, created with the header "Node,Direction,Force,Time," at the first
// write.  The force is the step's internal force (intForce; the explicit solver's inertial term is zero,
// solverExplicit.cpp:606-613).
class classPrintForces {
  public:
    classPrintForces(const InpForces &f) : _f(f) {}
    const std::vector<int32_t> &getNodesToWrite() const { return _f.nodes; }
    const std::vector<int32_t> &getDirectionsToWrite() const { return _f.dirs; }
    double getPeriod() const { return _f.period; }
    int getIterationsCompleted() const { return _iterationsCompleted; }
    // Whether the step at `time` with step `dt` writes (the test of printForcesOut).
    bool due(double time, double dt, double t0) const;
    // Appends the rows of one step: forces[i] is the internal force of pair i.
    bool write(const std::string &dir, double t0, double tf, double time, const std::vector<double> &forces);

  private:
    InpForces _f;
    int _iterationsCompleted = 0;
    bool _firstTime = true;
};

class classExtractData {
  public:
    // Validates the extractions against what the device path provides; false + err otherwise.
    bool init(const InpModel &model, std::string &err);
    bool empty() const { return _node.empty() && _element.empty(); }
    bool needsNodal() const { return !_node.empty(); }
    bool needsEnergies() const { return !_element.empty(); }
    std::map<std::string, std::unique_ptr<classNodeData>> &nodeData() { return _node; }
    std::map<std::string, std::unique_ptr<classElementData>> &elementData() { return _element; }
    void finalise();

  private:
    std::map<std::string, std::unique_ptr<classNodeData>> _node;
    std::map<std::string, std::unique_ptr<classElementData>> _element;
};

} // namespace mps

#endif
