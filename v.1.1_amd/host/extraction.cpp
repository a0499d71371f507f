// extraction.cpp -- CSV time series of the explicit path (extraction.h).
#include "extraction.h"

#include <algorithm>
#include <numeric>

namespace mps {

void classDataBase::closeFile() {
    if (_pFile) {
        std::fclose(_pFile);
        _pFile = nullptr;
    }
}

bool classDataBase::openFile(const std::string &dir, const std::string &header) {
    closeFile();
    const std::string fname = dir + "/" + _name + ".csv";
    _pFile = std::fopen(fname.c_str(), "w");
    if (!_pFile) return false;
    std::fputs(header.c_str(), _pFile);
    return true;
}

classNodeData::classNodeData(const InpNodeExtraction &x)
    : classDataBase(x.prefix + "_" + x.dataType + std::to_string(x.comp) + "_" + x.operation, x.operation),
      _nodes(x.nodes), _dataType(x.dataType), _comp(x.comp) {}

double classNodeData::getValue(const std::vector<double> &vals) const {
    if (_operation == "Mean") return std::accumulate(vals.begin(), vals.end(), 0.) / (double)vals.size();
    if (_operation == "Max") return *std::max_element(vals.begin(), vals.end());
    if (_operation == "Min") return *std::min_element(vals.begin(), vals.end());
    return std::accumulate(vals.begin(), vals.end(), 0.); // Sum
}

bool classNodeData::writeData(const std::string &dir, double time, const std::vector<double> &vals) {
    if (time <= _lastTime) return true;
    _lastTime = time;
    if (!_pFile) {
        std::string header = "Time,Value\n";
        if (_operation == "Rough") {
            header = "Time";
            for (int32_t n : _nodes) header += ",Node" + std::to_string(n + 1) + "Comp" + std::to_string(_comp);
            header += "\n";
        }
        if (!openFile(dir, header)) return false;
    }
    if (_operation == "Rough") {
        std::fprintf(_pFile, "%.16g", time);
        for (double v : vals) std::fprintf(_pFile, ",%.16g", v);
        std::fprintf(_pFile, "\n");
    } else {
        std::fprintf(_pFile, "%.16g,%.16g\n", time, getValue(vals));
    }
    std::fflush(_pFile);
    return true;
}

classElementData::classElementData(const InpElementExtraction &x)
    : classDataBase((x.all ? std::string("All") : x.prefix) + "_" + x.dataType + "_" + x.operation, x.operation),
      _dataType(x.dataType) {}

bool classElementData::writeData(const std::string &dir, double time, double val) {
    if (time <= _lastTime) return true;
    _lastTime = time;
    if (!_pFile && !openFile(dir, "Time,Value\n")) return false;
    std::fprintf(_pFile, "%.16g,%.16g\n", time, val);
    std::fflush(_pFile);
    return true;
}

bool classExtractData::init(const InpModel &model, std::string &err) {
    for (const InpNodeExtraction &x : model.nodeExtractions) {
        if (x.dataType != "Unknown" && x.dataType != "InternalForce" && x.dataType != "ExternalForce") {
            err = "node extraction type " + x.dataType + " is not defined (Unknown, InternalForce, ExternalForce)";
            return false;
        }
        if (x.operation != "Mean" && x.operation != "Max" && x.operation != "Min" && x.operation != "Sum" &&
            x.operation != "Rough") {
            err = "operation " + x.operation + " has not been defined";
            return false;
        }
        if (x.comp < 0 || x.comp > 2) {
            err = "component " + std::to_string(x.comp) + " does not exist";
            return false;
        }
        std::unique_ptr<classNodeData> d(new classNodeData(x));
        if (!_node.count(d->getName())) _node[d->getName()] = std::move(d);
    }
    for (const InpElementExtraction &x : model.elementExtractions) {
        if (x.dataType != "DEFO_ENERGY" && x.dataType != "KIN_ENERGY" && x.dataType != "VOLUME" &&
            x.dataType != "EXTERNAL_ENERGY") {
            err = "element extraction type " + x.dataType +
                  " is not on this path (DEFO_ENERGY, KIN_ENERGY, VOLUME, EXTERNAL_ENERGY are)";
            return false;
        }
        if (x.operation != "Sum") {
            err = "element extraction operation " + x.operation + " is not on this path (Sum is)";
            return false;
        }
        if (!x.all) {
            err = "element extraction over an element subset is not on this path (All is)";
            return false;
        }
        std::unique_ptr<classElementData> d(new classElementData(x));
        if (!_element.count(d->getName())) _element[d->getName()] = std::move(d);
    }
    return true;
}

void classExtractData::finalise() {
    for (auto &kv : _node) kv.second->closeFile();
    for (auto &kv : _element) kv.second->closeFile();
}

bool classPrintForces::due(double time, double dt, double t0) const {
    double period = _f.period;
    if (period < dt) period = dt;
    return time >= _iterationsCompleted * period + t0;
}

bool classPrintForces::write(const std::string &dir, double t0, double tf, double time,
                             const std::vector<double> &forces) {
    const std::string name = dir + "/externalForces" + std::to_string(t0) + "-" + std::to_string(tf) + ".csv";
    FILE *f = std::fopen(name.c_str(), _firstTime ? "w" : "a");
    if (!f) return false;
    if (_firstTime) std::fputs("Node,Direction,Force,Time,\n", f);
    for (size_t i = 0; i < _f.nodes.size(); i++) {
        const std::string row = std::to_string(_f.nodes[i] + 1) + "," + std::to_string(_f.dirs[i]) + "," +
                                std::to_string(forces[i]) + "," + std::to_string(time) + ",\n";
        std::fputs(row.c_str(), f);
    }
    std::fclose(f);
    _firstTime = false;
    _iterationsCompleted++;
    return true;
}

} // namespace mps
