// vtk_output.cpp -- VTK series of the explicit path (vtk_output.h).
#include "vtk_output.h"

#include <cstdio>
#include <fstream>

namespace mps {
namespace {

// Buffered text writer; numbers as `stream << double` with the default precision (%.6g).
struct Out {
    std::string s;
    void str(const char *t) { s += t; }
    void str(const std::string &t) { s += t; }
    void num(double x) {
        char b[32];
        std::snprintf(b, sizeof b, "%.6g", x);
        s += b;
    }
    void integer(long long x) { s += std::to_string(x); }
    void nl() { s += '\n'; }
};

} // namespace

// allOutputs_<rank>.vtk.series: output.cpp:323-372.
bool VtkSeriesWriter::updateSeries(double timeRun, std::string &err) {
    const std::string name = _dir + "/allOutputs_" + std::to_string(_rank) + ".vtk.series";
    Out entry;
    entry.str("{ \"name\" : \"MuPhiSimOutputs_" + std::to_string(_rank) + "_." + std::to_string(_cont - 1) +
              ".vtk\", \"time\" :");
    entry.num(timeRun);
    entry.str("}");
    std::vector<std::string> lines;
    {
        std::ifstream in(name);
        if (in.is_open() && _cont != 1) {
            std::string l;
            while (std::getline(in, l)) lines.push_back(l);
        }
    }
    std::ofstream out(name);
    if (!out) {
        err = "cannot write " + name;
        return false;
    }
    if (lines.empty()) {  // new series (no file, or the first output of the run)
        out << "{\n\"file-series-version\" : \"1.0\",\n\"files\" : [\n" << entry.s << "\n]\n}\n";
        return true;
    }
    if (lines.size() < 5) {
        err = "wrong output file " + name;
        return false;
    }
    for (size_t i = 0; i + 2 < lines.size(); i++) out << lines[i] << (i + 3 == lines.size() ? ",\n" : "\n");
    out << entry.s << "\n";
    for (size_t i = lines.size() - 2; i < lines.size(); i++) out << lines[i] << "\n";
    return true;
}

bool VtkSeriesWriter::write(const InpModel &m, const std::vector<double> &u, const std::vector<double> &v,
                            const std::vector<double> &a, const std::vector<double> &eout,
                            const std::vector<int> &outVars, double timeRun, std::string &err) {
    const size_t N = (size_t)m.nNodes, Ne = (size_t)m.nElem, nen = (size_t)m.nen;
    const std::string fname =
        _dir + "/MuPhiSimOutputs_" + std::to_string(_rank) + "_." + std::to_string(_cont) + ".vtk";
    _cont++;
    if (!updateSeries(timeRun, err)) return false;
    Out o;
    o.s.reserve(N * 160 + Ne * 400 + 1024);
    o.str("# vtk DataFile Version 1.0\n3D Unstructured Grid of ");
    o.str(m.elemType);
    o.str(" elements\nASCII\n\nDATASET UNSTRUCTURED_GRID\nPOINTS ");
    o.integer((long long)N);
    o.str(" float\n");
    for (size_t n = 0; n < N; n++) {
        for (int d = 0; d < 3; d++) {
            o.num(m.X[d * N + n]);
            o.str(" ");
        }
        o.nl();
    }
    o.nl();
    o.str("CELLS ");
    o.integer((long long)Ne);
    o.str(" ");
    o.integer((long long)(Ne * (nen + 1)));
    o.nl();
    for (size_t e = 0; e < Ne; e++) {
        o.integer((long long)nen);
        o.str(" ");
        for (size_t k = 0; k < nen; k++) {
            o.integer(m.conn[e * nen + k]);
            o.str(" ");
        }
        o.nl();
    }
    o.nl();
    o.str("CELL_TYPES ");
    o.integer((long long)Ne);
    o.nl();
    const char *ctype = nen == 8 ? "12\n" : "10\n";  // VTK_HEXAHEDRON, VTK_TETRA
    for (size_t e = 0; e < Ne; e++) o.str(ctype);
    o.nl();
    o.str("POINT_DATA ");
    o.integer((long long)N);
    o.nl();
    const char *vecs[3] = {"VECTORS Displacement float \n", "VECTORS Velocity float \n",
                           "VECTORS Acceleration float \n"};
    const std::vector<double> *fields[3] = {&u, &v, &a};
    for (int f = 0; f < 3; f++) {
        o.str(vecs[f]);
        for (size_t n = 0; n < N; n++) {
            for (int d = 0; d < 3; d++) {
                o.num((*fields[f])[d * N + n]);
                o.str(" ");
            }
            o.nl();
        }
        o.nl();
    }
    o.str("CELL_DATA ");
    o.integer((long long)Ne);
    o.nl();
    const char *tens[2] = {"TENSORS Cauchy float \n", "TENSORS E float \n"};
    for (int t = 0; t < 2; t++) {
        o.str(tens[t]);
        for (size_t e = 0; e < Ne; e++) {
            for (int i = 0; i < 3; i++) {
                for (int j = 0; j < 3; j++) {
                    o.num(eout[e * 21 + t * 9 + i * 3 + j]);
                    o.str(" ");
                }
                o.nl();
            }
            o.nl();
        }
    }
    const char *scal[3] = {"SCALARS VMS float \n", "SCALARS VolumetricStrain float \n",
                           "SCALARS EquivalentStrain float \n"};
    for (int k = 0; k < 3; k++) {
        if (k > 0) o.nl();
        o.str(scal[k]);
        o.str("LOOKUP_TABLE default\n");
        for (size_t e = 0; e < Ne; e++) {
            o.num(eout[e * 21 + 18 + k]);
            o.nl();
        }
    }
    o.nl();
    o.str("SCALARS Pressure float \nLOOKUP_TABLE default\n");
    for (size_t e = 0; e < Ne; e++) {
        const double *c = &eout[e * 21];
        o.num((c[0] + c[4] + c[8]) / 3);
        o.nl();
    }
    o.nl();
    for (int k : outVars) {
        o.str("SCALARS IntVars");
        o.integer(k + 1);
        o.str(" float \nLOOKUP_TABLE default\n");
        for (size_t e = 0; e < Ne; e++) o.str("0\n");
        o.nl();
    }
    FILE *fp = std::fopen(fname.c_str(), "w");
    if (!fp) {
        err = "cannot write " + fname;
        return false;
    }
    const bool ok = std::fwrite(o.s.data(), 1, o.s.size(), fp) == o.s.size();
    std::fclose(fp);
    if (!ok) err = "short write " + fname;
    return ok;
}

} // namespace mps
