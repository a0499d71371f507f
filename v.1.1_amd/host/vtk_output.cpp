// vtk_output.cpp -- VTK series of the explicit path (vtk_output.h).
#include "vtk_output.h"

#include <algorithm>
#include <charconv>
#include <cstdio>
#include <fstream>
#include <thread>

namespace mps {
namespace {

// Buffered text writer; numbers as `stream << double` with the default precision, i.e. printf's %.6g, which
// std::to_chars in general format with precision 6 reproduces character for character (and faster).
struct Out {
    std::string s;
    void str(const char *t) { s += t; }
    void str(const std::string &t) { s += t; }
    void num(double x) {
        char b[32];
        const auto r = std::to_chars(b, b + sizeof b, x, std::chars_format::general, 6);
        s.append(b, r.ptr);
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

// Items [0, n) formatted by item(o, i) on up to 16 threads, each into its own buffer, joined in order into out:
// the text is the same as one pass over the items would give.
template <typename F>
void formatItems(std::string &out, size_t n, size_t bytesPerItem, F item) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t nt = n < 20000 ? 1 : hw;
    std::vector<Out> part(nt);
    auto work = [&](size_t t) {
        const size_t b = n * t / nt, e = n * (t + 1) / nt;
        part[t].s.reserve((e - b) * bytesPerItem);
        for (size_t i = b; i < e; i++) item(part[t], i);
    };
    std::vector<std::thread> th;
    for (size_t t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    for (auto &p : part) out += p.s;
}

const std::string &VtkSeriesWriter::geometry(const InpModel &m) {
    if (_geomModel == &m && !_geom.empty()) return _geom;
    const size_t N = (size_t)m.nNodes, Ne = (size_t)m.nElem, nen = (size_t)m.nen;
    Out o;
    o.str("# vtk DataFile Version 1.0\n3D Unstructured Grid of ");
    o.str(m.elemType);
    o.str(" elements\nASCII\n\nDATASET UNSTRUCTURED_GRID\nPOINTS ");
    o.integer((long long)N);
    o.str(" float\n");
    formatItems(o.s, N, 40, [&](Out &q, size_t n) {
        for (int d = 0; d < 3; d++) {
            q.num(m.X[d * N + n]);
            q.str(" ");
        }
        q.nl();
    });
    o.nl();
    o.str("CELLS ");
    o.integer((long long)Ne);
    o.str(" ");
    o.integer((long long)(Ne * (nen + 1)));
    o.nl();
    formatItems(o.s, Ne, 8 * (nen + 1), [&](Out &q, size_t e) {
        q.integer((long long)nen);
        q.str(" ");
        for (size_t k = 0; k < nen; k++) {
            q.integer(m.conn[e * nen + k]);
            q.str(" ");
        }
        q.nl();
    });
    o.nl();
    o.str("CELL_TYPES ");
    o.integer((long long)Ne);
    o.nl();
    const char *ctype = nen == 8 ? "12\n" : "10\n";  // VTK_HEXAHEDRON, VTK_TETRA
    for (size_t e = 0; e < Ne; e++) o.str(ctype);
    o.nl();
    _geom.swap(o.s);
    _geomModel = &m;
    return _geom;
}

bool VtkSeriesWriter::write(const InpModel &m, const double *u, const double *v, const double *a,
                            const double *eout, const std::vector<int> &outVars, double timeRun, std::string &err) {
    const size_t N = (size_t)m.nNodes, Ne = (size_t)m.nElem;
    const std::string fname =
        _dir + "/MuPhiSimOutputs_" + std::to_string(_rank) + "_." + std::to_string(_cont) + ".vtk";
    _cont++;
    if (!updateSeries(timeRun, err)) return false;
    const std::string &geom = geometry(m);  // the same for every output of the run
    Out o;
    o.s.reserve(N * 120 + Ne * 300 + 1024);
    o.str("POINT_DATA ");
    o.integer((long long)N);
    o.nl();
    const char *vecs[3] = {"VECTORS Displacement float \n", "VECTORS Velocity float \n",
                           "VECTORS Acceleration float \n"};
    const double *fields[3] = {u, v, a};
    for (int f = 0; f < 3; f++) {
        o.str(vecs[f]);
        const double *x = fields[f];
        formatItems(o.s, N, 40, [&](Out &q, size_t n) {
            for (int d = 0; d < 3; d++) {
                q.num(x[d * N + n]);
                q.str(" ");
            }
            q.nl();
        });
        o.nl();
    }
    o.str("CELL_DATA ");
    o.integer((long long)Ne);
    o.nl();
    const char *tens[2] = {"TENSORS Cauchy float \n", "TENSORS E float \n"};
    for (int t = 0; t < 2; t++) {
        o.str(tens[t]);
        formatItems(o.s, Ne, 130, [&](Out &q, size_t e) {
            for (int i = 0; i < 3; i++) {
                for (int j = 0; j < 3; j++) {
                    q.num(eout[e * 21 + t * 9 + i * 3 + j]);
                    q.str(" ");
                }
                q.nl();
            }
            q.nl();
        });
    }
    const char *scal[3] = {"SCALARS VMS float \n", "SCALARS VolumetricStrain float \n",
                           "SCALARS EquivalentStrain float \n"};
    for (int k = 0; k < 3; k++) {
        if (k > 0) o.nl();
        o.str(scal[k]);
        o.str("LOOKUP_TABLE default\n");
        formatItems(o.s, Ne, 14, [&](Out &q, size_t e) {
            q.num(eout[e * 21 + 18 + k]);
            q.nl();
        });
    }
    o.nl();
    o.str("SCALARS Pressure float \nLOOKUP_TABLE default\n");
    formatItems(o.s, Ne, 14, [&](Out &q, size_t e) {
        const double *c = &eout[e * 21];
        q.num((c[0] + c[4] + c[8]) / 3);
        q.nl();
    });
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
    const bool ok = std::fwrite(geom.data(), 1, geom.size(), fp) == geom.size() &&
                    std::fwrite(o.s.data(), 1, o.s.size(), fp) == o.s.size();
    std::fclose(fp);
    if (!ok) err = "short write " + fname;
    return ok;
}

} // namespace mps
