// inp_reader.cpp -- .inp reader of the explicit solid path (see inp_reader.h
// for the keyword list and the reference functions each part follows).
#include "inp_reader.h"

#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <set>
#include <sstream>
#include <string_view>
#include <thread>

#include "mps_gpu.h"

namespace mps {
namespace {

bool starts(std::string_view line, const char *kw) {
    const size_t n = std::strlen(kw);
    return line.size() >= n && std::memcmp(line.data(), kw, n) == 0;
}

// Text after column `col` (the reference's line.compare(col, n, "...") tests).
std::string tail(std::string_view line, size_t col) { return line.size() > col ? std::string(line.substr(col)) : std::string(); }

std::vector<std::string> fields(std::string_view line) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : line) {
        if (c == ',') {
            out.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(c);
        }
    }
    out.push_back(cur);
    return out;
}

bool isBlank(std::string_view s) { return s.find_first_not_of(" \t") == std::string_view::npos; }

// Numeric fields of the node and element blocks, parsed in place (the blocks of a 1M-element deck hold some
// 13 million numbers): atoi/atof semantics -- leading blanks and a '+' skipped, the number ends at the first
// character that cannot continue it -- with std::from_chars, which rounds as strtod does.  *p advances past
// the number; false when no number starts there.
bool parseInt(const char *&p, const char *end, long long &v) {
    while (p < end && (*p == ' ' || *p == '\t')) p++;
    if (p < end && *p == '+') p++;
    const auto r = std::from_chars(p, end, v);
    if (r.ec != std::errc()) return false;
    p = r.ptr;
    return true;
}
bool parseDouble(const char *&p, const char *end, double &v) {
    while (p < end && (*p == ' ' || *p == '\t')) p++;
    if (p < end && *p == '+') p++;
    const auto r = std::from_chars(p, end, v);
    if (r.ec != std::errc()) return false;
    p = r.ptr;
    return true;
}
// The field that starts at p ends at the next comma (or the line's end): skip what is left of it.
bool nextField(const char *&p, const char *end) {
    const void *c = std::memchr(p, ',', (size_t)(end - p));
    if (!c) return false;
    p = static_cast<const char *>(c) + 1;
    return true;
}
// atoi/atof semantics on one field, with a check that a number was present.
bool toInt(std::string_view s, long long &v) {
    const char *p = s.data();
    return parseInt(p, s.data() + s.size(), v);
}
bool toDouble(std::string_view s, double &v) {
    const char *p = s.data();
    return parseDouble(p, s.data() + s.size(), v);
}
int countCommas(std::string_view s) {
    int n = 0;
    for (char c : s) n += c == ',';
    return n;
}

struct Reader {
    std::string buf;                    // the file
    std::vector<std::string_view> L;   // its lines, trailing blanks removed
    std::string err;
    std::string where(size_t i) const {
        std::ostringstream o;
        o << "line " << (i + 1) << " \"" << L[i] << "\"";
        return o.str();
    }
    bool fail(size_t i, const std::string &msg) {
        err = msg + " (" + where(i) + ")";
        return false;
    }
};

// Data lines [first, last) of a large block parsed in chunks on up to 16 threads: parse(i, out) appends line i's
// values to out (a blank line appends nothing) or returns an error message.  The chunks' values are joined in
// line order; on errors the first failing line in file order is reported.
template <typename T, typename F>
bool parseBlock(Reader &R, size_t first, size_t last, std::vector<T> &vals, F parse) {
    const size_t n = last - first;
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const size_t nt = n < 65536 ? 1 : hw;
    std::vector<std::vector<T>> part(nt);
    std::vector<size_t> badLine(nt, SIZE_MAX);
    std::vector<std::string> badMsg(nt);
    auto work = [&](size_t t) {
        const size_t b = first + n * t / nt, e = first + n * (t + 1) / nt;
        part[t].reserve(e - b);
        for (size_t i = b; i < e; i++) {
            const char *msg = parse(i, part[t]);
            if (msg) {
                badLine[t] = i;
                badMsg[t] = msg;
                return;
            }
        }
    };
    std::vector<std::thread> th;
    for (size_t t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    for (size_t t = 0; t < nt; t++)
        if (badLine[t] != SIZE_MAX) return R.fail(badLine[t], badMsg[t]);
    size_t total = 0;
    for (auto &p : part) total += p.size();
    vals.clear();
    vals.reserve(total);
    for (auto &p : part) vals.insert(vals.end(), p.begin(), p.end());
    return true;
}

// nodesManagement (inputFile.cpp:20-111): first block after a line starting
// "*Node", data until the next keyword; the comma count sets ndim.
bool readNodes(Reader &R, InpModel &m) {
    size_t i = 0;
    while (i < R.L.size() && !starts(R.L[i], "*Node")) i++;
    if (i == R.L.size()) {
        R.err = "no *Node block";
        return false;
    }
    const size_t first = i + 1;
    size_t last = first;
    while (last < R.L.size() && !starts(R.L[last], "*")) last++;
    struct NodeRow {
        long long id;
        double x[3];
    };
    std::vector<NodeRow> rows;
    const bool ok = parseBlock(R, first, last, rows, [&](size_t i, std::vector<NodeRow> &out) -> const char * {
        std::string_view l = R.L[i];
        if (isBlank(l)) return nullptr;
        if (countCommas(l) != 3) return "this path is 3-D: a node line needs id, x, y, z";
        const char *p = l.data(), *e = l.data() + l.size();
        NodeRow r;
        if (!parseInt(p, e, r.id) || !nextField(p, e) || !parseDouble(p, e, r.x[0]) || !nextField(p, e) ||
            !parseDouble(p, e, r.x[1]) || !nextField(p, e) || !parseDouble(p, e, r.x[2]))
            return "malformed node line";
        out.push_back(r);
        return nullptr;
    });
    if (!ok) return false;
    std::vector<long long> ids(rows.size());
    std::vector<double> xyz(3 * rows.size());
    for (size_t r = 0; r < rows.size(); r++) {
        ids[r] = rows[r].id;
        for (int d = 0; d < 3; d++) xyz[3 * r + d] = rows[r].x[d];
    }
    const long long N = (long long)ids.size();
    if (N == 0) {
        R.err = "empty *Node block";
        return false;
    }
    m.ndim = 3;
    m.nNodes = (int32_t)N;
    m.X.assign(3 * (size_t)N, 0.0);
    std::vector<char> seen((size_t)N, 0);
    for (size_t r = 0; r < ids.size(); r++) {
        if (ids[r] < 1 || ids[r] > N)
            return (R.err = "node id " + std::to_string(ids[r]) + " outside 1.." + std::to_string(N) +
                            " (node ids must number the nodes 1..N)",
                    false);
        size_t k = (size_t)(ids[r] - 1);
        if (seen[k]) return (R.err = "duplicate node id " + std::to_string(ids[r]), false);
        seen[k] = 1;
        for (int d = 0; d < 3; d++) m.X[(size_t)d * N + k] = xyz[3 * r + d];
    }
    return true;
}

// elementsManagement (inputFile.cpp:121-330): first "*Element" block; the
// type is the text after column 15 ("*Element, type=").
bool readElements(Reader &R, InpModel &m) {
    size_t i = 0;
    while (i < R.L.size() && !starts(R.L[i], "*Element")) i++;
    if (i == R.L.size()) {
        R.err = "There is not a list of elements in the input file";
        return false;
    }
    std::string type = tail(R.L[i], 15);
    if (type == "C3D8")
        m.nen = 8;
    else if (type == "C3D4")
        m.nen = 4;
    else
        return R.fail(i, "element type \"" + type + "\" is not on this path (C3D8 and C3D4 are)");
    m.elemType = type;
    const size_t first = i + 1;
    size_t last = first;
    while (last < R.L.size() && !starts(R.L[last], "*")) last++;
    struct ElemRow {
        long long id;
        int32_t n[8];
    };
    const int nen = m.nen;
    const long long nNodes = m.nNodes;
    const std::string countMsg = "element line needs id and " + std::to_string(nen) + " nodes";
    std::vector<ElemRow> rows;
    const bool ok = parseBlock(R, first, last, rows, [&](size_t i, std::vector<ElemRow> &out) -> const char * {
        std::string_view l = R.L[i];
        if (isBlank(l)) return nullptr;
        if (countCommas(l) != nen) return countMsg.c_str();
        const char *p = l.data(), *e = l.data() + l.size();
        ElemRow r;
        if (!parseInt(p, e, r.id)) return "malformed element id";
        for (int a = 0; a < nen; a++) {
            long long n;
            if (!nextField(p, e) || !parseInt(p, e, n)) return "malformed node id";
            if (n < 1 || n > nNodes) return "element node id outside 1..N";
            r.n[a] = (int32_t)(n - 1);
        }
        out.push_back(r);
        return nullptr;
    });
    if (!ok) return false;
    i = last;
    std::vector<long long> ids(rows.size());
    std::vector<int32_t> nodes((size_t)nen * rows.size());
    for (size_t r = 0; r < rows.size(); r++) {
        ids[r] = rows[r].id;
        for (int a = 0; a < nen; a++) nodes[r * nen + a] = rows[r].n[a];
    }
    for (size_t j = i; j < R.L.size(); j++)
        if (starts(R.L[j], "*Element"))
            m.warnings.push_back("only the first *Element block is read, as the reference does (" + R.where(j) + ")");
    const long long E = (long long)ids.size();
    if (E == 0) return (R.err = "empty *Element block", false);
    m.nElem = (int32_t)E;
    m.conn.assign((size_t)E * m.nen, 0);
    std::vector<char> seen((size_t)E, 0);
    for (size_t r = 0; r < ids.size(); r++) {
        if (ids[r] < 1 || ids[r] > E)
            return (R.err = "element id " + std::to_string(ids[r]) + " outside 1.." + std::to_string(E) +
                            " (element ids must number the elements 1..N_e)",
                    false);
        size_t k = (size_t)(ids[r] - 1);
        if (seen[k]) return (R.err = "duplicate element id " + std::to_string(ids[r]), false);
        seen[k] = 1;
        for (int a = 0; a < m.nen; a++) m.conn[k * m.nen + a] = nodes[r * m.nen + a];
    }
    return true;
}

// *FEM Nodes / *FEM Elements (inputFile.cpp:331-378): the path is pure FEM,
// so both must be present and read "All".
bool readFemSets(Reader &R) {
    const char *kws[2] = {"*FEM Nodes", "*FEM Elements"};
    for (const char *kw : kws) {
        size_t i = 0;
        while (i < R.L.size() && !starts(R.L[i], kw)) i++;
        if (i == R.L.size()) return (R.err = std::string("no ") + kw + " block (meshless domains are not on this path)", false);
        if (i + 1 >= R.L.size() || !starts(R.L[i + 1], "All"))
            return R.fail(i, std::string(kw) + " must be \"All\" on this path");
    }
    return true;
}

// materialManagement (inputFile.cpp:838-1018, 2010-2054, 2118-2150).
bool readMaterials(Reader &R, InpModel &m) {
    bool inMat = false, listing = false, haveLaw = false;
    double rho = 0; // function-scope in the reference: carries over between *MATERIAL blocks
    InpMaterial cur;
    auto flush = [&]() {
        if (inMat && haveLaw) m.materials.push_back(cur);
    };
    for (size_t i = 0; i < R.L.size(); i++) {
        std::string_view line = R.L[i];
        if (starts(line, "*MATERIAL")) {
            flush();
            cur = InpMaterial();
            cur.name = tail(line, 16);
            inMat = true;
            listing = false;
            haveLaw = false;
            continue;
        }
        if (!inMat) continue;
        if (starts(line, "*Density")) {
            if (i + 1 >= R.L.size() || !toDouble(R.L[i + 1], rho)) return R.fail(i, "*Density needs a value on the next line");
            i++;
        } else if (starts(line, "*HyperElastic") || starts(line, "*Linear-Elastic")) {
            // "*HyperElastic, <model>" (model = text after column 15, inputFile.cpp:930-1024) or
            // "*Linear-Elastic" (inputFile.cpp:886-928); next line "E, nu[, stress state]"
            std::string model = starts(line, "*Linear-Elastic") ? std::string("Linear-Elastic") : tail(line, 15);
            if (model != "Neo-Hookean" && model != "3D-Compressible-Neo-Hookean" && model != "St-Venant-Kirchhoff" &&
                model != "Linear-Elastic")
                return R.fail(i, "constitutive model \"" + model +
                                     "\" has no device descriptor (Neo-Hookean, 3D-Compressible-Neo-Hookean, "
                                     "St-Venant-Kirchhoff and Linear-Elastic do)");
            if (i + 1 >= R.L.size()) return R.fail(i, "*HyperElastic needs \"E, nu\" on the next line");
            std::vector<std::string> f = fields(R.L[i + 1]);
            if (f.size() < 2 || !toDouble(f[0], cur.young) || !toDouble(f[1], cur.nu))
                return R.fail(i + 1, "*HyperElastic needs \"E, nu\"");
            if (rho <= 0) return R.fail(i, "density must be positive and given (*Density) before the law keyword");
            cur.nameCons = model;
            cur.rho = rho;
            haveLaw = true;
            i++;
        } else if (starts(line, "*List of Elements")) {
            if (!haveLaw) return R.fail(i, "The constitutive model has not been properly defined");
            if (i + 1 >= R.L.size()) return R.fail(i, "*List of Elements needs a list");
            std::string_view t = R.L[i + 1];
            if (starts(t, "All")) {
                cur.all = true;
            } else {
                listing = true;
                long long id;
                if (!toInt(t, id) || id < 1 || id > m.nElem) return R.fail(i + 1, "element id outside 1..N_e");
                cur.elems.push_back((int32_t)(id - 1));
            }
            i++;
        } else if (starts(line, "*END") || starts(line, "*End")) {
            break;
        } else if (starts(line, "*")) {
            return R.fail(i, "keyword inside a *MATERIAL block that is not on this path");
        } else if (listing) {
            if (isBlank(line)) continue;
            long long id;
            if (!toInt(line, id) || id < 1 || id > m.nElem) return R.fail(i, "element id outside 1..N_e");
            cur.elems.push_back((int32_t)(id - 1));
        }
    }
    flush();
    if (m.materials.empty()) return (R.err = "no *MATERIAL with a constitutive law", false);
    return true;
}

// dispBCManagement (inputFile.cpp:2960-3075): node,dir,U lines until the
// next keyword; node 1-based, dir 0-based.
bool readDirichlet(Reader &R, size_t &i, const InpModel &m, InpSolver &s) {
    for (i++; i < R.L.size() && !starts(R.L[i], "*"); i++) {
        if (isBlank(R.L[i])) continue;
        std::vector<std::string> f = fields(R.L[i]);
        long long node, dir;
        double U;
        if (f.size() != 3 || !toInt(f[0], node) || !toInt(f[1], dir) || !toDouble(f[2], U))
            return R.fail(i, "boundary line needs node, direction, value");
        if (node < 1 || node > m.nNodes) return R.fail(i, "boundary node outside 1..N");
        if (dir < 0 || dir >= 3) return R.fail(i, "The direction has not been recognized (0-x, 1-y, 2-z)");
        s.dirichlet.push_back({(int32_t)(node - 1), (int32_t)dir, U});
    }
    i--; // the caller's loop advances to the keyword line
    return true;
}

bool nextValue(Reader &R, size_t i, double &v) {
    return i + 1 < R.L.size() && toDouble(R.L[i + 1], v);
}

// solverManagement (inputFile.cpp:2253-2830).  Scale factor, time window,
// *Outputs and the instantaneous-BC flag are function-scope variables in the
// reference and carry over from one *SOLVER block to the next.
bool readSolvers(Reader &R, InpModel &m) {
    double scale = 1.0, t0 = 0.0, tf = 0.0;
    int numFiles = -1;
    std::vector<int> outVars;
    bool instantaneous = false;
    for (size_t i = 0; i < R.L.size(); i++) {
        if (!starts(R.L[i], "*SOLVER")) continue;
        std::string type = tail(R.L[i], 9);
        if (type != "EXPLICIT")
            return R.fail(i, "solver \"" + type + "\" is not on this path (EXPLICIT is)");
        InpSolver s;
        bool timeSeen = false, velocity = false, displacement = false, closed = false;
        for (i++; i < R.L.size(); i++) {
            std::string_view l = R.L[i];
            double v;
            if (starts(l, "*END SOLVER")) {
                closed = true;
                break;
            } else if (starts(l, "*Scale Factor")) {
                if (!nextValue(R, i, v)) return R.fail(i, "*Scale Factor needs a value");
                scale = v;
                i++;
            } else if (starts(l, "*Number Of Steps")) {
                if (!nextValue(R, i, v)) return R.fail(i, "*Number Of Steps needs a value");
                s.numStepsIgnored = (int)v;
                m.warnings.push_back("*Number Of Steps is ignored by the EXPLICIT solver (" + R.where(i) + ")");
                i++;
            } else if (starts(l, "*Time")) {
                if (i + 1 >= R.L.size()) return R.fail(i, "*Time needs \"t0, tf\"");
                std::vector<std::string> f = fields(R.L[i + 1]);
                if (f.size() < 2 || !toDouble(f[0], t0) || !toDouble(f[1], tf)) return R.fail(i + 1, "*Time needs \"t0, tf\"");
                timeSeen = true;
                i++;
            } else if (starts(l, "*BOUNDARY, TYPE=DISPLACEMENT RAMP") || starts(l, "*BOUNDARY, TYPE=DISPLACEMENT INST")) {
                if (starts(l, "*BOUNDARY, TYPE=DISPLACEMENT INST")) instantaneous = true;
                displacement = true;
                if (!timeSeen)
                    return R.fail(i, "*Time must precede *BOUNDARY inside *SOLVER (the ramp duration is taken at parse time)");
                if (!readDirichlet(R, i, m, s)) return false;
            } else if (starts(l, "*BOUNDARY") &&
                       (tail(l, 16) == "FORCE" || tail(l, 16) == "PRESSURE RAMP" || tail(l, 16) == "PRESSURE INST" ||
                        tail(l, 16) == "TRACTION RAMP" || tail(l, 16) == "TRACTION INST")) {
                // a Neumann block: read with the whole file by readLoads, for every solver
                for (i++; i + 1 < R.L.size() && !starts(R.L[i + 1], "*"); i++) {
                }
            } else if (starts(l, "*BOUNDARY, TYPE=VELOCITY")) {
                velocity = true;
                if (!readDirichlet(R, i, m, s)) return false;
            } else if (starts(l, "*Outputs")) {
                // outManagement (inputFile.cpp:4769-4829): number of files, then optionally the
                // internal variables to print: "All, n" (0..n-1), one index k (printed as is), or
                // "max, i1, i2, ..." (1-based, each <= max)
                if (!nextValue(R, i, v)) return R.fail(i, "*Outputs needs the number of output files");
                numFiles = (int)v;
                if (numFiles < 1) return R.fail(i, "*Outputs needs at least one output file");
                outVars.clear();
                if (i + 2 < R.L.size() && !starts(R.L[i + 2], "*")) {
                    std::string_view t = R.L[i + 2];
                    std::vector<std::string> f = fields(t);
                    long long x;
                    if (starts(t, "All")) {
                        // atoi of the text from column 4: a word there ("All,max") asks for no variable
                        if (!toInt(tail(t, 4), x)) x = 0;
                        for (long long k = 0; k < x; k++) outVars.push_back((int)k);
                    } else if (f.size() == 1) {
                        if (!toInt(t, x)) return R.fail(i + 2, "*Outputs: malformed internal variable");
                        outVars.push_back((int)x);
                    } else {
                        long long mx;
                        if (!toInt(f[0], mx)) return R.fail(i + 2, "*Outputs: malformed internal variable list");
                        for (size_t k = 1; k < f.size(); k++) {
                            if (!toInt(f[k], x)) return R.fail(i + 2, "*Outputs: malformed internal variable list");
                            if (x - 1 > mx) return R.fail(i + 2, "*Outputs: internal variable beyond the maximum");
                            outVars.push_back((int)(x - 1));
                        }
                    }
                }
            } else if (starts(l, "*MPS Fixed Steps")) {
                if (!nextValue(R, i, v) || v < 1) return R.fail(i, "*MPS Fixed Steps needs a positive count");
                s.fixedSteps = (int64_t)v;
                i++;
            } else if (starts(l, "*MPS Velocity Ramp Steps")) {
                if (!nextValue(R, i, v) || v < 0) return R.fail(i, "*MPS Velocity Ramp Steps needs a count >= 0");
                s.velRampSteps = (int)v;
                i++;
            } else if (starts(l, "*MPS Precision")) {
                std::string p = i + 1 < R.L.size() ? std::string(R.L[i + 1]) : std::string();
                if (p == "FP64")
                    s.precision = MPS_PREC_FP64;
                else if (p == "FP32_ELEM")
                    s.precision = MPS_PREC_FP32_ELEM;
                else
                    return R.fail(i, "*MPS Precision is FP64 or FP32_ELEM");
                i++;
            } else if (starts(l, "*Bulk viscosity") || starts(l, "*Dev viscosity")) {
                // solverManagement (inputFile.cpp:2462-2510): defaults, then the next line's characters:
                // spaces skipped, '*' ends it (no parameters), ',' closes c1, the last character closes cl.
                // The reference consumes that line even when it is a keyword; here a keyword is left to be read.
                const bool bulk = starts(l, "*Bulk viscosity");
                s.visc = bulk ? MPS_VISC_BULK : MPS_VISC_DEV;
                s.viscC1 = bulk ? 0.06 : 1;
                s.viscCl = bulk ? 1.2 : 0.5;
                if (i + 1 < R.L.size() && !starts(R.L[i + 1], "*")) {
                    std::string_view t = R.L[++i];
                    std::string tok;
                    for (size_t c = 0; c < t.size(); c++) {
                        if (t[c] == '*') break;
                        if (t[c] == ' ') continue;
                        if (t[c] == ',') {
                            s.viscC1 = std::atof(tok.c_str());
                            tok.clear();
                        } else {
                            tok += t[c];
                            if (c == t.size() - 1) s.viscCl = std::atof(tok.c_str());
                        }
                    }
                }
            } else if (starts(l, "*Forces")) {
                // forceOutManagement (inputFile.cpp:4830-4887): the output frequency (period 1/f), then
                // "node, direction" lines up to the next keyword; a new *Forces replaces the block's earlier one
                // (inputFile.cpp:2537-2539).  The direction is the field after the first comma; printForcesOut
                // uses it as the 0-based direction of Forces[node + direction*N] and rejects one >= ndim
                // (output.cpp:119-126).  (The reference takes it up to a second comma it assumes: with one comma
                // per line, as in its own decks, it reads past its list of comma positions.)
                if (!nextValue(R, i, v) || !(v > 0)) return R.fail(i, "*Forces needs a positive output frequency");
                InpForces F;
                F.on = true;
                F.period = 1 / v;
                for (i += 2; i < R.L.size() && !starts(R.L[i], "*"); i++) {
                    if (isBlank(R.L[i])) continue;
                    std::vector<std::string> f = fields(R.L[i]);
                    long long node, dir;
                    if (f.size() < 2 || !toInt(f[0], node) || !toInt(f[1], dir))
                        return R.fail(i, "*Forces needs \"node, direction\" lines");
                    if (node < 1 || node > m.nNodes) return R.fail(i, "*Forces: node outside the mesh");
                    if (dir < 0 || dir >= m.ndim)
                        return R.fail(i, "*Forces: direction of force requested exceeds the simulation dimension");
                    F.nodes.push_back((int32_t)(node - 1));
                    F.dirs.push_back((int32_t)dir);
                }
                i--;
                s.forces = F;
            } else if (starts(l, "*ACTIVATION") ||
                       starts(l, "*FSI") || starts(l, "*ADAPTIVE TIME STEP") || starts(l, "*TIME STEPPING PLAN")) {
                return R.fail(i, "solver option not on this path");
            } else if (starts(l, "*")) {
                m.warnings.push_back("solver option ignored by the EXPLICIT solver (" + R.where(i) + ")");
            }
        }
        if (!closed) return (R.err = "*SOLVER without *END SOLVER", false);
        if (velocity && displacement)
            return (R.err = "TYPE=VELOCITY cannot be mixed with TYPE=DISPLACEMENT in one solver", false);
        s.scale = scale;
        s.t0 = t0;
        s.tf = tf;
        s.timeSet = timeSeen;
        s.numFiles = numFiles;
        s.outVars = outVars;
        s.bcMode = velocity ? MPS_BC_VELOCITY : (instantaneous ? MPS_BC_INST : MPS_BC_RAMP);
        if (s.numFiles < 0) return (R.err = "*Outputs is mandatory for the EXPLICIT solver", false);
        if (s.dirichlet.empty())
            return (R.err = "the EXPLICIT solver needs at least one mechanical Dirichlet DOF", false);
        std::set<int64_t> dofs;
        for (const InpDirichlet &d : s.dirichlet)
            if (!dofs.insert((int64_t)d.node + (int64_t)m.nNodes * d.dir).second)
                return (R.err = "duplicate Dirichlet DOF: node " + std::to_string(d.node + 1) + " direction " +
                                std::to_string(d.dir),
                        false);
        if (s.fixedSteps == 0 && !(s.tf > s.t0))
            return (R.err = "*Time needs tf > t0 (or *MPS Fixed Steps)", false);
        m.solvers.push_back(s);
    }
    if (m.solvers.empty()) return (R.err = "no *SOLVER, EXPLICIT block", false);
    return true;
}

// dataExtractionManagement (inputFile.cpp:2229-2243, 4950-5120).
bool readExtractions(Reader &R, InpModel &m) {
    for (size_t i = 0; i < R.L.size(); i++) {
        const bool node = starts(R.L[i], "*EXTRACTION, NODE");
        const bool elem = starts(R.L[i], "*EXTRACTION, ELEMENT");
        if (!node && !elem) continue;
        if (i + 1 >= R.L.size()) return R.fail(i, "extraction needs a definition line");
        std::vector<std::string> f = fields(R.L[i + 1]);
        if (node) {
            InpNodeExtraction x;
            long long comp;
            if (f.size() != 4 || !toInt(f[3], comp)) return R.fail(i + 1, "node extraction needs name,type,operation,comp");
            x.prefix = f[0];
            x.dataType = f[1];
            x.operation = f[2];
            x.comp = (int)comp;
            for (i += 2; i < R.L.size() && !starts(R.L[i], "*"); i++) {
                long long id;
                if (isBlank(R.L[i])) continue;
                if (!toInt(R.L[i], id) || id < 1 || id > m.nNodes) return R.fail(i, "node id outside 1..N");
                x.nodes.push_back((int32_t)(id - 1));
            }
            i--;
            if (!x.nodes.empty()) m.nodeExtractions.push_back(x);
        } else {
            InpElementExtraction x;
            if (f.size() != 3) return R.fail(i + 1, "element extraction needs name,type,operation");
            x.prefix = f[0];
            x.dataType = f[1];
            x.operation = f[2];
            for (i += 2; i < R.L.size() && !starts(R.L[i], "*"); i++) {
                long long id;
                if (isBlank(R.L[i])) continue;
                if (starts(R.L[i], "All")) {
                    x.all = true;
                    continue;
                }
                if (!toInt(R.L[i], id) || id < 1 || id > m.nElem) return R.fail(i, "element id outside 1..N_e");
                x.elems.push_back((int32_t)(id - 1));
            }
            i--;
            if (x.all || !x.elems.empty()) m.elementExtractions.push_back(x);
        }
    }
    return true;
}

// NeumannCommonManagement (inputFile.cpp:4703-4760): the whole file, in order.
//   FORCE (forceBCManagement, :2841-2941): a "t0, tf, Ff" line, then "node, Fx, Fy, Fz" lines up to the next
//     keyword, each entry F * Ff over the window [t0, tf].
//   PRESSURE RAMP / INST (pressureRampBCManagement :3580-3684, pressureInstBCManagement :3686-3790): a
//     "t0, tf, P, S" line, then one element id per line (its leading integer); face S of each element.
//   *GRAVITY: the value g on the next line.
// Traction, contact and flux loads are not on this path and are rejected.
bool readLoads(Reader &R, InpModel &m) {
    bool gravity = false;
    for (size_t i = 0; i < R.L.size(); i++) {
        std::string_view l = R.L[i];
        if (starts(l, "*BOUNDARY")) {
            const std::string type = tail(l, 16);
            if (type == "FORCE") {
                InpLoadBlock b;
                b.type = MPS_LOAD_FORCE;
                if (i + 1 >= R.L.size()) return R.fail(i, "FORCE needs a \"t0, tf, Ff\" line");
                std::vector<std::string> f = fields(R.L[i + 1]);
                double Ff;
                if (f.size() != 3 || !toDouble(f[0], b.t0) || !toDouble(f[1], b.tf) || !toDouble(f[2], Ff))
                    return R.fail(i + 1, "FORCE needs a \"t0, tf, Ff\" line");
                if (!(b.tf > b.t0)) return R.fail(i + 1, "FORCE window needs tf > t0 (the load grows as F/(tf-t0) * t)");
                for (i += 2; i < R.L.size() && !starts(R.L[i], "*"); i++) {
                    if (isBlank(R.L[i])) continue;
                    std::vector<std::string> g = fields(R.L[i]);
                    long long id;
                    double F[3];
                    if (g.size() != 4 || !toInt(g[0], id) || !toDouble(g[1], F[0]) || !toDouble(g[2], F[1]) ||
                        !toDouble(g[3], F[2]))
                        return R.fail(i, "FORCE entry needs \"node, Fx, Fy, Fz\"");
                    if (id < 1 || id > m.nNodes) return R.fail(i, "node id outside 1..N");
                    b.index.push_back((int32_t)(id - 1));
                    for (int d = 0; d < 3; d++) b.force.push_back(F[d] * Ff);
                }
                i--;
                m.loads.blocks.push_back(b);
            } else if (type == "PRESSURE RAMP" || type == "PRESSURE INST" || type == "TRACTION RAMP" ||
                       type == "TRACTION INST") {
                // pressure: "t0, tf, P, S"; traction (tractionRampBCManagement :3342-3460): "t0, tf, Tx, Ty, Tz, S"
                InpLoadBlock b;
                const bool trac = type[0] == 'T';
                b.type = type == "PRESSURE RAMP"   ? MPS_LOAD_PRESSURE_RAMP
                         : type == "PRESSURE INST" ? MPS_LOAD_PRESSURE_INST
                         : type == "TRACTION RAMP" ? MPS_LOAD_TRACTION_RAMP
                                                   : MPS_LOAD_TRACTION_INST;
                const char *form = trac ? "TRACTION needs a \"t0, tf, Tx, Ty, Tz, S\" line"
                                        : "PRESSURE needs a \"t0, tf, P, S\" line";
                if (i + 1 >= R.L.size()) return R.fail(i, form);
                std::vector<std::string> f = fields(R.L[i + 1]);
                double S;
                if (f.size() != (trac ? 6u : 4u) || !toDouble(f[0], b.t0) || !toDouble(f[1], b.tf) ||
                    !toDouble(f.back(), S))
                    return R.fail(i + 1, form);
                if (trac) {
                    b.force.resize(3);
                    for (int d = 0; d < 3; d++)
                        if (!toDouble(f[2 + d], b.force[d])) return R.fail(i + 1, form);
                } else if (!toDouble(f[2], b.value)) {
                    return R.fail(i + 1, form);
                }
                const int nf = m.nen == 8 ? 6 : 4;
                if ((int)S < 1 || (int)S > nf) return R.fail(i + 1, "face S outside 1.." + std::to_string(nf));
                const bool ramp = b.type == MPS_LOAD_PRESSURE_RAMP || b.type == MPS_LOAD_TRACTION_RAMP;
                if (ramp ? !(b.tf > b.t0) : !(b.tf >= b.t0))
                    return R.fail(i + 1, "window needs tf > t0 (tf >= t0 for INST)");
                for (i += 2; i < R.L.size() && !starts(R.L[i], "*"); i++) {
                    if (isBlank(R.L[i])) continue;
                    long long id;
                    if (!toInt(fields(R.L[i])[0], id) || id < 1 || id > m.nElem)
                        return R.fail(i, "surface load element id outside 1..N_e");
                    b.index.push_back((int32_t)(id - 1));
                    b.face.push_back((int32_t)S);
                }
                i--;
                m.loads.blocks.push_back(b);
            }
        } else if (starts(l, "*GRAVITY")) {
            InpLoadBlock b;
            b.type = MPS_LOAD_GRAVITY;
            if (i + 1 >= R.L.size() || !toDouble(R.L[i + 1], b.value)) return R.fail(i, "*GRAVITY needs the value g");
            if (gravity) return R.fail(i, "a second *GRAVITY: one gravity load per model");
            gravity = true;
            m.loads.blocks.push_back(b);
            i++;
        } else if (starts(l, "*HERTZIAN CONTACT") || starts(l, "*STOCHASTIC HERTZIAN CONTACT") ||
                   starts(l, "*FLUX EXTRADOF")) {
            return R.fail(i, "contact and flux loads are not on this path");
        }
    }
    return true;
}

// Keywords outside *SOLVER that the path does not read: Dirichlet conditions belong in a *SOLVER block.
bool rejectUnsupported(Reader &R) {
    bool inSolver = false;
    for (size_t i = 0; i < R.L.size(); i++) {
        std::string_view l = R.L[i];
        if (starts(l, "*SOLVER")) inSolver = true;
        if (starts(l, "*END SOLVER")) inSolver = false;
        if (inSolver || !starts(l, "*BOUNDARY")) continue;
        const std::string type = tail(l, 16);
        if (type != "FORCE" && type != "PRESSURE RAMP" && type != "PRESSURE INST" &&
            type != "TRACTION RAMP" && type != "TRACTION INST")
            return R.fail(i, "Dirichlet conditions outside *SOLVER are not on this path");
    }
    return true;
}

} // namespace

bool readInp(const std::string &path, InpModel &m, std::string &err) {
    // the whole file in one read, split into lines with trailing blanks and '\r' removed
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = "cannot open input file " + path;
        return false;
    }
    Reader R;
    std::string &buf = R.buf;
    if (std::fseek(f, 0, SEEK_END) == 0) {
        const long size = std::ftell(f);
        if (size > 0) buf.reserve((size_t)size);
        std::rewind(f);
    }
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.append(chunk, got);
    std::fclose(f);
    R.L.reserve((size_t)std::count(buf.begin(), buf.end(), '\n') + 1);
    for (size_t b = 0; b < buf.size();) {
        const void *nl = std::memchr(buf.data() + b, '\n', buf.size() - b);
        const size_t e = nl ? (size_t)(static_cast<const char *>(nl) - buf.data()) : buf.size();
        size_t t = e;
        while (t > b && (buf[t - 1] == '\r' || buf[t - 1] == ' ' || buf[t - 1] == '\t')) t--;
        R.L.emplace_back(buf.data() + b, t - b);
        b = e + 1;
    }
    m = InpModel();
    bool ok = readNodes(R, m) && readElements(R, m) && readFemSets(R) && readMaterials(R, m) && readSolvers(R, m) &&
              readExtractions(R, m) && readLoads(R, m) && rejectUnsupported(R);
    if (!ok) err = R.err;
    return ok;
}

bool elementMaterials(const InpModel &m, std::vector<int32_t> &matOfElem, std::string &err) {
    matOfElem.assign((size_t)m.nElem, -1);
    for (size_t k = 0; k < m.materials.size(); k++) {
        const InpMaterial &mat = m.materials[k];
        if (mat.all)
            for (auto &x : matOfElem) x = (int32_t)k;
        for (int32_t e : mat.elems) matOfElem[(size_t)e] = (int32_t)k;
    }
    for (int32_t e = 0; e < m.nElem; e++)
        if (matOfElem[(size_t)e] < 0) {
            err = "element " + std::to_string(e + 1) + " has no material";
            return false;
        }
    return true;
}

} // namespace mps
