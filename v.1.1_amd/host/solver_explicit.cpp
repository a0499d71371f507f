// solver_explicit.cpp -- classSolverExplicit on the device engine (solver_explicit.h).
#include "solver_explicit.h"

#include <sys/stat.h>
#include <sys/time.h>

#include <algorithm>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <ctime>
#include <deque>
#include <mutex>
#include <thread>

#include "mps_gpu.h"
#include "vtk_output.h"

namespace mps {
namespace {

double wallNow() {
    timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}
double cpuNow() { return (double)std::clock() / CLOCKS_PER_SEC; }

// VTK outputs written on a thread of their own while the solve goes on.  Each output is a device snapshot of
// the step (mps_snapshot_begin) in one of two slots; the thread waits for its copy to pinned memory, formats
// and writes the file, then releases the slot.  An output whose slot is still being written waits for that
// write first, so at most two outputs are in flight and the files are written in order.
class AsyncVtkWriter {
  public:
    AsyncVtkWriter(mps_handle h, VtkSeriesWriter &w, const InpModel &m, const std::vector<int> &outVars)
        : _h(h), _w(w), _m(m), _outVars(outVars) {}
    ~AsyncVtkWriter() {
        std::string ignored;
        finish(ignored);
    }
    bool submit(double time, std::string &err) {
        const int slot = _next;
        _next ^= 1;
        {
            std::unique_lock<std::mutex> lk(_mu);
            _cv.wait(lk, [&] { return !_inUse[slot] || _failed; });
            if (_failed) {
                err = _werr;
                return false;
            }
        }
        if (mps_snapshot_begin(_h, slot, 1) != MPS_OK) {
            err = std::string("mps_snapshot_begin: ") + mps_last_error(_h);
            return false;
        }
        {
            std::lock_guard<std::mutex> lk(_mu);
            _inUse[slot] = true;
            _q.push_back({slot, time});
        }
        if (!_th.joinable()) _th = std::thread([this] { run(); });
        _cv.notify_all();
        return true;
    }
    // Waits for every queued output; false (with err) if one failed.
    bool finish(std::string &err) {
        {
            std::lock_guard<std::mutex> lk(_mu);
            _stop = true;
        }
        _cv.notify_all();
        if (_th.joinable()) _th.join();
        if (_failed) err = _werr;
        return !_failed;
    }

  private:
    struct Job {
        int slot;
        double time;
    };
    void run() {
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(_mu);
                _cv.wait(lk, [&] { return !_q.empty() || _stop; });
                if (_q.empty()) return;
                j = _q.front();
                _q.pop_front();
            }
            const double *u, *v, *a, *e;
            std::string err;
            bool ok = mps_snapshot_wait(_h, j.slot, &u, &v, &a, &e) == MPS_OK;
            if (!ok) err = "mps_snapshot_wait failed";
            if (ok) ok = _w.write(_m, u, v, a, e, _outVars, j.time, err);
            mps_snapshot_release(_h, j.slot);
            {
                std::lock_guard<std::mutex> lk(_mu);
                _inUse[j.slot] = false;
                if (!ok && !_failed) {
                    _failed = true;
                    _werr = err;
                }
            }
            _cv.notify_all();
        }
    }
    mps_handle _h;
    VtkSeriesWriter &_w;
    const InpModel &_m;
    const std::vector<int> &_outVars;
    std::thread _th;
    std::mutex _mu;
    std::condition_variable _cv;
    std::deque<Job> _q;
    bool _inUse[2] = {false, false};
    bool _stop = false, _failed = false;
    std::string _werr;
    int _next = 0;
};

struct Log {
    bool quiet;
    __attribute__((format(printf, 2, 3))) void info(const char *f, ...) const {
        if (quiet) return;
        va_list ap;
        va_start(ap, f);
        std::printf("INFO ::: ");
        std::vprintf(f, ap);
        std::printf("\n");
        std::fflush(stdout);
        va_end(ap);
    }
};

struct Handle {
    mps_handle h = nullptr;
    ~Handle() {
        if (h) mps_destroy(h);
    }
};

constexpr int kHistoryCap = 4096;

} // namespace

int classSolverExplicit::calculate(const InpModel &m, const std::vector<std::unique_ptr<constitutiveModels>> &laws,
                                   const std::vector<int32_t> &lawOfElem, classExtractData &extract,
                                   const RunOptions &opt, std::string &err) {
    const Log log{opt.quiet};
    const double startWall = wallNow(), startCpu = cpuNow();
    const int32_t N = m.nNodes, Ne = m.nElem;

    // Device material descriptors: one law per solve; E, nu, rho per element
    // when the elements carry different materials.
    const DeviceMaterial d0 = laws[(size_t)lawOfElem[0]]->deviceDescriptor();
    bool uniform = true;
    for (auto &l : laws) {
        DeviceMaterial d = l->deviceDescriptor();
        if (d.law != d0.law) {
            err = "elements carry different constitutive laws; one solve runs one law (" + laws[0]->getNameCons() +
                  " and " + l->getNameCons() + ")";
            return -1;
        }
        uniform = uniform && d.young == d0.young && d.nu == d0.nu && d.rho == d0.rho;
    }
    std::vector<double> E, nu, rho;
    if (uniform) {
        E = {d0.young};
        nu = {d0.nu};
        rho = {d0.rho};
    } else {
        E.resize((size_t)Ne);
        nu.resize((size_t)Ne);
        rho.resize((size_t)Ne);
        for (int32_t e = 0; e < Ne; e++) {
            DeviceMaterial d = laws[(size_t)lawOfElem[(size_t)e]]->deviceDescriptor();
            E[(size_t)e] = d.young;
            nu[(size_t)e] = d.nu;
            rho[(size_t)e] = d.rho;
        }
    }
    std::vector<int32_t> dof;
    std::vector<double> val;
    for (const InpDirichlet &d : _s.dirichlet) {
        dof.push_back(d.node + N * d.dir);
        val.push_back(d.value);
    }

    mps_mesh mesh{N, Ne, m.nen, 0, m.X.data(), m.conn.data()};
    mps_material mat{d0.law, uniform ? 0 : 1, E.data(), nu.data(), rho.data()};
    mps_bcs bcs{(int32_t)dof.size(), _s.bcMode, _s.velRampSteps, 0, dof.data(), val.data()};
    mps_solver_cfg cfg{};
    cfg.t0 = _s.t0;
    cfg.tf = _s.tf;
    cfg.scale = _s.scale;
    cfg.n_steps = _s.fixedSteps;
    cfg.precision = opt.precision >= 0 ? opt.precision : _s.precision;
    cfg.energy_every = extract.needsEnergies() ? 1 : 0;
    cfg.history_cap = kHistoryCap;
    cfg.device = opt.device;

    Handle H;
    if (mps_create(&mesh, &mat, &bcs, &cfg, &H.h) != MPS_OK) {
        err = std::string("mps_create: ") + mps_last_error(nullptr);
        return -1;
    }
    mps_handle h = H.h;
    auto fail = [&](const char *what) {
        err = std::string(what) + ": " + mps_last_error(h);
        return -1;
    };
    if (!_u0.empty() || !_v0.empty() || !_a0.empty()) {
        if (mps_set_state(h, _u0.empty() ? nullptr : _u0.data(), _v0.empty() ? nullptr : _v0.data(),
                          _a0.empty() ? nullptr : _a0.data()) != MPS_OK)
            return fail("mps_set_state");
    }
    // artificial viscosity of this solver (solverExplicit.cpp:381-421)
    if (_s.visc != 0 && mps_set_viscosity(h, _s.visc, _s.viscC1, _s.viscCl, 0.0, 0.0) != MPS_OK)
        return fail("mps_set_viscosity");
    // Neumann loads, global for all solvers (solverExplicit.cpp:47; NeumannBCManagement at every step)
    const InpLoads &ld = m.loads;
    if (!ld.empty()) {
        std::vector<mps_load> L;
        for (const InpLoadBlock &b : ld.blocks)
            L.push_back(mps_load{b.type, (int32_t)(b.type == MPS_LOAD_GRAVITY ? 0 : b.index.size()), b.t0, b.tf, b.value,
                                 b.index.data(), b.face.data(), b.force.data()});  // TRACTION: force = T
        if (mps_set_loads(h, (int32_t)L.size(), L.data()) != MPS_OK) return fail("mps_set_loads");
    }
    double dt0 = 0;
    if (mps_critical_dt(h, nullptr, &dt0) != MPS_OK) return fail("mps_critical_dt");

    // Progress sampler (solverExplicit.cpp:168-170).
    long long numStepsT = _s.fixedSteps > 0 ? _s.fixedSteps : (long long)((_s.tf - _s.t0) / dt0);
    const int times = std::max(1, _s.numFiles);
    const long long sampler = std::max(1LL, numStepsT / times);
    log.info("Solver used is EXPLICIT");
    log.info("The initial time step is %g", dt0);
    log.info("done preprocessing in solver (Wall %gs, CPU %gs)", wallNow() - startWall, cpuNow() - startCpu);

    // Nodal extractions (Unknown, InternalForce, ExternalForce at the listed nodes and component) and the
    // *Forces pairs (printForcesOut, solverExplicit.cpp:304: the step's internal force at the listed (node,
    // direction) pairs whenever the output is due) are recorded by the device inside every step
    // (mps_set_dof_series) and read once per chunk of at most kHistoryCap steps, with the steps' times and dts.
    std::vector<int32_t> serKind, serDof;
    std::vector<size_t> nodeOff;  // first series entry of each node extraction
    for (auto &kv : extract.nodeData()) {
        const classNodeData &d = *kv.second;
        const std::string &t = d.getDataType();
        const int32_t kind = t == "Unknown" ? MPS_SERIES_U : t == "InternalForce" ? MPS_SERIES_FINT : MPS_SERIES_FEXT;
        nodeOff.push_back(serDof.size());
        for (int32_t n : d.getNodes()) {
            serKind.push_back(kind);
            serDof.push_back(n + N * d.getComp());
        }
    }
    std::unique_ptr<classPrintForces> forcesOut;
    const size_t forceOff = serDof.size();
    std::vector<double> forceVal;
    if (_s.forces.on) {
        forcesOut.reset(new classPrintForces(_s.forces));
        for (size_t j = 0; j < _s.forces.nodes.size(); j++) {
            serKind.push_back(MPS_SERIES_FINT);
            serDof.push_back(_s.forces.nodes[j] + N * _s.forces.dirs[j]);
        }
        forceVal.resize(_s.forces.nodes.size());
    }
    const size_t nSer = serDof.size();
    if (mps_set_dof_series(h, (int64_t)nSer, serKind.data(), serDof.data()) != MPS_OK) return fail("mps_set_dof_series");
    const bool perStep = nSer > 0 || extract.needsEnergies();
    std::vector<double> series(nSer * (size_t)kHistoryCap);
    std::vector<double> ht(kHistoryCap), hdt(kHistoryCap), hdef(kHistoryCap), hkin(kHistoryCap), hvol(kHistoryCap),
        hext(kHistoryCap);

    const long long chunk = perStep ? kHistoryCap : 1LL << 40;
    // VTK series (outputsManagement, solverExplicit.cpp:173-186, 310-322): at t0 when t0 < 1e-12, then
    // whenever `sampler` steps have passed since the last output, and at the last step.
    // The files are written on a thread of their own from device snapshots, while the solve goes on.
    std::unique_ptr<AsyncVtkWriter> vtk;
    if (opt.vtk) vtk.reset(new AsyncVtkWriter(h, *opt.vtk, m, _s.outVars));
    auto output = [&](double time) -> int { return !vtk || vtk->submit(time, err) ? 0 : -1; };
    if (_s.t0 < 1E-12 && output(_s.t0) != 0) return -1;
    long long sampler2 = 0;
    mps_report rep{};
    for (;;) {
        const long long before = rep.steps;
        const long long n = std::min(chunk, std::max(1LL, sampler - sampler2));
        const int rc = mps_run(h, n, &rep);
        if (rc == MPS_ERR_NAN || rep.status == MPS_ERR_NAN) {
            err = "There are NANs in the displacements. Try a smaller scale factor";
            return -1;
        }
        if (rc == MPS_ERR_DT || rep.status == MPS_ERR_DT) {
            err = "Negative jacobian";
            return -1;
        }
        if (rc != MPS_OK) return fail("mps_run");
        const long long done = rep.steps - before;
        if (done > 0 && perStep) {
            if (mps_get_history(h, before, done, ht.data(), hdt.data(), hdef.data(), hkin.data(), hvol.data()) != MPS_OK)
                return fail("mps_get_history");
            if (extract.needsEnergies() &&
                mps_get_energy_balance(h, before, done, hext.data(), nullptr, nullptr, nullptr) != MPS_OK)
                return fail("mps_get_energy_balance");
            if (nSer > 0 && mps_get_dof_series(h, before, done, series.data()) != MPS_OK)
                return fail("mps_get_dof_series");
        }
        for (long long k = 0; perStep && k < done; k++) {
            const double *row = series.data() + (size_t)k * nSer;
            if (forcesOut && forcesOut->due(ht[k], hdt[k], _s.t0)) {
                for (size_t j = 0; j < forceVal.size(); j++) forceVal[j] = row[forceOff + j];
                if (!forcesOut->write(opt.outputDir, _s.t0, _s.tf, ht[k], forceVal)) {
                    err = "cannot write the *Forces output in " + opt.outputDir;
                    return -1;
                }
            }
            for (auto &kv : extract.elementData()) {
                const std::string &t = kv.second->getDataType();
                const double v = t == "DEFO_ENERGY"  ? hdef[k]
                                 : t == "KIN_ENERGY" ? hkin[k]
                                 : t == "VOLUME"     ? hvol[k]
                                                     : hext[k];  // EXTERNAL_ENERGY
                if (!kv.second->writeData(opt.outputDir, ht[k], v)) {
                    err = "cannot write " + kv.first + ".csv in " + opt.outputDir;
                    return -1;
                }
            }
            size_t e = 0;
            for (auto &kv : extract.nodeData()) {
                classNodeData &d = *kv.second;
                const std::vector<double> vals(row + nodeOff[e], row + nodeOff[e] + d.getNodes().size());
                e++;
                if (!d.writeData(opt.outputDir, ht[k], vals)) {
                    err = "cannot write " + kv.first + ".csv in " + opt.outputDir;
                    return -1;
                }
            }
        }
        sampler2 += done;
        if (sampler2 >= sampler || rep.status == MPS_DONE) {
            if (output(rep.time) != 0) return -1;
            sampler2 = 0;
            log.info("Time %g out of %g", rep.time, _s.fixedSteps > 0 ? rep.time : _s.tf);
        }
        if (rep.status == MPS_DONE) break;
        if (done == 0) {
            err = "mps_run made no progress";
            return -1;
        }
    }
    if (vtk && !vtk->finish(err)) return -1;
    _steps = rep.steps;
    _time = rep.time;
    _u.resize(3 * (size_t)N);
    _v.resize(3 * (size_t)N);
    _a.resize(3 * (size_t)N);
    if (mps_get_state(h, _u.data(), _v.data(), _a.data()) != MPS_OK) return fail("mps_get_state");
    log.info("done solving (Wall %gs, CPU %gs)", wallNow() - startWall, cpuNow() - startCpu);
    return 0;
}

int runInputFile(const std::string &path, const RunOptions &opt, std::string &err, std::vector<double> *finalU) {
    const Log log{opt.quiet};
    InpModel model;
    if (!readInp(path, model, err)) return -1;
    for (const std::string &w : model.warnings)
        if (!opt.quiet) std::fprintf(stderr, "WARNING ::: %s\n", w.c_str());
    log.info("Type of elements %s", model.elemType.c_str());
    std::vector<int32_t> matOfElem;
    if (!elementMaterials(model, matOfElem, err)) return -1;
    std::vector<std::unique_ptr<constitutiveModels>> laws;
    for (const InpMaterial &mat : model.materials) {
        laws.push_back(makeConstitutiveModel(mat.name, mat.nameCons, mat.rho, mat.young, mat.nu));
        if (!laws.back()) {
            err = "constitutive model " + mat.nameCons + " has no device descriptor";
            return -1;
        }
    }
    classExtractData extract;
    if (!extract.init(model, err)) return -1;
    VtkSeriesWriter vtk(opt.outputDir);
    RunOptions o2 = opt;
    if (opt.writeVtk) o2.vtk = &vtk;
    if (!extract.empty() || opt.writeVtk) {
        struct stat st;
        if (stat(opt.outputDir.c_str(), &st) != 0 && mkdir(opt.outputDir.c_str(), 0755) != 0) {
            err = "cannot create output directory " + opt.outputDir;
            return -1;
        }
    }
    std::vector<double> u, v, a;
    for (size_t k = 0; k < model.solvers.size(); k++) {
        classSolverExplicit s(model.solvers[k]);
        if (k > 0) s.initValues(u, v, a);
        if (s.calculate(model, laws, matOfElem, extract, o2, err) != 0) return -1;
        s.getFinalValues(u, v, a);
    }
    extract.finalise();
    if (finalU) *finalU = u;
    return 0;
}

} // namespace mps
