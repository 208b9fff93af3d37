// A C++ consumer of the adapters in omodels_opm.hpp, built against
// include/ and libomodels_amd.so the way reference-side code would be: a
// stand-in Simulator with its Glue specialization, one Newton iteration
// through the linearizer, the BiCGStab backend and the black-oil update, and
// a second linearization.  Reads a case from a directory of raw arrays and
// writes what it computed there.
//
// usage: omodels_consumer <dir>
//   <dir>/dims.txt: nx ny nz n_src n_oil n_gas dx dy dz gravity dt
//   <dir>/<name>.bin: perm poro oil_tab gas_tab consts src_cell src_type src_a
//                     src_b pv0 mean0 [trans depth volume thpres]
// writes jac0 res0 x pv1 mean1 jac1 res1 flows jac2 res2 (.out) and prints one
// summary line.
// Exit codes: 0 done, 3 an error of the library (message on stderr).
#include <cstdio>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "omodels_opm.hpp"

namespace {

template <class T>
bool readBin(const std::string& path, std::vector<T>& v, bool optional = false)
{
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) {
        if (optional) { v.clear(); return false; }
        throw std::runtime_error("missing " + path);
    }
    const std::streamsize n = f.tellg();
    f.seekg(0);
    v.resize((size_t)n / sizeof(T));
    f.read(reinterpret_cast<char*>(v.data()), n);
    return true;
}

template <class T>
void writeBin(const std::string& path, const std::vector<T>& v)
{
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

// the reference's Simulator, reduced to what the glue reads
struct StandInSimulator {
    int nx = 0, ny = 0, nz = 0, nSrc = 0, nOil = 0, nGas = 0;
    double dx = 0, dy = 0, dz = 0, gravity = 0, dt = 0;
    std::vector<double> perm, poro, oilTab, gasTab, consts, srcA, srcB;
    std::vector<int32_t> srcCell, srcType;
    std::vector<double> pv;            // solution(0)
    std::vector<uint8_t> meaning;
    int newtonIteration = 0;
    std::vector<double> trans, depth, volume, thpres;
};

// an auxiliary module (BaseAuxiliaryModule's role): one DOF whose first
// unknown follows cell 0's pressure, coupled into cell 0's water equation
struct PressureProbe : omodels_opm::GpuAuxiliaryModule {
    unsigned numDofs() const override { return 1; }
    void addNeighbors(std::vector<std::pair<int32_t, int32_t>>& e) const override
    {
        const int32_t a = localToGlobalDof(0);
        e.push_back({a, a});
        e.push_back({0, a});
        e.push_back({a, 0});
    }
    void applyInitial(std::vector<double>& x) override { x = {1.0e5, 0.0, 0.0}; }
    void linearize(const std::vector<double>& pv, const std::vector<uint8_t>&, const std::vector<double>& x,
                   omodels_opm::AuxTerms& t) override
    {
        const int32_t a = localToGlobalDof(0);
        const double k = 1.0e-6, dp = x[0] - pv[1];
        t.resRow = {a, 0};
        t.res = {dp, x[1], x[2], 0.0, k * dp, 0.0};
        t.blockRow = {a, a, 0, 0};
        t.blockCol = {a, 0, 0, a};
        t.blocks = {1, 0, 0, 0, 1, 0, 0, 0, 1,          // d(row a)/d(x)
                    0, -1, 0, 0, 0, 0, 0, 0, 0,         // d(row a)/d(cell 0): -dp0
                    0, 0, 0, 0, -k, 0, 0, 0, 0,         // d(cell 0 water)/d(p0)
                    0, 0, 0, k, 0, 0, 0, 0, 0};         // d(cell 0 water)/d(x0)
    }
};

}  // namespace

template <>
struct omodels_opm::Glue<StandInSimulator> {
    static omodels_blackoil_problem describe(const StandInSimulator& s)
    {
        omodels_blackoil_problem p{};
        p.nx = s.nx; p.ny = s.ny; p.nz = s.nz; p.n_src = s.nSrc;
        p.dx = s.dx; p.dy = s.dy; p.dz = s.dz; p.gravity = s.gravity;
        p.perm = s.perm.data(); p.poro = s.poro.data();
        p.n_oil = s.nOil; p.n_gas = s.nGas;
        p.oil_tab = s.oilTab.data(); p.gas_tab = s.gasTab.data();
        p.oil_comp = s.consts[0]; p.water_pref = s.consts[1]; p.water_comp = s.consts[2]; p.water_mu = s.consts[3];
        p.rho_ref[0] = s.consts[4]; p.rho_ref[1] = s.consts[5]; p.rho_ref[2] = s.consts[6];
        p.src_cell = s.srcCell.data(); p.src_type = s.srcType.data();
        p.src_a = s.srcA.data(); p.src_b = s.srcB.data();
        return p;
    }
    static void packSolution(const StandInSimulator& s, std::vector<double>& pv, std::vector<uint8_t>& meaning)
    {
        pv = s.pv;
        meaning = s.meaning;
    }
    static void unpackSolution(StandInSimulator& s, const std::vector<double>& pv, const std::vector<uint8_t>& meaning)
    {
        s.pv = pv;
        s.meaning = meaning;
    }
    // the reservoir problem's boundaries are closed: no face has a condition
    static void boundaryConditions(const StandInSimulator&, std::vector<int32_t>& cell, std::vector<int32_t>& face,
                                   std::vector<int32_t>& type, std::vector<double>& rate, std::vector<double>& state,
                                   std::vector<uint8_t>& meaning)
    {
        cell.clear(); face.clear(); type.clear(); rate.clear(); state.clear(); meaning.clear();
    }
    static double timeStepSize(const StandInSimulator& s) { return s.dt; }
    static int newtonIteration(const StandInSimulator& s) { return s.newtonIteration; }
    static void discretization(const StandInSimulator& s, std::vector<double>& trans, std::vector<double>& depth,
                               std::vector<double>& volume, std::vector<double>& thpres)
    {
        trans = s.trans;
        depth = s.depth;
        volume = s.volume;
        thpres = s.thpres;
    }
};

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::cerr << "usage: omodels_consumer <dir>\n";
        return 1;
    }
    const std::string dir = argv[1];
    StandInSimulator sim;
    try {
        std::ifstream d(dir + "/dims.txt");
        d >> sim.nx >> sim.ny >> sim.nz >> sim.nSrc >> sim.nOil >> sim.nGas >> sim.dx >> sim.dy >> sim.dz
            >> sim.gravity >> sim.dt;
        if (!d) throw std::runtime_error("bad dims.txt");
        readBin(dir + "/perm.bin", sim.perm);
        readBin(dir + "/poro.bin", sim.poro);
        readBin(dir + "/oil_tab.bin", sim.oilTab);
        readBin(dir + "/gas_tab.bin", sim.gasTab);
        readBin(dir + "/consts.bin", sim.consts);
        readBin(dir + "/src_cell.bin", sim.srcCell);
        readBin(dir + "/src_type.bin", sim.srcType);
        readBin(dir + "/src_a.bin", sim.srcA);
        readBin(dir + "/src_b.bin", sim.srcB);
        readBin(dir + "/pv0.bin", sim.pv);
        readBin(dir + "/mean0.bin", sim.meaning);
        readBin(dir + "/trans.bin", sim.trans, true);
        readBin(dir + "/depth.bin", sim.depth, true);
        readBin(dir + "/volume.bin", sim.volume, true);
        readBin(dir + "/thpres.bin", sim.thpres, true);
    } catch (const std::exception& e) {
        std::cerr << "input: " << e.what() << "\n";
        return 1;
    }
    {
        // the host-only entry points: defaults and the reference problem's tables
        omodels_params prm;
        omodels_default_params(&prm);
        std::vector<double> oil(4 * 16), gas(3 * 16), consts(7);
        int32_t nOil = 0, nGas = 0;
        if (omodels_blackoil_default_fluid(oil.data(), &nOil, gas.data(), &nGas, consts.data()) != 0) {
            std::cerr << "error: default_fluid: " << omodels_last_error() << "\n";
            return 3;
        }
        std::printf("{\"host\": {\"lin_tol\": %g, \"newton_max_iter\": %d, \"n_oil\": %d, \"n_gas\": %d}}\n",
                    prm.lin_tol, prm.newton_max_iter, nOil, nGas);
        std::fflush(stdout);
    }
    try {
        omodels_opm::GpuTpfaLinearizer<StandInSimulator> lin;
        lin.init(sim);
        omodels_params prm;
        omodels_default_params(&prm);
        omodels_opm::GpuBiCGStabBackend backend(lin.handle(), prm);

        lin.updateBoundaryConditionData();
        omodels_opm::LinearizationType lt;
        lt.time = 1;
        lin.setLinearizationType(lt);
        if (lin.getLinearizationType().time != 1 || !lin.constraintsMap().empty())
            throw std::runtime_error("linearization type / constraints");
        sim.newtonIteration = 0;
        lin.linearizeDomain();
        writeBin(dir + "/jac0.out", lin.jacobian());
        writeBin(dir + "/res0.out", lin.residual());
        backend.prepare(lin.jacobian(), lin.residual());
        std::vector<double> x(lin.residual().size(), 0.0);
        const bool converged = backend.solve(x);
        writeBin(dir + "/x.out", x);
        const int switched = omodels_opm::newtonUpdate(lin, sim, prm, x);
        writeBin(dir + "/pv1.out", sim.pv);
        writeBin(dir + "/mean1.out", sim.meaning);
        sim.newtonIteration = 1;
        lin.linearizeDomain();
        writeBin(dir + "/jac1.out", lin.jacobian());
        writeBin(dir + "/res1.out", lin.residual());
        // linearizeDomain(domain) over every cell: the full linearization again
        struct { std::vector<int> cells; } all;
        for (long long c = 0; c < (long long)lin.numCells(); ++c) all.cells.push_back((int)c);
        const std::vector<double> j1 = lin.jacobian(), r1 = lin.residual();
        lin.linearizeDomain(all);
        const bool same = lin.jacobian() == j1 && lin.residual() == r1;
        // FLOWS of the current state, then linearize() with an auxiliary module
        lin.updateFlowsInfo();
        writeBin(dir + "/flows.out", lin.getFlowsInfo());
        writeBin(dir + "/velocity.out", lin.getVelocityInfo());
        PressureProbe probe;
        lin.addAuxiliaryModule(&probe);
        lin.linearize();
        writeBin(dir + "/jac2.out", lin.jacobian());
        writeBin(dir + "/res2.out", lin.residual());
        std::printf("{\"converged\": %d, \"iterations\": %d, \"switched\": %d, \"cells\": %lld, "
                    "\"subdomain_same\": %d}\n",
                    converged ? 1 : 0, backend.iterations(), switched, (long long)lin.numCells(), same ? 1 : 0);
        return 0;
    } catch (const omodels_opm::NumericalProblem& e) {
        std::cerr << "numerical problem: " << e.what() << "\n";
        return 3;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 3;
    }
}
