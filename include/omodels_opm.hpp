/* omodels_opm.hpp -- header-only C++ adapters that put the C ABI of omodels.h
 * behind the interfaces opm-models' TypeTag properties expect of a linearizer
 * (TpfaLinearizer, tpfalinearizer.hh:116-379 and :868-883), a linear-solver
 * backend (ParallelBaseBackend, parallelbasebackend.hh:137-290) and the
 * black-oil Newton update (BlackOilNewtonMethod::updatePrimaryVariables_,
 * blackoilnewtonmethod.hh:226-427).
 *
 * The simulator side is reached through one customization point, Glue<Sim>,
 * which the integrating code specializes for its Simulator type:
 *
 *   template <> struct omodels_opm::Glue<MySimulator> {
 *     // the problem description (grid, permeability, porosity, PVT tables,
 *     // wells as source terms); the arrays must outlive the call
 *     static omodels_blackoil_problem describe(const MySimulator&);
 *     // solution(0) as [N][3] primary variables and [N] gas meanings
 *     static void packSolution(const MySimulator&, std::vector<double>& pv, std::vector<uint8_t>& meaning);
 *     static void unpackSolution(MySimulator&, const std::vector<double>& pv, const std::vector<uint8_t>& meaning);
 *     static double timeStepSize(const MySimulator&);          // simulator.timeStepSize()
 *     static int newtonIteration(const MySimulator&);          // newtonMethod().numIterations()
 *     // optional per-neighbour inputs (tpfalinearizer.hh:443-452); leave a
 *     // vector empty for the structured grid's own values
 *     static void discretization(const MySimulator&, std::vector<double>& trans, std::vector<double>& depth,
 *                                std::vector<double>& volume, std::vector<double>& thpres);
 *     // only for updateBoundaryConditionData: the faces with a condition
 *     static void boundaryConditions(const MySimulator&, std::vector<int32_t>& cell, std::vector<int32_t>& face,
 *                                    std::vector<int32_t>& type, std::vector<double>& massRate,
 *                                    std::vector<double>& state, std::vector<uint8_t>& meaning);
 *   };
 *
 * Errors: a negative return code of the C ABI throws std::runtime_error with
 * omodels_last_error(); a numerical failure (return code 2) throws
 * omodels_opm::NumericalProblem, the reference's Opm::NumericalProblem role.
 */
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "omodels.h"

namespace omodels_opm {

struct NumericalProblem : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline int check(int rc, const char* what)
{
    if (rc < 0) throw std::runtime_error(std::string(what) + ": " + omodels_last_error());
    if (rc == 2) throw NumericalProblem(std::string(what) + ": " + omodels_last_error());
    return rc;
}

template <class Simulator>
struct Glue;    // specialized by the integrating code

/* LinearizationType (linearizationtype.hh:34-39): the linearizer stores it
 * (tpfalinearizer.hh:300-306); its TPFA path linearizes fully implicitly
 * whatever it holds, and so does this one. */
struct LinearizationType {
    enum VarType { implicit, pressure, seqtransport };
    VarType type = implicit;
    unsigned int time = 0;
};

/* An auxiliary module (BaseAuxiliaryModule, baseauxiliarymodule.hh:55-121):
 * its DOFs follow the cells (setDofOffset), addNeighbors lists the pattern
 * entries it needs as (row, column) pairs of global DOFs, applyInitial
 * writes its initial solution ([numDofs][3]) and linearize adds its terms
 * for the current cell state and its own solution. */
struct AuxTerms {
    std::vector<int32_t> blockRow, blockCol;   // blocks (row-major 3x3) added at (row, col)
    std::vector<double> blocks;
    std::vector<int32_t> resRow;               // residual rows (3 values each) added at row
    std::vector<double> res;
};
class GpuAuxiliaryModule {
public:
    virtual ~GpuAuxiliaryModule() = default;
    virtual unsigned numDofs() const = 0;
    virtual void addNeighbors(std::vector<std::pair<int32_t, int32_t>>& entries) const = 0;
    virtual void applyInitial(std::vector<double>& x) = 0;
    /* pv / meaning: the cells' solution(0); x: this module's solution(0) */
    virtual void linearize(const std::vector<double>& pv, const std::vector<uint8_t>& meaning,
                           const std::vector<double>& x, AuxTerms& terms) = 0;
    void setDofOffset(int value) { dofOffset_ = value; }
    int dofOffset() const { return dofOffset_; }
    int localToGlobalDof(unsigned localDofIdx) const { return dofOffset_ + (int)localDofIdx; }

private:
    int dofOffset_ = 0;
};

/* Linearizer: createMatrix_ (init), linearize, linearizeDomain (whole and
 * sub-domain), linearizeAuxiliaryEquations with the auxiliary modules added,
 * jacobian() / residual() as host arrays in the Dune layouts (row-major 3x3
 * blocks in the sorted BCRS pattern, residual [rows][3]), eraseMatrix,
 * updateStoredTransmissibilities, updateBoundaryConditionData, the flows. */
template <class Simulator, class G = Glue<Simulator>>
class GpuTpfaLinearizer {
public:
    GpuTpfaLinearizer() = default;
    GpuTpfaLinearizer(const GpuTpfaLinearizer&) = delete;
    GpuTpfaLinearizer& operator=(const GpuTpfaLinearizer&) = delete;
    ~GpuTpfaLinearizer()
    {
        if (model_) omodels_model_destroy(model_);
    }

    void init(Simulator& simulator)
    {
        simulator_ = &simulator;
        const omodels_blackoil_problem p = G::describe(simulator);
        model_ = omodels_model_create(&p);
        if (!model_) throw std::runtime_error(std::string("model_create: ") + omodels_last_error());
        int64_t n = 0, nnzb = 0;
        check(omodels_model_num_cells(model_, &n, &nnzb), "num_cells");
        n_ = n;
        nnzb_ = nnzb;
        gridNnzb_ = nnzb;
        rowPtr_.resize(n + 1);
        colIdx_.resize(nnzb);
        check(omodels_linearizer_pattern(model_, rowPtr_.data(), colIdx_.data(), nullptr, nullptr, nullptr,
                                         nullptr, nullptr, nullptr),
              "pattern");
        jacobian_.assign(9 * (size_t)nnzb, 0.0);
        residual_.assign(3 * (size_t)n, 0.0);
        updateStoredTransmissibilities();
    }

    /* tpfalinearizer.hh:189-245 at the simulator's current solution */
    void linearizeDomain()
    {
        std::vector<double> pv;
        std::vector<uint8_t> meaning;
        G::packSolution(*simulator_, pv, meaning);
        if ((int64_t)pv.size() != 3 * n_ || (int64_t)meaning.size() != n_)
            throw std::invalid_argument("packSolution: wrong sizes");
        check(omodels_model_set_solution(model_, pv.data(), meaning.data(), 0), "set_solution");
        check(omodels_linearizer_linearize_domain(model_, G::timeStepSize(*simulator_),
                                                  G::newtonIteration(*simulator_)),
              "linearize_domain");
        check(omodels_linearizer_get_jacobian(model_, jacobian_.data()), "get_jacobian");
        check(omodels_linearizer_get_residual(model_, residual_.data()), "get_residual");
    }

    /* linearizeDomain(domain) (tpfalinearizer.hh:226-245): the cells of
     * domain.cells only, over the previous Jacobian and residual */
    template <class SubDomain>
    void linearizeDomain(const SubDomain& domain)
    {
        std::vector<double> pv;
        std::vector<uint8_t> meaning;
        G::packSolution(*simulator_, pv, meaning);
        if ((int64_t)pv.size() != 3 * n_ || (int64_t)meaning.size() != n_)
            throw std::invalid_argument("packSolution: wrong sizes");
        std::vector<int32_t> cells(domain.cells.begin(), domain.cells.end());
        check(omodels_model_set_solution(model_, pv.data(), meaning.data(), 0), "set_solution");
        check(omodels_linearizer_linearize_subdomain(model_, G::timeStepSize(*simulator_),
                                                     G::newtonIteration(*simulator_), (int32_t)cells.size(),
                                                     cells.data()),
              "linearize_subdomain");
        check(omodels_linearizer_get_jacobian(model_, jacobian_.data()), "get_jacobian");
        check(omodels_linearizer_get_residual(model_, residual_.data()), "get_residual");
    }

    /* updateStoredTransmissibilities (tpfalinearizer.hh:868-883), with the
     * depths, volumes and threshold pressures createMatrix_ reads (:443-452) */
    void updateStoredTransmissibilities()
    {
        std::vector<double> t, z, v, th;
        G::discretization(*simulator_, t, z, v, th);
        auto arr = [](const std::vector<double>& a, int64_t n) -> const double* {
            if (a.empty()) return nullptr;
            if ((int64_t)a.size() != n) throw std::invalid_argument("discretization: wrong sizes");
            return a.data();
        };
        check(omodels_linearizer_update_discretization(model_, arr(t, 3 * n_), arr(z, n_), arr(v, n_),
                                                       arr(th, 3 * n_)),
              "update_discretization");
    }

    void eraseMatrix() { check(omodels_linearizer_erase_matrix(model_), "erase_matrix"); }

    /* linearize() (tpfalinearizer.hh:171-177): the domain, then the
     * auxiliary equations; finalize() has nothing left to do here */
    void linearize()
    {
        linearizeDomain();
        linearizeAuxiliaryEquations();
    }
    void finalize() {}

    /* model.addAuxiliaryModule: the module's DOFs follow the cells and the
     * modules added before; the pattern grows by its entries (createMatrix_,
     * :503-507) and its initial solution is set (applyInitial).  Create the
     * solver backend after the last module. */
    void addAuxiliaryModule(GpuAuxiliaryModule* module)
    {
        modules_.push_back(module);
        int32_t off = (int32_t)n_;
        std::vector<std::pair<int32_t, int32_t>> entries;
        std::vector<double> x0;
        for (GpuAuxiliaryModule* m : modules_) {
            m->setDofOffset(off);
            off += (int32_t)m->numDofs();
            m->addNeighbors(entries);
            std::vector<double> x(3 * (size_t)m->numDofs(), 0.0);
            m->applyInitial(x);
            x0.insert(x0.end(), x.begin(), x.end());
        }
        nAux_ = off - n_;
        std::vector<int32_t> rows, cols;
        for (const auto& e : entries) { rows.push_back(e.first); cols.push_back(e.second); }
        check(omodels_model_set_auxiliary(model_, (int32_t)nAux_, (int32_t)rows.size(), rows.data(), cols.data()),
              "set_auxiliary");
        check(omodels_model_set_auxiliary_solution(model_, x0.data(), 1), "set_auxiliary_solution");
        refreshPattern_();
    }

    /* linearizeAuxiliaryEquations (tpfalinearizer.hh:254-281): every
     * module's terms added to the linearized system, in module order */
    void linearizeAuxiliaryEquations()
    {
        if (modules_.empty()) return;
        std::vector<double> pv, xall(3 * (size_t)nAux_);
        std::vector<uint8_t> meaning;
        G::packSolution(*simulator_, pv, meaning);
        check(omodels_model_get_auxiliary_solution(model_, xall.data()), "get_auxiliary_solution");
        for (GpuAuxiliaryModule* m : modules_) {
            const size_t b = 3 * (size_t)(m->dofOffset() - n_);
            std::vector<double> x(xall.begin() + b, xall.begin() + b + 3 * m->numDofs());
            AuxTerms t;
            m->linearize(pv, meaning, x, t);
            check(omodels_linearizer_add_auxiliary(model_, (int32_t)t.blockRow.size(), t.blockRow.data(),
                                                   t.blockCol.data(), t.blocks.data(), (int32_t)t.resRow.size(),
                                                   t.resRow.data(), t.res.data()),
                  "add_auxiliary");
        }
        check(omodels_linearizer_get_jacobian(model_, jacobian_.data()), "get_jacobian");
        check(omodels_linearizer_get_residual(model_, residual_.data()), "get_residual");
    }

    /* updateFlowsInfo / getFlowsInfo / getFloresInfo (tpfalinearizer.hh:
     * 647-705): per cell and neighbour in intersection order, [faces][3] */
    void updateFlowsInfo()
    {
        const size_t nf = (size_t)(gridNnzb_ - n_);
        flows_.assign(3 * nf, 0.0);
        flores_.assign(3 * nf, 0.0);
        check(omodels_linearizer_flows(model_, flows_.data(), flores_.data()), "flows");
        // getVelocityInfo: the reference's darcyFlux value (FLORES, the
        // reservoir rate through the face) divided by the face area
        // (tpfalinearizer.hh:759-762)
        const omodels_blackoil_problem p = G::describe(*simulator_);
        const double area[3] = {p.dy * p.dz, p.dx * p.dz, p.dx * p.dy};
        std::vector<int32_t> nbPtr((size_t)n_ + 1), nbCell(nf);
        check(omodels_linearizer_pattern(model_, nullptr, nullptr, nullptr, nbPtr.data(), nbCell.data(), nullptr,
                                         nullptr, nullptr),
              "pattern");
        velocities_.assign(3 * nf, 0.0);
        for (int64_t I = 0; I < n_; ++I)
            for (int32_t k = nbPtr[I]; k < nbPtr[I + 1]; ++k) {
                const int64_t d = nbCell[k] > I ? nbCell[k] - I : I - nbCell[k];
                const int ax = (p.nx > 1 && d == 1) ? 0 : ((p.ny > 1 && d == p.nx) ? 1 : 2);
                for (int e = 0; e < 3; ++e) velocities_[3 * (size_t)k + e] = flores_[3 * (size_t)k + e] / area[ax];
            }
    }
    const std::vector<double>& getFlowsInfo() const { return flows_; }
    const std::vector<double>& getFloresInfo() const { return flores_; }
    /* per neighbour entry and equation: FLORES / face area, the velocities
     * the reference's dispersion reads (it fills them when dispersion is
     * enabled; here after every updateFlowsInfo) */
    const std::vector<double>& getVelocityInfo() const { return velocities_; }

    /* updateBoundaryConditionData: the faces with a condition from
     * Glue::boundaryConditions(sim, cell, face, type, massRate, state,
     * meaning) (see omodels_model_set_boundary_conditions) */
    void updateBoundaryConditionData()
    {
        std::vector<int32_t> cell, face, type;
        std::vector<double> rate, state;
        std::vector<uint8_t> meaning;
        G::boundaryConditions(*simulator_, cell, face, type, rate, state, meaning);
        const size_t n = cell.size();
        if (face.size() != n || type.size() != n)
            throw std::invalid_argument("boundaryConditions: wrong sizes");
        check(omodels_model_set_boundary_conditions(model_, (int32_t)n, cell.data(), face.data(), type.data(),
                                                    rate.empty() ? nullptr : rate.data(),
                                                    state.empty() ? nullptr : state.data(),
                                                    meaning.empty() ? nullptr : meaning.data(), nullptr, nullptr),
              "set_boundary_conditions");
    }

    void setLinearizationType(LinearizationType t) { linearizationType_ = t; }
    const LinearizationType& getLinearizationType() const { return linearizationType_; }
    /* the TPFA linearizer has no constrained DOFs (tpfalinearizer.hh:366-367) */
    template <class Constraints = int>
    std::map<unsigned, Constraints> constraintsMap() const { return {}; }

    const std::vector<int32_t>& rowPtr() const { return rowPtr_; }
    const std::vector<int32_t>& colIdx() const { return colIdx_; }
    const std::vector<double>& jacobian() const { return jacobian_; }
    const std::vector<double>& residual() const { return residual_; }
    int64_t numCells() const { return n_; }
    /* rows of the system: the cells and the auxiliary modules' DOFs */
    int64_t numTotalDof() const { return n_ + nAux_; }
    omodels_model* handle() const { return model_; }

private:
    void refreshPattern_()
    {
        int64_t n = 0;
        check(omodels_model_num_cells(model_, &n, &nnzb_), "num_cells");
        rowPtr_.assign((size_t)(n_ + nAux_ + 1), 0);
        colIdx_.assign((size_t)nnzb_, 0);
        check(omodels_linearizer_pattern(model_, rowPtr_.data(), colIdx_.data(), nullptr, nullptr, nullptr, nullptr,
                                         nullptr, nullptr),
              "pattern");
        jacobian_.assign(9 * (size_t)nnzb_, 0.0);
        residual_.assign(3 * (size_t)(n_ + nAux_), 0.0);
    }

    Simulator* simulator_ = nullptr;
    omodels_model* model_ = nullptr;
    int64_t n_ = 0, nnzb_ = 0, gridNnzb_ = 0, nAux_ = 0;
    std::vector<int32_t> rowPtr_, colIdx_;
    std::vector<double> jacobian_, residual_, flows_, flores_, velocities_;
    std::vector<GpuAuxiliaryModule*> modules_;
    LinearizationType linearizationType_;
};

/* Linear-solver backend: prepare(M, b), setResidual, getResidual, solve(x),
 * iterations (parallelbasebackend.hh:137-290) with ISTL BiCGStab and the
 * block ILU(0) / Jacobi preconditioner the parameters select. */
class GpuBiCGStabBackend {
public:
    GpuBiCGStabBackend(omodels_model* model, const omodels_params& prm)
        : solver_(omodels_solver_create(model, &prm))
    {
        if (!solver_) throw std::runtime_error(std::string("solver_create: ") + omodels_last_error());
    }
    GpuBiCGStabBackend(const GpuBiCGStabBackend&) = delete;
    GpuBiCGStabBackend& operator=(const GpuBiCGStabBackend&) = delete;
    ~GpuBiCGStabBackend() { omodels_solver_destroy(solver_); }

    void prepare(const std::vector<double>& values, const std::vector<double>& b)
    {
        check(omodels_solver_prepare(solver_, values.data(), b.data()), "prepare");
    }
    void setResidual(const std::vector<double>& b) { check(omodels_solver_prepare(solver_, nullptr, b.data()), "set_residual"); }
    void getResidual(std::vector<double>& b) const { check(omodels_solver_get_residual(solver_, b.data()), "get_residual"); }
    /* false when the solver did not converge, as the reference's solve() */
    bool solve(std::vector<double>& x) { return check(omodels_solver_solve(solver_, x.data()), "solve") == 0; }
    int iterations() const { return omodels_solver_iterations(solver_); }

private:
    omodels_solver* solver_;
};

/* The black-oil Newton update of the linearizer's model state
 * (blackoilnewtonmethod.hh:226-427): solution(0) -= delta with the chopping
 * and the switching of the gas meaning; returns the number of switched cells.
 * The caller copies the new state back with Glue::unpackSolution. */
template <class Simulator, class G = Glue<Simulator>>
int newtonUpdate(GpuTpfaLinearizer<Simulator, G>& lin, Simulator& sim, const omodels_params& prm,
                 const std::vector<double>& delta)
{
    int32_t switched = 0;
    check(omodels_newton_update(lin.handle(), &prm, delta.data(), &switched), "newton_update");
    std::vector<double> pv(3 * (size_t)lin.numCells());
    std::vector<uint8_t> meaning((size_t)lin.numCells());
    check(omodels_model_get_solution(lin.handle(), pv.data(), meaning.data()), "get_solution");
    G::unpackSolution(sim, pv, meaning);
    return switched;
}

}  // namespace omodels_opm
