// lens_host.h -- the two-phase immiscible element path behind the C ABI:
// LensModel (FvBaseDiscretization state + FvBaseLinearizer for the lens
// problem) and its Newton method / time loop.
#pragma once

#include <vector>

#include "../../include/omodels.h"
#include "lens_physics.h"
#include "omodels_host.h"

namespace omodels {

LensParams makeLensParams(const omodels_lens_problem* p);

class LensModel : public LinearSystem {
public:
    explicit LensModel(const omodels_lens_problem* p, hipStream_t s = nullptr);

    LensParams P{};
    std::vector<int32_t> hRowPtr, hColIdx, hDiag;
    std::vector<uint8_t> hInLens;
    DevBuf<double> pv0, pv1, st1;
    DevBuf<uint8_t> inLens;
    DevBuf<int> nbOff, bad;

    void initial(double* pv) const;
    // FvBaseLinearizer::linearize (fvbaselinearizer.hh:510-583) at solution(0)
    void linearizeDomain(double dt, int iterationIndex);
    void advanceTimeLevel();
    void restoreTimeLevel();
    // NewtonMethod::update_: x <- x - dx; false when dx has non-finite entries
    bool update(const double* dx);
};

// Simulator::run until tEnd (simulator.hh:481-512,697-810) with
// FvBaseProblem::timeIntegration / nextTimeStepSize (fvbaseproblem.hh:512-606)
// and NewtonMethod::apply (newtonmethod.hh:228-465)
int simulateLens(LensModel& m, const omodels_params& prm, double tEnd, double dt0, int maxSteps, double* tHist,
                 double* dtHist, int* nSteps, std::vector<omodels_iter_log>* log, omodels_run_stats* stats);

}  // namespace omodels
