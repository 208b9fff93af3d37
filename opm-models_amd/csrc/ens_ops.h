// ens_ops.h -- host launchers of the ensemble vector kernels (ens_kernels.hip)
// and of the ensemble assembly kernels (bo_kernels.hip).
#pragma once

#include "ensemble.h"

namespace omodels {

// fdev: the same fluid in device memory (read by the assembly)
void launchEnsAssemble(const BoFluid& f, const BoFluid* fdev, const EnsInst* insts, const int2* table, int nBlocks,
                       const int* active, const int2* srcTable, int nSrc, hipStream_t s);
void launchEnsStorage(const BoFluid& f, const EnsInst* insts, const int2* table, int nChunks, const int* refresh,
                      hipStream_t s);
void launchEnsUpdate(const BoFluid& f, const EnsInst* insts, const int2* table, int nChunks, const int* active,
                     const NewtonUpdateParams& prm, hipStream_t s);
void assemblyGrid(const BoGrid& G, int& gx, int& gy, int& gz, int& kChunk, int targetBlocks);

// Work of the exact scalar products: items (member, workgroup index within the
// member, the member's workgroup count, unused), two device accumulators and
// a ticket per member (zero between products).
struct EnsDots {
    const int4* items;
    int nItems;
    ExactAcc* acc;
    unsigned* tickets;
};

// BiCGStab steps over all active instances (the same sequence as the
// single-reservoir solver, one launch per step for the whole ensemble)
class EnsOps {
public:
    EnsOps(const EnsInst* insts, const int2* table, int nChunks, const int* first, int nInst, double* part,
           const EnsDots& dots, KrylovScalars* ks, hipStream_t s);
    template <class S> void copy(const int* act, int dst, int src);   // src < 0: zero
    template <class S> void dot(const int* act, int ia, int ib, int op, int tagOff);
    template <class S> void updateP(const int* act);
    template <class S> void updateXR(const int* act);
    template <class S> void dot2(const int* act, int tagOff);
    template <class S> void updateXR2(const int* act, int tagOff);
    template <class S> void spmv(const int* act, int ix, int iy);
    template <class S> void maxAbs(const int* act, int id, int useRes, double* out);
    void copyMatrix(const int* act, int toFloat);
    void factorToFloat(const int* act);
    void rhsToFloat(const int* act);
    void addCorrection(const int* act);
    void residual64(const int* act);

private:
    const EnsInst* insts_;
    const int2* table_;
    int nChunks_;
    const int* first_;
    int nInst_;
    double* part_;
    EnsDots dots_;
    KrylovScalars* ks_;
    hipStream_t s_;
};

}  // namespace omodels
