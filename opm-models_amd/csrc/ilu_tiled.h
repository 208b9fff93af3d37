// ilu_tiled.h -- persistent tile-wavefront ILU(0) kernels for the 7-point
// structured pattern (see ilu_tiled.hip).
#pragma once

#include <vector>

#include "omodels_internal.h"

namespace omodels {

struct TileGeom {
    int nx, ny, nz;
    int BX, BY, BZ;     // tile shape, BX*BY*BZ <= 240
    int TX, TY, TZ, T;  // tiles per axis and in total
};

TileGeom makeTileGeom(int nx, int ny, int nz);
// a grid cut into tiles of the given shape (0 = the default shape)
TileGeom makeTileGeom(int nx, int ny, int nz, int bx, int by, int bz);
// tile shapes the autotuner tries for a grid (valid for the sweep kernels)
std::vector<TileGeom> tileShapeCandidates(int nx, int ny, int nz);

class TiledIluSet;

// Tile shapes chosen by the autotuner, per grid size and solver precision
// (shared by every model and ensemble of the process).  `level` = the
// level-scheduled kernels were fastest.
struct TunedIlu {
    TileGeom shape{};
    bool level = false;
    int sweep = 0;      // sweep form of the tiled kernels: 1 streaming, 2 batched
};
bool lookupTunedIlu(int nx, int ny, int nz, int precision, TunedIlu& out);
void storeTunedIlu(int nx, int ny, int nz, int precision, const TunedIlu& t);

// One reservoir (an ensemble of one instance).
class TiledIlu {
public:
    TiledIlu(int nx, int ny, int nz, hipStream_t s);
    // sweepMode: 0 chosen from the tiles per tile level, 1 streaming, 2 batched
    TiledIlu(const TileGeom& shape, hipStream_t s, int sweepMode = 0);
    void setSweepMode(int mode);
    bool streaming() const;
    ~TiledIlu();
    TiledIlu(const TiledIlu&) = delete;
    TiledIlu& operator=(const TiledIlu&) = delete;
    // in-place factorization of lu (a copy of A); inverted pivots also row-wise in dinv
    void factor(const int* rp, const int* diag, double* lu, double* dinv, int* fail, hipStream_t s);
    // v = (LU)^-1 d; active: nullptr, or a device flag, the application is skipped while it is 0.
    // With A and y, also y = A v inside the backward sweep where the sweeps stream;
    // returns whether it did (otherwise y is untouched)
    template <class S, class P>
    bool apply(const int* rp, const int* diag, const S* lu, const P* dinv, const S* d, S* v, hipStream_t s,
               const int* active = nullptr, const int* ci = nullptr, const S* A = nullptr, S* y = nullptr);
    // upload the descriptor of this application ahead of a stream capture
    template <class S, class P>
    void primeApply(const int* rp, const int* diag, const S* lu, const P* dinv, const S* d, S* v, hipStream_t s,
                    const int* active = nullptr);
    long generation() const;
    // throws if a dependency wait ran into its bound
    void checkError(hipStream_t s);
    const TileGeom& geom() const { return g; }

private:
    TileGeom g{};
    TiledIluSet* set_;
};

}  // namespace omodels
