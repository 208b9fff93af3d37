// io.cpp -- the data formats on either side of a run: restart files in the
// reference's text format and VTK unstructured-grid files of cell fields.
//
// Restart (Restart, restart.hh:40-280, with Simulator::serialize,
// simulator.hh:928-940, and the black-oil entity lines of
// BlackOilModel::serializeEntity, blackoilmodel.hh:468-492): a cookie line
// naming the grid, a "Simulator" section with the time manager's state, then
// an "Entities: Codim 0" section with one line per cell: the three primary
// variables, the gas, water and pressure meanings and the PVT region, every
// section closed by an empty line.  Numbers are written with 20 significant
// digits, which read back to the same doubles.
//
// VTK: one .vtu (XML, ASCII) with the cells as hexahedra of the structured
// grid and one Float64 array per cell field, the layout VtkMultiWriter
// produces for element-centred data (vtkmultiwriter.hh).
#include <cctype>
#include <cfloat>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "omodels_io.h"

namespace omodels {

namespace {

std::string restartCookie(long nx, long ny, long nz)
{
    const long numElements = nx * ny * nz;
    const long numVertices = (nx + 1) * (ny + 1) * (nz + 1);
    const long numEdges = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz;
    std::ostringstream oss;
    oss << "eWoms restart file: gridName='blubb' numCPUs=1 myRank=0 numElements=" << numElements
        << " numEdges=" << numEdges << " numVertices=" << numVertices;
    return oss.str();
}

void expectLine(std::istream& in, const std::string& want, const char* what)
{
    std::string line;
    if (!std::getline(in, line)) throw std::runtime_error(std::string("restart file ends before ") + what);
    if (line != want) throw std::runtime_error(std::string("restart file: could not start section '") + want + "'");
}

// the empty line that closes a section: nothing but white space is left on it
void expectSectionEnd(std::istream& in)
{
    std::string line;
    if (!std::getline(in, line)) throw std::runtime_error("restart file ends inside a section");
    for (char c : line)
        if (!std::isspace((unsigned char)c)) throw std::runtime_error("restart file: unread values in a section");
}

}  // namespace

void writeRestart(const char* path, long nx, long ny, long nz, const double* pv, const uint8_t* meaning, double time,
                  int timeStepIdx)
{
    std::ofstream out(path);
    if (!out.good()) throw std::runtime_error(std::string("cannot open restart file '") + path + "' for writing");
    out.precision(20);
    out << restartCookie(nx, ny, nz) << "\n" << "\n";
    // episode index, episode start time, episode length, start time, time, time step index
    out << "Simulator" << "\n";
    out << 0 << " " << 0.0 << " " << DBL_MAX << " " << 0.0 << " " << time << " " << timeStepIdx << " " << "\n";
    out << "Entities: Codim 0" << "\n";
    const long N = nx * ny * nz;
    for (long I = 0; I < N; ++I) {
        for (int e = 0; e < 3; ++e) out << pv[3 * I + e] << " ";
        // gas meaning (0 Sg, 1 Rs), water meaning (0 Sw), pressure meaning (0 Po), PVT region
        out << (int)meaning[I] << " " << 0 << " " << 0 << " " << 0 << " " << "\n";
    }
    out << "\n";
    if (!out.good()) throw std::runtime_error(std::string("writing restart file '") + path + "' failed");
}

void readRestart(const char* path, long nx, long ny, long nz, double* pv, uint8_t* meaning, double* time,
                 int* timeStepIdx)
{
    std::ifstream in(path);
    if (!in.good()) throw std::runtime_error(std::string("restart file '") + path + "' could not be opened");
    expectLine(in, restartCookie(nx, ny, nz), "the grid cookie");
    expectSectionEnd(in);
    expectLine(in, "Simulator", "the simulator section");
    {
        std::string line;
        if (!std::getline(in, line)) throw std::runtime_error("restart file ends inside the simulator section");
        std::istringstream ls(line);
        int episodeIdx = 0, step = 0;
        double episodeStart = 0, episodeLength = 0, startTime = 0, t = 0;
        if (!(ls >> episodeIdx >> episodeStart >> episodeLength >> startTime >> t >> step))
            throw std::runtime_error("restart file: malformed simulator section");
        if (time) *time = t;
        if (timeStepIdx) *timeStepIdx = step;
    }
    expectLine(in, "Entities: Codim 0", "the cell section");
    const long N = nx * ny * nz;
    std::string line;
    for (long I = 0; I < N; ++I) {
        if (!std::getline(in, line)) throw std::runtime_error("restart file is corrupted (too few cells)");
        std::istringstream ls(line);
        double x[3];
        unsigned mg = 0, mw = 0, mp = 0, region = 0;
        if (!(ls >> x[0] >> x[1] >> x[2] >> mg >> mw >> mp >> region))
            throw std::runtime_error("restart file: malformed line of cell " + std::to_string(I));
        if (mg > 1 || mw != 0 || mp != 0 || region != 0)
            throw std::runtime_error("restart file: cell " + std::to_string(I) +
                                     " uses primary-variable meanings or a PVT region this model does not have");
        for (int e = 0; e < 3; ++e) pv[3 * I + e] = x[e];
        meaning[I] = (uint8_t)mg;
    }
    expectSectionEnd(in);
}

void writeVtu(const char* path, long nx, long ny, long nz, double dx, double dy, double dz, int nfields,
              const char* const* names, const double* fields)
{
    std::ofstream out(path);
    if (!out.good()) throw std::runtime_error(std::string("cannot open VTK file '") + path + "' for writing");
    out.precision(17);
    const long N = nx * ny * nz;
    const long V = (nx + 1) * (ny + 1) * (nz + 1);
    out << "<?xml version=\"1.0\"?>\n"
        << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
        << "  <UnstructuredGrid>\n"
        << "    <Piece NumberOfPoints=\"" << V << "\" NumberOfCells=\"" << N << "\">\n"
        << "      <CellData>\n";
    for (int f = 0; f < nfields; ++f) {
        out << "        <DataArray type=\"Float64\" Name=\"" << names[f]
            << "\" NumberOfComponents=\"1\" format=\"ascii\">\n";
        const double* a = fields + (long)f * N;
        for (long I = 0; I < N; ++I) out << a[I] << ((I % 8 == 7 || I + 1 == N) ? "\n" : " ");
        out << "        </DataArray>\n";
    }
    out << "      </CellData>\n"
        << "      <Points>\n"
        << "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (long k = 0; k <= nz; ++k)
        for (long j = 0; j <= ny; ++j)
            for (long i = 0; i <= nx; ++i) out << i * dx << " " << j * dy << " " << k * dz << "\n";
    out << "        </DataArray>\n"
        << "      </Points>\n"
        << "      <Cells>\n"
        << "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
    auto vid = [&](long i, long j, long k) { return i + (nx + 1) * (j + (ny + 1) * k); };
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j)
            for (long i = 0; i < nx; ++i) {
                // VTK_HEXAHEDRON: the bottom face counter-clockwise, then the top face
                out << vid(i, j, k) << " " << vid(i + 1, j, k) << " " << vid(i + 1, j + 1, k) << " "
                    << vid(i, j + 1, k) << " " << vid(i, j, k + 1) << " " << vid(i + 1, j, k + 1) << " "
                    << vid(i + 1, j + 1, k + 1) << " " << vid(i, j + 1, k + 1) << "\n";
            }
    out << "        </DataArray>\n"
        << "        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
    for (long I = 0; I < N; ++I) out << 8 * (I + 1) << ((I % 16 == 15 || I + 1 == N) ? "\n" : " ");
    out << "        </DataArray>\n"
        << "        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n";
    for (long I = 0; I < N; ++I) out << 12 << ((I % 32 == 31 || I + 1 == N) ? "\n" : " ");
    out << "        </DataArray>\n"
        << "      </Cells>\n"
        << "    </Piece>\n"
        << "  </UnstructuredGrid>\n"
        << "</VTKFile>\n";
    if (!out.good()) throw std::runtime_error(std::string("writing VTK file '") + path + "' failed");
}

}  // namespace omodels
