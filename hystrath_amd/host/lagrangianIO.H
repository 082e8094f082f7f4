// dsmcFoam+'s lagrangian cloud files (<time>/lagrangian/dsmc/) in OpenFOAM
// ASCII format, the files a dsmcParcel cloud reads and writes
// (dsmcParcelIO.C:133-519, through Cloud<dsmcParcel> and IOPosition):
//
//   positions       "(x y z) celli" per parcel      (Cloud<passiveParticle>)
//   U               vectorField
//   ERot            scalarField
//   typeId          labelField (index into dsmcProperties typeIdList)
//   vibLevel        labelFieldField, "k(l0 ... lk-1)" per parcel
//   newParcel, classification, origProcId, origId   labelFields
//
// Lists are read in full form "N ( ... )" and uniform form "N{value}".
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "dictionary.H"

namespace hs {

// A cloud in the engine's structure-of-arrays parcel layout.
struct LagrangianCloud {
    std::vector<double> px, py, pz;
    std::vector<float> ux, uy, uz, erot;
    std::vector<int32_t> cell;
    std::vector<uint8_t> type;
    std::vector<uint16_t> vib;   // modes x n, one row per vibrational mode
    int modes = 0;
    int64_t size() const { return (int64_t)px.size(); }
    void resize(int64_t n, int modes);
};

// true if <timeDir>/lagrangian/dsmc/positions exists
bool hasLagrangianCloud(const std::string& timeDir);

// Reads <timeDir>/lagrangian/dsmc with vibLevel lists of up to `modes`
// entries (shorter lists are padded with zeros). ERot, typeId and vibLevel
// are zero when their file is absent; a field whose length differs from
// positions, a cell label outside [0, nCells), a typeId outside
// [0, nSpecies) or a vibLevel list longer than `modes` is a FatalError.
LagrangianCloud readLagrangianCloud(const std::string& timeDir, int modes, int64_t nCells, int nSpecies);

// Writes every file of the cloud; the files' location is
// "<timeName>/lagrangian/dsmc". Reals are written in the shortest form that
// reads back to the same double, so a write/read round trip is exact.
void writeLagrangianCloud(const std::string& timeDir, const std::string& timeName, const LagrangianCloud& c);

// internalField of a volScalarField file ("uniform v" or "nonuniform
// List<scalar> N (...)") as n values (dsmcSigmaTcRMax, dsmcCloud.C:625).
std::vector<double> readVolScalarInternalField(const std::string& path, int64_t n);

}  // namespace hs
