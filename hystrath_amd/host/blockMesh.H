// The engine's mesh: one axis-aligned block of uniform hex cells, numbered
// i + nx*(j + ny*k) as blockMesh numbers a single hex block, with one patch
// on each of its six sides.  It is read from constant/polyMesh when the case
// holds one (as OpenFOAM's fvMesh reads every case, polyMesh.C) and from
// system/blockMeshDict otherwise.
#pragma once

#include <string>
#include <vector>

#include "dictionary.H"

namespace hs {

struct BlockMesh {
    int n[3] = { 1, 1, 1 };
    double lo[3] = { 0, 0, 0 }, hi[3] = { 1, 1, 1 };
    std::string sidePatch[6];      // patch name of each side (hs_side order)
    std::string sideType[6];       // its polyPatch type (wall, patch, cyclic, empty, symmetryPlane)
    std::vector<std::string> patchNames;
};

// The block of system/blockMeshDict (one hex, simpleGrading (1 1 1)).
BlockMesh readBlockMesh(const dictionary& blockMeshDict);

// true if <caseDir>/constant/polyMesh holds points, faces, owner, neighbour
// and boundary
bool hasPolyMesh(const std::string& caseDir);

// The block that <caseDir>/constant/polyMesh describes.  The mesh is
// accepted when its cells are hexes whose centres lie on a uniform lattice,
// numbered as blockMesh numbers one block, and every boundary patch lies on
// whole sides of the block with each side held by one patch; otherwise a
// FatalError names what does not fit.  patchNames lists the boundary file's
// patches in its order.
BlockMesh readPolyMesh(const std::string& caseDir);

}  // namespace hs
