// dsmcCloud.H -- host-side mirror of hyStrath's dsmcCloud for the MI355X
// engine: reads a dsmcFoam+ case directory, selects the models by the names
// the reference registers, and drives the engine through the C ABI
// (include/hystrath_dsmc.h).
//
// Reference interface mirrored: dsmcCloud (src/lagrangian/dsmc/clouds/
// dsmcCloud.H:87-188, dsmcCloud.C:586-691 constructor, :819-926 evolve,
// :935-985 info), BinaryCollisionModel::New (BinaryCollisionModel.C:59-90),
// collisionPartnerSelection::New (collisionPartnerSelection.C:64-97),
// dsmcBoundaries (dsmcBoundaries.C:105-119), dsmcFieldProperties /
// dsmcVolFields (dsmcFieldProperties.C:52-252, dsmcVolFields.C:1071-2382),
// dsmcMeshFill (dsmcMeshFill.C:68-~230).
#pragma once

#include <string>
#include <vector>

#include "../../include/hystrath_dsmc.h"
#include "blockMesh.H"
#include "dictionary.H"
#include "lagrangianIO.H"

namespace hs {

// A dsmcVolFields model of system/fieldPropertiesDict
struct FieldModel {
    std::string fieldName;
    std::vector<int> typeIds;
    bool measureMeanFreePath = true;
    bool measureHeatFluxShearStress = false;
    bool resetAtOutput = true;
    double resetAtOutputUntilTime = 1e300;
};

// Exact fluid volume of every cell with an embedded cylinder (axis z) removed.
std::vector<double> cellVolumes(const BlockMesh& m, bool cylinder, double cx, double cy, double r);

class dsmcCloud {
public:
    // dsmcCloud(Time&, const word&, const dynamicFvMesh&) for dsmcFoam+ when
    // initialise is false; for dsmcInitialise+ when true (no state is read)
    dsmcCloud(const std::string& caseDir, bool initialise = false);
    ~dsmcCloud();
    dsmcCloud(const dsmcCloud&) = delete;
    dsmcCloud& operator=(const dsmcCloud&) = delete;

    // dsmcInitialise+: every dsmcMeshFill configuration of dsmcInitialiseDict
    void initialise();
    // one time step (dsmcCloud::evolve)
    void evolve();
    // terminal summary of the last step (dsmcCloud::info)
    void info();
    int nTerminalOutputs() const { return nTerminalOutputs_; }

    // time control from system/controlDict
    double time() const { return time_; }
    double deltaT() const { return cfg_.dt; }
    double endTime() const { return endTime_; }
    bool loop();                 // runTime.loop(): advance the time if not past endTime
    bool writeTime() const;      // runTime.outputTime()
    std::string timeName() const;
    // fields of every dsmcVolFields model + a restart checkpoint into <case>/<timeName>
    void write();

    void writeCheckpoint(const std::string& path) const;
    void readCheckpoint(const std::string& path);
    // the engine's parcels, and the lagrangian cloud files of <timeDir>
    LagrangianCloud downloadCloud() const;
    void writeLagrangian(const std::string& timeDir, const std::string& timeName) const;
    // start from <timeDir>'s lagrangian cloud and dsmcSigmaTcRMax field
    void readLagrangian(const std::string& timeDir, double t, LagrangianCloud& c);
    int64_t nParticles() const;
    int vibModes() const;         // most vibrational modes of any species
    bool rotational() const;      // any species with rotational DoF

    hs_dsmc_engine* engine() { return eng_; }
    const hs_config& config() const { return cfg_; }
    const std::vector<std::string>& typeIdList() const { return typeIdList_; }
    const std::vector<FieldModel>& fieldModels() const { return fields_; }
    const std::vector<double>& cellVolumes() const { return vol_; }
    const BlockMesh& mesh() const { return mesh_; }

private:
    void readProperties(const dictionary& d);
    void readBoundaries(const dictionary& d);
    void readFields(const dictionary& d);
    void readReactions(const dictionary& d);
    void check(int rc, const char* what) const;
    // latest time directory of the case holding <time>/<rel>, and its time
    std::string latestTimeWith(const std::string& rel, double& t) const;

    std::string case_;
    hs_config cfg_{};
    BlockMesh mesh_;
    std::vector<double> vol_;
    std::vector<std::string> typeIdList_;
    std::vector<hs_reaction> reactions_;  // system/chemReactDict
    std::vector<FieldModel> fields_;
    hs_dsmc_engine* eng_ = nullptr;
    int nTerminalOutputs_ = 1;
    double time_ = 0.0, startTime_ = 0.0, endTime_ = 0.0, writeInterval_ = 0.0;
    bool writeByStep_ = false;
    long long stepIndex_ = 0;
    int precision_ = 10;
    std::vector<double> prevSums_;
    long long prevSampled_ = 0;
    std::string cylPatch_ = "cylinder";   // patch name of the embedded cylinder's wall faces
};

// Derived dsmcVolFields quantities of one field model from the cumulative
// per-species sums (dsmcVolFields.C:1242-1790).
struct VolFields {
    std::vector<double> rhoN, rhoM, Ttra, Trot, Tvib, Tov, p, Ma, mfp, mcr, measuredCollisionRate, dsmcNMean;
    std::vector<double> Ux, Uy, Uz;
};
VolFields computeVolFields(const dsmcCloud& cloud, const std::vector<double>& sums, long long nSteps,
                           const FieldModel& fm);

// Wall-face values of one field model (dsmcVolFields.C:1890-2126) from the
// engine's wall-face sums [species*HS_WS_COUNT + slot][face]; normal is
// [3*face + axis], pointing out of the gas.
struct WallFields {
    std::vector<double> rhoN, rhoM, Ttra, Trot, Tvib, Tov, p, tau, q;
    std::vector<double> U, fD;   // 3 per face
};
WallFields computeWallFields(const dsmcCloud& cloud, const std::vector<double>& wallSums, int64_t nFaces,
                             const std::vector<double>& normal, long long nSteps, const FieldModel& fm);

}  // namespace hs
