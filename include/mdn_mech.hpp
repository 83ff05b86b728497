// mdn_mech.hpp -- header-only C++17 adapters over the C ABI with the reference's class shapes,
// for dropping the gfx950 path into MEDYAN's MController (src/Controller/MController.h:108-211).
//
//   mdn::ForceFieldManager  ~ medyan::ForceFieldManager (ForceFieldManager.h:64-123):
//       vectorizeAllForceFields, computeEnergy (+inf sentinel), computeIndividualEnergies,
//       computeForces (zero-fill + accumulate), assignallforcemags, computeLoadForces, plus
//       resetNeighborLists.
//   mdn::ForceField         ~ medyan::ForceField (ForceField.h:34-99): getName, vectorize,
//       computeEnergy (-1 sentinel), computeForces (accumulating), assignforcemags,
//       computeLoadForces, getNeighborLists; one adapter per force field on the device
//       (GpuFilamentStretching, GpuFilamentBending, GpuLinkerStretching, GpuMotorGhostStretching,
//       GpuCylinderExclVolume, GpuBoundaryCylinderRepulsion), from ForceFieldManager::forceFields.
//   mdn::CGMethod           ~ medyan::CGMethod::minimize (CGMethod.hpp:1746-1759) with the
//       ConjugateGradientParams fields (CGMethod.hpp:44-63).
//
// Errors raise mdn::Error (the reference throws std::runtime_error on the same conditions).
#pragma once

#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mdn_mech.h"

namespace mdn {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct EnergyReport {
    double total = 0;
    std::vector<std::pair<std::string, double>> individual;
};

inline const char* ff_name(int ff) {
    static const char* names[MDN_NUM_FF] = {"FilamentStretching", "FilamentBending", "LinkerStretching",
                                            "MotorGhostStretching", "CylinderExclVolume",
                                            "BoundaryCylinderRepulsion"};
    return (ff >= 0 && ff < MDN_NUM_FF) ? names[ff] : "?";
}

class ForceFieldManager;

// The reference's force-field interface (ForceField.h:34-99). The six force fields of a system
// live in one device context (vectorized together by ForceFieldManager::vectorizeAllForceFields
// and resetNeighborLists), so vectorize() here only checks that the context holds a system, and a
// force field's energy is read from the context's per-force-field energy evaluation (which
// evaluates all six at once: callers that need the total use ForceFieldManager::computeEnergy).
class ForceField {
public:
    virtual ~ForceField() = default;
    virtual std::string getName() = 0;
    virtual void vectorize() = 0;
    // this force field's energy at coord, or -1 when it is not finite even after the numerical
    // fallback (ForceField.h:55-58)
    virtual double computeEnergy(double* coord) = 0;
    // this force field's forces at coord ADDED to f (3 * n_beads values)
    virtual void computeForces(double* coord, double* f) = 0;
    virtual void computeLoadForces() {}
    virtual void assignforcemags() {}
    // the neighbor lists live on the device (ForceFieldManager::resetNeighborLists)
    virtual std::vector<void*> getNeighborLists() { return {}; }
};

class ForceFieldManager {
public:
    explicit ForceFieldManager(int device = 0) {
        int st = mdn_ctx_create(device, &ctx_);
        if (st != MDN_OK) throw Error(st, "mdn_ctx_create failed (no HIP device?)");
    }
    ~ForceFieldManager() { mdn_ctx_destroy(ctx_); }
    ForceFieldManager(const ForceFieldManager&) = delete;
    ForceFieldManager& operator=(const ForceFieldManager&) = delete;

    // serializeDof + vectorizeAllForceFields for the bonded and crosslink force fields
    void vectorizeAllForceFields(const mdn_system_desc& sys, const std::vector<double>& coord) {
        check(mdn_system_upload(ctx_, &sys, coord.data()));
        nvar_ = 3 * sys.n_beads;
    }
    // serializeDof (DofSerializer.hpp:64-92) on the device, for the next system's beads in
    // Bead::getBeads() order: src[i] = the bead's index in the current system, or -1 for a new
    // bead (its coordinates the next three values of newCoord); then vectorizeAllForceFields(sys)
    // uploads the next system with those device coordinates
    void serializeDof(const std::vector<int32_t>& src, const std::vector<double>& newCoord) {
        check(mdn_dof_serialize(ctx_, (int32_t)src.size(), src.data(), newCoord.empty() ? nullptr : newCoord.data()));
    }
    void vectorizeAllForceFields(const mdn_system_desc& sys) {
        check(mdn_system_upload(ctx_, &sys, nullptr));
        nvar_ = 3 * sys.n_beads;
    }
    // LinkerStretching / MotorGhostStretching::vectorize alone, between minimizations whose
    // chemistry only bound or unbound crosslinks (filaments, neighbor lists kept)
    void vectorizeCrosslinks(const std::vector<mdn_crosslink>& linkers, const std::vector<mdn_crosslink>& motors) {
        check(mdn_crosslinks_update(ctx_, linkers.data(), (int32_t)linkers.size(), motors.data(),
                                    (int32_t)motors.size()));
    }
    // SubSystem::resetNeighborLists (hybrid cylinder NL + boundary NL) and the excluded-volume /
    // boundary vectorization; returns the number of excluded-volume pairs
    long resetNeighborLists(const std::vector<double>& coord) {
        int64_t n = 0;
        check(mdn_nl_build(ctx_, coord.data(), &n));
        return (long)n;
    }
    // the same from the context's device coordinates (the last upload or minimization result)
    long resetNeighborLists() {
        int64_t n = 0;
        check(mdn_nl_build(ctx_, nullptr, &n));
        return (long)n;
    }
    // the device coordinates: read back, replace, displace by seeded Gaussian bead noise
    void getCoords(std::vector<double>& coord) {
        coord.resize((size_t)nvar_);
        check(mdn_get_coords(ctx_, coord.data()));
    }
    void setCoords(const std::vector<double>& coord) { check(mdn_set_coords(ctx_, coord.data())); }
    void perturb(uint64_t seed, uint64_t stream, double sigma) { check(mdn_perturb(ctx_, seed, stream, sigma)); }
    EnergyReport computeIndividualEnergies(const double* coord) {
        mdn_energy_report r{};
        check(mdn_energy(ctx_, coord, &r));
        EnergyReport out;
        out.total = r.total;
        for (int ff = 0; ff < MDN_NUM_FF; ff++) out.individual.emplace_back(ff_name(ff), r.ff[ff]);
        return out;
    }
    double computeEnergy(const double* coord) {
        mdn_energy_report r{};
        check(mdn_energy(ctx_, coord, &r));
        double e = 0;
        for (double v : r.ff) {
            if (v <= -1) return std::numeric_limits<double>::infinity();
            e += v;
        }
        return e;
    }
    void computeForces(const double* coord, double* force, int numVar) {
        check(mdn_forces(ctx_, coord, force, numVar));
    }
    // ForceField::computeForces of one force field (accumulates)
    void computeForces(int ff, const double* coord, double* force, int numVar) {
        check(mdn_ff_forces(ctx_, ff, coord, force, numVar));
    }
    // ForceFieldManager::assignallforcemags (ForceFieldManager.cpp:347-351): the linkers' and
    // motors' stretchForce after the last force evaluation, in vectorization order
    void assignallforcemags(std::vector<double>& linkerStretchForce, std::vector<double>& motorStretchForce) {
        int64_t counts[2 + MDN_NUM_FF];
        check(mdn_system_counts(ctx_, counts));
        linkerStretchForce.assign((size_t)counts[2 + MDN_FF_LINKER], 0.0);
        motorStretchForce.assign((size_t)counts[2 + MDN_FF_MOTOR], 0.0);
        check(mdn_crosslink_stretch_forces(ctx_, linkerStretchForce.data(), motorStretchForce.data()));
    }
    // ForceFieldManager::computeLoadForces (ForceFieldManager.cpp:292-310) at the current
    // coordinates: per filament f, loadForcesP[f * stride + i] of its plus-end bead and
    // loadForcesM[f * stride + i] of its minus-end bead, stride = max(cylinderNumMon);
    // monomerSize and cylinderNumMon are per filament type (SysParams::Geometry())
    int computeLoadForces(int nFilaments, const std::vector<double>& monomerSize, const std::vector<int>& cylinderNumMon,
                          std::vector<double>& loadForcesP, std::vector<double>& loadForcesM) {
        int stride = 0;
        for (int n : cylinderNumMon) stride = n > stride ? n : stride;
        loadForcesP.assign((size_t)nFilaments * stride, 0.0);
        loadForcesM.assign((size_t)nFilaments * stride, 0.0);
        std::vector<int32_t> nm(cylinderNumMon.begin(), cylinderNumMon.end());
        check(mdn_boundary_load_forces(ctx_, nullptr, (int32_t)monomerSize.size(), monomerSize.data(), nm.data(), stride,
                                       loadForcesP.data(), loadForcesM.data()));
        return stride;
    }
    mdn_ctx* ctx() const { return ctx_; }
    int numVar() const { return nvar_; }

    // the force fields in ForceFieldManager order (MController.cpp:70-178), as ForceField
    // adapters on this manager's context
    inline const std::vector<std::unique_ptr<ForceField>>& forceFields();
    // one force field's energy and (accumulated) forces
    double ffEnergy(int ff, const double* coord) {
        mdn_energy_report r{};
        check(mdn_energy(ctx_, coord, &r));
        return r.ff[ff];
    }
    void checkVectorized() {
        int64_t counts[2 + MDN_NUM_FF];
        check(mdn_system_counts(ctx_, counts));
        if (counts[0] <= 0) throw Error(MDN_ERR_STATE, "no system vectorized (vectorizeAllForceFields)");
    }

private:
    void check(int st) const {
        if (st != MDN_OK) throw Error(st, mdn_last_error(ctx_));
    }
    mdn_ctx* ctx_ = nullptr;
    int nvar_ = 0;
    std::vector<std::unique_ptr<ForceField>> ffs_;
};

// A force field of the device context (MDN_FF_*).
class DeviceForceField : public ForceField {
public:
    DeviceForceField(ForceFieldManager& m, int ff) : m_(m), ff_(ff) {}
    std::string getName() override { return ff_name(ff_); }
    void vectorize() override { m_.checkVectorized(); }
    double computeEnergy(double* coord) override { return m_.ffEnergy(ff_, coord); }
    void computeForces(double* coord, double* f) override { m_.computeForces(ff_, coord, f, m_.numVar()); }

protected:
    ForceFieldManager& m_;
    int ff_;
};

struct GpuFilamentStretching : DeviceForceField {
    explicit GpuFilamentStretching(ForceFieldManager& m) : DeviceForceField(m, MDN_FF_STRETCH) {}
};
struct GpuFilamentBending : DeviceForceField {
    explicit GpuFilamentBending(ForceFieldManager& m) : DeviceForceField(m, MDN_FF_BEND) {}
};
// LinkerStretching / MotorGhostStretching: assignforcemags reads each crosslink's stretchForce of
// the last force evaluation (LinkerStretchingHarmonic.cpp:291) into stretchForce().
struct GpuCrosslinkStretching : DeviceForceField {
    GpuCrosslinkStretching(ForceFieldManager& m, int ff) : DeviceForceField(m, ff) {}
    void assignforcemags() override {
        std::vector<double> l, mo;
        m_.assignallforcemags(l, mo);
        sf_ = ff_ == MDN_FF_LINKER ? l : mo;
    }
    const std::vector<double>& stretchForce() const { return sf_; }

private:
    std::vector<double> sf_;
};
struct GpuLinkerStretching : GpuCrosslinkStretching {
    explicit GpuLinkerStretching(ForceFieldManager& m) : GpuCrosslinkStretching(m, MDN_FF_LINKER) {}
};
struct GpuMotorGhostStretching : GpuCrosslinkStretching {
    explicit GpuMotorGhostStretching(ForceFieldManager& m) : GpuCrosslinkStretching(m, MDN_FF_MOTOR) {}
};
struct GpuCylinderExclVolume : DeviceForceField {
    explicit GpuCylinderExclVolume(ForceFieldManager& m) : DeviceForceField(m, MDN_FF_EXVOL) {}
};
// BoundaryCylinderRepulsion: computeLoadForces for filaments of one type with the given monomer
// size and monomers per cylinder (BoundaryCylinderRepulsion.cpp:122-187), into loadForcesP/M().
struct GpuBoundaryCylinderRepulsion : DeviceForceField {
    explicit GpuBoundaryCylinderRepulsion(ForceFieldManager& m, int nFilaments = 0, double monomerSize = 2.7,
                                          int cylinderNumMon = 40)
        : DeviceForceField(m, MDN_FF_BOUNDARY), nfil_(nFilaments), msize_(monomerSize), nmon_(cylinderNumMon) {}
    void computeLoadForces() override {
        if (nfil_ > 0) m_.computeLoadForces(nfil_, {msize_}, {nmon_}, p_, m_load_);
    }
    const std::vector<double>& loadForcesP() const { return p_; }
    const std::vector<double>& loadForcesM() const { return m_load_; }

private:
    int nfil_;
    double msize_;
    int nmon_;
    std::vector<double> p_, m_load_;
};

inline const std::vector<std::unique_ptr<ForceField>>& ForceFieldManager::forceFields() {
    if (ffs_.empty()) {
        ffs_.emplace_back(new GpuFilamentStretching(*this));
        ffs_.emplace_back(new GpuFilamentBending(*this));
        ffs_.emplace_back(new GpuLinkerStretching(*this));
        ffs_.emplace_back(new GpuMotorGhostStretching(*this));
        ffs_.emplace_back(new GpuCylinderExclVolume(*this));
        ffs_.emplace_back(new GpuBoundaryCylinderRepulsion(*this));
    }
    return ffs_;
}

// ConjugateGradientParams (CGMethod.hpp:44-63) with MController's values (MController.cpp:34-63)
struct ConjugateGradientParams {
    int descentSearchMethod = MDN_CG_POLAK_RIBIERE;
    double maxDist = 1.0;
    double lambdaMax = 0.01;
    double gradTol = 1.0;
    double energyRelTol = 1e-9;
    int energyCallLimit = 0;
    int forceCallLimit = 10000;
    bool reportOnCallLimit = true;
    bool tryToRecoverInLineSearchError = true;
    int precision = MDN_PREC_FP64;
    int lineSearchAlgorithm = MDN_LS_BACKTRACKING;   // "BACKTRACKING" / "QUADRATIC" (CGMethod.hpp:51)

    mdn_cg_params c() const {
        mdn_cg_params p{};
        p.method = descentSearchMethod;
        p.line_search = lineSearchAlgorithm;
        p.try_recover = tryToRecoverInLineSearchError;
        p.max_dist = maxDist;
        p.lambda_max = lambdaMax;
        p.grad_tol = gradTol;
        p.energy_rel_tol = energyRelTol;
        p.force_call_limit = forceCallLimit;
        p.energy_call_limit = energyCallLimit;
        p.report_on_call_limit = reportOnCallLimit;
        p.precision = precision;
        return p;
    }
};

struct MinimizationResult : mdn_min_result {
    bool success() const { return error_flags == 0; }
    bool converged() const { return converged_grad || converged_energy; }
};

class CGMethod {
public:
    // minimize from coord (in/out); forces of the final state into *outForce. The neighbor lists
    // of ffm stay frozen during the minimization, as in the reference.
    MinimizationResult minimize(ForceFieldManager& ffm, const ConjugateGradientParams& p, std::vector<double>& coord,
                                std::vector<double>* outForce) {
        MinimizationResult res{};
        mdn_cg_params cp = p.c();
        if (outForce) outForce->assign(coord.size(), 0.0);
        int st = mdn_minimize(ffm.ctx(), &cp, coord.data(), outForce ? outForce->data() : nullptr, &res);
        if (st == MDN_ERR_LINESEARCH) return res;   // MController runs its diagnosis and throws
        if (st != MDN_OK) throw Error(st, mdn_last_error(ffm.ctx()));
        return res;
    }
    // minimize from the context's device coordinates and leave the result there (no coordinate
    // transfer: successive minimizations between chemistry steps that change no topology)
    MinimizationResult minimize(ForceFieldManager& ffm, const ConjugateGradientParams& p) {
        MinimizationResult res{};
        mdn_cg_params cp = p.c();
        int st = mdn_minimize(ffm.ctx(), &cp, nullptr, nullptr, &res);
        if (st == MDN_ERR_LINESEARCH) return res;
        if (st != MDN_OK) throw Error(st, mdn_last_error(ffm.ctx()));
        return res;
    }
};

}  // namespace mdn
