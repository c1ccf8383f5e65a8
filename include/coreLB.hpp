// Drop-in counterpart of the reference's include/coreLB.hpp (class LB2D,
// coreLB.hpp:57-227).  Every method runs on the GPU through the C ABI of
// mantiis_hip.h; the public population and moment vectors are device fields
// with the reference's names.  Assigning one field to another
// (usq_old = usq, rho_old = rho, ftemp = f, fb = f) is a device copy, element
// reads download the field once per change, and assigning a std::vector
// uploads it.  Cell order is the reference's compact order.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "globalParameters.hpp"
#include "grid.hpp"
#include "mantiis_hip.h"
#include "shape.hpp"
#include "utils.hpp"

class LB2D;

// A device-resident LB2D vector.
class DeviceField {
public:
    DeviceField(LB2D* owner, int id) : owner_(owner), id_(id) {}
    DeviceField& operator=(const DeviceField& o);
    DeviceField& operator=(const std::vector<double>& v);
    double operator[](size_t i) const { return host()[i]; }
    size_t size() const;
    const std::vector<double>& host() const;
    operator std::vector<double>() const { return host(); }
    std::vector<double>::const_iterator begin() const { return host().begin(); }
    std::vector<double>::const_iterator end() const { return host().end(); }
    int id() const { return id_; }

private:
    LB2D* owner_;
    int id_;
    mutable std::vector<double> cache_;
    mutable uint64_t version_ = ~0ull;
};

class LB2D {
public:
    using MethodPtr = void (LB2D::*)(int);
    MethodPtr collisionMethod = nullptr;
    MethodPtr wallBoundaryMethod = nullptr;
    MethodPtr openBoundaryMethod = nullptr;

    int NX, NY;
    double resolution = _GLOBAL_::Cl;
    int t = 0;
    lattice latt;
    const int NC = 9;
    const std::vector<double> W{4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                                1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

    DeviceField f{this, MTS_F_F}, fb{this, MTS_F_FB}, ftemp{this, MTS_F_FTEMP}, feq{this, MTS_F_FEQ},
        fneq{this, MTS_F_FNEQ}, rho{this, MTS_F_RHO}, rho_old{this, MTS_F_RHO_OLD}, ux{this, MTS_F_UX},
        uy{this, MTS_F_UY}, usq{this, MTS_F_USQ}, usq_old{this, MTS_F_USQ_OLD}, psi{this, MTS_F_PSI},
        Fx{this, MTS_F_FX}, Fy{this, MTS_F_FY}, taus{this, MTS_F_TAUS};

    double diff = 1.0;
    std::vector<double> diff_over_time;
    Grid2D& grid;

    // coreLB.hpp:231-378 (+ the device state)
    LB2D(int& x, int& y, lattice& l, Grid2D& G, Shape& shape) : NX(x), NY(y), latt(l), grid(G) {
        mts_params p;
        mts_default_params(&p);
        p.dx = _GLOBAL_::Cl;
        p.Fbody = _GLOBAL_::Fbody;
        p.T = _GLOBAL_::T_phy;
        p.P = _GLOBAL_::P_phy;
        p.Gff = _GLOBAL_::Gff;
        p.Gfs = _GLOBAL_::Gfs;
        p.precision = precision_from_env();
        p.periodic_y = mantiis_torus_from_env();
        p.multiblock = G.isMultigrid ? mantiis_levels_from_env() : 0;
        const std::vector<int32_t> pore = shape.flatDomain();
        const std::vector<double> kn = shape.flat(shape.Kn), lp = shape.flat(shape.LocPore);
        mts_sim h = nullptr;
        check(mts_create(shape.Nx, shape.Ny, pore.data(), kn.data(), lp.data(), &p, &h));
        h_ = std::shared_ptr<mts_sim_s>(h, [](mts_sim s) { mts_destroy(s); });
        grid.mirror(h);
    }

    LB2D(const LB2D&) = delete;
    LB2D& operator=(const LB2D&) = delete;

    mts_sim handle() const { return h_.get(); }
    uint64_t version() const { return version_; }
    void touched() { version_++; }
    static void check(int rc) {
        if (rc != MTS_OK) throw std::runtime_error(mts_last_error());
    }

    // coreLB.hpp:381-417
    void initialize() {
        check(mts_initialize(handle()));
        touched();
    }
    // coreLB.hpp:420-430 and the per-method operators
    void equilibrium() { run(MTS_M_EQUILIBRIUM); }
    void regularize() { run(MTS_M_REGULARIZE); }
    void SRTCollision(int lvl = 1) { run(MTS_M_SRT, lvl); }
    void MRTCollisionNonRegularized(int lvl = 1) { run(MTS_M_MRT, lvl); }
    void MRTCollisionRegularized(int lvl = 1) { run(MTS_M_RMRT, lvl); }
    void setCollision(MethodPtr m) {
        collisionMethod = m;
        sync_operators();
    }
    void setWallBoundary(MethodPtr m) {
        wallBoundaryMethod = m;
        sync_operators();
    }
    void setOpenBoundary(MethodPtr m) { openBoundaryMethod = m; }
    void collide(int lvl = 1) { (this->*collisionMethod)(lvl); }
    void stream(int lvl = 1) { run(MTS_M_STREAM, lvl); }
    void BBWall(int lvl = 1) { run(MTS_M_BBWALL, lvl); }
    void SRBBWall(int lvl = 1) { run(MTS_M_SRBBWALL, lvl); }
    void MDBBWall(int lvl = 1) { run(MTS_M_MDBBWALL, lvl); }
    void periodicBoundary(int lvl = 1) { run(MTS_M_PERIODIC, lvl); }
    void generalizedPeriodicBoundaryAdsorption(int = 1) { run(MTS_M_GENPERIODIC); }
    void generalizedPeriodicBoundaryTransport(int = 1) {}
    void applyBoundaryConditions(int lvl = 1) {
        (this->*wallBoundaryMethod)(lvl);
        (this->*openBoundaryMethod)(lvl);
    }
    void calculateMacroscopicPorperties() { run(MTS_M_MACROSCOPIC); }
    void calculateDensity() { run(MTS_M_DENSITY); }
    void calculateVelocity() { run(MTS_M_VELOCITY); }
    void getPseudoPotential() { run(MTS_M_PSEUDOPOTENTIAL); }
    void calculateForces() { run(MTS_M_FORCES); }
    void applySourceTerm() { run(MTS_M_SOURCE); }
    void zeroDown(int lvl = 1) { run(MTS_M_ZERODOWN, lvl); }
    void calculateVelocityDifference() {
        run(MTS_M_VELOCITY_DIFF);
        record_residual();
    }
    void calculateDensityDifference() {
        run(MTS_M_DENSITY_DIFF);
        record_residual();
    }
    // coreLB.hpp:1101-1231: the buffer cells of levels l1 and l2 = l1 + 1
    void interpolateBlockInterface(int l1, int l2) {
        if (l2 != l1 + 1) throw std::invalid_argument("interpolateBlockInterface: levels l1 and l1 + 1");
        run(MTS_M_INTERPOLATE, l1);
    }
    // coreLB.hpp:1310-1331 (fb = f, collide, ftemp = f, zeroDown, stream, BC)
    void evolutionStepOfMultiBlock(int lvl) {
        const int rep = 1 << (grid.maxLevel - lvl);
        for (int i = 0; i < rep; i++) {
            fb = f;
            collide(lvl);
            ftemp = f;
            zeroDown(lvl);
            stream(lvl);
            applyBoundaryConditions(lvl);
        }
        if (lvl > 1) evolutionStepOfMultiBlock(lvl - 1);
        if (lvl != grid.maxLevel) interpolateBlockInterface(lvl, lvl + 1);
    }

    // Fused whole steps (one HIP pass per step): n steps of the transport,
    // adsorption or coupled loop of MantiisApp; the residual of the last step
    // is appended to diff_over_time.
    void stepTransport(int n) { fused(mts_step_transport, n); }
    void stepAdsorption(int n) { fused(mts_step_adsorption, n); }
    void stepCoupled(int n) { fused(mts_step_coupled, n); }
    void setBodyForce(double F) { check(mts_set_body_force(handle(), F)); }

    // coreLB.hpp:1344-1411
    std::tuple<std::vector<double>, std::vector<double>, std::vector<double>> completeSingleGridDomain() {
        const size_t n = (size_t)NX * NY;
        std::vector<double> a(n), b(n), c(n);
        check(mts_complete_domain(handle(), a.data(), b.data(), c.data()));
        return std::make_tuple(a, b, c);
    }
    // FP64 permeability report (no reference counterpart): Darcy velocity
    // U_D = sum u_x / (NX NY) over completeSingleGridDomain's pixels, mean u_x
    // of the fluid pixels, sum rho u_x, sum rho, porosity, the finest level's
    // mean tau_s, k = nu U_D / Fbody (lattice units, nu = (tau - 0.5) / 3) and
    // k Cl^2 [m^2].
    struct Permeability {
        double darcy_velocity, mean_ux, flux, mass, porosity, tau, k_lattice, k_m2;
    };
    Permeability permeability() {
        double o[8];
        check(mts_permeability(handle(), o));
        return Permeability{o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7]};
    }

    // coreLB.hpp:1619-1641: f re-expanded to NX*NY*9 with zeros at solids.
    std::vector<double> prepareDistributions() {
        const std::vector<double>& fc = f.host();
        std::vector<double> out((size_t)NX * NY * NC, 0.0);
        for (size_t c = 0; c < grid.gridIJ.size(); c++) {
            const size_t p = (size_t)grid.gridIJ[c][1] * NX + grid.gridIJ[c][0];
            for (int ic = 0; ic < NC; ic++) out[p * NC + ic] = fc[c * NC + ic];
        }
        return out;
    }

    // Inverse of prepareDistributions (no reference counterpart): sets f from
    // an NX*NY*9 pixel-major vector (a dump of f.txt / f.bin) and recomputes
    // the moments from it (the loops take usq_old, rho_old from them), which
    // is the whole state a step continues from -- so a run resumed from a
    // lossless dump continues exactly as the uninterrupted one.  Single-block
    // lattices only (a multi-block dump keeps one cell per pixel).
    void loadDistributions(const std::vector<double>& fGlobal) {
        if (grid.maxLevel > 1) throw std::invalid_argument("loadDistributions: single-block lattices only");
        if (fGlobal.size() != (size_t)NX * NY * NC) throw std::invalid_argument("loadDistributions: expected NX*NY*9 values");
        std::vector<double> fc(grid.gridIJ.size() * NC);
        for (size_t c = 0; c < grid.gridIJ.size(); c++) {
            const size_t p = (size_t)grid.gridIJ[c][1] * NX + grid.gridIJ[c][0];
            for (int ic = 0; ic < NC; ic++) fc[c * NC + ic] = fGlobal[p * NC + ic];
        }
        f = fc;
        calculateMacroscopicPorperties();  // rho, ux, uy, usq of the loaded f
    }

private:
    std::shared_ptr<mts_sim_s> h_;
    uint64_t version_ = 0;

    static int precision_from_env() {
        const char* e = std::getenv("MANTIIS_PRECISION");
        return (e && std::string(e) == "fp32") ? MTS_FP32DEV : MTS_FP64;
    }
    void run(int method, int lvl = 1) {
        check(mts_method(handle(), method, lvl));
        touched();
    }
    void record_residual() {
        check(mts_residual(handle(), &diff));
        diff_over_time.push_back(diff);
    }
    void fused(int (*fn)(mts_sim, int), int n) {
        check(fn(handle(), n));
        touched();
        record_residual();
    }
    int code_of(MethodPtr m, bool wall) const {
        if (!wall) {
            if (m == &LB2D::MRTCollisionRegularized) return MTS_COLL_RMRT;
            if (m == &LB2D::MRTCollisionNonRegularized) return MTS_COLL_MRT;
            return MTS_COLL_SRT;
        }
        if (m == &LB2D::SRBBWall) return MTS_WALL_SRBB;
        if (m == &LB2D::MDBBWall) return MTS_WALL_MDBB;
        return MTS_WALL_BB;
    }
    void sync_operators() {
        const int c = collisionMethod ? code_of(collisionMethod, false) : MTS_COLL_RMRT;
        const int w = wallBoundaryMethod ? code_of(wallBoundaryMethod, true) : MTS_WALL_SRBB;
        check(mts_set_operators(handle(), c, w));
    }
};

inline size_t DeviceField::size() const {
    int64_t n = 0;
    LB2D::check(mts_field_size(owner_->handle(), id_, &n));
    return (size_t)n;
}
inline const std::vector<double>& DeviceField::host() const {
    if (version_ != owner_->version()) {
        cache_.resize(size());
        LB2D::check(mts_get_field(owner_->handle(), id_, cache_.data()));
        version_ = owner_->version();
    }
    return cache_;
}
inline DeviceField& DeviceField::operator=(const DeviceField& o) {
    if (this == &o) return *this;
    const int src = o.id_, dst = id_;
    int m = 0;
    if (src == MTS_F_F && dst == MTS_F_FTEMP) m = MTS_M_COPY_F_FTEMP;
    else if (src == MTS_F_F && dst == MTS_F_FB) m = MTS_M_COPY_F_FB;
    else if (src == MTS_F_USQ && dst == MTS_F_USQ_OLD) m = MTS_M_COPY_USQ_OLD;
    else if (src == MTS_F_RHO && dst == MTS_F_RHO_OLD) m = MTS_M_COPY_RHO_OLD;
    if (m) LB2D::check(mts_method(owner_->handle(), m, 1));
    else {
        std::vector<double> v = o.host();
        LB2D::check(mts_set_field(owner_->handle(), id_, v.data()));
    }
    owner_->touched();
    return *this;
}
inline DeviceField& DeviceField::operator=(const std::vector<double>& v) {
    if (v.size() != size()) throw std::invalid_argument("field size mismatch");
    LB2D::check(mts_set_field(owner_->handle(), id_, v.data()));
    owner_->touched();
    return *this;
}
