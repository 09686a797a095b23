// sdfibm_hip.hpp -- header-only C++ mirror of the reference's SolidCloud
// coupling interface over the C ABI of sdfibm_hip.h.
//
// The reference's solver calls `solidcloud.interact(t, dt)` once per step
// (src/main.cpp:66 -> src/solidcloud.cpp:435) and reads each solid's fluid
// force and torque afterwards (src/solid.h:64-65, 136-140).  This class keeps
// that shape: construct once with the mesh, hand over solids and U, call
// interact(time, dt), read forces/torques and the fields.  Errors throw
// sdfibm::hip::Error (the C ABI itself never throws).
#ifndef SDFIBM_HIP_HPP
#define SDFIBM_HIP_HPP

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "sdfibm_hip.h"

namespace sdfibm {
namespace hip {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

class SolidCloud {
public:
    SolidCloud(const sdfibm_mesh_desc& mesh, const sdfibm_config& cfg) {
        const int st = sdfibm_create(&mesh, &cfg, &ctx_);
        if (st != SDFIBM_OK) throw Error(st, std::string("sdfibm_create: ") + sdfibm_last_error(nullptr));
        n_cells_ = (long long)mesh.nx * mesh.ny * mesh.nz;
    }
    ~SolidCloud() {
        if (ctx_) sdfibm_destroy(ctx_);
    }
    SolidCloud(const SolidCloud&) = delete;
    SolidCloud& operator=(const SolidCloud&) = delete;

    void setSolids(const std::vector<sdfibm_solid>& solids) {
        check(sdfibm_set_solids(ctx_, solids.data(), (int64_t)solids.size()), "set_solids");
        ft_.assign(solids.size() * 6, 0.0);
    }
    // U as OpenFOAM stores a volVectorField's internal field: n_cells x 3 doubles.
    void setVelocity(const double* U) { check(sdfibm_set_velocity(ctx_, U), "set_velocity"); }

    // SolidCloud::interact(time, dt)
    void interact(double time, double dt) { check(sdfibm_interact(ctx_, time, dt, ft_.data()), "interact"); }

    std::array<double, 3> fluidForce(size_t s) const { return {ft_[6 * s], ft_[6 * s + 1], ft_[6 * s + 2]}; }
    std::array<double, 3> fluidTorque(size_t s) const { return {ft_[6 * s + 3], ft_[6 * s + 4], ft_[6 * s + 5]}; }

    // Host copies of As, Fs (AoS), Ts, Ct (as doubles), like the reference's fields.
    void fields(std::vector<double>* As, std::vector<double>* Fs, std::vector<double>* Ts,
                std::vector<double>* Ct) const {
        if (As) As->resize(n_cells_);
        if (Fs) Fs->resize(3 * n_cells_);
        if (Ts) Ts->resize(n_cells_);
        if (Ct) Ct->resize(n_cells_);
        check(sdfibm_get_fields(ctx_, As ? As->data() : nullptr, Fs ? Fs->data() : nullptr,
                                Ts ? Ts->data() : nullptr, Ct ? Ct->data() : nullptr),
              "get_fields");
    }

    sdfibm_ctx* handle() const { return ctx_; }

private:
    void check(int st, const char* what) const {
        if (st != SDFIBM_OK) throw Error(st, std::string(what) + ": " + sdfibm_last_error(ctx_));
    }
    sdfibm_ctx* ctx_ = nullptr;
    long long n_cells_ = 0;
    std::vector<double> ft_;
};

}  // namespace hip
}  // namespace sdfibm

#endif  // SDFIBM_HIP_HPP
