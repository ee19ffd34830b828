// BD simulation box driven by config / model / XYZ files: the counterpart of
// BDSimBox(config_file, model_file, xyz_file) (src/simbox/simbox_bd.h,
// src/simbox/simbox_bd.cc:31-294) on the device BD step.
#pragma once
#include <cstdio>
#include <memory>
#include <string>
#include "mob_gpu.hpp"
#include "molecule_io.hpp"

namespace sdt {

struct SimBoxParams {
    // mobility (InitMob, simbox_bd.cc:88-124): SPME or the dense Ewald sum
    bool ewald = false;            // mob-method ewald (MobEwald)
    double ewald_tol = 1e-4;
    double spme_xi = 0.5, spme_rmax = 4.0;
    int spme_dim = 128, spme_porder = 4, mob_interval = 100;
    // Brownian displacements (InitBrwn, :127-182)
    int lanczos_maxiters = 50, lanczos_maxnrhs = 20;
    double lanczos_tol = 1.0e-3;
    bool random = false;           // brownian-method random (BrwnRandom)
    bool chol = false;             // brownian-method chol (BrwnChol, dense mobility only)
    int random_numvecs = 30, random_maxnrhs = 50;
    // forces (InitForces, :185-234)
    bool steric = false;
    double steric_r0 = 0.0, steric_k0 = 0.0;
    // control (Init, :237-289)
    double delta_t = 0.001;
    int traj_interval = -1, traj_start = 0;
    std::string traj_output = "./trajectory.out";
    // additions of this library: arithmetic and Gaussian-stream seed
    int precision = 0;             // SDT_FP64 (reference arithmetic) or SDT_FP32
    uint64_t seed = 1234;          // RndStream(1234) in the reference
};

class SimBoxGpu {
  public:
    // parses the three files (host only); throws Error on invalid input
    SimBoxGpu(const char* config_file, const char* model_file, const char* xyz_file);
    ~SimBoxGpu();
    void init();                  // device state, trajectory file
    void advance(int nsteps);     // BDSimBox::Advance with trajectory frames
    void write_xyz(const char* path) const;   // a restart frame of the current state
    const SimBoxParams& params() const { return prm_; }
    int npos() const { return io_.npos(); }
    long long step() const { return step_; }
    double time() const { return time_; }
    void positions(double* out) const;
    void radii(double* out) const;
    const std::vector<std::string>& warnings() const { return io_.warnings(); }
    BDGpu* bd() { return bd_.get(); }

  private:
    MoleculeIO io_;
    SimBoxParams prm_;
    std::vector<double> pos_;     // host copy (before init, and after each advance)
    long long step_ = 0, start_step_ = 0;
    double time_ = 0.0, start_time_ = 0.0;
    std::unique_ptr<BDGpu> bd_;
    FILE* traj_ = nullptr;
};

}  // namespace sdt
