// BD simulation box from config / model / XYZ files (BDSimBox,
// src/simbox/simbox_bd.cc:31-366) over the device BD step.
#include "simbox.hpp"
#include <cfloat>
#include <climits>
#include <cstdio>
#include <cstring>

namespace sdt {

SimBoxGpu::SimBoxGpu(const char* config_file, const char* model_file, const char* xyz_file) {
    SDT_REQUIRE(config_file && model_file && xyz_file, SDT_EINVAL, "null file name");
    // ImportMolecule (simbox_bd.cc:62-85)
    SDT_REQUIRE(io_.parse_config(config_file), SDT_EINVAL, io_.err());
    SDT_REQUIRE(io_.parse_model(model_file), SDT_EINVAL, io_.err());
    SDT_REQUIRE(io_.parse_xyz(xyz_file), SDT_EINVAL, io_.err());
    SDT_REQUIRE(io_.box[0] == io_.box[1] && io_.box[1] == io_.box[2], SDT_EINVAL,
                "the periodic box must be cubic (Lx = Ly = Lz)");
    // InitMob (:88-124): "spme", else the dense Ewald sum (the reference's
    // default for any other value, :91-98)
    const char* mob = io_.get_string_key("mob-method");
    prm_.ewald = !(mob && strcmp(mob, "spme") == 0);
    if (prm_.ewald) {
        prm_.ewald_tol = io_.get_float_key("ewald-tol", 0.0, DBL_MAX, 1e-4);
    } else {
        prm_.spme_xi = io_.get_float_key("spme-xi", 0.0, DBL_MAX, 0.5);
        prm_.spme_rmax = io_.get_float_key("spme-rmax", 0.0, DBL_MAX, 4.0);
        prm_.spme_dim = io_.get_int_key("spme-nfft-grids", 1, 1024, 128);
        prm_.spme_porder = io_.get_int_key("spme-porder", 4, 6, 4);
    }
    prm_.mob_interval = io_.get_int_key("mob-interval", 1, INT_MAX, 100);
    // InitBrwn (:127-182): "lanczos", "random", else Cholesky (the
    // reference's default, :130-139), which needs the dense mobility
    const char* brwn = io_.get_string_key("brownian-method");
    prm_.random = brwn && strcmp(brwn, "random") == 0;
    prm_.chol = !(brwn && (strcmp(brwn, "lanczos") == 0 || prm_.random));
    SDT_REQUIRE(!(prm_.chol && !prm_.ewald), SDT_EINVAL,
                "Cholesky requires using the explicit matrix approaches (mob-method ewald)");
    SDT_REQUIRE(!(prm_.random && prm_.ewald), SDT_EINVAL,
                "brownian-method random runs on the SPME mobility here");
    if (prm_.random) {
        prm_.random_numvecs = io_.get_int_key("random-numvecs", 1, INT_MAX, 30);
        prm_.random_maxnrhs = io_.get_int_key("random-maxnrhs", 1, INT_MAX, 50);
    } else if (!prm_.chol) {
        prm_.lanczos_maxiters = io_.get_int_key("lanczos-maxiters", 1, INT_MAX, 50);
        prm_.lanczos_maxnrhs = io_.get_int_key("lanczos-maxnrhs", 1, INT_MAX, 20);
        prm_.lanczos_tol = io_.get_float_key("lanczos-tol", 0.0, DBL_MAX, 1.0e-3);
    }
    // InitForces (:185-234): steric; bonded forces are not on this path
    SDT_REQUIRE(io_.line_count("bond") == 0, SDT_EINVAL, "bond lines are not supported");
    const int nst = io_.line_count("steric");
    SDT_REQUIRE(nst <= 1, SDT_EINVAL, "duplicate \"steric\" lines");
    if (nst == 1) {
        double r0 = 0.0, k0 = 0.0;
        SDT_REQUIRE(sscanf(io_.get_line("steric", 0), "%le %le", &r0, &k0) == 2, SDT_EINVAL,
                    "invalid steric line");
        prm_.steric = true;
        prm_.steric_r0 = r0;
        prm_.steric_k0 = k0;
    }
    // Init (:237-289)
    prm_.delta_t = io_.get_float_key("delta-t", 0.0, DBL_MAX, 0.001);
    prm_.traj_interval = io_.get_int_key("traj-interval", INT_MIN, INT_MAX, -1);
    prm_.traj_start = io_.get_int_key("traj-startstep", 0, INT_MAX, 0);
    if (const char* t = io_.get_string_key("traj-output")) prm_.traj_output = t;
    const char* pr = io_.get_string_key("precision");
    if (pr) {
        SDT_REQUIRE(strcmp(pr, "fp64") == 0 || strcmp(pr, "fp32") == 0, SDT_EINVAL,
                    "precision must be fp64 or fp32");
        prm_.precision = strcmp(pr, "fp32") == 0 ? SDT_FP32 : SDT_FP64;
    }
    if (const char* sd = io_.get_string_key("seed")) prm_.seed = strtoull(sd, nullptr, 10);
    if (io_.has_seed) prm_.seed = io_.seed;   // a restart frame continues its stream
    pos_ = io_.pos();
    step_ = start_step_ = io_.start_step;
    time_ = start_time_ = io_.start_time;
}

SimBoxGpu::~SimBoxGpu() {
    if (traj_) fclose(traj_);
}

void SimBoxGpu::init() {
    SDT_REQUIRE(!bd_, SDT_ESTATE, "init called twice");
    const int n = io_.npos();
    int maxnrhs = prm_.lanczos_maxnrhs;
    if (!prm_.random && !prm_.chol && maxnrhs > 16) {
        // the device's groups are at most 16 columns wide
        std::fprintf(stderr,
                     "WARNING: lanczos-maxnrhs %d > 16: the displacements are computed in groups "
                     "of 16 columns (the reference's groups of %d give other samples of the same "
                     "distribution)\n",
                     maxnrhs, maxnrhs);
        maxnrhs = 16;
    }
    if (prm_.ewald)
        bd_.reset(new BDGpu(n, pos_.data(), io_.rdi().data(), io_.box[0], prm_.ewald_tol,
                            prm_.chol, prm_.lanczos_maxiters, maxnrhs, prm_.lanczos_tol,
                            prm_.steric ? prm_.steric_r0 : 0.0,
                            prm_.steric ? prm_.steric_k0 : 0.0, prm_.delta_t, prm_.mob_interval,
                            prm_.seed));
    else
        bd_.reset(new BDGpu(n, pos_.data(), io_.rdi().data(), io_.box[0], prm_.spme_xi,
                            prm_.spme_rmax, prm_.spme_dim, prm_.spme_porder, prm_.precision,
                            prm_.lanczos_maxiters, maxnrhs, prm_.lanczos_tol,
                            prm_.steric ? prm_.steric_r0 : 0.0,
                            prm_.steric ? prm_.steric_k0 : 0.0, prm_.delta_t, prm_.mob_interval,
                            prm_.seed));
    if (prm_.random) bd_->use_random(prm_.random_numvecs, prm_.random_maxnrhs);
    bd_->init();
    bd_->set_step(step_);
    if (prm_.traj_interval > 0) {
        traj_ = fopen(prm_.traj_output.c_str(), "w+");
        SDT_REQUIRE(traj_ != nullptr, SDT_EINVAL, "cannot open the trajectory file " + prm_.traj_output);
    }
}

void SimBoxGpu::advance(int nsteps) {
    SDT_REQUIRE(bd_, SDT_ESTATE, "advance called before init");
    SDT_REQUIRE(nsteps >= 0, SDT_EINVAL, "nsteps must be non-negative");
    // frames after steps s with s > traj-startstep and s % traj-interval == 0
    // (simbox_bd.cc:356-362); the device runs the steps in between unbroken
    long long left = nsteps;
    while (left > 0) {
        long long chunk = left;
        if (traj_) {
            const long long ti = prm_.traj_interval;
            long long next = (step_ / ti + 1) * ti;   // next multiple of ti after step_
            while (next <= prm_.traj_start) next += ti;
            chunk = std::min(left, next - step_);
        }
        bd_->advance((int)chunk);
        step_ += chunk;
        time_ = start_time_ + (double)(step_ - start_step_) * prm_.delta_t;
        left -= chunk;
        if (traj_ && step_ > prm_.traj_start && step_ % prm_.traj_interval == 0) {
            bd_->get_positions(pos_.data());
            SDT_REQUIRE(write_xyz_frame(traj_, npos(), io_.box, step_, time_, prm_.seed, io_.types(),
                                        pos_.data()),
                        SDT_EINVAL, "trajectory write failed");
            fflush(traj_);
        }
    }
    bd_->get_positions(pos_.data());
}

void SimBoxGpu::positions(double* out) const { memcpy(out, pos_.data(), sizeof(double) * pos_.size()); }

void SimBoxGpu::radii(double* out) const {
    memcpy(out, io_.rdi().data(), sizeof(double) * io_.rdi().size());
}

void SimBoxGpu::write_xyz(const char* path) const {
    FILE* fp = fopen(path, "w");
    SDT_REQUIRE(fp != nullptr, SDT_EINVAL, std::string("cannot open ") + path);
    const bool ok = write_xyz_frame(fp, npos(), io_.box, step_, time_, prm_.seed, io_.types(),
                                    pos_.data());
    fclose(fp);
    SDT_REQUIRE(ok, SDT_EINVAL, std::string("write failed: ") + path);
}

}  // namespace sdt
