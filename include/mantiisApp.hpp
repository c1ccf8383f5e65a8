// Drop-in counterpart of the reference's include/mantiisApp.hpp: Settings and
// MantiisApp with the same construction (argv: <num_threads> <is_multiblock>
// <Nx> <input_folder/> <Nt>), init() and run() (mantiisApp.hpp:46-432).
// The step loops issue one fused HIP pass per time step through LB2D.
#pragma once

#include <chrono>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "comm.hpp"
#include "lb2d.hpp"

struct Settings {
    bool load_existing_model = 1;
    bool dump_final_results = 1;
    int verbose = 1;
    int dump_every_timestep = 1;
    int dump_distributions_first_timestep = 0;
    int dump_macroscopic_every_timestep = 0;
    double threshold = 1.0e-10;
    double mfp = _GLOBAL_::mfp;
    double Fbody = _GLOBAL_::Fbody;
    double dx = _GLOBAL_::Cl;
    double T = _GLOBAL_::T_phy;
    double P = _GLOBAL_::P_phy;
    int collision_method = 2;
    int wall_boundary = 2;
    std::string simulation_type = "transport";
    // Steps fused between two residual checks (no reference counterpart; 1
    // reproduces the reference's per-step residual and break test exactly).
    int check_every = 1;
    // No reference counterpart: start from the populations of a dump
    // (f.txt, or f.bin written with binary_output) instead of rest -- a resume
    // or a warm start -- and also write the final fields as raw binary.
    std::string initial_distributions = "";
    bool binary_output = false;
};

// mantiisApp.hpp:66-73: dumps are written by threads while the run goes on.
template <typename T>
struct WriterThread {
    static void write(const std::string& fName, const std::vector<T>& data) { IO::writeVectorToFile(fName, data); }
};

class MantiisApp {
public:
    bool is_multiblock;
    int Nx;
    int Ny;
    std::string folder;
    int iters;

private:
    bool load_existing_model = 1;
    bool dump_final_results = 1;
    double tol = 1.0e-10;
    int verbose = 1;
    int dump_every_timestep = 1;
    int dump_distributions_first_timestep = 0;
    int dump_macroscopic_every_timestep = 0;
    int check_every = 1;
    std::string initial_distributions;
    bool binary_output = false;
    std::vector<std::thread> writer_threads;
    int proc_id = 0;
    int num_procs = 1;
    int num_threads = 1;
    int collision_method;
    int wall_boundary;
    std::string simulation_type = "transport";
    lattice latt;
    std::unique_ptr<Shape> shape;
    std::unique_ptr<Grid2D> grid;
    std::unique_ptr<LB2D> lb;

public:
    // mantiisApp.hpp:109-156
    MantiisApp(int argc, char* argv[], const Settings& settings) {
        load_existing_model = settings.load_existing_model;
        dump_final_results = settings.dump_final_results;
        verbose = settings.verbose > 0 ? settings.verbose : 1;
        dump_every_timestep = settings.dump_every_timestep;
        dump_distributions_first_timestep = settings.dump_distributions_first_timestep;
        dump_macroscopic_every_timestep = settings.dump_macroscopic_every_timestep;
        tol = settings.threshold;
        check_every = settings.check_every > 0 ? settings.check_every : 1;
        initial_distributions = settings.initial_distributions;
        binary_output = settings.binary_output;
        _GLOBAL_::mfp = settings.mfp;
        _GLOBAL_::Fbody = settings.Fbody;
        _GLOBAL_::Cl = settings.dx;
        _GLOBAL_::T_phy = settings.T;
        _GLOBAL_::P_phy = settings.P;
        _GLOBAL_::recompute();
        collision_method = settings.collision_method;
        wall_boundary = settings.wall_boundary;
        simulation_type = settings.simulation_type;
        if (argc < 6) {
            std::cerr << "Usage: " << argv[0] << " <num_threads> <is_multiblock> <Nx> <input_folder/> <Nt>\n";
            std::exit(EXIT_FAILURE);
        }
        num_threads = std::stoi(argv[1]);
        is_multiblock = (std::string(argv[2]) == "true");
        Nx = std::stoi(argv[3]);
        Ny = Nx;
        folder = argv[4];
        iters = std::stoi(argv[5]);
        mantiis_parallel::launchMPI(proc_id, num_procs, num_threads);
        std::cout << "Mean free path: " << _GLOBAL_::mfp << "\n";
        std::cout << "Cl: " << _GLOBAL_::Cl << "\n";
        std::cout << "Temperature: " << _GLOBAL_::T_phy << "\n";
        std::cout << "Pressure: " << _GLOBAL_::P_phy << "\n";
        std::cout << "Fbody: " << _GLOBAL_::Fbody << "\n";
        std::cout << "Force conversion factor: " << std::setprecision(15) << _GLOBAL_::Cf << "\n";
        std::cout << "Input density: " << std::setprecision(15) << _GLOBAL_::rhoin << "\n";
    }

    // mantiisApp.hpp:158-237
    void init() {
        shape = std::make_unique<Shape>(Nx, Ny, _GLOBAL_::Cl);
        if (load_existing_model) shape->loadExistingModel(folder);
        grid = std::make_unique<Grid2D>(*shape, latt, is_multiblock);
        grid->initialize(*shape);
        std::cout << "The number of active cells: " << grid->globalGridSize << "\n";
        std::cout << "Initializing Simulation...\n";
        lb = std::make_unique<LB2D>(Nx, Ny, latt, *grid, *shape);
        lb->initialize();
        switch (collision_method) {
            case 2: lb->setCollision(&LB2D::MRTCollisionRegularized); break;
            case 1: lb->setCollision(&LB2D::MRTCollisionNonRegularized); break;
            case 0: lb->setCollision(&LB2D::SRTCollision); break;
            default:
                throw std::runtime_error(
                    "Incorrect keyword for collision: available options are\n"
                    "\t 2 for regularized MRT\n\t 1 for simple MRT\n\t 0 for simple BGK\n");
        }
        switch (wall_boundary) {
            case 2: lb->setWallBoundary(&LB2D::SRBBWall); break;
            case 1: lb->setWallBoundary(&LB2D::MDBBWall); break;
            case 0: lb->setWallBoundary(&LB2D::BBWall); break;
            default:
                throw std::runtime_error(
                    "Incorrect keyword for wall boundary method: available options are\n"
                    "\t 2 for SRBBWall\n\t 1 for MDBBWall\n\t 0 for BBWall\n");
        }
        lb->setOpenBoundary(&LB2D::periodicBoundary);
        if (!initial_distributions.empty()) {
            std::vector<double> f0((size_t)Nx * Ny * 9);
            const std::string& p = initial_distributions;
            if (p.size() > 4 && p.compare(p.size() - 4, 4, ".bin") == 0) IO::loadBinary(p, f0);
            else IO::loadDataFromFile(p, f0);
            lb->loadDistributions(f0);
        }
    }

    ~MantiisApp() { join_writers(); }

    // mantiisApp.hpp:239-284 as fused steps
    void transport_simulation_loop(int n = 1) { lb->stepTransport(n); }
    void adsorption_simulation_loop(int n = 1) { lb->stepAdsorption(n); }
    void coupled_simulation_loop(int n = 1) { lb->stepCoupled(n); }

    // mantiisApp.hpp:309-320, 334-348: finish the earlier writes, then hand
    // this dump's vectors to new threads
    void join_writers() {
        for (auto& th : writer_threads)
            if (th.joinable()) th.join();
        writer_threads.clear();
    }
    void write_async(const std::string& name, std::vector<double>&& v) {
        for (auto it = writer_threads.begin(); it != writer_threads.end();) {
            if (it->joinable()) {
                it->join();
                it = writer_threads.erase(it);
            } else {
                ++it;
            }
        }
        writer_threads.emplace_back(&WriterThread<double>::write, name, std::move(v));
    }

    // mantiisApp.hpp:286-430
    void run() {
        using namespace std::chrono;
        const auto start = high_resolution_clock::now();
        auto run_loop = [&](auto step) {
            for (int i = 0; i < iters;) {
                const int n = std::min(check_every, iters - i);
                step(n);
                i += n;
                lb->t += n - 1;
                if (lb->t % verbose == 0)
                    std::cout << "Timestep: " << lb->t << ", residual difference: " << lb->diff << "\n";
                if (lb->diff < tol) break;
                if (lb->t < dump_distributions_first_timestep)
                    write_async(folder + "f_" + std::to_string(lb->t) + ".txt", lb->prepareDistributions());
                if (dump_macroscopic_every_timestep > 0 && lb->t % dump_macroscopic_every_timestep == 0) {
                    auto d = lb->completeSingleGridDomain();
                    write_async(folder + "ux_" + std::to_string(lb->t) + ".txt", std::move(std::get<0>(d)));
                    write_async(folder + "uy_" + std::to_string(lb->t) + ".txt", std::move(std::get<1>(d)));
                }
                lb->t++;
            }
            join_writers();
        };
        if (simulation_type == "transport") {
            run_loop([&](int n) { transport_simulation_loop(n); });
        } else if (simulation_type == "adsorption") {
            run_loop([&](int n) { adsorption_simulation_loop(n); });
        } else if (simulation_type == "coupled") {
            std::cout << "======================= Adsorption ONLY for initial condition =======================\n";
            const double fb = _GLOBAL_::Fbody;
            lb->setBodyForce(0.0);
            const int it = iters;
            iters = 100;
            run_loop([&](int n) { adsorption_simulation_loop(n); });
            iters = it;
            lb->setBodyForce(fb);
            std::cout << "======================= Coupling =======================\n";
            run_loop([&](int n) { coupled_simulation_loop(n); });
        } else {
            throw std::runtime_error("Unknown simulation type: " + simulation_type);
        }
        if (dump_final_results) {
            auto d = lb->completeSingleGridDomain();
            const std::vector<double> fd = lb->prepareDistributions();
            IO::writeVectorToFile(folder + "f.txt", fd);
            IO::writeVectorToFile(folder + "ux.txt", std::get<0>(d));
            IO::writeVectorToFile(folder + "uy.txt", std::get<1>(d));
            IO::writeVectorToFile(folder + "rho.txt", std::get<2>(d));
            IO::writeVectorToFile(folder + "convergence.txt", lb->diff_over_time);
            if (binary_output) {
                IO::writeBinary(folder + "f.bin", fd);
                IO::writeBinary(folder + "ux.bin", std::get<0>(d));
                IO::writeBinary(folder + "uy.bin", std::get<1>(d));
                IO::writeBinary(folder + "rho.bin", std::get<2>(d));
            }
            // FP64 permeability report (no reference counterpart), for the
            // flow runs: permeability.txt, one "key value" per line
            LB2D::Permeability perm{};
            const bool flow = simulation_type != "adsorption";
            if (flow) {
                perm = lb->permeability();
                std::ofstream pf(folder + "permeability.txt");
                pf << std::setprecision(15);
                pf << "darcy_velocity " << perm.darcy_velocity << "\n";
                pf << "mean_ux " << perm.mean_ux << "\n";
                pf << "flux " << perm.flux << "\n";
                pf << "mass " << perm.mass << "\n";
                pf << "porosity " << perm.porosity << "\n";
                pf << "tau " << perm.tau << "\n";
                pf << "k_lattice " << perm.k_lattice << "\n";
                pf << "k_m2 " << perm.k_m2 << "\n";
                std::cout << "Permeability (m^2): " << std::setprecision(15) << perm.k_m2 << "\n";
            }
            std::ofstream details(folder + "SimulationDetails.txt");
            details << "Domain dimension: " << shape->Nx << " x " << shape->Ny << "\n";
            details << "Tolerance was set to " << tol << "\n";
            details << "Number timesteps executed: " << lb->t << "\n";
            details << "The number of active cells: " << grid->globalGridSize << "\n";
            if (flow) details << "Permeability (m^2): " << std::setprecision(15) << perm.k_m2 << "\n";
            const auto ms = duration_cast<milliseconds>(high_resolution_clock::now() - start).count();
            std::cout << "Total run time (milliseconds) " << std::setprecision(15) << ms << std::endl;
            details << "Total run time (milliseconds): " << std::setprecision(15) << ms;
        }
    }
};
