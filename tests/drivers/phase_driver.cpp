// Uses LB2D method by method, as the reference's own
// MantiisApp::transport_simulation_loop does (mantiisApp.hpp:239-246):
// usq_old = usq, equilibrium, evolutionStepOfMultiBlock,
// calculateMacroscopicPorperties, calculateVelocityDifference.
// argv: <Nx> <input_folder/> <Nt> <out_prefix> [<is_multiblock>]
#include <iostream>
#include "include/mantiisApp.hpp"

int main(int argc, char* argv[]) {
    if (argc < 5) return 2;
    const int N = std::stoi(argv[1]);
    std::string folder = argv[2];
    const int iters = std::stoi(argv[3]);
    const std::string out = argv[4];
    const bool multi = argc > 5 && std::string(argv[5]) == "true";
    _GLOBAL_::Cl = 1.0e-9;
    _GLOBAL_::Fbody = 1.0e-8;
    _GLOBAL_::recompute();
    lattice latt;
    Shape shape(N, N, _GLOBAL_::Cl);
    shape.loadExistingModel(folder);
    Grid2D grid(shape, latt, multi);
    grid.initialize(shape);
    int nx = N, ny = N;
    LB2D lb(nx, ny, latt, grid, shape);
    lb.initialize();
    lb.setCollision(&LB2D::MRTCollisionRegularized);
    lb.setWallBoundary(&LB2D::SRBBWall);
    lb.setOpenBoundary(&LB2D::periodicBoundary);
    for (int t = 0; t < iters; t++) {
        lb.usq_old = lb.usq;
        lb.equilibrium();
        lb.evolutionStepOfMultiBlock(grid.maxLevel);
        lb.calculateMacroscopicPorperties();
        lb.calculateVelocityDifference();
    }
    auto [ux, uy, rho] = lb.completeSingleGridDomain();
    IO::writeVectorToFile(out + "ux.txt", ux);
    IO::writeVectorToFile(out + "uy.txt", uy);
    IO::writeVectorToFile(out + "rho.txt", rho);
    IO::writeVectorToFile(out + "convergence.txt", lb.diff_over_time);
    std::cout << "cells " << grid.globalGridSize << " last residual " << lb.diff << "\n";
    return 0;
}
