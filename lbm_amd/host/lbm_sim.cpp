// lbm_sim.cpp -- `lbm_sim [params.in]`: the reference's solver entry point
// (CPU/main.cpp:4-20) with the time step on the MI355X.
#include <exception>
#include <iostream>

#include "InamuroSolver.hpp"

int main(int argc, char* argv[])
{
    try {
        lbm::InamuroSolver solver(argc > 1 ? argv[1] : "params.in");
        solver.run();
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
