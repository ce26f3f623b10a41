// LBMBase.hpp -- the abstract LBM step interface of the reference
// (CPU/LBMBase.hpp:12-49), kept with the same virtual members so that a
// driver written against it runs the GPU implementation unchanged.
#pragma once

#include <iostream>
#include <string>
#include <vector>

namespace lbm {

class LBMBase {
protected:
    // Nested containers of the reference's interface ([q][x][y][z] and
    // [x][y][z]); the GPU implementation does not use them for its state.
    using Vector4D = std::vector<std::vector<std::vector<std::vector<double>>>>;
    using Vector3D = std::vector<std::vector<std::vector<double>>>;

public:
    LBMBase() = default;
    virtual ~LBMBase() = default;

    virtual void collision() = 0;
    virtual void stream(Vector4D& dist) = 0;
    virtual void applyBoundaryConditions(Vector4D& dist) = 0;
    virtual void getMacro() = 0;
    virtual void performTimeStep() = 0;
    virtual void writeResults(int timeStep) = 0;
    virtual void getGridSize(int& nx, int& ny, int& nz) const = 0;
    virtual std::string getAlgorithmName() const = 0;

    virtual void printInfo() const
    {
        int nx, ny, nz;
        getGridSize(nx, ny, nz);
        std::cout << "\n=== LBM ===\nalgorithm: " << getAlgorithmName() << "\ngrid: " << nx << " x "
                  << ny << " x " << nz << "\n===========" << std::endl;
    }
};

} // namespace lbm
