// Shape::calculateProperties through this repository's include/shape.hpp,
// used the way LB2D::calculateMultigridTauS uses it (coreLB.hpp:447-456):
// a Shape of the pore image, resolution and _GLOBAL_::mfp, then the maps.
// Usage: shape_driver <input folder/> <output folder/> <n>; writes
// localporesize.dat and Kn.dat with Shape::writeToText.
#include <cstdlib>
#include <string>

#include "include/shape.hpp"

int main(int argc, char* argv[]) {
    if (argc < 4) return 2;
    std::string in = argv[1], out = argv[2];
    const int n = std::atoi(argv[3]);
    Shape shape(n, n, 1.0e-9);
    shape.loadExistingModel(in);
    _GLOBAL_::recompute();
    shape.calculateProperties(1.0e-9, _GLOBAL_::mfp);
    shape.writeToText(out);
    return 0;
}
