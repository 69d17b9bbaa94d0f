// titan::Mass of the C++ facade (include/Titan/mass.h:16-76).  Host only.
#ifndef TITAN_MI355X_MASS_H
#define TITAN_MI355X_MASS_H

#include <cstdint>
#include <vector>

#include "object.h"
#include "vec.h"

namespace titan {

class Mass {
   public:
    // m defaults to 0.1 kg; the `fixed` argument is ignored, as in the
    // reference (src/mass.cu:42-55) -- use fix()
    Mass(const Vec &position, double mass = 0.1, bool fixed = false);

    double m;  // mass in kg
    double T;  // local time
    Vec pos;   // position in m
    Vec vel;   // velocity in m/s

    void setExternalForce(const Vec &v) { extern_force = v; }
    Vec acceleration() { return acc; }

    void addConstraint(CONSTRAINT_TYPE type, const Vec &vec, double num);
    void clearConstraints(CONSTRAINT_TYPE type);
    void clearConstraints();
    void setDrag(double C);
    void fix();
    void unfix();

    struct Local {  // one local constraint, already normalised as the reference's constructors do
        int kind;
        Vec v;
        double num;
    };

   private:
    Mass();  // m = 1.0 (src/mass.cu:8-9)

    // what the device holds for this mass (the reference's CUDA_MASS)
    struct Record {
        double m = 0, T = 0;
        Vec pos, vel, acc, ext;
        bool fixed = false;
        int32_t drag = 0;
        std::vector<Local> locals;
    };

    Vec acc;
    Vec extern_force;
    bool valid = true;
    bool fixed = false;
    double drag_coefficient = 0;
    std::vector<Local> locals;  // insertion order
    Record dev;
    int64_t index = -1;         // slot in the engine body, -1 = not on the device

    friend class Simulation;
    friend class Spring;
    friend class Container;
    friend class Lattice;
    friend class Beam;
    friend class Cube;
};

}  // namespace titan

#endif
