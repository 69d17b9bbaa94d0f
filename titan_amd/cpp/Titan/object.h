// Constraints and containers of the C++ facade, with the reference's API
// (include/Titan/object.h:45-304, src/object.cu:26-515).  Host only.
#ifndef TITAN_MI355X_OBJECT_H
#define TITAN_MI355X_OBJECT_H

#include <vector>

#include "vec.h"

namespace titan {

class Mass;
class Spring;

// include/Titan/object.h:231-233
enum CONSTRAINT_TYPE { CONSTRAINT_PLANE, CONTACT_PLANE, BALL, DIRECTION };

class Constraint {  // a global shape (plane or ball)
   public:
    virtual ~Constraint() = default;
};

struct Ball : public Constraint {
    Ball(const Vec &center, double radius) : _radius(radius), _center(center) {}
    double _radius;
    Vec _center;
};

// object.h:99-121: the normal is made unit length on construction
struct ContactPlane : public Constraint {
    ContactPlane(const Vec &normal, double offset)
        : _normal(normal / normal.norm()), _offset(offset), _FRICTION_K(0.0), _FRICTION_S(0.0) {}
    Vec _normal;
    double _offset;
    double _FRICTION_K;
    double _FRICTION_S;
};

class Container {  // object.h:237-262
   public:
    virtual ~Container() {}
    void translate(const Vec &displ);
    void rotate(const Vec &axis, double angle);
    void setMassValues(double m);   // adds m to every mass, as the reference does (object.cu:143-147)
    void setSpringConstants(double k);
    void setRestLengths(double len);
    void addConstraint(CONSTRAINT_TYPE type, const Vec &v, double d);
    void clearConstraints();
    void fix();
    void add(Mass *m);
    void add(Spring *s);
    void add(Container *c);

    std::vector<Mass *> masses;
    std::vector<Spring *> springs;
};

class Cube : public Container {
   public:
    ~Cube() {}
    Cube(const Vec &center, double side_length = 1.0);
    double _side_length;
    Vec _center;
};

class Lattice : public Container {
   public:
    ~Lattice() {}
    Lattice(const Vec &center, const Vec &dims, int nx = 10, int ny = 10, int nz = 10);
    int nx, ny, nz;
    Vec _center, _dims;
};

class Beam : public Container {
   public:
    ~Beam() {}
    Beam(const Vec &center, const Vec &dims, int nx = 10, int ny = 10, int nz = 10);
    int nx, ny, nz;
    Vec _center, _dims;
};

}  // namespace titan

#endif
