// titan::Spring of the C++ facade (include/Titan/spring.h:17-66).  Host only.
#ifndef TITAN_MI355X_SPRING_H
#define TITAN_MI355X_SPRING_H

#include "mass.h"
#include "vec.h"

namespace titan {

enum SpringType { PASSIVE_SOFT, PASSIVE_STIFF, ACTIVE_CONTRACT_THEN_EXPAND, ACTIVE_EXPAND_THEN_CONTRACT };

class Spring {
   public:
    Spring()
        : _left(nullptr), _right(nullptr), _k(10000.0), _rest(1.0), _type(PASSIVE_SOFT), _omega(0.0), _damping(0.0) {}
    Spring(Mass *left, Mass *right)
        : _left(left), _right(right), _k(10000.0), _type(PASSIVE_SOFT), _omega(0.0), _damping(0.0) {
        defaultLength();
    }
    Spring(Mass *left, Mass *right, double k, double rest_length)
        : _left(left), _right(right), _k(k), _rest(rest_length), _type(PASSIVE_SOFT), _omega(0.0), _damping(0.0) {}
    Spring(Mass *left, Mass *right, double k, double rest_length, SpringType type, double omega)
        : _left(left), _right(right), _k(k), _rest(rest_length), _type(type), _omega(omega), _damping(0.0) {}

    void setRestLength(double rest_length) { _rest = rest_length; }
    void defaultLength() { _rest = (_left->pos - _right->pos).norm(); }  // src/spring.cu:14
    void changeType(SpringType type, double omega) {
        _type = type;
        _omega = omega;
    }
    void addDamping(double constant) { _damping = constant; }
    void setLeft(Mass *left) { _left = left; }
    void setRight(Mass *right) { _right = right; }
    void setMasses(Mass *left, Mass *right) {
        _left = left;
        _right = right;
    }

    Mass *_left;
    Mass *_right;
    double _k;        // N/m
    double _rest;     // m
    SpringType _type;
    double _omega;    // rad/s
    double _damping;

   private:
    // what the device holds for this spring (the reference's CUDA_SPRING)
    struct Record {
        Mass *left = nullptr, *right = nullptr;
        double k = 0, rest = 0, damping = 0, omega = 0;
        int type = 0;
    } dev;
    int64_t index = -1;  // position in the engine body's spring list, -1 = not on the device
    friend class Simulation;
};

}  // namespace titan

#endif
