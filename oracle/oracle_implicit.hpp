// Implicit step families for the oracle (test infrastructure only).
#pragma once

namespace oracle {

template <typename T>
struct UnsupportedStepper {
    void init(const Problem<T>*) {}
    int32_t step(const T*, T*, const T*, T*, T*, T*, T*, T*, T, T, bool, bool, int32_t* counters) {
        counters[0] = counters[1] = 0;
        return -1;
    }
};
template <typename T> using DIRKStepper = UnsupportedStepper<T>;
template <typename T> using FIRKStepper = UnsupportedStepper<T>;
template <typename T> using RosenbrockStepper = UnsupportedStepper<T>;

}  // namespace oracle
